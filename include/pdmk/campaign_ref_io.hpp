// Reference-data files ("PDMC"): oracle potentials at sampled targets together with the
// problem they belong to, protected by a SHA-256 of the body.  Same names, constants and
// on-disk layout as cpp/include/pdmk/campaign_ref_io.hpp (header lines 12-44, body order
// pack_body 114-129, checks in read_ref 151-169), so files move between the two libraries.
//
// File = 80-byte little-endian header, then the body
//   cell[9] (FP64, column-major) | ref_idx[M] (uint32) | ref_u[M] (FP64) | pos[3N] (FP64) | chg[N] (FP64)
// with header.body_len = body bytes and header.body_sha256 = SHA-256(body).
//
// The SHA-256 here is written from FIPS 180-4: its round constants and initial hash value
// are computed exactly from their definition (the first 32 fractional bits of the cube
// roots of the first 64 primes and of the square roots of the first 8 primes) by integer
// root extraction.
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace pdmk::campaign {

constexpr uint32_t MAGIC = 0x50444D43u;  // bytes 'C' 'M' 'D' 'P' on disk, "PDMC" as a little-endian word
constexpr uint32_t VERSION = 1u;
constexpr uint32_t ENDIAN_TAG = 0x01020304u;
constexpr uint32_t HEADER_LEN = 80u;
constexpr uint32_t DEFAULT_M = 1000u;

#pragma pack(push, 1)
struct RefHeader {
    uint32_t magic;
    uint32_t version;
    uint32_t endian_tag;
    uint32_t header_len;
    uint32_t config_id;
    uint32_t N;
    uint8_t digits;
    uint8_t prec;
    uint8_t _pad[2];
    uint32_t M;
    uint64_t seed;
    uint64_t body_len;
    uint8_t body_sha256[32];
};
#pragma pack(pop)
static_assert(sizeof(RefHeader) == 80, "RefHeader must be 80 bytes");

struct RefData {
    RefHeader header;
    double cell[9];                 // column-major
    std::vector<uint32_t> ref_idx;  // M sampled target indices
    std::vector<double> ref_u;      // M reference potentials
    std::vector<double> pos;        // 3N, column-major 3 x N
    std::vector<double> chg;        // N
};

namespace detail {

using u128 = unsigned __int128;

// floor(v^(1/k)) for k = 2, 3 by bisection on 128-bit integers (v < 2^110).
inline uint64_t iroot(u128 v, int k) {
    uint64_t lo = 0, hi = uint64_t(1) << 40;
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        u128 p = mid;
        for (int i = 1; i < k; ++i) p *= mid;
        (p <= v ? lo : hi) = mid;
    }
    return lo;
}

// First 32 fractional bits of the k-th root of each of the first n primes.
template <int N> std::array<uint32_t, N> prime_root_bits(int k) {
    std::array<uint32_t, N> out{};
    int found = 0;
    for (uint32_t p = 2; found < N; ++p) {
        bool prime = true;
        for (uint32_t d = 2; d * d <= p; ++d)
            if (p % d == 0) prime = false;
        if (!prime) continue;
        // floor(root(p) * 2^32) = floor(root(p * 2^(32 k)))
        const u128 v = u128(p) << (32 * k);
        out[found++] = uint32_t(iroot(v, k));  // the integer part drops out of the low 32 bits
    }
    return out;
}

inline const std::array<uint32_t, 64> &round_constants() {
    static const std::array<uint32_t, 64> k = prime_root_bits<64>(3);
    return k;
}

inline const std::array<uint32_t, 8> &initial_hash() {
    static const std::array<uint32_t, 8> h = prime_root_bits<8>(2);
    return h;
}

inline uint32_t ror(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// One 64-byte block (FIPS 180-4, 6.2.2).
inline void compress(uint32_t st[8], const uint8_t *blk) {
    const auto &K = round_constants();
    uint32_t w[64];
    for (int t = 0; t < 16; ++t)
        w[t] = uint32_t(blk[4 * t]) << 24 | uint32_t(blk[4 * t + 1]) << 16 | uint32_t(blk[4 * t + 2]) << 8 |
               uint32_t(blk[4 * t + 3]);
    for (int t = 16; t < 64; ++t) {
        const uint32_t s0 = ror(w[t - 15], 7) ^ ror(w[t - 15], 18) ^ (w[t - 15] >> 3);
        const uint32_t s1 = ror(w[t - 2], 17) ^ ror(w[t - 2], 19) ^ (w[t - 2] >> 10);
        w[t] = s1 + w[t - 7] + s0 + w[t - 16];
    }
    uint32_t v[8];
    std::memcpy(v, st, sizeof(v));
    for (int t = 0; t < 64; ++t) {
        // v = (a, b, c, d, e, f, g, h)
        const uint32_t S1 = ror(v[4], 6) ^ ror(v[4], 11) ^ ror(v[4], 25);
        const uint32_t Ch = (v[4] & v[5]) ^ (~v[4] & v[6]);
        const uint32_t T1 = v[7] + S1 + Ch + K[t] + w[t];
        const uint32_t S0 = ror(v[0], 2) ^ ror(v[0], 13) ^ ror(v[0], 22);
        const uint32_t Maj = (v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]);
        for (int i = 7; i > 0; --i) v[i] = v[i - 1];
        v[4] += T1;
        v[0] = T1 + S0 + Maj;
    }
    for (int i = 0; i < 8; ++i) st[i] += v[i];
}

}  // namespace detail

// SHA-256 of data[0 .. len) into out[32] (big-endian words).
inline void sha256(const uint8_t *data, size_t len, uint8_t out[32]) {
    uint32_t st[8];
    const auto &H0 = detail::initial_hash();
    for (int i = 0; i < 8; ++i) st[i] = H0[i];
    size_t done = 0;
    for (; len - done >= 64; done += 64) detail::compress(st, data + done);
    // padding: 0x80, zeros, then the bit length as a 64-bit big-endian integer
    uint8_t last[128] = {0};
    const size_t rest = len - done;
    if (rest) std::memcpy(last, data + done, rest);
    last[rest] = 0x80;
    const size_t nb = rest + 9 <= 64 ? 64 : 128;
    const uint64_t bits = uint64_t(len) * 8;
    for (int i = 0; i < 8; ++i) last[nb - 8 + i] = uint8_t(bits >> (56 - 8 * i));
    detail::compress(st, last);
    if (nb == 128) detail::compress(st, last + 64);
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = uint8_t(st[i] >> (24 - 8 * j));
}

// The body bytes of d (cell, ref_idx, ref_u, pos, chg, in that order).
inline std::vector<uint8_t> pack_body(const RefData &d) {
    const size_t parts[5] = {sizeof(d.cell), d.ref_idx.size() * sizeof(uint32_t), d.ref_u.size() * sizeof(double),
                             d.pos.size() * sizeof(double), d.chg.size() * sizeof(double)};
    const void *src[5] = {d.cell, d.ref_idx.data(), d.ref_u.data(), d.pos.data(), d.chg.data()};
    std::vector<uint8_t> body(parts[0] + parts[1] + parts[2] + parts[3] + parts[4]);
    size_t at = 0;
    for (int i = 0; i < 5; ++i) {
        if (parts[i]) std::memcpy(body.data() + at, src[i], parts[i]);
        at += parts[i];
    }
    return body;
}

// Writes d to path; the fixed header fields, M, body_len and the digest are filled in here.
inline void write_ref(const std::string &path, const RefData &d) {
    RefHeader h = d.header;
    h.magic = MAGIC;
    h.version = VERSION;
    h.endian_tag = ENDIAN_TAG;
    h.header_len = HEADER_LEN;
    h.M = uint32_t(d.ref_idx.size());
    const std::vector<uint8_t> body = pack_body(d);
    h.body_len = body.size();
    sha256(body.data(), body.size(), h.body_sha256);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open for write: " + path);
    const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(body.data(), 1, body.size(), f) == body.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) throw std::runtime_error("write failed: " + path);
}

// Reads and verifies path (magic, version, endian tag, header length, body length and
// SHA-256); throws std::runtime_error on any mismatch.
inline RefData read_ref(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open for read: " + path);
    RefData d;
    auto fail = [&](const char *why) {
        std::fclose(f);
        throw std::runtime_error(std::string(why) + ": " + path);
    };
    if (std::fread(&d.header, sizeof(RefHeader), 1, f) != 1) fail("header short read");
    const RefHeader &h = d.header;
    if (h.magic != MAGIC) fail("bad magic");
    if (h.version != VERSION) fail("bad version");
    if (h.endian_tag != ENDIAN_TAG) fail("bad endian tag");
    if (h.header_len != HEADER_LEN) fail("bad header_len");
    const uint64_t want = sizeof(d.cell) + uint64_t(h.M) * (sizeof(uint32_t) + sizeof(double)) + uint64_t(h.N) * 4 * sizeof(double);
    if (h.body_len != want) fail("body length does not match N and M");
    std::vector<uint8_t> body(h.body_len);
    if (!body.empty() && std::fread(body.data(), 1, body.size(), f) != body.size()) fail("body short read");
    std::fclose(f);
    uint8_t dig[32];
    sha256(body.data(), body.size(), dig);
    if (std::memcmp(dig, h.body_sha256, 32) != 0) throw std::runtime_error("sha256 mismatch: " + path);
    d.ref_idx.resize(h.M);
    d.ref_u.resize(h.M);
    d.pos.resize(size_t(3) * h.N);
    d.chg.resize(h.N);
    void *dst[5] = {d.cell, d.ref_idx.data(), d.ref_u.data(), d.pos.data(), d.chg.data()};
    const size_t parts[5] = {sizeof(d.cell), h.M * sizeof(uint32_t), h.M * sizeof(double), d.pos.size() * sizeof(double),
                             d.chg.size() * sizeof(double)};
    size_t at = 0;
    for (int i = 0; i < 5; ++i) {
        if (parts[i]) std::memcpy(dst[i], body.data() + at, parts[i]);
        at += parts[i];
    }
    return d;
}

}  // namespace pdmk::campaign
