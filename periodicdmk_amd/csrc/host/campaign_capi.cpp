// C entry points over the header-only reference-data file format (pdmk_campaign.h).
#include <pdmk/campaign_ref_io.hpp>
#include <pdmk/pdmk_campaign.h>

#include <cstdio>
#include <cstring>
#include <exception>

using namespace pdmk::campaign;

namespace {

int fail(const char *what, const std::exception &e) {
    std::fprintf(stderr, "%s: %s\n", what, e.what());
    return -1;
}

}  // namespace

extern "C" {

int pdmk_campaign_sha256(const unsigned char *data, size_t len, unsigned char *out32) {
    if ((!data && len) || !out32) return -1;
    sha256(data, len, out32);
    return 0;
}

int pdmk_campaign_write(const char *path, unsigned config_id, unsigned N, int digits, int prec, unsigned long long seed,
                        const double *cell9, unsigned M, const unsigned *ref_idx, const double *ref_u,
                        const double *pos_3xN, const double *chg) {
    if (!path || !cell9 || (M && (!ref_idx || !ref_u)) || (N && (!pos_3xN || !chg))) return -1;
    if (digits < 0 || digits > 255 || prec < 0 || prec > 255) return -1;
    try {
        RefData d{};
        d.header.config_id = config_id;
        d.header.N = N;
        d.header.digits = uint8_t(digits);
        d.header.prec = uint8_t(prec);
        d.header.seed = seed;
        std::memcpy(d.cell, cell9, sizeof(d.cell));
        d.ref_idx.assign(ref_idx, ref_idx + M);
        d.ref_u.assign(ref_u, ref_u + M);
        d.pos.assign(pos_3xN, pos_3xN + size_t(3) * N);
        d.chg.assign(chg, chg + N);
        write_ref(path, d);
        return 0;
    } catch (const std::exception &e) {
        return fail("pdmk_campaign_write", e);
    }
}

int pdmk_campaign_read_header(const char *path, unsigned *config_id, unsigned *N, int *digits, int *prec,
                              unsigned long long *seed, unsigned *M) {
    if (!path) return -1;
    try {
        const RefData d = read_ref(path);
        if (config_id) *config_id = d.header.config_id;
        if (N) *N = d.header.N;
        if (digits) *digits = d.header.digits;
        if (prec) *prec = d.header.prec;
        if (seed) *seed = d.header.seed;
        if (M) *M = d.header.M;
        return 0;
    } catch (const std::exception &e) {
        return fail("pdmk_campaign_read_header", e);
    }
}

int pdmk_campaign_read(const char *path, double *cell9, unsigned *ref_idx, double *ref_u, double *pos_3xN, double *chg) {
    if (!path) return -1;
    try {
        const RefData d = read_ref(path);
        if (cell9) std::memcpy(cell9, d.cell, sizeof(d.cell));
        if (ref_idx && !d.ref_idx.empty()) std::memcpy(ref_idx, d.ref_idx.data(), d.ref_idx.size() * sizeof(uint32_t));
        if (ref_u && !d.ref_u.empty()) std::memcpy(ref_u, d.ref_u.data(), d.ref_u.size() * sizeof(double));
        if (pos_3xN && !d.pos.empty()) std::memcpy(pos_3xN, d.pos.data(), d.pos.size() * sizeof(double));
        if (chg && !d.chg.empty()) std::memcpy(chg, d.chg.data(), d.chg.size() * sizeof(double));
        return 0;
    } catch (const std::exception &e) {
        return fail("pdmk_campaign_read", e);
    }
}

}  // extern "C"
