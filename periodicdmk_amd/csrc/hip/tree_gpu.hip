// Tree construction with the per-point work on the GPU (see tree_gpu.hpp).
#include "tree_gpu.hpp"

#include "plan.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>

namespace pdmk::gpu {
namespace {

constexpr int TB = 256;

struct Bounds {
    double lo[3], hi[3];
};

// Image of point p under translation t: kept iff inside the extended grid (the comparison
// sequence of host::make_geometry).
__device__ __forceinline__ bool image_inside(const double *__restrict__ r, int p, const double *__restrict__ t,
                                             const Bounds &b) {
    const double x = r[3 * (size_t)p] + t[0], y = r[3 * (size_t)p + 1] + t[1], z = r[3 * (size_t)p + 2] + t[2];
    return !(x < b.lo[0] || x > b.hi[0] || y < b.lo[1] || y > b.hi[1] || z < b.lo[2] || z > b.hi[2]);
}

// cnt[s][blk] = kept images of translation s among the points of block blk.
__global__ __launch_bounds__(TB) void k_image_count(const double *__restrict__ r, int n, const double *__restrict__ shifts,
                                                    Bounds b, int *__restrict__ cnt) {
    const int p = blockIdx.x * TB + threadIdx.x;
    const bool in = p < n && image_inside(r, p, shifts + 3 * blockIdx.y, b);
    const int c = __syncthreads_count(in);
    if (threadIdx.x == 0) cnt[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = c;
}

// Writes the kept images at n + off[s][blk] + (rank within the block, in point order).
__global__ __launch_bounds__(TB) void k_image_write(const double *__restrict__ r, int n, const double *__restrict__ shifts,
                                                    Bounds b, const int *__restrict__ off, double *__restrict__ pos,
                                                    int *__restrict__ img_src) {
    __shared__ int wsum[TB / 64];
    const int p = blockIdx.x * TB + threadIdx.x;
    const double *t = shifts + 3 * blockIdx.y;
    const bool in = p < n && image_inside(r, p, t, b);
    const unsigned long long m = __ballot(in);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int k = off[(size_t)blockIdx.y * gridDim.x + blockIdx.x] + __popcll(m & ((1ull << lane) - 1));
    for (int i = 0; i < w; ++i) k += wsum[i];
    if (in) {
        img_src[k] = p;
        for (int d = 0; d < 3; ++d) pos[3 * ((size_t)n + k) + d] = r[3 * (size_t)p + d] + t[d];
    }
}

struct Grid0 {
    double xelo[3], inv;
    int m[3];
};

// Level-0 box of every point (host grid_sort's expression), its path key and the counts.
__global__ __launch_bounds__(TB) void k_level0(const double *__restrict__ pos, int npts, Grid0 g, int kshift,
                                               int *__restrict__ box, unsigned long long *__restrict__ key,
                                               int *__restrict__ cnt) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= npts) return;
    int ii[3];
    for (int d = 0; d < 3; ++d) {
        const int v = (int)((pos[3 * (size_t)i + d] - g.xelo[d]) * g.inv);
        ii[d] = min(max(v, 0), g.m[d] - 1);
    }
    const int b = (ii[2] * g.m[1] + ii[1]) * g.m[0] + ii[0];
    box[i] = b;
    key[i] = (unsigned long long)b << kshift;
    atomicAdd(&cnt[b], 1);
}

// Points whose box was just split move to the child (bit d: x_d >= centre_d); the child
// index enters the path key at `shift`.
__global__ __launch_bounds__(TB) void k_distribute(const double *__restrict__ pos, int npts, int *__restrict__ box,
                                                   unsigned long long *__restrict__ key, const int *__restrict__ first_child,
                                                   const double *__restrict__ center, int shift, int *__restrict__ cnt) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= npts) return;
    const int b = box[i], fc = first_child[b];
    if (fc < 0) return;
    const double *x = pos + 3 * (size_t)i, *c = center + 3 * (size_t)b;
    const int oct = (x[0] >= c[0] ? 1 : 0) | (x[1] >= c[1] ? 2 : 0) | (x[2] >= c[2] ? 4 : 0);
    box[i] = fc + oct;
    key[i] |= (unsigned long long)oct << shift;
    atomicAdd(&cnt[fc + oct], 1);
}

__global__ __launch_bounds__(TB) void k_iota(int *__restrict__ v, int n) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < n) v[i] = i;
}

__global__ __launch_bounds__(TB) void k_zero_int(int *__restrict__ v, int n) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < n) v[i] = 0;
}

__global__ __launch_bounds__(TB) void k_gather3(const double *__restrict__ pos, const int *__restrict__ order, int n,
                                                double *__restrict__ out) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const size_t j = 3 * (size_t)order[i];
    for (int d = 0; d < 3; ++d) out[3 * (size_t)i + d] = pos[j + d];
}

inline unsigned blocks(long n) { return (unsigned)((n + TB - 1) / TB); }

struct TooDeep : std::runtime_error {
    TooDeep() : std::runtime_error("tree deeper than the path keys") {}
};

// Sources are the nsall points (originals, then images); targets are the n originals, so
// they share the first n positions.
struct GpuPoints final : host::PointBackend {
    hipStream_t s;
    int n = 0, nsall = 0, D = 0;  // D: tree levels the path keys hold
    DevBuf pos, sbox, tbox, skey, tkey, fc, cen, cs, ct;
    size_t cap = 0;

    void ensure_boxes(int nb) {
        if ((size_t)nb <= cap) return;
        cap = std::max((size_t)nb, 2 * cap);
        fc.alloc(sizeof(int) * cap);
        cen.alloc(sizeof(double) * 3 * cap);
        cs.alloc(sizeof(int) * cap);
        ct.alloc(sizeof(int) * cap);
    }

    void zero_counts(int from, int to) {
        hipLaunchKernelGGL(k_zero_int, dim3(blocks(to - from)), dim3(TB), 0, s, cs.as<int>() + from, to - from);
        hipLaunchKernelGGL(k_zero_int, dim3(blocks(to - from)), dim3(TB), 0, s, ct.as<int>() + from, to - from);
        check(hipGetLastError(), "k_zero_int");
    }

    void download_counts(int from, int to, std::vector<int> &sc, std::vector<int> &tc) {
        sc.resize(to - from);
        tc.resize(to - from);
        check(hipMemcpyAsync(sc.data(), cs.as<int>() + from, sizeof(int) * (to - from), hipMemcpyDeviceToHost, s), "D2H");
        check(hipMemcpyAsync(tc.data(), ct.as<int>() + from, sizeof(int) * (to - from), hipMemcpyDeviceToHost, s), "D2H");
        check(hipStreamSynchronize(s), "tree counts");
    }

    void level0(const host::Tree &t, std::vector<int> &sc, std::vector<int> &tc) override {
        const int nb0 = t.nboxes;
        int bits = 1;
        while ((1L << bits) < nb0) ++bits;
        D = (64 - bits) / 3;
        ensure_boxes(nb0);
        zero_counts(0, nb0);
        Grid0 g{};
        for (int d = 0; d < 3; ++d) {
            g.xelo[d] = t.xelo[d];
            g.m[d] = t.mle[d];
        }
        g.inv = 1.0 / t.rc;
        hipLaunchKernelGGL(k_level0, dim3(blocks(nsall)), dim3(TB), 0, s, pos.as<double>(), nsall, g, 3 * D, sbox.as<int>(),
                           skey.as<unsigned long long>(), cs.as<int>());
        hipLaunchKernelGGL(k_level0, dim3(blocks(n)), dim3(TB), 0, s, pos.as<double>(), n, g, 3 * D, tbox.as<int>(),
                           tkey.as<unsigned long long>(), ct.as<int>());
        check(hipGetLastError(), "k_level0");
        download_counts(0, nb0, sc, tc);
    }

    void distribute(host::Tree &t, const std::vector<int> &split) override {
        const int L = t.level[split.front()];
        if (L >= D) throw TooDeep();
        const int from = t.first_child[split.front()], to = t.nboxes;
        ensure_boxes(to);
        check(hipMemcpyAsync(fc.p, t.first_child.data(), sizeof(int) * to, hipMemcpyHostToDevice, s), "H2D");
        check(hipMemcpyAsync(cen.p, t.center.data(), sizeof(double) * 3 * to, hipMemcpyHostToDevice, s), "H2D");
        zero_counts(from, to);
        const int shift = 3 * (D - 1 - L);
        hipLaunchKernelGGL(k_distribute, dim3(blocks(nsall)), dim3(TB), 0, s, pos.as<double>(), nsall, sbox.as<int>(),
                           skey.as<unsigned long long>(), fc.as<int>(), cen.as<double>(), shift, cs.as<int>());
        hipLaunchKernelGGL(k_distribute, dim3(blocks(n)), dim3(TB), 0, s, pos.as<double>(), n, tbox.as<int>(),
                           tkey.as<unsigned long long>(), fc.as<int>(), cen.as<double>(), shift, ct.as<int>());
        check(hipGetLastError(), "k_distribute");
        std::vector<int> sc, tc;
        download_counts(from, to, sc, tc);
        // children ranges follow the parent's in child order (host::build_tree layout)
        for (int b : split) {
            const int f = t.first_child[b];
            if (t.src_count[b] > 0)
                for (int k = 0, at = t.src_begin[b]; k < 8; ++k) {
                    t.src_begin[f + k] = at;
                    t.src_count[f + k] = sc[f + k - from];
                    at += sc[f + k - from];
                }
            if (t.trg_count[b] > 0)
                for (int k = 0, at = t.trg_begin[b]; k < 8; ++k) {
                    t.trg_begin[f + k] = at;
                    t.trg_count[f + k] = tc[f + k - from];
                    at += tc[f + k - from];
                }
        }
    }

    // stable radix sort of (path key, index); the sorted index order is the host's
    void sort(const DevBuf &key, int npts, std::vector<int> &order, std::vector<double> &sorted) {
        DevBuf kout, vin, vout, tmp, spos;
        kout.alloc(sizeof(unsigned long long) * npts);
        vin.alloc(sizeof(int) * npts);
        vout.alloc(sizeof(int) * npts);
        hipLaunchKernelGGL(k_iota, dim3(blocks(npts)), dim3(TB), 0, s, vin.as<int>(), npts);
        check(hipGetLastError(), "k_iota");
        size_t bytes = 0;
        check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key.as<unsigned long long>(), kout.as<unsigned long long>(),
                                                 vin.as<int>(), vout.as<int>(), npts, 0, 64, s),
              "radix sort size");
        tmp.alloc(std::max(bytes, (size_t)1));
        check(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, key.as<unsigned long long>(), kout.as<unsigned long long>(),
                                                 vin.as<int>(), vout.as<int>(), npts, 0, 64, s),
              "radix sort");
        spos.alloc(sizeof(double) * 3 * npts);
        hipLaunchKernelGGL(k_gather3, dim3(blocks(npts)), dim3(TB), 0, s, pos.as<double>(), vout.as<int>(), npts,
                           spos.as<double>());
        check(hipGetLastError(), "k_gather3");
        order.resize(npts);
        sorted.resize(3 * (size_t)npts);
        check(hipMemcpyAsync(order.data(), vout.p, sizeof(int) * npts, hipMemcpyDeviceToHost, s), "D2H");
        check(hipMemcpyAsync(sorted.data(), spos.p, sizeof(double) * 3 * npts, hipMemcpyDeviceToHost, s), "D2H");
        check(hipStreamSynchronize(s), "tree sort");
    }

    void finish(host::Tree &t) override {
        sort(skey, nsall, t.src_order, t.src_sorted);
        sort(tkey, n, t.trg_order, t.trg_sorted);
    }
};

}  // namespace

void stable_order(const int *key, int n, int end_bit, int *order, hipStream_t s) {
    if (n <= 0) return;
    DevBuf kout, vin, tmp;
    kout.alloc(sizeof(int) * n);
    vin.alloc(sizeof(int) * n);
    hipLaunchKernelGGL(k_iota, dim3(blocks(n)), dim3(TB), 0, s, vin.as<int>(), n);
    check(hipGetLastError(), "k_iota");
    size_t bytes = 0;
    check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key, kout.as<int>(), vin.as<int>(), order, n, 0, end_bit, s),
          "radix sort size");
    tmp.alloc(std::max(bytes, (size_t)1));
    check(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, key, kout.as<int>(), vin.as<int>(), order, n, 0, end_bit, s),
          "radix sort");
    check(hipStreamSynchronize(s), "stable_order");
}

bool build_tree_gpu(const double *cell, const double *pos, int n, int nz, int n_per_leaf, hipStream_t s,
                    host::Geometry &g, host::Tree &t) {
    g = host::make_frame(cell, pos, n, nz);
    Bounds b{};
    const std::vector<std::array<double, 3>> shifts = host::image_shifts(g, b.lo, b.hi);
    const int ns = (int)shifts.size(), nblk = (int)blocks(n);
    GpuPoints pts;
    pts.s = s;
    pts.n = n;
    DevBuf r, sh, cnt, off, isrc;
    r.upload(g.r_src);
    std::vector<double> shv(3 * (size_t)std::max(ns, 1));
    for (int i = 0; i < ns; ++i)
        for (int d = 0; d < 3; ++d) shv[3 * (size_t)i + d] = shifts[i][d];
    sh.upload(shv);
    // image counts per (translation, block) -> offsets in enumeration order
    std::vector<int> c((size_t)ns * nblk), o((size_t)ns * nblk + 1, 0);
    if (ns > 0) {
        cnt.alloc(sizeof(int) * c.size());
        hipLaunchKernelGGL(k_image_count, dim3(nblk, ns), dim3(TB), 0, s, r.as<double>(), n, sh.as<double>(), b, cnt.as<int>());
        check(hipGetLastError(), "k_image_count");
        check(hipMemcpyAsync(c.data(), cnt.p, sizeof(int) * c.size(), hipMemcpyDeviceToHost, s), "D2H");
        check(hipStreamSynchronize(s), "image counts");
        for (size_t i = 0; i < c.size(); ++i) o[i + 1] = o[i] + c[i];
    }
    const long n_img = o.back();
    if ((long)n + n_img > 0x7fffffffL) throw std::runtime_error("too many images");
    pts.nsall = n + (int)n_img;
    pts.pos.alloc(sizeof(double) * 3 * (size_t)pts.nsall);
    check(hipMemcpyAsync(pts.pos.p, r.p, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToDevice, s), "D2D");
    g.img_src.resize(n_img);
    if (n_img > 0) {
        off.upload(o);
        isrc.alloc(sizeof(int) * n_img);
        hipLaunchKernelGGL(k_image_write, dim3(nblk, ns), dim3(TB), 0, s, r.as<double>(), n, sh.as<double>(), b,
                           off.as<int>(), pts.pos.as<double>(), isrc.as<int>());
        check(hipGetLastError(), "k_image_write");
        check(hipMemcpyAsync(g.img_src.data(), isrc.p, sizeof(int) * n_img, hipMemcpyDeviceToHost, s), "D2H");
    }
    pts.sbox.alloc(sizeof(int) * pts.nsall);
    pts.skey.alloc(sizeof(unsigned long long) * pts.nsall);
    pts.tbox.alloc(sizeof(int) * n);
    pts.tkey.alloc(sizeof(unsigned long long) * n);
    try {
        t = host::build_tree_with(g, n_per_leaf, pts);
    } catch (const TooDeep &) {
        return false;
    }
    return true;
}

}  // namespace pdmk::gpu
