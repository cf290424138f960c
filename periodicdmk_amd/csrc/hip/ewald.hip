// FP64 Ewald sum on the GPU: the computation behind pdmk_ewald_nufft
// (ewald_nufft, cpp/src/ewald_nufft.cpp:445-654; CellList, cpp/include/pdmk/cell_list.hpp).
//
//   phi_i = sum_{j, images n, 0 < r < r_cut} q_j erfc(alpha r)/r            (real space)
//         + sum_{k != 0, |k| <= k_max} (4 pi/V) e^{-k^2/4 alpha^2}/k^2 Re(S(k) e^{i k.x_i})
//         - 2 alpha/sqrt(pi) q_i,
//   r_cut = s/alpha, k_max = 2 s alpha, S(k) = sum_j q_j e^{-i k.x_j},
//   energy = 1/2 sum_i q_i phi_i - pi Q^2/(2 V alpha^2).
//
// The sum does not depend on alpha at a given s except, for a system with net charge Q, by
// the uniform shift pi Q/(V alpha^2) that the omitted k = 0 term leaves in the real-space
// part.  The GPU picks the alpha that balances its real-space and reciprocal work and then
// adds pi Q/V (1/alpha_ref^2 - 1/alpha^2), alpha_ref being the reference's choose_alpha
// (cpp/src/ewald_nufft.cpp:40-56), so potentials and energy are the reference's for any Q.
// The reciprocal part is the exact truncated sum (direct structure factors, FP64), which the
// reference approximates with a NUFFT at tolerance 1e-12.  Positions are wrapped into the
// cell first (the periodic sum is invariant under that).
#include "ewald.hpp"
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace pdmk::gpu {

namespace {

constexpr int EW_THREADS = 256;

// Real-space sum for the targets of one bin (workgroup (bin, chunk of EW_THREADS targets)).
// Neighbour list of bin b: entries nb_ptr[b] .. nb_ptr[b+1] of (source bin, image shift).
// Sources of a neighbour bin are read with wave-uniform addresses.
__global__ __launch_bounds__(EW_THREADS) void k_ewald_real(const double4 *__restrict__ pts, const int *__restrict__ bin_start,
                                                           const int *__restrict__ nb_ptr, const int *__restrict__ nb_bin,
                                                           const double4 *__restrict__ nb_shift, double alpha, double rc2,
                                                           double *__restrict__ out) {
    const int b = blockIdx.x;
    const int t = bin_start[b] + blockIdx.y * EW_THREADS + (int)threadIdx.x;
    const int tend = bin_start[b + 1];
    if (bin_start[b] + (int)blockIdx.y * EW_THREADS >= tend) return;  // whole workgroup
    const bool active = t < tend;
    const double4 me = active ? pts[t] : make_double4(0, 0, 0, 0);
    double acc = 0;
    for (int e = nb_ptr[b]; e < nb_ptr[b + 1]; ++e) {
        const int sb = nb_bin[e];
        const double4 sh = nb_shift[e];
        const double tx = me.x - sh.x, ty = me.y - sh.y, tz = me.z - sh.z;  // = r_t - shift
        for (int j = bin_start[sb]; j < bin_start[sb + 1]; ++j) {
            const double4 q = pts[j];
            const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (active && r2 < rc2 && r2 > 1e-30) {
                const double r = sqrt(r2);
                acc += q.w * erfc(alpha * r) / r;
            }
        }
    }
    if (active) out[t] = acc;
}

// Real-space sum at one target per workgroup (ewald_nufft_targets): the workgroup's threads
// stride over the sources of the target bin's neighbour entries; the partial sums are added
// in a fixed order (lanes by xor shuffles, then the wavefronts in index order).
__global__ __launch_bounds__(EW_THREADS) void k_ewald_real_targets(
    const double4 *__restrict__ pts, const int *__restrict__ bin_start, const int *__restrict__ nb_ptr,
    const int *__restrict__ nb_bin, const double4 *__restrict__ nb_shift, const int *__restrict__ trg_sorted,
    const int *__restrict__ trg_bin, double alpha, double rc2, double *__restrict__ out) {
    __shared__ double part[EW_THREADS / 64];
    const int t = blockIdx.x, b = trg_bin[t];
    const double4 me = pts[trg_sorted[t]];
    double acc = 0;
    for (int e = nb_ptr[b]; e < nb_ptr[b + 1]; ++e) {
        const int sb = nb_bin[e];
        const double4 sh = nb_shift[e];
        const double tx = me.x - sh.x, ty = me.y - sh.y, tz = me.z - sh.z;
        for (int j = bin_start[sb] + (int)threadIdx.x; j < bin_start[sb + 1]; j += EW_THREADS) {
            const double4 q = pts[j];
            const double dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (r2 < rc2 && r2 > 1e-30) {
                const double r = sqrt(r2);
                acc += q.w * erfc(alpha * r) / r;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0;
        for (int w = 0; w < EW_THREADS / 64; ++w) v += part[w];
        out[t] = v;
    }
}

// phi[u] = real[sorted(u)] + far[u] - 2 alpha/sqrt(pi) q_u + shift
__global__ void k_ewald_finalize(const double *__restrict__ real_sorted, const int *__restrict__ sorted_of,
                                 const double *__restrict__ far, const double *__restrict__ q, double self, double shift,
                                 double *__restrict__ phi, int n) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    phi[u] = real_sorted[sorted_of[u]] + far[u] - self * q[u] + shift;
}

struct Lat {
    double a[3][3];    // a[j] = lattice vector j
    double inv[3][3];  // fractional f = inv x: f_d = sum_k inv[d][k] x_k
    double vol, h[3];  // volume and heights
};

Lat make_lat(const double *cell) {
    Lat L;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) L.a[j][i] = cell[i + 3 * j];
    auto cross = [](const double *u, const double *v, double *w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    double c[3][3];
    for (int d = 0; d < 3; ++d) cross(L.a[(d + 1) % 3], L.a[(d + 2) % 3], c[d]);
    const double det = L.a[0][0] * c[0][0] + L.a[0][1] * c[0][1] + L.a[0][2] * c[0][2];
    if (std::fabs(det) < 1e-30) throw std::invalid_argument("cell matrix is singular");
    L.vol = std::fabs(det);
    for (int d = 0; d < 3; ++d) {
        for (int k = 0; k < 3; ++k) L.inv[d][k] = c[d][k] / det;  // rows of A^{-1} = (a_{d+1} x a_{d+2}) / det
        L.h[d] = L.vol / std::sqrt(c[d][0] * c[d][0] + c[d][1] * c[d][1] + c[d][2] * c[d][2]);
    }
    return L;
}

}  // namespace

double ewald_reference_alpha(int n, double vol, double s) {
    const double balanced = std::pow(n * M_PI * M_PI * M_PI / (vol * vol), 1.0 / 6.0);
    double target_nbr = 5000.0;
    if (const char *e = std::getenv("DMK_EWALD_TARGET_NBR")) target_nbr = std::atof(e);
    const double density = s * std::cbrt(4.0 * M_PI * (n / vol) / (3.0 * target_nbr));
    return std::max(balanced, density);
}

double ewald_gpu_alpha(int n, double vol) {
    // real-space work ~ n rho (s/alpha)^3, reciprocal work ~ n V (s alpha)^3: balanced at
    // alpha^6 = c n / V^2, with c ~ 500 from the FP64 operation counts per pair and per mode
    return std::pow(500.0 * n / (vol * vol), 1.0 / 6.0);
}

namespace {

// Positions wrapped into the cell: fractional coordinates in [0, 1) and theta = 2 pi frac.
void wrap_points(const Lat &L, int n, const double *pos, std::vector<double> &frac, std::vector<double> &theta) {
    frac.resize(3 * (size_t)n);
    theta.resize(3 * (size_t)n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double f = L.inv[d][0] * pos[3 * (size_t)i] + L.inv[d][1] * pos[3 * (size_t)i + 1] +
                       L.inv[d][2] * pos[3 * (size_t)i + 2];
            f -= std::floor(f);
            if (f >= 1.0) f = 0.0;
            frac[3 * (size_t)i + d] = f;
            theta[3 * (size_t)i + d] = 2 * M_PI * f;
        }
}

// Real-space bins of side >= r_cut along each height (nimg images per axis when one bin
// spans the cell): points counting-sorted by bin, and per bin the (source bin, image shift)
// entries with offsets -nimg..nimg per axis, wrapped.
struct Bins {
    int nb[3], nimg[3], nbins = 0, max_occ = 0;
    std::vector<int> bin_of, bstart, sorted_of, nptr, nbin;
    std::vector<double4> pts, nshift;
};

void bin_shape(const Lat &L, double rcut, int nb[3], int nimg[3]) {
    for (int d = 0; d < 3; ++d) {
        nb[d] = std::max(1, (int)std::floor(L.h[d] / rcut));
        nimg[d] = nb[d] == 1 ? (int)std::ceil(rcut / L.h[d]) : 1;
    }
}

int bin_index(const int nb[3], const double *f) {
    int c[3];
    for (int d = 0; d < 3; ++d) c[d] = std::min(nb[d] - 1, (int)(f[d] * nb[d]));
    return (c[2] * nb[1] + c[1]) * nb[0] + c[0];
}

// The entries of bin b: calls visit(source bin, shift d0, d1, d2) in a fixed order.
template <class F>
void for_each_neighbour(const Lat &L, const int nb[3], const int nimg[3], int b, F &&visit) {
    const int c[3] = {b % nb[0], (b / nb[0]) % nb[1], b / (nb[0] * nb[1])};
    for (int oz = -nimg[2]; oz <= nimg[2]; ++oz)
        for (int oy = -nimg[1]; oy <= nimg[1]; ++oy)
            for (int ox = -nimg[0]; ox <= nimg[0]; ++ox) {
                const int o[3] = {ox, oy, oz};
                int cc[3], sh[3];
                for (int d = 0; d < 3; ++d) {
                    cc[d] = c[d] + o[d];
                    sh[d] = 0;
                    while (cc[d] < 0) { cc[d] += nb[d]; --sh[d]; }
                    while (cc[d] >= nb[d]) { cc[d] -= nb[d]; ++sh[d]; }
                }
                double v[3];
                for (int d = 0; d < 3; ++d) v[d] = sh[0] * L.a[0][d] + sh[1] * L.a[1][d] + sh[2] * L.a[2][d];
                visit((cc[2] * nb[1] + cc[1]) * nb[0] + cc[0], v);
            }
}

Bins make_bins(const Lat &L, const std::vector<double> &frac, const double *q, int n, double rcut) {
    Bins B;
    bin_shape(L, rcut, B.nb, B.nimg);
    B.nbins = B.nb[0] * B.nb[1] * B.nb[2];
    B.bin_of.resize(n);
    B.bstart.assign(B.nbins + 1, 0);
    B.sorted_of.resize(n);
    for (int i = 0; i < n; ++i) {
        B.bin_of[i] = bin_index(B.nb, &frac[3 * (size_t)i]);
        B.bstart[B.bin_of[i] + 1]++;
    }
    for (int b = 0; b < B.nbins; ++b) B.bstart[b + 1] += B.bstart[b];
    B.pts.resize(n);
    std::vector<int> fill(B.bstart.begin(), B.bstart.end() - 1);
    for (int i = 0; i < n; ++i) {
        const int k = fill[B.bin_of[i]]++;
        B.sorted_of[i] = k;
        double x[3];
        for (int dd = 0; dd < 3; ++dd)
            x[dd] = L.a[0][dd] * frac[3 * (size_t)i] + L.a[1][dd] * frac[3 * (size_t)i + 1] +
                    L.a[2][dd] * frac[3 * (size_t)i + 2];
        B.pts[k] = make_double4(x[0], x[1], x[2], q[i]);
    }
    B.nptr.push_back(0);
    for (int b = 0; b < B.nbins; ++b) {
        B.max_occ = std::max(B.max_occ, B.bstart[b + 1] - B.bstart[b]);
        for_each_neighbour(L, B.nb, B.nimg, b, [&](int sb, const double *v) {
            if (B.bstart[sb + 1] == B.bstart[sb]) return;
            B.nbin.push_back(sb);
            B.nshift.push_back(make_double4(v[0], v[1], v[2], 0));
        });
        B.nptr.push_back((int)B.nbin.size());
    }
    return B;
}

// Reciprocal modes k = 2 pi A^{-T} m, 0 < |k| <= k_max, in the half space (k > 0, or k = 0
// and j > 0, or k = j = 0 and i > 0) with weight 2 (4 pi/V) e^{-k^2/4 alpha^2}/k^2.
struct Modes {
    int mmax[3];
    std::vector<int> mijk;
    std::vector<double> mw;
};

Modes make_modes(const Lat &L, double alpha, double kmax) {
    Modes M;
    for (int d = 0; d < 3; ++d) M.mmax[d] = (int)std::ceil(kmax * L.h[d] / (2 * M_PI));
    const double pref = 4 * M_PI / L.vol, inv4a2 = 1 / (4 * alpha * alpha), kmax2 = kmax * kmax;
    for (int k = 0; k <= M.mmax[2]; ++k)
        for (int j = (k == 0 ? 0 : -M.mmax[1]); j <= M.mmax[1]; ++j)
            for (int i = ((k == 0 && j == 0) ? 1 : -M.mmax[0]); i <= M.mmax[0]; ++i) {
                double kv[3];
                for (int d = 0; d < 3; ++d) kv[d] = 2 * M_PI * (i * L.inv[0][d] + j * L.inv[1][d] + k * L.inv[2][d]);
                const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                if (k2 > kmax2) continue;
                M.mijk.push_back(i);
                M.mijk.push_back(j);
                M.mijk.push_back(k);
                M.mw.push_back(2 * pref * std::exp(-k2 * inv4a2) / k2);
            }
    return M;
}

struct StreamGuard {
    hipStream_t s = nullptr;
    StreamGuard() { check(hipStreamCreate(&s), "hipStreamCreate"); }
    ~StreamGuard() {
        if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
    }
};

}  // namespace

namespace {

// Device arrays of one Ewald call: owned buffers, or slices of an arena.
struct Store {
    Arena *arena = nullptr;
    std::vector<DevBuf> owned;
    template <class T> T *put(const std::vector<T> &v) {
        if (arena) return arena->put(v);
        owned.emplace_back();
        owned.back().upload(v);
        return owned.back().template as<T>();
    }
    template <class T> T *make(size_t n) {
        if (arena) return arena->take_n<T>(n);
        owned.emplace_back();
        owned.back().alloc(sizeof(T) * n);
        return owned.back().template as<T>();
    }
};

struct EwaldDev {
    double4 *pts = nullptr, *nshift = nullptr;
    int *bstart = nullptr, *nptr = nullptr, *nbin = nullptr, *sorted = nullptr;
    double *q = nullptr, *real = nullptr, *far = nullptr, *phi = nullptr;
};

EwaldDev stage_ewald(Store &S, FourierSum &fs, const Bins &B, const Modes &M, const std::vector<double> &theta,
                     const double *q, int n, int max_blocks) {
    EwaldDev d;
    d.pts = S.put(B.pts);
    d.bstart = S.put(B.bstart);
    d.nptr = S.put(B.nptr);
    d.nbin = S.put(B.nbin);
    d.nshift = S.put(B.nshift);
    d.sorted = S.put(B.sorted_of);
    d.q = S.put(std::vector<double>(q, q + n));
    d.real = S.make<double>(n);
    d.far = S.make<double>(n);
    d.phi = S.make<double>(n);
    fs.setup(theta, n, 2 * M.mmax[0] + 1, 2 * M.mmax[1] + 1, 2 * M.mmax[2] + 1, M.mijk, M.mw, S.arena, max_blocks);
    return d;
}

// The Ewald sum with its device arrays owned (arena == nullptr) or carved from `arena`.
void ewald_sum_in(const double *cell, int n, const double *pos, const double *q, double s, double *pot,
                  double *energy, Arena *arena) {
    if (n <= 0) throw std::invalid_argument("N must be positive");
    if (!(s > 0)) throw std::invalid_argument("s must be positive");
    const Lat L = make_lat(cell);
    const double alpha_ref = ewald_reference_alpha(n, L.vol, s);
    const double alpha = ewald_gpu_alpha(n, L.vol);
    const double rcut = s / alpha, kmax = 2 * s * alpha;
    std::vector<double> frac, theta;
    wrap_points(L, n, pos, frac, theta);
    const Bins B = make_bins(L, frac, q, n, rcut);
    const Modes M = make_modes(L, alpha, kmax);

    int max_blocks = 2048;
    if (arena) {
        // size the request before touching the memory: everything but the structure factor
        // partial sums, then as many partial-sum blocks (point tiles) as the rest holds
        Arena dry;
        Store ds{&dry, {}};
        FourierSum f0;
        stage_ewald(ds, f0, B, M, theta, q, n, 1);
        const size_t per_block = sizeof(double2) * (size_t)std::max<size_t>(1, M.mw.size());
        const size_t fixed = dry.used - per_block + 256;
        if (fixed > arena->cap || (arena->cap - fixed) / per_block < 1)
            throw std::length_error("Ewald sum of N=" + std::to_string(n) + " points and " + std::to_string(M.mw.size()) +
                                    " modes needs " + std::to_string(fixed + per_block) +
                                    " bytes of device scratch, more than the " + std::to_string(arena->cap) +
                                    " available");
        max_blocks = (int)std::min<size_t>(2048, (arena->cap - fixed) / per_block);
        arena->used = 0;
    }
    Store S{arena, {}};
    FourierSum fs;
    const EwaldDev d = stage_ewald(S, fs, B, M, theta, q, n, max_blocks);

    // the arena path uses the null stream (nothing to create)
    std::unique_ptr<StreamGuard> guard;
    if (!arena) guard = std::make_unique<StreamGuard>();
    const hipStream_t st = guard ? guard->s : nullptr;
    fs.run(d.q, d.far, st);
    const dim3 grid(B.nbins, (B.max_occ + EW_THREADS - 1) / EW_THREADS);
    hipLaunchKernelGGL(k_ewald_real, grid, dim3(EW_THREADS), 0, st, d.pts, d.bstart, d.nptr, d.nbin, d.nshift, alpha,
                       rcut * rcut, d.real);
    check(hipGetLastError(), "k_ewald_real");
    double Q = 0;
    for (int i = 0; i < n; ++i) Q += q[i];
    const double shift = M_PI * Q / L.vol * (1 / (alpha_ref * alpha_ref) - 1 / (alpha * alpha));
    hipLaunchKernelGGL(k_ewald_finalize, dim3((n + 255) / 256), dim3(256), 0, st, d.real, d.sorted, d.far, d.q,
                       2 * alpha / std::sqrt(M_PI), shift, d.phi, n);
    check(hipGetLastError(), "k_ewald_finalize");
    check(hipMemcpyAsync(pot, d.phi, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (energy) {
        double e = 0;
        for (int i = 0; i < n; ++i) e += q[i] * pot[i];
        *energy = 0.5 * e - M_PI * Q * Q / (2 * L.vol * alpha_ref * alpha_ref);
    }
}

// Device scratch of the allocation-free Ewald entry point: one copy per GPU, held until the
// process ends.  Host code uses the per-device address hipGetSymbolAddress returns, never
// the array's name (a host-side stand-in).
constexpr size_t kEwaldScratchBytes = size_t(128) << 20;
__device__ __align__(256) char ewald_scratch[kEwaldScratchBytes];

char *ewald_scratch_ptr() {
    static std::mutex mu;
    static std::vector<char *> ptr;  // per device
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptr.size() <= dev) ptr.resize(dev + 1, nullptr);
    if (!ptr[dev]) {
        void *p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(ewald_scratch)), "hipGetSymbolAddress(ewald_scratch)");
        ptr[dev] = static_cast<char *>(p);
    }
    return ptr[dev];
}

}  // namespace

void ewald_sum(const double *cell, int n, const double *pos, const double *q, double s, double *pot, double *energy) {
    ewald_sum_in(cell, n, pos, q, s, pot, energy, nullptr);
}

size_t ewald_scratch_bytes() { return kEwaldScratchBytes; }

void ewald_sum_noalloc(const double *cell, int n, const double *pos, const double *q, double s, double *pot,
                       double *energy) {
    Arena a;
    a.base = ewald_scratch_ptr();
    a.cap = kEwaldScratchBytes;
    ewald_sum_in(cell, n, pos, q, s, pot, energy, &a);
}

double ewald_targets_alpha(const double *cell, int n, const double *pos, int m, const int *targets, double s) {
    const Lat L = make_lat(cell);
    std::vector<double> frac, theta;
    wrap_points(L, n, pos, frac, theta);
    double best = 0, best_cost = 0;
    const double hmin = std::min(L.h[0], std::min(L.h[1], L.h[2]));
    std::vector<long long> cnt;
    for (int div = 1; div <= 32; ++div) {
        const double rcut = hmin / div, alpha = s / rcut, kmax = 2 * s * alpha;
        int nb[3], nimg[3];
        bin_shape(L, rcut, nb, nimg);
        const long long nbins = (long long)nb[0] * nb[1] * nb[2];
        if (nbins > 4 * (long long)n + 64) break;
        cnt.assign(nbins, 0);
        for (int i = 0; i < n; ++i) cnt[bin_index(nb, &frac[3 * (size_t)i])]++;
        // points the target workgroups visit, and (point, mode) products of the structure
        // factors and the evaluation (half-space mode count from the volume of the k-ball)
        double visits = 0;
        for (int t = 0; t < m; ++t)
            for_each_neighbour(L, nb, nimg, bin_index(nb, &frac[3 * (size_t)targets[t]]),
                               [&](int sb, const double *) { visits += (double)cnt[sb]; });
        const double modes = 0.5 * (4.0 / 3.0) * M_PI * kmax * kmax * kmax * L.vol / std::pow(2 * M_PI, 3);
        const double cost = visits + 2.0 * modes * ((double)n + m);
        if (best == 0 || cost < best_cost) {
            best = alpha;
            best_cost = cost;
        }
    }
    return best;
}

void ewald_sum_targets(const double *cell, int n, const double *pos, const double *q, int m, const int *targets,
                       double s, double *pot) {
    if (n <= 0) throw std::invalid_argument("N must be positive");
    if (m < 0) throw std::invalid_argument("M must be non-negative");
    if (!(s > 0)) throw std::invalid_argument("s must be positive");
    for (int t = 0; t < m; ++t)
        if (targets[t] < 0 || targets[t] >= n) throw std::invalid_argument("target index out of range");
    if (m == 0) return;
    const Lat L = make_lat(cell);
    const double alpha_ref = ewald_reference_alpha(n, L.vol, s);
    const double alpha = ewald_targets_alpha(cell, n, pos, m, targets, s);
    const double rcut = s / alpha, kmax = 2 * s * alpha;
    std::vector<double> frac, theta;
    wrap_points(L, n, pos, frac, theta);
    const Bins B = make_bins(L, frac, q, n, rcut);
    const Modes M = make_modes(L, alpha, kmax);
    std::vector<int> t_sorted(m), t_bin(m);
    std::vector<double> t_theta(3 * (size_t)m);
    for (int t = 0; t < m; ++t) {
        t_sorted[t] = B.sorted_of[targets[t]];
        t_bin[t] = B.bin_of[targets[t]];
        for (int d = 0; d < 3; ++d) t_theta[3 * (size_t)t + d] = theta[3 * (size_t)targets[t] + d];
    }

    StreamGuard guard;
    const hipStream_t st = guard.s;
    DevBuf d_pts, d_bstart, d_nptr, d_nbin, d_nshift, d_tsorted, d_tbin, d_ttheta, d_q, d_real, d_far;
    d_pts.upload(B.pts);
    d_bstart.upload(B.bstart);
    d_nptr.upload(B.nptr);
    d_nbin.upload(B.nbin);
    d_nshift.upload(B.nshift);
    d_tsorted.upload(t_sorted);
    d_tbin.upload(t_bin);
    d_ttheta.upload(t_theta);
    d_q.upload(std::vector<double>(q, q + n));
    d_real.alloc(sizeof(double) * (size_t)m);
    d_far.alloc(sizeof(double) * (size_t)m);
    FourierSum fs;
    fs.setup(theta, n, 2 * M.mmax[0] + 1, 2 * M.mmax[1] + 1, 2 * M.mmax[2] + 1, M.mijk, M.mw);
    fs.run_at(d_q.as<double>(), d_ttheta.as<double>(), m, d_far.as<double>(), st);
    hipLaunchKernelGGL(k_ewald_real_targets, dim3(m), dim3(EW_THREADS), 0, st, d_pts.as<double4>(), d_bstart.as<int>(),
                       d_nptr.as<int>(), d_nbin.as<int>(), d_nshift.as<double4>(), d_tsorted.as<int>(),
                       d_tbin.as<int>(), alpha, rcut * rcut, d_real.as<double>());
    check(hipGetLastError(), "k_ewald_real_targets");
    std::vector<double> real(m), far(m);
    check(hipMemcpyAsync(real.data(), d_real.p, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, st), "D2H");
    check(hipMemcpyAsync(far.data(), d_far.p, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, st), "D2H");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    double Q = 0;
    for (int i = 0; i < n; ++i) Q += q[i];
    const double shift = M_PI * Q / L.vol * (1 / (alpha_ref * alpha_ref) - 1 / (alpha * alpha));
    const double self = 2 * alpha / std::sqrt(M_PI);
    for (int t = 0; t < m; ++t) pot[t] = real[t] + far[t] - self * q[targets[t]] + shift;
}

}  // namespace pdmk::gpu
