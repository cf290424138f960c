// gfx950 kernels of the periodic DMK evaluate() path.  Included by plan.hip only.
//
// Stage map (reference functions each kernel replaces):
//   k_gather_charges ... update_charges + image-charge gather
//                        (cpp/src/periodic_dmk.cpp:164-170, cpp/src/dmk/pbc_tree.cpp:1807-1821)
//   k_charge2proxy ..... proxy::charge2proxycharge_3d (cpp/src/dmk/proxy.cpp:153-212)
//   k_tensor3 .......... tensorprod::transform_3d c2p / p2c (cpp/src/dmk/tensorprod.cpp:39-75)
//   k_proxy2pw ......... proxy::proxycharge2pw_3d (cpp/src/dmk/proxy.cpp:42-90)
//   k_pw_shift ......... form_eval_expansions shifts with fused radialft
//                        (cpp/src/dmk/pbc_tree.cpp:1294-1369, cpp/include/dmk/tree.hpp:287-316)
//   k_pw2proxy ......... planewave_to_proxy_potential (cpp/src/dmk/planewave.cpp:49-105)
//   k_eval_targets ..... proxy::eval_targets_3d (cpp/src/dmk/proxy.cpp:307-448)
//   k_near_field ....... evaluate_direct_interactions + laplace_3d_poly_all_pairs
//                        (cpp/src/dmk/pbc_tree.cpp:1442-1681, cpp/include/dmk/vector_kernels.hpp:354-421)
//   k_ff_* ............. far_field_nufft, 3-D non-QP (cpp/src/dmk/pbc.cpp:378-541)
//   k_finalize ......... apply_target_self_correction + desort_potentials + far-field add
//                        (cpp/src/dmk/pbc_tree.cpp:1759-1805, cpp/src/periodic_dmk.cpp:172-182)
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "dpp_fma.hpp"

namespace pdmk::gpu {

template <class R>
struct alignas(4 * sizeof(R)) P4 { R x, y, z, w; };

template <class R>
struct alignas(2 * sizeof(R)) Cx { R re, im; };

// Workgroups are dispatched to the 8 XCDs round-robin; this bijection of [0, nb) gives each
// XCD one contiguous range of workgroup indices, so work items that share inputs (items are
// laid out in tree order, i.e. spatially) run under the same L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    constexpr int X = 8;
    const int x = b % X, i = b / X, q = nb / X, r = nb % X;
    return x * q + min(x, r) + i;
}

// Kernels of the proxy / plane-wave chain raise their waves' issue priority: evaluate()
// runs them beside the near-field kernel, and the chain is the longer path.
__device__ __forceinline__ void chain_priority() { __builtin_amdgcn_s_setprio(2); }

// A zero the compiler cannot see through: adding it moves a wave-uniform value into a VGPR.
__device__ __forceinline__ int vgpr_zero() {
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

// 1 / sqrt(x) for the near-field distances: every caller passes 1e-30 < x < 1e30 (a positive
// normal number).  FP64: the hardware estimate y = v_rsq_f64(x) and one third-order
// correction y (1 + e/2 + 3e^2/8), e = 1 - x y^2 -- the library rsqrt's own arithmetic, bit
// for bit on this range, without its branches for zero, infinity and NaN inputs.
template <class R> __device__ __forceinline__ R rsqrt_full(R x);
template <> __device__ __forceinline__ float rsqrt_full<float>(float x) { return __frsqrt_rn(x); }
template <> __device__ __forceinline__ double rsqrt_full<double>(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(y * -x, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// ------------------------------------------------------------------ O(N) stages
// Charges of the sorted sources.  Quasi-periodic passes multiply each by the real (comp 0)
// or imaginary (comp 1) part of its image phase e^{i b.R} (originals: phase 1).
template <class R>
__global__ void k_gather_charges(const double *__restrict__ q, const int *__restrict__ src_user, P4<R> *__restrict__ src,
                                 int nsall, R *__restrict__ pot_sorted, R *__restrict__ pot_near, int n,
                                 const double2 *__restrict__ phase, int comp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsall) {
        double v = q[src_user[i]];
        if (phase) v *= comp == 0 ? phase[i].x : phase[i].y;
        src[i].w = (R)v;
    }
    if (i < n) {
        pot_sorted[i] = R(0);
        pot_near[i] = R(0);
    }
}

// pot[stride u] = pot_sorted[i] - w0_i q_u + far[u], u = trg_user[i] (stride 2: one component
// of an interleaved complex output).  Vector v = blockIdx.z of a many-vector launch uses the
// arrays offset by v times the vs_* element strides (all zero for one vector).
template <class R>
__global__ void k_finalize(const R *__restrict__ pot_sorted, const R *__restrict__ pot_near, const double *__restrict__ w0,
                           const int *__restrict__ trg_user, const double *__restrict__ q, const double *__restrict__ far,
                           const int *__restrict__ far_idx, R *__restrict__ pot, int n, int use_self, int use_far,
                           int stride, long long vs_sorted = 0, long long vs_near = 0, long long vs_q = 0,
                           long long vs_far = 0, long long vs_pot = 0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long vz = blockIdx.z;
    pot_sorted += vz * vs_sorted;
    pot_near += vz * vs_near;
    q += vz * vs_q;
    far += vz * vs_far;
    pot += vz * vs_pot;
    const int u = trg_user[i];
    double v = (double)pot_sorted[i] + (double)pot_near[i];
    if (use_self) v -= w0[i] * q[u];
    if (use_far) v += far[far_idx[i]];
    pot[(size_t)stride * u] = (R)v;
}

// ------------------------------------------------------------------ upward pass
// One workgroup per destination box:
//   C[c][b][a] = sum_s T_a(x_s) T_b(y_s) T_c(z_s) q_s,  x_s = sc (r_s - centre).
// Sources hold coordinates relative to their leaf centre; each range lies in one leaf and
// carries rng_off = leaf centre - box centre, so x_s = sc (local_s + off).  Per tile of
// C2P_TS sources the Chebyshev values are tabulated in LDS (q folded into T_c); thread
// (c, bb) then owns the RB x P output block C[c][bb RB .. bb RB + RB - 1][0 .. P-1] in
// registers, so each source costs RB P FMAs per thread against ceil(P/4) + 2 LDS reads.
constexpr int C2P_TS = 64;

// register block RB (b values per thread) and workgroup size per order; output blocks beyond
// one workgroup's threads are processed in further passes over the sources
template <int P> struct C2pCfg {
    // (RB 2 at P = 18: 162 columns, 3 wavefronts per box instead of 2 with RB 3; the 6-digit chain
    // alone at N=1e6 1.21 -> 1.12 ms in FP32, 1.87 -> 1.62 ms in FP64, the sums unchanged)
    static constexpr int RB = P >= 18 ? 2 : 3;
    static constexpr int NB = (P + RB - 1) / RB;
    static constexpr int cols = NB * P;
    static constexpr int threads = P >= 38 ? 256 : (cols + 63) / 64 * 64;
};

template <class R, int P>
__global__ __launch_bounds__(C2pCfg<P>::threads) void k_charge2proxy(
    const P4<R> *__restrict__ src, const int *__restrict__ grp_slot, const R *__restrict__ grp_geom,
    const int *__restrict__ grp_rng, const int *__restrict__ rng_begin, const int *__restrict__ rng_count,
    const R *__restrict__ rng_off, R *__restrict__ proxy, const R *__restrict__ qv = nullptr, int qv_ld = 0,
    long long vs_out = 0) {
    chain_priority();
    constexpr int RB = C2pCfg<P>::RB, NB = C2pCfg<P>::NB, NT = C2pCfg<P>::threads;
    // many-vector launch: vector blockIdx.z takes its charges from qv[source * qv_ld + v]
    // (src[].w otherwise) and writes its own proxy array, vs_out elements further on
    proxy += (long long)blockIdx.z * vs_out;
    constexpr int PX = (P + 3) & ~3, PY = NB * RB;  // padded table rows
    __shared__ R Tx[C2P_TS * PX], Ty[C2P_TS * PY], Tz[C2P_TS * P];
    const int g = blockIdx.x;
    const R sc = grp_geom[4 * g + 3];
    for (int col0 = 0; col0 < NB * P; col0 += NT) {
        const int col = col0 + (int)threadIdx.x;
        const int c = col / NB, bb = col - c * NB;
        const bool active = col < NB * P;
        R acc[RB][P];
#pragma unroll
        for (int j = 0; j < RB; ++j)
#pragma unroll
            for (int a = 0; a < P; ++a) acc[j][a] = R(0);
        for (int r = grp_rng[g]; r < grp_rng[g + 1]; ++r) {
            const int beg = rng_begin[r], cnt = rng_count[r];
            for (int s0 = 0; s0 < cnt; s0 += C2P_TS) {
                const int ns = min(C2P_TS, cnt - s0);
                __syncthreads();
                for (int t = threadIdx.x; t < 3 * C2P_TS; t += NT) {
                    const int s = t % C2P_TS, d = t / C2P_TS;
                    R *T = d == 0 ? Tx + s * PX : (d == 1 ? Ty + s * PY : Tz + s * P);
                    const int len = d == 0 ? PX : (d == 1 ? PY : P);
                    if (s < ns) {
                        const P4<R> pt = src[beg + s0 + s];
                        const R x = sc * ((d == 0 ? pt.x : (d == 1 ? pt.y : pt.z)) + rng_off[4 * r + d]);
                        const R m = (d == 2) ? (qv ? qv[(size_t)(beg + s0 + s) * qv_ld + blockIdx.z] : pt.w) : R(1);
                        R t0 = m, t1 = x * m;
                        T[0] = t0;
                        if (P > 1) T[1] = t1;
                        for (int k = 2; k < P; ++k) {
                            const R t2 = R(2) * x * t1 - t0;
                            T[k] = t2;
                            t0 = t1;
                            t1 = t2;
                        }
                        for (int k = P; k < len; ++k) T[k] = R(0);
                    } else {
                        for (int k = 0; k < len; ++k) T[k] = R(0);
                    }
                }
                __syncthreads();
                if (active) {
                    for (int s = 0; s < ns; ++s) {
                        R tx[P];
#pragma unroll
                        for (int a = 0; a < P; ++a) tx[a] = Tx[s * PX + a];
                        const R z = Tz[s * P + c];
#pragma unroll
                        for (int j = 0; j < RB; ++j) {
                            const R w = Ty[s * PY + bb * RB + j] * z;
#pragma unroll
                            for (int a = 0; a < P; ++a) acc[j][a] = fma(tx[a], w, acc[j][a]);
                        }
                    }
                }
            }
        }
        if (active) {
            R *out = proxy + (size_t)grp_slot[g] * (P * P * P) + (size_t)P * P * c;
#pragma unroll
            for (int j = 0; j < RB; ++j) {
                const int b = bb * RB + j;
                if (b < P)
#pragma unroll
                    for (int a = 0; a < P; ++a) out[a + P * b] = acc[j][a];
            }
        }
    }
}

// The same proxy charges on the FP64 matrix cores, for p >= 28.  Per destination box
//   C[(c, b), a] = sum_s A[(c, b), s] B[s, a],  A = q_s T_c(z_s) T_b(y_s),  B = T_a(x_s),
// is a GEMM of P^2 rows by PA columns (a padded to 16s) over the box's sources, computed with
// v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds
// D[l / 16 + 4 r][l % 16] in accumulator element r (layout read off gfx950 hardware).  The
// row tiles of a box are split over RS workgroups; wave w of a workgroup owns its row tiles
// w, w + NW, ... and every column tile.  Two shapes at p = 28: for trees with many boxes
// (WIDE) two workgroups of 4 waves per box (7 row tiles per wave, 158 VGPRs: three
// workgroups share a CU, so one's table building overlaps the others' matrix products);
// otherwise one workgroup of 16 waves (4 row tiles per wave, 98 VGPRs).
typedef double c2p_d4 __attribute__((ext_vector_type(4)));
typedef float c2p_f4 __attribute__((ext_vector_type(4)));

// 16 x 16 x 4 matrix-core products.  Operands: lane l holds A[l%16][l/16] and B[l/16][l%16].
// Accumulator element e of lane l is D[row(l/16, e)][l%16]: rows l/16 + 4e for FP64 and
// 4(l/16) + e for FP32 (both read off the hardware with in-bounds probes).
template <class R> struct Mfma16;
template <> struct Mfma16<double> {
    typedef c2p_d4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kk, int e) { return kk + 4 * e; }
};
template <> struct Mfma16<float> {
    typedef c2p_f4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kk, int e) { return 4 * kk + e; }
};
template <int P, bool WIDE> struct C2pMfmaCfg {
    static constexpr int RT = (P * P + 15) / 16;                  // row tiles
    static constexpr int RS = P >= 38 || WIDE ? 2 : 1;            // workgroups per box
    static constexpr int RTS = (RT + RS - 1) / RS;                // row tiles per workgroup
    static constexpr int NW = WIDE ? 4 : (P >= 38 ? 8 : 16);      // waves per workgroup
    static constexpr int TPW = (RTS + NW - 1) / NW;   // row tiles per wave
    static constexpr int TS = 64;                     // sources per LDS table
    static constexpr int PA = (P + 15) / 16 * 16;     // a, padded
    static constexpr int CT = PA / 16;                // column tiles
    // LDS row strides of the tables (odd numbers of doubles: a thread writing its source's row
    // and the 4 sources of a fragment read fall on different banks; PA and P put them on one)
    static constexpr int SX = PA + 1, SY = P % 2 ? P : P + 1;
};

template <class R, int P, bool WIDE>
__global__ __launch_bounds__((64 * C2pMfmaCfg<P, WIDE>::NW)) void k_charge2proxy_mfma(
    const P4<R> *__restrict__ src, const int *__restrict__ grp_slot, const R *__restrict__ grp_geom,
    const int *__restrict__ grp_rng, const int *__restrict__ rng_begin, const int *__restrict__ rng_count,
    const R *__restrict__ rng_off, R *__restrict__ proxy) {
    chain_priority();
    using C = C2pMfmaCfg<P, WIDE>;
    using M = Mfma16<R>;
    constexpr int TS = C::TS, PA = C::PA, RT = C::RT, RS = C::RS, RTS = C::RTS, NW = C::NW, TPW = C::TPW, CT = C::CT,
                  SX = C::SX, SY = C::SY;
    static_assert(P <= PA, "order above the padded column count");
    __shared__ R Tx[TS * SX], Ty[TS * SY], Tz[TS * SY];
    const int g = blockIdx.x / RS, part = blockIdx.x - g * RS, lane = threadIdx.x & 63;
    const int rt0 = part * RTS, rt1 = min(RT, rt0 + RTS);  // this workgroup's row tiles
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int kk = lane >> 4, jj = lane & 15;
    const R sc = grp_geom[4 * g + 3];
    int ci[TPW], bi[TPW];
    R rv[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int row = (rt0 + w + NW * t) * 16 + jj;
        const bool ok = row < P * P;
        rv[t] = ok ? R(1) : R(0);
        ci[t] = ok ? row / P : 0;
        bi[t] = ok ? row % P : 0;
    }
    typename M::acc_t acc[TPW][CT];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[t][ct] = typename M::acc_t{R(0), R(0), R(0), R(0)};
    for (int r = grp_rng[g]; r < grp_rng[g + 1]; ++r) {
        const int beg = rng_begin[r], cnt = rng_count[r];
        for (int s0 = 0; s0 < cnt; s0 += TS) {
            const int ns = min(TS, cnt - s0);
            __syncthreads();
            for (int t = threadIdx.x; t < 3 * TS; t += blockDim.x) {
                const int s = t % TS, d = t / TS;
                R *T = d == 0 ? Tx + s * SX : (d == 1 ? Ty + s * SY : Tz + s * SY);
                const int len = d == 0 ? PA : P;
                if (s < ns) {
                    const P4<R> pt = src[beg + s0 + s];
                    const R x = sc * ((d == 0 ? pt.x : (d == 1 ? pt.y : pt.z)) + rng_off[4 * r + d]);
                    const R m = (d == 2) ? pt.w : R(1);
                    R t0 = m, t1 = x * m;
                    T[0] = t0;
                    T[1] = t1;
                    for (int k = 2; k < P; ++k) {
                        const R t2 = R(2) * x * t1 - t0;
                        T[k] = t2;
                        t0 = t1;
                        t1 = t2;
                    }
                    for (int k = P; k < len; ++k) T[k] = R(0);
                } else {
                    for (int k = 0; k < len; ++k) T[k] = R(0);
                }
            }
            __syncthreads();
            for (int k0 = 0; k0 < ns; k0 += 4) {
                const int s = k0 + kk;
                R bv[CT];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) bv[ct] = Tx[s * SX + 16 * ct + jj];
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    if (rt0 + w + NW * t < rt1) {
                        const R a = Tz[s * SY + ci[t]] * Ty[s * SY + bi[t]] * rv[t];
#pragma unroll
                        for (int ct = 0; ct < CT; ++ct) acc[t][ct] = M::mma(a, bv[ct], acc[t][ct]);
                    }
                }
            }
        }
    }
    R *out = proxy + (size_t)grp_slot[g] * (P * P * P);
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        if (rt0 + w + NW * t >= rt1) continue;
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int a = ct * 16 + jj;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = (rt0 + w + NW * t) * 16 + M::row(kk, e);
                if (row < P * P && a < P) out[(size_t)row * P + a] = acc[t][ct][e];
            }
        }
    }
}

// ------------------------------------------------------------------ tensor transforms
// All transforms are sequences of mode products of a small 3-D tensor with small dense
// matrices.  One workgroup owns one output box; a thread owns one column of the mode being
// contracted (the column sits in registers), the matrix entries are read with wave-uniform
// addresses (scalar loads), and intermediates live in LDS with row strides of an odd number
// of elements, so 64-bit LDS accesses by consecutive lanes hit distinct banks.  When the
// intermediates of a whole box do not fit the LDS budget the box is processed in slices of
// the first contraction's output index (upward, c2p/p2c) or of the first output index of the
// downward contraction; each slice then runs all three steps on its own.
constexpr int kTfLdsBytes = 64 * 1024;

template <int N> struct OddStride { static constexpr int v = (N % 2) ? N : N + 1; };

// LDS budget of a tensor-transform workgroup: 64 KiB, or all 160 KiB of a CU for FP64 at
// p >= 38 (12 digits), whose registers allow one such workgroup per CU anyway: 6 output planes
// per workgroup instead of 1, so each input is read 7 times instead of 38 (measured on the
// N=4e6 clustered workload: child -> parent 431 -> 128 ms, parent -> child 162 -> 73 ms per
// evaluate).  The proxy <-> plane-wave transforms keep 64 KiB: with 160 KiB they hold one
// wavefront per SIMD and ran 1.7x slower.
constexpr int tf_lds_budget(int p, int real_bytes) { return p >= 38 && real_bytes == 8 ? 160 * 1024 : kTfLdsBytes; }

// slices per pass so that `reserved` + `per_slice` bytes x slices stays within the LDS budget
constexpr int tf_slices(int total, int per_slice, int reserved = 0, int budget = kTfLdsBytes) {
    const int k = (budget - reserved) / per_slice;
    return k < 1 ? 1 : (k > total ? total : k);
}

// Slices of the proxy <-> plane-wave transforms that fit the LDS budget; each slice is a
// workgroup of its own (grid y), so a level with few boxes still fills the GPU.
// BK: the LDS budget in KiB per workgroup (0 = the default below); the autotuner picks
// among 0, 16, 32 and 64 per plan.
// slices per box from which the proxy <-> plane-wave transforms take their first contraction
// of whole boxes from HBM (two passes) rather than re-reading each box's input per slice
constexpr int kTwoPassSlices = 8;

template <class R, int P, int NPW, int BK = 0> struct P2pwCfg {
    static constexpr int H = (NPW + 1) / 2;
    // (FP64 at p = 28: one k3 plane per workgroup, 31 KB, five workgroups per CU; two planes at
    // 61 KB fit two: N=1e6 9-digit evaluate() 9.06 -> 8.95 ms.  At p = 18 a 32 KB budget, on
    // the plane-wave -> proxy side too: FP32 4 planes instead of 9, 6-digit evaluate()
    // 3.06 -> 2.95 ms; FP64 2 instead of 4, 4.72 -> 4.60 ms)
    static constexpr int K3 = tf_slices(H, (P * P + NPW * OddStride<P>::v) * (int)sizeof(Cx<R>), 0,
                                        BK ? BK * 1024 : (((P == 28 && sizeof(R) == 8) || P == 18) ? 32 * 1024 : kTfLdsBytes));
    static constexpr int slices = (H + K3 - 1) / K3;
    // the y and x steps have K3 P and K3 n columns of n outputs for 256 threads; the outputs
    // of a column are split over SGY / SGX wavefronts when that fills more of them
    static constexpr int NT = 256;
    static constexpr int cap(int v) { return v < 1 ? 1 : (v > 8 ? 8 : v); }
    static constexpr int SGY = cap(NT / ((K3 * P + 63) / 64 * 64)), SGX = cap(NT / ((K3 * NPW + 63) / 64 * 64));
    static constexpr int PSY = (NPW + SGY - 1) / SGY, PSX = (NPW + SGX - 1) / SGX;
};
template <class R, int P, int NPW, int BK = 0> struct Pw2pCfg {
    static constexpr int H = (NPW + 1) / 2;
    // (FP64 at p = 28: 2 slices of a per workgroup, 43 KB, so three workgroups share a CU;
    // 3 per workgroup at 65 KB fit two: N=1e6 9-digit evaluate() 9.12 -> 9.07 ms)
    static constexpr int AC = tf_slices(P, (H * NPW + H * OddStride<P>::v) * (int)sizeof(Cx<R>), 0,
                                        BK ? BK * 1024 : ((P == 28 && sizeof(R) == 8) ? 48 * 1024 : (P == 18 ? 32 * 1024 : kTfLdsBytes)));
    static constexpr int slices = (P + AC - 1) / AC;
    // the y and z steps have AC H and AC P columns of P outputs for 256 threads; the outputs of
    // a column are split over SGY / SGZ wavefronts when that fills more of them
    static constexpr int NT = 256;
    static constexpr int cap(int v) { return v < 1 ? 1 : (v > 8 ? 8 : v); }
    static constexpr int SGY = cap(NT / ((AC * H + 63) / 64 * 64)), SGZ = cap(NT / ((AC * P + 63) / 64 * 64));
    static constexpr int PSY = (P + SGY - 1) / SGY, PSZ = (P + SGZ - 1) / SGZ;
};

// 3-mode tensor transform out(a',b',c') (+)= sum_children U0(a',a) U1(b',b) U2(c',c) in(a,b,c)
// (tensor layout [c][b][a], a fastest).  U_d is the minus/plus parent<->child matrix selected
// by bit d of the child code, transposed for child -> parent.  Item i writes out slot
// item_out[i] from the CSR range item_ptr[i]..item_ptr[i+1] of (in slot, child code);
// item_add[i] = accumulate into the existing output.  Workgroup (item, slice) computes the
// output planes c' in [slice kc, slice kc + kc), kc <= T3Cfg::KC; the grid is 1-D, workgroup
// (item, slice) at xcd_block index item * slices + slice, so an item's slices run together
// under one L2 and its input is fetched from memory about once.
template <class R, int P> struct T3Cfg {
    static constexpr int PP = P * P, TS = OddStride<P>::v;
    // output planes per workgroup (the budget keeps 2 p^2 values spare, so FP64 at p = 28
    // fits three workgroups per CU)
    static constexpr int KC = tf_slices(P, (PP + P * TS) * (int)sizeof(R), 2 * PP * (int)sizeof(R),
                                        tf_lds_budget(P, (int)sizeof(R)));
    // at most 4 waves below p = 38: evaluate() runs these beside the near field, whose
    // retiring 4-wave workgroups free 4 wave slots at a time
    static constexpr int threads = PP <= 256 ? (PP + 63) / 64 * 64 : (P >= 38 ? 512 : 256);
    // the y and x steps have KC P columns of P outputs; when that leaves more than half the
    // threads idle, each column's outputs are split between SG wavefronts
    static constexpr int SG = KC * P * 2 <= threads ? 2 : 1;
    static constexpr int PS = (P + SG - 1) / SG;  // outputs per thread in the y and x steps
    static constexpr int slots(int nk) { return (nk * P + 63) / 64 * 64 * SG; }  // y / x work slots
};

// Work slot w of the y and x steps -> (column, segment): consecutive lanes take consecutive
// columns, and a wavefront's lanes share one segment (so its matrix rows are wave-uniform).
template <int SG> __device__ __forceinline__ void t3_slot(int w, int &col, int &sg) {
    col = (w & 63) + 64 * (w / (64 * SG));
    sg = __builtin_amdgcn_readfirstlane((w >> 6) % SG);  // wave-uniform (w = lane + multiple of 64)
}

template <class R, int P, bool TRANSPOSE>
__global__ __launch_bounds__((T3Cfg<R, P>::threads)) void k_tensor3(
    const R *__restrict__ in_base, R *__restrict__ out_base, const int *__restrict__ item_out,
    const int *__restrict__ item_add, const int *__restrict__ item_ptr, const int *__restrict__ in_slot,
    const int *__restrict__ in_code, const R *__restrict__ mats_rm, int n_items, int kc, long long vs_in = 0,
    long long vs_out = 0) {
    chain_priority();
    in_base += (long long)blockIdx.z * vs_in;  // vector blockIdx.z of a many-vector launch
    out_base += (long long)blockIdx.z * vs_out;
    using Cfg = T3Cfg<R, P>;
    constexpr int PP = P * P, TS = Cfg::TS, KC = Cfg::KC, NT = Cfg::threads, SG = Cfg::SG, PS = Cfg::PS;
    constexpr int PR = P;  // row stride of the matrices
    constexpr int XC = (Cfg::slots(KC) + NT - 1) / NT;
    __shared__ R T1[KC * PP];       // [k][b][a]
    __shared__ R T2[KC * P * TS];   // [k][b'][a] rows of TS
    // U_s(r, c) = M_s[r + P c] (parent -> child) or M_s[c + P r] (child -> parent), read
    // row-major from mats_rm[(transposed) 2 + s][r][c] at wave-uniform addresses: scalar
    // loads into SGPRs (a broadcast LDS read of 16 bytes would feed only 2 FP64 / 4 FP32 FMAs
    // per lane)
    const int ns = (P + kc - 1) / kc, lg = xcd_block(blockIdx.x, gridDim.x);
    const int it = lg / ns;
    if (it >= n_items) return;
    const int k0 = (lg - it * ns) * kc;
    const R *Ubase = mats_rm + (TRANSPOSE ? 2 : 0) * PP;
    const int nk = min(kc, P - k0), nslot = Cfg::slots(nk);
    R *out = out_base + (size_t)item_out[it] * P * PP;
    const bool add = item_add[it] != 0;
    const int e0 = item_ptr[it], e1 = item_ptr[it + 1];
    // s += sum_c U(r, c) x[c] for the row of U starting at urow
    auto row_dot = [](const R *urow, const R (&x)[P], R s) {
#pragma unroll
        for (int c = 0; c < P; ++c) s = fma(urow[c], x[c], s);
        return s;
    };
    // the same sums for four consecutive rows at once (four independent dependency chains)
    auto row_dot4 = [](const R *urow, const R (&x)[P], R (&o)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = R(0);
#pragma unroll
        for (int c = 0; c < P; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) o[i] = fma(urow[i * PR + c], x[c], o[i]);
    };
    // x-step accumulators: work slot (column (k, b'), segment of a') -- persistent over children
    R acc[XC][PS];
#pragma unroll
    for (int q = 0; q < XC; ++q)
#pragma unroll
        for (int a = 0; a < PS; ++a) acc[q][a] = R(0);
    for (int e = e0; e < e1; ++e) {
        const R *in = in_base + (size_t)in_slot[e] * P * PP;
        const int code = in_code[e];
        const R *U0 = Ubase + (code & 1) * P * PR, *U1 = Ubase + ((code >> 1) & 1) * P * PR,
                *U2 = Ubase + ((code >> 2) & 1) * P * PR;
        __syncthreads();
        // z: T1[k][b][a] = sum_c U2(k0 + k, c) in[c][b][a]
        for (int j = threadIdx.x; j < PP; j += NT) {
            R x[P];
#pragma unroll
            for (int c = 0; c < P; ++c) x[c] = in[c * PP + j];
            int k = 0;
            for (; k + 4 <= nk; k += 4) {  // four independent sums at a time
                R o[4];
                row_dot4(U2 + (k0 + k) * PR, x, o);
#pragma unroll
                for (int i = 0; i < 4; ++i) T1[(k + i) * PP + j] = o[i];
            }
            for (; k < nk; ++k) T1[k * PP + j] = row_dot(U2 + (k0 + k) * PR, x, R(0));
        }
        __syncthreads();
        // y: T2[k][b'][a] = sum_b U1(b', b) T1[k][b][a]; slot = (column (k, a), segment of b')
        for (int w = threadIdx.x; w < nslot; w += NT) {
            int col, sg;
            t3_slot<SG>(w, col, sg);
            if (col >= nk * P) continue;
            const int k = col / P, a = col - k * P;
            R x[P];
#pragma unroll
            for (int b = 0; b < P; ++b) x[b] = T1[k * PP + b * P + a];
            const int b2e = min(P, sg * PS + PS);
            int b2 = sg * PS;
            for (; b2 + 4 <= b2e; b2 += 4) {
                R o[4];
                row_dot4(U1 + b2 * PR, x, o);
#pragma unroll
                for (int i = 0; i < 4; ++i) T2[(k * P + b2 + i) * TS + a] = o[i];
            }
            for (; b2 < b2e; ++b2) T2[(k * P + b2) * TS + a] = row_dot(U1 + b2 * PR, x, R(0));
        }
        __syncthreads();
        // x: acc[a'] += sum_a U0(a', a) T2[k][b'][a]; slot = (column (k, b'), segment of a')
#pragma unroll
        for (int q = 0; q < XC; ++q) {
            const int w = threadIdx.x + q * NT;
            int col, sg;
            t3_slot<SG>(w, col, sg);
            if (w < nslot && col < nk * P) {
                R x[P];
#pragma unroll
                for (int a = 0; a < P; ++a) x[a] = T2[col * TS + a];
#pragma unroll
                for (int i = 0; i < PS; ++i)
                    if (sg * PS + i < P) acc[q][i] = row_dot(U0 + (sg * PS + i) * PR, x, acc[q][i]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < XC; ++q) {
        const int w = threadIdx.x + q * NT;
        int col, sg;
        t3_slot<SG>(w, col, sg);
        if (w < nslot && col < nk * P) {
            R *dst = out + (size_t)k0 * PP + (size_t)col * P + sg * PS;  // [c' = k0 + k][b'][a']
#pragma unroll
            for (int i = 0; i < PS; ++i)
                if (sg * PS + i < P) dst[i] = add ? dst[i] + acc[q][i] : acc[q][i];
        }
    }
}

// The same transform for FP64 at p = 28 on the matrix cores (v_mfma_f64_16x16x4f64,
// operand and accumulator layout as in k_charge2proxy_mfma).  Workgroup (item, ct) computes
// the output planes c' in [16 ct, 16 ct + 16), the item's inputs taken in slices of 4 values
// of a:
//   z:  T1[c'][(b, al)]  = sum_c U2(c', c) in[c][b][4 s + al]        (16 rows, 4 p columns)
//   y:  T2[(c', b')][al] = sum_b U1(b', b) T1[c'][(b, al)]            (rows b', 64 columns)
//   x:  out[(c', b')][a'] += sum_al U0(a', 4 s + al) T2[(c', b')][al] (rows a', 16 p columns)
// z and y go through LDS, x into accumulators that persist over the slices and the item's
// inputs (a wave owns x's column tiles w, w + NW, ...).  Every output sums its terms in the
// order of k_tensor3 (c, then b, then a ascending, inputs in order), so the two kernels agree
// to the last bit.  LDS holds both matrices transposed (Ut[c][r], zero-padded, so fragment
// reads need no bounds checks) and the two slice buffers (zero where c' or b is past p).
template <int P> struct T3MfmaCfg {
    static constexpr int PP = P * P;
    static constexpr int KS = (P + 3) / 4;        // 4-steps of a contraction = slices of a
    static constexpr int CT = (P + 15) / 16;      // 16-row tiles of an output index
    static constexpr int NW = P >= 38 ? 8 : 4;    // waves per workgroup
    static constexpr int XT = (P + NW - 1) / NW;  // x column tiles per wave (16 p columns = p tiles)
    static constexpr int LP = 16 * CT + 8;        // row stride of Ut (8 mod 16: a fragment's rows
    static constexpr int S1 = 16 * KS + 8;        // and those of T1 fall on different banks)
};

template <int P, bool TRANSPOSE>
__global__ __launch_bounds__(64 * T3MfmaCfg<P>::NW) void k_tensor3_mfma(
    const double *__restrict__ in_base, double *__restrict__ out_base, const int *__restrict__ item_out,
    const int *__restrict__ item_add, const int *__restrict__ item_ptr, const int *__restrict__ in_slot,
    const int *__restrict__ in_code, const double *__restrict__ mats_rm, int n_items, long long vs_in = 0,
    long long vs_out = 0) {
    chain_priority();
    in_base += (long long)blockIdx.z * vs_in;  // vector blockIdx.z of a many-vector launch
    out_base += (long long)blockIdx.z * vs_out;
    using Cfg = T3MfmaCfg<P>;
    using M = Mfma16<double>;
    typedef M::acc_t acc_t;
    constexpr int PP = Cfg::PP, KS = Cfg::KS, CT = Cfg::CT, NW = Cfg::NW, XT = Cfg::XT, LP = Cfg::LP, S1 = Cfg::S1,
                  NT = 64 * NW;
    __shared__ double Ut[2 * 4 * KS * LP];  // Ut[s][c][r] = U_s(r, c), zero outside r, c < p
    __shared__ double T1[16 * S1];          // [c' - 16 ct][(b, al)]
    __shared__ double T2[16 * P * 4];       // [(c' - 16 ct, b')][al]
    const int lg = xcd_block(blockIdx.x, gridDim.x), it = lg / CT, ct = lg - it * CT;
    if (it >= n_items) return;
    const double *Ubase = mats_rm + (TRANSPOSE ? 2 : 0) * PP;
    for (int i = threadIdx.x; i < 2 * 4 * KS * LP; i += NT) {
        const int sd = i / (4 * KS * LP), c = (i / LP) % (4 * KS), r = i % LP;
        Ut[i] = (r < P && c < P) ? Ubase[sd * PP + r * P + c] : 0.0;
    }
    for (int i = threadIdx.x; i < 16 * S1; i += NT) T1[i] = 0.0;
    for (int i = threadIdx.x; i < 16 * P * 4; i += NT) T2[i] = 0.0;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), l = threadIdx.x & 63, kk = l >> 4,
              jj = l & 15;
    acc_t acc[XT][CT];
#pragma unroll
    for (int t = 0; t < XT; ++t)
#pragma unroll
        for (int mt = 0; mt < CT; ++mt) acc[t][mt] = acc_t{0.0, 0.0, 0.0, 0.0};
    const int e0 = item_ptr[it], e1 = item_ptr[it + 1];
    for (int e = e0; e < e1; ++e) {
        const double *in = in_base + (size_t)in_slot[e] * P * PP;
        const int code = in_code[e];
        const double *U0 = Ut + (code & 1) * 4 * KS * LP, *U1 = Ut + ((code >> 1) & 1) * 4 * KS * LP,
                     *U2 = Ut + ((code >> 2) & 1) * 4 * KS * LP;
        for (int s = 0; s < KS; ++s) {
            __syncthreads();  // Ut and the zero padding written; the previous y done with T1
            // z: one job per 16 columns (b, al)
            for (int nt = w; nt < KS; nt += NW) {
                const int n = 16 * nt + jj, b = n >> 2, a = 4 * s + (n & 3);
                acc_t d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const int c = 4 * ks + kk;
                    const double bv = (c < P && b < P && a < P) ? in[(size_t)c * PP + b * P + a] : 0.0;
                    d = M::mma(U2[c * LP + 16 * ct + jj], bv, d);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int cl = M::row(kk, q);
                    if (16 * ct + cl < P && b < P) T1[cl * S1 + n] = d[q];
                }
            }
            __syncthreads();
            // y: job = (row tile of b', one of the 4 column tiles of (c', al))
            for (int job = w; job < 4 * CT; job += NW) {
                const int mt = job % CT, nt = job / CT;
                const int n = 16 * nt + jj, cl = n >> 2, al = n & 3;
                acc_t d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const int b = 4 * ks + kk;
                    d = M::mma(U1[b * LP + 16 * mt + jj], T1[cl * S1 + 4 * b + al], d);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int bp = 16 * mt + M::row(kk, q);
                    if (bp < P && 16 * ct + cl < P) T2[(cl * P + bp) * 4 + al] = d[q];
                }
            }
            __syncthreads();
            // x: this wave's column tiles of (c', b'), one 4-step over a = 4 s + al
#pragma unroll
            for (int t = 0; t < XT; ++t) {
                const int nt = w + NW * t;
                if (nt < P) {
                    const double bv = T2[(16 * nt + jj) * 4 + kk];
#pragma unroll
                    for (int mt = 0; mt < CT; ++mt)
                        acc[t][mt] = M::mma(U0[(4 * s + kk) * LP + 16 * mt + jj], bv, acc[t][mt]);
                }
            }
        }
    }
    double *out = out_base + (size_t)item_out[it] * P * PP + (size_t)16 * ct * PP;
    const bool add = item_add[it] != 0;
#pragma unroll
    for (int t = 0; t < XT; ++t) {
        const int n = 16 * (w + NW * t) + jj;  // (c' - 16 ct) p + b'
        if (w + NW * t >= P || 16 * ct + n / P >= P) continue;
#pragma unroll
        for (int mt = 0; mt < CT; ++mt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int ap = 16 * mt + M::row(kk, q);
                if (ap < P) {
                    double *dst = out + (size_t)n * P + ap;
                    *dst = add ? *dst + acc[t][mt][q] : acc[t][mt][q];
                }
            }
    }
}

// proxy[parent_slot[i]] += sum_{k in child_ptr[i] .. child_ptr[i+1]} scratch[k], element-wise, the
// children added in order (deterministic).
template <class R, int P>
__global__ __launch_bounds__(256) void k_sum_children(R *__restrict__ proxy, const R *__restrict__ scratch,
                                                       const int *__restrict__ parent_slot, const int *__restrict__ child_ptr,
                                                       int n_parents, long long vs_proxy = 0, long long vs_scratch = 0) {
    chain_priority();
    proxy += (long long)blockIdx.z * vs_proxy;  // vector blockIdx.z of a many-vector launch
    scratch += (long long)blockIdx.z * vs_scratch;
    constexpr int P3 = P * P * P;
    const int it = blockIdx.x, e = blockIdx.y * blockDim.x + threadIdx.x;
    if (it >= n_parents || e >= P3) return;
    R *dst = proxy + (size_t)parent_slot[it] * P3 + e;
    R v = *dst;
    for (int k = child_ptr[it]; k < child_ptr[it + 1]; ++k) v += scratch[(size_t)k * P3 + e];
    *dst = v;
}

// ------------------------------------------------------------------ plane waves
// Plane-wave mode layout: m = k1 * (H n) + k3 * n + k2 ([x][z][y], y fastest), k3 < H.
// The polynomial <-> plane-wave matrices have the parity of the Chebyshev polynomials:
// W(k, j) = sum_i Vinv(j, i) e^{-i (k - n/2) x_i h} over nodes x_i symmetric about 0 with
// T_j(-x) = (-1)^j T_j(x), so W(k, j) is real for even j and imaginary for odd j (the other
// part is rounding, below 1e-15 of the largest entry).  The transforms use that: a product
// with W(k, j) costs one real multiply-add per real component instead of two.
// Outgoing expansion pw(k1,k2,k3) = sum W(k1,a) W(k2,b) W(k3,c) C[c][b][a], W = poly2pw
// (W(k, j) at W[k + n j]); one workgroup per (box, slice of k3), 1-D grid, (box, slice) at
// xcd_block index box * slices + slice (a box's slices share one L2).  Box x reads proxy
// slot slot0 + x and writes plane-wave slot x.
// The z step of whole boxes into HBM, for orders where the slices of k_proxy2pw would each
// re-read the box's proxy charges (20 slices per box for FP64 at p = 28, 28 at p = 38):
// t1[box][k][b][a] = sum_c W(k, c) C[c][b][a] for k < H, one thread per column (b, a), the
// same sums as k_proxy2pw's z step.  k_p2pw_yx then takes each plane of t1 through the y and
// x steps.
template <class R, int P, int NPW>
__global__ __launch_bounds__(256) void k_p2pw_z(const R *__restrict__ proxy, Cx<R> *__restrict__ t1,
                                                const Cx<R> *__restrict__ W, int slot0, int n_boxes) {
    chain_priority();
    constexpr int H = (NPW + 1) / 2, PP = P * P, PER = (PP + 255) / 256;  // workgroups per box
    const int lg = xcd_block(blockIdx.x, gridDim.x), bx = lg / PER, j = (lg - bx * PER) * 256 + (int)threadIdx.x;
    if (bx >= n_boxes || j >= PP) return;
    const R *C = proxy + (size_t)(slot0 + bx) * P * PP;
    Cx<R> *out = t1 + (size_t)bx * H * PP;
    R x[P];
#pragma unroll
    for (int c = 0; c < P; ++c) x[c] = C[c * PP + j];
    for (int k = 0; k < H; ++k) {
        R sr = 0, si = 0;
#pragma unroll
        for (int c = 0; c < P; c += 2) sr = fma(W[k + NPW * c].re, x[c], sr);
#pragma unroll
        for (int c = 1; c < P; c += 2) si = fma(W[k + NPW * c].im, x[c], si);
        out[(size_t)k * PP + j] = {sr, si};
    }
}

template <class R, int P, int NPW, int BK = 0>
__global__ __launch_bounds__(256) void k_proxy2pw(const R *__restrict__ proxy, Cx<R> *__restrict__ pw_out,
                                                   const Cx<R> *__restrict__ W, int slot0, int n_boxes,
                                                   long long vs_in = 0, long long vs_out = 0) {
    chain_priority();
    proxy += (long long)blockIdx.z * vs_in;  // vector blockIdx.z of a many-vector launch
    pw_out += (long long)blockIdx.z * vs_out;
    constexpr int H = (NPW + 1) / 2, COL = H * NPW, PP = P * P, TS = OddStride<P>::v;
    constexpr int K3 = P2pwCfg<R, P, NPW, BK>::K3;
    __shared__ Cx<R> T1[K3 * PP];        // [k][b][a]
    __shared__ Cx<R> T2[K3 * NPW * TS];  // [k][k2][a] rows of TS
    constexpr int NS = P2pwCfg<R, P, NPW, BK>::slices;
    const int lg = xcd_block(blockIdx.x, gridDim.x), bx = lg / NS, slice = lg - bx * NS;
    if (bx >= n_boxes) return;
    const R *C = proxy + (size_t)(slot0 + bx) * P * PP;
    Cx<R> *pw = pw_out + (size_t)bx * NPW * COL;
    {
        const int k30 = slice * K3;  // this workgroup's slice of k3
        const int nk = min(K3, H - k30);
        // z: T1[k][b][a] = sum_c W(k30 + k, c) C[c][b][a]
        for (int j = threadIdx.x; j < PP; j += blockDim.x) {
            R x[P];
#pragma unroll
            for (int c = 0; c < P; ++c) x[c] = C[c * PP + j];
            for (int k = 0; k < nk; ++k) {
                R sr = 0, si = 0;
#pragma unroll
                for (int c = 0; c < P; c += 2) sr = fma(W[k30 + k + NPW * c].re, x[c], sr);
#pragma unroll
                for (int c = 1; c < P; c += 2) si = fma(W[k30 + k + NPW * c].im, x[c], si);
                T1[k * PP + j] = {sr, si};
            }
        }
        __syncthreads();
        // y: T2[k][k2][a] = sum_b W(k2, b) T1[k][b][a]; slot = (column (k, a), segment of k2,
        // uniform across a wavefront)
        constexpr int SGY = P2pwCfg<R, P, NPW, BK>::SGY, PSY = P2pwCfg<R, P, NPW, BK>::PSY;
        const int ny = (nk * P + 63) / 64 * 64 * SGY;
        for (int t = threadIdx.x; t < ny; t += blockDim.x) {
            int w, sg;
            t3_slot<SGY>(t, w, sg);
            if (w >= nk * P) continue;
            const int k = w / P, a = w - k * P;
            Cx<R> v[P];
#pragma unroll
            for (int b = 0; b < P; ++b) v[b] = T1[k * PP + b * P + a];
            for (int k2 = sg * PSY; k2 < min(NPW, sg * PSY + PSY); ++k2) {
                R sr = 0, si = 0;
#pragma unroll
                for (int b = 0; b < P; ++b) {
                    if (b % 2 == 0) {
                        const R m = W[k2 + NPW * b].re;
                        sr = fma(m, v[b].re, sr);
                        si = fma(m, v[b].im, si);
                    } else {
                        const R m = W[k2 + NPW * b].im;
                        sr = fma(-m, v[b].im, sr);
                        si = fma(m, v[b].re, si);
                    }
                }
                T2[(k * NPW + k2) * TS + a] = {sr, si};
            }
        }
        __syncthreads();
        // x: pw[k1][k30 + k][k2] = sum_a W(k1, a) T2[k][k2][a]; slot = (column w = k n + k2,
        // segment of k1, uniform across a wavefront)
        constexpr int SGX = P2pwCfg<R, P, NPW, BK>::SGX, PSX = P2pwCfg<R, P, NPW, BK>::PSX;
        const int nx = (nk * NPW + 63) / 64 * 64 * SGX;
        for (int t = threadIdx.x; t < nx; t += blockDim.x) {
            int w, sg;
            t3_slot<SGX>(t, w, sg);
            if (w >= nk * NPW) continue;
            Cx<R> v[P];
#pragma unroll
            for (int a = 0; a < P; ++a) v[a] = T2[w * TS + a];
            Cx<R> *dst = pw + (size_t)k30 * NPW + w;
            for (int k1 = sg * PSX; k1 < min(NPW, sg * PSX + PSX); ++k1) {
                R sr = 0, si = 0;
#pragma unroll
                for (int a = 0; a < P; ++a) {
                    if (a % 2 == 0) {
                        const R m = W[k1 + NPW * a].re;
                        sr = fma(m, v[a].re, sr);
                        si = fma(m, v[a].im, si);
                    } else {
                        const R m = W[k1 + NPW * a].im;
                        sr = fma(-m, v[a].im, sr);
                        si = fma(m, v[a].re, si);
                    }
                }
                dst[(size_t)k1 * COL] = {sr, si};
            }
        }
    }
}

// Incoming expansion of one level: pin[i][m] = lscale sum_e pw_out[slot_e - slot0][m] shift0[sten_e][m]
// over box i's CSR range of (neighbour slot, stencil); shift0 holds the level-0 translation
// times the radial Fourier factor, lscale = 2^level scales the latter to this level.
template <class R>
__global__ __launch_bounds__(256) void k_pw_shift(const Cx<R> *__restrict__ pw_out, const Cx<R> *__restrict__ shift0,
                                                   R lscale, const int *__restrict__ nbr_ptr,
                                                   const int *__restrict__ nbr_slot, const int *__restrict__ nbr_sten,
                                                   int slot0, int nm, int item0, Cx<R> *__restrict__ pin,
                                                   long long vs_in = 0, long long vs_out = 0) {
    chain_priority();
    pw_out += (long long)blockIdx.z * vs_in;  // vector blockIdx.z of a many-vector launch
    pin += (long long)blockIdx.z * vs_out;
    // x = target box (consecutive workgroups share neighbours; each XCD takes a contiguous,
    // spatially coherent range of boxes), y = mode chunk: the neighbours' chunk is read from
    // L2 by the 27 boxes around it instead of from memory
    const int bi = xcd_block(blockIdx.x, gridDim.x), it = item0 + bi;
    const int m = blockIdx.y * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    R sr = 0, si = 0;
    for (int e = nbr_ptr[it]; e < nbr_ptr[it + 1]; ++e) {
        const Cx<R> a = pw_out[(size_t)(nbr_slot[e] - slot0) * nm + m];
        const Cx<R> s = shift0[(size_t)nbr_sten[e] * nm + m];
        sr = fma(a.re, s.re, fma(-a.im, s.im, sr));
        si = fma(a.re, s.im, fma(a.im, s.re, si));
    }
    pin[(size_t)bi * nm + m] = {sr * lscale, si * lscale};
}

// The same sum in FP64 with the shift factors formed from their separable parts instead of
// read from the 27-stencil table (27 n^2 h complex values, 13 MB at p = 28, more than an XCD's
// L2, so every target box's shifts streamed it from memory beside the neighbours' plane
// waves): shift0[t][m] = radial0[m] e^{i (x_c dx + x_a dz + x_b dy)} for the mode
// m = (c h + a) n + b and stencil t = 9 (dz + 1) + 3 (dy + 1) + (dx + 1), x_j = (j - n/2) sphase;
// ephase[j] = e^{i x_j}, taken conjugated for an offset of -1 and left out for 0.  The radial
// factor multiplies the sum once.
__global__ __launch_bounds__(256) void k_pw_shift_sep(const Cx<double> *__restrict__ pw_out,
                                                      const double *__restrict__ radial0,
                                                      const Cx<double> *__restrict__ ephase, double lscale,
                                                      const int *__restrict__ nbr_ptr, const int *__restrict__ nbr_slot,
                                                      const int *__restrict__ nbr_sten, int slot0, int nm, int item0,
                                                      int n, int h, Cx<double> *__restrict__ pin, long long vs_in = 0,
                                                      long long vs_out = 0) {
    chain_priority();
    pw_out += (long long)blockIdx.z * vs_in;  // vector blockIdx.z of a many-vector launch
    pin += (long long)blockIdx.z * vs_out;
    const int bi = xcd_block(blockIdx.x, gridDim.x), it = item0 + bi;
    const int m = blockIdx.y * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const int b = m % n, a = (m / n) % h, c = m / (n * h);
    const Cx<double> ex = ephase[c], ez = ephase[a], ey = ephase[b];
    double sr = 0, si = 0;
    for (int e = nbr_ptr[it]; e < nbr_ptr[it + 1]; ++e) {
        const int t = nbr_sten[e], dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;  // uniform
        double fr = 1, fi = 0;
        if (dz) {
            fr = ez.re;
            fi = dz > 0 ? ez.im : -ez.im;
        }
        if (dy) {
            const double gi = dy > 0 ? ey.im : -ey.im, r = fr * ey.re - fi * gi;
            fi = fma(fr, gi, fi * ey.re);
            fr = r;
        }
        if (dx) {
            const double gi = dx > 0 ? ex.im : -ex.im, r = fr * ex.re - fi * gi;
            fi = fma(fr, gi, fi * ex.re);
            fr = r;
        }
        const Cx<double> av = pw_out[(size_t)(nbr_slot[e] - slot0) * nm + m];
        sr = fma(av.re, fr, fma(-av.im, fi, sr));
        si = fma(av.re, fi, fma(av.im, fr, si));
    }
    const double rr = radial0[m] * lscale;
    pin[(size_t)bi * nm + m] = {sr * rr, si * rr};
}

// Folded matrix table of the plane-wave transforms: F[j * 64 + k] = W(k, j).re for even j,
// W(k, j).im for odd j (the entry's only non-zero part), 0 for k >= n; k in a row of 64.
//
// Second pass of the two-pass proxy -> plane-wave transform (after k_p2pw_z), FP64 on the
// matrix cores: the y and x steps of one k3 plane per wavefront.  With the parity folded into
// the operands, each step is a pair of real products with F:
//   y:  T2t[a][k2] = sum_b X[b][a] F[b][k2],  X = (T.re, T.im) for even b, (-T.im, T.re) for odd b
//   x:  pw[k1][k2] = sum_a F[a][k1] Z[a][k2], Z = (T2.re, T2.im) for even a, (-T2.im, T2.re) odd a
// (T[b][a] = the plane of t1).  v_mfma_f64_16x16x4f64 leaves y's output tile with lane l
// holding rows a = 4 e + l/16 of column l%16, which is the B operand x needs for the
// 4-row steps e = 0..3 of that tile, so the plane goes from y to x in registers, one 16-column
// tile of k2 at a time.  F sits in LDS (A and B fragments of both steps); the plane's t1 is
// read once per wavefront into registers.  Each output sums its terms in the same order as
// k_proxy2pw (b, then a, ascending).
template <int P, int NPW>
__global__ __launch_bounds__(256, 2) void k_p2pw_yx(const Cx<double> *__restrict__ t1, const double *__restrict__ F,
                                                 Cx<double> *__restrict__ pw_out, int n_boxes) {
    chain_priority();
    static_assert(NPW <= 64, "k2 and k1 within a table row of 64");
    using M = Mfma16<double>;
    typedef M::acc_t acc_t;
    constexpr int H = (NPW + 1) / 2, COL = H * NPW, PP = P * P, PG = (H + 3) / 4;  // 4 planes per workgroup
    constexpr int KS = (P + 3) / 4, AT = (P + 15) / 16, CT = (NPW + 15) / 16;     // 4-steps, a tiles, k tiles
    __shared__ double Fs[4 * KS * 64];
    for (int i = threadIdx.x; i < 4 * KS * 64; i += 256) Fs[i] = i < P * 64 ? F[i] : 0.0;
    __syncthreads();
    const int lg = xcd_block(blockIdx.x, gridDim.x), bx = lg / PG;
    const int k = (lg - bx * PG) * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    if (bx >= n_boxes || k >= H) return;
    const int l = threadIdx.x & 63, kk = l >> 4, jj = l & 15;
    const bool odd = kk & 1;  // parity of b = 4 ks + kk (y) and of a = 16 at + 4 e + kk (x)
    const Cx<double> *T = t1 + ((size_t)bx * H + k) * PP;
    // y's A fragments: X[b = 4 ks + kk][a = 16 at + jj], real-output and imaginary-output parts
    double xr[KS][AT], xi[KS][AT];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int at = 0; at < AT; ++at) {
            const int b = 4 * ks + kk, a = 16 * at + jj;
            Cx<double> x = {0.0, 0.0};
            if (b < P && a < P) x = T[b * P + a];
            xr[ks][at] = odd ? -x.im : x.re;
            xi[ks][at] = odd ? x.re : x.im;
        }
    Cx<double> *dst = pw_out + (size_t)bx * NPW * COL + (size_t)k * NPW;
    // one k2 tile at a time (unrolled, the tiles' accumulators would take the registers of a
    // second wavefront per SIMD at p = 38)
#pragma unroll 1
    for (int ct = 0; ct < CT; ++ct) {  // k2 tile
        acc_t yr[AT], yi[AT];
#pragma unroll
        for (int at = 0; at < AT; ++at) yr[at] = yi[at] = acc_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const double f = Fs[(4 * ks + kk) * 64 + 16 * ct + jj];  // B[b][k2] = F[b][k2]
#pragma unroll
            for (int at = 0; at < AT; ++at) {
                yr[at] = M::mma(xr[ks][at], f, yr[at]);
                yi[at] = M::mma(xi[ks][at], f, yi[at]);
            }
        }
        // yr/yi[at][e] = T2 (re, im) at a = 16 at + 4 e + kk, k2 = 16 ct + jj: fold into Z
#pragma unroll
        for (int at = 0; at < AT; ++at)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double re = yr[at][e], im = yi[at][e];
                yr[at][e] = odd ? -im : re;
                yi[at][e] = odd ? re : im;
            }
        for (int mt = 0; mt < CT; ++mt) {  // k1 tile
            acc_t pr = {0.0, 0.0, 0.0, 0.0}, pi = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int at = 0; at < AT; ++at)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (4 * at + e < KS) {
                        const double fa = Fs[(4 * (4 * at + e) + kk) * 64 + 16 * mt + jj];  // A[k1][a] = F[a][k1]
                        pr = M::mma(fa, yr[at][e], pr);
                        pi = M::mma(fa, yi[at][e], pi);
                    }
                }
            const int k2 = 16 * ct + jj;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k1 = 16 * mt + M::row(kk, e);
                if (k1 < NPW && k2 < NPW) dst[(size_t)k1 * COL + k2] = {pr[e], pi[e]};
            }
        }
    }
}

// Shifts of groups of kShiftGroup consecutive target boxes (siblings, mostly): the group's
// neighbour lists merged by source slot, so each neighbour's plane waves are read once for
// the whole group (a 2x2x2 group of boxes has 64 distinct neighbours instead of 8 x 27).
// Entry e of group g: source slot gslot[e], and for target t of the group the stencil in
// byte t of gsten[e] (255: not a neighbour of t).  The shift factor comes from the
// 27-stencil table (FP32) or from the separable phases (FP64, as k_pw_shift_sep).  Targets
// sum their neighbours in slot order.
constexpr int kShiftGroup = 8;
template <class R>
__global__ __launch_bounds__(256) void k_pw_shift_grp(const Cx<R> *__restrict__ pw_out, const Cx<R> *__restrict__ shift0,
                                                      const double *__restrict__ radial0,
                                                      const Cx<double> *__restrict__ ephase, R lscale,
                                                      const int *__restrict__ gptr, const int *__restrict__ gslot,
                                                      const unsigned long long *__restrict__ gsten, int slot0, int nm,
                                                      int item0, int n_items, int n, int h, Cx<R> *__restrict__ pin,
                                                      long long vs_in = 0, long long vs_out = 0) {
    chain_priority();
    constexpr int G = kShiftGroup;
    pw_out += (long long)blockIdx.z * vs_in;  // vector blockIdx.z of a many-vector launch
    pin += (long long)blockIdx.z * vs_out;
    // workgroup x = group of the batch (item0 is a multiple of G); y = mode chunk
    const int g = item0 / G + xcd_block(blockIdx.x, gridDim.x);
    const int m = blockIdx.y * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    Cx<R> ex{}, ey{}, ez{};
    if constexpr (sizeof(R) == 8) {
        const int b = m % n, a = (m / n) % h, c = m / (n * h);
        ex = ephase[c];
        ez = ephase[a];
        ey = ephase[b];
    }
    R sr[G], si[G];
#pragma unroll
    for (int t = 0; t < G; ++t) sr[t] = si[t] = R(0);
    for (int e = gptr[g]; e < gptr[g + 1]; ++e) {
        const unsigned long long st = gsten[e];  // uniform
        const Cx<R> av = pw_out[(size_t)(gslot[e] - slot0) * nm + m];
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const int sten = (int)((st >> (8 * t)) & 255);
            if (sten == 255) continue;  // uniform
            R fr, fi;
            if constexpr (sizeof(R) == 8) {
                const int dz = sten / 9 - 1, dy = (sten / 3) % 3 - 1, dx = sten % 3 - 1;
                fr = 1;
                fi = 0;
                if (dz) {
                    fr = ez.re;
                    fi = dz > 0 ? ez.im : -ez.im;
                }
                if (dy) {
                    const R gi = dy > 0 ? ey.im : -ey.im, r = fr * ey.re - fi * gi;
                    fi = fma(fr, gi, fi * ey.re);
                    fr = r;
                }
                if (dx) {
                    const R gi = dx > 0 ? ex.im : -ex.im, r = fr * ex.re - fi * gi;
                    fi = fma(fr, gi, fi * ex.re);
                    fr = r;
                }
            } else {
                const Cx<R> f = shift0[(size_t)sten * nm + m];
                fr = f.re;
                fi = f.im;
            }
            sr[t] = fma(av.re, fr, fma(-av.im, fi, sr[t]));
            si[t] = fma(av.re, fi, fma(av.im, fr, si[t]));
        }
    }
    R rr = lscale;
    if constexpr (sizeof(R) == 8) rr = radial0[m] * lscale;
#pragma unroll
    for (int t = 0; t < G; ++t) {
        const int it = g * G + t - item0;  // item within the batch
        if (it < n_items) pin[(size_t)it * nm + m] = {sr[t] * rr, si[t] * rr};
    }
}

// down[c][b][a] (+)= 2 Re sum V(k1,a) V(k2,b) h(k3) V(k3,c) pin[k1][k3][k2], V = pw2poly
// (V(k, j) at V[k + n j]), h = 1/2 for k3 >= n/2 (planewave_to_proxy_potential).  One
// workgroup per (box, slice of a).
// The x step of whole boxes into HBM, for orders where the slices of k_pw2proxy would each
// re-read the box's incoming expansion (14 slices per box for FP64 at p = 28, 38 at p = 38):
// u1[it][a][k3][k2] = sum_k1 V(k1, a) pin[it][k1][k3][k2] for every a, one thread per column
// (k3, k2), the same sums as k_pw2proxy's x step.  k_pw2p_yz then takes each plane a of u1
// through the y and z steps.
template <class R, int P, int NPW>
__global__ __launch_bounds__(256) void k_pw2p_x(const Cx<R> *__restrict__ pin, const Cx<R> *__restrict__ V,
                                                int n_items, Cx<R> *__restrict__ u1) {
    chain_priority();
    constexpr int H = (NPW + 1) / 2, COL = H * NPW, PER = (COL + 255) / 256;  // workgroups per box
    const int lg = xcd_block(blockIdx.x, gridDim.x), it = lg / PER, col = (lg - it * PER) * 256 + (int)threadIdx.x;
    if (it >= n_items || col >= COL) return;
    const Cx<R> *X = pin + (size_t)it * NPW * COL;
    Cx<R> acc[P];
#pragma unroll
    for (int a = 0; a < P; ++a) acc[a] = {R(0), R(0)};
    for (int k1 = 0; k1 < NPW; ++k1) {
        const Cx<R> x = X[(size_t)k1 * COL + col];
#pragma unroll
        for (int a = 0; a < P; ++a) {
            if (a % 2 == 0) {
                const R m = V[k1 + NPW * a].re;
                acc[a].re = fma(m, x.re, acc[a].re);
                acc[a].im = fma(m, x.im, acc[a].im);
            } else {
                const R m = V[k1 + NPW * a].im;
                acc[a].re = fma(-m, x.im, acc[a].re);
                acc[a].im = fma(m, x.re, acc[a].im);
            }
        }
    }
    Cx<R> *out = u1 + (size_t)it * P * COL + col;
#pragma unroll
    for (int a = 0; a < P; ++a) out[(size_t)a * COL] = acc[a];
}

// Second pass of the two-pass plane-wave -> proxy transform (after k_pw2p_x), FP64 on the
// matrix cores: the y and z steps of one plane a per wavefront.  With G[k][j] the folded
// table (the non-zero part of V(k, j): real for even j, imaginary for odd j):
//   y:  R[k3][b] = sum_k2 u1.re[a][k3][k2] G[k2][b],  I[k3][b] = the same of u1.im;
//       U2[k3][b] = h(k3) (R, I) for even b, h(k3) (-I, R) for odd b
//   z:  Zr[c][b] = sum_k3 G[k3][c] U2.re[k3][b],  Zi[c][b] = the same of U2.im;
//       down[c][b][a] (+)= 2 Zr for even c, 2 (-Zi) for odd c
// (the parity of the output index picks the part, so each step forms both products and
// keeps one; a negated sum of fused multiply-adds is the sum of the negated terms, bit for
// bit).  y's output tile, lane l holding rows k3 = 4 q + l/16 of column b = l%16, is z's B
// operand for the 4-steps q of that tile, so U2 stays in registers.  LDS holds G transposed
// (Gt[k][j], zero-padded).  Every output sums its terms in the order of k_pw2proxy (k2, then
// k3 ascending).
template <int P, int NPW> struct Pw2pMfmaCfg {
    static constexpr int H = (NPW + 1) / 2;
    static constexpr int KS2 = (NPW + 3) / 4;     // y's 4-steps over k2
    static constexpr int MT = (H + 15) / 16;      // y's row tiles over k3
    static constexpr int KSZ = (H + 3) / 4;       // z's 4-steps over k3
    static constexpr int BT = (P + 15) / 16;      // column tiles over b, row tiles over c
    static constexpr int GR = 4 * (KS2 > 4 * MT ? KS2 : 4 * MT);  // rows of Gt
    static constexpr int LG = 16 * BT + 8;        // row stride of Gt (8 mod 16)
};

template <int P, int NPW>
__global__ __launch_bounds__(256) void k_pw2p_yz(const Cx<double> *__restrict__ u1, const double *__restrict__ F,
                                                 const int *__restrict__ box_down, const int *__restrict__ box_zero,
                                                 int item0, int n_items, double *__restrict__ down_base) {
    chain_priority();
    using Cfg = Pw2pMfmaCfg<P, NPW>;
    using M = Mfma16<double>;
    typedef M::acc_t acc_t;
    constexpr int H = Cfg::H, COL = H * NPW, PP = P * P, KS2 = Cfg::KS2, MT = Cfg::MT, KSZ = Cfg::KSZ, BT = Cfg::BT,
                  GR = Cfg::GR, LG = Cfg::LG, AG = (P + 3) / 4;  // 4 planes a per workgroup
    static_assert(NPW <= 64 && 16 * BT <= 64, "table rows of 64");
    __shared__ double Gt[GR * LG];  // Gt[k][j] = G[k][j] for k < n, j < p, else 0
    for (int i = threadIdx.x; i < GR * LG; i += 256) {
        const int k = i / LG, j = i - k * LG;
        Gt[i] = (k < NPW && j < P) ? F[j * 64 + k] : 0.0;
    }
    __syncthreads();
    const int lg = xcd_block(blockIdx.x, gridDim.x), it = lg / AG;
    const int a = (lg - it * AG) * 4 + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    if (it >= n_items || a >= P) return;
    const int l = threadIdx.x & 63, kk = l >> 4, jj = l & 15;
    const Cx<double> *X = u1 + ((size_t)it * P + a) * COL;  // X[k3][k2]
    // y: R, I tiles (k3 tile mt, b tile nt)
    acc_t yr[MT][BT], yi[MT][BT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < BT; ++nt) yr[mt][nt] = yi[mt][nt] = acc_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < KS2; ++ks) {
        const int k2 = 4 * ks + kk;
        double g[BT];
#pragma unroll
        for (int nt = 0; nt < BT; ++nt) g[nt] = Gt[k2 * LG + 16 * nt + jj];  // B[k2][b]
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int k3 = 16 * mt + jj;
            Cx<double> x = {0.0, 0.0};
            if (k3 < H && k2 < NPW) x = X[k3 * NPW + k2];  // A[k3][k2]
#pragma unroll
            for (int nt = 0; nt < BT; ++nt) {
                yr[mt][nt] = M::mma(x.re, g[nt], yr[mt][nt]);
                yi[mt][nt] = M::mma(x.im, g[nt], yi[mt][nt]);
            }
        }
    }
    // U2 in place: element q of tile (mt, nt) is k3 = 16 mt + 4 q + kk, b = 16 nt + jj
    const bool bodd = jj & 1;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double hw = (16 * mt + M::row(kk, q) >= NPW / 2) ? 0.5 : 1.0;
#pragma unroll
            for (int nt = 0; nt < BT; ++nt) {
                const double r = yr[mt][nt][q], i = yi[mt][nt][q];
                yr[mt][nt][q] = (bodd ? -i : r) * hw;
                yi[mt][nt][q] = (bodd ? r : i) * hw;
            }
        }
    // z: c tile ct, b tile nt; 4-step ks = 4 mt + q over k3
    double *down = down_base + (size_t)box_down[item0 + it] * P * PP;
    const bool zero = box_zero[item0 + it] != 0;
    const bool codd = kk & 1;  // parity of c = 16 ct + 4 q + kk
#pragma unroll
    for (int ct = 0; ct < BT; ++ct) {
        acc_t zr[BT], zi[BT];
#pragma unroll
        for (int nt = 0; nt < BT; ++nt) zr[nt] = zi[nt] = acc_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (4 * mt + q < KSZ) {
                    const double gz = Gt[(16 * mt + 4 * q + kk) * LG + 16 * ct + jj];  // A[c][k3] = G[k3][c]
#pragma unroll
                    for (int nt = 0; nt < BT; ++nt) {
                        zr[nt] = M::mma(gz, yr[mt][nt][q], zr[nt]);
                        zi[nt] = M::mma(gz, yi[mt][nt][q], zi[nt]);
                    }
                }
            }
#pragma unroll
        for (int nt = 0; nt < BT; ++nt) {
            const int b = 16 * nt + jj;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = 16 * ct + M::row(kk, q);
                if (c < P && b < P) {
                    double *dst = down + ((size_t)c * P + b) * P + a;
                    const double v = codd ? -zi[nt][q] : zr[nt][q];
                    *dst = (zero ? 0.0 : *dst) + 2.0 * v;
                }
            }
        }
    }
}

template <class R, int P, int NPW, int BK = 0>
__global__ __launch_bounds__(256) void k_pw2proxy(const Cx<R> *__restrict__ pin, const Cx<R> *__restrict__ V,
                                                   const int *__restrict__ box_down, const int *__restrict__ box_zero,
                                                   int item0, int n_items, R *__restrict__ down_base,
                                                   long long vs_in = 0, long long vs_out = 0) {
    chain_priority();
    pin += (long long)blockIdx.z * vs_in;  // vector blockIdx.z of a many-vector launch
    down_base += (long long)blockIdx.z * vs_out;
    constexpr int H = (NPW + 1) / 2, COL = H * NPW, PP = P * P, half = NPW / 2, TS2 = OddStride<P>::v;
    constexpr int AC = Pw2pCfg<R, P, NPW, BK>::AC;
    __shared__ Cx<R> U1[AC * COL];       // [a][k3][k2]
    __shared__ Cx<R> U2[AC * H * TS2];   // [a][k3][b] rows of TS2
    // 1-D grid: (box, slice of a) at xcd_block index box * slices + slice
    constexpr int NS = Pw2pCfg<R, P, NPW, BK>::slices;
    const int lg = xcd_block(blockIdx.x, gridDim.x), it = lg / NS, slice = lg - it * NS;
    if (it >= n_items) return;
    const Cx<R> *X = pin + (size_t)it * NPW * COL;
    R *down = down_base + (size_t)box_down[item0 + it] * P * PP;
    const bool zero = box_zero[item0 + it] != 0;
    {
        const int a0 = slice * AC;
        const int na = min(AC, P - a0);
        // Each step's outputs are polynomial indices (a0 + a, b0 + b, c0 + c): the matrix row of
        // an output is real (even index) or imaginary (odd).  The parity of the first output
        // (uniform across a workgroup or a wavefront) selects a template instance, in which the
        // unrolled loops fold every output's parity.
        using Even = std::integral_constant<int, 0>;
        using Odd = std::integral_constant<int, 1>;
        // x: U1[a][k3][k2] = sum_k1 V(k1, a0 + a) pin[k1][k3][k2], column = (k3, k2)
        auto xstep = [&](auto par) {
            constexpr int A0P = decltype(par)::value;
            for (int col = threadIdx.x; col < COL; col += blockDim.x) {
                Cx<R> acc[AC];
#pragma unroll
                for (int a = 0; a < AC; ++a) acc[a] = {R(0), R(0)};
                for (int k1 = 0; k1 < NPW; ++k1) {
                    const Cx<R> x = X[(size_t)k1 * COL + col];
#pragma unroll
                    for (int a = 0; a < AC; ++a) {
                        if (a < na) {
                            if ((A0P + a) % 2 == 0) {
                                const R m = V[k1 + NPW * (a0 + a)].re;
                                acc[a].re = fma(m, x.re, acc[a].re);
                                acc[a].im = fma(m, x.im, acc[a].im);
                            } else {
                                const R m = V[k1 + NPW * (a0 + a)].im;
                                acc[a].re = fma(-m, x.im, acc[a].re);
                                acc[a].im = fma(m, x.re, acc[a].im);
                            }
                        }
                    }
                }
#pragma unroll
                for (int a = 0; a < AC; ++a)
                    if (a < na) U1[a * COL + col] = acc[a];
            }
        };
        if (a0 & 1) {
            xstep(Odd{});
        } else {
            xstep(Even{});
        }
        __syncthreads();
        // y: U2[a][k3][b] = h(k3) sum_k2 V(k2, b) U1[a][k3][k2]; slot = (column w = a H + k3,
        // segment of b, uniform across a wavefront)
        constexpr int SGY = Pw2pCfg<R, P, NPW, BK>::SGY, PSY = Pw2pCfg<R, P, NPW, BK>::PSY;
        const int ny = (na * H + 63) / 64 * 64 * SGY;
        for (int t = threadIdx.x; t < ny; t += blockDim.x) {
            int w, sg;
            t3_slot<SGY>(t, w, sg);
            if (w >= na * H) continue;
            const int k3 = w % H, b0 = sg * PSY;
            auto ystep = [&](auto par) {
                constexpr int B0P = decltype(par)::value;
                Cx<R> acc[PSY];
#pragma unroll
                for (int b = 0; b < PSY; ++b) acc[b] = {R(0), R(0)};
                for (int k2 = 0; k2 < NPW; ++k2) {
                    const Cx<R> x = U1[w * NPW + k2];
#pragma unroll
                    for (int b = 0; b < PSY; ++b) {
                        if (b0 + b < P) {
                            if ((B0P + b) % 2 == 0) {
                                const R m = V[k2 + NPW * (b0 + b)].re;
                                acc[b].re = fma(m, x.re, acc[b].re);
                                acc[b].im = fma(m, x.im, acc[b].im);
                            } else {
                                const R m = V[k2 + NPW * (b0 + b)].im;
                                acc[b].re = fma(-m, x.im, acc[b].re);
                                acc[b].im = fma(m, x.re, acc[b].im);
                            }
                        }
                    }
                }
                const R hw = (k3 >= half) ? R(0.5) : R(1);
#pragma unroll
                for (int b = 0; b < PSY; ++b)
                    if (b0 + b < P) U2[w * TS2 + b0 + b] = {acc[b].re * hw, acc[b].im * hw};
            };
            if (b0 & 1) ystep(Odd{});
            else ystep(Even{});
        }
        __syncthreads();
        // z: down[c][b][a0 + a] (+)= 2 Re sum_k3 V(k3, c) U2[a][k3][b]; slot = (column w = b na + a,
        // segment of c, uniform across a wavefront)
        constexpr int SGZ = Pw2pCfg<R, P, NPW, BK>::SGZ, PSZ = Pw2pCfg<R, P, NPW, BK>::PSZ;
        const int nz = (na * P + 63) / 64 * 64 * SGZ;
        for (int t = threadIdx.x; t < nz; t += blockDim.x) {
            int w, sg;
            t3_slot<SGZ>(t, w, sg);
            if (w >= na * P) continue;
            const int b = w / na, a = w - b * na, c0 = sg * PSZ;
            auto zstep = [&](auto par) {
                constexpr int C0P = decltype(par)::value;
                R acc[PSZ];
#pragma unroll
                for (int c = 0; c < PSZ; ++c) acc[c] = R(0);
                for (int k3 = 0; k3 < H; ++k3) {
                    const Cx<R> x = U2[(a * H + k3) * TS2 + b];
#pragma unroll
                    for (int c = 0; c < PSZ; ++c) {
                        if (c0 + c < P) {
                            if ((C0P + c) % 2 == 0) acc[c] = fma(V[k3 + NPW * (c0 + c)].re, x.re, acc[c]);
                            else acc[c] = fma(-V[k3 + NPW * (c0 + c)].im, x.im, acc[c]);
                        }
                    }
                }
                R *dst = down + (size_t)b * P + a0 + a;
#pragma unroll
                for (int c = 0; c < PSZ; ++c)
                    if (c0 + c < P)
                        dst[(size_t)(c0 + c) * PP] = (zero ? R(0) : dst[(size_t)(c0 + c) * PP]) + R(2) * acc[c];
            };
            if (c0 & 1) zstep(Odd{});
            else zstep(Even{});
        }
        __syncthreads();
    }
}

// pot_sorted[t] += sum C(a,b,c) T_a T_b T_c at the targets of one leaf per workgroup.  The
// coefficient planes C[c][.][.] are staged through LDS one at a time and read as broadcasts;
// each thread evaluates EV_TPT targets, T_a sits in registers and T_b, T_c follow the
// Chebyshev recurrence along their loops.
constexpr int EV_THREADS = 128, EV_TPT = 2;

template <class R, int P>
__global__ __launch_bounds__(EV_THREADS) void k_eval_targets(const R *__restrict__ down_base, const int *__restrict__ leaf_down,
                                                              const R *__restrict__ leaf_geom, const int *__restrict__ leaf_trg,
                                                              const P4<R> *__restrict__ trg, R *__restrict__ pot_sorted,
                                                              long long vs_in = 0, long long vs_out = 0) {
    chain_priority();
    down_base += (long long)blockIdx.z * vs_in;  // vector blockIdx.z of a many-vector launch
    pot_sorted += (long long)blockIdx.z * vs_out;
    constexpr int PP = P * P;
    __shared__ R plane[PP];
    const int lf = blockIdx.x;
    const R *C = down_base + (size_t)leaf_down[lf] * P * PP;
    const R cx = leaf_geom[4 * lf], cy = leaf_geom[4 * lf + 1], cz = leaf_geom[4 * lf + 2], sc = leaf_geom[4 * lf + 3];
    const int beg = leaf_trg[2 * lf], cnt = leaf_trg[2 * lf + 1];
    for (int t0 = 0; t0 < cnt; t0 += EV_THREADS * EV_TPT) {
        R tx[EV_TPT][P], y[EV_TPT], z[EV_TPT], s[EV_TPT], tcm1[EV_TPT], tc[EV_TPT];
#pragma unroll
        for (int i = 0; i < EV_TPT; ++i) {
            const int t = t0 + (int)threadIdx.x + EV_THREADS * i;
            const P4<R> pt = t < cnt ? trg[beg + t] : P4<R>{cx, cy, cz, R(0)};
            const R x = sc * (pt.x - cx);
            y[i] = sc * (pt.y - cy);
            z[i] = sc * (pt.z - cz);
            tx[i][0] = R(1);
            if (P > 1) tx[i][1] = x;
#pragma unroll
            for (int k = 2; k < P; ++k) tx[i][k] = R(2) * x * tx[i][k - 1] - tx[i][k - 2];
            s[i] = R(0);
            tcm1[i] = R(1);
            tc[i] = R(1);
        }
        for (int c = 0; c < P; ++c) {
            __syncthreads();
            for (int e = threadIdx.x; e < PP; e += EV_THREADS) plane[e] = C[(size_t)c * PP + e];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < EV_TPT; ++i) {
                if (c == 1) {
                    tc[i] = z[i];
                } else if (c > 1) {
                    const R tn = R(2) * z[i] * tc[i] - tcm1[i];
                    tcm1[i] = tc[i];
                    tc[i] = tn;
                }
            }
            R sb[EV_TPT], tbm1[EV_TPT], tb[EV_TPT];
#pragma unroll
            for (int i = 0; i < EV_TPT; ++i) { sb[i] = R(0); tbm1[i] = R(1); tb[i] = R(1); }
            for (int b = 0; b < P; ++b) {
#pragma unroll
                for (int i = 0; i < EV_TPT; ++i) {
                    if (b == 1) {
                        tb[i] = y[i];
                    } else if (b > 1) {
                        const R tn = R(2) * y[i] * tb[i] - tbm1[i];
                        tbm1[i] = tb[i];
                        tb[i] = tn;
                    }
                }
                R sa[EV_TPT];
#pragma unroll
                for (int i = 0; i < EV_TPT; ++i) sa[i] = R(0);
#pragma unroll
                for (int a = 0; a < P; ++a) {
                    const R cv = plane[b * P + a];
#pragma unroll
                    for (int i = 0; i < EV_TPT; ++i) sa[i] = fma(cv, tx[i][a], sa[i]);
                }
#pragma unroll
                for (int i = 0; i < EV_TPT; ++i) sb[i] = fma(sa[i], tb[i], sb[i]);
            }
#pragma unroll
            for (int i = 0; i < EV_TPT; ++i) s[i] = fma(sb[i], tc[i], s[i]);
        }
#pragma unroll
        for (int i = 0; i < EV_TPT; ++i) {
            const int t = t0 + (int)threadIdx.x + EV_THREADS * i;
            if (t < cnt) pot_sorted[beg + t] += s[i];
        }
    }
}

// The same evaluation on the FP64 matrix cores, for p >= 28.  Per leaf and stage of
// targets
//   R[t, (c, b)] = sum_a T_a(x_t) C[(c, b), a]          (GEMM, a padded to 16s)
//   pot_t += sum_(c,b) R[t, (c, b)] T_b(y_t) T_c(z_t)
// with v_mfma_f64_16x16x4f64 (operand layout as in k_charge2proxy_mfma: A[l % 16][l / 16],
// B[l / 16][l % 16], D[l / 16 + 4 r][l % 16]).  Wave w takes the (c, b) tiles w, w + NW, ...,
// whose B operands stay in registers for the whole leaf; each stage runs the TT target tiles
// as independent accumulators.  Lane partials are reduced over each 16-lane row, then over
// the waves in a fixed order.  (Reloading the B operands per stage of 64 targets read each
// leaf's proxy 4 times: 1846 us at N=1e6, 9 digits.)  At p = 38 a wave's B operands would not
// fit its registers: the leaf's (c, b) tiles are taken in RH = 2 halves, each half a pass over
// the leaf's targets that adds its part of the sum to the potentials.
template <int P> struct EvMfmaCfg {
    static constexpr int RT = (P * P + 15) / 16;
    static constexpr int NW = 8;
    static constexpr int RH = P >= 38 ? 2 : 1;              // passes over the leaf's targets
    static constexpr int MT = (RT + NW * RH - 1) / (NW * RH);  // (c, b) tiles per wave and pass
    static constexpr int TT = 2;                   // target tiles per stage
    static constexpr int NT = 16 * TT;             // targets per stage
    static constexpr int PA = (P + 15) / 16 * 16;
    // LDS row strides, in elements: the A reads (16 targets x 2 a per half wavefront) at a
    // stride of 34 and the T_b / T_c reads (2 targets x 16 (c, b) columns) at 48 fall on
    // distinct banks; the stride of 32 put the 16 targets of an A read on the same banks
    // (measured at N=1e6, 9 digits: 3.3x fewer LDS bank-conflict cycles, 20% fewer busy
    // cycles; an XOR-swizzled layout removed more conflicts but its address arithmetic cost
    // more than it saved)
    static constexpr int LX = PA + 2, LY = 48;
    static_assert(P <= LY, "order above the table stride");
};
// One workgroup per leaf.  Wave w owns the (c, b) tiles w, w + NW, ...: their B operands
// (the leaf's proxy rows) are loaded into registers once per leaf, and the targets pass by
// in stages of NT, so the proxy is read from memory once per leaf.
template <int P>
__global__ __launch_bounds__(64 * EvMfmaCfg<P>::NW) void k_eval_targets_mfma(
    const double *__restrict__ down_base, const int *__restrict__ leaf_down, const double *__restrict__ leaf_geom,
    const int *__restrict__ leaf_trg, const P4<double> *__restrict__ trg, double *__restrict__ pot_sorted) {
    chain_priority();
    using C = EvMfmaCfg<P>;
    constexpr int RT = C::RT, NW = C::NW, RH = C::RH, MT = C::MT, TT = C::TT, NT = C::NT, PA = C::PA, PP = P * P;
    static_assert(P <= PA, "order above the padded a count");
    constexpr int LX = C::LX, LY = C::LY;
    __shared__ double Tx[NT * LX], Ty[NT * LY], Tz[NT * LY], red[NW * NT];
    const int lf = blockIdx.x, lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int kk = lane >> 4, jj = lane & 15;
    const double *Cm = down_base + (size_t)leaf_down[lf] * P * PP;
    const double cx = leaf_geom[4 * lf], cy = leaf_geom[4 * lf + 1], cz = leaf_geom[4 * lf + 2], sc = leaf_geom[4 * lf + 3];
    const int beg = leaf_trg[2 * lf], cnt = leaf_trg[2 * lf + 1];
    for (int h = 0; h < RH; ++h) {
        // B operands of tile rt = w + NW (h MT + m): lane (kk, jj) supplies C[rt 16 + jj][4 ks + kk]
        double B[MT][PA / 4];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int row = (w + NW * (h * MT + m)) * 16 + jj;
#pragma unroll
            for (int ks = 0; ks < PA / 4; ++ks) {
                const int a = 4 * ks + kk;
                B[m][ks] = (row < PP && a < P) ? Cm[(size_t)row * P + a] : 0.0;
            }
        }
        for (int t0 = 0; t0 < cnt; t0 += NT) {
            __syncthreads();
            for (int task = threadIdx.x; task < 3 * NT; task += blockDim.x) {
                const int t = task % NT, d = task / NT;
                double *T = d == 0 ? Tx + t * LX : (d == 1 ? Ty + t * LY : Tz + t * LY);
                const int len = d == 0 ? PA : P;
                if (t0 + t < cnt) {
                    const P4<double> pt = trg[beg + t0 + t];
                    const double x = sc * (d == 0 ? pt.x - cx : (d == 1 ? pt.y - cy : pt.z - cz));
                    double ta = 1.0, tb = x;
                    T[0] = 1.0;
                    T[1] = x;
                    for (int k = 2; k < P; ++k) {
                        const double tn = 2.0 * x * tb - ta;
                        T[k] = tn;
                        ta = tb;
                        tb = tn;
                    }
                    for (int k = P; k < len; ++k) T[k] = 0.0;
                } else {
                    for (int k = 0; k < len; ++k) T[k] = 0.0;
                }
            }
            __syncthreads();
            double A[TT][PA / 4];
#pragma unroll
            for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                for (int ks = 0; ks < PA / 4; ++ks) A[tt][ks] = Tx[(tt * 16 + jj) * LX + 4 * ks + kk];
            double part[TT][4];
#pragma unroll
            for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                for (int e = 0; e < 4; ++e) part[tt][e] = 0.0;
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const int rt = w + NW * (h * MT + m);
                if (rt >= RT) break;  // wave-uniform
                const int row = rt * 16 + jj;  // (c, b) of the B column this lane supplies
                c2p_d4 D[TT];
#pragma unroll
                for (int tt = 0; tt < TT; ++tt) D[tt] = c2p_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < PA / 4; ++ks)
#pragma unroll
                    for (int tt = 0; tt < TT; ++tt)
                        D[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[tt][ks], B[m][ks], D[tt], 0, 0, 0);
                // D[tt][e] = R[target tt 16 + kk + 4 e, (c, b) = rt 16 + jj]
                const int cb = row < PP ? row : 0, c = cb / P, b = cb % P;
#pragma unroll
                for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int t = tt * 16 + kk + 4 * e;
                        part[tt][e] = fma(D[tt][e], Ty[t * LY + b] * Tz[t * LY + c], part[tt][e]);
                    }
            }
            // sum over the 16 lanes of each row (the (c, b) columns), then over the waves
#pragma unroll
            for (int tt = 0; tt < TT; ++tt)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    double v = part[tt][e];
#pragma unroll
                    for (int m = 8; m >= 1; m >>= 1) v += __shfl_xor(v, m, 16);
                    if (jj == 0) red[w * NT + tt * 16 + kk + 4 * e] = v;
                }
            __syncthreads();
            for (int t = threadIdx.x; t < NT; t += blockDim.x) {
                if (t0 + t < cnt) {
                    double v = 0;
                    for (int q = 0; q < NW; ++q) v += red[q * NT + t];
                    pot_sorted[beg + t0 + t] += v;
                }
            }
        }
    }
}

// ------------------------------------------------------------------ near field
// Residual u = P((r - bs/2) 2/bs) / r for 1e-30 < r^2 < bs^2, summed over the list1
// sources of every target.  For fewer than 6 digits P is applied in r^2 with an even/odd
// split (monomial coefficients per bs, prepared on the host), as laplace_3d_poly_all_pairs
// does.
//
// Work item = one wavefront = up to 64 targets of one leaf (consecutive in the leaf's
// Morton order, so the item is spatially compact).  The geometry is fixed for the life of a
// tree, so which sources can reach which wavefront is decided once, at construction
// (k_near_steps): a step is one source of one list1 entry for which at least one lane of
// the wavefront is in support.  The evaluation walks only the steps, in the order the
// construction found them, and applies the per-lane support test to each; sources are read
// with wave-uniform addresses.
//
// Coordinates are relative to each point's leaf centre; geom[e] = (centre of the target
// leaf - centre of the source leaf, bs), so r = |(t + off) - s|.  Each list entry is
// summed in R and the per-target total in FP64.
constexpr int NF_CHUNK = 4;
constexpr int NF_SPLIT_STRIDE = 16;  // coefficient stride of the even/odd table
// steps whose sources are loaded together (scalar loads: U points in SGPRs); the step list
// is padded by NF_STEP_PAD valid indices so a group may read past the end of its run
template <class R> constexpr int nf_unroll() { return sizeof(R) == 4 ? 8 : 4; }
constexpr int NF_STEP_PAD = 16;

struct NfEntry { int sbeg, scnt, cbase, cfidx; };

// Step lists, once per tree.  Sources of every leaf are cut into chunks of NF_CHUNK
// consecutive (Morton-ordered) points with bounding boxes cbox; for each list1 entry the 64
// lanes test 64 chunks at once against the item's box wbox (shifted into the source leaf's
// frame), evaluated with the same rounded operations as the per-pair r^2 so that it never
// drops a pair the per-pair test keeps; the sources of kept chunks are then tested per lane.
// PASS 0 writes the step count of the k-th entry of item i to cnt[ie_base[i] + k] and adds
// the in-support pairs to *pairs; PASS 1 writes the steps' source indices from
// ie_ptr[ie_base[i] + k] on.
template <class R, int PASS>
__global__ __launch_bounds__(256) void k_near_steps(const P4<R> *__restrict__ src, const P4<R> *__restrict__ trg,
                                                     const int4 *__restrict__ items, const P4<R> *__restrict__ wbox,
                                                     const int *__restrict__ l1_ptr, const NfEntry *__restrict__ l1,
                                                     const P4<R> *__restrict__ geom, const P4<R> *__restrict__ cbox,
                                                     const int *__restrict__ ie_base, int *__restrict__ cnt,
                                                     const long long *__restrict__ ie_ptr, int *__restrict__ step_src,
                                                     unsigned long long *__restrict__ pairs, int n_items) {
    const int item = blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= n_items) return;  // whole wavefront
    const int lane = threadIdx.x & 63;
    const int4 it = items[item];
    const bool active = lane < it.y;
    const P4<R> me = active ? trg[it.x + lane] : P4<R>{R(1e18), R(1e18), R(1e18), R(0)};
    const P4<R> blo = wbox[2 * item], bhi = wbox[2 * item + 1];
    const int eb = l1_ptr[it.z], base = ie_base[item];
    unsigned long long np = 0;
    for (int e = eb; e < l1_ptr[it.z + 1]; ++e) {
        const NfEntry E = l1[e];
        const P4<R> gm = geom[e];
        const R bs = gm.w, d2max = bs * bs;
        const R tx = me.x + gm.x, ty = me.y + gm.y, tz = me.z + gm.z;
        const R lx = blo.x + gm.x, ly = blo.y + gm.y, lz = blo.z + gm.z;
        const R hx = bhi.x + gm.x, hy = bhi.y + gm.y, hz = bhi.z + gm.z;
        const int nch = (E.scnt + NF_CHUNK - 1) / NF_CHUNK;
        long long out = PASS == 1 ? ie_ptr[base + e - eb] : 0;
        int steps = 0;
        for (int c0 = 0; c0 < nch; c0 += 64) {
            bool keep = false;
            if (c0 + lane < nch) {
                const P4<R> cl = cbox[2 * (E.cbase + c0 + lane)], ch = cbox[2 * (E.cbase + c0 + lane) + 1];
                const R dx = fmax(fmax(cl.x - hx, lx - ch.x), R(0));
                const R dy = fmax(fmax(cl.y - hy, ly - ch.y), R(0));
                const R dz = fmax(fmax(cl.z - hz, lz - ch.z), R(0));
                keep = fma(dx, dx, fma(dy, dy, dz * dz)) < d2max;
            }
            unsigned long long mask = __ballot(keep);
            while (mask) {
                const int k = __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                const int s0 = E.sbeg + (c0 + k) * NF_CHUNK;
                const int ns = min(NF_CHUNK, E.sbeg + E.scnt - s0);
                for (int j = 0; j < ns; ++j) {
                    const P4<R> q = src[s0 + j];
                    const R dx = tx - q.x, dy = ty - q.y, dz = tz - q.z;
                    const R r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    const unsigned long long in = __ballot(r2 < d2max && r2 > R(1e-30));
                    if (in) {
                        if (PASS == 1 && lane == 0) step_src[out] = s0 + j;
                        out += 1;
                        steps += 1;
                        np += __popcll(in);
                    }
                }
            }
        }
        if (PASS == 0 && lane == 0) cnt[base + e - eb] = steps;
    }
    if (PASS == 0 && lane == 0) atomicAdd(pairs, np);
}

// Residual evaluation over the step lists (steps of the k-th entry of item i:
// step_src[ie_ptr[ie_base[i] + k] .. ie_ptr[ie_base[i] + k + 1])).
template <class R, int NC, bool SPLIT>
__global__ __launch_bounds__(1024) void k_near_field(const P4<R> *__restrict__ src, const P4<R> *__restrict__ trg,
                                                     const int4 *__restrict__ items, const int *__restrict__ l1_ptr,
                                                     const NfEntry *__restrict__ l1, const P4<R> *__restrict__ geom,
                                                     const int *__restrict__ ie_base, const long long *__restrict__ ie_ptr,
                                                     const int *__restrict__ step_src, const R *__restrict__ coeffs,
                                                     R *__restrict__ pot_sorted, int item0, int item1) {
    // wave index made wave-uniform explicitly, so everything derived from it lives in SGPRs
    const int item = item0 + xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                     __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= item1) return;  // whole wavefront
    const int lane = threadIdx.x & 63;
    const int4 it = items[item];  // (first target, count, target-leaf list index, -)
    const bool active = lane < it.y;
    const P4<R> me = active ? trg[it.x + lane] : P4<R>{R(1e18), R(1e18), R(1e18), R(0)};
    R c[NC];
    if (!SPLIT) {
#pragma unroll
        for (int i = 0; i < NC; ++i) c[i] = coeffs[i];
    }
    // FP64: leading coefficient in a VGPR, so each Horner FMA reads at most one SGPR (no copy
    // per step), and the charge product inside the support branch.  Measured at N=1e6: the
    // FP64 near field 3.60 -> 3.43 ms and evaluate() 4.82 -> 4.67 ms at 6 digits; in FP32 the
    // same change slows evaluate() (3.02 -> 3.10 ms) although the kernel alone gains 1%.
    constexpr bool trim = sizeof(R) == 8;
    const R ctop = SPLIT ? R(0) : (trim ? c[NC - 1] + (R)vgpr_zero() : c[NC - 1]);
    const int eb = l1_ptr[it.z];
    const long long *iep = ie_ptr + ie_base[item] - eb;
    double acc = 0;
    for (int e = eb; e < l1_ptr[it.z + 1]; ++e) {
        const long long sb = iep[e], se = iep[e + 1];
        if (sb == se) continue;
        const P4<R> gm = geom[e];
        const R bs = gm.w, d2max = bs * bs, rsc = R(2) / bs, cen = -bs / R(2);
        const R tx = me.x + gm.x, ty = me.y + gm.y, tz = me.z + gm.z;
        if (SPLIT) {
            const int cf = l1[e].cfidx;
#pragma unroll
            for (int i = 0; i < NC; ++i) c[i] = coeffs[cf * NF_SPLIT_STRIDE + i];
        }
        constexpr int U = nf_unroll<R>();
        const int *sp = step_src + sb;
        const int cnt = (int)(se - sb);
        R ea = 0;
        for (int g = 0; g < cnt; g += U) {
            // all U indices and then all U points are requested before the first use
            int id[U];
#pragma unroll
            for (int j = 0; j < U; ++j) id[j] = sp[g + j];
            P4<R> qs[U];
#pragma unroll
            for (int j = 0; j < U; ++j) qs[j] = src[id[j]];
            // distances of all U points first and every point used unconditionally, so that
            // all point loads are in flight before the first wait; steps past the run get
            // r^2 = bs^2, outside the support
            R r2s[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const R dx = tx - qs[j].x, dy = ty - qs[j].y, dz = tz - qs[j].z;
                const R r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                r2s[j] = g + j < cnt ? r2 : d2max;
            }
            // (empty statements that keep the U distances and charges ahead of the first step:
            // the compiler would otherwise sink them, and with them the point loads, into
            // later steps)
#pragma unroll
            for (int j = 0; j < U; ++j) {
                if (trim) asm volatile("" ::"v"(r2s[j]), "s"(qs[j].w));
                else asm volatile("" ::"v"(r2s[j]));
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const R r2 = r2s[j];
                R u = 0;
                if (r2 < d2max && r2 > R(1e-30)) {
                    const R rinv = rsqrt_full<R>(r2);
                    if (SPLIT) {
                        constexpr int SE = (NC - 1) & ~1, SO = ((NC - 2) & ~1) + 1;
                        R Ev = c[SE];
#pragma unroll
                        for (int i = SE - 2; i >= 0; i -= 2) Ev = fma(Ev, r2, c[i]);
                        R Od = c[SO];
#pragma unroll
                        for (int i = SO - 2; i >= 0; i -= 2) Od = fma(Od, r2, c[i]);
                        u = fma(Ev, rinv, Od);
                    } else {
                        const R x = fma(r2, rinv, cen) * rsc;
                        R p = ctop;
#pragma unroll
                        for (int i = NC - 2; i >= 0; --i) p = fma(p, x, c[i]);
                        u = p * rinv;
                    }
                    if (trim) ea = fma(u, qs[j].w, ea);
                }
                if (!trim) ea = fma(u, qs[j].w, ea);
            }
        }
        acc += (double)ea;
    }
    if (active) pot_sorted[it.x + lane] += (R)acc;
}

// ---------------------------------------------------------- near field in quad rows (FP32)
// The wavefront's 64 targets form 16 rows of 4 (lane l = 4 g + j: row g, target it.x + l).
// Every row walks a step list of its own: the sources (over all list1 entries of the target
// leaf, in entry order, and in walk order inside an entry) that reach at least one of its 4
// targets.  The 4 targets of a row are much closer together than the 64 of the wavefront, so
// fewer of a step's lanes fall outside the support: at N=1e6 (uniform) about 54% of the lanes
// of a step are in support against 34% when the whole wavefront shares one list.
//
// Step lists (built once per tree, k_near_rows_build): per item, K steps per row (K = the
// longest row's count rounded up to 4; shorter rows are padded with a far sentinel point),
// stored as source indices in blocks of 4 steps: block b holds at element 4 g + j the source
// of step 4 b + j of row g, so lane l loads element l of each block (one coalesced 256-byte
// read) and the 4 points of a row's block reach the row's lanes by DPP quad broadcasts.
// Per (item, row) the non-empty entries are listed as slots (entry ordinal, end step), ending
// with a terminator slot (nothing in support).
//
// For each target the in-support sources appear in its row's list in the same order as in
// the wavefront's list of k_near_field, each entry's sum is formed the same way and added to
// the FP64 total in entry order (entries without steps add nothing, there as 0.0 here), so
// the potentials equal k_near_field's bit for bit.
constexpr int NF_ROWS = 16;          // rows of 4 targets per wavefront
constexpr int NF_ROW_PAD = 2 * 64;   // sentinel indices past the last block (loads run one block ahead)
// Bit 31 of a step index flags a row's entry switch at that step (the first step of each of
// the row's entries but its first).  Readers form byte offsets as (unsigned)index * size in
// 32 bits, which drops the flag (source indices stay below 2^26).
constexpr int NF_STEP_ENTRY = (int)0x80000000u;

template <int S>
__device__ __forceinline__ float quad_bcast(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), S * 0x55, 0xF, 0xF, true));
}

// ea += q(lane 4 g + S) * u in one DPP-broadcast FMA (the same fused multiply-add as
// fmaf(quad_bcast<S>(q), u, ea)), led by s_nop 1: the two wait states a DPP read needs after a
// VALU write of its source (the compiler does not check hazards inside inline asm).
#define PDMK_FMAC_QUAD1(SEL) \
    asm("s_nop 1\n"          \
        "v_fmac_f32_dpp %0, %1, %2 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1" : "+v"(ea) : "v"(q), "v"(u))
template <int S>
__device__ __forceinline__ void fmac_quad1(float &ea, float q, float u) {
    if constexpr (S == 0) PDMK_FMAC_QUAD1("0,0,0,0");
    if constexpr (S == 1) PDMK_FMAC_QUAD1("1,1,1,1");
    if constexpr (S == 2) PDMK_FMAC_QUAD1("2,2,2,2");
    if constexpr (S == 3) PDMK_FMAC_QUAD1("3,3,3,3");
}
#undef PDMK_FMAC_QUAD1

// *(T *)((char *)base + off) with a 32-bit unsigned byte offset: a wave-uniform base and a
// per-lane offset that the address computation needs no 64-bit arithmetic for
template <class T, class B>
__device__ __forceinline__ T nf_load_u32(const B *base, unsigned off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// PASS 0: steps of row g in the k-th entry of item i -> cnt[(ie_base[i] + k) * 16 + g], and the
// in-support pairs added to *pairs.  PASS 1: the source indices into the row blocks from
// steps + wave_base[i] (the item's K, as built on the host, in items[i].w).
template <class R, int PASS>
__global__ __launch_bounds__(256) void k_near_rows_build(
    const P4<R> *__restrict__ src, const P4<R> *__restrict__ trg, const int4 *__restrict__ items,
    const P4<R> *__restrict__ wbox, const int *__restrict__ l1_ptr, const NfEntry *__restrict__ l1,
    const P4<R> *__restrict__ geom, const P4<R> *__restrict__ cbox, const int *__restrict__ ie_base,
    int *__restrict__ cnt, const long long *__restrict__ wave_base, int *__restrict__ steps,
    unsigned long long *__restrict__ pairs, int n_items) {
    const int item = blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= n_items) return;  // whole wavefront
    const int lane = threadIdx.x & 63;
    const int4 it = items[item];
    const bool active = lane < it.y;
    const P4<R> me = active ? trg[it.x + lane] : P4<R>{R(1e18), R(1e18), R(1e18), R(0)};
    const P4<R> blo = wbox[2 * item], bhi = wbox[2 * item + 1];
    const int eb = l1_ptr[it.z], base = ie_base[item];
    int *out = PASS == 1 ? steps + wave_base[item] + lane : nullptr;
    unsigned long long np = 0;
    int k_row = 0;  // steps of this lane's row so far (the same in the row's 4 lanes)
    for (int e = eb; e < l1_ptr[it.z + 1]; ++e) {
        const NfEntry E = l1[e];
        const P4<R> gm = geom[e];
        const R bs = gm.w, d2max = bs * bs;
        const R tx = me.x + gm.x, ty = me.y + gm.y, tz = me.z + gm.z;
        const R lx = blo.x + gm.x, ly = blo.y + gm.y, lz = blo.z + gm.z;
        const R hx = bhi.x + gm.x, hy = bhi.y + gm.y, hz = bhi.z + gm.z;
        const int nch = (E.scnt + NF_CHUNK - 1) / NF_CHUNK;
        int c_row = 0;
        for (int c0 = 0; c0 < nch; c0 += 64) {
            bool keep = false;
            if (c0 + lane < nch) {
                const P4<R> cl = cbox[2 * (E.cbase + c0 + lane)], ch = cbox[2 * (E.cbase + c0 + lane) + 1];
                const R dx = fmax(fmax(cl.x - hx, lx - ch.x), R(0));
                const R dy = fmax(fmax(cl.y - hy, ly - ch.y), R(0));
                const R dz = fmax(fmax(cl.z - hz, lz - ch.z), R(0));
                keep = fma(dx, dx, fma(dy, dy, dz * dz)) < d2max;
            }
            unsigned long long mask = __ballot(keep);
            while (mask) {
                const int k = __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                const int s0 = E.sbeg + (c0 + k) * NF_CHUNK;
                const int ns = min(NF_CHUNK, E.sbeg + E.scnt - s0);
                for (int j = 0; j < ns; ++j) {
                    const P4<R> q = src[s0 + j];
                    const R dx = tx - q.x, dy = ty - q.y, dz = tz - q.z;
                    const R r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    const unsigned long long in = __ballot(r2 < d2max && r2 > R(1e-30));
                    np += __popcll(in);
                    if ((in >> (lane & ~3)) & 0xFull) {
                        // bit 31: the row's step starts an entry after its first
                        if (PASS == 1 && (k_row & 3) == (lane & 3))
                            out[(size_t)(k_row >> 2) * 64] = (s0 + j) | (c_row == 0 && k_row > 0 ? NF_STEP_ENTRY : 0);
                        ++k_row;
                        ++c_row;
                    }
                }
            }
        }
        if (PASS == 0 && (lane & 3) == 0) cnt[(size_t)(base + e - eb) * NF_ROWS + (lane >> 2)] = c_row;
    }
    if (PASS == 0 && lane == 0) atomicAdd(pairs, np);
}

// Residual evaluation over the row step lists (FP32, NC coefficients in the x form).  Row g
// of item i lists its non-empty entries in slots rslot[(rs_base[i] + s) * 16 + g] =
// (end step << 8) | ordinal of the entry in the target leaf's list1, followed by a terminator
// (ordinal = number of entries: bs = 0, nothing in support).  Each wavefront copies its leaf's
// entry geometry (offset, bs, 2/bs, -bs/2, bs^2) and its rows' slots into LDS (ne_max + 1
// entries per wavefront), so an entry switch reads LDS only.  In LDS a slot is decoded ahead:
// (end step << 16) | byte offset of the entry's geometry from the LDS base (end steps below
// 2^16, and the workgroup's LDS below 64 KiB, both checked on the host; the terminator's end
// is clamped to 2^16 - 1, past every step).
template <int NC>
__global__ __launch_bounds__(1024) void k_near_rows(const P4<float> *__restrict__ src, const P4<float> *__restrict__ trg,
                                                   const int4 *__restrict__ items, const int *__restrict__ l1_ptr,
                                                   const P4<float> *__restrict__ geom, const int *__restrict__ rs_base,
                                                   const int *__restrict__ rslot, const long long *__restrict__ wave_base,
                                                   const int *__restrict__ steps, const float *__restrict__ coeffs,
                                                   float *__restrict__ pot_sorted, int item0, int item1, int ne_max) {
    extern __shared__ float4 nf_lds[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = item0 + xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + w;
    if (item >= item1) return;  // whole wavefront (no workgroup barrier below)
    const int lane = threadIdx.x & 63, g = lane >> 2;
    const int4 it = items[item];  // (first target, count, target-leaf list index, K)
    // this wavefront's LDS: ne_max + 1 entries x (offset+bs, consts), then (ne_max + 1) x 16 slots
    float4 *geo = nf_lds + (size_t)w * (ne_max + 1) * 6;
    int *slots = reinterpret_cast<int *>(geo + 2 * (ne_max + 1));
    const int eb = l1_ptr[it.z], ne = l1_ptr[it.z + 1] - eb;
    for (int e = lane; e <= ne; e += 64) {
        const P4<float> gm = e < ne ? geom[eb + e] : P4<float>{0.0f, 0.0f, 0.0f, 0.0f};
        const float bs = gm.w;
        geo[2 * e] = make_float4(gm.x, gm.y, gm.z, bs * bs);
        geo[2 * e + 1] = make_float4(e < ne ? 2.0f / bs : 0.0f, -bs / 2.0f, 0.0f, 0.0f);
    }
    const int *rs = rslot + (size_t)rs_base[item] * NF_ROWS;
    const unsigned geo_off = (unsigned)((const char *)geo - (const char *)nf_lds);
    for (int k = lane; k < (ne + 1) * NF_ROWS; k += 64) {
        const int v = rs[k];
        slots[k] = (int)(((unsigned)min(v >> 8, 0xffff) << 16) | (geo_off + 32u * (unsigned)(v & 255)));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool active = lane < it.y;
    const P4<float> me = active ? trg[it.x + lane] : P4<float>{1e18f, 1e18f, 1e18f, 0.0f};
    float c[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) c[i] = coeffs[i];
    // the second-highest coefficient in a vector register: the first Horner step then takes
    // the highest from a scalar register and needs no copy per pair (one scalar operand per
    // FMA)
    float c_next;
    asm("v_mov_b32 %0, %1" : "=v"(c_next) : "s"(c[NC - 2]));
    float tx = 0, ty = 0, tz = 0, d2max = 0, rsc = 0, cen = 0;
    int end = 0;
    const int *slot = slots + g;  // the row's next slot
    double acc = 0;
    float ea = 0;
    auto enter = [&]() {  // the row's next non-empty entry (or the terminator)
        acc += (double)ea;
        ea = 0;
        const unsigned v = (unsigned)*slot;
        slot += NF_ROWS;
        end = (int)(v >> 16);
        const float4 *ge = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(nf_lds) + (v & 0xffffu));
        const float4 a = ge[0], b = ge[1];
        tx = me.x + a.x;
        ty = me.y + a.y;
        tz = me.z + a.z;
        d2max = a.w;
        rsc = b.x;
        cen = b.y;
    };
    enter();
    const int K = it.w;
    // step indices: block b at sb + 64 b, element `lane`; the loads below take a wave-uniform
    // base and a 32-bit lane offset (one address instruction per load)
    const int *sb = steps + wave_base[item];
    auto index = [&](int b) { return nf_load_u32<int>(sb + (size_t)b * 64, (unsigned)lane * 4u); };
    auto point = [&](int id) { return nf_load_u32<P4<float>>(src, (unsigned)id * 16u); };
    // m: ballot of the block's entry flags, bit 4 g + s = row g switches entries at step s
    auto block = [&](const P4<float> &pc, unsigned long long m, int k) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            // rows reach their entries' ends at different steps; the wave-wide test is scalar
            // (the empty asm keeps the compiler from folding the per-lane test into it, which
            // would issue that test at every step)
            if ((m >> s) & 0x1111111111111111ull) {
                asm volatile("");
                if (k + s >= end) enter();
            }
            float px, py, pz;
            switch (s) {
            case 0: px = quad_bcast<0>(pc.x); py = quad_bcast<0>(pc.y); pz = quad_bcast<0>(pc.z); break;
            case 1: px = quad_bcast<1>(pc.x); py = quad_bcast<1>(pc.y); pz = quad_bcast<1>(pc.z); break;
            case 2: px = quad_bcast<2>(pc.x); py = quad_bcast<2>(pc.y); pz = quad_bcast<2>(pc.z); break;
            default: px = quad_bcast<3>(pc.x); py = quad_bcast<3>(pc.y); pz = quad_bcast<3>(pc.z); break;
            }
            // (source - target: the negated difference, exactly, so r^2 is the same)
            const float dx = px - tx, dy = py - ty, dz = pz - tz;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            float u = 0;
            if (r2 < d2max && r2 > 1e-30f) {
                const float rinv = rsqrt_full<float>(r2);
                const float x = fmaf(r2, rinv, cen) * rsc;
                float p = fmaf(c[NC - 1], x, c_next);
#pragma unroll
                for (int i = NC - 3; i >= 0; --i) p = fmaf(p, x, c[i]);
                u = p * rinv;
            }
            switch (s) {
            case 0: fmac_quad1<0>(ea, pc.w, u); break;
            case 1: fmac_quad1<1>(ea, pc.w, u); break;
            case 2: fmac_quad1<2>(ea, pc.w, u); break;
            default: fmac_quad1<3>(ea, pc.w, u); break;
            }
        }
    };
    // two blocks per iteration, each block's point loaded one block ahead and its index two
    // blocks ahead (the padding keeps both in bounds); the two point registers alternate, so
    // no copies between blocks
    const int nb = K >> 2;
    int id = index(0);
    unsigned long long ma = __ballot(id < 0), mb;
    P4<float> pa = point(id), pb;
    id = index(1);
    int b = 0;
    for (; b + 2 <= nb; b += 2) {
        mb = __ballot(id < 0);
        pb = point(id);
        id = index(b + 2);
        block(pa, ma, 4 * b);
        ma = __ballot(id < 0);
        pa = point(id);
        id = index(b + 3);
        block(pb, mb, 4 * b + 4);
    }
    if (b < nb) block(pa, ma, 4 * b);
    acc += (double)ea;
    if (active) pot_sorted[it.x + lane] += (float)acc;
}

// ------------------------------------------------ near field of many charge vectors (FP32)
// qv[i * VC + c] = charge c of sorted source i, for the VC vectors of a pass: vector v0 + c
// of Q (vector-major, N per vector; zero past the last vector).
template <int VC>
__global__ void k_gather_many(const double *__restrict__ Q, int n_vec, int v0, int n, const int *__restrict__ src_user,
                              int nsall, float *__restrict__ qv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsall) return;
    const int u = src_user[i];
#pragma unroll
    for (int c = 0; c < VC; ++c) qv[(size_t)i * VC + c] = v0 + c < n_vec ? (float)Q[(size_t)(v0 + c) * n + u] : 0.0f;
}

// ea[c] += q[c](lane 4 g + S) * u for 8 charges of a row's step: DPP-broadcast FMAs in one
// asm statement that starts with s_nop 1, the two wait states a DPP read needs after a VALU
// write of its source (the compiler does not check hazards inside inline asm).  ea and q point
// at 8 consecutive elements of register arrays (constant offsets once inlined).
#define PDMK_FMAC_QUAD(SEL)                                                                                    \
    asm("s_nop 1\n"                                                                                            \
        "v_fmac_f32_dpp %0, %8, %16 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"            \
        "v_fmac_f32_dpp %1, %9, %16 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"            \
        "v_fmac_f32_dpp %2, %10, %16 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"           \
        "v_fmac_f32_dpp %3, %11, %16 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"           \
        "v_fmac_f32_dpp %4, %12, %16 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"           \
        "v_fmac_f32_dpp %5, %13, %16 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"           \
        "v_fmac_f32_dpp %6, %14, %16 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"           \
        "v_fmac_f32_dpp %7, %15, %16 quad_perm:[" SEL "] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"           \
        : "+v"(ea[0]), "+v"(ea[1]), "+v"(ea[2]), "+v"(ea[3]), "+v"(ea[4]), "+v"(ea[5]), "+v"(ea[6]), "+v"(ea[7]) \
        : "v"(q[0]), "v"(q[1]), "v"(q[2]), "v"(q[3]), "v"(q[4]), "v"(q[5]), "v"(q[6]), "v"(q[7]), "v"(u))

template <int S>
__device__ __forceinline__ void fmac_quad8(float *ea, const float *q, float u) {
    if constexpr (S == 0) PDMK_FMAC_QUAD("0,0,0,0");
    if constexpr (S == 1) PDMK_FMAC_QUAD("1,1,1,1");
    if constexpr (S == 2) PDMK_FMAC_QUAD("2,2,2,2");
    if constexpr (S == 3) PDMK_FMAC_QUAD("3,3,3,3");
}
#undef PDMK_FMAC_QUAD

// all VC charges of a step, 8 per asm block
template <int S, int VC>
__device__ __forceinline__ void fmac_quads(float (&ea)[VC], const float (&q)[VC], float u) {
#pragma unroll
    for (int v = 0; v < VC; v += 8) fmac_quad8<S>(ea + v, q + v, u);
}

// k_near_rows for VC charge vectors at once (a multiple of 8): the residual of each (target, source) pair is
// formed once and multiplied into VC sums.  pot_near[c * n + t] (c < vc) receives vector c's
// near field of sorted target t; each vector's sums are those of k_near_rows on that vector,
// bit for bit.
template <int NC, int VC>
__global__ __launch_bounds__(256) void k_near_rows_many(
    const P4<float> *__restrict__ src, const float *__restrict__ qv, const P4<float> *__restrict__ trg,
    const int4 *__restrict__ items, const int *__restrict__ l1_ptr, const P4<float> *__restrict__ geom,
    const int *__restrict__ rs_base, const int *__restrict__ rslot, const long long *__restrict__ wave_base,
    const int *__restrict__ steps, const float *__restrict__ coeffs, float *__restrict__ pot_near, int n, int vc,
    int item0, int item1, int ne_max) {
    static_assert(VC % 8 == 0, "the DPP FMA blocks take 8 vectors each");
    extern __shared__ float4 nf_lds[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = item0 + xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + w;
    if (item >= item1) return;  // whole wavefront (no workgroup barrier below)
    const int lane = threadIdx.x & 63, g = lane >> 2;
    const int4 it = items[item];
    float4 *geo = nf_lds + (size_t)w * (ne_max + 1) * 6;
    int *slots = reinterpret_cast<int *>(geo + 2 * (ne_max + 1));
    const int eb = l1_ptr[it.z], ne = l1_ptr[it.z + 1] - eb;
    for (int e = lane; e <= ne; e += 64) {
        const P4<float> gm = e < ne ? geom[eb + e] : P4<float>{0.0f, 0.0f, 0.0f, 0.0f};
        const float bs = gm.w;
        geo[2 * e] = make_float4(gm.x, gm.y, gm.z, bs * bs);
        geo[2 * e + 1] = make_float4(e < ne ? 2.0f / bs : 0.0f, -bs / 2.0f, 0.0f, 0.0f);
    }
    const int *rs = rslot + (size_t)rs_base[item] * NF_ROWS;
    for (int k = lane; k < (ne + 1) * NF_ROWS; k += 64) slots[k] = rs[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool active = lane < it.y;
    const P4<float> me = active ? trg[it.x + lane] : P4<float>{1e18f, 1e18f, 1e18f, 0.0f};
    float c[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) c[i] = coeffs[i];
    float c_next;  // as in k_near_rows
    asm("v_mov_b32 %0, %1" : "=v"(c_next) : "s"(c[NC - 2]));
    float tx = 0, ty = 0, tz = 0, d2max = 0, rsc = 0, cen = 0;
    int end = 0, ns = 0;
    double acc[VC];
    float ea[VC];
#pragma unroll
    for (int v = 0; v < VC; ++v) {
        acc[v] = 0;
        ea[v] = 0;
    }
    auto enter = [&]() {
#pragma unroll
        for (int v = 0; v < VC; ++v) {
            acc[v] += (double)ea[v];
            ea[v] = 0;
        }
        const int sv = slots[ns * NF_ROWS + g];
        ++ns;
        const int o = sv & 255;
        end = sv >> 8;
        const float4 a = geo[2 * o], b = geo[2 * o + 1];
        tx = me.x + a.x;
        ty = me.y + a.y;
        tz = me.z + a.z;
        d2max = a.w;
        rsc = b.x;
        cen = b.y;
    };
    enter();
    const int K = it.w;
    const int *sp = steps + wave_base[item] + lane;
    int id = sp[0];
    P4<float> pc = nf_load_u32<P4<float>>(src, (unsigned)id * 16u);
    float qc[VC];
#pragma unroll
    for (int v = 0; v < VC; v += 4)
        *reinterpret_cast<float4 *>(&qc[v]) = nf_load_u32<float4>(qv + v, (unsigned)id * (unsigned)(VC * sizeof(float)));
    id = sp[64];
    for (int k = 0; k < K; k += 4) {
        const P4<float> pn = nf_load_u32<P4<float>>(src, (unsigned)id * 16u);
        float qn[VC];
#pragma unroll
        for (int v = 0; v < VC; v += 4)
            *reinterpret_cast<float4 *>(&qn[v]) = nf_load_u32<float4>(qv + v, (unsigned)id * (unsigned)(VC * sizeof(float)));
        id = sp[(size_t)(k + 8) * NF_ROWS];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (k + s >= end) enter();
            float px, py, pz;
            switch (s) {
            case 0: px = quad_bcast<0>(pc.x); py = quad_bcast<0>(pc.y); pz = quad_bcast<0>(pc.z); break;
            case 1: px = quad_bcast<1>(pc.x); py = quad_bcast<1>(pc.y); pz = quad_bcast<1>(pc.z); break;
            case 2: px = quad_bcast<2>(pc.x); py = quad_bcast<2>(pc.y); pz = quad_bcast<2>(pc.z); break;
            default: px = quad_bcast<3>(pc.x); py = quad_bcast<3>(pc.y); pz = quad_bcast<3>(pc.z); break;
            }
            const float dx = px - tx, dy = py - ty, dz = pz - tz;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            float u = 0;
            if (r2 < d2max && r2 > 1e-30f) {
                const float rinv = rsqrt_full<float>(r2);
                const float x = fmaf(r2, rinv, cen) * rsc;
                float p = fmaf(c[NC - 1], x, c_next);
#pragma unroll
                for (int i = NC - 3; i >= 0; --i) p = fmaf(p, x, c[i]);
                u = p * rinv;
            }
            switch (s) {
            case 0: fmac_quads<0>(ea, qc, u); break;
            case 1: fmac_quads<1>(ea, qc, u); break;
            case 2: fmac_quads<2>(ea, qc, u); break;
            default: fmac_quads<3>(ea, qc, u); break;
            }
        }
        pc = pn;
#pragma unroll
        for (int v = 0; v < VC; ++v) qc[v] = qn[v];
    }
    if (active) {
#pragma unroll
        for (int v = 0; v < VC; ++v)
            if (v < vc) pot_near[(size_t)v * n + it.x + lane] = (float)(acc[v] + (double)ea[v]);
    }
}

// The FP64 row kernel (NC coefficients, x form).  The same step lists and slots; a row's 4
// points of a block reach its lanes through LDS (64-bit DPP has no quad broadcast): each lane
// writes the point it loaded, then every lane reads its row's point of the step.  Per
// wavefront LDS: 64 points (2 KiB), then (ne_max + 1) entries x 48 bytes of geometry and
// (ne_max + 1) x 16 slots, decoded in LDS as in the FP32 kernel.  The same loop structure as
// the FP32 kernel (two blocks per iteration, scalar entry-switch tests).  Bit-identical to
// k_near_field<double, NC, false> for the same reason as the FP32 kernel.
__host__ __device__ inline size_t near_rows64_lds_per_wave(int ne_max) {
    return 64 * 2 * sizeof(double2) + (size_t)(ne_max + 1) * 3 * sizeof(double2) + (size_t)(ne_max + 1) * NF_ROWS * 4;
}

template <int NC>
__global__ __launch_bounds__(1024) void k_near_rows64(const P4<double> *__restrict__ src,
                                                      const P4<double> *__restrict__ trg,
                                                      const int4 *__restrict__ items, const int *__restrict__ l1_ptr,
                                                      const P4<double> *__restrict__ geom,
                                                      const int *__restrict__ rs_base, const int *__restrict__ rslot,
                                                      const long long *__restrict__ wave_base,
                                                      const int *__restrict__ steps, const double *__restrict__ coeffs,
                                                      double *__restrict__ pot_sorted, int item0, int item1, int ne_max) {
    extern __shared__ double2 nf_lds64[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = item0 + xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + w;
    if (item >= item1) return;  // whole wavefront (no workgroup barrier below)
    const int lane = threadIdx.x & 63, g = lane >> 2;
    const int4 it = items[item];  // (first target, count, target-leaf list index, K)
    double2 *pts = reinterpret_cast<double2 *>(reinterpret_cast<char *>(nf_lds64) + (size_t)w * near_rows64_lds_per_wave(ne_max));
    double2 *geo = pts + 128;
    int *slots = reinterpret_cast<int *>(geo + 3 * (ne_max + 1));
    const int eb = l1_ptr[it.z], ne = l1_ptr[it.z + 1] - eb;
    for (int e = lane; e <= ne; e += 64) {
        const P4<double> gm = e < ne ? geom[eb + e] : P4<double>{0.0, 0.0, 0.0, 0.0};
        const double bs = gm.w;
        geo[3 * e] = make_double2(gm.x, gm.y);
        geo[3 * e + 1] = make_double2(gm.z, bs * bs);
        geo[3 * e + 2] = make_double2(e < ne ? 2.0 / bs : 0.0, -bs / 2.0);
    }
    const int *rs = rslot + (size_t)rs_base[item] * NF_ROWS;
    const unsigned geo_off = (unsigned)((const char *)geo - (const char *)nf_lds64);
    for (int k = lane; k < (ne + 1) * NF_ROWS; k += 64) {
        const int v = rs[k];
        slots[k] = (int)(((unsigned)min(v >> 8, 0xffff) << 16) | (geo_off + 48u * (unsigned)(v & 255)));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool active = lane < it.y;
    const P4<double> me = active ? trg[it.x + lane] : P4<double>{1e18, 1e18, 1e18, 0.0};
    double c[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) c[i] = coeffs[i];
    const double ctop = c[NC - 1] + (double)vgpr_zero();  // leading coefficient in a VGPR (see k_near_field)
    double tx = 0, ty = 0, tz = 0, d2max = 0, rsc = 0, cen = 0;
    int end = 0;
    const int *slot = slots + g;  // the row's next slot
    double acc = 0, ea = 0;
    auto enter = [&]() {  // the row's next non-empty entry (or the terminator)
        acc += ea;
        ea = 0;
        const unsigned v = (unsigned)*slot;
        slot += NF_ROWS;
        end = (int)(v >> 16);
        const double2 *ge = reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(nf_lds64) + (v & 0xffffu));
        const double2 a = ge[0], b = ge[1], cc = ge[2];
        tx = me.x + a.x;
        ty = me.y + a.y;
        tz = me.z + b.x;
        d2max = b.y;
        rsc = cc.x;
        cen = cc.y;
    };
    enter();
    const int K = it.w;
    const int *sb = steps + wave_base[item];
    auto index = [&](int b) { return nf_load_u32<int>(sb + (size_t)b * 64, (unsigned)lane * 4u); };
    auto point = [&](int id) { return nf_load_u32<P4<double>>(src, (unsigned)id * 32u); };
    // m: ballot of the block's entry flags, bit 4 g + s = row g switches entries at step s
    auto block = [&](const P4<double> &pc, unsigned long long m, int k) {
        pts[2 * lane] = make_double2(pc.x, pc.y);
        pts[2 * lane + 1] = make_double2(pc.z, pc.w);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            // the wave-wide switch test is scalar; the empty asm keeps the per-lane test inside
            if ((m >> s) & 0x1111111111111111ull) {
                asm volatile("");
                if (k + s >= end) enter();
            }
            const double2 a = pts[2 * (4 * g + s)], b = pts[2 * (4 * g + s) + 1];
            // (source - target: the negated difference, exactly, so r^2 is the same)
            const double dx = a.x - tx, dy = a.y - ty, dz = b.x - tz;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (r2 < d2max && r2 > 1e-30) {
                const double rinv = rsqrt_full<double>(r2);
                const double x = fma(r2, rinv, cen) * rsc;
                double p = ctop;
#pragma unroll
                for (int i = NC - 2; i >= 0; --i) p = fma(p, x, c[i]);
                ea = fma(p * rinv, b.y, ea);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // every lane has read the block before it is overwritten
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // two blocks per iteration, points one block ahead, indices two (see k_near_rows)
    const int nb = K >> 2;
    int id = index(0);
    unsigned long long ma = __ballot(id < 0), mb;
    P4<double> pa = point(id), pb;
    id = index(1);
    int b = 0;
    for (; b + 2 <= nb; b += 2) {
        mb = __ballot(id < 0);
        pb = point(id);
        id = index(b + 2);
        block(pa, ma, 4 * b);
        ma = __ballot(id < 0);
        pa = point(id);
        id = index(b + 3);
        block(pb, mb, 4 * b + 4);
    }
    if (b < nb) block(pa, ma, 4 * b);
    acc += ea;
    if (active) pot_sorted[it.x + lane] += acc;
}

template <class R>
__global__ void k_fill(R *__restrict__ p, R v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ far field
// Half-space modes with nonzero weight are grouped in lines of constant (j, k) with i
// contiguous; each line is cut into chunks of up to FF_CHUNK consecutive i.  Chunk c
// covers modes chunk_m0[c] .. + chunk_len[c] - 1 with indices (chunk_i0[c] + r, j, k).
// All far-field arithmetic is FP64.
constexpr int FF_CHUNK = 8;
constexpr int FF_TP = 16;  // points per LDS tile
constexpr int FF_QUAD = 4;  // structure factors per scalar load in the evaluation
constexpr int FF_MANY_S = 4;  // vectors per workgroup of a many-vector structure-factor launch
constexpr int FF_MANY_E = 2;  // vectors per workgroup of a many-vector evaluation

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

// Partial structure factors per workgroup of points:
//   Sp[blk][m] = sum_{points of blk} q e^{-i(i th1 + j th2 + k th3)}.
// Per point tile, e^{-i m th} for every axis index m is tabulated in LDS (FP64 recurrence
// from one sincos per axis), with the tile's charges beside it.  A lane owns one chunk: per
// point it gathers its first mode's factor from the three tables, walks the chunk by
// multiplying with e^{-i th1} and adds q times each factor.  Workgroup z of a many-vector
// launch takes vectors V z .. V z + V - 1 (< nv; vector v's charges at q + v vs_q, its partial
// sums at Sp + v vs_sp): the factors are formed once for its V vectors, and each vector's sums
// are those of a one-vector launch, bit for bit.
template <int V>
__global__ __launch_bounds__(256) void k_ff_structure(const double *__restrict__ theta, const double *__restrict__ q, int n,
                                                       int ms, int mt, int mu, const int *__restrict__ chunk_ijk,
                                                       const int *__restrict__ chunk_m0, const int *__restrict__ chunk_len,
                                                       int n_chunks, int n_modes, double2 *__restrict__ Sp,
                                                       long long vs_q = 0, long long vs_sp = 0, int nv = 1) {
    extern __shared__ double2 ff_lds[];
    const int v0 = blockIdx.z * V;
    double2 *E1 = ff_lds, *E2 = E1 + FF_TP * ms, *E3 = E2 + FF_TP * mt, *Est = E3 + FF_TP * mu;
    double *Qs = reinterpret_cast<double *>(Est + FF_TP);  // [V][FF_TP]
    const int c = blockIdx.y * blockDim.x + threadIdx.x;
    const bool active = c < n_chunks;
    const int ci = active ? chunk_ijk[3 * c] + ms / 2 : 0;
    const int cj = active ? chunk_ijk[3 * c + 1] + mt / 2 : 0;
    const int ck = active ? chunk_ijk[3 * c + 2] + mu / 2 : 0;
    const int len = active ? chunk_len[c] : 0;
    double2 acc[V][FF_CHUNK];
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
        for (int r = 0; r < FF_CHUNK; ++r) acc[v][r] = make_double2(0, 0);
    for (int p0 = blockIdx.x * FF_TP; p0 < n; p0 += gridDim.x * FF_TP) {
        __syncthreads();
        for (int t = threadIdx.x; t < FF_TP * 3; t += blockDim.x) {
            const int pp = t % FF_TP, ax = t / FF_TP, gp = p0 + pp;
            const int cnt = ax == 0 ? ms : (ax == 1 ? mt : mu);
            const int lo = -(cnt / 2);
            double2 *row = (ax == 0 ? E1 : (ax == 1 ? E2 : E3)) + pp * cnt;
            const double th = gp < n ? theta[3 * (size_t)gp + ax] : 0.0;
            double s1, c1, sl, cl;
            sincos(th, &s1, &c1);
            sincos(-th * lo, &sl, &cl);
            const double2 step = make_double2(c1, -s1);  // e^{-i th}
            if (ax == 0) Est[pp] = step;
            double2 cur = make_double2(cl, sl);  // e^{-i lo th}
            for (int m = 0; m < cnt; ++m) {
                row[m] = cur;
                cur = cmul(cur, step);
            }
        }
        for (int t = threadIdx.x; t < FF_TP * V; t += blockDim.x) {
            const int pp = t % FF_TP, v = t / FF_TP, gp = p0 + pp;
            Qs[t] = gp < n && v0 + v < nv ? q[(long long)(v0 + v) * vs_q + gp] : 0.0;
        }
        __syncthreads();
        if (active) {
            const int np = min(FF_TP, n - p0);
            for (int pp = 0; pp < np; ++pp) {
                const double2 st = Est[pp];
                double qq[V];
#pragma unroll
                for (int v = 0; v < V; ++v) qq[v] = Qs[v * FF_TP + pp];
                double2 cur = cmul(cmul(E1[pp * ms + ci], E2[pp * mt + cj]), E3[pp * mu + ck]);
#pragma unroll
                for (int r = 0; r < FF_CHUNK; ++r) {
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        acc[v][r].x = fma(qq[v], cur.x, acc[v][r].x);
                        acc[v][r].y = fma(qq[v], cur.y, acc[v][r].y);
                    }
                    cur = cmul(cur, st);
                }
            }
        }
    }
    if (active) {
        const int m0 = chunk_m0[c];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (v0 + v >= nv) break;
            double2 *out = Sp + (long long)(v0 + v) * vs_sp + (size_t)blockIdx.x * n_modes + m0;
#pragma unroll
            for (int r = 0; r < FF_CHUNK; ++r)
                if (r < len) out[r] = acc[v][r];
        }
    }
}

// S[pidx[m]] = w[m] * sum_blk Sp[blk][m]: workgroup = 32 modes x 8 partial-sum lanes.
__global__ __launch_bounds__(256) void k_ff_reduce(const double2 *__restrict__ Sp, int nblk, const double *__restrict__ w,
                                                   const int *__restrict__ pidx, int n_modes, double2 *__restrict__ S,
                                                   long long vs_sp = 0, long long vs_s = 0) {
    __shared__ double2 part[8][32];
    Sp += (long long)blockIdx.z * vs_sp;  // vector blockIdx.z of a many-vector launch
    S += (long long)blockIdx.z * vs_s;
    const int ml = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int m = blockIdx.x * 32 + ml;
    double sr = 0, si = 0;
    if (m < n_modes)
        for (int b = g; b < nblk; b += 8) {
            const double2 v = Sp[(size_t)b * n_modes + m];
            sr += v.x;
            si += v.y;
        }
    part[g][ml] = make_double2(sr, si);
    __syncthreads();
    if (g == 0 && m < n_modes) {
        double tr = 0, ti = 0;
        for (int k = 0; k < 8; ++k) {
            tr += part[k][ml].x;
            ti += part[k][ml].y;
        }
        S[pidx[m]] = make_double2(tr * w[m], ti * w[m]);
    }
}

// e *= z^d for a wave-uniform exponent d by square-and-multiply (|z| = 1, z^-1 = conj z).
__device__ __forceinline__ void ff_advance(double2 &e, int d, double2 z) {
    if (d < 0) {
        z.y = -z.y;
        d = -d;
    }
    while (d) {
        if (d & 1) e = cmul(e, z);
        d >>= 1;
        if (d) z = cmul(z, z);
    }
}

// far[u] = sum_m Re( S[m] e^{+i(i th1 + j th2 + k th3)} ), one point per lane.  Line L =
// (j, k, ilo, nq) holds its weighted S at S[base[L] ..], padded to nq * FF_QUAD entries; its
// sum over i is a Horner polynomial in z1 = e^{i th1} with wave-uniform coefficients (read
// FF_QUAD at a time), multiplied by z1^ilo z2^j z3^k, which is advanced from the previous
// line by (small) exponent steps.  Workgroup z of a many-vector launch evaluates vectors
// V z .. V z + V - 1 (< nv; S at S + v vs_s, output at far + v vs_far) with one set of
// phases; each vector's sums are those of a one-vector launch, bit for bit.
template <int V>
__global__ __launch_bounds__(256) void k_ff_eval(const double *__restrict__ theta, int n, const int4 *__restrict__ lines,
                                                 const int *__restrict__ base, int n_lines, const double2 *__restrict__ S,
                                                 double *__restrict__ far, long long vs_s = 0, long long vs_far = 0,
                                                 int nv = 1) {
    const int v0 = blockIdx.z * V;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int uc = u < n ? u : n - 1;
    double2 z1, z2, z3;
    {
        double s, c;
        sincos(theta[3 * (size_t)uc], &s, &c);
        z1 = make_double2(c, s);
        sincos(theta[3 * (size_t)uc + 1], &s, &c);
        z2 = make_double2(c, s);
        sincos(theta[3 * (size_t)uc + 2], &s, &c);
        z3 = make_double2(c, s);
    }
    // vectors past nv read vector v0's sums (in bounds) and store nothing
    const double2 *Sv[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Sv[v] = S + (long long)(v0 + (v0 + v < nv ? v : 0)) * vs_s;
    double2 e1 = make_double2(1, 0), e2 = make_double2(1, 0), e3 = make_double2(1, 0);
    int ic = 0, jc = 0, kc = 0;
    double acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0;
    for (int L = 0; L < n_lines; ++L) {
        const int4 d = lines[L];
        const int bL = base[L];
        ff_advance(e1, d.z - ic, z1);
        ff_advance(e2, d.x - jc, z2);
        ff_advance(e3, d.y - kc, z3);
        ic = d.z;
        jc = d.x;
        kc = d.y;
        double2 h[V];
#pragma unroll
        for (int v = 0; v < V; ++v) h[v] = make_double2(0, 0);
        for (int qd = d.w - 1; qd >= 0; --qd) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                double2 sv[FF_QUAD];
#pragma unroll
                for (int r = 0; r < FF_QUAD; ++r) sv[r] = Sv[v][bL + FF_QUAD * qd + r];
#pragma unroll
                for (int r = FF_QUAD - 1; r >= 0; --r)
                    h[v] = make_double2(fma(h[v].x, z1.x, fma(-h[v].y, z1.y, sv[r].x)),
                                        fma(h[v].x, z1.y, fma(h[v].y, z1.x, sv[r].y)));
            }
        }
        const double2 bs = cmul(cmul(e1, e2), e3);
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = fma(h[v].x, bs.x, fma(-h[v].y, bs.y, acc[v]));
    }
    if (u < n) {
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (v0 + v < nv) far[(long long)(v0 + v) * vs_far + u] = acc[v];
    }
}

// ------------------------------------------------------------------ far field by NUFFT
// Spreading, grid DFTs, deconvolution and interpolation of the non-uniform FFT evaluation of
// the same truncated Fourier sum (kernel values and point order prepared on the host).
// Grid layout [g2][g1][g0] (axis 0 fastest); mode box [k2][k1][k0], k = m + M/2.

// Lane l of a wavefront owns the kernel offsets (dy, dz) = (j % W, j / W), j = l + 64 ps
// (NP passes when W^2 > 64), and keeps its subgrid row (y, z) = (by + dy, bz + dz), the E
// cells along x, in registers.  A point's kernel along x is stored already placed in the row
// (phi_x at bx .. bx + W - 1, zeros elsewhere), held by lanes 0..E-1 of every 16-lane row and
// broadcast with DPP, so every point costs E FMAs per pass whatever its base bx; ky[dy],
// kz[dz], the charge and the row id are per-lane loads, issued D points ahead.
template <int W, int E> struct NufftCfg {
    static constexpr int NP = (W * W + 63) / 64;
    static constexpr int KS = E + 2 * W;  // kernel values per point
    static constexpr int D = NP == 1 ? 8 : 4;
};

template <int NP> struct NuPoint {
    double ka, kb, q, y[NP], z[NP];
    int pb;
};

// A zero the compiler cannot see through: indexing a uniform load with it keeps the load on
// the vector path, whose in-order completion lets the loads of later points stay in flight
// (a scalar load would need a full wait on every use).
__device__ __forceinline__ int nu_vzero() {
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

// Loads point p (kernel row, ky[dy], kz[dz], charge if Q, packed base); vz = nu_vzero().
template <int W, int E, bool Q, int NP>
__device__ __forceinline__ void nu_load(NuPoint<NP> &s, const double *__restrict__ kval, const double *__restrict__ qs,
                                        const int *__restrict__ pbase, int p, int rl, int vz, const int (&oy)[NP],
                                        const int (&oz)[NP]) {
    constexpr int KS = NufftCfg<W, E>::KS;
    const double *kv = kval + (size_t)p * KS;
    s.ka = kv[min(rl, KS - 1)];
    s.kb = E > 16 ? kv[min(16 + rl, KS - 1)] : 0.0;
#pragma unroll
    for (int ps = 0; ps < NP; ++ps) {
        s.y[ps] = kv[E + oy[ps]];
        s.z[ps] = kv[E + W + oz[ps]];
    }
    s.q = Q ? qs[p + vz] : 0.0;
    s.pb = pbase[p + vz];
}

// ---- per-point plan data of the NUFFT (built once per tree)
struct NufftGridArgs {
    int n[3], ntile[3], T, lo, B, W, E, KS;
    double h[3], beta;
};

__device__ __forceinline__ double nu_es_kernel(double z, double beta) {
    return fabs(z) < 1.0 ? exp(beta * (sqrt(1.0 - z * z) - 1.0)) : 0.0;
}

// Grid coordinate x in [0, n), tile coordinate and first grid point of the support along
// axis d of the point with angle theta.
__device__ __forceinline__ void nu_axis(double theta, const NufftGridArgs &a, int d, double &x, int &tc, int &g0) {
    const int n = a.n[d];
    x = theta / a.h[d];
    if (x < 0) x = 0;
    if (x >= n) x = nextafter((double)n, 0.0);
    tc = min(n - 1, (int)floor(x)) / a.T;
    g0 = (int)floor(x - 0.5 * a.W) + 1;
}

// key[u] = tile B^3 + (bz B + by) B + bx; bad[0] = 1 if a base leaves [0, B).
__global__ __launch_bounds__(256) void k_nufft_keys(const double *__restrict__ theta, int n, NufftGridArgs a,
                                                    int *__restrict__ key, int *__restrict__ bad) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    int tc[3], b[3];
    for (int d = 0; d < 3; ++d) {
        double x;
        int g0;
        nu_axis(theta[3 * (size_t)u + d], a, d, x, tc[d], g0);
        b[d] = g0 - a.T * tc[d] - a.lo;
        if (b[d] < 0 || b[d] >= a.B) {
            bad[0] = 1;
            b[d] = 0;
        }
    }
    const int tile = (tc[2] * a.ntile[1] + tc[1]) * a.ntile[0] + tc[0];
    key[u] = tile * a.B * a.B * a.B + (b[2] * a.B + b[1]) * a.B + b[0];
}

// Plan-order rows: kval[i] = (phi_x placed at bx .. bx+W-1 of the E-cell row, phi_y, phi_z)
// and pbase[i] for the point u = order[i].
__global__ __launch_bounds__(256) void k_nufft_rows(const double *__restrict__ theta, const int *__restrict__ order, int n,
                                                    NufftGridArgs a, double *__restrict__ kval, int *__restrict__ pbase) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int u = order[i];
    double *row = kval + (size_t)i * a.KS;
    int b[3];
    for (int d = 0; d < 3; ++d) {
        double x;
        int tc, g0;
        nu_axis(theta[3 * (size_t)u + d], a, d, x, tc, g0);
        b[d] = min(max(g0 - a.T * tc - a.lo, 0), a.B - 1);
        double *dst = d == 0 ? row + b[0] : row + a.E + (d - 1) * a.W;
        if (d == 0)
            for (int k = 0; k < a.E; ++k) row[k] = 0.0;
        for (int k = 0; k < a.W; ++k) dst[k] = nu_es_kernel((x - (g0 + k)) / (0.5 * a.W), a.beta);
    }
    pbase[i] = b[0] | (b[1] << 8) | (b[2] << 16);
}

// cnt[tile]++ for every point (keys of k_nufft_keys).
__global__ __launch_bounds__(256) void k_nufft_tile_hist(const int *__restrict__ key, int n, int B3, int *__restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[key[i] / B3], 1);
}

// qs[i] = q[order[i]]: the charges in plan order.
__global__ __launch_bounds__(256) void k_nufft_permute(const double *__restrict__ q, const int *__restrict__ order, int n,
                                                       double *__restrict__ qs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) qs[i] = q[order[i]];
}

// One wavefront per work unit (the points of one tile, or a part of them).  The lanes sum
// q phi_x phi_y phi_z over the points of a subgrid row (consecutive points with one (by, bz))
// in registers and add the row into the tile subgrid in LDS when the row changes; the
// subgrid (E^3, compact) goes to scratch[slot].  The sum order is fixed by the plan, so the
// result is deterministic.
template <int W, int E>
__global__ __launch_bounds__(64) void k_nufft_spread(const int *__restrict__ unit_ptr, const int *__restrict__ unit_slot,
                                                     const double *__restrict__ kval, const double *__restrict__ qs,
                                                     const int *__restrict__ pbase, int SY, int SZ,
                                                     double *__restrict__ scratch) {
    using C = NufftCfg<W, E>;
    constexpr int NP = C::NP, D = C::D;
    extern __shared__ double nu_lds[];
    double *sub = nu_lds;
    const int unit = blockIdx.x, lane = threadIdx.x;
    for (int e = lane; e < E * SZ; e += 64) sub[e] = 0.0;
    const int rl = lane & 15, vz = nu_vzero();
    int off[NP], oy[NP], oz[NP];
    bool ok[NP];
    double okf[NP];
#pragma unroll
    for (int ps = 0; ps < NP; ++ps) {
        const int j = lane + 64 * ps;
        ok[ps] = j < W * W;
        okf[ps] = ok[ps] ? 1.0 : 0.0;
        oy[ps] = ok[ps] ? j % W : 0;
        oz[ps] = ok[ps] ? j / W : 0;
        off[ps] = oz[ps] * SZ + oy[ps] * SY;
    }
    const int pbeg = unit_ptr[unit], pend = unit_ptr[unit + 1];
    double acc[NP][E];
#pragma unroll
    for (int ps = 0; ps < NP; ++ps)
#pragma unroll
        for (int k = 0; k < E; ++k) acc[ps][k] = 0.0;
    // adds the register rows into the subgrid row (by, bz) = (rw & 255, rw >> 8)
    auto flush = [&](int rw) {
        asm volatile("" ::: "memory");
        double *blk = sub + (rw >> 8) * SZ + (rw & 255) * SY;
#pragma unroll
        for (int ps = 0; ps < NP; ++ps)
            if (ok[ps]) {
                double *r = blk + off[ps];
#pragma unroll
                for (int k = 0; k < E; ++k) r[k] += acc[ps][k];
            }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int ps = 0; ps < NP; ++ps)
#pragma unroll
            for (int k = 0; k < E; ++k) acc[ps][k] = 0.0;
    };
    NuPoint<NP> ring[D];
#pragma unroll
    for (int j = 0; j < D; ++j) nu_load<W, E, true>(ring[j], kval, qs, pbase, min(pbeg + j, pend - 1), rl, vz, oy, oz);
    int row = -1;
    for (int p0 = pbeg; p0 < pend; p0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            if (p0 + j < pend) {
                const NuPoint<NP> cur = ring[j];
                nu_load<W, E, true>(ring[j], kval, qs, pbase, min(p0 + j + D, pend - 1), rl, vz, oy, oz);
                const int rw = __builtin_amdgcn_readfirstlane(cur.pb) >> 8;
                if (rw != row) {
                    if (row >= 0) flush(row);
                    row = rw;
                }
#pragma unroll
                for (int ps = 0; ps < NP; ++ps)
                    dpp_row_axpy<E>(acc[ps], cur.ka, cur.kb, cur.q * cur.y[ps] * cur.z[ps] * okf[ps]);
            }
        }
    }
    if (row >= 0) flush(row);
    __syncthreads();
    constexpr int E2 = E * E, E3 = E2 * E;
    double *dst = scratch + (size_t)unit_slot[unit] * E3;
    for (int e = lane; e < E3; e += 64) dst[e] = sub[(e / E2) * SZ + ((e / E) % E) * SY + e % E];
}

// scratch[tile] += scratch[extra slots of the tile] (in unit order), for tiles split into
// several units.  Workgroup (chunk of 256 cells, listed tile).
__global__ __launch_bounds__(256) void k_nufft_extra_reduce(double *__restrict__ scratch, const int *__restrict__ extra_tile,
                                                            const int *__restrict__ extra_ptr, int n_tiles, int E3) {
    const int i = blockIdx.y, e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E3) return;
    double v = scratch[(size_t)extra_tile[i] * E3 + e];
    for (int k = extra_ptr[i]; k < extra_ptr[i + 1]; ++k) v += scratch[(size_t)(n_tiles + k) * E3 + e];
    scratch[(size_t)extra_tile[i] * E3 + e] = v;
}

// grid[g] = sum over the tiles whose subgrid covers g (fixed order); tiles without points
// hold zeros.
__global__ __launch_bounds__(64) void k_nufft_gather(const double *__restrict__ scratch, const int *__restrict__ gp0,
                                                     const int *__restrict__ gt0, const int *__restrict__ gl0,
                                                     const int *__restrict__ gp1, const int *__restrict__ gt1,
                                                     const int *__restrict__ gl1, const int *__restrict__ gp2,
                                                     const int *__restrict__ gt2, const int *__restrict__ gl2, int n0, int n1,
                                                     int n2, int nt0, int nt1, int E, double *__restrict__ grid) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n0 * n1 * n2) return;
    const int g0 = g % n0, g1 = (g / n0) % n1, g2 = g / (n0 * n1);
    const size_t E3 = (size_t)E * E * E;
    double v = 0;
    for (int a = gp2[g2]; a < gp2[g2 + 1]; ++a)
        for (int b = gp1[g1]; b < gp1[g1 + 1]; ++b) {
            const int tzy = gt2[a] * nt1 + gt1[b];
            const int lzy = gl2[a] * E + gl1[b];
            for (int c = gp0[g0]; c < gp0[g0 + 1]; ++c)
                v += scratch[(size_t)(tzy * nt0 + gt0[c]) * E3 + (size_t)lzy * E + gl0[c]];
        }
    grid[g] = v;
}

// out[a][k][b] = sum_l in[a][l][b] mat[k L + l] (complex matrix; real or complex input and
// output, the real output keeping the real part).  A lane owns one line (a, b) (b fastest, so
// the loads of a wavefront are contiguous when B > 1) and DFT_KC outputs k of it, grid y =
// the chunk of k: each input element read feeds DFT_KC complex multiply-adds, and the matrix
// entries are wave-uniform loads.  Each output is summed over l in order.
constexpr int DFT_KC = 8;

template <bool RIN, bool ROUT>
__global__ __launch_bounds__(256) void k_dft_axis(const void *__restrict__ in, void *__restrict__ out,
                                                  const double2 *__restrict__ mat, int A, int L, int K, int B) {
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long)A * B) return;
    const int b = (int)(col % B);
    const long a = col / B;
    const int k0 = blockIdx.y * DFT_KC;
    int row[DFT_KC];
#pragma unroll
    for (int j = 0; j < DFT_KC; ++j) row[j] = min(k0 + j, K - 1) * L;  // clamped rows are not stored
    double sr[DFT_KC], si[DFT_KC];
#pragma unroll
    for (int j = 0; j < DFT_KC; ++j) sr[j] = si[j] = 0.0;
    for (int l = 0; l < L; ++l) {
        const size_t ii = ((size_t)a * L + l) * B + b;
        if (RIN) {
            const double x = static_cast<const double *>(in)[ii];
#pragma unroll
            for (int j = 0; j < DFT_KC; ++j) {
                const double2 m = mat[row[j] + l];
                sr[j] = fma(x, m.x, sr[j]);
                si[j] = fma(x, m.y, si[j]);
            }
        } else {
            const double2 x = static_cast<const double2 *>(in)[ii];
#pragma unroll
            for (int j = 0; j < DFT_KC; ++j) {
                const double2 m = mat[row[j] + l];
                sr[j] = fma(x.x, m.x, fma(-x.y, m.y, sr[j]));
                si[j] = fma(x.x, m.y, fma(x.y, m.x, si[j]));
            }
        }
    }
#pragma unroll
    for (int j = 0; j < DFT_KC; ++j) {
        if (k0 + j >= K) break;
        const size_t o = ((size_t)a * K + k0 + j) * B + b;
        if (ROUT) static_cast<double *>(out)[o] = sr[j];
        else static_cast<double2 *>(out)[o] = make_double2(sr[j], si[j]);
    }
}

// C[pos[i]] = cw[i] B[pos[i]] for the listed modes (the other entries of C stay zero).
__global__ void k_nufft_coef(const double2 *__restrict__ Bbox, const int *__restrict__ pos, const double *__restrict__ cw,
                             int n_modes, double2 *__restrict__ Cbox) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_modes) return;
    const double2 v = Bbox[pos[i]];
    Cbox[pos[i]] = make_double2(cw[i] * v.x, cw[i] * v.y);
}

// Row stride of the interpolation kernel's partial sums (odd, so that lane i reading row i
// spreads over the banks), and points per partial-sum batch.
constexpr int NU_PS = 65, NU_IB = 32;

// One wavefront per work unit: the tile subgrid (periodic wrap of the grid) is staged in
// LDS; for each subgrid row (by, bz) the lanes load their rows into registers and, per point,
// form sum_x phi_x row[x] phi_y phi_z over their offsets (dy, dz); the 64 lane partials of
// each point are added in lane order.  far[i] = the interpolated value at the point in plan
// position i.
template <int W, int E>
__global__ __launch_bounds__(64) void k_nufft_interp(const double *__restrict__ grid, const int *__restrict__ unit_ptr,
                                                     const int *__restrict__ unit_tile, const double *__restrict__ kval,
                                                     const int *__restrict__ pbase, int n0, int n1, int n2, int nt0, int nt1,
                                                     int T, int lo, int SY, int SZ, double *__restrict__ far) {
    using C = NufftCfg<W, E>;
    constexpr int NP = C::NP, D = C::D;
    extern __shared__ double nu_lds[];
    const int unit = blockIdx.x, lane = threadIdx.x;
    double *sub = nu_lds, *part = nu_lds + E * SZ;
    int *wrap = reinterpret_cast<int *>(part + NU_IB * NU_PS);  // 3 E grid indices
    const int rl = lane & 15, vz = nu_vzero();
    int off[NP], oy[NP], oz[NP];
    double okf[NP];
#pragma unroll
    for (int ps = 0; ps < NP; ++ps) {
        const int j = lane + 64 * ps;
        const bool ok = j < W * W;
        okf[ps] = ok ? 1.0 : 0.0;
        oy[ps] = ok ? j % W : 0;
        oz[ps] = ok ? j / W : 0;
        off[ps] = oz[ps] * SZ + oy[ps] * SY;
    }
    const int pbeg = unit_ptr[unit], pend = unit_ptr[unit + 1];
    NuPoint<NP> ring[D];
#pragma unroll
    for (int j = 0; j < D; ++j) nu_load<W, E, false>(ring[j], kval, nullptr, pbase, min(pbeg + j, pend - 1), rl, vz, oy, oz);
    const int tile = unit_tile[unit];
    const int org[3] = {T * (tile % nt0) + lo, T * ((tile / nt0) % nt1) + lo, T * (tile / (nt0 * nt1)) + lo};
    const int nn[3] = {n0, n1, n2};
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (lane < E) wrap[d * E + lane] = ((org[d] + lane) % nn[d] + nn[d]) % nn[d];
    __syncthreads();
    {
        constexpr int E2 = E * E, E3 = E2 * E;
        for (int e0 = 0; e0 < E3; e0 += 256) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = min(e0 + 64 * u + lane, E3 - 1);
                v[u] = grid[((size_t)wrap[2 * E + e / E2] * n1 + wrap[E + (e / E) % E]) * n0 + wrap[e % E]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + 64 * u + lane;
                if (e < E3) sub[(e / E2) * SZ + ((e / E) % E) * SY + e % E] = v[u];
            }
        }
    }
    __syncthreads();
    double col[NP][E];
    int row = -1, bstart = pbeg, nb = 0;
    // far for the nb points from bstart: lane i adds row i of the partial sums
    auto reduce = [&](int n_pts) {
        asm volatile("" ::: "memory");
        if (lane < n_pts) {
            double s = 0;
            for (int l = 0; l < 64; ++l) s += part[lane * NU_PS + l];
            far[bstart + lane] = s;
        }
        asm volatile("" ::: "memory");
    };
    for (int p0 = pbeg; p0 < pend; p0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            if (p0 + j < pend) {
                const NuPoint<NP> cur = ring[j];
                nu_load<W, E, false>(ring[j], kval, nullptr, pbase, min(p0 + j + D, pend - 1), rl, vz, oy, oz);
                const int rw = __builtin_amdgcn_readfirstlane(cur.pb) >> 8;
                if (rw != row) {
                    row = rw;
                    const double *blk = sub + (row >> 8) * SZ + (row & 255) * SY;
#pragma unroll
                    for (int ps = 0; ps < NP; ++ps)
#pragma unroll
                        for (int k = 0; k < E; ++k) col[ps][k] = blk[off[ps] + k];
                }
                double acc = 0;
#pragma unroll
                for (int ps = 0; ps < NP; ++ps) {
                    double s0 = 0, s1 = 0;
                    dpp_row_dot<E>(s0, s1, cur.ka, cur.kb, col[ps]);
                    acc = fma(s0 + s1, cur.y[ps] * cur.z[ps] * okf[ps], acc);
                }
                part[nb * NU_PS + lane] = acc;
                if (++nb == NU_IB) {
                    reduce(NU_IB);
                    bstart += NU_IB;
                    nb = 0;
                }
            }
        }
    }
    reduce(nb);
}


// ------------------------------------------------------------------ quasi-periodic far field
// psi(x_u) = e^{i b.x_u} sum_m w_m S(m) e^{i m.theta_u},  S(m) = sum_l q_l e^{-i b.x_l}
// e^{-i m.theta_l}, over the full mode box m = (i, j, k) - (ms/2, mt/2, mu/2), index
// i + ms (j + mt k); w_m = 0 outside the window (cpp/src/dmk/pbc.cpp:249-355).  FP64.
// Sp[y][m]: partial structure factors of the points of workgroup column y.  Per tile of tp
// points the per-axis factors e^{-i m_d theta_d} are tabulated in LDS by recurrence
// (tp (ms + mt + mu + 2) complex values); a thread owns QP_CHUNK consecutive i of one (j, k)
// line and walks them by multiplying with e^{-i theta_1}, so a (point, mode) pair costs one
// complex multiply-add.
constexpr int QP_CHUNK = 8;
__global__ __launch_bounds__(256) void k_qp_structure(const double *__restrict__ theta, const double *__restrict__ bx,
                                                       const double *__restrict__ q, int n, int ms, int mt, int mu, int tp,
                                                       double2 *__restrict__ Sp) {
    extern __shared__ double2 qp_lds[];
    double2 *E1 = qp_lds, *E2 = E1 + tp * ms, *E3 = E2 + tp * mt, *Q = E3 + tp * mu, *Est = Q + tp;
    const int M = ms * mt * mu, nci = (ms + QP_CHUNK - 1) / QP_CHUNK;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = c < nci * mt * mu;
    const int i0 = active ? QP_CHUNK * (c % nci) : 0, j = active ? (c / nci) % mt : 0, k = active ? c / (nci * mt) : 0;
    double2 acc[QP_CHUNK];
#pragma unroll
    for (int r = 0; r < QP_CHUNK; ++r) acc[r] = make_double2(0, 0);
    for (int p0 = blockIdx.y * tp; p0 < n; p0 += gridDim.y * tp) {
        __syncthreads();
        for (int t = threadIdx.x; t < tp * 4; t += blockDim.x) {
            const int pp = t % tp, ax = t / tp, gp = p0 + pp;
            if (ax == 3) {  // q e^{-i b.x}
                double sn = 0, cs = 0;
                if (gp < n) {
                    sincos(-bx[gp], &sn, &cs);
                    sn *= q[gp];
                    cs *= q[gp];
                }
                Q[pp] = make_double2(cs, sn);
                continue;
            }
            const int cnt = ax == 0 ? ms : (ax == 1 ? mt : mu);
            double2 *row = (ax == 0 ? E1 : (ax == 1 ? E2 : E3)) + pp * cnt;
            const double th = gp < n ? theta[3 * (size_t)gp + ax] : 0.0;
            double s1, c1, sl, cl;
            sincos(th, &s1, &c1);
            sincos(th * (cnt / 2), &sl, &cl);
            const double2 step = make_double2(c1, -s1);  // e^{-i th}
            if (ax == 0) Est[pp] = step;
            double2 cur = make_double2(cl, sl);  // e^{-i (-(cnt/2)) th}
            for (int mm = 0; mm < cnt; ++mm) {
                row[mm] = cur;
                cur = cmul(cur, step);
            }
        }
        __syncthreads();
        if (active) {
            const int np = min(tp, n - p0);
            for (int pp = 0; pp < np; ++pp) {
                const double2 st = Est[pp];
                double2 cur = cmul(cmul(cmul(E1[pp * ms + i0], E2[pp * mt + j]), E3[pp * mu + k]), Q[pp]);
#pragma unroll
                for (int r = 0; r < QP_CHUNK; ++r) {
                    acc[r].x += cur.x;
                    acc[r].y += cur.y;
                    cur = cmul(cur, st);
                }
            }
        }
    }
    if (active) {
        const int m0 = i0 + ms * (j + mt * k);
#pragma unroll
        for (int r = 0; r < QP_CHUNK; ++r)
            if (i0 + r < ms) Sp[(size_t)blockIdx.y * M + m0 + r] = acc[r];
    }
}

// S[m] = w[m] sum_y Sp[y][m], partials added in column order (deterministic).
__global__ __launch_bounds__(256) void k_qp_reduce(const double2 *__restrict__ Sp, int nblk, const double *__restrict__ w,
                                                    int M, double2 *__restrict__ S) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    double2 a = make_double2(0, 0);
    for (int b = 0; b < nblk; ++b) {
        const double2 v = Sp[(size_t)b * M + m];
        a.x += v.x;
        a.y += v.y;
    }
    S[m] = make_double2(a.x * w[m], a.y * w[m]);
}

// out[2u], out[2u + 1] += Re, Im psi(x_u).  The i axis is taken in chunks of QP_CHUNK: the
// chunk's factors e^{i m_1 theta_1} sit in registers and serve every (j, k) line, whose
// factors follow recurrences in e^{i theta_2}, e^{i theta_3}; the weighted structure factors
// are read at wave-uniform addresses.
template <class R>
__global__ __launch_bounds__(256) void k_qp_eval(const double *__restrict__ theta, const double *__restrict__ bx, int n,
                                                  int ms, int mt, int mu, const double2 *__restrict__ S, R *__restrict__ out) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    double2 st[3], e[3];
    const int cnt[3] = {ms, mt, mu};
    for (int d = 0; d < 3; ++d) {
        const double th = theta[3 * (size_t)u + d];
        double sn, cs;
        sincos(th, &sn, &cs);
        e[d] = make_double2(cs, sn);  // e^{i th}
        sincos(-th * (cnt[d] / 2), &sn, &cs);
        st[d] = make_double2(cs, sn);  // e^{i (-(cnt/2)) th}
    }
    double2 acc = make_double2(0, 0), c1next = st[0];
    for (int i0 = 0; i0 < ms; i0 += QP_CHUNK) {
        double2 c1[QP_CHUNK];
#pragma unroll
        for (int r = 0; r < QP_CHUNK; ++r) {
            c1[r] = c1next;
            c1next = cmul(c1next, e[0]);
        }
        const int nr = min(QP_CHUNK, ms - i0);
        double2 c3 = st[2];
        for (int k = 0; k < mu; ++k) {
            double2 acc2 = make_double2(0, 0), c2 = st[1];
            for (int j = 0; j < mt; ++j) {
                const double2 *row = S + ((size_t)k * mt + j) * ms + i0;
                double2 l = make_double2(0, 0);
#pragma unroll
                for (int r = 0; r < QP_CHUNK; ++r) {
                    if (r < nr) {
                        const double2 v = row[r];
                        l.x = fma(v.x, c1[r].x, fma(-v.y, c1[r].y, l.x));
                        l.y = fma(v.x, c1[r].y, fma(v.y, c1[r].x, l.y));
                    }
                }
                const double2 t2 = cmul(l, c2);
                acc2.x += t2.x;
                acc2.y += t2.y;
                c2 = cmul(c2, e[1]);
            }
            const double2 t3 = cmul(acc2, c3);
            acc.x += t3.x;
            acc.y += t3.y;
            c3 = cmul(c3, e[2]);
        }
    }
    double sn, cs;
    sincos(bx[u], &sn, &cs);
    const double2 psi = cmul(acc, make_double2(cs, sn));
    out[2 * (size_t)u] = (R)((double)out[2 * (size_t)u] + psi.x);
    out[2 * (size_t)u + 1] = (R)((double)out[2 * (size_t)u + 1] + psi.y);
}

}  // namespace pdmk::gpu
