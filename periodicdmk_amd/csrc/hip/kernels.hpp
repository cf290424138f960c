// gfx950 kernels of the periodic DMK evaluate() path.  Included by plan.hip only.
//
// Stage map (reference functions each kernel replaces):
//   k_gather_charges ... update_charges + image-charge gather
//                        (cpp/src/periodic_dmk.cpp:164-170, cpp/src/dmk/pbc_tree.cpp:1807-1821)
//   k_charge2proxy ..... proxy::charge2proxycharge_3d (cpp/src/dmk/proxy.cpp:153-212)
//   k_tensor3 .......... tensorprod::transform_3d c2p / p2c (cpp/src/dmk/tensorprod.cpp:39-75)
//   k_proxy2pw ......... proxy::proxycharge2pw_3d (cpp/src/dmk/proxy.cpp:42-90)
//   k_shift_pw2proxy ... form_eval_expansions shift + planewave_to_proxy_potential
//                        (cpp/src/dmk/pbc_tree.cpp:1294-1369, cpp/include/dmk/tree.hpp:287-316,
//                         cpp/src/dmk/planewave.cpp:49-105)
//   k_eval_targets ..... proxy::eval_targets_3d (cpp/src/dmk/proxy.cpp:307-448)
//   k_near_field ....... evaluate_direct_interactions + laplace_3d_poly_all_pairs
//                        (cpp/src/dmk/pbc_tree.cpp:1442-1681, cpp/include/dmk/vector_kernels.hpp:354-421)
//   k_ff_* ............. far_field_nufft, 3-D non-QP (cpp/src/dmk/pbc.cpp:378-541)
//   k_finalize ......... apply_target_self_correction + desort_potentials + far-field add
//                        (cpp/src/dmk/pbc_tree.cpp:1759-1805, cpp/src/periodic_dmk.cpp:172-182)
#pragma once

#include <hip/hip_runtime.h>

namespace pdmk::gpu {

template <class R>
struct alignas(4 * sizeof(R)) P4 { R x, y, z, w; };

template <class R>
struct alignas(2 * sizeof(R)) Cx { R re, im; };

template <class R> __device__ __forceinline__ R rsqrt_full(R x);
template <> __device__ __forceinline__ float rsqrt_full<float>(float x) { return __frsqrt_rn(x); }
template <> __device__ __forceinline__ double rsqrt_full<double>(double x) { return rsqrt(x); }

// ------------------------------------------------------------------ O(N) stages
template <class R>
__global__ void k_gather_charges(const double *__restrict__ q, const int *__restrict__ src_user, P4<R> *__restrict__ src,
                                 int nsall, R *__restrict__ pot_sorted, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsall) src[i].w = (R)q[src_user[i]];
    if (i < n) pot_sorted[i] = R(0);
}

// pot[u] = pot_sorted[i] - w0_i q_u + far[u], u = trg_user[i]
template <class R>
__global__ void k_finalize(const R *__restrict__ pot_sorted, const double *__restrict__ w0, const int *__restrict__ trg_user,
                           const double *__restrict__ q, const double *__restrict__ far, R *__restrict__ pot, int n,
                           int use_self, int use_far) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int u = trg_user[i];
    double v = (double)pot_sorted[i];
    if (use_self) v -= w0[i] * q[u];
    if (use_far) v += far[u];
    pot[u] = (R)v;
}

// ------------------------------------------------------------------ upward pass
// One workgroup per destination box.  Thread t owns the (b, c) column
// C[0..P-1, b, c] of the proxy coefficients and accumulates
//   C[a,b,c] = sum_s T_a(x_s) T_b(y_s) T_c(z_s) q_s,  x_s = sc (r_s - centre).
// Sources hold coordinates relative to their leaf centre; each range lies in one
// leaf and carries rng_off = leaf centre - box centre, so x_s = sc (local_s + off).
template <class R, int P>
__global__ __launch_bounds__(1024) void k_charge2proxy(const P4<R> *__restrict__ src, const int *__restrict__ grp_slot,
                                                        const R *__restrict__ grp_geom, const int *__restrict__ grp_rng,
                                                        const int *__restrict__ rng_begin, const int *__restrict__ rng_count,
                                                        const R *__restrict__ rng_off, R *__restrict__ proxy) {
    constexpr int TS = 32;
    __shared__ R Tx[TS * P], Ty[TS * P], Tz[TS * P];
    const int g = blockIdx.x;
    const R sc = grp_geom[4 * g + 3];
    R *out = proxy + (size_t)grp_slot[g] * (P * P * P);
    for (int pc0 = 0; pc0 < P * P; pc0 += blockDim.x) {
        const int pc = pc0 + threadIdx.x;
        const bool active = pc < P * P;
        const int b = active ? pc % P : 0, c = active ? pc / P : 0;
        R acc[P];
#pragma unroll
        for (int a = 0; a < P; ++a) acc[a] = R(0);
        for (int r = grp_rng[g]; r < grp_rng[g + 1]; ++r) {
            const int beg = rng_begin[r], cnt = rng_count[r];
            for (int s0 = 0; s0 < cnt; s0 += TS) {
                __syncthreads();
                for (int t = threadIdx.x; t < 3 * TS; t += blockDim.x) {
                    const int s = t % TS, d = t / TS;
                    R *T = (d == 0 ? Tx : (d == 1 ? Ty : Tz)) + s * P;
                    if (s0 + s < cnt) {
                        const P4<R> pt = src[beg + s0 + s];
                        const R x = sc * ((d == 0 ? pt.x : (d == 1 ? pt.y : pt.z)) + rng_off[4 * r + d]);
                        const R m = (d == 2) ? pt.w : R(1);
                        R t0 = m, t1 = x * m;
                        T[0] = t0;
                        if (P > 1) T[1] = t1;
                        for (int k = 2; k < P; ++k) {
                            const R t2 = R(2) * x * t1 - t0;
                            T[k] = t2;
                            t0 = t1;
                            t1 = t2;
                        }
                    } else {
                        for (int k = 0; k < P; ++k) T[k] = R(0);
                    }
                }
                __syncthreads();
                if (active) {
                    const int ns = min(TS, cnt - s0);
                    for (int s = 0; s < ns; ++s) {
                        const R w = Ty[s * P + b] * Tz[s * P + c];
#pragma unroll
                        for (int a = 0; a < P; ++a) acc[a] = fma(Tx[s * P + a], w, acc[a]);
                    }
                }
            }
        }
        if (active)
#pragma unroll
            for (int a = 0; a < P; ++a) out[a + P * b + P * P * c] = acc[a];
    }
}

// 3-mode tensor transform out(a',b',c') (+)= sum U0(a',a) U1(b',b) U2(c',c) in(a,b,c), with
// U_d the minus/plus parent<->child matrix selected by bit d of the child code (transposed for
// child -> parent).  Item i writes out slot item_out[i]; its inputs are the CSR range
// item_ptr[i]..item_ptr[i+1] of (in slot, child code); item_add[i] = accumulate into the output.
template <class R, int P>
__global__ __launch_bounds__(256) void k_tensor3(const R *__restrict__ in_base, R *__restrict__ out_base,
                                                  const int *__restrict__ item_out, const int *__restrict__ item_add,
                                                  const int *__restrict__ item_ptr, const int *__restrict__ in_slot,
                                                  const int *__restrict__ in_code, const R *__restrict__ mats,
                                                  int transpose, int n_items, R *__restrict__ scratch) {
    constexpr int P3 = P * P * P;
    __shared__ R U[3][P * P];
    R *t1 = scratch + (size_t)blockIdx.x * 2 * P3;
    R *t2 = t1 + P3;
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        R *out = out_base + (size_t)item_out[it] * P3;
        bool add = item_add[it] != 0;
        for (int e = item_ptr[it]; e < item_ptr[it + 1]; ++e) {
            const R *in = in_base + (size_t)in_slot[e] * P3;
            const int code = in_code[e];
            __syncthreads();
            for (int t = threadIdx.x; t < 3 * P * P; t += blockDim.x) {
                const int d = t / (P * P), rc = t % (P * P), r = rc % P, c = rc / P;
                const R *M = mats + ((code >> d) & 1) * P * P;
                U[d][r + P * c] = transpose ? M[c + P * r] : M[r + P * c];
            }
            __syncthreads();
            for (int o = threadIdx.x; o < P3; o += blockDim.x) {  // z mode
                const int ab = o % (P * P), c2 = o / (P * P);
                R s = 0;
                for (int c = 0; c < P; ++c) s = fma(U[2][c2 + P * c], in[ab + P * P * c], s);
                t1[o] = s;
            }
            __syncthreads();
            for (int o = threadIdx.x; o < P3; o += blockDim.x) {  // y mode
                const int a = o % P, b2 = (o / P) % P, c2 = o / (P * P);
                R s = 0;
                for (int b = 0; b < P; ++b) s = fma(U[1][b2 + P * b], t1[a + P * b + P * P * c2], s);
                t2[o] = s;
            }
            __syncthreads();
            for (int o = threadIdx.x; o < P3; o += blockDim.x) {  // x mode
                const int a2 = o % P, bc = o / P;
                R s = 0;
                for (int a = 0; a < P; ++a) s = fma(U[0][a2 + P * a], t2[a + P * bc], s);
                out[o] = add ? out[o] + s : s;
            }
            add = true;
        }
    }
}

// ------------------------------------------------------------------ plane waves
// pw(k1,k2,k3) = sum P(k1,m1) P(k2,m2) P(k3,m3) C(m1,m2,m3), k3 < H  (P = poly2pw, n_pw x p)
template <class R, int P>
__global__ __launch_bounds__(256) void k_proxy2pw(const R *__restrict__ proxy, Cx<R> *__restrict__ pw_out,
                                                   const Cx<R> *__restrict__ poly2pw, int n_pw, int n_boxes,
                                                   Cx<R> *__restrict__ scratch) {
    const int H = (n_pw + 1) / 2, nm = n_pw * n_pw * H;
    Cx<R> *t1 = scratch + (size_t)blockIdx.x * (P * P * H + P * n_pw * H);
    Cx<R> *t2 = t1 + P * P * H;
    for (int bx = blockIdx.x; bx < n_boxes; bx += gridDim.x) {
        const R *C = proxy + (size_t)bx * P * P * P;
        Cx<R> *pw = pw_out + (size_t)bx * nm;
        __syncthreads();
        for (int o = threadIdx.x; o < P * P * H; o += blockDim.x) {
            const int m12 = o % (P * P), k3 = o / (P * P);
            R sr = 0, si = 0;
            for (int m3 = 0; m3 < P; ++m3) {
                const Cx<R> w = poly2pw[k3 + n_pw * m3];
                const R c = C[m12 + P * P * m3];
                sr = fma(w.re, c, sr);
                si = fma(w.im, c, si);
            }
            t1[o] = {sr, si};
        }
        __syncthreads();
        for (int o = threadIdx.x; o < P * n_pw * H; o += blockDim.x) {
            const int m1 = o % P, k2 = (o / P) % n_pw, k3 = o / (P * n_pw);
            R sr = 0, si = 0;
            for (int m2 = 0; m2 < P; ++m2) {
                const Cx<R> w = poly2pw[k2 + n_pw * m2], v = t1[m1 + P * m2 + P * P * k3];
                sr = fma(w.re, v.re, fma(-w.im, v.im, sr));
                si = fma(w.re, v.im, fma(w.im, v.re, si));
            }
            t2[o] = {sr, si};
        }
        __syncthreads();
        for (int o = threadIdx.x; o < nm; o += blockDim.x) {
            const int k1 = o % n_pw, k23 = o / n_pw;
            R sr = 0, si = 0;
            for (int m1 = 0; m1 < P; ++m1) {
                const Cx<R> w = poly2pw[k1 + n_pw * m1], v = t2[m1 + P * k23];
                sr = fma(w.re, v.re, fma(-w.im, v.im, sr));
                si = fma(w.re, v.im, fma(w.im, v.re, si));
            }
            pw[o] = {sr, si};
        }
    }
}

// For each box of one level: pw_in = sum over (neighbour, stencil) of pw_out[nbr] * shift[stencil] * lscale
// (the self entry has a real shift), then
//   down(k1,k2,k3) (+)= 2 Re sum W(m1,k1) W(m2,k2) h(m3) W(m3,k3) pw_in(m1,m2,m3), h = 1/2 for m3 >= n_pw/2.
template <class R, int P>
__global__ __launch_bounds__(256) void k_shift_pw2proxy(const Cx<R> *__restrict__ pw_out, const Cx<R> *__restrict__ shift0,
                                                         R lscale, const int *__restrict__ box_down,
                                                         const int *__restrict__ box_zero, const int *__restrict__ nbr_ptr,
                                                         const int *__restrict__ nbr_slot, const int *__restrict__ nbr_stencil,
                                                         const Cx<R> *__restrict__ pw2poly, int n_pw, int n_items,
                                                         R *__restrict__ down_base, Cx<R> *__restrict__ scratch) {
    const int H = (n_pw + 1) / 2, nm = n_pw * n_pw * H, half = n_pw / 2;
    Cx<R> *pin = scratch + (size_t)blockIdx.x * (nm + P * n_pw * H + P * P * H);
    Cx<R> *t1 = pin + nm;
    Cx<R> *t2 = t1 + P * n_pw * H;
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        __syncthreads();
        for (int m = threadIdx.x; m < nm; m += blockDim.x) {
            R sr = 0, si = 0;
            for (int e = nbr_ptr[it]; e < nbr_ptr[it + 1]; ++e) {
                const Cx<R> a = pw_out[(size_t)nbr_slot[e] * nm + m];
                const Cx<R> s = shift0[(size_t)nbr_stencil[e] * nm + m];
                sr = fma(a.re, s.re, fma(-a.im, s.im, sr));
                si = fma(a.re, s.im, fma(a.im, s.re, si));
            }
            pin[m] = {sr * lscale, si * lscale};
        }
        __syncthreads();
        for (int o = threadIdx.x; o < P * n_pw * H; o += blockDim.x) {  // contract m1 -> k1
            const int k1 = o % P, m2 = (o / P) % n_pw, m3 = o / (P * n_pw);
            R sr = 0, si = 0;
            for (int m1 = 0; m1 < n_pw; ++m1) {
                const Cx<R> w = pw2poly[m1 + n_pw * k1], v = pin[m1 + n_pw * m2 + n_pw * n_pw * m3];
                sr = fma(w.re, v.re, fma(-w.im, v.im, sr));
                si = fma(w.re, v.im, fma(w.im, v.re, si));
            }
            t1[o] = {sr, si};
        }
        __syncthreads();
        for (int o = threadIdx.x; o < P * P * H; o += blockDim.x) {  // contract m2 -> k2, half weight
            const int k1 = o % P, k2 = (o / P) % P, m3 = o / (P * P);
            R sr = 0, si = 0;
            for (int m2 = 0; m2 < n_pw; ++m2) {
                const Cx<R> w = pw2poly[m2 + n_pw * k2], v = t1[k1 + P * m2 + P * n_pw * m3];
                sr = fma(w.re, v.re, fma(-w.im, v.im, sr));
                si = fma(w.re, v.im, fma(w.im, v.re, si));
            }
            const R hw = (m3 >= half) ? R(0.5) : R(1);
            t2[o] = {sr * hw, si * hw};
        }
        __syncthreads();
        R *down = down_base + (size_t)box_down[it] * P * P * P;
        const bool zero = box_zero[it] != 0;
        for (int o = threadIdx.x; o < P * P * P; o += blockDim.x) {  // contract m3 -> k3, 2 Re
            const int k12 = o % (P * P), k3 = o / (P * P);
            R s = 0;
            for (int m3 = 0; m3 < H; ++m3) {
                const Cx<R> w = pw2poly[m3 + n_pw * k3], v = t2[k12 + P * P * m3];
                s = fma(v.re, w.re, fma(-v.im, w.im, s));
            }
            down[o] = (zero ? R(0) : down[o]) + R(2) * s;
        }
    }
}

// pot_sorted[t] += sum C(a,b,c) T_a T_b T_c at the targets of one leaf per workgroup.
template <class R, int P>
__global__ __launch_bounds__(256) void k_eval_targets(const R *__restrict__ down_base, const int *__restrict__ leaf_down,
                                                       const R *__restrict__ leaf_geom, const int *__restrict__ leaf_trg,
                                                       const P4<R> *__restrict__ trg, R *__restrict__ pot_sorted) {
    const int lf = blockIdx.x;
    const R *C = down_base + (size_t)leaf_down[lf] * P * P * P;
    const R cx = leaf_geom[4 * lf], cy = leaf_geom[4 * lf + 1], cz = leaf_geom[4 * lf + 2], sc = leaf_geom[4 * lf + 3];
    const int beg = leaf_trg[2 * lf], cnt = leaf_trg[2 * lf + 1];
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
        const P4<R> pt = trg[beg + t];
        R tx[P], ty[P], tz[P];
        const R x = sc * (pt.x - cx), y = sc * (pt.y - cy), z = sc * (pt.z - cz);
        tx[0] = ty[0] = tz[0] = R(1);
        tx[1] = x; ty[1] = y; tz[1] = z;
#pragma unroll
        for (int k = 2; k < P; ++k) {
            tx[k] = R(2) * x * tx[k - 1] - tx[k - 2];
            ty[k] = R(2) * y * ty[k - 1] - ty[k - 2];
            tz[k] = R(2) * z * tz[k - 1] - tz[k - 2];
        }
        R s = 0;
        for (int c = 0; c < P; ++c) {
            R sb = 0;
            for (int b = 0; b < P; ++b) {
                const R *row = C + P * b + P * P * c;
                R sa = 0;
#pragma unroll
                for (int a = 0; a < P; ++a) sa = fma(row[a], tx[a], sa);
                sb = fma(sa, ty[b], sb);
            }
            s = fma(sb, tz[c], s);
        }
        pot_sorted[beg + t] += s;
    }
}

// ------------------------------------------------------------------ near field
// One workgroup per target leaf; each thread owns one target (loop if more).  For
// every source box in list1, sources are staged through LDS and the residual
// u = P((r - bs/2) 2/bs) / r, 1e-30 < r^2 < bs^2, is accumulated.  For fewer than 6
// digits P is used in r (even/odd split), as in laplace_3d_poly_all_pairs.
// Coordinates are relative to each point's leaf centre; l1_geom[e] = (centre of the
// target leaf - centre of the source leaf, bs), so r = |(t + off) - s|.  Each list entry
// is summed in R and the per-target total in FP64.
template <class R, int NC, bool SPLIT>
__global__ __launch_bounds__(256) void k_near_field(const P4<R> *__restrict__ src, const P4<R> *__restrict__ trg,
                                                     const int *__restrict__ leaf_trg, const int *__restrict__ l1_ptr,
                                                     const int *__restrict__ l1_src, const P4<R> *__restrict__ l1_geom,
                                                     const R *__restrict__ coeffs, R *__restrict__ pot_sorted) {
    constexpr int TILE = 256;
    __shared__ P4<R> tile[TILE];
    __shared__ R cf[64];
    R c[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) c[i] = coeffs[i];
    const int lf = blockIdx.x;
    const int tbeg = leaf_trg[2 * lf], tcnt = leaf_trg[2 * lf + 1];
    for (int rep = 0; rep * blockDim.x < tcnt; ++rep) {
        const int t = rep * blockDim.x + threadIdx.x;
        const bool active = t < tcnt;
        const P4<R> me = active ? trg[tbeg + t] : P4<R>{R(0), R(0), R(0), R(0)};
        double acc = 0;
        for (int e = l1_ptr[lf]; e < l1_ptr[lf + 1]; ++e) {
            const int sbeg = l1_src[2 * e], scnt = l1_src[2 * e + 1];
            const P4<R> gm = l1_geom[e];
            const R bs = gm.w;
            const R tx = me.x + gm.x, ty = me.y + gm.y, tz = me.z + gm.z;
            const R d2max = bs * bs, rsc = R(2) / bs, cen = -bs / R(2);
            __syncthreads();
            if (SPLIT && threadIdx.x == 0) {
                // monomial coefficients in r: sum c_i rsc^i (r + cen)^i
                double cd[NC];
                double rp = 1;
                for (int i = 0; i < NC; ++i) { cd[i] = (double)coeffs[i] * rp; rp *= (double)rsc; }
                for (int i = 0; i < NC; ++i)
                    for (int j = NC - 1; j > i; --j) cd[j - 1] += (double)cen * cd[j];
                for (int i = 0; i < NC; ++i) cf[i] = (R)cd[i];
            }
            R ea = 0;
            for (int s0 = 0; s0 < scnt; s0 += TILE) {
                const int ns = min(TILE, scnt - s0);
                __syncthreads();
                if ((int)threadIdx.x < ns) tile[threadIdx.x] = src[sbeg + s0 + threadIdx.x];
                __syncthreads();
                if (active) {
                    for (int s = 0; s < ns; ++s) {
                        const P4<R> q = tile[s];
                        const R dx = tx - q.x, dy = ty - q.y, dz = tz - q.z;
                        const R r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                        const R rinv = rsqrt_full<R>(r2);
                        R u;
                        if (SPLIT) {
                            constexpr int SE = (NC - 1) & ~1, SO = ((NC - 2) & ~1) + 1;
                            R E = cf[SE];
#pragma unroll
                            for (int i = SE - 2; i >= 0; i -= 2) E = fma(E, r2, cf[i]);
                            R O = cf[SO];
#pragma unroll
                            for (int i = SO - 2; i >= 0; i -= 2) O = fma(O, r2, cf[i]);
                            u = fma(E, rinv, O);
                        } else {
                            const R x = fma(r2, rinv, cen) * rsc;
                            R p = c[NC - 1];
#pragma unroll
                            for (int i = NC - 2; i >= 0; --i) p = fma(p, x, c[i]);
                            u = p * rinv;
                        }
                        const bool in = (r2 > R(1e-30)) && (r2 < d2max);
                        ea = in ? fma(u, q.w, ea) : ea;
                    }
                }
            }
            acc += (double)ea;
        }
        if (active) pot_sorted[tbeg + t] += (R)acc;
    }
}

// ------------------------------------------------------------------ far field
// Half-space modes with nonzero weight are grouped in lines of constant (j, k) with i
// contiguous; each line is cut into chunks of up to FF_CHUNK consecutive i.  Chunk c
// covers modes chunk_m0[c] .. + chunk_len[c] - 1 with indices (chunk_i0[c] + r, j, k).
// Structure factors per block of points:
//   Sp[blk][m] = sum_{points of blk} q e^{-i(i th1 + j th2 + k th3)}.
// Per-point axis factors e^{-i m th} are built in FP64 by recurrence in LDS.
constexpr int FF_CHUNK = 4;
constexpr int FF_TP = 16;  // points per LDS tile

__global__ __launch_bounds__(256) void k_ff_structure(const double *__restrict__ theta, const double *__restrict__ q, int n,
                                                       int ms, int mt, int mu, const int *__restrict__ chunk_ijk,
                                                       const int *__restrict__ chunk_m0, const int *__restrict__ chunk_len,
                                                       int n_chunks, int n_modes, double2 *__restrict__ Sp) {
    extern __shared__ double2 ff_lds[];
    const int ni = ms, nj = mt, nk = mu;  // index ranges [-m/2, (m-1)/2]
    double2 *E1 = ff_lds, *E2 = E1 + FF_TP * ni, *E3 = E2 + FF_TP * nj;
    const int c = blockIdx.y * blockDim.x + threadIdx.x;
    const bool active = c < n_chunks;
    const int ci = active ? chunk_ijk[3 * c] + ms / 2 : 0;
    const int cj = active ? chunk_ijk[3 * c + 1] + mt / 2 : 0;
    const int ck = active ? chunk_ijk[3 * c + 2] + mu / 2 : 0;
    const int len = active ? chunk_len[c] : 0;
    double2 acc[FF_CHUNK];
#pragma unroll
    for (int r = 0; r < FF_CHUNK; ++r) acc[r] = make_double2(0, 0);
    for (int p0 = blockIdx.x * FF_TP; p0 < n; p0 += gridDim.x * FF_TP) {
        __syncthreads();
        for (int t = threadIdx.x; t < FF_TP * 3; t += blockDim.x) {
            const int pp = t % FF_TP, ax = t / FF_TP, gp = p0 + pp;
            const int cnt = ax == 0 ? ni : (ax == 1 ? nj : nk);
            const int lo = -((ax == 0 ? ms : (ax == 1 ? mt : mu)) / 2);
            double2 *row = (ax == 0 ? E1 : (ax == 1 ? E2 : E3)) + pp * cnt;
            const double th = gp < n ? theta[3 * (size_t)gp + ax] : 0.0;
            const double qq = gp < n ? (ax == 2 ? q[gp] : 1.0) : 0.0;
            double s1, c1, sl, cl;
            sincos(th, &s1, &c1);
            sincos(-th * lo, &sl, &cl);
            double2 cur = make_double2(cl * qq, sl * qq);  // q e^{-i lo th}
            for (int m = 0; m < cnt; ++m) {
                row[m] = cur;
                cur = make_double2(cur.x * c1 + cur.y * s1, cur.y * c1 - cur.x * s1);  // x e^{-i th}
            }
        }
        __syncthreads();
        if (active) {
            const int np = min(FF_TP, n - p0);
            for (int pp = 0; pp < np; ++pp) {
                const double2 a = E2[pp * nj + cj], b = E3[pp * nk + ck];
                const double2 base = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
#pragma unroll
                for (int r = 0; r < FF_CHUNK; ++r) {
                    const double2 e = E1[pp * ni + min(ci + r, ni - 1)];
                    acc[r].x = fma(e.x, base.x, fma(-e.y, base.y, acc[r].x));
                    acc[r].y = fma(e.x, base.y, fma(e.y, base.x, acc[r].y));
                }
            }
        }
    }
    if (active) {
        const int m0 = chunk_m0[c];
#pragma unroll
        for (int r = 0; r < FF_CHUNK; ++r)
            if (r < len) Sp[(size_t)blockIdx.x * n_modes + m0 + r] = acc[r];
    }
}

// S[m] = w[m] * sum_blk Sp[blk][m]
__global__ void k_ff_reduce(const double2 *__restrict__ Sp, int nblk, const double *__restrict__ w, int n_modes,
                            double2 *__restrict__ S) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n_modes) return;
    double sr = 0, si = 0;
    for (int b = 0; b < nblk; ++b) {
        const double2 v = Sp[(size_t)b * n_modes + m];
        sr += v.x;
        si += v.y;
    }
    S[m] = make_double2(sr * w[m], si * w[m]);
}

// far[u] = sum_m Re( S[m] e^{+i(i th1 + j th2 + k th3)} )
__global__ __launch_bounds__(256) void k_ff_eval(const double *__restrict__ theta, int n, const int *__restrict__ line_jk,
                                                 const int *__restrict__ line_i, const int *__restrict__ line_ptr,
                                                 int n_lines, const double2 *__restrict__ S, double *__restrict__ far) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const double t1 = theta[3 * (size_t)u], t2 = theta[3 * (size_t)u + 1], t3 = theta[3 * (size_t)u + 2];
    double s1, c1;
    sincos(t1, &s1, &c1);
    double acc = 0;
    for (int L = 0; L < n_lines; ++L) {
        const int j = line_jk[2 * L], k = line_jk[2 * L + 1], ilo = line_i[2 * L], ihi = line_i[2 * L + 1];
        double sb, cb;
        sincos(ilo * t1 + j * t2 + k * t3, &sb, &cb);
        double br = cb, bi = sb;
        const double2 *Sl = S + line_ptr[L];
        for (int i = 0; i <= ihi - ilo; ++i) {
            const double2 sv = Sl[i];
            acc += sv.x * br - sv.y * bi;
            const double nr = br * c1 - bi * s1;
            bi = br * s1 + bi * c1;
            br = nr;
        }
    }
    far[u] = acc;
}

}  // namespace pdmk::gpu
