// CPU oracle -- test infrastructure only.
//
// DMK parameter tables and per-level Fourier data for the 3-D Laplace kernel.
//
// Reference behaviour followed:
//   * beta table .............................. cpp/src/dmk/util.cpp:12-37 (calc_bandlimiting)
//   * n_pw / n_order table .................... cpp/include/dmk/tree.hpp:30-89 (ExpansionConstants)
//   * difference-kernel plane-wave terms ...... cpp/src/dmk/fourier_data.cpp:19-28
//   * Laplace-3D difference kernel FT ......... cpp/src/dmk/fourier_data.cpp:415-463, 614-644
//   * radial FT -> tensor-product modes ....... cpp/src/dmk/util.cpp:126-141
//   * PW <-> Chebyshev matrices ............... cpp/src/dmk/planewave.cpp:118-146
//   * PW translation (shift) matrices ......... cpp/src/dmk/planewave.cpp:148-223
//   * radialft fused into the shifts .......... cpp/src/dmk/pbc_tree.cpp:1069-1088
//   * c2p / p2c Chebyshev matrices ............ cpp/include/dmk/chebychev.hpp:212-302
//   * self-interaction constant W0 ............ cpp/include/dmk/tree.hpp:348-387
#pragma once

#include "omath.hpp"

#include <array>
#include <complex>
#include <vector>

namespace oracle {

// Near-field residual polynomial coefficients (monomial basis in
// x = (r - bs/2) * 2/bs, highest order last).  These are fitted data that the
// GPU path must share with the reference to reproduce its near field; they are
// restated here as data.  Source: cpp/src/dmk/aot_kernels.cpp:229-317.
#include "residual_coeffs.inc"

inline double beta_for_eps(double eps) {
    if (eps >= 1e-3) return 7.2462000846862793;
    if (eps >= 1e-6) return 13.739999771118164;
    if (eps >= 1e-9) return 20.736000061035156;
    if (eps >= 1e-12) return 27.870000839233398;
    return 2.664 * (-std::log10(eps)) + 0.306;
}

inline void npw_norder_for_eps(double eps, double beta, int &n_pw, int &n_order) {
    if (eps >= 1e-3) { n_pw = 13; n_order = 9; return; }
    if (eps >= 1e-6) { n_pw = 27; n_order = 18; return; }
    if (eps >= 1e-9) { n_pw = 39; n_order = 28; return; }
    if (eps >= 1e-12) { n_pw = 55; n_order = 38; return; }
    int nf = (int)std::ceil(3.0 * beta / M_PI * (1.0 - eps));
    nf -= 1;
    n_pw = 2 * nf + 1;
    n_order = (int)std::ceil(1.43 * beta - 3.26);
}

// Residual coefficient table for the tree's digit count (cpp/src/dmk/aot_kernels.cpp:320-470).
inline std::vector<double> residual_coeffs(int n_digits) {
    const double *t = nullptr;
    int nc = 0;
    switch (n_digits <= 2 ? 2 : n_digits) {
    case 2: t = kResidual3D_2; nc = 5; break;
    case 3: t = kResidual3D_3; nc = 6; break;
    case 4: t = kResidual3D_4; nc = 9; break;
    case 5: t = kResidual3D_5; nc = 11; break;
    case 6: t = kResidual3D_6; nc = 12; break;
    case 7: t = kResidual3D_7; nc = 15; break;
    case 8: t = kResidual3D_8; nc = 17; break;
    case 9: t = kResidual3D_9; nc = 18; break;
    case 10: t = kResidual3D_10; nc = 22; break;
    case 11: t = kResidual3D_11; nc = 23; break;
    case 12: t = kResidual3D_12; nc = 25; break;
    default: throw std::runtime_error("unsupported n_digits");
    }
    return std::vector<double>(t, t + nc);
}

// All per-tree tables, in double (cast to Real by the tree).
struct Tables {
    double beta = 0;
    int n_pw = 0, n_order = 0, n_modes = 0;  // n_modes = n_pw*n_pw*((n_pw+1)/2)
    int n_levels = 0;
    Prolate0 pf;
    double psi0 = 0, c0 = 0, c1 = 0, g0d2 = 0, c4 = 0;
    std::vector<double> boxsize;  // n_levels + 1
    std::vector<double> hpw;      // per level
    // per level
    std::vector<std::vector<double>> radialft;                  // n_modes
    std::vector<std::vector<std::complex<double>>> poly2pw;     // F-layout (n_pw x p)
    std::vector<std::vector<std::complex<double>>> pw2poly;     // F-layout (n_pw x p)
    std::vector<std::vector<std::complex<double>>> wpwshift;    // [27][n_modes], radialft fused
    std::vector<double> c2p, p2c;                               // [8][3][p x p] F-layout
    std::vector<double> w0;                                      // n_levels + 1

    void build(double eps, double beta_, double rc, int n_lev) {
        beta = beta_;
        npw_norder_for_eps(eps, beta, n_pw, n_order);
        n_levels = n_lev;
        n_modes = n_pw * n_pw * ((n_pw + 1) / 2);
        pf = Prolate0(beta);
        psi0 = (double)pf.eval(0);
        ld lc0, lc1, lc2, lc4;
        pf.intvals(lc0, lc1, lc2, lc4);
        c0 = (double)lc0; c1 = (double)lc1; g0d2 = (double)lc2; c4 = (double)lc4;

        boxsize.resize(n_lev + 1);
        boxsize[0] = rc;
        for (int l = 1; l <= n_lev; ++l) boxsize[l] = boxsize[l - 1] / 2;

        const int p = n_order;
        const int nf = (n_pw - 1) / 2;
        hpw.resize(n_lev);
        for (int l = 0; l < n_lev; ++l) hpw[l] = 2 * beta / nf / boxsize[l];

        // Level-0 difference kernel FT (radial samples k^2 = 0..3*(n_pw/2)^2)
        const int n_fourier = 3 * (n_pw / 2) * (n_pw / 2) + 1;
        std::vector<double> kft0(n_fourier);
        {
            const double bs = boxsize[0], small = 0.5 * bs, big = bs;
            const double h0 = hpw[0];
            const double ws = h0 * h0 * h0 * 0.5 / (M_PI * M_PI);
            std::vector<ld> xq, wq;
            gauss_legendre(100, xq, wq);
            for (int i = 0; i < n_fourier; ++i) {
                const ld rk = std::sqrt((ld)i) * h0;
                if (i == 0) {
                    kft0[i] = 0.5 * ws * g0d2 * (big * big - small * small);
                    continue;
                }
                ld s = 0;
                for (size_t j = 0; j < xq.size(); ++j) {
                    const ld t = 0.5L * (xq[j] + 1);
                    const ld f = 0.5L * pf.eval(t) * wq[j] / lc0;
                    s += (std::cos(rk * t * small) - std::cos(rk * t * big)) * f;
                }
                kft0[i] = (double)(s * ws / (rk * rk));
            }
        }

        // Chebyshev data
        auto xs = cheb_nodes(p);
        auto Vinv = cheb_vandermonde_inverse(p);  // Vinv[k*p+i]

        radialft.assign(n_lev, {});
        poly2pw.assign(n_lev, {});
        pw2poly.assign(n_lev, {});
        wpwshift.assign(n_lev, {});
        const int h = (n_pw + 1) / 2, half = n_pw / 2;
        for (int l = 0; l < n_lev; ++l) {
            const double scale = std::ldexp(1.0, l);  // Laplace-3D: x2 per level
            auto &rft = radialft[l];
            rft.resize(n_modes);
            int ind = 0;
            for (int j3 = -half; j3 <= 0; ++j3)
                for (int j2 = -half; j2 <= (n_pw - 1) / 2; ++j2)
                    for (int j1 = -half; j1 <= (n_pw - 1) / 2; ++j1)
                        rft[ind++] = kft0[j1 * j1 + j2 * j2 + j3 * j3] * scale;

            // pw2poly(j,k) = sum_i exp(i (j - half) x_i bs/2 hpw) Vinv(k,i); poly2pw = conj
            auto &P2 = pw2poly[l];
            auto &P1 = poly2pw[l];
            P2.assign(n_pw * p, 0);
            P1.assign(n_pw * p, 0);
            const ld dsq = 0.5L * boxsize[l];
            for (int j = 0; j < n_pw; ++j)
                for (int k = 0; k < p; ++k) {
                    ld re = 0, im = 0;
                    for (int i = 0; i < p; ++i) {
                        const ld ph = (ld)(j - half) * xs[i] * dsq * hpw[l];
                        re += std::cos(ph) * Vinv[k * p + i];
                        im += std::sin(ph) * Vinv[k * p + i];
                    }
                    P2[j + k * n_pw] = {(double)re, (double)im};
                    P1[j + k * n_pw] = {(double)re, -(double)im};
                }

            // translations: t = k0*9 + k1*3 + k2, mode m = j0*n_pw*n_pw + j1*n_pw + j2,
            // factor exp(i ts_{j} * bs * (k - 1)) per axis, ts_j = (j - half) hpw
            auto &W = wpwshift[l];
            W.assign(27 * n_modes, 0);
            for (int t = 0; t < 27; ++t) {
                const int k0 = t / 9, k1 = (t / 3) % 3, k2 = t % 3;
                int m = 0;
                for (int j0 = 0; j0 < h; ++j0)
                    for (int j1 = 0; j1 < n_pw; ++j1)
                        for (int j2 = 0; j2 < n_pw; ++j2, ++m) {
                            const ld ph = ((ld)(j0 - half) * (k0 - 1) + (ld)(j1 - half) * (k1 - 1) +
                                           (ld)(j2 - half) * (k2 - 1)) * hpw[l] * boxsize[l];
                            W[t * n_modes + m] = {(double)(std::cos(ph) * rft[m]), (double)(std::sin(ph) * rft[m])};
                        }
            }
        }

        // c2p / p2c: p2c_minus = Vinv * V(xm), p2c_plus = Vinv * V(xp), xm = x/2 - 1/2, xp = x/2 + 1/2
        std::array<std::vector<double>, 2> pc;
        for (int s = 0; s < 2; ++s) {
            pc[s].assign(p * p, 0);
            for (int k = 0; k < p; ++k)
                for (int j = 0; j < p; ++j) {
                    ld acc = 0;
                    for (int i = 0; i < p; ++i) {
                        const ld xc = 0.5L * xs[i] + (s == 0 ? -0.5L : 0.5L);
                        acc += Vinv[k * p + i] * cheb_T(j, xc);
                    }
                    pc[s][k + j * p] = (double)acc;  // F-layout (row k, col j)
                }
        }
        c2p.assign(8 * 3 * p * p, 0);
        p2c.assign(8 * 3 * p * p, 0);
        for (int child = 0; child < 8; ++child)
            for (int d = 0; d < 3; ++d) {
                const int bit = (child >> d) & 1;
                double *dst_p2c = &p2c[(child * 3 + d) * p * p];
                double *dst_c2p = &c2p[(child * 3 + d) * p * p];
                for (int r = 0; r < p; ++r)
                    for (int c = 0; c < p; ++c) {
                        dst_p2c[r + c * p] = pc[bit][r + c * p];
                        dst_c2p[r + c * p] = pc[bit][c + r * p];
                    }
            }

        // W0 per depth
        w0.resize(n_lev + 1);
        for (int l = 0; l < n_lev; ++l) {
            const double bs = (l == 0) ? 0.5 * boxsize[0] : boxsize[l];
            w0[l] = psi0 / (c0 * bs);
        }
        w0[n_lev] = psi0 / (c0 * (boxsize[n_lev - 1] / 2));
    }
};

}  // namespace oracle
