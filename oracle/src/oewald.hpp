// CPU oracle -- test infrastructure only.
//
// FP64 Ewald sums used to pin the DMK oracle end to end.
//   * ewald_direct: restatement of cpp/src/ewald_nufft.cpp:777-910 (O(N^2) real space
//     over lattice shells, exact reciprocal sum, self term -2 alpha/sqrt(pi) q_i;
//     balanced alpha = (N pi^3 / V^2)^(1/6), r_cut = s/alpha, k_max = 2 s alpha).
//   * ewald_cells: the same sum with a fractional-coordinate cell list for the real
//     space part (nearest image only, r_cut below half the smallest cell height) and
//     the exact reciprocal sum with separable phase factors; potentials at a subset of
//     targets.  Alpha follows choose_alpha's density rule (cpp/src/ewald_nufft.cpp:40-56).
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

struct Cell {
    double a[3][3];  // a[j][i]: component i of lattice vector j
    double b[3][3];  // reciprocal vectors (rows), b_i . a_j = 2 pi delta_ij
    double vol;
    double h[3];     // heights
    explicit Cell(const double *cell_colmajor) {
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) a[j][i] = cell_colmajor[i + 3 * j];
        auto cross = [](const double *u, const double *v, double *w) {
            w[0] = u[1] * v[2] - u[2] * v[1];
            w[1] = u[2] * v[0] - u[0] * v[2];
            w[2] = u[0] * v[1] - u[1] * v[0];
        };
        double c[3][3];
        cross(a[1], a[2], c[0]);
        cross(a[2], a[0], c[1]);
        cross(a[0], a[1], c[2]);
        vol = a[0][0] * c[0][0] + a[0][1] * c[0][1] + a[0][2] * c[0][2];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) b[i][k] = 2 * M_PI * c[i][k] / vol;
        vol = std::fabs(vol);
        for (int i = 0; i < 3; ++i) {
            double n = std::sqrt(c[i][0] * c[i][0] + c[i][1] * c[i][1] + c[i][2] * c[i][2]);
            h[i] = vol / n;
        }
    }
    void frac(const double *x, double *f) const {
        for (int i = 0; i < 3; ++i) f[i] = (b[i][0] * x[0] + b[i][1] * x[1] + b[i][2] * x[2]) / (2 * M_PI);
    }
};

inline void ewald_direct(const double *cell, int N, const double *pos, const double *q, double s, double *pot,
                         double *energy) {
    Cell C(cell);
    const double V = C.vol;
    const double alpha = std::pow(N * M_PI * M_PI * M_PI / (V * V), 1.0 / 6.0);
    const double rcut = s / alpha, kmax = 2 * s * alpha;
    int nmax[3];
    for (int d = 0; d < 3; ++d) nmax[d] = (int)std::ceil(rcut / C.h[d]);
    for (int i = 0; i < N; ++i) pot[i] = 0;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < N; ++i) {
        double acc = 0;
        for (int j = 0; j < N; ++j)
            for (int n0 = -nmax[0]; n0 <= nmax[0]; ++n0)
                for (int n1 = -nmax[1]; n1 <= nmax[1]; ++n1)
                    for (int n2 = -nmax[2]; n2 <= nmax[2]; ++n2) {
                        if (i == j && n0 == 0 && n1 == 0 && n2 == 0) continue;
                        double dr[3];
                        for (int d = 0; d < 3; ++d)
                            dr[d] = pos[3 * i + d] - pos[3 * j + d] + n0 * C.a[0][d] + n1 * C.a[1][d] + n2 * C.a[2][d];
                        const double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                        if (r2 > rcut * rcut) continue;
                        const double r = std::sqrt(r2);
                        acc += q[j] * std::erfc(alpha * r) / r;
                    }
        pot[i] = acc;
    }
    int mmax[3];
    for (int d = 0; d < 3; ++d) mmax[d] = (int)std::ceil(kmax / (2 * M_PI / C.h[d]));
    const double pref = 4 * M_PI / V, inv4a2 = 1 / (4 * alpha * alpha);
    std::vector<std::array<double, 4>> ks;
    for (int m3 = -mmax[2]; m3 <= mmax[2]; ++m3)
        for (int m2 = -mmax[1]; m2 <= mmax[1]; ++m2)
            for (int m1 = -mmax[0]; m1 <= mmax[0]; ++m1) {
                if (!m1 && !m2 && !m3) continue;
                double kv[3];
                for (int d = 0; d < 3; ++d) kv[d] = m1 * C.b[0][d] + m2 * C.b[1][d] + m3 * C.b[2][d];
                const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                if (k2 > kmax * kmax) continue;
                ks.push_back({kv[0], kv[1], kv[2], pref * std::exp(-k2 * inv4a2) / k2});
            }
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < N; ++i) {
        double acc = 0;
        for (auto &k : ks) {
            // Re(w S(k) e^{ik.x_i}) = w sum_j q_j cos(k.(x_i - x_j))
            for (int j = 0; j < N; ++j) {
                const double ph = k[0] * (pos[3 * i] - pos[3 * j]) + k[1] * (pos[3 * i + 1] - pos[3 * j + 1]) +
                                  k[2] * (pos[3 * i + 2] - pos[3 * j + 2]);
                acc += k[3] * q[j] * std::cos(ph);
            }
        }
        pot[i] += acc;
    }
    double e = 0, Q = 0;
    for (int i = 0; i < N; ++i) {
        pot[i] -= 2 * alpha / std::sqrt(M_PI) * q[i];
        e += q[i] * pot[i];
        Q += q[i];
    }
    if (energy) *energy = 0.5 * e - M_PI * Q * Q / (2 * V * alpha * alpha);
}

// Potentials at targets tidx[0..M) (or all when tidx == nullptr and M == N).  alpha_in > 0
// overrides the splitting parameter (the sum does not depend on it; it moves work between
// the real-space and reciprocal parts), still bounded so that r_cut < hmin/2.
inline void ewald_cells(const double *cell, int N, const double *pos, const double *q, double s, int M,
                        const int *tidx, double *pot, double alpha_in = 0) {
    Cell C(cell);
    const double V = C.vol;
    const double hmin = std::min(C.h[0], std::min(C.h[1], C.h[2]));
    // density rule with ~400 neighbours, bounded so that r_cut < hmin/2
    double alpha = alpha_in > 0 ? alpha_in : s * std::cbrt(4 * M_PI * (N / V) / (3 * 400.0));
    alpha = std::max(alpha, 2.05 * s / hmin);
    const double rcut = s / alpha, kmax = 2 * s * alpha;
    // fractional coordinates wrapped to [0,1)
    std::vector<double> f(3 * (size_t)N);
    for (int i = 0; i < N; ++i) {
        C.frac(&pos[3 * (size_t)i], &f[3 * (size_t)i]);
        for (int d = 0; d < 3; ++d) f[3 * (size_t)i + d] -= std::floor(f[3 * (size_t)i + d]);
    }
    int nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = std::max(1, (int)std::floor(C.h[d] / rcut));
    const int ncell = nc[0] * nc[1] * nc[2];
    std::vector<int> head(ncell + 1, 0), order(N), cellof(N);
    for (int i = 0; i < N; ++i) {
        int ci[3];
        for (int d = 0; d < 3; ++d) ci[d] = std::min(nc[d] - 1, (int)(f[3 * (size_t)i + d] * nc[d]));
        cellof[i] = (ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
        head[cellof[i] + 1]++;
    }
    for (int c = 0; c < ncell; ++c) head[c + 1] += head[c];
    {
        std::vector<int> fill(head.begin(), head.end() - 1);
        for (int i = 0; i < N; ++i) order[fill[cellof[i]]++] = i;
    }
    std::vector<int> targets(M);
    for (int t = 0; t < M; ++t) targets[t] = tidx ? tidx[t] : t;
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < M; ++t) {
        const int i = targets[t];
        int ci[3];
        for (int d = 0; d < 3; ++d) ci[d] = std::min(nc[d] - 1, (int)(f[3 * (size_t)i + d] * nc[d]));
        double acc = 0;
        // distinct neighbouring cells along each axis (periodic)
        int nb[3][3], nn[3];
        for (int d = 0; d < 3; ++d) {
            nn[d] = 0;
            for (int o = -1; o <= 1; ++o) {
                const int c = ((ci[d] + o) % nc[d] + nc[d]) % nc[d];
                bool seen = false;
                for (int u = 0; u < nn[d]; ++u) seen |= (nb[d][u] == c);
                if (!seen) nb[d][nn[d]++] = c;
            }
        }
        for (int iz = 0; iz < nn[2]; ++iz)
            for (int iy = 0; iy < nn[1]; ++iy)
                for (int ix = 0; ix < nn[0]; ++ix) {
                    const int c = (nb[2][iz] * nc[1] + nb[1][iy]) * nc[0] + nb[0][ix];
                    for (int k = head[c]; k < head[c + 1]; ++k) {
                        const int j = order[k];
                        // minimum image in fractional space, then check all neighbouring images
                        double df[3];
                        for (int d = 0; d < 3; ++d) {
                            df[d] = f[3 * (size_t)i + d] - f[3 * (size_t)j + d];
                            df[d] -= std::round(df[d]);
                        }
                        for (int n0 = -1; n0 <= 1; ++n0)
                            for (int n1 = -1; n1 <= 1; ++n1)
                                for (int n2 = -1; n2 <= 1; ++n2) {
                                    const double g0 = df[0] + n0, g1 = df[1] + n1, g2 = df[2] + n2;
                                    double dr[3];
                                    for (int d = 0; d < 3; ++d) dr[d] = g0 * C.a[0][d] + g1 * C.a[1][d] + g2 * C.a[2][d];
                                    const double rr = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                                    if (rr > rcut * rcut || rr == 0) continue;
                                    const double r = std::sqrt(rr);
                                    acc += q[j] * std::erfc(alpha * r) / r;
                                }
                    }
                }
        pot[t] = acc - 2 * alpha / std::sqrt(M_PI) * q[i];
    }
    // reciprocal: k = m1 b0 + m2 b1 + m3 b2, half space, exact structure factors
    int mmax[3];
    for (int d = 0; d < 3; ++d) mmax[d] = (int)std::ceil(kmax / (2 * M_PI / C.h[d]));
    struct K { int m1, m2, m3; double w; };
    std::vector<K> ks;
    const double pref = 4 * M_PI / V, inv4a2 = 1 / (4 * alpha * alpha);
    for (int m3 = 0; m3 <= mmax[2]; ++m3)
        for (int m2 = -mmax[1]; m2 <= mmax[1]; ++m2)
            for (int m1 = -mmax[0]; m1 <= mmax[0]; ++m1) {
                const bool pos_half = m3 > 0 || (m3 == 0 && (m2 > 0 || (m2 == 0 && m1 > 0)));
                if (!pos_half) continue;
                double kv[3];
                for (int d = 0; d < 3; ++d) kv[d] = m1 * C.b[0][d] + m2 * C.b[1][d] + m3 * C.b[2][d];
                const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                if (k2 > kmax * kmax) continue;
                ks.push_back({m1, m2, m3, 2 * pref * std::exp(-k2 * inv4a2) / k2});
            }
    const int NK = (int)ks.size();
    // phase of mode m at point j = 2 pi (m1 f0 + m2 f1 + m3 f2)
    std::vector<std::complex<double>> S(NK, 0);
    const int nthr =
#ifdef _OPENMP
        omp_get_max_threads();
#else
        1;
#endif
    std::vector<std::vector<std::complex<double>>> Sp(nthr, std::vector<std::complex<double>>(NK, 0));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        std::vector<std::complex<double>> e0(2 * mmax[0] + 1), e1(2 * mmax[1] + 1), e2(mmax[2] + 1);
#pragma omp for schedule(static)
        for (int j = 0; j < N; ++j) {
            const double *fj = &f[3 * (size_t)j];
            for (int m = -mmax[0]; m <= mmax[0]; ++m) e0[m + mmax[0]] = std::polar(1.0, -2 * M_PI * m * fj[0]);
            for (int m = -mmax[1]; m <= mmax[1]; ++m) e1[m + mmax[1]] = std::polar(1.0, -2 * M_PI * m * fj[1]);
            for (int m = 0; m <= mmax[2]; ++m) e2[m] = std::polar(1.0, -2 * M_PI * m * fj[2]);
            for (int k = 0; k < NK; ++k)
                Sp[tid][k] += q[j] * e0[ks[k].m1 + mmax[0]] * e1[ks[k].m2 + mmax[1]] * e2[ks[k].m3];
        }
    }
    for (int t = 0; t < nthr; ++t)
        for (int k = 0; k < NK; ++k) S[k] += Sp[t][k];
#pragma omp parallel for schedule(static)
    for (int t = 0; t < M; ++t) {
        const double *fi = &f[3 * (size_t)targets[t]];
        double acc = 0;
        for (int k = 0; k < NK; ++k) {
            const double ph = 2 * M_PI * (ks[k].m1 * fi[0] + ks[k].m2 * fi[1] + ks[k].m3 * fi[2]);
            acc += ks[k].w * (S[k].real() * std::cos(ph) - S[k].imag() * std::sin(ph));
        }
        pot[t] += acc;
    }
}

}  // namespace oracle
