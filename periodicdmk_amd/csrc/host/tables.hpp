// Host-side DMK tables for the 3-D Laplace kernel, built once per tree in FP64.
//
// Reference behaviour followed:
//   * eps = 10^-max(1,n_digits), tree digits round(log10(1/eps) - 0.1)
//                                   cpp/src/periodic_dmk.cpp:52, cpp/src/dmk/pbc_tree.cpp:148
//   * beta, n_pw, n_order tables .. cpp/src/dmk/util.cpp:12-37, cpp/include/dmk/tree.hpp:30-89
//   * psi_0^c (Legendre series, unit L2 norm) and its integrals
//                                   cpp/src/dmk/prolate.cpp:187-250, 479-510
//   * difference-kernel FT ........ cpp/src/dmk/fourier_data.cpp:19-28, 415-463, 614-644
//   * tensor-product radial FT .... cpp/src/dmk/util.cpp:126-141
//   * PW <-> Chebyshev matrices ... cpp/src/dmk/planewave.cpp:118-146
//   * PW shift matrices (radial FT fused) cpp/src/dmk/planewave.cpp:148-223, pbc_tree.cpp:1069-1088
//   * p2c / c2p ................... cpp/include/dmk/chebychev.hpp:212-302
//   * W0 .......................... cpp/include/dmk/tree.hpp:348-387
//   * far-field weights ........... cpp/src/dmk/pbc.cpp:378-438
#pragma once

#include <complex>
#include <vector>

namespace pdmk::host {

// psi_0^c on [-1, 1] as an even normalised-Legendre series.
struct Pswf {
    double c = 0;
    std::vector<double> coef;  // coefficient of sqrt(k + 1/2) P_k(x), k = 0, 2, 4, ...
    explicit Pswf(double c_ = 0);
    double operator()(double x) const;
};

struct DmkTables {
    int n_digits = 0;      // digits of the tree
    double eps = 0, beta = 0;
    int n_pw = 0, p = 0, h = 0, n_modes = 0;  // h = (n_pw+1)/2, n_modes = n_pw^2 h
    std::vector<double> residual;             // NC monomial coefficients
    double psi0 = 0, c0 = 0, g0d2 = 0;
    int nlevels = 0;
    std::vector<double> boxsize;              // nlevels + 1
    std::vector<double> w0;                   // nlevels + 1, indexed by depth + pwexp
    // plane-wave modes (z index a < h, y index b < n_pw, x index c < n_pw) are stored at
    // m = c (h n_pw) + a n_pw + b, see mode_index()
    std::vector<double> radial0;              // n_modes, level-0 difference kernel FT (level l: x 2^l)
    std::vector<std::complex<double>> poly2pw, pw2poly;  // [n_pw x p] column-major, same at every level
    std::vector<std::complex<double>> shift0;  // [27][n_modes]: translation phase x radial0
    std::vector<double> p2c[2];               // [p x p] column-major, child on the minus / plus side
    // far field
    int ms = 0, mt = 0, mu = 0;
    double ainv[6] = {0, 0, 0, 0, 0, 0};      // a11, a21, a22, a31, a32, a33 of U^{-1}
    std::vector<int> mode_ijk;                // 3 per half-space mode with nonzero weight
    std::vector<double> mode_w;               // 2 * weight (conjugate mode folded in)
};

// Gauss-Legendre nodes (ascending) and weights on [-1, 1].
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w);

inline int mode_index(const DmkTables &T, int a, int b, int c) { return (c * T.h + a) * T.n_pw + b; }

// n_digits in 1..12; rc and nlevels describe the tree; umat is the lattice frame.
DmkTables make_tables(int n_digits, double rc, int nlevels, const double *umat);

}  // namespace pdmk::host
