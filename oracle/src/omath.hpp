// CPU oracle -- test infrastructure only.
//
// Special functions and quadrature used to build the DMK tables: Gauss-Legendre
// rules, the prolate spheroidal wave function psi_0^c on [-1,1], and Chebyshev
// interpolation matrices.  Everything here is evaluated in long double so that the
// oracle's tables are a more precise restatement than the double-precision
// tables the product computes.
//
// Reference behaviour followed:
//   * Gauss-Legendre nodes/weights ........ cpp/src/dmk/legeexps.cpp:77 (legerts)
//   * psi_0^c as an even Legendre series with unit L2 norm on [-1,1]
//     ....................................... cpp/src/dmk/prolate.cpp:187-250 (prolfun0)
//   * integral values c0,c1,c2/c0,c4 ....... cpp/src/dmk/prolate.cpp:479-510 (prolate_intvals)
//   * Chebyshev nodes / Vandermonde / p2c ... cpp/include/dmk/chebychev.hpp:38-226
#pragma once

#include <cmath>
#include <vector>
#include <stdexcept>

namespace oracle {

using ld = long double;

// Gauss-Legendre rule on [-1,1], nodes ascending.  Newton iteration on P_n.
inline void gauss_legendre(int n, std::vector<ld> &x, std::vector<ld> &w) {
    x.assign(n, 0);
    w.assign(n, 0);
    const ld pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        ld z = std::cos(pi * (i + 0.75L) / (n + 0.5L));
        ld dp = 0;
        for (int it = 0; it < 100; ++it) {
            ld p0 = 1, p1 = z;
            for (int k = 1; k < n; ++k) {
                ld p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            ld dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-19L) {
                // one more evaluation for the derivative at the converged node
                p0 = 1; p1 = z;
                for (int k = 1; k < n; ++k) {
                    ld p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
                    p0 = p1;
                    p1 = p2;
                }
                dp = n * (z * p1 - p0) / (z * z - 1);
                break;
            }
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        ld wi = 2 / ((1 - z * z) * dp * dp);
        w[i] = wi;
        w[n - 1 - i] = wi;
    }
}

// psi_0^c(x) = sum_i a_i * Pbar_{2i}(x), Pbar_k = sqrt(k+1/2) P_k, sum a_i^2 = 1,
// sign chosen so that psi_0(0) > 0.  The coefficients are the eigenvector of the
// smallest eigenvalue of the prolate operator restricted to even normalised
// Legendre functions (a symmetric tridiagonal matrix), found by Sturm bisection
// followed by inverse iteration.
struct Prolate0 {
    ld c = 0;
    std::vector<ld> a;  // coefficients of Pbar_{2i}

    Prolate0() = default;
    explicit Prolate0(double c_) : c(c_) {
        const int m = 40 + (int)(1.5 * c_);  // number of even terms
        std::vector<ld> d(m), e(m, 0);     // diagonal, super-diagonal
        const ld c2 = c * c;
        for (int i = 0; i < m; ++i) {
            const ld k = 2 * i;
            d[i] = k * (k + 1) + c2 * (2 * k * (k + 1) - 1) / ((2 * k + 3) * (2 * k - 1));
            e[i] = c2 * (k + 2) * (k + 1) / ((2 * k + 3) * std::sqrt((2 * k + 1) * (2 * k + 5)));
        }
        // Sturm count: number of eigenvalues < lam
        auto count_below = [&](ld lam) {
            int cnt = 0;
            ld q = d[0] - lam;
            if (q < 0) cnt++;
            for (int i = 1; i < m; ++i) {
                if (q == 0) q = 1e-300L;
                q = d[i] - lam - e[i - 1] * e[i - 1] / q;
                if (q < 0) cnt++;
            }
            return cnt;
        };
        ld lo = -1, hi = d[0] + std::fabs(e[0]) + 1;
        while (count_below(hi) < 1) hi *= 2;
        for (int it = 0; it < 200; ++it) {
            ld mid = 0.5L * (lo + hi);
            if (count_below(mid) >= 1) hi = mid; else lo = mid;
        }
        const ld lam = lo - 1e-14L * (1 + std::fabs(lo));
        // inverse iteration with the tridiagonal (T - lam I)
        a.assign(m, 1);
        for (int it = 0; it < 6; ++it) {
            std::vector<ld> cp(m), dp(m), rhs = a;
            // Thomas algorithm
            ld denom = d[0] - lam;
            cp[0] = e[0] / denom;
            dp[0] = rhs[0] / denom;
            for (int i = 1; i < m; ++i) {
                denom = (d[i] - lam) - e[i - 1] * cp[i - 1];
                cp[i] = (i < m - 1) ? e[i] / denom : 0;
                dp[i] = (rhs[i] - e[i - 1] * dp[i - 1]) / denom;
            }
            a[m - 1] = dp[m - 1];
            for (int i = m - 2; i >= 0; --i) a[i] = dp[i] - cp[i] * a[i + 1];
            ld nrm = 0;
            for (auto v : a) nrm += v * v;
            nrm = std::sqrt(nrm);
            for (auto &v : a) v /= nrm;
        }
        if (eval(0) < 0)
            for (auto &v : a) v = -v;
    }

    // psi_0(x) for |x| <= 1 (the DMK path never evaluates outside [-1,1]).
    ld eval(ld x) const {
        const int kmax = 2 * (int)a.size() - 2;
        ld pm1 = 0, p = 1, s = 0;  // P_{k-1}, P_k
        for (int k = 0; k <= kmax; ++k) {
            if ((k & 1) == 0) s += a[k / 2] * std::sqrt((ld)k + 0.5L) * p;
            const ld pn = ((2 * k + 1) * x * p - k * pm1) / (k + 1);
            pm1 = p;
            p = pn;
        }
        return s;
    }

    // int_0^r psi_0(t) dt, using int P_k = (P_{k+1} - P_{k-1})/(2k+1).
    ld integral(ld r) const {
        const int kmax = 2 * (int)a.size();
        std::vector<ld> P(kmax + 2), P0(kmax + 2);
        P[0] = 1; P[1] = r; P0[0] = 1; P0[1] = 0;
        for (int k = 1; k <= kmax; ++k) {
            P[k + 1] = ((2 * k + 1) * r * P[k] - k * P[k - 1]) / (k + 1);
            P0[k + 1] = ((2 * k + 1) * 0 * P0[k] - k * P0[k - 1]) / (k + 1);
        }
        ld s = a[0] * std::sqrt(0.5L) * r;  // k = 0 term
        for (int i = 1; i < (int)a.size(); ++i) {
            const int k = 2 * i;
            const ld anti_r = (P[k + 1] - P[k - 1]) / (2 * k + 1);
            const ld anti_0 = (P0[k + 1] - P0[k - 1]) / (2 * k + 1);
            s += a[i] * std::sqrt((ld)k + 0.5L) * (anti_r - anti_0);
        }
        return s;
    }

    // c0 = int_0^1 psi0, c1 = int_0^1 x psi0, c2 = (int_0^1 x^2 psi0)/c0, c4 = int_0^1 x^4 psi0.
    void intvals(ld &c0, ld &c1, ld &c2, ld &c4) const {
        std::vector<ld> x, w;
        gauss_legendre(120, x, w);
        c0 = c1 = c2 = c4 = 0;
        for (size_t i = 0; i < x.size(); ++i) {
            const ld t = 0.5L * (x[i] + 1), wt = 0.5L * w[i];
            const ld f = eval(t) * wt;
            c0 += f;
            c1 += f * t;
            c2 += f * t * t;
            c4 += f * t * t * t * t;
        }
        c2 /= c0;
    }
};

// Chebyshev nodes on [-1,1], ascending: x_i = cos(pi*((p-1-i)+1/2)/p).
inline std::vector<ld> cheb_nodes(int p) {
    const ld pi = 3.141592653589793238462643383279502884L;
    std::vector<ld> x(p);
    for (int i = 0; i < p; ++i) x[i] = std::cos(pi * ((p - 1 - i) + 0.5L) / p);
    return x;
}

// Dense inverse by Gauss-Jordan with partial pivoting (row-major n x n).
inline std::vector<ld> invert(std::vector<ld> A, int n) {
    std::vector<ld> B(n * n, 0);
    for (int i = 0; i < n; ++i) B[i * n + i] = 1;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A[r * n + col]) > std::fabs(A[piv * n + col])) piv = r;
        if (A[piv * n + col] == 0) throw std::runtime_error("singular matrix");
        for (int j = 0; j < n; ++j) {
            std::swap(A[col * n + j], A[piv * n + j]);
            std::swap(B[col * n + j], B[piv * n + j]);
        }
        const ld inv = 1 / A[col * n + col];
        for (int j = 0; j < n; ++j) { A[col * n + j] *= inv; B[col * n + j] *= inv; }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const ld f = A[r * n + col];
            if (f == 0) continue;
            for (int j = 0; j < n; ++j) { A[r * n + j] -= f * A[col * n + j]; B[r * n + j] -= f * B[col * n + j]; }
        }
    }
    return B;
}

// Vandermonde V(i,j) = T_j(x_i) at the Chebyshev nodes; returns V^{-1} row-major
// (Vinv[k*p+i]) computed by explicit elimination.
inline std::vector<ld> cheb_vandermonde_inverse(int p) {
    auto x = cheb_nodes(p);
    std::vector<ld> V(p * p);
    for (int i = 0; i < p; ++i) {
        ld t0 = 1, t1 = x[i];
        V[i * p + 0] = 1;
        if (p > 1) V[i * p + 1] = x[i];
        for (int j = 2; j < p; ++j) {
            ld t2 = 2 * x[i] * t1 - t0;
            V[i * p + j] = t2;
            t0 = t1;
            t1 = t2;
        }
    }
    return invert(V, p);
}

inline ld cheb_T(int j, ld x) {
    ld t0 = 1, t1 = x;
    if (j == 0) return 1;
    for (int k = 2; k <= j; ++k) { ld t2 = 2 * x * t1 - t0; t0 = t1; t1 = t2; }
    return t1;
}

}  // namespace oracle
