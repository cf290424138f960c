// Host-side DMK tables; see tables.hpp for the reference behaviour followed.
#include "tables.hpp"

#include "residual_coeffs.hpp"

#include <cmath>
#include <stdexcept>

namespace pdmk::host {

namespace {

constexpr double kPi = 3.14159265358979323846;


double cheb(int k, double x) {
    double a = 1, b = x;
    if (k == 0) return 1;
    for (int i = 1; i < k; ++i) { const double c = 2 * x * b - a; a = b; b = c; }
    return b;
}

struct Params { double beta; int n_pw, p; };

Params params_for(double eps) {
    // bracket tables for 3/6/9/12 digits (util.cpp:22-27, tree.hpp:40-44)
    if (eps >= 1e-3) return {7.2462000846862793, 13, 9};
    if (eps >= 1e-6) return {13.739999771118164, 27, 18};
    if (eps >= 1e-9) return {20.736000061035156, 39, 28};
    if (eps >= 1e-12) return {27.870000839233398, 55, 38};
    throw std::invalid_argument("PeriodicDMK: n_digits must be in 1..12");
}

}  // namespace

// Gauss-Legendre nodes (ascending) and weights on [-1, 1].
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 1;
        for (int it = 0; it < 60; ++it) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            const double dz = p1 / dp;
            z -= dz;
            if (std::abs(dz) < 1e-16) break;
        }
        x[n - 1 - i] = z;
        x[i] = -z;
        w[i] = w[n - 1 - i] = 2.0 / ((1 - z * z) * dp * dp);
    }
}

// The prolate operator -(1-x^2) d^2/dx^2 + 2x d/dx + c^2 x^2 restricted to even
// normalised Legendre functions is symmetric tridiagonal; psi_0 is the eigenvector of
// its smallest eigenvalue (bisection on the Sturm count, then inverse iteration).
Pswf::Pswf(double c_) : c(c_) {
    if (c <= 0) return;
    const int m = 48 + (int)(1.2 * c);
    std::vector<double> d(m), e(m);
    for (int i = 0; i < m; ++i) {
        const double k = 2.0 * i;
        d[i] = k * (k + 1) + c * c * (2 * k * (k + 1) - 1) / ((2 * k + 3) * (2 * k - 1));
        e[i] = c * c * (k + 1) * (k + 2) / ((2 * k + 3) * std::sqrt((2 * k + 1) * (2 * k + 5)));
    }
    auto below = [&](double lam) {  // number of eigenvalues < lam
        int n = 0;
        double q = d[0] - lam;
        n += q < 0;
        for (int i = 1; i < m; ++i) {
            q = d[i] - lam - e[i - 1] * e[i - 1] / (q != 0 ? q : 1e-300);
            n += q < 0;
        }
        return n;
    };
    double lo = 0, hi = d[0] + std::abs(e[0]);
    while (below(hi) == 0) hi *= 2;
    for (int it = 0; it < 120 && hi - lo > 1e-15 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        (below(mid) >= 1 ? hi : lo) = mid;
    }
    const double lam = lo * (1 - 1e-13);
    coef.assign(m, 1.0);
    std::vector<double> cp(m), rp(m);
    for (int it = 0; it < 5; ++it) {
        double den = d[0] - lam;
        cp[0] = e[0] / den;
        rp[0] = coef[0] / den;
        for (int i = 1; i < m; ++i) {
            den = d[i] - lam - e[i - 1] * cp[i - 1];
            cp[i] = e[i] / den;
            rp[i] = (coef[i] - e[i - 1] * rp[i - 1]) / den;
        }
        coef[m - 1] = rp[m - 1];
        for (int i = m - 2; i >= 0; --i) coef[i] = rp[i] - cp[i] * coef[i + 1];
        double nrm = 0;
        for (double v : coef) nrm += v * v;
        nrm = 1.0 / std::sqrt(nrm);
        for (double &v : coef) v *= nrm;
    }
    if ((*this)(0.0) < 0)
        for (double &v : coef) v = -v;
}

double Pswf::operator()(double x) const {
    double pk = 1, pkm1 = 0, s = 0;
    const int kmax = 2 * (int)coef.size() - 2;
    for (int k = 0; k <= kmax; ++k) {
        if (!(k & 1)) s += coef[k / 2] * std::sqrt(k + 0.5) * pk;
        const double pn = ((2 * k + 1) * x * pk - k * pkm1) / (k + 1);
        pkm1 = pk;
        pk = pn;
    }
    return s;
}

DmkTables make_tables(int n_digits_user, double rc, int nlevels, const double *U) {
    DmkTables T;
    T.eps = std::pow(10.0, -std::max(1, n_digits_user));
    T.n_digits = std::max(1, (int)std::lround(std::log10(1.0 / T.eps) - 0.1));
    const Params P = params_for(T.eps);
    T.beta = P.beta;
    T.n_pw = P.n_pw;
    T.p = P.p;
    T.h = (T.n_pw + 1) / 2;
    T.n_modes = T.n_pw * T.n_pw * T.h;
    const ResidualTable rt = residual_table(T.n_digits);
    if (!rt.c) throw std::invalid_argument("PeriodicDMK: unsupported digit count");
    T.residual.assign(rt.c, rt.c + rt.nc);

    const Pswf pf(T.beta);
    std::vector<double> gx, gw;
    gauss_legendre(128, gx, gw);
    double c0 = 0, c2 = 0;
    std::vector<double> psi_q(gx.size());  // psi0 at the nodes mapped to [0, 1]
    for (size_t i = 0; i < gx.size(); ++i) {
        const double t = 0.5 * (gx[i] + 1), wt = 0.5 * gw[i];
        psi_q[i] = pf(t);
        c0 += wt * psi_q[i];
        c2 += wt * psi_q[i] * t * t;
    }
    T.psi0 = pf(0.0);
    T.c0 = c0;
    T.g0d2 = c2 / c0;

    T.nlevels = nlevels;
    T.boxsize.resize(nlevels + 1);
    for (int l = 0; l <= nlevels; ++l) T.boxsize[l] = std::ldexp(rc, -l);
    T.w0.resize(nlevels + 1);
    for (int l = 0; l < nlevels; ++l) T.w0[l] = T.psi0 / (T.c0 * (l == 0 ? 0.5 * rc : T.boxsize[l]));
    T.w0[nlevels] = T.psi0 / (T.c0 * 0.5 * T.boxsize[nlevels - 1]);

    // level-0 difference kernel FT on k^2 = 0 .. 3 (n_pw/2)^2 (Laplace 3-D)
    const int half = T.n_pw / 2, nf = (T.n_pw - 1) / 2;
    const double bs = rc, hpw = 2 * T.beta / nf / bs, ws = hpw * hpw * hpw * 0.5 / (kPi * kPi);
    const int n_k2 = 3 * half * half + 1;
    std::vector<double> fk(n_k2);
    fk[0] = 0.5 * ws * T.g0d2 * (bs * bs - 0.25 * bs * bs);
    for (int i = 1; i < n_k2; ++i) {
        const double k = std::sqrt((double)i) * hpw;
        double s = 0;
        for (size_t j = 0; j < gx.size(); ++j) {
            const double t = 0.5 * (gx[j] + 1);
            s += (std::cos(k * t * 0.5 * bs) - std::cos(k * t * bs)) * 0.5 * gw[j] * psi_q[j];
        }
        fk[i] = s / T.c0 * ws / (k * k);
    }
    // mode layout m = c (h n) + a n + b: c = x index (n), a = z index (h), b = y index (n)
    T.radial0.resize(T.n_modes);
    for (int a = 0; a < T.h; ++a)
        for (int b = 0; b < T.n_pw; ++b)
            for (int c = 0; c < T.n_pw; ++c) {
                const int ja = a - half, jb = b - half, jc = c - half;
                T.radial0[mode_index(T, a, b, c)] = fk[ja * ja + jb * jb + jc * jc];
            }

    // Chebyshev nodes (ascending) and the closed-form inverse Vandermonde
    // Vinv(k, i) = (k ? 2 : 1)/p * T_k(x_i)
    const int p = T.p;
    std::vector<double> xs(p);
    for (int i = 0; i < p; ++i) xs[i] = std::cos(kPi * ((p - 1 - i) + 0.5) / p);
    std::vector<double> vinv(p * p);
    for (int k = 0; k < p; ++k)
        for (int i = 0; i < p; ++i) vinv[k * p + i] = (k ? 2.0 : 1.0) / p * cheb(k, xs[i]);

    // PW <-> Chebyshev: exp(i (j - half) x (bs/2) hpw) with (bs/2) hpw = beta / nf at every level
    const double dphase = T.beta / nf;
    T.pw2poly.assign(T.n_pw * p, 0);
    T.poly2pw.assign(T.n_pw * p, 0);
    for (int j = 0; j < T.n_pw; ++j)
        for (int k = 0; k < p; ++k) {
            std::complex<double> s = 0;
            for (int i = 0; i < p; ++i) s += std::polar(vinv[k * p + i], (j - half) * xs[i] * dphase);
            T.pw2poly[j + T.n_pw * k] = s;
            T.poly2pw[j + T.n_pw * k] = std::conj(s);
        }

    // shifts: t = 9 kz + 3 ky + kx (k = offset + 1), factor exp(i (j - half) hpw bs (k - 1)) per axis
    const double sphase = 2 * T.beta / nf;  // hpw * boxsize, the same at every level
    T.shift0.assign(27 * (size_t)T.n_modes, 0);
    for (int t = 0; t < 27; ++t) {
        const int kz = t / 9 - 1, ky = (t / 3) % 3 - 1, kx = t % 3 - 1;
        for (int a = 0; a < T.h; ++a)
            for (int b = 0; b < T.n_pw; ++b)
                for (int c = 0; c < T.n_pw; ++c) {
                    const int m = mode_index(T, a, b, c);
                    const double ph = ((a - half) * kz + (b - half) * ky + (c - half) * kx) * sphase;
                    T.shift0[(size_t)t * T.n_modes + m] = std::polar(T.radial0[m], ph);
                }
    }

    // parent -> child Chebyshev re-expansion, child on the minus / plus side
    for (int s = 0; s < 2; ++s) {
        T.p2c[s].assign(p * p, 0);
        for (int k = 0; k < p; ++k)
            for (int j = 0; j < p; ++j) {
                double acc = 0;
                for (int i = 0; i < p; ++i) acc += vinv[k * p + i] * cheb(j, 0.5 * xs[i] + (s ? 0.5 : -0.5));
                T.p2c[s][k + p * j] = acc;
            }
    }

    // far field: truncated lattice sum of the periodised PSWF window
    const double twopi = 2 * kPi, u0 = U[0], u3 = U[3], u4 = U[4], u6 = U[6], u7 = U[7], u8 = U[8];
    const double vol = u0 * u4 * u8, kmax = T.beta / rc;
    const double n1 = u0, n2 = std::sqrt(u3 * u3 + u4 * u4), n3 = std::sqrt(u6 * u6 + u7 * u7 + u8 * u8);
    T.ms = std::max(2, 2 * (int)std::ceil(kmax * n1 / twopi));
    T.mt = std::max(2, 2 * (int)std::ceil(kmax * n2 / twopi));
    T.mu = std::max(2, 2 * (int)std::ceil(kmax * n3 / twopi));
    const double a11 = 1 / u0, a21 = -u3 / (u0 * u4), a22 = 1 / u4;
    const double a31 = (u3 * u7 - u6 * u4) / (u0 * u4 * u8), a32 = -u7 / (u4 * u8), a33 = 1 / u8;
    T.ainv[0] = a11; T.ainv[1] = a21; T.ainv[2] = a22; T.ainv[3] = a31; T.ainv[4] = a32; T.ainv[5] = a33;
    const double dfac = 4 * kPi / (T.psi0 * vol);
    auto weight = [&](int i, int j, int k) {
        if (i < -T.ms / 2 || i > (T.ms - 1) / 2 || j < -T.mt / 2 || j > (T.mt - 1) / 2 || k < -T.mu / 2 ||
            k > (T.mu - 1) / 2)
            return 0.0;
        const double k1 = twopi * a11 * i, k2 = twopi * (a21 * i + a22 * j), k3 = twopi * (a31 * i + a32 * j + a33 * k);
        const double kk = k1 * k1 + k2 * k2 + k3 * k3;
        if (kk < 1e-30) return 0.0;
        const double x = std::sqrt(kk) * rc / T.beta;
        return x <= 1 ? dfac * pf(x) / kk : 0.0;
    };
    for (int k = -T.mu / 2; k <= (T.mu - 1) / 2; ++k)
        for (int j = -T.mt / 2; j <= (T.mt - 1) / 2; ++j)
            for (int i = -T.ms / 2; i <= (T.ms - 1) / 2; ++i) {
                const double w = weight(i, j, k);
                if (w == 0) continue;
                const bool upper = k > 0 || (k == 0 && (j > 0 || (j == 0 && i > 0)));
                const double wc = weight(-i, -j, -k);
                if (!upper && wc != 0) continue;  // folded into its conjugate
                T.mode_ijk.push_back(i);
                T.mode_ijk.push_back(j);
                T.mode_ijk.push_back(k);
                T.mode_w.push_back(upper && wc != 0 ? 2 * w : w);
            }
    return T;
}

}  // namespace pdmk::host
