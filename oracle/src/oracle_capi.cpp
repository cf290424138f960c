// CPU oracle -- test infrastructure only.  Loaded by tests/, smoke() and the
// cpu_baseline leg of bench.py through ctypes; never linked into the product.
#include "oewald.hpp"
#include "otree.hpp"

#include <cstdio>
#include <cstring>
#include <exception>
#include <omp.h>

using namespace oracle;

namespace {
struct Handle {
    int is_float;
    void *tree;
};
template <typename Real>
Tree<Real> *T(Handle *h) { return static_cast<Tree<Real> *>(h->tree); }
}  // namespace

extern "C" {

void *oracle_create(int is_float, const double *cell, int N, const double *pos, int n_digits, int npl, int nz) {
    try {
        auto *h = new Handle{is_float, nullptr};
        if (is_float) h->tree = new Tree<float>(cell, pos, N, n_digits, npl, nz);
        else h->tree = new Tree<double>(cell, pos, N, n_digits, npl, nz);
        return h;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle_create: %s\n", e.what());
        return nullptr;
    }
}

void oracle_destroy(void *hv) {
    auto *h = static_cast<Handle *>(hv);
    if (!h) return;
    if (h->is_float) delete T<float>(h); else delete T<double>(h);
    delete h;
}

// stages: bit0 proxy/plane-wave part, bit1 near field, bit2 self correction, bit3 far field
int oracle_evaluate(void *hv, const double *q, int stages, double *pot_out) {
    auto *h = static_cast<Handle *>(hv);
    if (!h) return -1;
    try {
        if (h->is_float) {
            auto p = T<float>(h)->evaluate(q, stages);
            for (size_t i = 0; i < p.size(); ++i) pot_out[i] = p[i];
        } else {
            auto p = T<double>(h)->evaluate(q, stages);
            for (size_t i = 0; i < p.size(); ++i) pot_out[i] = p[i];
        }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle_evaluate: %s\n", e.what());
        return -1;
    }
}

// Quasi-periodic evaluate: pot_out[2i], pot_out[2i+1] = Re, Im at point i (user order).
int oracle_evaluate_complex(void *hv, const double *q, const double *bloch, int stages, double *pot_out) {
    auto *h = static_cast<Handle *>(hv);
    if (!h || !bloch) return -1;
    try {
        if (h->is_float) {
            auto p = T<float>(h)->evaluate_complex(q, bloch, stages);
            for (size_t i = 0; i < p.size(); ++i) pot_out[i] = p[i];
        } else {
            auto p = T<double>(h)->evaluate_complex(q, bloch, stages);
            for (size_t i = 0; i < p.size(); ++i) pot_out[i] = p[i];
        }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle_evaluate_complex: %s\n", e.what());
        return -1;
    }
}

// ints: [n_src, nsall, nboxes, n_levels, p, n_pw, ms, mt, mu, n_leaves, n_list1_total, n_direct_work,
//        n_c2p_work, n_digits]
// dbls: [rc, beta, psi0, c0, g0d2, w0[0]]
}  // extern "C"
template <typename Real>
static void info_impl(Tree<Real> *t, int *ints, double *dbls) {
    int leaves = 0;
    long l1 = 0;
    for (int b = 0; b < t->nboxes; ++b) {
        leaves += t->is_leaf[b];
        l1 += (long)t->list1[b].size();
    }
    int v[] = {t->n_src, t->nsall, t->nboxes, t->n_levels(), t->p, t->n_pw, t->ms, t->mt, t->mu, leaves, (int)l1,
               (int)t->direct_work.size(), (int)t->c2p_work.size(), t->n_digits};
    std::memcpy(ints, v, sizeof(v));
    double d[] = {(double)t->rc, (double)t->beta, t->tab.psi0, t->tab.c0, t->tab.g0d2, t->tab.w0[0]};
    std::memcpy(dbls, d, sizeof(d));
}
extern "C" int oracle_info(void *hv, int *ints, double *dbls) {
    auto *h = static_cast<Handle *>(hv);
    if (!h) return -1;
    if (h->is_float) info_impl(T<float>(h), ints, dbls); else info_impl(T<double>(h), ints, dbls);
    return 0;
}

// Per-box arrays for topology comparison: depth, parent, nchild, src_cnt, trg_cnt, ifpwexp, nlist1 (7*nboxes ints),
// and the sorted-to-user permutations src_perm[nsall], trg_perm[n_src].
template <typename Real>
static void arrays_impl(Tree<Real> *t, int *box7, int *src_perm, int *trg_perm, double *centers) {
    const int n = t->nboxes;
    for (int b = 0; b < n; ++b) {
        box7[7 * b + 0] = t->depth[b];
        box7[7 * b + 1] = t->parent[b];
        box7[7 * b + 2] = t->nchild[b];
        box7[7 * b + 3] = t->src_cnt[b];
        box7[7 * b + 4] = t->trg_cnt[b];
        box7[7 * b + 5] = t->ifpwexp[b];
        box7[7 * b + 6] = (int)t->list1[b].size();
        for (int d = 0; d < 3; ++d) centers[3 * b + d] = t->centers[3 * b + d];
    }
    if (src_perm) std::memcpy(src_perm, t->src_perm.data(), sizeof(int) * t->src_perm.size());
    if (trg_perm) std::memcpy(trg_perm, t->trg_perm.data(), sizeof(int) * t->trg_perm.size());
}
extern "C" {
int oracle_tree_arrays(void *hv, int *box7, int *src_perm, int *trg_perm, double *centers) {
    auto *h = static_cast<Handle *>(hv);
    if (!h) return -1;
    if (h->is_float) arrays_impl(T<float>(h), box7, src_perm, trg_perm, centers);
    else arrays_impl(T<double>(h), box7, src_perm, trg_perm, centers);
    return 0;
}

// ---- tables and special functions
double oracle_pswf(double c, double x) {
    static thread_local double cached_c = -1;
    static thread_local Prolate0 pf;
    if (c != cached_c) { pf = Prolate0(c); cached_c = c; }
    return (double)pf.eval(x);
}
double oracle_pswf_integral(double c, double r) { return (double)Prolate0(c).integral(r); }
void oracle_pswf_intvals(double c, double *out4) {
    ld c0, c1, c2, c4;
    Prolate0(c).intvals(c0, c1, c2, c4);
    out4[0] = (double)c0; out4[1] = (double)c1; out4[2] = (double)c2; out4[3] = (double)c4;
}
void oracle_gauss_legendre(int n, double *x, double *w) {
    std::vector<ld> xx, ww;
    gauss_legendre(n, xx, ww);
    for (int i = 0; i < n; ++i) { x[i] = (double)xx[i]; w[i] = (double)ww[i]; }
}
int oracle_residual_coeffs(int n_digits, double *out) {
    auto c = residual_coeffs(n_digits);
    if (out) std::memcpy(out, c.data(), sizeof(double) * c.size());
    return (int)c.size();
}
double oracle_beta(double eps) { return beta_for_eps(eps); }
// c2p/p2c (8*3*p*p each), Vandermonde inverse (p*p, row k col i), pw2poly for a level (2*n_pw*p, re/im interleaved)
int oracle_cheb_tables(int p, double *c2p, double *p2c, double *vinv) {
    Tables t;
    (void)t;
    auto V = cheb_vandermonde_inverse(p);
    if (vinv)
        for (int i = 0; i < p * p; ++i) vinv[i] = (double)V[i];
    // reuse the table builder for c2p/p2c with a dummy configuration of order p
    auto xs = cheb_nodes(p);
    for (int s = 0; s < 2; ++s) {
        for (int k = 0; k < p; ++k)
            for (int j = 0; j < p; ++j) {
                ld acc = 0;
                for (int i = 0; i < p; ++i) acc += V[k * p + i] * cheb_T(j, 0.5L * xs[i] + (s == 0 ? -0.5L : 0.5L));
                for (int child = 0; child < 8; ++child)
                    for (int d = 0; d < 3; ++d)
                        if (((child >> d) & 1) == s) {
                            if (p2c) p2c[(child * 3 + d) * p * p + k + j * p] = (double)acc;
                            if (c2p) c2p[(child * 3 + d) * p * p + j + k * p] = (double)acc;
                        }
            }
    }
    return 0;
}

// ---- Ewald
int oracle_ewald_direct(const double *cell, int N, const double *pos, const double *q, double s, double *pot,
                        double *energy) {
    try { ewald_direct(cell, N, pos, q, s, pot, energy); return 0; }
    catch (const std::exception &e) { std::fprintf(stderr, "oracle_ewald_direct: %s\n", e.what()); return -1; }
}
int oracle_ewald_cells(const double *cell, int N, const double *pos, const double *q, double s, int M,
                       const int *tidx, double *pot, double alpha) {
    try { ewald_cells(cell, N, pos, q, s, M, tidx, pot, alpha); return 0; }
    catch (const std::exception &e) { std::fprintf(stderr, "oracle_ewald_cells: %s\n", e.what()); return -1; }
}

// OpenMP threads of the oracle's parallel loops (the CPU baseline is timed at all threads and
// at one).  Returns the previous count.
int oracle_set_threads(int n) {
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}

}  // extern "C"
