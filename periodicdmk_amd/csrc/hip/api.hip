// C++ API (PeriodicDMKTreeT, AutoPeriodicDMKTree) and the C ABI of libpdmk.so.
//
// Reference interfaces replaced:
//   PeriodicDMKTreeT ctor / evaluate / evaluate_timed ... cpp/src/periodic_dmk.cpp:141-308
//   AutoPeriodicDMKTree .................................. cpp/src/periodic_dmk.cpp:319-357
//   pdmk_* C ABI ......................................... cpp/src/pdmk_capi.cpp:20-166
#include <pdmk/pdmk_capi.h>
#include <pdmk/pdmk_gpu.h>
#include <pdmk/periodic_dmk.hpp>

#include "ewald.hpp"
#include "plan.hpp"
#include "../host/tree_build.hpp"

#include <array>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <optional>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace pdmk {

template <typename Real, int DIM>
PeriodicDMKTreeT<Real, DIM>::PeriodicDMKTreeT(const MatND<DIM> &cell, const std::vector<VecND<DIM>> &positions,
                                              int n_digits, int n_per_leaf, int nz,
                                              std::optional<std::array<double, DIM>> bloch) {
    double c[9];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) c[i + 3 * j] = cell[j][i];
    std::vector<double> pos(3 * positions.size());
    for (size_t i = 0; i < positions.size(); ++i)
        for (int d = 0; d < 3; ++d) pos[3 * i + d] = positions[i][d];
    p_ = std::make_unique<gpu::Plan<Real>>(c, pos.data(), (int)positions.size(), n_digits, n_per_leaf, nz,
                                           bloch.has_value() ? bloch->data() : nullptr);
    n_src_ = (int)positions.size();
    rc_ = p_->rc;
}

template <typename Real, int DIM>
PeriodicDMKTreeT<Real, DIM>::~PeriodicDMKTreeT() = default;

template <typename Real, int DIM>
std::vector<Real> PeriodicDMKTreeT<Real, DIM>::evaluate(const std::vector<double> &charges) {
    if (p_->qp) throw std::logic_error("evaluate: bloch is set - use evaluate_complex()");
    if ((int)charges.size() != n_src_) throw std::invalid_argument("charges size != n_src");
    std::vector<Real> pot(n_src_);
    p_->evaluate_host(charges.data(), pot.data(), nullptr);
    return pot;
}

template <typename Real, int DIM>
std::vector<Real> PeriodicDMKTreeT<Real, DIM>::evaluate_timed(const std::vector<double> &charges, EvalTiming &timing) {
    if (p_->qp) throw std::logic_error("evaluate_timed: bloch is set - use evaluate_complex()");
    if ((int)charges.size() != n_src_) throw std::invalid_argument("charges size != n_src");
    std::vector<Real> pot(n_src_);
    double t[3] = {0, 0, 0};
    p_->evaluate_host(charges.data(), pot.data(), t);
    timing.nufft_ms = t[0];
    timing.dmk_ms = t[1];
    timing.direct_ms = t[2];
    timing.build_ms = 0;
    return pot;
}

// cpp/src/periodic_dmk.cpp:251-277: logic_error without bloch, invalid_argument for a size
// mismatch, (Re, Im) per point.
template <typename Real, int DIM>
std::vector<std::complex<Real>> PeriodicDMKTreeT<Real, DIM>::evaluate_complex(const std::vector<double> &charges) {
    if (!p_->qp) throw std::logic_error("evaluate_complex: bloch not set at construction");
    if ((int)charges.size() != n_src_) throw std::invalid_argument("charges size != n_src");
    std::vector<Real> flat(2 * (size_t)n_src_);
    p_->evaluate_complex_host(charges.data(), flat.data());
    std::vector<std::complex<Real>> out(n_src_);
    for (int i = 0; i < n_src_; ++i) out[i] = std::complex<Real>(flat[2 * (size_t)i], flat[2 * (size_t)i + 1]);
    return out;
}

template <typename Real, int DIM>
void PeriodicDMKTreeT<Real, DIM>::evaluate_device(const double *d_charges, Real *d_pot, void *stream) {
    if (p_->qp) throw std::logic_error("evaluate: bloch is set - use evaluate_complex()");
    p_->evaluate(d_charges, d_pot, static_cast<hipStream_t>(stream));
}

template class PeriodicDMKTreeT<double, 3>;
template class PeriodicDMKTreeT<float, 3>;

AutoPeriodicDMKTree::AutoPeriodicDMKTree(const Mat3 &cell, const std::vector<Vec3> &positions, int n_digits,
                                         int n_per_leaf, int nz) {
    if (n_digits <= 6) {
        single_precision_ = true;
        impl_ = std::make_unique<PeriodicDMKTreeT<float, 3>>(cell, positions, n_digits, n_per_leaf, nz);
    } else {
        single_precision_ = false;
        impl_ = std::make_unique<PeriodicDMKTreeT<double, 3>>(cell, positions, n_digits, n_per_leaf, nz);
    }
}

std::variant<std::vector<float>, std::vector<double>> AutoPeriodicDMKTree::evaluate(const std::vector<double> &q) {
    if (single_precision_) return std::get<std::unique_ptr<PeriodicDMKTreeT<float, 3>>>(impl_)->evaluate(q);
    return std::get<std::unique_ptr<PeriodicDMKTreeT<double, 3>>>(impl_)->evaluate(q);
}

AutoPeriodicDMKTree make_periodic_dmk_tree(const Mat3 &cell, const std::vector<Vec3> &positions, int n_digits,
                                           int n_per_leaf, int nz) {
    return AutoPeriodicDMKTree(cell, positions, n_digits, n_per_leaf, nz);
}

}  // namespace pdmk

// ====================================================================== C ABI
namespace {

// Handles carry a precision tag so that device-pointer entry points can check them.
// A captured graph holds a handle's device buffers, streams and events: `graphs` counts the
// live graphs that use it, and a destroy call while one lives is deferred (`dead`) until the
// last of them is destroyed, so a graph never replays into freed memory.
struct Handle {
    int is_float;
    void *tree;
    int graphs = 0;
    bool dead = false;
};

void free_handle(Handle *H) {
    if (H->is_float)
        delete static_cast<pdmk::PeriodicDMKTreeT<float, 3> *>(H->tree);
    else
        delete static_cast<pdmk::PeriodicDMKTreeT<double, 3> *>(H->tree);
    delete H;
}

template <typename Real>
pdmk::PeriodicDMKTreeT<Real, 3> *tree_of(void *h) {
    auto *H = static_cast<Handle *>(h);
    return static_cast<pdmk::PeriodicDMKTreeT<Real, 3> *>(H->tree);
}

pdmk::Mat3 to_mat3(const double *c) {
    pdmk::Mat3 h;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) h[j][i] = c[i + 3 * j];
    return h;
}

std::vector<pdmk::Vec3> to_positions(const double *p, int N) {
    if (N < 0) throw std::invalid_argument("negative N");
    std::vector<pdmk::Vec3> v(N);
    for (int i = 0; i < N; ++i) v[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    return v;
}

template <typename Real>
void *create_impl(const double *cell, int N, const double *pos, int n_digits, int n_per_leaf, int nz,
                  const double *bloch = nullptr) {
    try {
        if (!cell || !pos) throw std::invalid_argument("null cell or positions");
        std::optional<std::array<double, 3>> b;
        if (bloch) b = std::array<double, 3>{bloch[0], bloch[1], bloch[2]};
        auto *t = new pdmk::PeriodicDMKTreeT<Real, 3>(to_mat3(cell), to_positions(pos, N), n_digits, n_per_leaf, nz, b);
        return new Handle{sizeof(Real) == 4, t};
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_create: %s\n", e.what());
        return nullptr;
    }
}

template <typename Real, typename Out>
int evaluate_impl(void *h, const double *charges, Out *pot_out) {
    if (!h) return -1;
    try {
        if (static_cast<Handle *>(h)->is_float != (sizeof(Real) == 4)) throw std::invalid_argument("handle precision mismatch");
        auto *t = tree_of<Real>(h);
        const int N = t->n_points();
        std::vector<double> q(charges, charges + N);
        auto pot = t->evaluate(q);
        for (int i = 0; i < N; ++i) pot_out[i] = static_cast<Out>(pot[i]);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_evaluate: %s\n", e.what());
        return -1;
    }
}

template <typename Real, typename Out>
int evaluate_complex_impl(void *h, const double *charges, Out *pot_out) {
    if (!h || !charges || !pot_out) return -1;
    try {
        if (static_cast<Handle *>(h)->is_float != (sizeof(Real) == 4)) throw std::invalid_argument("handle precision mismatch");
        auto *t = tree_of<Real>(h);
        const int N = t->n_points();
        std::vector<double> q(charges, charges + N);
        auto pot = t->evaluate_complex(q);
        for (int i = 0; i < N; ++i) {
            pot_out[2 * (size_t)i] = static_cast<Out>(pot[i].real());
            pot_out[2 * (size_t)i + 1] = static_cast<Out>(pot[i].imag());
        }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_evaluate_complex: %s\n", e.what());
        return -1;
    }
}

template <typename Real>
void destroy_impl(void *h) {
    if (!h) return;
    auto *H = static_cast<Handle *>(h);
    if (H->graphs > 0) {
        H->dead = true;  // freed with the last graph that uses it
        return;
    }
    free_handle(H);
}

// A captured list of evaluate() calls on device pointers.  Trees are dealt to n_branches
// branches in order of first appearance; each branch's calls (one tree's calls in their
// order) are captured into a graph of its own, launched on the branch's stream, so the
// branches run concurrently on the device's hardware queues.  A launch forks the branches
// from the caller's stream and joins them back.
struct BatchGraph {
    hipStream_t origin = nullptr;
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> ev;  // ev[0]: fork; ev[1 + k]: join of branch k
    std::vector<hipGraph_t> graph;
    std::vector<hipGraphExec_t> exec;
    std::vector<Handle *> users;  // distinct handles the graph captured (each counted once)
    ~BatchGraph() {
        if (origin) (void)hipStreamSynchronize(origin);
        for (hipStream_t b : branch)
            if (b) (void)hipStreamSynchronize(b);
        for (hipGraphExec_t x : exec)
            if (x) (void)hipGraphExecDestroy(x);
        for (hipGraph_t g : graph)
            if (g) (void)hipGraphDestroy(g);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t b : branch)
            if (b) (void)hipStreamDestroy(b);
        if (origin) (void)hipStreamDestroy(origin);
        for (Handle *H : users)
            if (--H->graphs == 0 && H->dead) free_handle(H);
    }
    void launch(hipStream_t s) {
        using pdmk::gpu::check;
        check(hipEventRecord(ev[0], s), "hipEventRecord");
        for (size_t k = 0; k < exec.size(); ++k) {
            check(hipStreamWaitEvent(branch[k], ev[0], 0), "hipStreamWaitEvent");
            check(hipGraphLaunch(exec[k], branch[k]), "hipGraphLaunch");
            check(hipEventRecord(ev[1 + k], branch[k]), "hipEventRecord");
            check(hipStreamWaitEvent(s, ev[1 + k], 0), "hipStreamWaitEvent");
        }
    }
};

BatchGraph *make_batch_graph(int n_calls, void *const *handles, const double *const *d_charges, void *const *d_pots,
                             int n_branches) {
    using pdmk::gpu::check;
    if (n_calls <= 0 || n_branches <= 0 || !handles || !d_charges || !d_pots)
        throw std::invalid_argument("need n_calls > 0, n_branches > 0 and the three arrays");
    for (int i = 0; i < n_calls; ++i) {
        auto *H = static_cast<Handle *>(handles[i]);
        if (!H || !d_charges[i] || !d_pots[i]) throw std::invalid_argument("null handle or buffer in call list");
        const bool qp = H->is_float ? tree_of<float>(H)->plan().qp : tree_of<double>(H)->plan().qp;
        if (qp) throw std::invalid_argument("quasi-periodic handle in call list");
    }
    std::unordered_map<void *, int> branch_of;
    std::vector<std::vector<int>> calls_of(n_branches);
    for (int i = 0; i < n_calls; ++i) {
        const int k = branch_of.emplace(handles[i], (int)branch_of.size() % n_branches).first->second;
        calls_of[k].push_back(i);
    }
    while (!calls_of.empty() && calls_of.back().empty()) calls_of.pop_back();  // fewer trees than branches
    const int nb = (int)calls_of.size();
    auto G = std::make_unique<BatchGraph>();
    for (const auto &kv : branch_of) {
        auto *H = static_cast<Handle *>(kv.first);
        if (H->dead) throw std::invalid_argument("destroyed handle in call list");
        ++H->graphs;
        G->users.push_back(H);
    }
    check(hipStreamCreate(&G->origin), "hipStreamCreate");
    G->branch.assign(nb, nullptr);
    for (auto &b : G->branch) check(hipStreamCreate(&b), "hipStreamCreate");
    G->ev.assign(nb + 1, nullptr);
    for (auto &e : G->ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    G->graph.assign(nb, nullptr);
    G->exec.assign(nb, nullptr);
    for (int k = 0; k < nb; ++k) {
        check(hipStreamBeginCapture(G->branch[k], hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            for (int i : calls_of[k]) {
                auto *H = static_cast<Handle *>(handles[i]);
                if (H->is_float)
                    tree_of<float>(H)->evaluate_device(d_charges[i], static_cast<float *>(d_pots[i]), G->branch[k]);
                else
                    tree_of<double>(H)->evaluate_device(d_charges[i], static_cast<double *>(d_pots[i]), G->branch[k]);
            }
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(G->branch[k], &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(G->branch[k], &G->graph[k]), "hipStreamEndCapture");
        check(hipGraphInstantiate(&G->exec[k], G->graph[k], nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    return G.release();
}

template <typename F>
int guarded(const char *what, F &&f) {
    try {
        return f();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s: %s\n", what, e.what());
        return -1;
    }
}

}  // namespace

extern "C" {

int pdmk_ewald_nufft(const double *cell, int N, const double *pos, const double *charges, double s, double *pot_out,
                     double *energy_out) {
    try {
        if (!cell || !pos || !charges || !pot_out) throw std::invalid_argument("null argument");
        pdmk::gpu::ewald_sum(cell, N, pos, charges, s, pot_out, energy_out);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_ewald_nufft: %s\n", e.what());
        return -1;
    }
}

int pdmk_ewald_nufft_noalloc(const double *cell, int N, const double *pos, const double *charges, double s,
                             double *pot_out, double *energy_out) {
    try {
        if (!cell || !pos || !charges || !pot_out) throw std::invalid_argument("null argument");
        pdmk::gpu::ewald_sum_noalloc(cell, N, pos, charges, s, pot_out, energy_out);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_ewald_nufft_noalloc: %s\n", e.what());
        return -1;
    }
}

unsigned long long pdmk_ewald_noalloc_scratch_bytes(void) { return pdmk::gpu::ewald_scratch_bytes(); }

int pdmk_gpu_ewald_targets(const double *cell, int N, const double *pos, const double *charges, int M,
                           const int *target_indices, double s, double *pot_out) {
    try {
        if (!cell || !pos || !charges || (M > 0 && (!target_indices || !pot_out)))
            throw std::invalid_argument("null argument");
        pdmk::gpu::ewald_sum_targets(cell, N, pos, charges, M, target_indices, s, pot_out);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_gpu_ewald_targets: %s\n", e.what());
        return -1;
    }
}

int pdmk_optimize_cubic_cover(const double *cell, int nz, int *m, int *n, int *p, double *e1, double *e2, double *e3,
                              double *s, double *cost) {
    if (nz <= 0) return 1;
    try {
        if (!cell || !m || !n || !p || !e1 || !e2 || !e3 || !s || !cost)
            throw std::invalid_argument("null argument");
        const pdmk::host::CubicCover c = pdmk::host::cubic_cover(cell, nz);
        *m = c.m;
        *n = c.n;
        *p = c.p;
        for (int d = 0; d < 3; ++d) {
            e1[d] = c.e1[d];
            e2[d] = c.e2[d];
            e3[d] = c.e3[d];
        }
        *s = c.s;
        *cost = c.cost;
        return 0;
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "pdmk_optimize_cubic_cover: %s\n", ex.what());
        return -1;
    }
}

void *pdmk_create_double(const double *cell, int N, const double *pos, int n_digits, int n_per_leaf, int nz) {
    return create_impl<double>(cell, N, pos, n_digits, n_per_leaf, nz);
}
int pdmk_evaluate_double(void *h, int, const double *, const double *charges, double *pot_out) {
    return evaluate_impl<double, double>(h, charges, pot_out);
}
void pdmk_destroy_double(void *h) { destroy_impl<double>(h); }

void *pdmk_create_float(const double *cell, int N, const double *pos, int n_digits, int n_per_leaf, int nz) {
    return create_impl<float>(cell, N, pos, n_digits, n_per_leaf, nz);
}
int pdmk_evaluate_float(void *h, int, const double *, const double *charges, float *pot_out) {
    return evaluate_impl<float, float>(h, charges, pot_out);
}
void pdmk_destroy_float(void *h) { destroy_impl<float>(h); }

// ---------------------------------------------------------------- quasi-periodic
void *pdmk_create_qp_double(const double *cell, int N, const double *pos, int n_digits, int n_per_leaf, int nz,
                            const double *bloch) {
    if (!bloch) return nullptr;
    return create_impl<double>(cell, N, pos, n_digits, n_per_leaf, nz, bloch);
}
void *pdmk_create_qp_float(const double *cell, int N, const double *pos, int n_digits, int n_per_leaf, int nz,
                           const double *bloch) {
    if (!bloch) return nullptr;
    return create_impl<float>(cell, N, pos, n_digits, n_per_leaf, nz, bloch);
}
int pdmk_evaluate_complex_double(void *h, int, const double *, const double *charges, double *pot_out) {
    return evaluate_complex_impl<double, double>(h, charges, pot_out);
}
int pdmk_evaluate_complex_float(void *h, int, const double *, const double *charges, float *pot_out) {
    return evaluate_complex_impl<float, float>(h, charges, pot_out);
}
int pdmk_gpu_evaluate_complex(void *h, const double *d_charges, void *d_pot, int stages) {
    if (!h) return -1;
    return guarded("pdmk_gpu_evaluate_complex", [&] {
        hipStream_t s = nullptr;
        if (static_cast<Handle *>(h)->is_float) {
            auto &pl = tree_of<float>(h)->plan();
            pl.evaluate_complex(d_charges, static_cast<float *>(d_pot), nullptr, stages);
            s = pl.stream;
        } else {
            auto &pl = tree_of<double>(h)->plan();
            pl.evaluate_complex(d_charges, static_cast<double *>(d_pot), nullptr, stages);
            s = pl.stream;
        }
        if (hipStreamSynchronize(s) != hipSuccess) throw std::runtime_error("hipStreamSynchronize failed");
        return 0;
    });
}

// ---------------------------------------------------------------- device memory helpers
int pdmk_gpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void *pdmk_gpu_malloc(unsigned long long bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        std::fprintf(stderr, "pdmk_gpu_malloc: hipMalloc of %llu bytes failed\n", bytes);
        return nullptr;
    }
    return p;
}

int pdmk_gpu_free(void *p) { return (p && hipFree(p) != hipSuccess) ? -1 : 0; }

int pdmk_gpu_memcpy(void *dst, const void *src, unsigned long long bytes, int kind) {
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                      : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    if (kind < 1 || kind > 3) return -1;
    const hipError_t e = hipMemcpy(dst, src, bytes, k);
    if (e != hipSuccess) {
        std::fprintf(stderr, "pdmk_gpu_memcpy: %s\n", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

int pdmk_gpu_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }

// ---------------------------------------------------------------- device-resident entry points
int pdmk_gpu_n_points(void *h) {
    if (!h) return -1;
    auto *H = static_cast<Handle *>(h);
    return H->is_float ? tree_of<float>(h)->n_points() : tree_of<double>(h)->n_points();
}

int pdmk_gpu_evaluate_double(void *h, const double *d_charges, double *d_pot, void *stream) {
    if (!h || static_cast<Handle *>(h)->is_float) return -1;
    return guarded("pdmk_gpu_evaluate_double", [&] {
        tree_of<double>(h)->evaluate_device(d_charges, d_pot, stream);
        return 0;
    });
}

int pdmk_gpu_evaluate_float(void *h, const double *d_charges, float *d_pot, void *stream) {
    if (!h || !static_cast<Handle *>(h)->is_float) return -1;
    return guarded("pdmk_gpu_evaluate_float", [&] {
        tree_of<float>(h)->evaluate_device(d_charges, d_pot, stream);
        return 0;
    });
}

void *pdmk_gpu_graph_create(int n_calls, void *const *handles, const double *const *d_charges, void *const *d_pots,
                            int n_branches) {
    try {
        return make_batch_graph(n_calls, handles, d_charges, d_pots, n_branches);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_gpu_graph_create: %s\n", e.what());
        return nullptr;
    }
}

int pdmk_gpu_set_single_stream(void *h, int on) {
    if (!h) return -1;
    if (static_cast<Handle *>(h)->is_float) tree_of<float>(h)->plan().single_stream = on != 0;
    else tree_of<double>(h)->plan().single_stream = on != 0;
    return 0;
}

int pdmk_gpu_graph_launch(void *g, void *stream) {
    if (!g) return -1;
    return guarded("pdmk_gpu_graph_launch", [&] {
        auto *G = static_cast<BatchGraph *>(g);
        G->launch(stream ? static_cast<hipStream_t>(stream) : G->origin);
        return 0;
    });
}

int pdmk_gpu_graph_destroy(void *g) {
    if (!g) return -1;
    delete static_cast<BatchGraph *>(g);
    return 0;
}

int pdmk_gpu_evaluate_timed(void *h, const double *d_charges, void *d_pot, double *times_ms) {
    if (!h || !times_ms) return -1;
    return guarded("pdmk_gpu_evaluate_timed", [&] {
        if (static_cast<Handle *>(h)->is_float)
            tree_of<float>(h)->plan().evaluate(d_charges, static_cast<float *>(d_pot), nullptr, 15, times_ms);
        else
            tree_of<double>(h)->plan().evaluate(d_charges, static_cast<double *>(d_pot), nullptr, 15, times_ms);
        return 0;
    });
}

int pdmk_gpu_evaluate_stages(void *h, const double *d_charges, void *d_pot, int stages) {
    if (!h) return -1;
    return guarded("pdmk_gpu_evaluate_stages", [&] {
        double t[3];
        if (static_cast<Handle *>(h)->is_float)
            tree_of<float>(h)->plan().evaluate(d_charges, static_cast<float *>(d_pot), nullptr, stages, t);
        else
            tree_of<double>(h)->plan().evaluate(d_charges, static_cast<double *>(d_pot), nullptr, stages, t);
        return 0;
    });
}

int pdmk_gpu_stats(void *h, long long *o) {
    if (!h || !o) return -1;
    const pdmk::gpu::Stats &s =
        static_cast<Handle *>(h)->is_float ? tree_of<float>(h)->plan().stats : tree_of<double>(h)->plan().stats;
    const long long v[16] = {s.n_src,    s.nsall,     s.nboxes,          s.n_levels,    s.p,        s.n_pw,
                             s.ms,       s.mt,        s.mu,              s.n_leaves,    s.n_list1,  s.n_direct_leaves,
                             s.n_far_modes, s.n_digits, s.is_float,      s.n_nf_steps};
    std::memcpy(o, v, sizeof(v));
    return 0;
}

int pdmk_gpu_evaluate_many(void *h, int n_vec, const void *d_charges, void *d_pot, void *stream) {
    if (!h || (n_vec > 0 && (!d_charges || !d_pot))) return -1;
    return guarded("pdmk_gpu_evaluate_many", [&] {
        auto *H = static_cast<Handle *>(h);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        if (H->is_float)
            tree_of<float>(h)->plan().evaluate_many(n_vec, static_cast<const double *>(d_charges),
                                                    static_cast<float *>(d_pot), st);
        else
            tree_of<double>(h)->plan().evaluate_many(n_vec, static_cast<const double *>(d_charges),
                                                     static_cast<double *>(d_pot), st);
        return 0;
    });
}

int pdmk_gpu_autotune(void *h, int repeats, double *out8) {
    if (!h) return -1;
    return guarded("pdmk_gpu_autotune", [&] {
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        auto run = [&](auto &pl) {
            pl.autotune(repeats, &v[0], &v[1]);
            const pdmk::gpu::Knobs &k = pl.knobs;
            v[2] = k.near_wpb;
            v[3] = k.near_late;
            v[4] = k.p2pw_bk;
            v[5] = k.pw2p_bk;
            v[6] = k.t3_div;
            v[7] = pl.knobs_cached ? 1 : 0;
        };
        if (static_cast<Handle *>(h)->is_float) run(tree_of<float>(h)->plan());
        else run(tree_of<double>(h)->plan());
        if (out8) std::memcpy(out8, v, sizeof(v));
        return 0;
    });
}

void pdmk_gpu_tune_cache_clear(void) { pdmk::gpu::tune_cache_clear(); }

int pdmk_gpu_knobs(void *h, const int *set5, int *get5) {
    if (!h) return -1;
    return guarded("pdmk_gpu_knobs", [&] {
        auto apply = [&](auto &pl) {
            pdmk::gpu::Knobs &k = pl.knobs;
            if (set5) {
                const int wpb = set5[0], late = set5[1], b1 = set5[2], b2 = set5[3], div = set5[4];
                auto ok_bk = [](int b) { return b == 0 || b == 16 || b == 32 || b == 64; };
                if (!(wpb == 0 || wpb == 2 || wpb == 4 || wpb == 8 || wpb == 16) || late < -1 || late > 1 ||
                    !ok_bk(b1) || !ok_bk(b2) || !(div == 1 || div == 2 || div == 4))
                    throw std::invalid_argument("knob value outside its candidates");
                k = pdmk::gpu::Knobs{wpb, late, b1, b2, div};
            }
            if (get5) {
                const int v[5] = {k.near_wpb, k.near_late, k.p2pw_bk, k.pw2p_bk, k.t3_div};
                std::memcpy(get5, v, sizeof(v));
            }
        };
        if (static_cast<Handle *>(h)->is_float) apply(tree_of<float>(h)->plan());
        else apply(tree_of<double>(h)->plan());
        return 0;
    });
}

int pdmk_gpu_build_info(void *h, double *o) {
    if (!h || !o) return -1;
    const bool f = static_cast<Handle *>(h)->is_float;
    const double *ms = f ? tree_of<float>(h)->plan().build_ms : tree_of<double>(h)->plan().build_ms;
    const bool gpu = f ? tree_of<float>(h)->plan().tree_on_gpu : tree_of<double>(h)->plan().tree_on_gpu;
    const int folded = f ? tree_of<float>(h)->plan().n_folded : tree_of<double>(h)->plan().n_folded;
    const int slots = f ? tree_of<float>(h)->plan().n_up : tree_of<double>(h)->plan().n_up;
    const bool by_level = f ? tree_of<float>(h)->plan().pw_by_level : tree_of<double>(h)->plan().pw_by_level;
    const bool rows = f ? tree_of<float>(h)->plan().nf_rows : tree_of<double>(h)->plan().nf_rows;
    const double v[8] = {ms[0], ms[1], ms[2], gpu ? 1.0 : 0.0, (double)folded, (double)slots, by_level ? 1.0 : 0.0,
                         rows ? 1.0 : 0.0};
    std::memcpy(o, v, sizeof(v));
    return 0;
}

int pdmk_gpu_work(void *h, double *o) {
    if (!h || !o) return -1;
    const pdmk::gpu::Work &w =
        static_cast<Handle *>(h)->is_float ? tree_of<float>(h)->plan().work : tree_of<double>(h)->plan().work;
    const double v[8] = {w.nf_pairs_support, w.nf_pairs_candidate, w.nf_flops, w.transform_flops,
                         w.ff_mode_points,   w.ff_flops,           w.on_bytes, w.nf_bytes};
    std::memcpy(o, v, sizeof(v));
    return 0;
}

int pdmk_gpu_enable_kernel_timing(void *h, int slots) {
    if (!h) return -1;
    return guarded("pdmk_gpu_enable_kernel_timing", [&] {
        if (static_cast<Handle *>(h)->is_float) tree_of<float>(h)->plan().enable_kernel_timing(slots);
        else tree_of<double>(h)->plan().enable_kernel_timing(slots);
        return 0;
    });
}

int pdmk_gpu_kernel_times(void *h, double *avg_ms5) {
    if (!h || !avg_ms5) return -1;
    return guarded("pdmk_gpu_kernel_times", [&] {
        int n = 0;
        if (static_cast<Handle *>(h)->is_float) tree_of<float>(h)->plan().kernel_times(avg_ms5, &n);
        else tree_of<double>(h)->plan().kernel_times(avg_ms5, &n);
        return n;
    });
}

int pdmk_gpu_tree_arrays(void *h, int *box7, double *centers, int *src_perm, int *trg_perm) {
    if (!h) return -1;
    auto copy = [&](auto &pl) {
        if (box7) std::memcpy(box7, pl.box7.data(), sizeof(int) * pl.box7.size());
        if (centers) std::memcpy(centers, pl.box_centers.data(), sizeof(double) * pl.box_centers.size());
        if (src_perm) std::memcpy(src_perm, pl.src_perm_host.data(), sizeof(int) * pl.src_perm_host.size());
        if (trg_perm) std::memcpy(trg_perm, pl.trg_perm_host.data(), sizeof(int) * pl.trg_perm_host.size());
    };
    if (static_cast<Handle *>(h)->is_float) copy(tree_of<float>(h)->plan());
    else copy(tree_of<double>(h)->plan());
    return 0;
}

}  // extern "C"
