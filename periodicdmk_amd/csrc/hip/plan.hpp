// Device plan of one periodic DMK tree: everything evaluate() needs, resident in
// HBM, plus the per-stage launch sequence.  Built once by the constructor of
// PeriodicDMKTreeT (host tree + tables are flattened and uploaded); evaluate() only
// moves charges in and potentials out.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

namespace pdmk::gpu {

void check(hipError_t e, const char *what);

// RAII device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            std::swap(p, o.p);
            std::swap(bytes, o.bytes);
        }
        return *this;
    }
    ~DevBuf();
    void alloc(size_t nbytes);
    template <class T> T *as() const { return static_cast<T *>(p); }
    template <class T> void upload(const std::vector<T> &v) {
        alloc(v.size() * sizeof(T));
        if (!v.empty()) check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
};

// Bump allocation over device memory the caller provides (256-byte aligned slices).  In a
// dry run (base == nullptr) it only counts the bytes, returns null pointers and copies
// nothing, so a caller can size a request before it touches the memory.  take() throws
// std::length_error past the capacity.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    bool dry() const { return base == nullptr; }
    void *take(size_t nbytes);
    template <class T> T *take_n(size_t n) { return static_cast<T *>(take(sizeof(T) * n)); }
    // synchronous host -> device copy into a fresh slice
    template <class T> T *put(const std::vector<T> &v) {
        T *d = take_n<T>(v.size());
        if (!dry() && !v.empty()) check(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
        return d;
    }
};

struct Stats {
    long long n_src = 0, nsall = 0, nboxes = 0, n_levels = 0, p = 0, n_pw = 0, ms = 0, mt = 0, mu = 0;
    long long n_leaves = 0, n_list1 = 0, n_direct_leaves = 0, n_far_modes = 0, n_digits = 0, is_float = 0,
              n_nf_steps = 0;
};

struct Work {
    double nf_pairs_support = 0, nf_pairs_candidate = 0, nf_flops = 0, transform_flops = 0;
    double ff_mode_points = 0, ff_flops = 0, on_bytes = 0, nf_bytes = 0;
};

// Truncated lattice Fourier sum in FP64:
//   out[u] = sum_{m in the half space} Re( w_m S(m) e^{+i m.theta_u} ),  S(m) = sum_u q_u e^{-i m.theta_u},
// theta = 2 pi x (fractional coordinates, wrapped to [0, 1)).  Modes (i, j, k) come in lines
// of constant (j, k) with consecutive i; index ranges are [-m/2, (m-1)/2] for m = ms, mt, mu.
// Its arrays are owned buffers, or slices of an Arena when setup() is given one.  The
// structure factors are summed in `blocks` partial sums of point tiles (at most max_blocks),
// then added in block order.
struct FourierSum {
    int n = 0, ms = 0, mt = 0, mu = 0, n_modes = 0, n_chunks = 0, n_lines = 0, blocks = 0;
    DevBuf theta, w, chunk_ijk, chunk_m0, chunk_len, lines, line_base, pidx, Sp, S;
    const double *theta_d = nullptr, *w_d = nullptr;
    const int *chunk_ijk_d = nullptr, *chunk_m0_d = nullptr, *chunk_len_d = nullptr, *line_base_d = nullptr,
              *pidx_d = nullptr;
    const int4 *lines_d = nullptr;
    double2 *Sp_d = nullptr, *S_d = nullptr;
    void setup(const std::vector<double> &theta_host, int n_points, int ms_, int mt_, int mu_,
               const std::vector<int> &mode_ijk, const std::vector<double> &mode_w, Arena *arena = nullptr,
               int max_blocks = 2048);
    void run(const double *d_q, double *d_out, hipStream_t s) const;  // enqueue only
    // S(m) from the setup points, then out[t] at n_eval other points (theta_eval[3t + d],
    // device, same convention); enqueue only
    void run_at(const double *d_q, const double *d_theta_eval, int n_eval, double *d_out, hipStream_t s) const;
    // run() for nv vectors (charges d_Q + v ldq, output d_out + v ldo), each equal to run() on
    // it: partial sums of vc vectors per launch in Sp_many (vc * sp_elems()), the weighted
    // sums of all nv in S_many (nv * s_elems(), padding zero), one evaluation launch
    void run_many(const double *d_Q, long long ldq, int nv, double *d_out, long long ldo, double2 *Sp_many, int vc,
                  double2 *S_many, hipStream_t s) const;
    size_t sp_elems() const { return (size_t)blocks * (size_t)std::max(1, n_modes); }
    size_t s_elems() const { return s_len; }
    size_t s_len = 1;
};

// The same Fourier sum by a non-uniform FFT (spreading with an exponential-of-semicircle
// kernel of width W onto a 2x oversampled grid, grid DFTs, interpolation), accurate to about
// 10^(2-W) relative; deterministic (no atomics).
struct NufftFarField {
    int n = 0, n_modes = 0, W = 0, T = 0, lo = 0, E = 0, SY = 0, SZ = 0, n_units = 0, n_tiles = 0, n_extra_tiles = 0;
    int g[3] = {0, 0, 0}, M[3] = {0, 0, 0}, nt[3] = {0, 0, 0};
    DevBuf order, kval, pbase, unit_ptr, unit_tile, unit_slot, extra_tile, extra_ptr, gp[3], gt[3], gl[3];
    DevBuf qs, scratch, grid, grid2, fwd[3], inv[3], t1, t2, bbox, cbox, pos, cw;
    std::vector<int> position;  // plan position of each user point (run() writes out[position[u]])
    void setup(const std::vector<double> &theta_host, int n_points, int ms_, int mt_, int mu_,
               const std::vector<int> &mode_ijk, const std::vector<double> &mode_w, int width);
    void run(const double *d_q, double *d_out, hipStream_t s) const;  // enqueue only
};

// Quasi-periodic far field, exact (all N x modes products): wave vectors 2 pi U^{-T} m + b
// over the full mode box, complex output added to an interleaved (Re, Im) array.
struct QpFarField {
    int n = 0, ms = 0, mt = 0, mu = 0, M = 0, nblk = 0, tp = 0;
    DevBuf theta, bx, w, Sp, S;
    void setup(const std::vector<double> &theta_host, const std::vector<double> &bx_host, int n_points, int ms_,
               int mt_, int mu_, const std::vector<double> &w_full);
    template <class R> void run(const double *d_q, R *d_out, hipStream_t s) const;  // enqueue only
};

// Launch shapes of evaluate() the autotuner chooses among (0 / -1 = the built-in rule).
struct Knobs {
    int near_wpb = 0;    // near-field wavefronts per workgroup: 0 (4, or 16 for a late start), 4, 8, 16
    int near_late = -1;  // near field after the upward pass: -1 (when p >= 28), 0, 1
    int p2pw_bk = 0;     // proxy -> plane-wave LDS budget, KiB: 0 (per order and precision), 16, 32, 64
    int pw2p_bk = 0;     // plane-wave -> proxy LDS budget, KiB: 0 (per order and precision), 16, 32, 64
    int t3_div = 1;      // tensor transform: output planes per workgroup = the budget's count / t3_div (1, 2)
};

// Forget the tuned knobs of this process (plans built afterwards use the built-in shapes).
void tune_cache_clear();

template <typename Real>
struct Plan {
    // sizes
    int n = 0, nsall = 0, nboxes = 0, nlevels = 0, p = 0, n_pw = 0, nc = 0, n_digits = 0;
    double rc = 0;
    int n_up = 0, n_down = 0;     // proxy slots (upward == outgoing plane-wave slots)
    int ms = 0, mt = 0, mu = 0, n_ff_modes = 0;
    Stats stats;
    Work work;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // concurrent evaluate(): far field and near field on their own streams (default
    // priority), joined before the final combination
    hipStream_t s_far = nullptr, s_near = nullptr;
    hipEvent_t e_in = nullptr, e_gather = nullptr, e_far = nullptr, e_near = nullptr, e_up = nullptr;
    hipEvent_t e_upward = nullptr;  // when set, recorded by the transform chain after the upward pass

    // O(N) data
    DevBuf src, src_user, trg, trg_user, trg_w0, pot_sorted, pot_near, q_dev, pot_dev, far, far_idx;
    // upward
    DevBuf up_slot, up_geom, up_rng, rng_begin, rng_count, rng_off, proxy_up;
    struct Tensor3Launch { int n_items; DevBuf out, add, ptr, in_slot, in_code; };
    // child -> parent per level (deepest first): fused = one item per parent adding its children
    // to the parent's proxy; else per-child transforms into c2p_scratch, then
    // parent_slot[i] += sum of scratch slots child_ptr[i] .. child_ptr[i+1]
    struct C2pLevel { Tensor3Launch t3; bool fused = false; int n_parents = 0; DevBuf parent_slot, child_ptr; };
    std::vector<C2pLevel> c2p_levels;
    DevBuf c2p_scratch;
    // Outgoing plane waves of all up slots (pw_out[slot]), or, when that does not fit in the
    // device memory, of one level's range at a time (pw_out[slot - pw_lo], formed just before
    // the level's shifts); c2p_scratch then shares pw_out's memory (the upward pass is done
    // before the first plane wave is formed, on the same stream).
    bool pw_by_level = false;
    bool shift_groups = true;  // k_pw_shift_grp (p >= 28; PDMK_SHIFT_GROUPS=0 / 1 forces either)
    bool t3_mfma = true;  // FP64 tensor transforms at p = 28 on the matrix cores (PDMK_T3_MFMA=0: vector ALUs)
    // the WIDE shape of k_charge2proxy_mfma (two 4-wave workgroups per box) for trees with at
    // least kC2pWideBoxes proxy slots, else one 16-wave workgroup (PDMK_C2P_WIDE=0 / 1 forces)
    bool c2p_wide = false;
    // evaluate() on the caller's stream only (the stages one after another): for callers that
    // run many trees side by side, where the trees, not the stages, fill the GPU
    bool single_stream = false;
    Real *c2p_buffer() { return pw_by_level ? pw_out.template as<Real>() : c2p_scratch.template as<Real>(); }
    DevBuf mats_rm;                         // parent <-> child matrices: [transposed][minus / plus][p][p]
    // plane waves
    DevBuf pw_out, poly2pw, pw2poly, shift0;  // shift0: FP32 plans
    DevBuf shift_r, shift_e;                   // FP64 plans: radial factor per mode, per-axis phase
    DevBuf poly2pw_f, pw2poly_f;  // folded tables of the two-pass transforms (64 entries a row)
    // pw_lo, pw_cnt: the range of up slots the level's shifts read
    // per level: the items' neighbour lists (CSR of slot, stencil) and the same merged per group
    // of kShiftGroup items (grp_ptr / grp_slot / grp_sten, byte t = stencil for item t or 255)
    struct EvalLevel {
        int level, n_items, pw_lo = 0, pw_cnt = 0;
        DevBuf down, zero, nbr_ptr, nbr_slot, nbr_sten, grp_ptr, grp_slot, grp_sten;
    };
    std::vector<EvalLevel> eval_levels;
    std::vector<Tensor3Launch> p2c_levels;  // one per level (may be empty)
    DevBuf proxy_down;
    int n_leaf_eval = 0;
    DevBuf leaf_down, leaf_geom, leaf_trg;
    // near field: per target leaf (nf_trg, nf_ptr -> list1 entries nf_l1/nf_geom), per
    // wavefront item (nf_items, nf_wbox), per source chunk (nf_cbox); step lists per
    // (item, entry): nf_ie_base[item] + k -> nf_ie_ptr -> source indices nf_steps
    int n_nf_leaves = 0, n_nf_items = 0;
    long long n_nf_steps = 0;
    DevBuf nf_trg, nf_ptr, nf_l1, nf_geom, nf_items, nf_wbox, nf_cbox, coeffs, nf_ie_base, nf_ie_ptr, nf_steps;
    // quad-row step lists (FP32 at 6 digits): per item the row blocks from nf_wave_base[i]
    // (K steps per row in nf_items[i].w), per (item, row) the entry slots from nf_rs_base[i]
    bool nf_rows = false;
    int nf_ne_max = 0;  // most list1 entries of a target leaf
    DevBuf nf_rs_base, nf_rslot, nf_wave_base, nf_row_steps;
    // far field: exact Fourier sum, or the NUFFT for large N x modes
    int far_nufft = 0;
    FourierSum ff;
    NufftFarField nf;
    // incoming plane waves of up to pin_batch boxes of one level
    DevBuf pin;
    int pin_batch = 0;
    // first contractions of the two-pass proxy <-> plane-wave transforms, for up to
    // tp_boxes_p2pw / tp_boxes_pw2p boxes at a time (empty when the plan does not use them)
    DevBuf tp_scratch;
    int tp_boxes_p2pw = 0, tp_boxes_pw2p = 0;

    // bloch (optional, 3 values): quasi-periodic tree; then only evaluate_complex applies
    Plan(const double *cell_colmajor, const double *pos, int N, int n_digits, int n_per_leaf, int nz,
         const double *bloch = nullptr);
    ~Plan();
    void release_handles();  // synchronize, then destroy the streams and events
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;

    // Enqueue evaluate() on `s` (nullptr = own stream).  d_q: FP64 charges (user order),
    // d_pot: Real potentials (user order).  stages: bit0 proxy/PW, bit1 near field,
    // bit2 self correction, bit3 far field.  times (optional, synchronous): EvalTiming split.
    void evaluate(const double *d_q, Real *d_pot, hipStream_t s, int stages = 15, double *times_ms = nullptr);

    // host buffers in, host buffers out
    void evaluate_host(const double *q, Real *pot, double *times_ms = nullptr);

    // evaluate() of n_vec charge vectors (d_Q, d_P: vector-major, n per vector) on `s`.  FP32
    // plans with row step lists take the vectors in passes of 16: the near field of the pass
    // (each pair's residual formed once for the 16 vectors) on s_near, the far fields on s_far
    // (the exact Fourier sum of several vectors per launch, the phases of a workgroup shared by
    // FF_MANY_S / FF_MANY_E vectors) and the proxy / plane-wave chain of
    // several vectors per launch on s, the three beside each other and beside the next pass;
    // other plans call evaluate() per vector.  Each vector's potentials equal evaluate()'s on
    // it bit for bit.  The per-vector buffers are allocated on first use (many_ensure).
    void evaluate_many(int n_vec, const double *d_Q, Real *d_P, hipStream_t s);
    static constexpr int kManyVec = 16;  // vectors per near-field pass
    // vectors per chain launch and per exact far-field launch (sized to a memory budget);
    // 0 until the buffers exist, -1 when they do not fit (evaluate_many then calls evaluate())
    int many_vc_chain = 0, many_vc_far = 0;
    bool many_ensure();
    // two pass slots (b = pass & 1): charges [nsall][kManyVec] (FP32), near [kManyVec][n],
    // far [kManyVec][n]
    DevBuf many_qv[2], many_near[2], many_far[2];
    // chain copies for many_vc_chain vectors, potentials of a pass, exact far-field partial
    // sums for many_vc_far vectors and weighted sums for kManyVec
    DevBuf many_up, many_c2p, many_pw, many_pin, many_down, many_pot, many_sp, many_s;
    // per slot: charges gathered, near field done, far field done, pass finished
    hipEvent_t many_ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

    // Launch shapes: set at construction from the per-device cache when a plan of the same
    // kind (order, precision, size buckets) was tuned before in this process (or in the file
    // PDMK_TUNE_CACHE names); autotune() times evaluate() on this plan over each knob's
    // candidates in turn (coordinate descent, keeping a change only when it is at least 2%
    // faster), stores the result in the cache and returns {ms default, ms tuned}.
    Knobs knobs;
    bool knobs_cached = false;
    std::string tune_key() const;
    void autotune(int repeats, double *ms_default, double *ms_tuned);

    // Quasi-periodic evaluate (bloch given at construction): d_out holds 2N Real, (Re, Im)
    // per point in user order.  Two passes of the proxy / plane-wave chain and the near field
    // (charges times the real, then the imaginary part of each image's phase), the self
    // correction on the real part, then the quasi-periodic far field.
    bool qp = false;
    double bloch[3] = {0, 0, 0};
    DevBuf src_phase;   // (cos, sin) of b.R per sorted source (R: the image's lattice translation)
    QpFarField qpf;
    int qp_comp = -1;   // component of the pass in progress (-1: not quasi-periodic)
    DevBuf pot_dev2;    // 2N Real for the host-buffer path
    void evaluate_complex(const double *d_q, Real *d_out, hipStream_t s, int stages = 15);
    void evaluate_complex_host(const double *q, Real *out);

    // Per-kernel-group timing with HIP events on the stream each group runs on (start and
    // end event per group).  Groups: 0 far field, 1 charge gather, 2 proxy/plane-wave
    // transforms, 3 near field, 4 finalize.  A ring of `slots` calls is kept;
    // kernel_times() averages the recorded calls (after the caller has synchronized).
    static constexpr int kGroups = 5;
    void enable_kernel_timing(int slots);
    int kernel_times(double *avg_ms, int *n_calls);
    std::vector<hipEvent_t> kev;
    int kslots = 0;
    long kcall = 0;

    // construction: total, tree (geometry, images, tree) and far-field plan milliseconds;
    // whether the tree's per-point work ran on the GPU
    double build_ms[3] = {0, 0, 0};
    bool tree_on_gpu = false;
    int n_folded = 0;  // boxes whose proxies and outgoing plane waves are a periodic partner's

    // topology export for parity tests
    std::vector<int> box7;
    std::vector<double> box_centers;
    std::vector<int> src_perm_host, trg_perm_host;
};

extern template struct Plan<float>;
extern template struct Plan<double>;

}  // namespace pdmk::gpu
