/// Device-resident entry points of libpdmk.so.
///
/// These extend the drop-in ABI (pdmk_capi.h) for callers that keep charges and
/// potentials in HBM.  They take handles created by pdmk_create_{double,float} and
/// run the same evaluate() path as pdmk_evaluate_* without host copies.  The
/// reference has no device memory; each function names the reference call it
/// corresponds to.  All return 0 on success and -1 on error (message on stderr).
#ifndef PDMK_GPU_H
#define PDMK_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/// Device-memory helpers for callers without a HIP runtime of their own (Python
/// tests and the benchmark use them so that one process holds one HIP runtime).
/// No reference counterpart.  malloc returns NULL on failure; the others return 0 on
/// success and -1 on error.  kind: 1 host->device, 2 device->host, 3 device->device.
int   pdmk_gpu_device_count(void);
void* pdmk_gpu_malloc(unsigned long long bytes);
int   pdmk_gpu_free(void* p);
int   pdmk_gpu_memcpy(void* dst, const void* src, unsigned long long bytes, int kind);
int   pdmk_gpu_synchronize(void);

/// Number of points owned by a handle (PeriodicDMKTreeT::n_points,
/// cpp/include/pdmk/periodic_dmk.hpp:52).  Returns -1 for a NULL handle.
int pdmk_gpu_n_points(void* h);

/// evaluate() with device pointers: d_charges[N] (FP64, user order) in,
/// d_pot[N] (double for *_double handles, float for *_float) out, enqueued on
/// `stream` (a hipStream_t; NULL = the handle's own stream).  Returns after
/// enqueueing; the caller synchronizes the stream.  Corresponds to
/// PeriodicDMKTreeT::evaluate, cpp/src/periodic_dmk.cpp:141-186.
int pdmk_gpu_evaluate_double(void* h, const double* d_charges, double* d_pot, void* stream);
int pdmk_gpu_evaluate_float(void* h, const double* d_charges, float* d_pot, void* stream);

/// A batch of evaluate() calls on device pointers captured once into a HIP graph and then
/// replayed with one launch, for many small trees (say 256 cells x 32 charge vectors) whose
/// evaluate() is a few dozen short kernels each.  No reference counterpart: the reference
/// runs trees one after another on the CPU.  Call i evaluates handles[i] on d_charges[i]
/// (FP64, N of that handle) into d_pots[i] (the handle's precision); the calls of one handle
/// keep their order, handles are dealt to n_branches branches, each a graph of its own
/// launched on its own stream (the branches run concurrently on the hardware queues: set
/// GPU_MAX_HW_QUEUES to about n_branches for the process to get them).  The buffers are
/// bound at creation: a launch reads the charges they hold then.  One handle's work must not
/// overlap another use of that handle (its scratch is shared).  create returns NULL on a
/// NULL or quasi-periodic handle, a NULL buffer, n_calls or n_branches <= 0 or a HIP error;
/// launch forks the branches from `stream` (NULL = the graph's own stream), joins them back
/// into it and returns; destroy waits for the graph's streams and frees it.  A graph uses its
/// handles' device memory: pdmk_destroy_* on a handle that a live graph uses only marks it,
/// and the handle is freed when the last such graph is destroyed (no other call may use a
/// handle after its destroy call).
/// on != 0: the handle's evaluate() runs its stages one after another on the caller's stream
/// instead of forking the far field and the near field onto the handle's own streams.  For
/// callers that evaluate many trees at once: their trees, not the stages, then fill the GPU
/// (256 trees of 2e3-2e5 points, 4 hardware queues: 96 -> 141 Mpts/s), while one tree alone
/// is slower (2e4 points: 0.56 -> 0.98 ms).  Results are unchanged.  No reference counterpart.
/// -1 for a NULL handle.
int pdmk_gpu_set_single_stream(void* h, int on);

void* pdmk_gpu_graph_create(int n_calls, void* const* handles, const double* const* d_charges, void* const* d_pots,
                            int n_branches);
int pdmk_gpu_graph_launch(void* g, void* stream);
int pdmk_gpu_graph_destroy(void* g);

/// evaluate_timed() on device pointers: runs synchronously and fills
/// times_ms[0..2] = {nufft_ms, dmk_ms, direct_ms} with the same stage grouping as
/// EvalTiming (cpp/include/pdmk/eval_timing.hpp:10-19,
/// cpp/src/periodic_dmk.cpp:188-249), measured with HIP events.
int pdmk_gpu_evaluate_timed(void* h, const double* d_charges, void* d_pot, double* times_ms);

/// Run only selected stages (bit 0: proxy/plane-wave part, bit 1: near field,
/// bit 2: self correction, bit 3: far field) on device pointers, synchronously.
/// Stage-level parity checks use it; stages=15 equals pdmk_gpu_evaluate_*.
int pdmk_gpu_evaluate_stages(void* h, const double* d_charges, void* d_pot, int stages);

/// Quasi-periodic trees (Bloch vector `bloch[3]`, in the frame the reference uses: it is
/// dotted with lattice-frame translations).  Replaces the C++ constructor with `bloch` and
/// PeriodicDMKTreeT::evaluate_complex (cpp/include/pdmk/periodic_dmk.hpp:17-144,
/// cpp/src/periodic_dmk.cpp:251-277), which the reference's C ABI does not expose.
/// create: NULL on error (message on stderr), or when bloch is NULL.  evaluate_complex:
/// pot_out[2i], pot_out[2i+1] = Re, Im at point i (user order); N and positions are ignored
/// as in pdmk_evaluate_*; -1 on error or on a handle without bloch.  The plain
/// pdmk_evaluate_* on a quasi-periodic handle return -1.
void* pdmk_create_qp_double(const double* cell_3x3, int N, const double* positions_3xN, int n_digits,
                            int n_per_leaf, int nz, const double* bloch);
void* pdmk_create_qp_float(const double* cell_3x3, int N, const double* positions_3xN, int n_digits,
                           int n_per_leaf, int nz, const double* bloch);
int pdmk_evaluate_complex_double(void* h, int N, const double* positions_3xN, const double* charges,
                                 double* pot_out);
int pdmk_evaluate_complex_float(void* h, int N, const double* positions_3xN, const double* charges,
                                float* pot_out);
/// Device pointers: d_charges FP64 (N), d_pot Real (2N); synchronous.  stages as in
/// pdmk_gpu_evaluate_stages (15 = everything).
int pdmk_gpu_evaluate_complex(void* h, const double* d_charges, void* d_pot, int stages);

/// pdmk_ewald_nufft without device allocation.  No counterpart in the reference: for callers
/// that cannot allocate device memory (serving loops, harnesses that forbid hipMalloc).  The
/// same arguments, potentials and energy (bit for bit, or to 1e-12 when the structure-factor
/// partial sums are cut into fewer point tiles to fit); every device array is a slice of a
/// static per-GPU scratch array of pdmk_ewald_noalloc_scratch_bytes() (128 MiB), so it
/// returns -1 (size in the message) when N and the mode count need more.  Calls that overlap
/// on one GPU share the array and race.
int pdmk_ewald_nufft_noalloc(const double* cell_3x3, int N, const double* positions_3xN, const double* charges,
                             double s, double* potentials_out, double* energy_out);
unsigned long long pdmk_ewald_noalloc_scratch_bytes(void);

/// FP64 Ewald potentials at M of the N points only, computed on the GPU: pot_out[t] is the
/// potential at point target_indices[t] (an index into positions_3xN / charges, which are laid
/// out as in pdmk_ewald_nufft), equal to pdmk_ewald_nufft's at that point.  Replaces
/// pdmk::ewald_nufft_targets (cpp/include/pdmk/ewald_nufft.hpp:42-48,
/// cpp/src/ewald_nufft.cpp:663-759), which the reference has in C++ only: an accuracy
/// reference at large N.  -1 on a NULL argument, N <= 0, M < 0, s <= 0 or an index outside
/// [0, N).
int pdmk_gpu_ewald_targets(const double* cell_3x3, int N, const double* positions_3xN, const double* charges,
                           int M, const int* target_indices, double s, double* pot_out);

/// Tree and work statistics: out[0..15] = {n_src, nsall, nboxes, n_levels, p, n_pw,
/// ms, mt, mu, n_leaves, n_list1, n_direct_leaves, n_far_modes, n_digits,
/// is_float, n_nf_steps}.
int pdmk_gpu_stats(void* h, long long* out16);

/// evaluate() of n_vec charge vectors on one tree: d_charges FP64 [n_vec][N], d_pot [n_vec][N]
/// in the handle's precision (N = pdmk_gpu_n_points), enqueued on `stream` (NULL = the
/// handle's stream).  No reference counterpart (the reference evaluates one vector per call).
/// FP32 handles at 6 digits (the row near field) form the near field of 16 vectors per pass,
/// each (target, source) residual once for all 16, beside the far fields, then run the proxy /
/// plane-wave chain per vector; other handles call evaluate() per vector.  Each vector's
/// potentials equal pdmk_gpu_evaluate_*'s on it.  The first batched call allocates the
/// 16-vector buffers (not capture-safe); later calls allocate nothing.  -1 on a NULL handle or
/// buffer, n_vec < 0, a quasi-periodic handle or a HIP error.
int pdmk_gpu_evaluate_many(void* h, int n_vec, const void* d_charges, void* d_pot, void* stream);

/// Autotuner of the launch shapes (project-written; the reference has no GPU kernels to tune):
/// times `repeats` evaluate() calls (after one warm-up) on the handle for each candidate of
/// each knob in turn -- near-field wavefronts per workgroup {4, 8, 16}, near-field start
/// (at once / after the upward pass), proxy -> plane-wave and plane-wave -> proxy LDS
/// budgets {16, 32, 64} KiB (slices per workgroup), tensor-transform planes per workgroup
/// (halved or not) -- keeping a change only when it is at least 2% faster, and applies the
/// result.  The choice is cached per device and plan kind (order, precision, log2 buckets of
/// N, proxy slots and near-field items) for the process, and appended to the file
/// PDMK_TUNE_CACHE names when set; new handles of a cached kind start with it.  Potentials do
/// not depend on the knobs beyond FP rounding.  out8 = {ms before (built-in rules), ms after,
/// the 5 knobs as in pdmk_gpu_knobs, 1 if the handle started from a cached choice}.  -1 on a
/// NULL handle or a HIP error.
int pdmk_gpu_autotune(void* h, int repeats, double* out8);
/// Knobs of a handle: set5 (NULL = unchanged) = {near-field wavefronts per workgroup (0 = rule,
/// 2, 4, 8, 16), near-field late start (-1 = rule, 0, 1), proxy -> plane-wave LDS KiB (0 = rule,
/// 16, 32, 64), plane-wave -> proxy LDS KiB (same), tensor-transform plane divisor (1, 2, 4)};
/// get5 (NULL = not read) receives the current ones.  -1 on a NULL handle or a value outside
/// those sets.
int pdmk_gpu_knobs(void* h, const int* set5, int* get5);
/// Forget this process's tuned choices (handles built afterwards start with the built-in
/// shapes; the PDMK_TUNE_CACHE file is not read again).
void pdmk_gpu_tune_cache_clear(void);

/// Construction of the handle: out[0..3] = {total ms, tree ms (geometry, images, tree),
/// far-field plan ms, 1 if the tree's per-point work (images, binning, distribution, sort)
/// ran on the GPU else 0, boxes folded onto a periodic counterpart (their proxies and
/// outgoing plane waves are not computed), proxy slots computed, 1 if the outgoing plane
/// waves are held one level at a time (when all of them do not fit in device memory) else 0,
/// 1 if the near field walks per-row step lists (k_near_rows: FP32 plans at 6 digits) else 0}.
/// No reference counterpart (measurement hook).
int pdmk_gpu_build_info(void* h, double* out8);

/// Algorithmic work per evaluate(): out[0..7] = {near-field pairs in support,
/// near-field candidate pairs, near-field flops (in-support pairs x (2*NC+12)),
/// proxy/PW transform flops, far-field mode-point products (of the exact sum), far-field
/// flops (of the exact sum, or of the NUFFT when the plan uses it: 4 W^3 per point plus the
/// pruned axis DFTs), bytes of O(N) traffic, near-field kernel HBM bytes}.  Near-field pair
/// counts are computed on the host at build time.
int pdmk_gpu_work(void* h, double* out8);

/// Kernel-group timing with HIP events on the evaluate stream (measurement aid; no
/// reference counterpart).  enable: keep the last `slots` evaluate() calls (0 turns it
/// off).  times: after the caller synchronized, avg_ms[0..4] = mean duration per call of
/// {far field, charge gather, proxy/plane-wave transforms, near field, finalize};
/// returns the number of calls averaged.
int pdmk_gpu_enable_kernel_timing(void* h, int slots);
int pdmk_gpu_kernel_times(void* h, double* avg_ms5);

/// Per-box topology (for comparison with a CPU restatement): box7[7*nboxes] =
/// {depth, parent, nchild, src_count, trg_count, ifpwexp, n_list1}, centers[3*nboxes],
/// src_perm[nsall] (sorted -> combined index), trg_perm[n_src] (sorted -> user index).
int pdmk_gpu_tree_arrays(void* h, int* box7, double* centers, int* src_perm, int* trg_perm);

#ifdef __cplusplus
}
#endif

#endif // PDMK_GPU_H
