/// C ABI of the MI355X PeriodicDMK library (libpdmk.so).
///
/// Drop-in for the reference C ABI: every symbol, argument and return convention
/// below is the one declared in cpp/include/pdmk/pdmk_capi.h:20-51 of
/// xuanzhaogao/PeriodicDMK and implemented in cpp/src/pdmk_capi.cpp:86-166, which the
/// Julia package binds with ccall (julia/PeriodicDMK/src/PeriodicDMK.jl:53, 120-126,
/// 135-137, 172-181, 198-207, 236).
///
/// Conventions (unchanged):
///   - cell_3x3: column-major 3x3, cell[i + 3*j] = component i of lattice vector j
///   - positions_3xN: column-major 3xN, positions[d + 3*i] = coordinate d of particle i
///   - int-returning functions return 0 on success, nonzero on error; errors are
///     reported on stderr.  create returns NULL on error.
///   - evaluate ignores its N and positions arguments (the tree owns the positions)
///     and writes n_points() potentials into pot_out.
///   - a handle is not re-entrant: calls on one handle must not overlap.
#ifndef PDMK_CAPI_H
#define PDMK_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/// FP64 Ewald reference potentials (and energy when energy_out != NULL), computed on the
/// GPU: exact truncated real-space and reciprocal sums (r_cut = s/alpha, k_max = 2 s alpha),
/// with the reference's splitting-parameter convention for non-neutral systems.
/// Replaces pdmk_ewald_nufft, cpp/include/pdmk/pdmk_capi.h:22-28.
int pdmk_ewald_nufft(const double* cell_3x3, int N,
                     const double* positions_3xN,
                     const double* charges,
                     double s,
                     double* potentials_out,
                     double* energy_out);

/// Cubic cover of a cell.  Replaces pdmk_optimize_cubic_cover,
/// cpp/include/pdmk/pdmk_capi.h:31-37.  Returns 1 for nz <= 0.
int pdmk_optimize_cubic_cover(const double* cell_3x3, int nz,
                              int* m_out, int* n_out, int* p_out,
                              double* e1_out, double* e2_out, double* e3_out,
                              double* s_out, double* cost_out);

/// Build a tree (FP64 kernels).  Replaces pdmk_create_double, pdmk_capi.h:41-43.
void* pdmk_create_double(const double* cell_3x3, int N,
                         const double* positions_3xN,
                         int n_digits, int n_per_leaf, int nz);
/// Evaluate potentials for new charges.  Replaces pdmk_evaluate_double, pdmk_capi.h:44-46.
int   pdmk_evaluate_double(void* h, int N,
                           const double* positions_3xN, const double* charges,
                           double* pot_out);
/// Free a tree.  Replaces pdmk_destroy_double, pdmk_capi.h:47.
void  pdmk_destroy_double(void* h);

/// Build a tree (FP32 kernels).  Replaces pdmk_create_float, pdmk_capi.h:49-51.
void* pdmk_create_float(const double* cell_3x3, int N,
                        const double* positions_3xN,
                        int n_digits, int n_per_leaf, int nz);
/// Replaces pdmk_evaluate_float, pdmk_capi.h:52-54.
int   pdmk_evaluate_float(void* h, int N,
                          const double* positions_3xN, const double* charges,
                          float* pot_out);
/// Replaces pdmk_destroy_float, pdmk_capi.h:55.
void  pdmk_destroy_float(void* h);

#ifdef __cplusplus
}
#endif

#endif // PDMK_CAPI_H
