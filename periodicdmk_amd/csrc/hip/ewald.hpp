// FP64 Ewald sum on the GPU (the computation behind pdmk_ewald_nufft).
#pragma once

#include <cstddef>

namespace pdmk::gpu {

// Potentials pot[N] (and the energy when energy != nullptr) of N charges q at positions
// pos[3N] (pos[3i + d]) in the cell whose columns cell[3j .. 3j+2] are the lattice vectors;
// s sets the truncation (r_cut = s/alpha, k_max = 2 s alpha).  Synchronous; throws on bad
// arguments or HIP errors.
void ewald_sum(const double *cell, int n, const double *pos, const double *q, double s, double *pot, double *energy);

// The same sum with every device array in a static per-device scratch array of
// ewald_scratch_bytes() (no device allocation): throws std::length_error when N and the mode
// count need more.  Calls that overlap on one GPU race (they share the array).
void ewald_sum_noalloc(const double *cell, int n, const double *pos, const double *q, double s, double *pot,
                       double *energy);
size_t ewald_scratch_bytes();

// Potentials pot[M] at the points targets[0 .. M) (indices into the N points) only
// (ewald_nufft_targets, cpp/src/ewald_nufft.cpp:663-759): the real-space sum is formed at the M
// targets, the structure factors over all N charges and evaluated at the M targets.  The
// splitting parameter balances those two costs; the result is that of ewald_sum at the same
// points.  Synchronous; throws on bad arguments (an index outside [0, N)) or HIP errors.
void ewald_sum_targets(const double *cell, int n, const double *pos, const double *q, int m, const int *targets,
                       double s, double *pot);

// The splitting parameter ewald_sum_targets uses: r_cut = h_min/k for the k in 1..32 that
// minimises the points the M target workgroups visit plus twice the (point, mode) products.
double ewald_targets_alpha(const double *cell, int n, const double *pos, int m, const int *targets, double s);

// Splitting parameter of the reference (choose_alpha, cpp/src/ewald_nufft.cpp:40-56).
double ewald_reference_alpha(int n, double vol, double s);

}  // namespace pdmk::gpu
