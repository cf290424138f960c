// FP64 Ewald sum on the GPU (the computation behind pdmk_ewald_nufft).
#pragma once

namespace pdmk::gpu {

// Potentials pot[N] (and the energy when energy != nullptr) of N charges q at positions
// pos[3N] (pos[3i + d]) in the cell whose columns cell[3j .. 3j+2] are the lattice vectors;
// s sets the truncation (r_cut = s/alpha, k_max = 2 s alpha).  Synchronous; throws on bad
// arguments or HIP errors.
void ewald_sum(const double *cell, int n, const double *pos, const double *q, double s, double *pot, double *energy);

// Splitting parameter of the reference (choose_alpha, cpp/src/ewald_nufft.cpp:40-56).
double ewald_reference_alpha(int n, double vol, double s);

}  // namespace pdmk::gpu
