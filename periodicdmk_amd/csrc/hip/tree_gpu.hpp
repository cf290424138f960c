// Tree construction with the per-point work on the GPU (the PBCTree constructor and image
// generation, cpp/src/dmk/pbc_tree.cpp:138-740 and cpp/include/pdmk/image_layout.hpp:14-88).
//
// The box-level decisions (refinement, 2:1 balance, flags, lists) stay in
// host::build_tree_with; the GPU holds the points: periodic images (generated in the host
// enumeration order), level-0 binning, the stable distribution of split boxes into their
// children, and the final order, a stable radix sort by box path.  The results are
// identical to host::make_geometry + host::build_tree.
#pragma once

#include "../host/tree_build.hpp"

#include <hip/hip_runtime.h>

namespace pdmk::gpu {

// Fills g (frame, rotated positions, img_src; img_pos is left empty) and t.  Returns false
// when the tree is deeper than the 64-bit path keys hold (about 14 levels for a
// 2^20-box level-0 grid); g and t are then unspecified.
bool build_tree_gpu(const double *cell_colmajor, const double *pos, int n, int nz, int n_per_leaf, hipStream_t s,
                    host::Geometry &g, host::Tree &t);

// order[i] = the index of the i-th smallest of key[0 .. n) (stable: equal keys keep index
// order), by an LSD radix sort over bits [0, end_bit).  Device pointers, enqueued on s.
void stable_order(const int *key, int n, int end_bit, int *order, hipStream_t s);

}  // namespace pdmk::gpu
