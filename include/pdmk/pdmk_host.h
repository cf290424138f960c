/// Host-side introspection of libpdmk.so (no GPU needed).
///
/// These expose the host construction steps that feed the GPU plan (tree build
/// and DMK tables) so that they can be compared with a CPU restatement of the
/// reference PBCTree (cpp/src/dmk/pbc_tree.cpp:138-1127) and FourierData
/// (cpp/src/dmk/fourier_data.cpp:646-673).  They have no counterpart in the
/// reference ABI.  Return 0 on success, -1 on error (message on stderr).
#ifndef PDMK_HOST_H
#define PDMK_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/// Build geometry + tree for the given inputs.  sizes8 = {n_src, nsall, nboxes,
/// n_levels, n_list1_total, n_up_groups, 0, 0} is always filled; each non-NULL array
/// is filled as in pdmk_gpu_tree_arrays (box7[7*nboxes], centers[3*nboxes],
/// src_perm[nsall], trg_perm[n_src]); geom16 = {umat[9], rc, mgrid[3], xgrid0[3]}.
int pdmk_host_tree(const double* cell_3x3, int N, const double* positions_3xN, int n_per_leaf, int nz,
                   long long* sizes8, int* box7, double* centers, int* src_perm, int* trg_perm, double* geom16);

/// Tables for n_digits at cube size rc, n_levels levels, lattice frame umat (column-major).
/// scalars8 = {beta, psi0, c0, g0d2, n_pw, p, n_modes, n_far_modes}; non-NULL arrays:
/// w0[n_levels+1], radial0[n_modes], poly2pw[2*n_pw*p] (re, im; column-major n_pw x p),
/// p2c[2*p*p] (minus, plus; column-major), far_ijkw[4*n_far_modes] (i, j, k, weight).
int pdmk_host_tables(int n_digits, double rc, int n_levels, const double* umat, double* scalars8, double* w0,
                     double* radial0, double* poly2pw, double* p2c, double* far_ijkw);

/// psi_0^c(x) for |x| <= 1 as computed by the library (unit L2 norm on [-1,1], psi(0) > 0).
double pdmk_host_pswf(double c, double x);

#ifdef __cplusplus
}
#endif

#endif  // PDMK_HOST_H
