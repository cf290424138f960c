// Host preparation of the far-field non-uniform FFT: the truncated lattice Fourier sum
//   far(x_u) = sum_{listed modes m} Re( w_m S(m) e^{+i m.theta_u} ),  S(m) = sum_u q_u e^{-i m.theta_u},
// evaluated by spreading the real charges onto a sigma = 2 periodic grid with an
// exponential-of-semicircle kernel phi(z) = exp(beta (sqrt(1 - z^2) - 1)) of W points,
// a separable DFT to the mode box, deconvolution and weighting, the inverse DFT back to the
// grid and interpolation with the same kernel (the type-1 / type-2 NUFFT pair the reference
// runs through DUCC, cpp/src/dmk/pbc.cpp:441-455, 507-521).  Everything here depends only on
// the positions, so it is built once per tree.
#pragma once

#include <complex>
#include <vector>

namespace pdmk::host {

struct NufftPlanHost {
    int n[3] = {0, 0, 0};      // grid points per axis (2 M, even, a multiple of kTile)
    int M[3] = {0, 0, 0};      // mode box per axis, indices -M/2 .. (M-1)/2
    int W = 0, E = 0;          // kernel width (grid points) and tile subgrid extent
    int ntile[3] = {0, 0, 0};  // tiles per axis
    double beta = 0, h[3] = {0, 0, 0};
    static constexpr int kTile = 4;  // grid cells per tile edge
    // points in tile order: user index, subgrid base (x, y, z) and kernel values kx, ky, kz
    std::vector<int> pinfo;     // 4 per point: bx, by, bz, user
    std::vector<double> kval;   // 3 W per point
    // work units: (tile, part) with a point range; parts split the points of a tile
    std::vector<int> unit_ptr;   // n_units + 1
    std::vector<int> unit_tile;  // tile index (z-major) per unit
    std::vector<int> tile_units; // per tile (z-major, all tiles): first unit, count (2 per tile)
    // per axis d and grid index g: covering (tile coordinate, local offset) pairs
    std::vector<int> gptr[3], gtile[3], gloc[3];
    // DFT matrices: forward F_d[k][g] = e^{-i (k - M/2) 2 pi g / n}, inverse B_d[g][k] = conj
    std::vector<std::complex<double>> fwd[3], inv[3];
    // deconvolution per axis and mode index k = m + M/2: h / Phi(m)
    std::vector<double> dcv[3];
};

// theta[3u + d] = 2 pi (fractional coordinate d of point u) in [0, 2 pi); modes per axis M[d];
// W = kernel width (6 .. 14).
NufftPlanHost make_nufft_plan(const std::vector<double> &theta, int n_points, const int M[3], int W);

}  // namespace pdmk::host
