// Host preparation of the far-field non-uniform FFT: the truncated lattice Fourier sum
//   far(x_u) = sum_{listed modes m} Re( w_m S(m) e^{+i m.theta_u} ),  S(m) = sum_u q_u e^{-i m.theta_u},
// evaluated by spreading the real charges onto a sigma >= 2 periodic grid with an
// exponential-of-semicircle kernel phi(z) = exp(beta (sqrt(1 - z^2) - 1)) of W points,
// a separable DFT to the mode box, deconvolution and weighting, the inverse DFT back to the
// grid and interpolation with the same kernel (the type-1 / type-2 NUFFT pair the reference
// runs through DUCC, cpp/src/dmk/pbc.cpp:441-455, 507-521).  Everything here depends only on
// the positions, so it is built once per tree.
//
// Work decomposition.  The grid is cut into tiles of T^3 cells.  A point whose cell lies in
// tile t touches the W^3 block of grid points starting at T t + lo + b (per axis), with
// lo = 1 - ceil(W/2) and base b in [0, B), B = T (W even) or T + 1 (W odd); every such block
// lies inside the tile's subgrid of E = B - 1 + W cells per axis.  Within a tile the points
// are ordered by base (bz, by, bx): the spreading and interpolation kernels keep the subgrid
// row (by, bz) in registers and touch the subgrid once per row instead of once per point.
#pragma once

#include <complex>
#include <vector>

namespace pdmk::host {

struct NufftPlanHost {
    int n[3] = {0, 0, 0};      // grid points per axis (>= 2 M, a multiple of T)
    int M[3] = {0, 0, 0};      // mode box per axis, indices -M/2 .. (M-1)/2
    int W = 0, T = 0, B = 0;   // kernel width, tile edge, bases per axis
    int lo = 0, E = 0;         // subgrid origin offset (T t + lo) and extent per axis
    int SY = 0, SZ = 0;        // LDS strides of a subgrid cell (x, y, z) -> z SZ + y SY + x, padded
                               // so that lanes owning (y, z) = (l % W, l / W) of a half
                               // wavefront hit distinct LDS banks
    int ntile[3] = {0, 0, 0};  // tiles per axis
    double beta = 0, h[3] = {0, 0, 0};
    // per point, in plan order (tile, then base (bz, by, bx), then user order), built on the
    // GPU: the user index, KS = E + 2 W kernel values (phi along x placed in its subgrid row,
    // phi_x at bx .. bx + W - 1 and zero elsewhere, then phi_y and phi_z at the W points of
    // its block) and the packed base bx | by << 8 | bz << 16
    int KS = 0;
    // work units: a point range of one tile; unit_ptr has n_units + 1 entries
    std::vector<int> unit_ptr, unit_tile;
    // subgrid slot per unit: the tile index for the first unit of a tile, n_tiles + k for the
    // k-th extra unit; extra slots are added into their tile slot (in unit order)
    std::vector<int> unit_slot;
    std::vector<int> extra_ptr, extra_tile;  // per tile with extra units: first extra slot ptr
    // per axis d and grid index g: covering (tile coordinate, local offset) pairs
    std::vector<int> gptr[3], gtile[3], gloc[3];
    // DFT matrices: forward F_d[k][g] = e^{-i (k - M/2) 2 pi g / n}, inverse B_d[g][k] = conj
    std::vector<std::complex<double>> fwd[3], inv[3];
    // deconvolution per axis and mode index k = m + M/2: h / Phi(m)
    std::vector<double> dcv[3];
};

// Tile edge for a mode box M and kernel width W: 4 when that still gives at least 2048 tiles
// (so that the one-wavefront-per-tile kernels fill the chip), else 2.  The spreading and
// interpolation kernels take at most 5 bases per axis (T + 1 for odd W).
int nufft_tile_edge(const int M[3], int W);

// Grid, tiles, subgrid strides, covering tables, DFT matrices and deconvolution for modes
// per axis M[d] and kernel width W (4 .. 14): everything that does not depend on the points.
NufftPlanHost make_nufft_grid(const int M[3], int W);

// Work units from the number of points per tile (tiles z-major, points in plan order): the
// points of a tile, split into equal parts when a tile holds far more than the average.
void nufft_units(NufftPlanHost &P, const std::vector<int> &tile_count);

}  // namespace pdmk::host
