// Host preparation of the far-field non-uniform FFT (see nufft.hpp).
#include "nufft.hpp"

#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace pdmk::host {

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

double es_kernel(double z, double beta) {
    if (!(std::fabs(z) < 1.0)) return 0.0;
    return std::exp(beta * (std::sqrt(1.0 - z * z) - 1.0));
}

// Subgrid strides: the smallest layout (SY >= E, SZ >= E SY) for which the 8-byte cells
// z SZ + y SY of the lanes l = 0..31 and 32..63, (y, z) = (l % W, l / W), fall in distinct
// double-banks (address mod 32 in 8-byte units).
void subgrid_strides(int E, int W, int &SY, int &SZ) {
    int best_w = 1 << 30, best_size = 1 << 30;
    for (int sy = E; sy < E + 9; ++sy)
        for (int sz = E * sy; sz < E * sy + 64; ++sz) {
            int worst = 0;
            for (int half = 0; half < 2; ++half) {
                int cnt[32] = {0};
                for (int l = 32 * half; l < 32 * half + 32 && l < W * W; ++l)
                    worst = std::max(worst, ++cnt[((l / W) * sz + (l % W) * sy) % 32]);
            }
            const int size = E * sz;
            if (worst < best_w || (worst == best_w && size < best_size)) {
                best_w = worst;
                best_size = size;
                SY = sy;
                SZ = sz;
            }
        }
}

}  // namespace

int nufft_tile_edge(const int M[3], int W) {
    (void)W;
    long tiles = 1;
    for (int d = 0; d < 3; ++d) tiles *= (2L * M[d] + 3) / 4;
    return tiles >= 2048 ? 4 : 2;
}

NufftPlanHost make_nufft_grid(const int Mm[3], int W) {
    if (W < 4 || W > 14) throw std::invalid_argument("NUFFT kernel width out of range");
    NufftPlanHost P;
    const int T = nufft_tile_edge(Mm, W);
    P.T = T;
    P.W = W;
    P.B = (W & 1) ? T + 1 : T;
    P.lo = 1 - (W + 1) / 2;
    P.E = P.B - 1 + W;
    P.KS = P.E + 2 * W;
    P.beta = 2.30 * W;  // exponential-of-semicircle shape for an upsampling factor of 2
    subgrid_strides(P.E, W, P.SY, P.SZ);
    for (int d = 0; d < 3; ++d) {
        P.M[d] = Mm[d];
        const int n = (2 * Mm[d] + T - 1) / T * T;
        P.n[d] = n;
        P.h[d] = kTwoPi / n;
        P.ntile[d] = n / T;
    }
    // ---- per axis: tiles whose subgrid covers grid index g, with the local offset
    for (int d = 0; d < 3; ++d) {
        const int n = P.n[d];
        P.gptr[d].assign(1, 0);
        for (int g = 0; g < n; ++g) {
            for (int t = 0; t < P.ntile[d]; ++t) {
                const int org = T * t + P.lo;
                int loc = ((g - org) % n + n) % n;
                for (; loc < P.E; loc += n) {
                    P.gtile[d].push_back(t);
                    P.gloc[d].push_back(loc);
                }
            }
            P.gptr[d].push_back((int)P.gtile[d].size());
        }
    }
    // ---- DFT matrices and deconvolution: Phi(m) = alpha int_{-1}^{1} phi(z) e^{-i m alpha z} dz,
    // alpha = h W / 2 (Gauss-Legendre on [0, 1], phi even)
    std::vector<double> gx, gw;
    gauss_legendre(256, gx, gw);
    for (int d = 0; d < 3; ++d) {
        const int n = P.n[d], M = P.M[d];
        P.fwd[d].resize((size_t)M * n);
        P.inv[d].resize((size_t)n * M);
        P.dcv[d].resize(M);
        const double alpha = P.h[d] * 0.5 * W;
        for (int k = 0; k < M; ++k) {
            const int m = k - M / 2;
            for (int g = 0; g < n; ++g) {
                // exact argument reduction of m g / n before the trigonometric evaluation
                const long r = ((long)m * g % n + n) % n;
                const double ph = kTwoPi * (double)r / n;
                P.fwd[d][(size_t)k * n + g] = std::complex<double>(std::cos(ph), -std::sin(ph));
                P.inv[d][(size_t)g * M + k] = std::complex<double>(std::cos(ph), std::sin(ph));
            }
            double phi = 0;
            for (size_t i = 0; i < gx.size(); ++i) {
                const double z = 0.5 * (gx[i] + 1);
                phi += gw[i] * es_kernel(z, P.beta) * std::cos(m * alpha * z);  // 0.5 (half interval) x 2 (even)
            }
            phi *= alpha;
            P.dcv[d][k] = P.h[d] / phi;
        }
    }
    return P;
}

void nufft_units(NufftPlanHost &P, const std::vector<int> &tile_count) {
    // ---- units: the points of a tile, split into equal parts of at most `part` points when a
    // tile holds many more than the average
    const int ntiles = P.ntile[0] * P.ntile[1] * P.ntile[2];
    if ((int)tile_count.size() != ntiles) throw std::logic_error("NUFFT tile count size");
    std::vector<int> tstart(ntiles + 1, 0);
    for (int t = 0; t < ntiles; ++t) tstart[t + 1] = tstart[t] + tile_count[t];
    const int np = tstart[ntiles];
    int nonempty = 0;
    for (int t = 0; t < ntiles; ++t) nonempty += tile_count[t] > 0;
    const int part = std::max(256, 4 * ((np + std::max(1, nonempty) - 1) / std::max(1, nonempty)));
    P.unit_ptr.push_back(0);
    int n_extra = 0;
    P.extra_ptr.push_back(0);
    for (int t = 0; t < ntiles; ++t) {
        const int a = tstart[t], e = tstart[t + 1];
        if (e == a) continue;
        const int nparts = (e - a + part - 1) / part;
        for (int k = 0; k < nparts; ++k) {
            P.unit_ptr.push_back(a + (int)((long)(e - a) * (k + 1) / nparts));
            P.unit_tile.push_back(t);
            P.unit_slot.push_back(k == 0 ? t : ntiles + n_extra++);
        }
        if (nparts > 1) {
            P.extra_tile.push_back(t);
            P.extra_ptr.push_back(n_extra);
        }
    }
}

}  // namespace pdmk::host
