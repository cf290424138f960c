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

}  // namespace

NufftPlanHost make_nufft_plan(const std::vector<double> &theta, int np, const int Mm[3], int W) {
    if (W < 2 || W > 14) throw std::invalid_argument("NUFFT kernel width out of range");
    NufftPlanHost P;
    constexpr int T = NufftPlanHost::kTile;
    P.W = W;
    P.beta = 2.30 * W;  // exponential-of-semicircle shape for an upsampling factor of 2
    P.E = T + W + 2;
    for (int d = 0; d < 3; ++d) {
        P.M[d] = Mm[d];
        int n = 2 * Mm[d];
        n = (n + T - 1) / T * T;
        P.n[d] = n;
        P.h[d] = kTwoPi / n;
        P.ntile[d] = n / T;
    }
    const int half = W / 2;
    // ---- per point: tile, subgrid base and kernel values
    const int ntiles = P.ntile[0] * P.ntile[1] * P.ntile[2];
    std::vector<int> tile_of(np), base(3 * (size_t)np);
    std::vector<double> kv(3 * (size_t)W * np);
    for (int u = 0; u < np; ++u) {
        int tc[3];
        for (int d = 0; d < 3; ++d) {
            const int n = P.n[d];
            double x = theta[3 * (size_t)u + d] / P.h[d];  // grid coordinate in [0, n)
            if (x < 0) x = 0;
            if (x >= n) x = std::nextafter((double)n, 0.0);
            const int c = std::min(n - 1, (int)std::floor(x));
            tc[d] = c / T;
            const int g0 = (int)std::floor(x - 0.5 * W) + 1;  // first grid point inside the support
            const int org = T * tc[d] - half - 1;
            const int b = g0 - org;
            if (b < 0 || b + W > P.E) throw std::logic_error("NUFFT tile subgrid too small");
            base[3 * (size_t)u + d] = b;
            for (int k = 0; k < W; ++k)
                kv[((size_t)u * 3 + d) * W + k] = es_kernel((x - (g0 + k)) / (0.5 * W), P.beta);
        }
        tile_of[u] = (tc[2] * P.ntile[1] + tc[1]) * P.ntile[0] + tc[0];
    }
    // ---- counting sort by tile; tiles split into parts of at most `part` points so that
    // there are enough work units for the GPU
    std::vector<int> tstart(ntiles + 1, 0);
    for (int u = 0; u < np; ++u) tstart[tile_of[u] + 1]++;
    for (int t = 0; t < ntiles; ++t) tstart[t + 1] += tstart[t];
    std::vector<int> order(np);
    {
        std::vector<int> fill(tstart.begin(), tstart.end() - 1);
        for (int u = 0; u < np; ++u) order[fill[tile_of[u]]++] = u;
    }
    P.pinfo.resize(4 * (size_t)np);
    P.kval.resize(3 * (size_t)W * np);
    for (int i = 0; i < np; ++i) {
        const int u = order[i];
        for (int d = 0; d < 3; ++d) P.pinfo[4 * (size_t)i + d] = base[3 * (size_t)u + d];
        P.pinfo[4 * (size_t)i + 3] = u;
        std::copy(kv.begin() + 3 * (size_t)W * u, kv.begin() + 3 * (size_t)W * (u + 1), P.kval.begin() + 3 * (size_t)W * i);
    }
    const int part = std::max(64, np / 4096 + 1);
    P.unit_ptr.push_back(0);
    P.tile_units.assign(2 * (size_t)ntiles, 0);
    for (int t = 0; t < ntiles; ++t) {
        P.tile_units[2 * t] = (int)P.unit_tile.size();
        for (int a = tstart[t]; a < tstart[t + 1]; a += part) {
            P.unit_tile.push_back(t);
            P.unit_ptr.push_back(std::min(tstart[t + 1], a + part));
        }
        P.tile_units[2 * t + 1] = (int)P.unit_tile.size() - P.tile_units[2 * t];
    }
    // ---- per axis: tiles whose subgrid covers grid index g, with the local offset
    for (int d = 0; d < 3; ++d) {
        const int n = P.n[d];
        P.gptr[d].assign(1, 0);
        for (int g = 0; g < n; ++g) {
            for (int t = 0; t < P.ntile[d]; ++t) {
                const int org = T * t - half - 1;
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

}  // namespace pdmk::host
