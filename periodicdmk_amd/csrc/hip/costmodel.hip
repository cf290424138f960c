// Learned cost model of the solver's construction parameters: a small FP32 MLP trained and
// evaluated with the kernels below (no library GEMMs: every layer is at most 64 wide).
//
//   x (7) -> GELU(W1 x + b1) (64) -> GELU(W2 . + b2) (64) -> GELU(W3 . + b3) (64) -> W4 . + b4 (2)
//
// x holds the problem's features and one candidate configuration (see pdmk_tune.h); the two
// outputs predict the standardized log10 evaluate() time and log10 error of that
// configuration.  GELU(z) = z Phi(z) with the normal CDF Phi(z) = (1 + erf(z / sqrt 2)) / 2.
//
// Parameters are one flat FP32 array, layer by layer: W_l (row-major, [out][in]) then b_l.
// Training minimises the masked mean squared error sum_{b,o} M[b,o] (Y[b,o] - T[b,o])^2 /
// sum M with Adam.  Weight gradients are summed over the batch rows in row order by one thread
// per weight, so a training run is deterministic.
#include "plan.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

namespace pdmk::gpu {
namespace {

constexpr int kLayers = 4;
constexpr int kDims[kLayers + 1] = {7, 64, 64, 64, 2};
constexpr int kIn = 7, kHid = 64, kOut = 2;
constexpr int kRowsPerBlock = 4;  // 4 rows x 64 units per 256-thread workgroup

constexpr int param_offset(int l) {  // start of W_l
    int o = 0;
    for (int i = 0; i < l; ++i) o += kDims[i + 1] * kDims[i] + kDims[i + 1];
    return o;
}
constexpr int kParams = param_offset(kLayers);

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Z[b][j] = b_j + sum_k W[j][k] X[b][k]; H = GELU(Z) (or Z for the output layer).  Thread
// (row, j); the row is wave-uniform.
template <bool ACT>
__global__ __launch_bounds__(256) void k_mlp_fwd(const float *__restrict__ X, int B, int nin,
                                                 const float *__restrict__ W, const float *__restrict__ bias, int nout,
                                                 float *__restrict__ Z, float *__restrict__ H) {
    const int row = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x >> 6), j = threadIdx.x & 63;
    if (row >= B || j >= nout) return;
    const float *x = X + (size_t)row * nin;
    const float *w = W + (size_t)j * nin;
    float z = bias[j];
    for (int k = 0; k < nin; ++k) z = fmaf(w[k], x[k], z);
    if (Z) Z[(size_t)row * nout + j] = z;
    H[(size_t)row * nout + j] = ACT ? gelu(z) : z;
}

// Output error: dY = 2 M (Y - T) / count; per-row loss terms M (Y - T)^2 / count.
__global__ __launch_bounds__(256) void k_mlp_out_grad(const float *__restrict__ Y, const float *__restrict__ T,
                                                      const float *__restrict__ M, int n, float inv_count,
                                                      float *__restrict__ dY, float *__restrict__ lterm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = M[i] * (Y[i] - T[i]);
    dY[i] = 2.0f * d * inv_count;
    lterm[i] = d * (Y[i] - T[i]) * inv_count;
}

// Back through a linear layer and the previous activation:
// dZ[b][k] = GELU'(Zp[b][k]) sum_j dZn[b][j] W[j][k]   (W: [nout][nin], nin = width of Zp).
__global__ __launch_bounds__(256) void k_mlp_bwd_data(const float *__restrict__ dZn, int B, int nout,
                                                      const float *__restrict__ W, int nin,
                                                      const float *__restrict__ Zp, float *__restrict__ dZ) {
    const int row = blockIdx.x * kRowsPerBlock + (int)(threadIdx.x >> 6), k = threadIdx.x & 63;
    if (row >= B || k >= nin) return;
    const float *d = dZn + (size_t)row * nout;
    float s = 0.0f;
    for (int j = 0; j < nout; ++j) s = fmaf(d[j], W[(size_t)j * nin + k], s);
    dZ[(size_t)row * nin + k] = s * gelu_grad(Zp[(size_t)row * nin + k]);
}

// Weight and bias gradients of one layer: gW[j][k] = sum_b dZ[b][j] Hin[b][k],
// gb[j] = sum_b dZ[b][j], summed over b in order (one thread per output).
__global__ __launch_bounds__(256) void k_mlp_wgrad(const float *__restrict__ dZ, const float *__restrict__ Hin, int B,
                                                   int nin, int nout, float *__restrict__ gW, float *__restrict__ gb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = nout * nin;
    if (t < nw) {
        const int j = t / nin, k = t % nin;
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s = fmaf(dZ[(size_t)b * nout + j], Hin[(size_t)b * nin + k], s);
        gW[t] = s;
    } else if (t < nw + nout) {
        const int j = t - nw;
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s += dZ[(size_t)b * nout + j];
        gb[j] = s;
    }
}

// Adam with decoupled weight decay (AdamW) on the flat parameter array; step t >= 1.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                                              float wd, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * ((mi * c1) / (sqrtf(vi * c2) + eps) + wd * p[i]);
}

// Sum of n values in index order by one thread (the loss of a step; n is small).
__global__ void k_sum_ordered(const float *__restrict__ x, int n, float *__restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float s = 0.0f;
        for (int i = 0; i < n; ++i) s += x[i];
        *out = s;
    }
}

// Device state of one batch: inputs, activations, gradients, optimizer moments.
struct Mlp {
    int B = 0;
    hipStream_t s = nullptr;
    DevBuf P, G, Mo, Vo, X, T, Msk, Y, dY, lterm, loss;
    DevBuf Z[kLayers - 1], H[kLayers - 1], dZ[kLayers - 1];
    explicit Mlp(int batch) : B(batch) {
        if (B <= 0) throw std::invalid_argument("batch must be positive");
        check(hipStreamCreate(&s), "hipStreamCreate");
        P.alloc(sizeof(float) * kParams);
        G.alloc(sizeof(float) * kParams);
        X.alloc(sizeof(float) * (size_t)B * kIn);
        Y.alloc(sizeof(float) * (size_t)B * kOut);
        for (int l = 0; l < kLayers - 1; ++l) {
            Z[l].alloc(sizeof(float) * (size_t)B * kHid);
            H[l].alloc(sizeof(float) * (size_t)B * kHid);
        }
    }
    ~Mlp() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
    Mlp(const Mlp &) = delete;
    Mlp &operator=(const Mlp &) = delete;
    float *w(int l) const { return P.as<float>() + param_offset(l); }
    float *bias(int l) const { return w(l) + kDims[l + 1] * kDims[l]; }
    float *gw(int l) const { return G.as<float>() + param_offset(l); }
    float *gbias(int l) const { return gw(l) + kDims[l + 1] * kDims[l]; }
    void upload(float *dst, const float *src, size_t n) {
        check(hipMemcpyAsync(dst, src, sizeof(float) * n, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    }
    void forward() {
        const dim3 grid((B + kRowsPerBlock - 1) / kRowsPerBlock);
        const float *in = X.as<float>();
        for (int l = 0; l < kLayers - 1; ++l) {
            hipLaunchKernelGGL(k_mlp_fwd<true>, grid, dim3(256), 0, s, in, B, kDims[l], w(l), bias(l), kDims[l + 1],
                               Z[l].as<float>(), H[l].as<float>());
            in = H[l].as<float>();
        }
        hipLaunchKernelGGL(k_mlp_fwd<false>, grid, dim3(256), 0, s, in, B, kHid, w(kLayers - 1), bias(kLayers - 1),
                           kOut, static_cast<float *>(nullptr), Y.as<float>());
        check(hipGetLastError(), "k_mlp_fwd");
    }
    void prepare_targets(const float *t, const float *m) {
        T.alloc(sizeof(float) * (size_t)B * kOut);
        Msk.alloc(sizeof(float) * (size_t)B * kOut);
        dY.alloc(sizeof(float) * (size_t)B * kOut);
        lterm.alloc(sizeof(float) * (size_t)B * kOut);
        loss.alloc(sizeof(float));
        for (auto &d : dZ) d.alloc(sizeof(float) * (size_t)B * kHid);
        upload(T.as<float>(), t, (size_t)B * kOut);
        upload(Msk.as<float>(), m, (size_t)B * kOut);
    }
    // forward, loss and the full gradient into G
    void backward(float inv_count) {
        forward();
        const int nY = B * kOut;
        hipLaunchKernelGGL(k_mlp_out_grad, dim3((nY + 255) / 256), dim3(256), 0, s, Y.as<float>(), T.as<float>(),
                           Msk.as<float>(), nY, inv_count, dY.as<float>(), lterm.as<float>());
        hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(64), 0, s, lterm.as<float>(), nY, loss.as<float>());
        const dim3 grid((B + kRowsPerBlock - 1) / kRowsPerBlock);
        const float *dnext = dY.as<float>();
        for (int l = kLayers - 1; l >= 0; --l) {
            const int nin = kDims[l], nout = kDims[l + 1];
            const float *hin = l == 0 ? X.as<float>() : H[l - 1].as<float>();
            const int nt = nout * nin + nout;
            hipLaunchKernelGGL(k_mlp_wgrad, dim3((nt + 255) / 256), dim3(256), 0, s, dnext, hin, B, nin, nout, gw(l),
                               gbias(l));
            if (l > 0) {
                hipLaunchKernelGGL(k_mlp_bwd_data, grid, dim3(256), 0, s, dnext, B, nout, w(l), nin, Z[l - 1].as<float>(),
                                   dZ[l - 1].as<float>());
                dnext = dZ[l - 1].as<float>();
            }
        }
        check(hipGetLastError(), "k_mlp backward");
    }
};

float inv_mask_count(const float *m, int n) {
    double c = 0;
    for (int i = 0; i < n; ++i) c += m[i];
    if (!(c > 0)) throw std::invalid_argument("mask selects no output");
    return (float)(1.0 / c);
}

template <class F>
int guarded(const char *what, F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s: %s\n", what, e.what());
        return -1;
    }
}

// ---------------------------------------------------------------- configuration choice
// A model blob is [x_mean (7), x_std (7), y_mean (2), y_std (2), parameters].
constexpr int kBlobHeader = 2 * kIn + 2 * kOut;
constexpr int kNplChoices[] = {100, 150, 200, 300, 400, 600, 800};
// nz below 4 widens rc and raised the error against the FP64 Ewald sum by up to 40x in the
// training solves (median 1.3x); nz >= 4 kept it within 1.1x of the defaults' (p90): the
// choice stays among nz >= 4 so that a faster configuration is one of equal accuracy
constexpr int kNzMin = 4, kNzMax = 8;
constexpr int kDefaultNpl = 300, kDefaultNz = 4;
constexpr double kGainMargin = 0.85;  // leave the defaults only for a predicted time below 0.85 x theirs

struct Candidate { int npl, nz, is_float; };

// The constructor parameters the model chooses among: n_per_leaf x nz x precision (FP32
// kernels only up to 6 digits, as AutoPeriodicDMKTree); the defaults come first.
std::vector<Candidate> candidates(int n_digits) {
    const int def_float = n_digits <= 6 ? 1 : 0;
    std::vector<Candidate> c{{kDefaultNpl, kDefaultNz, def_float}};
    for (int f = 1; f >= 0; --f) {
        if (f == 1 && n_digits > 6) continue;
        for (int npl : kNplChoices)
            for (int nz = kNzMin; nz <= kNzMax; ++nz)
                if (!(npl == kDefaultNpl && nz == kDefaultNz && f == def_float)) c.push_back({npl, nz, f});
    }
    return c;
}

void config_row(const double *f4, const Candidate &c, const float *xm, const float *xs, float *row) {
    const double raw[kIn] = {f4[0], f4[1], f4[2], f4[3], std::log2((double)c.npl), (double)c.nz, (double)c.is_float};
    for (int k = 0; k < kIn; ++k) row[k] = (float)((raw[k] - xm[k]) / xs[k]);
}

}  // namespace
}  // namespace pdmk::gpu

extern "C" {

// Features of a problem (host): log10 N, log10 of the cell matrix's 2-norm condition number,
// the clustering index log10(var / mean) of the point counts over a G^3 grid of fractional
// coordinates (G^3 about N / 64 cells; 0 for uniform points), and n_digits.
int pdmk_tune_features(const double *cell, int N, const double *pos, int n_digits, double *f4) {
    return pdmk::gpu::guarded("pdmk_tune_features", [&] {
        if (!cell || !pos || !f4 || N <= 0) throw std::invalid_argument("null argument or N <= 0");
        // A^T A (columns = lattice vectors) and its eigenvalues by Jacobi rotations
        double a[3][3], g[3][3];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) a[i][j] = cell[i + 3 * j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                g[i][j] = 0;
                for (int k = 0; k < 3; ++k) g[i][j] += a[k][i] * a[k][j];
            }
        for (int sweep = 0; sweep < 50; ++sweep) {
            double off = std::fabs(g[0][1]) + std::fabs(g[0][2]) + std::fabs(g[1][2]);
            if (off < 1e-300) break;
            for (int p = 0; p < 2; ++p)
                for (int q = p + 1; q < 3; ++q) {
                    if (std::fabs(g[p][q]) < 1e-300) continue;
                    const double th = 0.5 * std::atan2(2 * g[p][q], g[q][q] - g[p][p]);
                    const double c = std::cos(th), s = std::sin(th);
                    for (int k = 0; k < 3; ++k) {  // G <- G J
                        const double gp = g[k][p], gq = g[k][q];
                        g[k][p] = c * gp - s * gq;
                        g[k][q] = s * gp + c * gq;
                    }
                    for (int k = 0; k < 3; ++k) {  // G <- J^T G
                        const double gp = g[p][k], gq = g[q][k];
                        g[p][k] = c * gp - s * gq;
                        g[q][k] = s * gp + c * gq;
                    }
                }
        }
        const double ev[3] = {g[0][0], g[1][1], g[2][2]};
        const double emax = std::max(ev[0], std::max(ev[1], ev[2])), emin = std::min(ev[0], std::min(ev[1], ev[2]));
        if (!(emin > 0)) throw std::invalid_argument("cell matrix is singular");
        // fractional coordinates f = A^{-1} x by Cramer's rule
        double inv[3][3];
        const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
                inv[i][j] = (a[i1][j1] * a[i2][j2] - a[i1][j2] * a[i2][j1]) / det;
            }
        const int G = std::max(1, std::min(128, (int)std::lround(std::cbrt(N / 64.0))));
        std::vector<int> cnt((size_t)G * G * G, 0);
        for (int n = 0; n < N; ++n) {
            int c[3];
            for (int d = 0; d < 3; ++d) {
                double f = inv[d][0] * pos[3 * (size_t)n] + inv[d][1] * pos[3 * (size_t)n + 1] +
                           inv[d][2] * pos[3 * (size_t)n + 2];
                f -= std::floor(f);
                c[d] = std::min(G - 1, std::max(0, (int)(f * G)));
            }
            cnt[((size_t)c[2] * G + c[1]) * G + c[0]]++;
        }
        const double mean = (double)N / cnt.size();
        double var = 0;
        for (int v : cnt) var += (v - mean) * (v - mean);
        var /= cnt.size();
        f4[0] = std::log10((double)N);
        f4[1] = 0.5 * std::log10(emax / emin);
        f4[2] = std::log10(std::max(var / mean, 1e-3));
        f4[3] = n_digits;
    });
}

// The configuration the model predicts fastest among those whose predicted error is at most
// twice the default configuration's, kept only when its predicted time is below 0.85 x the
// default's: out3 = (n_per_leaf, nz, is_float); pred4 (optional) = predicted (ms, error) of the
// choice and of the default.  Returns the number of candidates.
int pdmk_tune_choose(const float *blob, int blob_len, const double *f4, int *out3, double *pred4) {
    using namespace pdmk::gpu;
    int n_cand = -1;
    const int rc = guarded("pdmk_tune_choose", [&] {
        if (!blob || !f4 || !out3) throw std::invalid_argument("null argument");
        if (blob_len != kBlobHeader + kParams) throw std::invalid_argument("model blob has the wrong length");
        const float *xm = blob, *xs = blob + kIn, *ym = blob + 2 * kIn, *ys = blob + 2 * kIn + kOut;
        const std::vector<Candidate> C = candidates((int)std::lround(f4[3]));
        const int B = (int)C.size();
        std::vector<float> X((size_t)B * kIn), Y((size_t)B * kOut);
        for (int b = 0; b < B; ++b) config_row(f4, C[b], xm, xs, &X[(size_t)b * kIn]);
        Mlp m(B);
        m.upload(m.P.as<float>(), blob + kBlobHeader, kParams);
        m.upload(m.X.as<float>(), X.data(), X.size());
        m.forward();
        check(hipMemcpyAsync(Y.data(), m.Y.p, sizeof(float) * Y.size(), hipMemcpyDeviceToHost, m.s), "D2H");
        check(hipStreamSynchronize(m.s), "hipStreamSynchronize");
        auto log_ms = [&](int b) { return (double)Y[2 * b] * ys[0] + ym[0]; };
        auto log_err = [&](int b) { return (double)Y[2 * b + 1] * ys[1] + ym[1]; };
        const double err_cap = log_err(0) + std::log10(2.0);
        int best = 0;
        for (int b = 1; b < B; ++b)
            if (log_err(b) <= err_cap && log_ms(b) < log_ms(best)) best = b;
        if (log_ms(best) > log_ms(0) + std::log10(kGainMargin)) best = 0;
        out3[0] = C[best].npl;
        out3[1] = C[best].nz;
        out3[2] = C[best].is_float;
        if (pred4) {
            pred4[0] = std::pow(10.0, log_ms(best));
            pred4[1] = std::pow(10.0, log_err(best));
            pred4[2] = std::pow(10.0, log_ms(0));
            pred4[3] = std::pow(10.0, log_err(0));
        }
        n_cand = B;
    });
    return rc == 0 ? n_cand : -1;
}

int pdmk_tune_blob_len(void) { return pdmk::gpu::kBlobHeader + pdmk::gpu::kParams; }

int pdmk_mlp_n_params(void) { return pdmk::gpu::kParams; }

int pdmk_mlp_forward(const float *params, int B, const float *X, float *Y) {
    using namespace pdmk::gpu;
    return guarded("pdmk_mlp_forward", [&] {
        if (!params || !X || !Y) throw std::invalid_argument("null argument");
        Mlp m(B);
        m.upload(m.P.as<float>(), params, kParams);
        m.upload(m.X.as<float>(), X, (size_t)B * kIn);
        m.forward();
        check(hipMemcpyAsync(Y, m.Y.p, sizeof(float) * (size_t)B * kOut, hipMemcpyDeviceToHost, m.s), "D2H");
        check(hipStreamSynchronize(m.s), "hipStreamSynchronize");
    });
}

int pdmk_mlp_loss_grad(const float *params, int B, const float *X, const float *T, const float *M, float *loss,
                       float *grad) {
    using namespace pdmk::gpu;
    return guarded("pdmk_mlp_loss_grad", [&] {
        if (!params || !X || !T || !M || !loss || !grad) throw std::invalid_argument("null argument");
        const float inv = inv_mask_count(M, B * kOut);
        Mlp m(B);
        m.upload(m.P.as<float>(), params, kParams);
        m.upload(m.X.as<float>(), X, (size_t)B * kIn);
        m.prepare_targets(T, M);
        m.backward(inv);
        check(hipMemcpyAsync(grad, m.G.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, m.s), "D2H");
        check(hipMemcpyAsync(loss, m.loss.p, sizeof(float), hipMemcpyDeviceToHost, m.s), "D2H");
        check(hipStreamSynchronize(m.s), "hipStreamSynchronize");
    });
}

int pdmk_mlp_train(float *params, int B, const float *X, const float *T, const float *M, int steps, float lr,
                   float weight_decay, float *loss_history) {
    using namespace pdmk::gpu;
    return guarded("pdmk_mlp_train", [&] {
        if (!params || !X || !T || !M) throw std::invalid_argument("null argument");
        if (steps < 0) throw std::invalid_argument("steps must be non-negative");
        const float inv = inv_mask_count(M, B * kOut);
        Mlp m(B);
        m.Mo.alloc(sizeof(float) * kParams);
        m.Vo.alloc(sizeof(float) * kParams);
        const std::vector<float> zeros(kParams, 0.0f);
        m.upload(m.Mo.as<float>(), zeros.data(), kParams);
        m.upload(m.Vo.as<float>(), zeros.data(), kParams);
        m.upload(m.P.as<float>(), params, kParams);
        m.upload(m.X.as<float>(), X, (size_t)B * kIn);
        m.prepare_targets(T, M);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        std::vector<float> hist(steps > 0 ? steps : 1);
        for (int t = 1; t <= steps; ++t) {
            m.backward(inv);
            const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, t)));
            const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, t)));
            hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, m.s, m.P.as<float>(),
                               m.G.as<float>(), m.Mo.as<float>(), m.Vo.as<float>(), kParams, lr, b1, b2, eps,
                               weight_decay, c1, c2);
            check(hipGetLastError(), "k_adam");
            if (loss_history)
                check(hipMemcpyAsync(&hist[t - 1], m.loss.p, sizeof(float), hipMemcpyDeviceToHost, m.s), "D2H");
        }
        check(hipMemcpyAsync(params, m.P.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, m.s), "D2H");
        check(hipStreamSynchronize(m.s), "hipStreamSynchronize");
        if (loss_history)
            for (int t = 0; t < steps; ++t) loss_history[t] = hist[t];
    });
}

}  // extern "C"
