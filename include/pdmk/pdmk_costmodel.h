/// Learned cost model: picks the tree constructor's n_per_leaf and nz and the precision of a
/// problem (the reference's callers pass them by hand; its only tuning is the hand-measured
/// K_opt table, cpp/include/dmk/pbc.hpp:463-495, kept as the defaults here).  No reference
/// counterpart.
///
/// The model is an FP32 MLP 7 -> 64 -> 64 -> 64 -> 2 (GELU) trained on solver-timed solves:
/// inputs are four problem features and one candidate configuration (log2 n_per_leaf, nz, FP32
/// kernels or not), outputs the standardised log10 evaluate() milliseconds and log10 relative
/// error against the FP64 Ewald sum.  A model blob is pdmk_tune_blob_len() floats: input means
/// and deviations (7 + 7), output means and deviations (2 + 2), then the pdmk_mlp_n_params()
/// network parameters.  The shipped blob is periodicdmk_amd/cost_model.json.
///
/// All functions return 0 on success (pdmk_tune_choose: the number of candidates) and -1 on a
/// NULL argument, a size out of range or a HIP error (message on stderr).
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/// Features of a problem, computed on the host: f4 = {log10 N, log10 of the cell matrix's
/// 2-norm condition number, clustering index log10(var / mean) of the point counts over a
/// G^3 grid of fractional coordinates (G^3 about N / 64; 0 for uniform points), n_digits}.
/// cell: 3x3 column-major lattice vectors; pos: N x 3 row-major.
int pdmk_tune_features(const double* cell, int N, const double* pos, int n_digits, double* f4);

/// The configuration the model predicts fastest among n_per_leaf in {100, 150, 200, 300, 400,
/// 600, 800}, nz in {4..8} and the precision (FP32 kernels only up to 6 digits) whose
/// predicted error is at most twice the default configuration's, kept only when its predicted
/// time is below 0.85 x the default's: out3 = {n_per_leaf, nz, 1 if FP32 kernels};
/// pred4 (optional) = predicted {ms, error} of the choice and of the default.
int pdmk_tune_choose(const float* blob, int blob_len, const double* f4, int* out3, double* pred4);

/// Length in floats of a model blob.
int pdmk_tune_blob_len(void);

/// Number of network parameters (weights and biases of the four layers).
int pdmk_mlp_n_params(void);

/// Forward pass on the GPU: Y[B][2] for X[B][7] (standardised inputs).
int pdmk_mlp_forward(const float* params, int B, const float* X, float* Y);

/// Masked mean-squared error of the forward pass against T[B][2] (M[B][2] = 1 where a target
/// is present) and its gradient with respect to every parameter (grad[pdmk_mlp_n_params()]),
/// summed over the batch in row order (deterministic).
int pdmk_mlp_loss_grad(const float* params, int B, const float* X, const float* T, const float* M, float* loss,
                       float* grad);

/// `steps` full-batch AdamW steps (beta1 0.9, beta2 0.999, eps 1e-8) on the GPU, params
/// updated in place; loss_history (optional, steps floats) receives the loss before each step.
int pdmk_mlp_train(float* params, int B, const float* X, const float* T, const float* M, int steps, float lr,
                   float weight_decay, float* loss_history);

#ifdef __cplusplus
}
#endif
