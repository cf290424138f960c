"""CPU checks of the learned cost model's pieces that need no GPU.

* The FP64 restatement of the MLP (tests/mlp_fp64.py), the checker of the GPU kernels, has
  the gradient of its loss: central differences in FP64 agree with backpropagation to 1e-7.
* Problem features (pdmk_tune_features, host code of libpdmk.so): log10 N, the cell's
  condition number (1 for a cubic cell, the singular values of a triangular cell by numpy),
  and the clustering index log10(var / mean) of grid counts, about 0 for uniform points
  (Poisson counts) and well above 0 for clustered ones.
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W
import mlp_fp64 as R


def test_fp64_mlp_gradient_is_the_derivative():
    rng = np.random.default_rng(3)
    from periodicdmk_amd import costmodel
    p = costmodel.init_params(7).astype(np.float64) * 0.5
    X = rng.normal(size=(12, 7))
    T = rng.normal(size=(12, 2))
    M = (rng.uniform(size=(12, 2)) < 0.8).astype(float)
    _, g = R.loss_grad(p, X, T, M)
    idx = rng.choice(p.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (R.loss_grad(p + e, X, T, M)[0] - R.loss_grad(p - e, X, T, M)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_features(pdmk):
    from periodicdmk_amd import costmodel
    n = 200_000
    f = costmodel.features(W.cubic_cell(), W.uniform_positions(n, 1, W.cubic_cell()), 6)
    assert abs(f[0] - np.log10(n)) < 1e-12 and abs(f[1]) < 1e-12 and f[3] == 6
    assert abs(f[2]) < 0.05  # Poisson counts: var / mean = 1
    cell = W.triclinic_cell()
    f = costmodel.features(cell, W.gaussian_mixture(n, 8, 0.02, 5, cell), 9)
    sv = np.linalg.svd(cell, compute_uv=False)
    assert abs(f[1] - np.log10(sv[0] / sv[-1])) < 1e-10
    assert f[2] > 1.0 and f[3] == 9


def test_defaults_without_a_model(pdmk, tmp_path):
    from periodicdmk_amd import costmodel
    for digits, prec in ((6, "float"), (9, "double")):
        ch = costmodel.choose(W.cubic_cell(), W.uniform_positions(1000, 1, W.cubic_cell()), digits,
                              path=str(tmp_path / "absent.json"))
        assert (ch["n_per_leaf"], ch["nz"], ch["precision"], ch["model"]) == (300, 4, prec, False)
