"""The learned cost model's GPU kernels against the FP64 CPU restatement (tests/mlp_fp64.py).

* forward (pdmk_mlp_forward) equals the FP64 forward to FP32 rounding (1e-5 of the output
  scale);
* the loss and its gradient (pdmk_mlp_loss_grad: forward, masked MSE, backward) equal the
  FP64 backpropagation to 1e-4 of the gradient's largest entry;
* Adam training (pdmk_mlp_train) drives the loss of a smooth regression down by 10x and is
  deterministic run to run.
"""
import numpy as np
import pytest

import mlp_fp64 as R

pytestmark = pytest.mark.gpu


def _data(rng, B):
    X = rng.normal(size=(B, 7))
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2], np.tanh(X[:, 3] - X[:, 4]) + 0.1 * X[:, 5]], axis=1)
    M = (rng.uniform(size=(B, 2)) < 0.9).astype(float)
    return X, T, M


def test_forward_matches_fp64(pdmk, cuda):
    from periodicdmk_amd import costmodel
    rng = np.random.default_rng(1)
    p = costmodel.init_params(3)
    X, _, _ = _data(rng, 333)
    y = costmodel.forward(p, X).astype(np.float64)
    ref, _, _ = R.forward(p.astype(np.float64), X.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(y - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("B", [1, 37, 1000])
def test_gradient_matches_fp64(pdmk, cuda, B):
    from periodicdmk_amd import costmodel
    rng = np.random.default_rng(B)
    p = costmodel.init_params(B)
    X, T, M = _data(rng, B)
    if M.sum() == 0:
        M[0, 0] = 1.0
    loss, g = costmodel.loss_grad(p, X, T, M)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    loss_ref, g_ref = R.loss_grad(p.astype(np.float64), f32(X), f32(T), M)
    assert abs(loss - loss_ref) <= 1e-5 * max(1.0, abs(loss_ref))
    err = np.max(np.abs(g.astype(np.float64) - g_ref)) / np.max(np.abs(g_ref))
    assert err <= 1e-4, err


def test_training_reduces_loss_and_is_deterministic(pdmk, cuda):
    from periodicdmk_amd import costmodel
    rng = np.random.default_rng(11)
    X, T, M = _data(rng, 2000)
    p0 = costmodel.init_params(5)
    p1, h1 = costmodel.train_params(p0, X, T, M, 600, lr=3e-3)
    p2, h2 = costmodel.train_params(p0, X, T, M, 600, lr=3e-3)
    assert np.array_equal(p1, p2) and np.array_equal(h1, h2)
    assert h1[-1] < 0.1 * h1[0], (h1[0], h1[-1])
    # the reported loss is the loss of the parameters before each step
    loss_end, _ = costmodel.loss_grad(p1, X, T, M)
    assert loss_end <= h1[-1] * 1.05


def test_trained_model_gain(pdmk, cuda, tmp_path):
    """What the trained network gains over the defaults: 6 new random problems (generator seed
    not used for the training solves, N >= 2e5 so that the times are stable), each solved
    with the model's choice and with the defaults; the model's total time is at least 5% lower
    and every choice keeps the error within 2x of the defaults' (against the FP64 Ewald sum
    at sampled points)."""
    from periodicdmk_amd import costmodel
    blob = costmodel.load_model()
    assert blob is not None, "the trained cost model ships with the package"
    recs = costmodel.heldout(6, 271828, str(tmp_path / "gain.jsonl"), blob=blob, n_min=2e5)
    assert len(recs) == 6 and all("failed" not in r for r in recs)
    t_def = sum(r["default"]["ms"] for r in recs)
    t_mod = sum(r["model"]["ms"] for r in recs)
    print(f"model {t_mod:.2f} ms, defaults {t_def:.2f} ms", flush=True)
    for r in recs:
        assert r["model"]["err"] <= 2.0 * r["default"]["err"] + 1e-15, r
    assert t_def >= 1.05 * t_mod
