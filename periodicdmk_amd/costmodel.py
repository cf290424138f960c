"""Learned cost model: picks n_per_leaf, nz and the precision mode of a solve.

A small FP32 MLP (7 inputs, three hidden layers of 64 with GELU, 2 outputs) predicts the
log10 evaluate() time and the log10 error against the FP64 Ewald sum of a solve from the
problem's features (log10 N, log10 cell condition number, clustering index, n_digits; see
``pdmk_tune_features``) and one candidate configuration (log2 n_per_leaf, nz, FP32 or FP64
kernels).  Training and inference run on the GPU in the library's own kernels
(``pdmk_mlp_train``, ``pdmk_mlp_forward``; csrc/hip/costmodel.hip).  ``choose`` returns the
configuration predicted fastest among those with nz >= 4 whose predicted error is at most
twice the default configuration's (n_per_leaf 300, nz 4, FP32 kernels up to 6 digits as
AutoPeriodicDMKTree), provided its predicted time is below 0.85 x the default's (else the
defaults); without a trained model it returns the defaults.  (The training solves sample
nz from 2; below 4 the measured error grew by up to 40x, so the choice keeps to nz >= 4.)

The solver-timed solves it was trained on are kept in ``cost_model_solves/`` (JSON lines).

The model ships as ``cost_model.json`` beside this file: the input and output
standardisation (means and standard deviations) and the flat parameter array.

Command line (all on the GPU):
  python -m periodicdmk_amd.costmodel generate --problems K --seed S --out FILE
      solver-timed solves of random problems (JSON lines, one per solve)
  python -m periodicdmk_amd.costmodel train --data FILE [FILE ...] --out cost_model.json
  python -m periodicdmk_amd.costmodel heldout --problems K --seed S --out FILE
      model choice against the defaults on new problems
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import lib, PDMKSolver, DeviceArray, ewald_targets, synchronize
from . import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "cost_model.json")
N_IN, N_OUT = 7, 2
DEFAULT = (300, 4)
NPL_CHOICES = (100, 150, 200, 300, 400, 600, 800)
NZ_CHOICES = tuple(range(2, 9))        # the training solves sample these
NZ_CHOICES_PICK = tuple(range(4, 9))   # and the choice keeps to nz >= 4 (equal accuracy)
GAIN_MARGIN = 0.85  # leave the defaults only for a predicted time below 0.85 x theirs
DATA_DIR = os.path.join(HERE, "cost_model_solves")

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


def _sig():
    L = lib()
    for name, res, args in (
            ("pdmk_mlp_n_params", C.c_int, []),
            ("pdmk_tune_blob_len", C.c_int, []),
            ("pdmk_mlp_forward", C.c_int, [_fp, C.c_int, _fp, _fp]),
            ("pdmk_mlp_loss_grad", C.c_int, [_fp, C.c_int, _fp, _fp, _fp, _fp, _fp]),
            ("pdmk_mlp_train", C.c_int, [_fp, C.c_int, _fp, _fp, _fp, C.c_int, C.c_float, C.c_float, _fp]),
            ("pdmk_tune_features", C.c_int, [_dp, C.c_int, _dp, C.c_int, _dp]),
            ("pdmk_tune_choose", C.c_int, [_fp, C.c_int, _dp, C.POINTER(C.c_int), _dp])):
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def n_params() -> int:
    return int(_sig().pdmk_mlp_n_params())


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initial weights (normal, std sqrt(2 / fan_in)) and zero biases, layer by layer."""
    rng = np.random.default_rng(seed)
    dims = (N_IN, 64, 64, 64, N_OUT)
    out = []
    for i in range(4):
        out.append(rng.normal(0.0, np.sqrt(2.0 / dims[i]), dims[i + 1] * dims[i]))
        out.append(np.zeros(dims[i + 1]))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == n_params()
    return p


def forward(params, X) -> np.ndarray:
    """MLP outputs (B, 2) for standardized inputs X (B, 7), on the GPU."""
    L = _sig()
    p, pp = _f32(params)
    x, xp = _f32(X)
    y = np.zeros((x.shape[0], N_OUT), dtype=np.float32)
    if L.pdmk_mlp_forward(pp, x.shape[0], xp, y.ctypes.data_as(_fp)) != 0:
        raise RuntimeError("pdmk_mlp_forward failed")
    return y


def loss_grad(params, X, T, M):
    """Masked mean squared error and its gradient with respect to the parameters, on the GPU."""
    L = _sig()
    p, pp = _f32(params)
    x, xp = _f32(X)
    t, tp = _f32(T)
    m, mp = _f32(M)
    g = np.zeros(p.size, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    if L.pdmk_mlp_loss_grad(pp, x.shape[0], xp, tp, mp, loss.ctypes.data_as(_fp), g.ctypes.data_as(_fp)) != 0:
        raise RuntimeError("pdmk_mlp_loss_grad failed")
    return float(loss[0]), g


def train_params(params, X, T, M, steps: int, lr: float = 2e-3, weight_decay: float = 1e-4):
    """Adam steps on the GPU (full batch); returns (params, loss per step)."""
    L = _sig()
    p = np.array(params, dtype=np.float32, copy=True)
    x, xp = _f32(X)
    t, tp = _f32(T)
    m, mp = _f32(M)
    hist = np.zeros(max(1, steps), dtype=np.float32)
    if L.pdmk_mlp_train(p.ctypes.data_as(_fp), x.shape[0], xp, tp, mp, int(steps), float(lr), float(weight_decay),
                        hist.ctypes.data_as(_fp)) != 0:
        raise RuntimeError("pdmk_mlp_train failed")
    return p, hist[:steps]


def features(cell, positions, n_digits) -> np.ndarray:
    """(log10 N, log10 cond(cell), clustering index, n_digits) of a problem (host code)."""
    L = _sig()
    cellc = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3).T)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
    f = np.zeros(4)
    if L.pdmk_tune_features(cellc.ctypes.data_as(_dp), pos.shape[0], pos.ctypes.data_as(_dp), int(n_digits),
                            f.ctypes.data_as(_dp)) != 0:
        raise RuntimeError("pdmk_tune_features failed")
    return f


def config_inputs(f4, npl, nz, is_float) -> np.ndarray:
    return np.array([f4[0], f4[1], f4[2], f4[3], np.log2(npl), nz, 1.0 if is_float else 0.0])


def targets(ms, err) -> np.ndarray:
    return np.array([np.log10(ms), np.log10(max(err, 1e-16))])


def fit(rows, steps: int = 800, seed: int = 0, lr: float = 1e-3):
    """Train on solve records (dicts with features, npl, nz, is_float, ms, err): returns the
    model blob (standardisation + parameters) and the loss history."""
    X = np.array([config_inputs(r["features"], r["npl"], r["nz"], r["is_float"]) for r in rows])
    Y = np.array([targets(r["ms"], r["err"]) for r in rows])
    M = np.ones_like(Y)
    xm, xs = X.mean(0), X.std(0)
    xs[xs < 1e-6] = 1.0
    ym, ys = Y.mean(0), Y.std(0)
    ys[ys < 1e-6] = 1.0
    p, hist = train_params(init_params(seed), (X - xm) / xs, (Y - ym) / ys, M, steps, lr)
    blob = np.concatenate([xm, xs, ym, ys, p]).astype(np.float32)
    return blob, hist


def save_model(blob, path=MODEL_PATH, meta=None):
    with open(path, "w") as f:
        json.dump({"layout": "x_mean[7] x_std[7] y_mean[2] y_std[2] params", "meta": meta or {},
                   "blob": [float(v) for v in blob]}, f)


def load_model(path=MODEL_PATH):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return np.array(json.load(f)["blob"], dtype=np.float32)


def predict(blob, f4, configs):
    """Predicted (ms, error) for each (npl, nz, is_float) of ``configs``."""
    xm, xs, ym, ys, p = blob[:7], blob[7:14], blob[14:16], blob[16:18], blob[18:]
    X = np.array([config_inputs(f4, *c) for c in configs])
    Y = forward(p, (X - xm) / xs) * ys + ym
    return 10.0 ** Y[:, 0], 10.0 ** Y[:, 1]


def choose(cell, positions, n_digits, blob=None, path=MODEL_PATH, precision=None):
    """The model's configuration for a problem: dict(n_per_leaf, nz, precision, predicted
    ms and error of the choice and of the defaults).  The model is ``blob``, else the one saved
    at ``path``; the defaults when there is none.  ``precision`` ("float" / "double") fixes the
    kernels' precision and chooses n_per_leaf and nz only (the same rule, on the host)."""
    blob = load_model(path) if blob is None else blob
    f4 = features(cell, positions, n_digits)
    prec_default = precision or ("float" if n_digits <= 6 else "double")
    if blob is None:
        return {"n_per_leaf": DEFAULT[0], "nz": DEFAULT[1], "precision": prec_default, "model": False}
    if precision is not None:
        fl = 1 if precision == "float" else 0
        cands = [(DEFAULT[0], DEFAULT[1], fl)] + [(a, b, fl) for a in NPL_CHOICES for b in NZ_CHOICES_PICK
                                                  if (a, b) != DEFAULT]
        ms, err = predict(blob, f4, cands)
        ok = [i for i in range(len(cands)) if err[i] <= 2.0 * err[0]]
        b = min(ok, key=lambda i: ms[i])
        if ms[b] > GAIN_MARGIN * ms[0]:
            b = 0
        return {"n_per_leaf": cands[b][0], "nz": cands[b][1], "precision": precision, "model": True,
                "pred_ms": float(ms[b]), "pred_err": float(err[b]), "pred_default_ms": float(ms[0]),
                "pred_default_err": float(err[0]), "features": f4.tolist()}
    L = _sig()
    b, bp = _f32(blob)
    out = (C.c_int * 3)()
    pred = np.zeros(4)
    if L.pdmk_tune_choose(bp, b.size, f4.ctypes.data_as(_dp), out, pred.ctypes.data_as(_dp)) < 0:
        raise RuntimeError("pdmk_tune_choose failed")
    return {"n_per_leaf": out[0], "nz": out[1], "precision": "float" if out[2] else "double", "model": True,
            "pred_ms": pred[0], "pred_err": pred[1], "pred_default_ms": pred[2], "pred_default_err": pred[3],
            "features": f4.tolist()}


# ------------------------------------------------------------------ solver-timed solves
def random_problem(rng, n_min: float = 1e4):
    """A random cell (cubic, or triclinic with off-diagonals U[0, 0.4] and diagonal U[0.7, 1.4]),
    N log-uniform in [n_min, 4e6] (at most 1e6 at 12 digits, 2e6 at 9), positions from a mixture of
    1, 8, 64 or 512 Gaussians (sigma log-uniform in [0.005, 0.05] of the cell; 1 = uniform),
    neutral U[-1, 1] charges and n_digits in {3, 6, 9, 12}."""
    digits = int(rng.choice([3, 6, 9, 12]))
    n_max = {3: 4e6, 6: 4e6, 9: 2e6, 12: 1e6}[digits]
    n = int(np.exp(rng.uniform(np.log(min(n_min, n_max)), np.log(n_max))))
    cell = np.eye(3)
    if rng.uniform() < 0.5:
        cell = np.diag(rng.uniform(0.7, 1.4, 3))
        cell[0, 1], cell[0, 2], cell[1, 2] = rng.uniform(0, 0.4, 3)
    k = int(rng.choice([1, 8, 64, 512]))
    seed = int(rng.integers(1, 2 ** 31))
    if k == 1:
        pos = W.uniform_positions(n, seed, cell)
    else:
        sigma = float(np.exp(rng.uniform(np.log(0.005), np.log(0.05))))
        pos = W.gaussian_mixture(n, k, sigma, seed, cell)
    q = W.neutral_uniform_charges(n, seed + 1)
    return {"cell": cell, "pos": pos, "q": q, "n_digits": digits, "clusters": k, "seed": seed}


def time_solve(prob, npl, nz, precision, calls: int = 3, ewald=None, targets_idx=None):
    """Build a tree with (npl, nz, precision), time evaluate() on device buffers (median of
    ``calls`` after one warm-up) and measure the error at the sampled targets against the FP64
    Ewald potentials ``ewald``."""
    t0 = time.perf_counter()
    s = PDMKSolver(prob["cell"], prob["pos"], n_digits=prob["n_digits"], n_per_leaf=npl, nz=nz, precision=precision)
    build_s = time.perf_counter() - t0
    n = s.n
    qd = DeviceArray.from_numpy(prob["q"])
    out = DeviceArray(n, s.dtype)
    try:
        s.evaluate_device(qd.ptr, out.ptr)
        synchronize()
        ts = []
        for _ in range(calls):
            t0 = time.perf_counter()
            s.evaluate_device(qd.ptr, out.ptr)
            synchronize()
            ts.append(time.perf_counter() - t0)
        pot = out.numpy().astype(np.float64)
    finally:
        qd.free()
        out.free()
        s.close()
    d = pot[targets_idx] - ewald
    err = float(np.linalg.norm(d) / np.linalg.norm(ewald))
    return {"ms": 1e3 * float(np.median(ts)), "err": err, "build_s": build_s}


def _configs_for(rng, digits, k):
    cands = [(npl, nz, f) for f in ((1, 0) if digits <= 6 else (0,)) for npl in NPL_CHOICES for nz in NZ_CHOICES]
    pick = [cands[i] for i in rng.choice(len(cands), size=min(k, len(cands)), replace=False)]
    default = (DEFAULT[0], DEFAULT[1], 1 if digits <= 6 else 0)
    return [default] + [c for c in pick if c != default]


def _reference(prob, m=64):
    n = prob["pos"].shape[0]
    tg = np.sort(np.random.default_rng(prob["seed"]).choice(n, min(n, m), replace=False)).astype(np.int32)
    return tg, ewald_targets(prob["cell"], prob["pos"], prob["q"], tg, s=5.5)


def generate(n_problems: int, seed: int, out_path: str, configs_per_problem: int = 5, budget_s: float = 1e9,
             n_min: float = 1e4):
    rng = np.random.default_rng(seed)
    t_start = time.time()
    with open(out_path, "a") as f:
        for i in range(n_problems):
            if time.time() - t_start > budget_s:
                print(f"[costmodel] time budget reached after {i} problems", flush=True)
                break
            prob = random_problem(rng, n_min)
            f4 = features(prob["cell"], prob["pos"], prob["n_digits"])
            tg, ew = _reference(prob)
            for npl, nz, fl in _configs_for(rng, prob["n_digits"], configs_per_problem):
                rec = {"problem": i, "seed": prob["seed"], "n": int(prob["pos"].shape[0]), "clusters": prob["clusters"],
                       "n_digits": prob["n_digits"], "features": f4.tolist(), "npl": npl, "nz": nz, "is_float": fl}
                try:
                    rec.update(time_solve(prob, npl, nz, "float" if fl else "double", ewald=ew, targets_idx=tg))
                except RuntimeError as e:
                    rec["failed"] = str(e)
                f.write(json.dumps(rec) + "\n")
                f.flush()
            print(f"[costmodel] problem {i}: N={prob['pos'].shape[0]} d={prob['n_digits']} "
                  f"k={prob['clusters']} ({time.time() - t_start:.0f} s)", flush=True)


def heldout(n_problems: int, seed: int, out_path: str, blob=None, budget_s: float = 1e9, n_min: float = 1e4):
    """Model choice against the defaults on new problems: time and error of both (one JSON
    line per problem appended to out_path; returns the records)."""
    blob = load_model() if blob is None else blob
    rng = np.random.default_rng(seed)
    t_start = time.time()
    recs = []
    with open(out_path, "a") as f:
        for i in range(n_problems):
            if time.time() - t_start > budget_s:
                print(f"[costmodel] time budget reached after {i} problems", flush=True)
                break
            prob = random_problem(rng, n_min)
            ch = choose(prob["cell"], prob["pos"], prob["n_digits"], blob)
            tg, ew = _reference(prob)
            prec_d = "float" if prob["n_digits"] <= 6 else "double"
            rec = {"problem": i, "seed": prob["seed"], "n": int(prob["pos"].shape[0]), "clusters": prob["clusters"],
                   "n_digits": prob["n_digits"], "choice": [ch["n_per_leaf"], ch["nz"], ch["precision"]]}
            try:
                rec["default"] = time_solve(prob, DEFAULT[0], DEFAULT[1], prec_d, ewald=ew, targets_idx=tg)
                rec["model"] = (rec["default"] if (ch["n_per_leaf"], ch["nz"], ch["precision"]) ==
                                (DEFAULT[0], DEFAULT[1], prec_d) else
                                time_solve(prob, ch["n_per_leaf"], ch["nz"], ch["precision"], ewald=ew, targets_idx=tg))
            except RuntimeError as e:
                rec["failed"] = str(e)
            f.write(json.dumps(rec) + "\n")
            f.flush()
            recs.append(rec)
            print(f"[costmodel] held-out {i}: {rec.get('choice')} "
                  f"default {rec.get('default', {}).get('ms', float('nan')):.2f} ms "
                  f"model {rec.get('model', {}).get('ms', float('nan')):.2f} ms", flush=True)
    return recs


def read_records(paths):
    rows = []
    for p in paths:
        with open(p) as f:
            for line in f:
                r = json.loads(line)
                if "failed" not in r and r.get("ms", 0) > 0 and r.get("err", 0) > 0:
                    rows.append(r)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--problems", type=int, default=100)
    g.add_argument("--seed", type=int, default=1)
    g.add_argument("--configs", type=int, default=5)
    g.add_argument("--budget-s", type=float, default=1e9)
    g.add_argument("--n-min", type=float, default=1e4, help="smallest N of the random problems")
    g.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--steps", type=int, default=800)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--out", default=MODEL_PATH)
    h = sub.add_parser("heldout")
    h.add_argument("--problems", type=int, default=200)
    h.add_argument("--seed", type=int, default=1000003)
    h.add_argument("--out", required=True)
    h.add_argument("--budget-s", type=float, default=1e9)
    h.add_argument("--n-min", type=float, default=1e4, help="smallest N of the random problems")
    h.add_argument("--model", default=MODEL_PATH, help="model file (default: the shipped one)")
    a = ap.parse_args(argv)
    if a.cmd == "generate":
        generate(a.problems, a.seed, a.out, a.configs, a.budget_s, a.n_min)
    elif a.cmd == "train":
        rows = read_records(a.data)
        blob, hist = fit(rows, a.steps, lr=a.lr)
        save_model(blob, a.out, {"solves": len(rows), "steps": a.steps, "lr": a.lr, "final_loss": float(hist[-1]),
                                 "data": [os.path.basename(d) for d in a.data]})
        print(f"[costmodel] trained on {len(rows)} solves: loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    else:
        blob = load_model(a.model)
        if blob is None:
            sys.exit(f"no model at {a.model}")
        heldout(a.problems, a.seed, a.out, blob=blob, budget_s=a.budget_s, n_min=a.n_min)


if __name__ == "__main__":
    sys.exit(main())
