"""The launch-shape autotuner (pdmk_gpu_autotune / pdmk_gpu_knobs).

* Every knob value gives the potentials of the built-in launch shapes (the knobs change how
  work is cut into workgroups, not the sums: FP64 to 1e-12, FP32 to 1e-6 relative);
* values outside a knob's candidates are refused;
* autotune() keeps the built-in shapes unless a candidate is faster, so its tuned time is at
  most the default time, and the potentials after tuning equal those before;
* the choice is cached per plan kind: a new handle of the same kind starts with it.
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision,digits,tol", [("float", 6, 1e-6), ("double", 9, 1e-12), ("double", 6, 1e-12)])
def test_knobs_keep_potentials(pdmk, cuda, precision, digits, tol):
    cell = W.triclinic_cell()
    pos = W.uniform_positions(20_000, 8, cell)
    q = W.neutral_uniform_charges(20_000, 9)
    s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=100, nz=4, precision=precision)
    assert s.knobs() == {"near_wpb": 0, "near_late": -1, "p2pw_bk": 0, "pw2p_bk": 0, "t3_div": 1}
    ref = s.evaluate(q).astype(np.float64)
    for knob, values in (("near_wpb", (2, 8, 16)), ("near_late", (0, 1)), ("p2pw_bk", (16, 32, 64)),
                         ("pw2p_bk", (16, 32, 64)), ("t3_div", (2, 4))):
        for v in values:
            s.knobs(**{knob: v})
            assert rel(s.evaluate(q), ref) <= tol, (knob, v)
        s.knobs(near_wpb=0, near_late=-1, p2pw_bk=0, pw2p_bk=0, t3_div=1)
    for bad in ({"near_wpb": 3}, {"p2pw_bk": 48}, {"t3_div": 3}, {"near_late": 2}):
        with pytest.raises(ValueError):
            s.knobs(**bad)
    s.close()


def test_autotune_never_slower_and_cached(pdmk, cuda):
    cell = W.cubic_cell()
    pos = W.uniform_positions(50_000, 10, cell)
    q = W.neutral_uniform_charges(50_000, 11)
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=300, nz=4, precision="float")
    ref = s.evaluate(q).astype(np.float64)
    r = s.autotune(repeats=3)
    print("autotune:", r, flush=True)
    assert r["ms_tuned"] <= r["ms_default"]
    assert s.knobs() == r["knobs"]
    assert rel(s.evaluate(q), ref) <= 1e-6
    t = pdmk.PDMKSolver(cell, W.uniform_positions(50_000, 12, cell), n_digits=6, n_per_leaf=300, nz=4,
                        precision="float")
    assert t.knobs() == r["knobs"]
    s.close()
    t.close()
    pdmk.lib().pdmk_gpu_tune_cache_clear()
    u = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=300, nz=4, precision="float")
    assert u.knobs()["t3_div"] == 1 and u.knobs()["near_wpb"] == 0
    u.close()
