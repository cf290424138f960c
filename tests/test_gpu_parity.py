"""GPU parity of the MI355X evaluate() path against the CPU oracle and FP64 Ewald.

Each stage of evaluate() (bit 0 proxy/plane-wave part, bit 1 near field, bit 2 self
correction, bit 3 far field, 15 = all) is compared with the oracle's restatement of the
reference stage on the same seeded inputs:
  * FP64 kernels: relative L2 difference < 1e-12 (same tree, same tables up to rounding,
    different summation order only);
  * FP32 kernels (n_digits <= 6): relative L2 difference < 1e-5 against the FP64 oracle
    (FP32 positions, transforms and near-field sums).
End-to-end checks against Ewald follow cpp/tests/test_periodic_dmk.cpp:20-68 (< 1e-4).
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu

STAGES = (1, 2, 4, 8, 15)


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def gpu_stages(solver, q, stages):
    return solver.run_stages(q, stages).astype(np.float64)


def case(name):
    if name == "cubic":
        c = W.cubic_cell()
        return c, W.uniform_positions(3000, 1, c), W.neutral_uniform_charges(3000, 2), 40
    if name == "triclinic":
        c = W.triclinic_cell()
        return c, W.uniform_positions(4000, 3, c), W.neutral_uniform_charges(4000, 4), 40
    if name == "clustered":
        c = W.triclinic_cell()
        return c, W.gaussian_mixture(6000, 8, 0.02, 21, c), W.pm1_charges(6000, 22), 30
    if name == "single_level":  # N = 1e4, n_per_leaf 300: every leaf is a level-0 box
        c = W.cubic_cell()
        return c, W.uniform_positions(10_000, 1, c), W.neutral_uniform_charges(10_000, 2), 300
    if name == "rectangular":  # orthorhombic, sides 1 x 1.5 x 2 (cpp/tests/test_nonuniform.cpp rectangular cells)
        c = np.diag([1.0, 1.5, 2.0])
        return c, W.uniform_positions(4000, 12, c), W.neutral_uniform_charges(4000, 13), 40
    if name == "long_cell":  # |a2| > |a1|: the lattice frame reorders the axes
        c = np.array([[1.0, 0.0, 0.1], [0.0, 1.7, 0.0], [0.0, 0.2, 1.2]])
        return c, W.uniform_positions(3000, 9, c), W.neutral_uniform_charges(3000, 10), 40
    raise KeyError(name)


PARAMS = [("cubic", "double", d) for d in (3, 6, 9, 12)] + [
    ("cubic", "float", 6), ("cubic", "float", 3),
    ("triclinic", "double", 6), ("triclinic", "double", 9), ("triclinic", "float", 6),
    ("clustered", "double", 6), ("clustered", "double", 9), ("clustered", "float", 6),
    ("single_level", "double", 6), ("single_level", "float", 6),
    ("long_cell", "double", 6), ("rectangular", "double", 6), ("rectangular", "float", 6),
    ("rectangular", "double", 9),
] + [
    # the other digit counts: each has its own residual polynomial (and, below 6 digits, the
    # even/odd split evaluation) on one of the four proxy orders
    ("cubic", "double", d) for d in (1, 2, 4, 5, 7, 8, 10, 11)
] + [("cubic", "float", d) for d in (1, 2, 4, 5)]


@pytest.mark.parametrize("name,precision,digits", PARAMS)
def test_stage_parity(pdmk, oracle_mod, cuda, name, precision, digits):
    cell, pos, q, npl = case(name)
    s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, nz=4, precision=precision)
    o = oracle_mod.OracleTree(cell, pos, digits, npl, 4, precision="double")
    tol = 1e-12 if precision == "double" else 1e-5
    for st in STAGES:
        g = gpu_stages(s, q, st)
        r = o.evaluate(q, st)
        assert rel(g, r) < tol, (st, rel(g, r))
    s.close()


def test_tree_topology_matches_oracle(pdmk, oracle_mod, cuda):
    cell, pos, q, npl = case("clustered")
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, precision="double")
    box7, sp, tp, cen = s.tree_arrays()
    o = oracle_mod.OracleTree(cell, pos, 6, npl, 4)
    ob7, osp, otp, ocen = o.tree_arrays()
    np.testing.assert_array_equal(box7, ob7)
    np.testing.assert_array_equal(sp, osp)
    np.testing.assert_array_equal(tp, otp)


@pytest.mark.parametrize("precision,tol", [("double", 1e-5), ("float", 1e-4)])
def test_reference_cubic_case_vs_ewald(pdmk, oracle_mod, cuda, precision, tol):
    L = 3.0
    rng = np.random.default_rng(42)
    pos = rng.uniform(0.03 * L, 0.97 * L, (200, 3))
    q = rng.normal(size=200)
    q[0] -= q.sum()
    ref, _ = oracle_mod.ewald_direct(np.eye(3) * L, pos, q, 5.0)
    s = pdmk.PDMKSolver(np.eye(3) * L, pos, n_digits=6, n_per_leaf=20, precision=precision)
    assert rel(s.evaluate(q), ref) < tol


def test_reference_triclinic_case_vs_ewald(pdmk, oracle_mod, cuda):
    cell = np.array([[5.0, 1.0, 0.5], [0.0, 4.5, 0.8], [0.0, 0.0, 4.0]])
    rng = np.random.default_rng(77)
    pos = rng.uniform(0.01, 0.99, (100, 3)) @ cell.T
    q = rng.normal(size=100)
    q[0] -= q.sum()
    ref, _ = oracle_mod.ewald_direct(cell, pos, q, 5.0)
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=20, precision="double")
    assert rel(s.evaluate(q), ref) < 1e-5


def test_madelung_on_gpu(pdmk, cuda):
    pos = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5], [.5, .5, .5]])
    q = np.array([1, 1, 1, 1, -1, -1, -1, -1.0])
    for prec in ("double", "float"):
        s = pdmk.PDMKSolver(np.eye(3), pos, n_digits=6, n_per_leaf=4, precision=prec)
        m = -0.5 * np.dot(q, s.evaluate(q).astype(np.float64)) / 8
        assert abs(m - 1.7475645946331822) < 1e-4


def test_9_digits_vs_golden_ewald_cfg1(pdmk, cuda):
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg1_ewald.npz"))
    cell, pos, q, npl = case("single_level")
    s = pdmk.PDMKSolver(cell, pos, n_digits=9, n_per_leaf=npl, precision="double")
    assert rel(s.evaluate(q), g["ewald"]) < 1e-7
    s6 = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, precision="float")
    assert rel(s6.evaluate(q).astype(np.float64), g["ewald"]) < 1e-5


def test_repeat_evaluate_is_deterministic(pdmk, cuda):
    cell, pos, q, npl = case("cubic")
    rng = np.random.default_rng(5)
    q2 = rng.normal(size=q.shape)
    for prec in ("double", "float"):
        s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, precision=prec)
        p1, p2, p1b = s.evaluate(q), s.evaluate(q2), s.evaluate(q)
        np.testing.assert_array_equal(p1, p1b)
        assert np.sum(np.abs(p1 - p2)) > 1e-6


@pytest.mark.parametrize("digits,tol", [(6, 1e-5), (9, 1e-8)])
def test_energy_vs_ewald(pdmk, cuda, digits, tol):
    """Energy 1/2 sum q phi of clustered charges against the Ewald energy
    (cpp/tests/test_nonuniform.cpp:369-377)."""
    cell, pos, q, npl = case("clustered")
    s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, precision="double")
    e = 0.5 * np.dot(q, s.evaluate(q))
    s.close()
    _, e_ref = pdmk.ewald(cell, pos, q, 6.0)
    assert abs(e - e_ref) < tol * abs(e_ref), (e, e_ref)


def test_reused_tree_equals_fresh_tree(pdmk, cuda):
    """A tree that evaluated other charges gives the potentials of a fresh tree
    (cpp/tests/test_pbc_tree_update_charges.cpp:11-76: update_charges + eval = new tree)."""
    cell, pos, q, npl = case("triclinic")
    q2 = W.neutral_uniform_charges(pos.shape[0], 77)
    for prec in ("double", "float"):
        a = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, precision=prec)
        a.evaluate(q)
        reused = a.evaluate(q2)
        a.close()
        b = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, precision=prec)
        np.testing.assert_array_equal(reused, b.evaluate(q2))
        b.close()


def test_evaluate_timed_parts_cover_wall_time(pdmk, cuda):
    """evaluate_timed's three parts (far field, proxy / plane waves, near field) add up to
    the call's wall time (cpp/tests/test_evaluate_timed.cpp:47-48: sum in [0.95, 1] wall);
    N = 1e6 so that the host call overhead is small against the GPU time."""
    import time
    n = 1_000_000
    cell = W.cubic_cell()
    s = pdmk.PDMKSolver(cell, W.uniform_positions(n, 11, cell), n_digits=6, n_per_leaf=300, precision="float")
    qd = pdmk.DeviceArray.from_numpy(W.neutral_uniform_charges(n, 100))
    out = pdmk.DeviceArray(n, s.dtype)
    try:
        s.evaluate_timed(qd.ptr, out.ptr)
        ratios = []
        for _ in range(3):
            t0 = time.perf_counter()
            t = s.evaluate_timed(qd.ptr, out.ptr)  # synchronous
            wall = 1e3 * (time.perf_counter() - t0)
            ratios.append(sum(t.values()) / wall)
        r = max(ratios)
        assert 0.95 <= r <= 1.0, ratios
    finally:
        qd.free()
        out.free()
        s.close()


def test_deep_tree_parity(pdmk, oracle_mod, cuda):
    """A 3e-7 cluster refines to 23 levels (built with the points on the host): the full
    evaluate against the oracle (cpp/tests/test_nonuniform.cpp covers deep trees)."""
    cub = W.cubic_cell()
    pos = np.vstack([0.3 + 3e-7 * np.random.default_rng(3).random((600, 3)), W.uniform_positions(400, 8, cub)])
    q = W.neutral_uniform_charges(pos.shape[0], 9)
    o = oracle_mod.OracleTree(cub, pos, 6, 20, 4, precision="double")
    ref = o.evaluate(q)
    o.close()
    for prec, tol in (("double", 1e-12), ("float", 1e-5)):
        s = pdmk.PDMKSolver(cub, pos, n_digits=6, n_per_leaf=20, nz=4, precision=prec)
        assert s.stats()["n_levels"] >= 20
        assert rel(s.evaluate(q).astype(np.float64), ref) < tol
        s.close()


def test_device_entry_points(pdmk, cuda):
    import ctypes as C
    cell, pos, q, npl = case("cubic")
    for prec in ("double", "float"):
        s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, precision=prec)
        host = s.evaluate(q)
        qd = pdmk.DeviceArray.from_numpy(q)
        out = pdmk.DeviceArray(s.n, s.dtype)
        s.evaluate_device(qd.ptr, out.ptr, 0)
        pdmk.synchronize()
        np.testing.assert_array_equal(out.numpy(), host)
        t = s.evaluate_timed(qd.ptr, out.ptr)
        assert all(v >= 0 for v in t.values()) and sum(t.values()) > 0
        np.testing.assert_array_equal(out.numpy(), host)
        # a handle of the other precision is rejected
        other = pdmk.lib().pdmk_gpu_evaluate_double if prec == "float" else pdmk.lib().pdmk_gpu_evaluate_float
        assert other(s._h, C.c_void_p(qd.ptr), C.c_void_p(out.ptr), None) == -1


def test_edge_inputs(pdmk, oracle_mod, cuda):
    # one charge; points on the cell boundary; coincident points; non-neutral charges
    s = pdmk.PDMKSolver(np.eye(3) * 2.0, np.array([[0.5, 0.5, 0.5]]), n_digits=6, n_per_leaf=20, precision="double")
    assert np.all(np.isfinite(s.evaluate(np.array([1.0]))))
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.25, 0.25, 0.25], [0.25, 0.25, 0.25], [0.999, 0.1, 0.0]])
    q = np.array([1.0, -0.5, 0.25, 0.5, 0.3])
    for prec, tol in (("double", 1e-12), ("float", 1e-5)):
        s = pdmk.PDMKSolver(np.eye(3), pos, n_digits=6, n_per_leaf=2, precision=prec)
        o = oracle_mod.OracleTree(np.eye(3), pos, 6, 2, 4)
        assert rel(s.evaluate(q).astype(np.float64), o.evaluate(q)) < tol


def test_n_digits_below_one_builds_one_digit_tree(pdmk, cuda):
    # eps = 10^-max(1, n_digits) (cpp/src/periodic_dmk.cpp:52): n_digits 0 or -3 is 1 digit
    cell, pos, q, npl = case("cubic")
    ref = pdmk.PDMKSolver(cell, pos, n_digits=1, n_per_leaf=npl, precision="double").evaluate(q)
    for nd in (0, -3):
        got = pdmk.PDMKSolver(cell, pos, n_digits=nd, n_per_leaf=npl, precision="double").evaluate(q)
        np.testing.assert_array_equal(got, ref)
    with pytest.raises(RuntimeError):
        pdmk.PDMKSolver(cell, pos, n_digits=13, n_per_leaf=npl, precision="double")


def test_wrong_charge_count_raises(pdmk, cuda):
    cell, pos, q, npl = case("cubic")
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, precision="double")
    with pytest.raises(ValueError):
        s.evaluate(q[:-1])


def test_full_size_workload_parity_and_linearity(pdmk, oracle_mod, cuda):
    """N = 1e6 cubic workload: the FP64 path equals the oracle stage by stage (the far field,
    evaluated by the NUFFT at this size, to the reference's NUFFT tolerance eps/4 of the exact
    sum); the FP32 path is linear in q and agrees with FP64."""
    n = 1_000_000
    cell = W.cubic_cell()
    pos = W.uniform_positions(n, 11, cell)
    q1 = W.neutral_uniform_charges(n, 100)
    q2 = W.neutral_uniform_charges(n, 101)
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=300, precision="double")
    o = oracle_mod.OracleTree(cell, pos, 6, 300, 4)
    assert rel(gpu_stages(s, q1, 7), o.evaluate(q1, 7)) < 1e-12
    eps4 = 0.25e-6  # the reference's far-field NUFFT tolerance at 6 digits
    assert rel(gpu_stages(s, q1, 8), o.evaluate(q1, 8)) < eps4
    g = s.evaluate(q1)
    assert rel(g, o.evaluate(q1)) < eps4
    s.close()
    o.close()
    f = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=300, precision="float")
    a, b = f.evaluate(q1).astype(np.float64), f.evaluate(q2).astype(np.float64)
    ab = f.evaluate(0.5 * q1 - 2.0 * q2).astype(np.float64)
    assert rel(ab, 0.5 * a - 2.0 * b) < 1e-5
    assert rel(a, g) < 1e-5


@pytest.mark.parametrize("name,digits", [("cubic", 3), ("cubic", 6), ("cubic", 9), ("cubic", 12), ("triclinic", 6),
                                         ("triclinic", 9), ("clustered", 6), ("long_cell", 6)])
def test_nufft_far_field(pdmk, oracle_mod, cuda, monkeypatch, name, digits):
    """Far field by the NUFFT (forced) against the oracle's exact truncated Fourier sum, to
    the reference's NUFFT tolerance eps/4 (cpp/src/dmk/pbc.cpp:440)."""
    cell, pos, q, npl = case(name)
    monkeypatch.setenv("PDMK_FAR_FIELD", "nufft")
    s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, nz=4, precision="double")
    monkeypatch.delenv("PDMK_FAR_FIELD")
    o = oracle_mod.OracleTree(cell, pos, digits, npl, 4, precision="double")
    far_g, far_o = gpu_stages(s, q, 8), o.evaluate(q, 8)
    assert rel(far_g, far_o) < 0.25 * 10.0 ** -digits, rel(far_g, far_o)
    # the NUFFT path is deterministic
    np.testing.assert_array_equal(gpu_stages(s, q, 8), far_g)
    s.close()
