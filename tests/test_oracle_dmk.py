"""CPU pins for the FP64 Ewald sums and the DMK oracle (end to end).

Known answers and tolerances follow the reference tests:
  * NaCl Madelung constant 1.7475645946331822 (cpp/tests/test_ewald_nufft.cpp:53-85, 1e-8;
    through DMK itself at rtol 0.01 in julia/PeriodicDMK/test/runtests.jl:8-30);
  * DMK vs Ewald, cubic L=3 N=200 and the triclinic cell of
    cpp/tests/test_periodic_dmk.cpp:20-68 (rel-L2 < 1e-4);
  * 9 digits < 1e-7 and float < 1e-4 (cpp/tests/test_nonuniform.cpp:282-304, 414-444);
  * repeat-evaluate stability and charge dependence (cpp/tests/test_reusable_tree.cpp:9-37).
"""
import os

import numpy as np
import pytest

from periodicdmk_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M_REF = 1.7475645946331822


def nacl():
    pos = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5], [.5, .5, .5]])
    q = np.array([1, 1, 1, 1, -1, -1, -1, -1.0])
    return np.eye(3), pos, q


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_madelung_ewald(oracle_mod):
    cell, pos, q = nacl()
    _, e = oracle_mod.ewald_direct(cell, pos, q, 5.0)
    assert abs(-e / 8 - M_REF) < 1e-8
    pc = oracle_mod.ewald_cells(cell, pos, q, 6.0)
    assert abs(-0.5 * np.dot(q, pc) / 8 - M_REF) < 1e-8


def test_madelung_dmk(oracle_mod):
    cell, pos, q = nacl()
    t = oracle_mod.OracleTree(cell, pos, 6, 4, 4)
    m = -0.5 * np.dot(q, t.evaluate(q)) / 8
    assert abs(m - M_REF) < 1e-4 * M_REF


def test_ewald_direct_vs_cells_triclinic(oracle_mod):
    cell = np.array([[5.0, 1.0, 0.5], [0.0, 4.5, 0.8], [0.0, 0.0, 4.0]])
    rng = np.random.default_rng(3)
    frac = rng.uniform(0.01, 0.99, (150, 3))
    pos = frac @ cell.T
    q = rng.normal(size=150)
    q -= q.mean()
    pd, _ = oracle_mod.ewald_direct(cell, pos, q, 5.5)
    pc = oracle_mod.ewald_cells(cell, pos, q, 6.0)
    assert rel(pc, pd) < 1e-10


def test_dmk_cubic_vs_ewald(oracle_mod):
    L = 3.0
    rng = np.random.default_rng(42)
    pos = rng.uniform(0.03 * L, 0.97 * L, (200, 3))
    q = rng.normal(size=200)
    q[0] -= q.sum()
    cell = np.eye(3) * L
    ref, _ = oracle_mod.ewald_direct(cell, pos, q, 5.0)
    for prec, tol in (("double", 1e-5), ("float", 1e-4)):
        t = oracle_mod.OracleTree(cell, pos, 6, 20, 4, precision=prec)
        assert rel(t.evaluate(q), ref) < tol


def test_dmk_triclinic_vs_ewald(oracle_mod):
    cell = np.array([[5.0, 1.0, 0.5], [0.0, 4.5, 0.8], [0.0, 0.0, 4.0]])
    rng = np.random.default_rng(77)
    pos = rng.uniform(0.01, 0.99, (100, 3)) @ cell.T
    q = rng.normal(size=100)
    q[0] -= q.sum()
    ref, _ = oracle_mod.ewald_direct(cell, pos, q, 5.0)
    t = oracle_mod.OracleTree(cell, pos, 6, 20, 4)
    assert rel(t.evaluate(q), ref) < 1e-5


def test_dmk_9_digits_clustered_vs_ewald(oracle_mod):
    cell = W.triclinic_cell()
    pos = W.gaussian_mixture(1500, 3, 0.03, 5, cell)
    q = W.pm1_charges(1500, 6)
    ref = oracle_mod.ewald_cells(cell, pos, q, 6.0)
    t = oracle_mod.OracleTree(cell, pos, 9, 20, 4)
    assert t.info()["n_levels"] >= 3
    assert rel(t.evaluate(q), ref) < 1e-7


def test_dmk_reuse_and_charge_dependence(oracle_mod):
    L = 3.0
    rng = np.random.default_rng(0)
    pos = rng.uniform(0.05 * L, 0.95 * L, (64, 3))
    q1, q2 = rng.normal(size=64), rng.normal(size=64)
    t = oracle_mod.OracleTree(np.eye(3) * L, pos, 6, 16, 4)
    p1, p2, p1b = t.evaluate(q1), t.evaluate(q2), t.evaluate(q1)
    np.testing.assert_allclose(p1b, p1, rtol=1e-12, atol=0)
    assert np.sum(np.abs(p1 - p2)) > 1e-10


def test_cfg1_against_golden_ewald(oracle_mod):
    """N = 1e4 uniform cubic, 6 digits, n_per_leaf 300 (the CPU-runnable workload)."""
    g = np.load(os.path.join(GOLDEN, "cfg1_ewald.npz"))
    cell = W.cubic_cell()
    pos = W.uniform_positions(10_000, 1, cell)
    q = W.neutral_uniform_charges(10_000, 2)
    chk = np.array([pos.sum(), pos[0, 0], pos[-1, 2], q.sum(), q[0], q[-1]])
    np.testing.assert_array_equal(chk, g["input_check"])
    t = oracle_mod.OracleTree(cell, pos, 6, 300, 4)
    err = rel(t.evaluate(q), g["ewald"])
    assert err < 2e-6, err
