"""Quasi-periodic (Bloch) evaluate on the GPU against the CPU oracle's restatement of
PeriodicDMKTreeT::evaluate_complex (cpp/src/periodic_dmk.cpp:251-277), stage by stage
(bit 0 proxy / plane waves, bit 1 near field, bit 2 self correction, bit 3 far field):
FP64 relative L2 difference < 1e-12, FP32 < 1e-5.  End to end against the Bloch Ewald sum
of tests/test_oracle_qp.py, and the reference's own checks
(cpp/tests/test_periodic_dmk_qp.cpp): bloch = 0 gives the periodic result with a zero
imaginary part, and the result does not depend on nz.
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W
from test_oracle_qp import ORDERED, qp_ewald

pytestmark = pytest.mark.gpu

B = np.array([2.0, -1.0, 0.5])


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def case(name):
    if name == "cubic":
        c = W.cubic_cell()
        return c, W.uniform_positions(3000, 1, c), W.neutral_uniform_charges(3000, 2), 40
    if name == "triclinic":
        c = W.triclinic_cell()
        return c, W.uniform_positions(2500, 3, c), W.neutral_uniform_charges(2500, 4), 40
    if name == "clustered":
        c = W.triclinic_cell()
        return c, W.gaussian_mixture(3000, 8, 0.02, 21, c), W.pm1_charges(3000, 22), 30
    raise KeyError(name)


@pytest.mark.parametrize("name,precision,digits", [
    ("cubic", "double", 6), ("cubic", "double", 9), ("cubic", "float", 6),
    ("triclinic", "double", 6), ("clustered", "double", 9), ("clustered", "float", 6),
])
def test_qp_stage_parity(pdmk, oracle_mod, cuda, name, precision, digits):
    cell, pos, q, npl = case(name)
    s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, nz=4, precision=precision, bloch=B)
    assert s.build_info()["folded_boxes"] == 0  # quasi-periodic trees never fold
    o = oracle_mod.OracleTree(cell, pos, digits, npl, 4, precision="double")
    tol = 1e-12 if precision == "double" else 1e-5
    for st in (1, 2, 4, 8, 15):
        g = s.run_complex_stages(q, st)
        r = o.evaluate_complex(q, B, st)
        assert rel(g, r) < tol, (st, rel(g, r))
    full = s.evaluate_complex(q)  # host-buffer path
    assert rel(full, o.evaluate_complex(q, B)) < tol
    s.close()
    o.close()


@pytest.mark.parametrize("cellname,precision,tol", [("cubic", "double", 2e-6), ("ordered", "double", 2e-6),
                                                    ("cubic", "float", 1e-5)])
def test_qp_matches_bloch_ewald(pdmk, cuda, cellname, precision, tol):
    cell = W.cubic_cell() if cellname == "cubic" else ORDERED
    n = 40
    pos = W.uniform_positions(n, 5, cell)
    q = W.neutral_uniform_charges(n, 6) + 0.1
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=10, nz=4, precision=precision, bloch=B)
    got = s.evaluate_complex(q)
    s.close()
    assert rel(got, qp_ewald(cell, pos, q, B)) < tol


def test_qp_bloch_zero_is_periodic(pdmk, cuda):
    cell = W.cubic_cell()
    pos = W.uniform_positions(500, 3, cell)
    q = W.neutral_uniform_charges(500, 4)
    a = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=40, precision="double", bloch=np.zeros(3))
    b = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=40, precision="double")
    c, p = a.evaluate_complex(q), b.evaluate(q)
    a.close(); b.close()
    assert np.abs(c.imag).max() < 1e-10 * np.abs(p).max()
    assert rel(c.real, p) < 1e-12


def test_qp_independent_of_nz(pdmk, cuda):
    cell = W.cubic_cell()
    pos = W.uniform_positions(40, 5, cell)
    q = W.neutral_uniform_charges(40, 6)
    r = []
    for nz in (4, 8):
        s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=300, nz=nz, precision="double", bloch=B)
        r.append(s.evaluate_complex(q))
        s.close()
    assert np.abs(r[0] - r[1]).max() < 1e-4


def test_qp_error_behaviour(pdmk, cuda):
    cell = W.cubic_cell()
    pos = W.uniform_positions(100, 5, cell)
    q = W.neutral_uniform_charges(100, 6)
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=40, precision="double", bloch=B)
    with pytest.raises(RuntimeError):  # evaluate() on a quasi-periodic tree
        s.evaluate(q)
    with pytest.raises(ValueError):
        s.evaluate_complex(q[:-1])
    s.close()
    t = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=40, precision="double")
    with pytest.raises(RuntimeError):  # evaluate_complex() without bloch
        t.evaluate_complex(q)
    t.close()


def test_qp_deterministic_and_large(pdmk, cuda):
    """N = 2e5 (cubic, 9 digits): repeatable to the bit, and linear in the charges."""
    cell = W.cubic_cell()
    pos = W.uniform_positions(200_000, 11, cell)
    q1, q2 = W.neutral_uniform_charges(200_000, 100), W.neutral_uniform_charges(200_000, 101)
    s = pdmk.PDMKSolver(cell, pos, n_digits=9, n_per_leaf=300, precision="double", bloch=B)
    a, b = s.evaluate_complex(q1), s.evaluate_complex(q2)
    np.testing.assert_array_equal(s.evaluate_complex(q1), a)
    ab = s.evaluate_complex(0.5 * q1 - 2.0 * q2)
    s.close()
    assert rel(ab, 0.5 * a - 2.0 * b) < 1e-12


@pytest.mark.parametrize("precision,tol", [("double", 1e-11), ("float", 1e-5)])
def test_qp_reciprocal_lattice_bloch_is_periodic(pdmk, cuda, precision, tol):
    """b = 2 pi e_1 on the unit cube gives the periodic potentials (tests/test_oracle_qp.py)."""
    cell = W.cubic_cell()
    pos = W.uniform_positions(3000, 9, cell)
    q = W.neutral_uniform_charges(3000, 10)
    a = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=40, precision=precision,
                        bloch=np.array([2 * np.pi, 0.0, 0.0]))
    b = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=40, precision=precision)
    c, p = a.evaluate_complex(q), b.evaluate(q).astype(np.float64)
    a.close(); b.close()
    assert np.abs(c.imag).max() < tol * np.abs(p).max()
    assert rel(c.real, p) < tol
