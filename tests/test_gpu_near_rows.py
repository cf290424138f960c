"""The near field over per-row step lists (k_near_rows in FP32, k_near_rows64 in FP64, at 6, 9
and 12 digits) against the wavefront-wide lists (k_near_field, PDMK_NEAR_ROWS=0).

For each target the row kernel visits the same in-support sources in the same order, forms
each list entry's sum the same way and adds the entries to the FP64 total in the same order,
so the near-field potentials, and with them the whole evaluate(), are bit-identical.  Checked
on uniform, triclinic, clustered, single-level, rectangular and partly filled inputs, and at
N = 1e6 (the reference metric's workload).  The in-support pair count (the roofline's work)
is the same for both lists.
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu


def solvers(pdmk, monkeypatch, cell, pos, npl, precision="float", digits=6):
    monkeypatch.setenv("PDMK_NEAR_ROWS", "0")
    old = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, nz=4, precision=precision)
    monkeypatch.delenv("PDMK_NEAR_ROWS")
    new = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, nz=4, precision=precision)
    assert not old.build_info()["near_rows"] and new.build_info()["near_rows"]
    return old, new


CASES = {
    "cubic": lambda: (W.cubic_cell(), W.uniform_positions(3000, 1, W.cubic_cell()), 40),
    "triclinic": lambda: (W.triclinic_cell(), W.uniform_positions(4000, 3, W.triclinic_cell()), 40),
    "clustered": lambda: (W.triclinic_cell(), W.gaussian_mixture(6000, 8, 0.02, 21, W.triclinic_cell()), 30),
    "single_level": lambda: (W.cubic_cell(), W.uniform_positions(10_000, 1, W.cubic_cell()), 300),
    "rectangular": lambda: (np.diag([1.0, 1.5, 2.0]), W.uniform_positions(4000, 12, np.diag([1.0, 1.5, 2.0])), 40),
    "sparse_leaves": lambda: (W.cubic_cell(), W.uniform_positions(700, 4, W.cubic_cell()), 7),
    "n1e5": lambda: (W.cubic_cell(), W.uniform_positions(100_000, 6, W.cubic_cell()), 300),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision,digits", [("float", 6), ("double", 6), ("double", 9), ("double", 12)])
def test_rows_equal_wavefront_lists(pdmk, cuda, monkeypatch, name, precision, digits):
    if digits == 12 and name in ("n1e5", "single_level"):
        pytest.skip("covered by the smaller inputs at 12 digits")
    cell, pos, npl = CASES[name]()
    q = W.neutral_uniform_charges(pos.shape[0], 77)
    old, new = solvers(pdmk, monkeypatch, cell, pos, npl, precision, digits)
    np.testing.assert_array_equal(new.run_stages(q, 2), old.run_stages(q, 2))
    np.testing.assert_array_equal(new.evaluate(q), old.evaluate(q))
    assert new.work()["nf_pairs_support"] == old.work()["nf_pairs_support"]
    old.close()
    new.close()


@pytest.mark.parametrize("precision,digits", [("float", 6), ("double", 9)])
def test_rows_full_size(pdmk, cuda, monkeypatch, precision, digits):
    n = 1_000_000
    cell = W.cubic_cell()
    pos = W.uniform_positions(n, 11, cell)
    q = W.neutral_uniform_charges(n, 100)
    old, new = solvers(pdmk, monkeypatch, cell, pos, 300, precision, digits)
    np.testing.assert_array_equal(new.run_stages(q, 2), old.run_stages(q, 2))
    np.testing.assert_array_equal(new.evaluate(q), old.evaluate(q))
    old.close()
    new.close()
