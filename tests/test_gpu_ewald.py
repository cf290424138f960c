"""GPU FP64 Ewald sum (pdmk_ewald_nufft) against the CPU oracle's Ewald sums.

Known answers follow the reference tests: the NaCl Madelung constant 1.7475645946331822
(cpp/tests/test_ewald_nufft.cpp:53-85, 1e-8) and Ewald vs direct-Ewald agreement at 1e-10
(test_ewald_nufft.cpp:78-84, 102-106).  For a non-neutral system the potentials follow the
reference's splitting-parameter convention (choose_alpha, cpp/src/ewald_nufft.cpp:40-56),
which for small N is the balanced alpha the oracle's ewald_direct uses.
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu
M_REF = 1.7475645946331822


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_madelung(pdmk, cuda):
    pos = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5], [.5, .5, .5]])
    q = np.array([1, 1, 1, 1, -1, -1, -1, -1.0])
    pot, e = pdmk.ewald(np.eye(3), pos, q, 5.0)
    assert abs(-e / 8 - M_REF) < 1e-8
    assert abs(-0.5 * np.dot(q, pot) / 8 - M_REF) < 1e-8


@pytest.mark.parametrize("name", ["cubic", "triclinic", "non_neutral", "outside_cell"])
def test_vs_direct_ewald(pdmk, oracle_mod, cuda, name):
    rng = np.random.default_rng(42)
    if name == "triclinic":  # cpp/tests/test_periodic_dmk.cpp:47-68 cell
        cell = np.array([[5.0, 1.0, 0.5], [0.0, 4.5, 0.8], [0.0, 0.0, 4.0]])
        pos = rng.uniform(0.01, 0.99, (100, 3)) @ cell.T
    else:
        cell = np.eye(3) * 3.0
        pos = rng.uniform(0.03 * 3, 0.97 * 3, (200, 3))
    q = rng.normal(size=pos.shape[0])
    if name != "non_neutral":
        q[0] -= q.sum()
    if name == "outside_cell":  # shifted by lattice vectors: the periodic sum is unchanged
        pos_in = pos.copy()
        pos = pos + np.array([3.0, -6.0, 9.0])
    ref, e_ref = oracle_mod.ewald_direct(cell, pos if name != "outside_cell" else pos_in, q, 5.0)
    pot, e = pdmk.ewald(cell, pos, q, 5.0)
    assert rel(pot, ref) < 1e-10
    assert abs(e - e_ref) < 1e-10 * abs(e_ref)


def test_larger_system_vs_cell_ewald(pdmk, oracle_mod, cuda):
    n = 20000
    cell = W.triclinic_cell()
    pos = W.uniform_positions(n, 3, cell)
    q = W.neutral_uniform_charges(n, 4)
    tg = np.sort(np.random.default_rng(1).choice(n, 200, replace=False)).astype(np.int32)
    ref = oracle_mod.ewald_cells(cell, pos, q, s=6.0, targets=tg)
    pot, _ = pdmk.ewald(cell, pos, q, 6.0)
    assert rel(pot[tg], ref) < 1e-10


@pytest.mark.parametrize("name", ["triclinic", "non_neutral", "clustered"])
def test_targets_equal_full_sum(pdmk, cuda, name):
    """pdmk_gpu_ewald_targets (ewald_nufft_targets, cpp/src/ewald_nufft.cpp:663-759) returns
    pdmk_ewald_nufft's potentials at the chosen points, although it picks its own splitting
    parameter (the sum does not depend on it)."""
    n = 6000
    cell = W.triclinic_cell()
    pos = (W.gaussian_mixture(n, 8, 0.03, 7, cell) if name == "clustered" else W.uniform_positions(n, 5, cell))
    q = W.neutral_uniform_charges(n, 6) + (0.01 if name == "non_neutral" else 0.0)
    tg = np.random.default_rng(2).choice(n, 150, replace=False).astype(np.int32)
    tg[-1] = tg[0]  # a repeated target
    full, _ = pdmk.ewald(cell, pos, q, 5.5)
    sub = pdmk.ewald_targets(cell, pos, q, tg, 5.5)
    assert rel(sub, full[tg]) < 1e-11, rel(sub, full[tg])
    assert sub[-1] == sub[0]


def test_targets_vs_cell_ewald_clustered(pdmk, oracle_mod, cuda):
    """At N = 2e5 clustered points the subset sum agrees with the oracle's cell-list Ewald."""
    n = 200_000
    cell = W.triclinic_cell()
    pos = W.gaussian_mixture(n, 64, 0.02, 21, cell)
    q = W.pm1_charges(n, 22)
    tg = np.sort(np.random.default_rng(3).choice(n, 40, replace=False)).astype(np.int32)
    # smallest splitting parameter the oracle allows: few reciprocal modes over the 2e5 charges
    ref = oracle_mod.ewald_cells(cell, pos, q, s=5.5, targets=tg, alpha=1e-9)
    assert rel(pdmk.ewald_targets(cell, pos, q, tg, 5.5), ref) < 1e-10


def test_targets_madelung(pdmk, cuda):
    pos = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5], [.5, .5, .5]])
    q = np.array([1, 1, 1, 1, -1, -1, -1, -1.0])
    pot = pdmk.ewald_targets(np.eye(3), pos, q, [0, 7], 5.0)
    np.testing.assert_allclose(-pot * q[[0, 7]] / 2, M_REF, rtol=1e-8)


def test_targets_errors(pdmk, cuda):
    pos = np.random.default_rng(0).random((10, 3))
    q = np.ones(10)
    assert pdmk.ewald_targets(np.eye(3), pos, q, np.zeros(0, dtype=np.int32)).shape == (0,)
    for bad in ([10], [-1]):
        with pytest.raises(RuntimeError):
            pdmk.ewald_targets(np.eye(3), pos, q, bad)
    with pytest.raises(RuntimeError):
        pdmk.ewald_targets(np.eye(3), pos, q, [0], s=0.0)


def test_errors(pdmk, cuda):
    pos = np.zeros((4, 3))
    q = np.ones(4)
    with pytest.raises(RuntimeError):
        pdmk.ewald(np.zeros((3, 3)), pos, q, 5.0)
    with pytest.raises(RuntimeError):
        pdmk.ewald(np.eye(3), pos, q, -1.0)


def test_reference_data_file_from_gpu_ewald(pdmk, cuda, tmp_path):
    """Reference-data files carry oracle potentials between runs: the FP64 Ewald sum of the
    N=1e4 cubic problem at M=1000 sampled targets goes into a PDMC file, is read back and
    verified, and the 6-digit evaluate() agrees with it to 1e-5 (the reference's float
    end-to-end bound is 1e-4)."""
    from periodicdmk_amd import campaign as K
    cell = W.cubic_cell()
    pos = W.uniform_positions(10_000, 1, cell)
    q = W.neutral_uniform_charges(10_000, 2)
    idx = np.random.default_rng(7).choice(10_000, K.DEFAULT_M, replace=False).astype(np.uint32)
    u_ewald, _ = pdmk.ewald(cell, pos, q, 5.5)
    path = tmp_path / "cfg1_d6.pdmc"
    K.write_ref(path, K.RefData(cell=cell, positions=pos, charges=q, ref_idx=idx, ref_u=u_ewald[idx],
                                config_id=1, digits=6, prec=1, seed=1))
    r = K.read_ref(path)
    assert (r.config_id, r.N, r.M, r.digits) == (1, 10_000, K.DEFAULT_M, 6)
    s = pdmk.PDMKSolver(r.cell, r.positions, n_digits=r.digits, n_per_leaf=300, nz=4, precision="double")
    pot = s.evaluate(r.charges)
    s.close()
    assert rel(pot[r.ref_idx], r.ref_u) < 1e-5
