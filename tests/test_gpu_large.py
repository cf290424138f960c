"""The largest workload and the memory layout it needs.

The clustered triclinic workload: N = 4,000,000 points in 64 isotropic Gaussian clusters
(standard deviation 0.02, centres uniform in fractional coordinates, splitmix64 seed 21), cell
columns a1=(1,0,0), a2=(0.3,1,0), a3=(0.2,0.25,1), charges +/-1 in equal numbers (seed 22),
n_per_leaf = 300, nz = 4.  Its tree has 333,928 boxes and 124,287 proxy slots; at 12 digits
the outgoing plane waves of all slots alone would take 168 GB, so the plan holds them one
level at a time.  The potentials are compared with the FP64 Ewald sum at 1000 sampled points
(pdmk_gpu_ewald_targets, pinned against the oracle's Ewald sums in test_gpu_ewald.py), to
10^(2 - n_digits) in relative L2 norm; the reference's own end-to-end tests allow 1e-4 at 6
digits (cpp/tests/test_periodic_dmk.cpp:42, 67).

The per-level layout runs the same kernels on the same values, so on smaller inputs it must
give exactly the potentials of the all-slots layout (PDMK_PW_LEVELS forces either).
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def build(pdmk, monkeypatch, cell, pos, digits, npl, precision, by_level, bloch=None):
    monkeypatch.setenv("PDMK_PW_LEVELS", "1" if by_level else "0")
    try:
        s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, nz=4, precision=precision, bloch=bloch)
    finally:
        monkeypatch.delenv("PDMK_PW_LEVELS")
    assert s.build_info()["pw_by_level"] == by_level
    return s


SMALL = {
    "cubic_6_double": (W.cubic_cell(), W.uniform_positions(20_000, 3, W.cubic_cell()), 40, 6, "double"),
    "clustered_9_double": (W.triclinic_cell(), W.gaussian_mixture(30_000, 8, 0.02, 21, W.triclinic_cell()), 30, 9,
                           "double"),
    "clustered_6_float": (W.triclinic_cell(), W.gaussian_mixture(30_000, 8, 0.02, 21, W.triclinic_cell()), 30, 6,
                          "float"),
}


@pytest.mark.parametrize("name", list(SMALL))
def test_plane_waves_by_level_identical(pdmk, cuda, monkeypatch, name):
    cell, pos, npl, digits, precision = SMALL[name]
    q = W.neutral_uniform_charges(pos.shape[0], 8)
    out = []
    for by_level in (False, True):
        s = build(pdmk, monkeypatch, cell, pos, digits, npl, precision, by_level)
        out.append(s.evaluate(q))
        s.close()
    np.testing.assert_array_equal(out[1], out[0])


def test_plane_waves_by_level_identical_quasi_periodic(pdmk, cuda, monkeypatch):
    cell = W.triclinic_cell()
    pos = W.gaussian_mixture(20_000, 8, 0.03, 4, cell)
    q = W.neutral_uniform_charges(pos.shape[0], 9)
    out = []
    for by_level in (False, True):
        s = build(pdmk, monkeypatch, cell, pos, 6, 40, "double", by_level, bloch=(0.3, -0.2, 0.1))
        out.append(s.evaluate_complex(q))
        s.close()
    np.testing.assert_array_equal(out[1], out[0])


@pytest.fixture(scope="module")
def clustered_4m():
    n = 4_000_000
    cell = W.triclinic_cell()
    pos = W.gaussian_mixture(n, 64, 0.02, 21, cell)
    q = W.pm1_charges(n, 22)
    tg = np.sort(np.random.default_rng(5).choice(n, 1000, replace=False)).astype(np.int32)
    return cell, pos, q, tg


@pytest.fixture(scope="module")
def clustered_4m_ewald(pdmk, clustered_4m):
    cell, pos, q, tg = clustered_4m
    return pdmk.ewald_targets(cell, pos, q, tg, 5.5)


@pytest.mark.parametrize("digits,precision", [(6, "float"), (9, "double"), (12, "double")])
def test_clustered_triclinic_4m(pdmk, cuda, clustered_4m, clustered_4m_ewald, digits, precision):
    cell, pos, q, tg = clustered_4m
    s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=300, nz=4, precision=precision)
    try:
        st = s.stats()
        assert (st["nsall"], st["nboxes"], s.build_info()["proxy_slots"]) == (23_423_948, 333_928, 124_287)
        pot = s.evaluate(q).astype(np.float64)
        err = rel(pot[tg], clustered_4m_ewald)
        print(f"N=4e6 clustered triclinic, {digits} digits {precision}: rel-L2 vs Ewald {err:.2e}, "
              f"plane waves by level: {s.build_info()['pw_by_level']}", flush=True)
        assert err < 10.0 ** (2 - digits), err
        # the same charges again give the same potentials
        np.testing.assert_array_equal(s.evaluate(q).astype(np.float64), pot)
    finally:
        s.close()


@pytest.mark.parametrize("digits,n,npl", [(9, 100_000, 40), (12, 40_000, 40)])
def test_two_pass_transforms_identical(pdmk, cuda, monkeypatch, digits, n, npl):
    """FP64 at 9 and 12 digits (p = 28 and 38) forms the first contraction of the proxy <->
    plane-wave transforms for whole boxes in an HBM scratch and takes each box's slices from it,
    several slices per workgroup when the boxes are many.  The sums are the one-pass kernels':
    the potentials must equal those with the two passes off (PDMK_TWO_PASS=0), also with a 1 MiB
    scratch that splits every level into chunks of a few boxes, and the many-vector path (which
    runs the one-pass kernels) must give each vector's single-vector potentials."""
    cell = W.cubic_cell()
    pos = W.uniform_positions(n, 31, cell)
    Q = np.stack([W.neutral_uniform_charges(n, 40 + k) for k in range(2)])
    out = {}
    for mode in ("0", None, "1"):
        if mode is None:
            monkeypatch.delenv("PDMK_TWO_PASS", raising=False)
        else:
            monkeypatch.setenv("PDMK_TWO_PASS", mode)
        s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, nz=4, precision="double")
        monkeypatch.delenv("PDMK_TWO_PASS", raising=False)
        try:
            out[mode] = s.evaluate(Q[0])
            if mode is None:
                many = s.evaluate_many(Q)
                np.testing.assert_array_equal(many[0], out[mode])
                np.testing.assert_array_equal(many[1], s.evaluate(Q[1]))
                print(f"{digits} digits, N={n}: {s.build_info()['proxy_slots']} proxy slots", flush=True)
        finally:
            s.close()
    np.testing.assert_array_equal(out[None], out["0"])
    np.testing.assert_array_equal(out["1"], out["0"])


@pytest.mark.parametrize("digits,n,npl", [(9, 60_000, 40)])
def test_tensor3_matrix_cores_identical(pdmk, cuda, monkeypatch, digits, n, npl):
    """FP64 at 9 digits (p = 28) runs the child <-> parent tensor transforms on the matrix
    cores, each output summed in the order of the vector-ALU kernel: the potentials must equal
    those with the vector-ALU kernel (PDMK_T3_MFMA=0), on a clustered triclinic tree whose
    levels have parents with up to 8 children."""
    cell = W.triclinic_cell()
    pos = W.gaussian_mixture(n, 8, 0.03, 7, cell)
    q = W.neutral_uniform_charges(n, 44)
    out = {}
    for mode in ("0", None):
        if mode is None:
            monkeypatch.delenv("PDMK_T3_MFMA", raising=False)
        else:
            monkeypatch.setenv("PDMK_T3_MFMA", mode)
        s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=npl, nz=4, precision="double")
        monkeypatch.delenv("PDMK_T3_MFMA", raising=False)
        try:
            out[mode] = s.evaluate(q)
        finally:
            s.close()
    np.testing.assert_array_equal(out[None], out["0"])


def test_charge_to_proxy_shapes_identical(pdmk, cuda, monkeypatch):
    """FP64 at 9 digits forms the proxy charges on the matrix cores in one of two workgroup
    shapes (two 4-wave workgroups per box for trees with many boxes, one 16-wave workgroup
    otherwise); every output sums its sources in the same order either way, so forcing each
    (PDMK_C2P_WIDE=1 / 0) on one clustered tree must give the same potentials bit for bit."""
    cell = W.triclinic_cell()
    pos = W.gaussian_mixture(60_000, 8, 0.03, 8, cell)
    q = W.neutral_uniform_charges(pos.shape[0], 46)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("PDMK_C2P_WIDE", mode)
        s = pdmk.PDMKSolver(cell, pos, n_digits=9, n_per_leaf=40, nz=4, precision="double")
        monkeypatch.delenv("PDMK_C2P_WIDE")
        try:
            out[mode] = s.evaluate(q)
        finally:
            s.close()
    np.testing.assert_array_equal(out["1"], out["0"])


@pytest.mark.parametrize("digits,precision", [(6, "float"), (6, "double"), (9, "double"), (12, "double")])
def test_child_to_parent_per_parent_items(pdmk, cuda, monkeypatch, digits, precision):
    """Levels with many parents run the child -> parent transform as one item per parent (its
    children's terms summed in one pass, then added to the parent's own proxy charges); levels
    with few run one item per child and a separate sum.  Forcing either everywhere
    (PDMK_C2P_FUSE=1 / a huge threshold) must give the same potentials to rounding: the two
    differ only in the order of the final additions."""
    cell = W.triclinic_cell()
    pos = W.gaussian_mixture(40_000, 8, 0.03, 9, cell)
    q = W.neutral_uniform_charges(pos.shape[0], 45)
    out = {}
    for mode in ("1", "1000000000"):
        monkeypatch.setenv("PDMK_C2P_FUSE", mode)
        s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=40, nz=4, precision=precision)
        monkeypatch.delenv("PDMK_C2P_FUSE")
        try:
            out[mode] = s.evaluate(q).astype(np.float64)
        finally:
            s.close()
    tol = 1e-6 if precision == "float" else 1e-13
    assert rel(out["1"], out["1000000000"]) < tol, rel(out["1"], out["1000000000"])


@pytest.mark.parametrize("digits,precision", [(6, "float"), (9, "double"), (12, "double")])
def test_grouped_shifts(pdmk, cuda, monkeypatch, digits, precision):
    """The plane-wave shifts of groups of 8 consecutive target boxes (their neighbour lists
    merged by source slot, each neighbour read once per group) against one box per workgroup
    (PDMK_SHIFT_GROUPS=1 / 0): the same terms, summed in another order per box, on a clustered
    triclinic tree (irregular lists, partial groups) and a cubic tree whose level-0 image boxes
    fold (a slot that appears twice in one box's list)."""
    cases = [(W.triclinic_cell(), W.gaussian_mixture(40_000, 8, 0.03, 12, W.triclinic_cell())),
             (W.cubic_cell(), W.uniform_positions(30_000, 13, W.cubic_cell()))]
    tol = 1e-6 if precision == "float" else 1e-13
    for cell, pos in cases:
        q = W.neutral_uniform_charges(pos.shape[0], 46)
        out = {}
        for mode in ("1", "0"):
            monkeypatch.setenv("PDMK_SHIFT_GROUPS", mode)
            s = pdmk.PDMKSolver(cell, pos, n_digits=digits, n_per_leaf=40, nz=4, precision=precision)
            monkeypatch.delenv("PDMK_SHIFT_GROUPS")
            try:
                out[mode] = s.evaluate(q).astype(np.float64)
            finally:
                s.close()
        assert rel(out["1"], out["0"]) < tol, rel(out["1"], out["0"])
