"""CPU tests of the library's host logic and C ABI (no GPU needed).

* The host tree of libpdmk.so equals the oracle's literal restatement of PBCTree
  (cpp/src/dmk/pbc_tree.cpp:138-959): box levels, parents, counts, plane-wave flags,
  list1 sizes, centres and both particle permutations, on uniform, triclinic and
  clustered inputs (2:1 balancing exercised).
* Library tables: PSWF vs the oracle, W0 and plane-wave matrices against closed forms.
* Every function declared in include/pdmk/*.h is exported; argument errors return the
  reference's error codes without touching a GPU.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from periodicdmk_amd import workloads as W

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


CASES = {
    "cubic": lambda: (W.cubic_cell(), W.uniform_positions(20_000, 1, W.cubic_cell()), 40),
    "triclinic": lambda: (W.triclinic_cell(), W.uniform_positions(20_000, 5, W.triclinic_cell()), 40),
    "clustered": lambda: (W.triclinic_cell(), W.gaussian_mixture(20_000, 8, 0.02, 21, W.triclinic_cell()), 40),
    "tight": lambda: (W.cubic_cell(), W.gaussian_mixture(8_000, 3, 0.005, 7, W.cubic_cell()), 20),
    "tiny": lambda: (np.eye(3) * 3.0, np.array([[1.5, 1.5, 1.5]]), 20),
}


@pytest.mark.parametrize("case", list(CASES))
def test_host_tree_equals_oracle(pdmk, oracle_mod, case):
    cell, pos, npl = CASES[case]()
    h = pdmk.host_tree(cell, pos, npl, 4)
    o = oracle_mod.OracleTree(cell, pos, 6, npl, 4)
    box7, sp, tp, cen = o.tree_arrays()
    np.testing.assert_array_equal(h["box7"], box7)
    np.testing.assert_array_equal(h["src_perm"], sp)
    np.testing.assert_array_equal(h["trg_perm"], tp)
    np.testing.assert_allclose(h["centers"], cen, rtol=0, atol=1e-14)
    assert abs(h["rc"] - o.info()["rc"]) < 1e-15


@pytest.mark.parametrize("beta", [7.2462000846862793, 13.739999771118164, 20.736000061035156, 27.870000839233398])
def test_library_pswf_matches_oracle(pdmk, oracle_mod, beta):
    for x in np.linspace(0, 1, 17):
        assert abs(pdmk.pswf(beta, x) - oracle_mod.pswf(beta, x)) < 1e-12


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_library_tables(pdmk, oracle_mod, digits):
    rc, nl = 0.25, 3
    umat = np.eye(3).reshape(9)
    t = pdmk.host_tables(digits, rc, nl, umat)
    beta = t["beta"]
    # W0 = psi0 / (c0 * bsize), bsize = rc/2 at level 0 and boxsize[l] below
    c0 = oracle_mod.pswf_integral(beta, 1.0)
    psi0 = oracle_mod.pswf(beta, 0.0)
    expect = [psi0 / (c0 * rc / 2), psi0 / (c0 * rc / 2), psi0 / (c0 * rc / 4), psi0 / (c0 * rc / 8)]
    np.testing.assert_allclose(t["w0"], expect, rtol=1e-12)
    # pw2poly = conj(poly2pw) interpolates exp(i (j - n/2) x beta/nf) in Chebyshev series on [-1,1];
    # checked for |frequency| <= beta/3, where the degree p-1 interpolant is accurate
    n_pw, p = t["n_pw"], t["p"]
    nf = (n_pw - 1) // 2
    x = np.linspace(-1, 1, 33)
    T = np.array([np.cos(k * np.arccos(x)) for k in range(p)])
    half = n_pw // 2
    for j in (half - nf // 3, half, half + nf // 3):
        approx = np.conj(t["poly2pw"][j]) @ T
        exact = np.exp(1j * (j - n_pw // 2) * x * beta / nf)
        assert np.max(np.abs(approx - exact)) < 10.0 ** (-digits)
    # far-field weights: w = 4 pi/(psi0 V) psi0(|k| rc/beta)/|k|^2, doubled for folded conjugates
    far = t["far"]
    k = 2 * np.pi * far[:, :3]
    kk = np.sum(k * k, axis=1)
    xv = np.sqrt(kk) * rc / beta
    w = np.array([4 * np.pi / psi0 * oracle_mod.pswf(beta, v) for v in xv]) / kk
    # near |k| rc/beta = 1 psi0 is ~1e-8, so compare against the largest weight
    np.testing.assert_allclose(far[:, 3], 2 * w, rtol=1e-11, atol=1e-13 * np.max(np.abs(w)))


def _declared_functions():
    names = []
    for h in glob.glob(os.path.join(REPO, "include", "pdmk", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"//[^\n]*", "", txt)
        names += re.findall(r"\b(pdmk_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_all_declared_symbols_exported(pdmk):
    L = C.CDLL(pdmk.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_reference_abi_symbols_present(pdmk):
    L = C.CDLL(pdmk.LIB_PATH)
    for n in ("pdmk_ewald_nufft", "pdmk_optimize_cubic_cover", "pdmk_create_double", "pdmk_evaluate_double",
              "pdmk_destroy_double", "pdmk_create_float", "pdmk_evaluate_float", "pdmk_destroy_float"):
        assert hasattr(L, n), n


def test_argument_errors_without_gpu(pdmk):
    L = pdmk.lib()
    cell = np.ascontiguousarray(np.eye(3))
    pos = np.ascontiguousarray(np.array([[0.5, 0.5, 0.5]]))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    # invalid digits / nz / N: create returns NULL (cpp/src/pdmk_capi.cpp:48-59)
    assert not L.pdmk_create_double(dp(cell), 1, dp(pos), 13, 20, 4)
    assert not L.pdmk_create_double(dp(cell), 1, dp(pos), 6, 20, 0)
    assert not L.pdmk_create_float(dp(cell), 0, dp(pos), 6, 20, 4)
    # NULL handle: evaluate returns -1 (pdmk_capi.cpp:65)
    q = np.zeros(1)
    out = np.zeros(1)
    assert L.pdmk_evaluate_double(None, 1, dp(pos), dp(q), dp(out)) == -1
    # nz <= 0: optimize_cubic_cover returns 1 (pdmk_capi.cpp:114)
    m, n, p = C.c_int(), C.c_int(), C.c_int()
    vec = np.zeros(12)
    assert L.pdmk_optimize_cubic_cover(dp(cell), 0, C.byref(m), C.byref(n), C.byref(p), dp(vec), dp(vec),
                                       dp(vec), dp(vec), dp(vec)) == 1
    L.pdmk_destroy_double(None)  # no-op
    # batched graphs: an empty call list or a NULL graph is rejected before any GPU work
    assert not L.pdmk_gpu_graph_create(0, None, None, None, 1)
    assert L.pdmk_gpu_graph_launch(None, None) == -1
    assert L.pdmk_gpu_graph_destroy(None) == -1
    assert L.pdmk_gpu_set_single_stream(None, 1) == -1
    # Ewald at targets: N <= 0 and s <= 0 are rejected before any GPU work
    idx = np.zeros(1, dtype=np.int32)
    ip = idx.ctypes.data_as(C.POINTER(C.c_int))
    assert L.pdmk_gpu_ewald_targets(dp(cell), 0, dp(pos), dp(q), 1, ip, 5.0, dp(out)) == -1
    assert L.pdmk_gpu_ewald_targets(dp(cell), 1, dp(pos), dp(q), 1, ip, 0.0, dp(out)) == -1
    assert L.pdmk_gpu_ewald_targets(dp(cell), 1, dp(pos), dp(q), 0, None, 5.0, None) == 0


def test_workload_generator_is_deterministic():
    a = W.uniform01(11, 5)
    b = W.uniform01(11, 5)
    np.testing.assert_array_equal(a, b)
    assert np.all((a >= 0) & (a < 1))
    # splitmix64 known answer for seed 0 (first output of the published generator)
    assert int(W.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    q = W.neutral_uniform_charges(1000, 2)
    assert abs(q.sum()) < 1e-12
    qpm = W.pm1_charges(1000, 3)
    assert qpm.sum() == 0 and set(np.unique(qpm)) == {-1.0, 1.0}
    cells = W.batch_cells(256)
    ns = np.array([n for _, n in cells])
    assert ns.min() >= 2000 and ns.max() <= 200_000
    assert 60 < sum(np.allclose(c, np.eye(3)) for c, _ in cells) < 196  # about half cubic
    for c, _ in cells:
        assert np.allclose(np.tril(c, -1), 0) and np.all(np.diag(c) == 1) and c.max() <= 1.0
    assert W.batch_cells(256)[7][1] == cells[7][1]


def _frame_ok(r):
    E = np.array([r["e1"], r["e2"], r["e3"]])
    np.testing.assert_allclose(E @ E.T, np.eye(3), atol=1e-12)
    assert np.linalg.det(E) > 0  # right-handed


@pytest.mark.parametrize("cell,nz,expect", [
    (np.eye(3), 4, (4, 4, 4, 0.25)),                      # cpp/tests/test_cubic_cover_optimizer.cpp:8-30
    (np.diag([4.0, 2.0, 1.0]), 4, (16, 8, 4, 0.25)),      # :32-45, julia runtests.jl:44-50
    (np.diag([16.0, 2.0, 2.0]), 4, (32, 4, 4, 0.5)),      # :47-59, runtests.jl:52-58
    (np.eye(3) * 2.0, 1, (1, 1, 1, 2.0)),                  # :61-72
])
def test_optimize_cubic_cover_known_answers(pdmk, cell, nz, expect):
    r = pdmk.optimize_cubic_cover(cell, nz)
    assert (r["m"], r["n"], r["p"]) == expect[:3]
    assert abs(r["s"] - expect[3]) < 1e-12
    _frame_ok(r)
    # cost = (m+2)(n+2)(p+2) s^3 / V
    V = abs(np.linalg.det(cell))
    assert abs(r["cost"] - (r["m"] + 2) * (r["n"] + 2) * (r["p"] + 2) * r["s"] ** 3 / V) < 1e-12 * r["cost"]


def test_optimize_cubic_cover_triclinic_frame(pdmk):
    # rows are lattice vectors in the reference test (cpp/tests/test_cubic_cover_optimizer.cpp:84-104)
    cell = np.array([[5.0, 0.0, 0.0], [1.0, 4.5, 0.0], [0.5, 0.8, 4.0]]).T
    r = pdmk.optimize_cubic_cover(cell, 4)
    _frame_ok(r)
    assert r["p"] == 4 and r["m"] >= 1 and r["n"] >= 1 and r["s"] > 0 and np.isfinite(r["cost"])
    # e1 along the longest lattice vector, e2 in the plane of the two longest
    a = cell.T
    order = np.argsort(-np.linalg.norm(a, axis=1), kind="stable")
    np.testing.assert_allclose(r["e1"], a[order[0]] / np.linalg.norm(a[order[0]]), atol=1e-14)
    assert abs(np.dot(np.cross(a[order[0]], a[order[1]]), r["e2"])) < 1e-12
    # p cubes of side s span the e3 width of the cell's bounding box; m, n the smallest covers
    width = [np.sum(np.abs(a @ r[k])) for k in ("e1", "e2", "e3")]
    assert abs(r["s"] * 4 - width[2]) < 1e-12
    assert (r["m"] - 1) * r["s"] < width[0] <= r["m"] * r["s"] + 1e-12
    assert (r["n"] - 1) * r["s"] < width[1] <= r["n"] * r["s"] + 1e-12


def test_optimize_cubic_cover_errors(pdmk):
    cell = np.array([[1.0, 0.0, 0.0], [0.3, 1.0, 0.0], [0.1, 0.2, 0.8]]).T
    for nz in (0, -1):  # the C ABI returns 1 (pdmk_capi.cpp:114)
        with pytest.raises(ValueError, match="code 1"):
            pdmk.optimize_cubic_cover(cell, nz)
    with pytest.raises(ValueError, match="code -1"):
        pdmk.optimize_cubic_cover(np.zeros((3, 3)), 4)


def test_host_tree_refuses_endless_refinement(pdmk):
    # more than ceil(1.5 * n_per_leaf) coincident points would split without end; the
    # builder stops before its integer box coordinates overflow and reports the input
    pos = np.full((40, 3), 0.3)
    with pytest.raises(RuntimeError, match="pdmk_host_tree failed"):
        pdmk.host_tree(np.eye(3), pos, 10, 4)
    # at the limit (15 coincident points, n_per_leaf 10) nothing splits endlessly
    h = pdmk.host_tree(np.eye(3), pos[:15], 10, 4)
    assert h["sizes"][2] > 0
