"""Structural invariants of the oracle's tree (CPU).

No fixture of the reference's PBCTree exists, so the restated tree is pinned by the
properties the reference code fixes (cpp/src/dmk/pbc_tree.cpp):
  * children sit at the parent's centre +- bs/2 of their own level, one level deeper, and
    counts are conserved from parent to children (bin sort, 52-85, 290-303);
  * no box holding more than ceil(1.5 n_per_leaf) sources is a leaf (hysteresis, 290-303);
  * a leaf holding sources touches no leaf two or more levels deeper (2:1 balance,
    438-598, applied to non-empty leaves only);
  * a box expands into plane waves iff it holds points and is at level 0, is not a leaf,
    or is a leaf with a non-leaf colleague (ifpwexp, 776-815);
  * list1 of a target leaf = every touching leaf holding sources, itself included
    (850-907 with the 1.05 tolerances);
checked against the geometry of the exported box centres, independently of the code
that built the lists.
"""
import math

import numpy as np
import pytest

from periodicdmk_amd import workloads as W

CASES = {
    "cubic": lambda: (W.cubic_cell(), W.uniform_positions(6_000, 1, W.cubic_cell()), 30),
    "triclinic": lambda: (W.triclinic_cell(), W.uniform_positions(6_000, 5, W.triclinic_cell()), 30),
    "clustered": lambda: (W.triclinic_cell(), W.gaussian_mixture(8_000, 8, 0.02, 21, W.triclinic_cell()), 20),
    "tight": lambda: (W.cubic_cell(), W.gaussian_mixture(4_000, 3, 0.005, 7, W.cubic_cell()), 20),
}


def _tree(oracle_mod, name):
    cell, pos, npl = CASES[name]()
    o = oracle_mod.OracleTree(cell, pos, 6, npl, 4)
    inf = o.info()
    box7, sp, tp, cen = o.tree_arrays()
    o.close()
    return inf, box7, sp, tp, cen, npl, len(pos)


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_tree_invariants(oracle_mod, name):
    inf, box7, sp, tp, cen, npl, n = _tree(oracle_mod, name)
    level, parent, nchild, nsrc, ntrg, pwexp, nl1 = (box7[:, k] for k in range(7))
    rc = inf["rc"]
    size = rc / 2.0 ** level
    nb = len(level)
    leaf = nchild == 0
    # permutations
    assert sorted(tp.tolist()) == list(range(n))
    assert len(np.unique(sp)) == len(sp)
    # counts and child geometry
    assert nsrc[level == 0].sum() == inf["nsall"] and ntrg[level == 0].sum() == n
    kids = {b: [] for b in range(nb)}
    for b in range(nb):
        p = parent[b]
        if p < 0:
            assert level[b] == 0
            continue
        kids[p].append(b)
        assert level[b] == level[p] + 1
        np.testing.assert_allclose(np.abs(cen[b] - cen[p]), size[b] / 2, rtol=0, atol=1e-12 * rc)
    for b in range(nb):
        if not leaf[b]:
            assert len(kids[b]) == 8
            assert nsrc[b] == sum(nsrc[k] for k in kids[b])
            assert ntrg[b] == sum(ntrg[k] for k in kids[b])
    # refinement threshold
    thr = math.ceil(1.5 * npl)
    assert not np.any(leaf & (nsrc > thr))
    # 2:1 balance (applied to leaves holding sources) and list1 = touching leaves with sources
    lv = np.nonzero(leaf)[0]
    d = np.abs(cen[lv][:, None, :] - cen[lv][None, :, :])
    touch = np.all(d <= 0.5 * (size[lv][:, None, None] + size[lv][None, :, None]) * (1 + 1e-9), axis=2)
    deeper = level[lv][None, :] - level[lv][:, None]
    assert not np.any(touch & (deeper >= 2) & (nsrc[lv] > 0)[:, None])
    has_t = ntrg[lv] > 0
    expect = (touch & (nsrc[lv] > 0)[None, :]).sum(axis=1)
    np.testing.assert_array_equal(nl1[lv][has_t], expect[has_t])
    # plane-wave flags: level 0, non-leaf, or a leaf with a non-leaf colleague (same level,
    # |centre offset| <= 1.05 bs per axis); cleared for boxes holding no points
    expect_pw = np.zeros(nb, dtype=bool)
    for lev in range(level.max() + 1):
        ids = np.nonzero(level == lev)[0]
        if lev == 0:
            expect_pw[ids] = True
            continue
        dd = np.abs(cen[ids][:, None, :] - cen[ids][None, :, :])
        coll = np.all(dd <= 1.05 * size[ids][0], axis=2)
        expect_pw[ids] = (~leaf[ids]) | (coll & (~leaf[ids])[None, :]).any(axis=1)
    expect_pw &= (nsrc + ntrg) > 0
    np.testing.assert_array_equal(pwexp.astype(bool), expect_pw)
    # a refined tree exercises the checks above
    assert level.max() >= 1
