"""Tree construction with the per-point work on the GPU (images, level-0 binning, stable
distribution into children, radix sort by box path) against the host construction: every
box (level, parent, leaf flag, source and target counts, plane-wave flag, list1 size),
every box centre and both permutations must be identical.  The host construction is itself
pinned to the oracle's restatement of the reference PBCTree (tests/test_host_logic.py).
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu


def cases():
    tri = W.triclinic_cell()
    cub = W.cubic_cell()
    long_cell = np.array([[1.0, 0.0, 0.1], [0.0, 1.7, 0.0], [0.0, 0.2, 1.2]])
    deep = np.vstack([0.3 + 3e-7 * np.random.default_rng(3).random((600, 3)), W.uniform_positions(400, 8, cub)])
    return {
        "cubic": (cub, W.uniform_positions(3000, 1, cub), 40),
        "triclinic": (tri, W.uniform_positions(4000, 3, tri), 40),
        "clustered": (tri, W.gaussian_mixture(6000, 8, 0.02, 21, tri), 30),
        "long_cell": (long_cell, W.uniform_positions(3000, 9, long_cell), 40),
        "single_level": (cub, W.uniform_positions(10_000, 1, cub), 300),
        "one_point": (cub, np.array([[0.5, 0.5, 0.5]]), 300),
        "uniform_2e5": (cub, W.uniform_positions(200_000, 4, cub), 300),
        # a 3e-7 cluster refines to 23 levels, past the 18 the GPU path keys hold for a 6^3
        # level-0 grid: built with the points on the host
        "deep": (cub, deep, 20),
    }


@pytest.mark.parametrize("name", list(cases()))
def test_gpu_tree_equals_host_tree(pdmk, cuda, name):
    cell, pos, npl = cases()[name]
    h = pdmk.host_tree(cell, pos, npl, 4)
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, precision="double")
    box7, sp, tp, cen = s.tree_arrays()
    assert s.build_info()["tree_on_gpu"] == (name != "deep")
    s.close()
    np.testing.assert_array_equal(box7, h["box7"])
    np.testing.assert_array_equal(cen, h["centers"])
    np.testing.assert_array_equal(sp, h["src_perm"])
    np.testing.assert_array_equal(tp, h["trg_perm"])
