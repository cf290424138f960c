"""evaluate() of many charge vectors per call (pdmk_gpu_evaluate_many) against one call per
vector: FP32 plans at 6 digits take the vectors in passes of 16 (the near field of a pass with
each residual computed once, the chain and the exact far field of several vectors per launch,
the far field's phases shared by 4 vectors per workgroup of structure factors and 2 of the
evaluation, passes overlapping); every vector's
potentials equal pdmk_gpu_evaluate_float's bit for bit (same sums in the same order).  FP64
plans call evaluate() per vector.  Vector counts below, at and above a pass of 16, groups of 4
cut short, three passes (the pass buffers reused), and 0; a tree large enough for the NUFFT far
field and for chain launches of fewer than 16 vectors, in two passes."""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = {
    "cubic": lambda: (W.cubic_cell(), W.uniform_positions(30_000, 21, W.cubic_cell()), 300),
    "triclinic": lambda: (W.triclinic_cell(), W.uniform_positions(20_000, 22, W.triclinic_cell()), 100),
    "clustered": lambda: (W.triclinic_cell(), W.gaussian_mixture(20_000, 8, 0.02, 23, W.triclinic_cell()), 60),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", ["float", "double"])
def test_many_equals_single(pdmk, cuda, name, precision):
    cell, pos, npl = CASES[name]()
    n = pos.shape[0]
    s = pdmk.PDMKSolver(cell, pos, n_digits=6, n_per_leaf=npl, nz=4, precision=precision)
    for nv in (1, 2, 11, 16, 37):
        Q = np.stack([W.neutral_uniform_charges(n, 500 + 7 * nv + v) for v in range(nv)])
        got = s.evaluate_many(Q)
        for v in range(nv):
            np.testing.assert_array_equal(got[v], s.evaluate(Q[v]), err_msg=f"{name} {precision} vector {v} of {nv}")
    assert s.evaluate_many(np.zeros((0, n))).shape == (0, n)
    s.close()


def test_many_large_tree(pdmk, cuda):
    n = 400_000
    cell = W.cubic_cell()
    s = pdmk.PDMKSolver(cell, W.uniform_positions(n, 24, cell), n_digits=6, n_per_leaf=300, nz=4, precision="float")
    Q = np.stack([W.neutral_uniform_charges(n, 900 + v) for v in range(18)])
    got = s.evaluate_many(Q)
    for v in range(18):
        np.testing.assert_array_equal(got[v], s.evaluate(Q[v]), err_msg=f"vector {v} of 18")
    s.close()
