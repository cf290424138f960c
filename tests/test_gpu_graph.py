"""Batched evaluate() through one HIP graph (pdmk_gpu_graph_create / _launch / _destroy).

A graph replays the same kernels on the same buffers as eager evaluate() calls, so every
replay must give exactly the eager potentials.  The replays are separated by eager work on
the same trees and buffers (other charge vectors, outputs cleared) and by synchronizations,
and each replay is checked.  The trees are a small version of the batched workload
(workloads.batch_cells): cubic and triclinic cells, N log-uniform in [2e3, 2e4], 6 digits in
FP32 and 9 digits in FP64.
"""
import numpy as np
import pytest

from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trees(pdmk):
    out = []
    for i, (cell, n) in enumerate(W.batch_cells(6, n_max=20_000)):
        prec, digits = ("float", 6) if i % 2 == 0 else ("double", 9)
        s = pdmk.PDMKSolver(cell, W.uniform_positions(n, 1000 + i, cell), n_digits=digits, n_per_leaf=300, nz=4,
                            precision=prec)
        qs = [pdmk.DeviceArray.from_numpy(W.neutral_uniform_charges(n, 100000 + 32 * i + v)) for v in range(3)]
        outs = [pdmk.DeviceArray(n, s.dtype) for _ in range(3)]
        out.append((s, qs, outs))
    yield out
    for s, qs, outs in out:
        for a in qs + outs:
            a.free()
        s.close()


def eager(pdmk, trees):
    res = []
    for s, qs, outs in trees:
        for q, o in zip(qs, outs):
            s.evaluate_device(q.ptr, o.ptr)
        pdmk.synchronize()
        res.append([o.numpy() for o in outs])
    return res


def test_graph_replays_equal_eager(pdmk, cuda, trees):
    ref = eager(pdmk, trees)
    calls = [(s, q, o) for s, qs, outs in trees for q, o in zip(qs, outs)]
    g = pdmk.BatchGraph(calls, n_branches=4)
    try:
        for rep in range(3):
            # eager work on the same trees between replays: another charge vector into the
            # first output, then every output cleared
            s0, qs0, outs0 = trees[0]
            tmp = pdmk.DeviceArray.from_numpy(W.neutral_uniform_charges(s0.n, 77 + rep))
            s0.evaluate_device(tmp.ptr, outs0[0].ptr)
            pdmk.synchronize()
            tmp.free()
            for _, _, outs in trees:
                for o in outs:
                    o.upload(np.zeros(o.n, dtype=o.dtype))
            g.launch()
            pdmk.synchronize()
            for t, (_, _, outs) in enumerate(trees):
                for v, o in enumerate(outs):
                    np.testing.assert_array_equal(o.numpy(), ref[t][v], err_msg=f"replay {rep} tree {t} vector {v}")
    finally:
        g.close()


def test_graph_single_branch_and_repeated_handle(pdmk, cuda, trees):
    """One branch, and one tree called twice on the same buffers: the calls keep their order."""
    ref = eager(pdmk, trees)
    s, qs, outs = trees[1]
    g = pdmk.BatchGraph([(s, qs[2], outs[0]), (s, qs[0], outs[0])], n_branches=1)
    try:
        outs[0].upload(np.zeros(outs[0].n, dtype=outs[0].dtype))
        g.launch()
        pdmk.synchronize()
        np.testing.assert_array_equal(outs[0].numpy(), ref[1][0])
    finally:
        g.close()


def test_single_stream_equals_concurrent(pdmk, cuda, trees):
    """pdmk_gpu_set_single_stream runs the same kernels on the same values, one stream."""
    ref = eager(pdmk, trees)
    for s, _, _ in trees:
        s.set_single_stream(True)
    try:
        got = eager(pdmk, trees)
    finally:
        for s, _, _ in trees:
            s.set_single_stream(False)
    for t in range(len(trees)):
        for v in range(3):
            np.testing.assert_array_equal(got[t][v], ref[t][v])


def test_graph_errors(pdmk, cuda, trees):
    s, qs, outs = trees[0]
    with pytest.raises(RuntimeError):
        pdmk.BatchGraph([], n_branches=2)
    with pytest.raises(RuntimeError):
        pdmk.BatchGraph([(s, qs[0], outs[0])], n_branches=0)
    cell = W.cubic_cell()
    qp = pdmk.PDMKSolver(cell, W.uniform_positions(3000, 5, cell), n_digits=6, n_per_leaf=300, nz=4,
                         precision="double", bloch=(0.1, 0.2, 0.3))
    q = pdmk.DeviceArray.from_numpy(np.zeros(3000))
    o = pdmk.DeviceArray(2 * 3000, np.float64)
    try:
        with pytest.raises(RuntimeError):
            pdmk.BatchGraph([(qp, q, o)], n_branches=1)
        # the library still works after the rejected capture
        s.evaluate_device(qs[0].ptr, outs[0].ptr)
        pdmk.synchronize()
    finally:
        q.free()
        o.free()
        qp.close()
