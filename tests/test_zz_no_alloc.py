"""The allocation-free Ewald entry point (pdmk_ewald_nufft_noalloc) against pdmk_ewald_nufft.

pdmk_ewald_nufft_noalloc keeps every device array of the Ewald sum in a static per-GPU
scratch array instead of allocating.  Its potentials and energy must equal those of
pdmk_ewald_nufft: bit for bit when the structure-factor partial sums fit in the scratch at
the regular block count, and to FP64 rounding (1e-12 relative) when the scratch caps the
block count (the partial sums are then added in fewer, larger groups).  Two calls in a row
share the scratch, the second on a smaller input; a request beyond the scratch is refused
on the host before any GPU work.
"""
import ctypes as C

import numpy as np
import pytest

import periodicdmk_amd as P
from periodicdmk_amd import workloads as W

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _fn():
    L = P.lib()
    f = L.pdmk_ewald_nufft_noalloc
    f.restype = C.c_int
    f.argtypes = [_dp, C.c_int, _dp, _dp, C.c_double, _dp, _dp]
    g = L.pdmk_ewald_noalloc_scratch_bytes
    g.restype = C.c_ulonglong
    g.argtypes = []
    return f, g


def _noalloc(cell, pos, q, s):
    f, _ = _fn()
    cellc = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).T)  # column j = lattice vector j
    p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64))
    qq = np.ascontiguousarray(np.asarray(q, dtype=np.float64))
    pot = np.zeros(p.shape[0])
    e = C.c_double()
    rc = f(cellc.ctypes.data_as(_dp), p.shape[0], p.ctypes.data_as(_dp), qq.ctypes.data_as(_dp), float(s),
           pot.ctypes.data_as(_dp), C.byref(e))
    return rc, pot, e.value


def test_noalloc_equals_ewald_two_calls():
    _, scratch = _fn()
    assert scratch() == 128 << 20
    # a large call whose partial sums the scratch caps, then a smaller one at the regular
    # block count on the scratch the first call left behind
    cases = ((100_000, W.triclinic_cell(), 1e-12), (3000, W.cubic_cell(), 0.0))
    for n, cell, tol in cases:
        print(f"no-alloc Ewald: N={n}", flush=True)
        pos = W.uniform_positions(n, 41, cell)
        q = W.neutral_uniform_charges(n, 42)
        ref, e_ref = P.ewald(cell, pos, q, s=5.0)
        rc, pot, e = _noalloc(cell, pos, q, 5.0)
        assert rc == 0
        err = np.max(np.abs(pot - ref)) / np.max(np.abs(ref))
        print(f"  max rel diff {err:.3e}, energy {e:.15e} vs {e_ref:.15e}", flush=True)
        if tol == 0.0:
            assert np.array_equal(pot, ref) and e == e_ref
        else:
            assert err <= tol and abs(e - e_ref) <= tol * abs(e_ref)


def test_noalloc_refuses_oversize():
    n = 2_000_000  # 92 bytes of per-point arrays alone exceed the 128 MiB scratch
    cell = W.cubic_cell()
    pos = W.uniform_positions(n, 43, cell)
    q = W.neutral_uniform_charges(n, 44)
    rc, _, _ = _noalloc(cell, pos, q, 5.0)
    assert rc == -1
