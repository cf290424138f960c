"""CPU oracle (test infrastructure only).

ctypes wrapper around ``oracle/_build/liboracle.so``, the CPU restatement of the
reference ``PeriodicDMKTreeT<Real,3>`` construction and ``evaluate()``
(cpp/src/periodic_dmk.cpp:41-186) plus the FP64 Ewald sums
(cpp/src/ewald_nufft.cpp:777-910).  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may use this module, and only as the checker.

Parity unpinned (checked end to end against the FP64 Ewald sum only): SCTL's
approx_rsqrt inside the reference residual kernel (replaced by an exact 1/sqrt), DUCC's
NUFFT far field (replaced by the exact sum it approximates to eps/4), and the reference's
exact tree and intermediate arrays, for which no fixture exists.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

STAGE_PW, STAGE_DIRECT, STAGE_SELF, STAGE_FAR = 1, 2, 4, 8
STAGE_ALL = 15


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ / OpenMP)."""
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def use_native() -> str:
    """Build (if needed) and load the -march=native variant of the oracle for this host
    instead of the portable one; call before the first use of the oracle.  Used by the timed
    CPU baseline, so that it runs at the host's full vector width."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("use_native() must come before the oracle library is loaded")
    subprocess.run(["make", "-C", _HERE, "native"], check=True, stdout=subprocess.DEVNULL)
    _LIB_PATH = os.path.join(_HERE, "_build", "native", "liboracle.so")
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, C.c_int, C.c_int]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_evaluate.argtypes = [C.c_void_p, dp, C.c_int, dp]
        L.oracle_evaluate_complex.argtypes = [C.c_void_p, dp, dp, C.c_int, dp]
        L.oracle_info.argtypes = [C.c_void_p, ip, dp]
        L.oracle_tree_arrays.argtypes = [C.c_void_p, ip, ip, ip, dp]
        L.oracle_pswf.restype = C.c_double
        L.oracle_pswf.argtypes = [C.c_double, C.c_double]
        L.oracle_pswf_integral.restype = C.c_double
        L.oracle_pswf_integral.argtypes = [C.c_double, C.c_double]
        L.oracle_pswf_intvals.argtypes = [C.c_double, dp]
        L.oracle_gauss_legendre.argtypes = [C.c_int, dp, dp]
        L.oracle_residual_coeffs.restype = C.c_int
        L.oracle_residual_coeffs.argtypes = [C.c_int, dp]
        L.oracle_beta.restype = C.c_double
        L.oracle_beta.argtypes = [C.c_double]
        L.oracle_cheb_tables.argtypes = [C.c_int, dp, dp, dp]
        L.oracle_ewald_direct.argtypes = [dp, C.c_int, dp, dp, C.c_double, dp, dp]
        L.oracle_ewald_cells.argtypes = [dp, C.c_int, dp, dp, C.c_double, C.c_int, ip, dp, C.c_double]
        L.oracle_set_threads.restype = C.c_int
        L.oracle_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's loops; returns the previous count."""
    return int(lib().oracle_set_threads(int(n)))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class OracleTree:
    """Mirror of PeriodicDMKTreeT<Real,3> on the CPU (Real = float or double)."""

    INFO_KEYS = ("n_src", "nsall", "nboxes", "n_levels", "p", "n_pw", "ms", "mt", "mu", "n_leaves",
                 "n_list1", "n_direct_work", "n_c2p_work", "n_digits")

    def __init__(self, cell, positions, n_digits=6, n_per_leaf=300, nz=4, precision="double"):
        cell = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3).T)  # column-major
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
        self.n = pos.shape[0]
        self._h = lib().oracle_create(1 if precision == "float" else 0, _d(cell), self.n, _d(pos),
                                      int(n_digits), int(n_per_leaf), int(nz))
        if not self._h:
            raise RuntimeError("oracle_create failed")

    def evaluate(self, charges, stages=STAGE_ALL):
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        if q.shape[0] != self.n:
            raise ValueError("charges size != n_src")
        out = np.zeros(self.n)
        if lib().oracle_evaluate(self._h, _d(q), int(stages), _d(out)) != 0:
            raise RuntimeError("oracle_evaluate failed")
        return out

    def evaluate_complex(self, charges, bloch, stages=STAGE_ALL):
        """Quasi-periodic potentials (complex, user order) for Bloch vector `bloch`
        (PeriodicDMKTreeT::evaluate_complex, cpp/src/periodic_dmk.cpp:251-277)."""
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        if q.shape[0] != self.n:
            raise ValueError("charges size != n_src")
        b = np.ascontiguousarray(np.asarray(bloch, dtype=np.float64).reshape(3))
        out = np.zeros(2 * self.n)
        if lib().oracle_evaluate_complex(self._h, _d(q), _d(b), int(stages), _d(out)) != 0:
            raise RuntimeError("oracle_evaluate_complex failed")
        return out[0::2] + 1j * out[1::2]

    def info(self):
        ints = np.zeros(len(self.INFO_KEYS), dtype=np.int32)
        dbls = np.zeros(6)
        lib().oracle_info(self._h, _i(ints), _d(dbls))
        d = dict(zip(self.INFO_KEYS, (int(v) for v in ints)))
        d.update(dict(zip(("rc", "beta", "psi0", "c0", "g0d2", "w0_0"), (float(v) for v in dbls))))
        return d

    def tree_arrays(self):
        inf = self.info()
        nb = inf["nboxes"]
        box7 = np.zeros(7 * nb, dtype=np.int32)
        sp = np.zeros(inf["nsall"], dtype=np.int32)
        tp = np.zeros(inf["n_src"], dtype=np.int32)
        cen = np.zeros(3 * nb)
        lib().oracle_tree_arrays(self._h, _i(box7), _i(sp), _i(tp), _d(cen))
        return box7.reshape(nb, 7), sp, tp, cen.reshape(nb, 3)

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.oracle_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pswf(c, x):
    return lib().oracle_pswf(float(c), float(x))


def pswf_integral(c, r):
    return lib().oracle_pswf_integral(float(c), float(r))


def pswf_intvals(c):
    out = np.zeros(4)
    lib().oracle_pswf_intvals(float(c), _d(out))
    return out


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().oracle_gauss_legendre(int(n), _d(x), _d(w))
    return x, w


def residual_coeffs(n_digits):
    nc = lib().oracle_residual_coeffs(int(n_digits), C.POINTER(C.c_double)())
    out = np.zeros(nc)
    lib().oracle_residual_coeffs(int(n_digits), _d(out))
    return out


def beta(eps):
    return lib().oracle_beta(float(eps))


def cheb_tables(p):
    c2p, p2c, vinv = np.zeros(24 * p * p), np.zeros(24 * p * p), np.zeros(p * p)
    lib().oracle_cheb_tables(int(p), _d(c2p), _d(p2c), _d(vinv))
    return c2p.reshape(8, 3, p, p), p2c.reshape(8, 3, p, p), vinv.reshape(p, p)


def ewald_direct(cell, positions, charges, s=4.8):
    cellc = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3).T)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
    q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
    pot = np.zeros(pos.shape[0])
    e = np.zeros(1)
    if lib().oracle_ewald_direct(_d(cellc), pos.shape[0], _d(pos), _d(q), float(s), _d(pot), _d(e)) != 0:
        raise RuntimeError("ewald_direct failed")
    return pot, float(e[0])


def ewald_cells(cell, positions, charges, s=5.0, targets=None, alpha=0.0):
    """FP64 Ewald potentials at `targets` (all points if None); alpha > 0 overrides the
    splitting parameter (bounded below so that r_cut < h_min / 2)."""
    cellc = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3).T)
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
    q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
    if targets is None:
        tidx = np.arange(pos.shape[0], dtype=np.int32)
    else:
        tidx = np.ascontiguousarray(np.asarray(targets, dtype=np.int32))
    pot = np.zeros(tidx.shape[0])
    if lib().oracle_ewald_cells(_d(cellc), pos.shape[0], _d(pos), _d(q), float(s), tidx.shape[0], _i(tidx),
                                _d(pot), float(alpha)) != 0:
        raise RuntimeError("ewald_cells failed")
    return pot
