"""MI355X periodic DMK solver: Python mirror of the reference bindings.

The product is ``periodicdmk_amd/lib/libpdmk.so`` (host C++ + gfx950 HIP kernels), whose
C ABI (``include/pdmk/pdmk_capi.h``) replaces the reference ``libpdmk_jl``.  This module
binds it with ctypes the way ``julia/PeriodicDMK/src/PeriodicDMK.jl`` binds the reference:

* ``PDMKSolver(cell, positions; n_digits, n_per_leaf, nz, precision)``
  (PeriodicDMK.jl:95-143) builds the tree once;
* ``evaluate(solver, charges)`` / ``solver.evaluate(charges)`` (PeriodicDMK.jl:188-211)
  evaluates potentials for new charges;
* ``solver.evaluate_device(d_charges, d_pot, stream)`` is the device-resident variant.

``cell`` has the lattice vectors as columns; ``positions`` is (N, 3).  There is no
fallback: if the library is not built, loading fails with an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpdmk.so")
_lib = None

STAGE_PW, STAGE_DIRECT, STAGE_SELF, STAGE_FAR = 1, 2, 4, 8
STATS_KEYS = ("n_src", "nsall", "nboxes", "n_levels", "p", "n_pw", "ms", "mt", "mu", "n_leaves", "n_list1",
              "n_direct_leaves", "n_far_modes", "n_digits", "is_float", "n_nf_steps")
WORK_KEYS = ("nf_pairs_support", "nf_pairs_candidate", "nf_flops", "transform_flops", "ff_mode_points", "ff_flops",
             "on_bytes", "nf_bytes")


def lib():
    """Load libpdmk.so from the package directory (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m periodicdmk_amd.build`")
    L = C.CDLL(LIB_PATH)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
    lp = C.POINTER(C.c_longlong)
    fp = C.POINTER(C.c_float)
    sig = {
        "pdmk_ewald_nufft": (C.c_int, [dp, C.c_int, dp, dp, C.c_double, dp, dp]),
        "pdmk_gpu_ewald_targets": (C.c_int, [dp, C.c_int, dp, dp, C.c_int, ip, C.c_double, dp]),
        "pdmk_optimize_cubic_cover": (C.c_int, [dp, C.c_int, ip, ip, ip, dp, dp, dp, dp, dp]),
        "pdmk_create_double": (vp, [dp, C.c_int, dp, C.c_int, C.c_int, C.c_int]),
        "pdmk_evaluate_double": (C.c_int, [vp, C.c_int, dp, dp, dp]),
        "pdmk_destroy_double": (None, [vp]),
        "pdmk_create_float": (vp, [dp, C.c_int, dp, C.c_int, C.c_int, C.c_int]),
        "pdmk_evaluate_float": (C.c_int, [vp, C.c_int, dp, dp, fp]),
        "pdmk_destroy_float": (None, [vp]),
        "pdmk_gpu_device_count": (C.c_int, []),
        "pdmk_gpu_malloc": (vp, [C.c_ulonglong]),
        "pdmk_gpu_free": (C.c_int, [vp]),
        "pdmk_gpu_memcpy": (C.c_int, [vp, vp, C.c_ulonglong, C.c_int]),
        "pdmk_gpu_synchronize": (C.c_int, []),
        "pdmk_gpu_n_points": (C.c_int, [vp]),
        "pdmk_gpu_evaluate_double": (C.c_int, [vp, vp, vp, vp]),
        "pdmk_gpu_evaluate_float": (C.c_int, [vp, vp, vp, vp]),
        "pdmk_gpu_evaluate_timed": (C.c_int, [vp, vp, vp, dp]),
        "pdmk_gpu_evaluate_stages": (C.c_int, [vp, vp, vp, C.c_int]),
        "pdmk_gpu_stats": (C.c_int, [vp, lp]),
        "pdmk_gpu_work": (C.c_int, [vp, dp]),
        "pdmk_gpu_tree_arrays": (C.c_int, [vp, ip, dp, ip, ip]),
        "pdmk_gpu_enable_kernel_timing": (C.c_int, [vp, C.c_int]),
        "pdmk_gpu_kernel_times": (C.c_int, [vp, dp]),
        "pdmk_host_tree": (C.c_int, [dp, C.c_int, dp, C.c_int, C.c_int, lp, ip, dp, ip, ip, dp]),
        "pdmk_host_tables": (C.c_int, [C.c_int, C.c_double, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "pdmk_create_qp_double": (C.c_void_p, [dp, C.c_int, dp, C.c_int, C.c_int, C.c_int, dp]),
        "pdmk_create_qp_float": (C.c_void_p, [dp, C.c_int, dp, C.c_int, C.c_int, C.c_int, dp]),
        "pdmk_evaluate_complex_double": (C.c_int, [vp, C.c_int, dp, dp, dp]),
        "pdmk_evaluate_complex_float": (C.c_int, [vp, C.c_int, dp, dp, C.POINTER(C.c_float)]),
        "pdmk_gpu_evaluate_complex": (C.c_int, [vp, vp, vp, C.c_int]),
        "pdmk_gpu_build_info": (C.c_int, [vp, dp]),
        "pdmk_gpu_autotune": (C.c_int, [vp, C.c_int, dp]),
        "pdmk_gpu_evaluate_many": (C.c_int, [vp, C.c_int, vp, vp, vp]),
        "pdmk_gpu_knobs": (C.c_int, [vp, ip, ip]),
        "pdmk_gpu_tune_cache_clear": (None, []),
        "pdmk_gpu_graph_create": (vp, [C.c_int, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.c_int]),
        "pdmk_gpu_graph_launch": (C.c_int, [vp, vp]),
        "pdmk_gpu_set_single_stream": (C.c_int, [vp, C.c_int]),
        "pdmk_gpu_graph_destroy": (C.c_int, [vp]),
        "pdmk_campaign_sha256": (C.c_int, [vp, C.c_size_t, vp]),
        "pdmk_campaign_write": (C.c_int, [C.c_char_p, C.c_uint, C.c_uint, C.c_int, C.c_int, C.c_ulonglong, dp,
                                          C.c_uint, C.POINTER(C.c_uint), dp, dp, dp]),
        "pdmk_campaign_read_header": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint), C.POINTER(C.c_uint), ip, ip,
                                                C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint)]),
        "pdmk_campaign_read": (C.c_int, [C.c_char_p, dp, C.POINTER(C.c_uint), dp, dp, dp]),
        "pdmk_host_pswf": (C.c_double, [C.c_double, C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else C.POINTER(C.c_double)()


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else C.POINTER(C.c_int)()


def _cell_colmajor(cell):
    cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    return np.ascontiguousarray(cell.T)  # memory: cell[i + 3 j] = component i of lattice vector j


def _positions(positions):
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    return pos


def device_count() -> int:
    return int(lib().pdmk_gpu_device_count())


def synchronize() -> None:
    if lib().pdmk_gpu_synchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def optimize_cubic_cover(cell, nz: int = 4) -> dict:
    """Cubic cover of a cell (pdmk_optimize_cubic_cover): m, n, p, orthonormal frame e1, e2,
    e3, cube side s and overhead cost.  Raises ValueError for nz <= 0 or a degenerate cell."""
    cellc = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3).T)  # column-major
    m, n, p = C.c_int(), C.c_int(), C.c_int()
    e = [np.zeros(3) for _ in range(3)]
    s, cost = C.c_double(), C.c_double()
    dp = C.POINTER(C.c_double)
    rc = lib().pdmk_optimize_cubic_cover(cellc.ctypes.data_as(dp), int(nz), C.byref(m), C.byref(n), C.byref(p),
                                         *(v.ctypes.data_as(dp) for v in e), C.byref(s), C.byref(cost))
    if rc != 0:
        raise ValueError(f"optimize_cubic_cover failed (code {rc})")
    return {"m": m.value, "n": n.value, "p": p.value, "e1": e[0], "e2": e[1], "e3": e[2], "s": s.value,
            "cost": cost.value}


def ewald(cell, positions, charges, s: float = 5.0):
    """FP64 Ewald potentials and energy on the GPU (pdmk_ewald_nufft); returns (pot, energy)."""
    cellc = _cell_colmajor(cell)
    pos = _positions(positions)
    q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
    if q.shape != (pos.shape[0],):
        raise ValueError("charges must have shape (N,)")
    pot = np.zeros(pos.shape[0])
    e = C.c_double()
    if lib().pdmk_ewald_nufft(_d(cellc), pos.shape[0], _d(pos), _d(q), float(s), _d(pot), C.byref(e)) != 0:
        raise RuntimeError("pdmk_ewald_nufft failed")
    return pot, e.value


def ewald_targets(cell, positions, charges, targets, s: float = 5.0):
    """FP64 Ewald potentials on the GPU at the points ``targets`` (indices into ``positions``)
    only (pdmk_gpu_ewald_targets); equal to ``ewald(...)[0][targets]``."""
    cellc = _cell_colmajor(cell)
    pos = _positions(positions)
    q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
    if q.shape != (pos.shape[0],):
        raise ValueError("charges must have shape (N,)")
    t = np.ascontiguousarray(np.asarray(targets, dtype=np.int32).ravel())
    pot = np.zeros(t.size)
    if lib().pdmk_gpu_ewald_targets(_d(cellc), pos.shape[0], _d(pos), _d(q), t.size, _i(t), float(s), _d(pot)) != 0:
        raise RuntimeError("pdmk_gpu_ewald_targets failed")
    return pot


class BatchGraph:
    """evaluate() calls on device buffers captured into one HIP graph (pdmk_gpu_graph_create).

    ``calls`` is a list of (solver, charges DeviceArray, potentials DeviceArray); the buffers
    are bound at creation and each ``launch()`` reads the charges they hold then."""

    def __init__(self, calls, n_branches: int = 8):
        n = len(calls)
        H = (C.c_void_p * max(1, n))(*[c[0]._h for c in calls])
        Q = (C.c_void_p * max(1, n))(*[c[1].ptr for c in calls])
        O = (C.c_void_p * max(1, n))(*[c[2].ptr for c in calls])
        self._g = lib().pdmk_gpu_graph_create(n, H, Q, O, int(n_branches))
        if not self._g:
            raise RuntimeError("pdmk_gpu_graph_create failed")

    def launch(self, stream: int = 0) -> None:
        """Enqueue one replay (on the graph's own stream unless ``stream`` is given)."""
        if lib().pdmk_gpu_graph_launch(self._g, C.c_void_p(stream) if stream else None) != 0:
            raise RuntimeError("pdmk_gpu_graph_launch failed")

    def close(self) -> None:
        if getattr(self, "_g", None):
            lib().pdmk_gpu_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def memcpy_d2d(dst: int, src: int, nbytes: int) -> None:
    """Device-to-device copy of nbytes between two device addresses."""
    if lib().pdmk_gpu_memcpy(C.c_void_p(dst), C.c_void_p(src), int(nbytes), 3) != 0:
        raise RuntimeError("device->device copy failed")


class DeviceArray:
    """A 1-D device buffer allocated through libpdmk.so (one HIP runtime per process)."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = lib().pdmk_gpu_malloc(self.nbytes)
        if not self.ptr:
            raise MemoryError(f"device allocation of {self.nbytes} bytes failed")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError("size mismatch")
        if lib().pdmk_gpu_memcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, 1) != 0:
            raise RuntimeError("host->device copy failed")

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        if lib().pdmk_gpu_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 2) != 0:
            raise RuntimeError("device->host copy failed")
        return out

    def offset(self, elements: int) -> int:
        return self.ptr + int(elements) * self.dtype.itemsize

    def free(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.pdmk_gpu_free(C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PDMKSolver:
    """Reusable periodic DMK tree on the GPU (mirror of the Julia ``PDMKSolver``)."""

    def __init__(self, cell, positions, n_digits=6, n_per_leaf=300, nz=4, precision="auto", bloch=None):
        if precision == "auto":
            precision = "float" if n_digits <= 6 else "double"
        if precision not in ("float", "double"):
            raise ValueError("precision must be 'float', 'double' or 'auto'")
        self.precision = precision
        self.dtype = np.float32 if precision == "float" else np.float64
        cellc = _cell_colmajor(cell)
        pos = _positions(positions)
        self.bloch = None if bloch is None else np.ascontiguousarray(np.asarray(bloch, dtype=np.float64).reshape(3))
        if self.bloch is None:
            create = lib().pdmk_create_float if precision == "float" else lib().pdmk_create_double
            self._h = create(_d(cellc), pos.shape[0], _d(pos), int(n_digits), int(n_per_leaf), int(nz))
        else:
            create = lib().pdmk_create_qp_float if precision == "float" else lib().pdmk_create_qp_double
            self._h = create(_d(cellc), pos.shape[0], _d(pos), int(n_digits), int(n_per_leaf), int(nz), _d(self.bloch))
        if not self._h:
            raise RuntimeError("pdmk_create failed (see stderr)")
        self.n = pos.shape[0]
        self.n_digits = n_digits

    def evaluate(self, charges, positions=None):
        """Potentials (user order) for ``charges``; ``positions`` is ignored as in the reference."""
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        if q.shape != (self.n,):
            raise ValueError("charges size != n_src")
        out = np.empty(self.n, dtype=self.dtype)
        if self.precision == "float":
            rc = lib().pdmk_evaluate_float(self._h, self.n, C.POINTER(C.c_double)(), _d(q),
                                           out.ctypes.data_as(C.POINTER(C.c_float)))
        else:
            rc = lib().pdmk_evaluate_double(self._h, self.n, C.POINTER(C.c_double)(), _d(q), _d(out))
        if rc != 0:
            raise RuntimeError("pdmk_evaluate failed (see stderr)")
        return out

    def evaluate_complex(self, charges, positions=None):
        """Quasi-periodic potentials (complex, user order); needs ``bloch`` at construction
        (PeriodicDMKTreeT::evaluate_complex)."""
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        if q.shape != (self.n,):
            raise ValueError("charges size != n_src")
        out = np.empty(2 * self.n, dtype=self.dtype)
        if self.precision == "float":
            rc = lib().pdmk_evaluate_complex_float(self._h, self.n, C.POINTER(C.c_double)(), _d(q),
                                                   out.ctypes.data_as(C.POINTER(C.c_float)))
        else:
            rc = lib().pdmk_evaluate_complex_double(self._h, self.n, C.POINTER(C.c_double)(), _d(q), _d(out))
        if rc != 0:
            raise RuntimeError("pdmk_evaluate_complex failed (see stderr)")
        return out[0::2].astype(np.float64) + 1j * out[1::2].astype(np.float64)

    def run_complex_stages(self, charges, stages: int):
        """Quasi-periodic, device path, selected stages only (bits as in STAGE_*)."""
        qd = DeviceArray.from_numpy(np.asarray(charges, dtype=np.float64))
        out = DeviceArray(2 * self.n, self.dtype)
        try:
            if lib().pdmk_gpu_evaluate_complex(self._h, C.c_void_p(qd.ptr), C.c_void_p(out.ptr), int(stages)) != 0:
                raise RuntimeError("pdmk_gpu_evaluate_complex failed (see stderr)")
            res = out.numpy().astype(np.float64)
        finally:
            qd.free()
            out.free()
        return res[0::2] + 1j * res[1::2]

    def evaluate_device(self, d_charges: int, d_pot: int, stream: int = 0):
        """Enqueue evaluate() on device pointers (FP64 charges in, Real potentials out)."""
        f = lib().pdmk_gpu_evaluate_float if self.precision == "float" else lib().pdmk_gpu_evaluate_double
        if f(self._h, C.c_void_p(d_charges), C.c_void_p(d_pot), C.c_void_p(stream)) != 0:
            raise RuntimeError("pdmk_gpu_evaluate failed (see stderr)")

    def evaluate_many_device(self, n_vec: int, d_charges: int, d_pot: int, stream: int = 0):
        """Enqueue evaluate() of n_vec vectors ([n_vec][N] FP64 charges in, Real potentials out,
        device pointers; pdmk_gpu_evaluate_many)."""
        if lib().pdmk_gpu_evaluate_many(self._h, int(n_vec), C.c_void_p(d_charges), C.c_void_p(d_pot),
                                        C.c_void_p(stream)) != 0:
            raise RuntimeError("pdmk_gpu_evaluate_many failed (see stderr)")

    def evaluate_many(self, charges):
        """Potentials (n_vec, N) for charges (n_vec, N), host arrays, through pdmk_gpu_evaluate_many."""
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        if q.ndim != 2 or q.shape[1] != self.n:
            raise ValueError("charges must have shape (n_vec, N)")
        qd = DeviceArray.from_numpy(q.reshape(-1))
        out = DeviceArray(q.size, self.dtype)
        try:
            self.evaluate_many_device(q.shape[0], qd.ptr, out.ptr)
            synchronize()
            return out.numpy().reshape(q.shape)
        finally:
            qd.free()
            out.free()

    def set_single_stream(self, on: bool = True) -> None:
        """evaluate() on the caller's stream only (pdmk_gpu_set_single_stream): for many trees
        evaluated side by side."""
        if lib().pdmk_gpu_set_single_stream(self._h, 1 if on else 0) != 0:
            raise RuntimeError("pdmk_gpu_set_single_stream failed")

    def evaluate_timed(self, d_charges: int, d_pot: int):
        t = np.zeros(3)
        if lib().pdmk_gpu_evaluate_timed(self._h, C.c_void_p(d_charges), C.c_void_p(d_pot), _d(t)) != 0:
            raise RuntimeError("pdmk_gpu_evaluate_timed failed (see stderr)")
        return {"nufft_ms": t[0], "dmk_ms": t[1], "direct_ms": t[2]}

    def evaluate_stages(self, d_charges: int, d_pot: int, stages: int):
        if lib().pdmk_gpu_evaluate_stages(self._h, C.c_void_p(d_charges), C.c_void_p(d_pot), int(stages)) != 0:
            raise RuntimeError("pdmk_gpu_evaluate_stages failed (see stderr)")

    def run_stages(self, charges, stages: int):
        """Host arrays in/out: run only the selected stages (bits as in STAGE_*)."""
        qd = DeviceArray.from_numpy(np.asarray(charges, dtype=np.float64))
        out = DeviceArray(self.n, self.dtype)
        self.evaluate_stages(qd.ptr, out.ptr, stages)
        res = out.numpy()
        qd.free()
        out.free()
        return res

    def stats(self):
        o = np.zeros(16, dtype=np.int64)
        lib().pdmk_gpu_stats(self._h, o.ctypes.data_as(C.POINTER(C.c_longlong)))
        return dict(zip(STATS_KEYS, (int(v) for v in o)))

    def build_info(self):
        """Construction times (ms) and whether the tree's per-point work ran on the GPU."""
        o = np.zeros(8)
        lib().pdmk_gpu_build_info(self._h, _d(o))
        return {"build_ms": float(o[0]), "tree_ms": float(o[1]), "far_plan_ms": float(o[2]), "tree_on_gpu": bool(o[3]),
                "folded_boxes": int(o[4]), "proxy_slots": int(o[5]), "pw_by_level": bool(o[6]),
                "near_rows": bool(o[7])}

    KNOBS = ("near_wpb", "near_late", "p2pw_bk", "pw2p_bk", "t3_div")

    def autotune(self, repeats: int = 3):
        """Time the launch-shape candidates on this tree and apply the fastest
        (pdmk_gpu_autotune); returns ms before and after and the chosen knobs."""
        o = np.zeros(8)
        if lib().pdmk_gpu_autotune(self._h, int(repeats), _d(o)) != 0:
            raise RuntimeError("pdmk_gpu_autotune failed (see stderr)")
        return {"ms_default": float(o[0]), "ms_tuned": float(o[1]),
                "knobs": dict(zip(self.KNOBS, (int(v) for v in o[2:7]))), "from_cache": bool(o[7])}

    def knobs(self, **set_values):
        """Current launch-shape knobs; keyword arguments set them first (pdmk_gpu_knobs)."""
        cur = np.zeros(5, dtype=np.int32)
        if lib().pdmk_gpu_knobs(self._h, None, _i(cur)) != 0:
            raise RuntimeError("pdmk_gpu_knobs failed")
        if set_values:
            new = dict(zip(self.KNOBS, (int(v) for v in cur)))
            new.update(set_values)
            arr = np.array([new[k] for k in self.KNOBS], dtype=np.int32)
            if lib().pdmk_gpu_knobs(self._h, _i(arr), _i(cur)) != 0:
                raise ValueError(f"pdmk_gpu_knobs rejected {set_values}")
        return dict(zip(self.KNOBS, (int(v) for v in cur)))

    def work(self):
        o = np.zeros(8)
        lib().pdmk_gpu_work(self._h, _d(o))
        return dict(zip(WORK_KEYS, (float(v) for v in o)))

    KERNEL_GROUPS = ("far_field", "gather", "transforms", "near_field", "finalize")

    def enable_kernel_timing(self, slots=64):
        if lib().pdmk_gpu_enable_kernel_timing(self._h, int(slots)) != 0:
            raise RuntimeError("pdmk_gpu_enable_kernel_timing failed")

    def kernel_times(self):
        """Mean ms per call of each kernel group over the recorded calls (stream must be synchronized)."""
        o = np.zeros(5)
        n = lib().pdmk_gpu_kernel_times(self._h, _d(o))
        if n < 0:
            raise RuntimeError("pdmk_gpu_kernel_times failed")
        return dict(zip(self.KERNEL_GROUPS, (float(v) for v in o))), int(n)

    def tree_arrays(self):
        s = self.stats()
        nb = s["nboxes"]
        box7 = np.zeros(7 * nb, dtype=np.int32)
        cen = np.zeros(3 * nb)
        sp = np.zeros(s["nsall"], dtype=np.int32)
        tp = np.zeros(s["n_src"], dtype=np.int32)
        lib().pdmk_gpu_tree_arrays(self._h, _i(box7), _d(cen), _i(sp), _i(tp))
        return box7.reshape(nb, 7), sp, tp, cen.reshape(nb, 3)

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            (_lib.pdmk_destroy_float if self.precision == "float" else _lib.pdmk_destroy_double)(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def evaluate(solver: PDMKSolver, *args):
    """``evaluate(solver, charges)`` or ``evaluate(solver, positions, charges)``."""
    return solver.evaluate(args[-1])


def host_tree(cell, positions, n_per_leaf=300, nz=4):
    """Host-side tree of the library (no GPU): box7, src_perm, trg_perm, centers, geometry."""
    cellc = _cell_colmajor(cell)
    pos = _positions(positions)
    sizes = np.zeros(8, dtype=np.int64)
    L = lib()
    if L.pdmk_host_tree(_d(cellc), pos.shape[0], _d(pos), int(n_per_leaf), int(nz),
                        sizes.ctypes.data_as(C.POINTER(C.c_longlong)), None, None, None, None, None) != 0:
        raise RuntimeError("pdmk_host_tree failed")
    nb, nsall, n = int(sizes[2]), int(sizes[1]), int(sizes[0])
    box7 = np.zeros(7 * nb, dtype=np.int32)
    cen = np.zeros(3 * nb)
    sp = np.zeros(nsall, dtype=np.int32)
    tp = np.zeros(n, dtype=np.int32)
    geom = np.zeros(16)
    L.pdmk_host_tree(_d(cellc), pos.shape[0], _d(pos), int(n_per_leaf), int(nz),
                     sizes.ctypes.data_as(C.POINTER(C.c_longlong)), _i(box7), _d(cen), _i(sp), _i(tp), _d(geom))
    return {"sizes": sizes, "box7": box7.reshape(nb, 7), "centers": cen.reshape(nb, 3), "src_perm": sp,
            "trg_perm": tp, "umat": geom[:9].reshape(3, 3).T, "rc": geom[9], "mgrid": geom[10:13].astype(int),
            "xgrid0": geom[13:16]}


def host_tables(n_digits, rc, n_levels, umat_colmajor):
    umat = np.ascontiguousarray(np.asarray(umat_colmajor, dtype=np.float64).reshape(9))
    s8 = np.zeros(8)
    L = lib()
    if L.pdmk_host_tables(int(n_digits), float(rc), int(n_levels), _d(umat), _d(s8), None, None, None, None, None) != 0:
        raise RuntimeError("pdmk_host_tables failed")
    n_pw, p, nm, nf = int(s8[4]), int(s8[5]), int(s8[6]), int(s8[7])
    w0, rad, p2w, p2c, far = np.zeros(n_levels + 1), np.zeros(nm), np.zeros(2 * n_pw * p), np.zeros(2 * p * p), \
        np.zeros(4 * nf)
    L.pdmk_host_tables(int(n_digits), float(rc), int(n_levels), _d(umat), _d(s8), _d(w0), _d(rad), _d(p2w), _d(p2c),
                       _d(far))
    return {"beta": s8[0], "psi0": s8[1], "c0": s8[2], "g0d2": s8[3], "n_pw": n_pw, "p": p, "n_modes": nm,
            "w0": w0, "radial0": rad, "poly2pw": (p2w[0::2] + 1j * p2w[1::2]).reshape(p, n_pw).T,
            "p2c": p2c.reshape(2, p, p).transpose(0, 2, 1), "far": far.reshape(nf, 4)}


def pswf(c, x):
    return lib().pdmk_host_pswf(float(c), float(x))
