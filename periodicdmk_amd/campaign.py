"""Reference-data ("PDMC") files: oracle potentials at sampled targets together with the
problem they belong to (cell, positions, charges), checked by a SHA-256 of the body.

The layout and the checks are those of the reference's ``cpp/include/pdmk/campaign_ref_io.hpp``
(80-byte header ``RefHeader``, body cell | ref_idx | ref_u | pos | chg; ``write_ref`` /
``read_ref``); files move between the two libraries.  Everything here goes through
``libpdmk.so`` (``pdmk_campaign.h``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import lib

MAGIC = 0x50444D43
VERSION = 1
ENDIAN_TAG = 0x01020304
HEADER_LEN = 80
DEFAULT_M = 1000


@dataclass
class RefData:
    """One reference-data file.  ``cell[:, j]`` is lattice vector j; ``positions`` is (N, 3)."""

    cell: np.ndarray
    positions: np.ndarray
    charges: np.ndarray
    ref_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.uint32))
    ref_u: np.ndarray = field(default_factory=lambda: np.zeros(0))
    config_id: int = 0
    digits: int = 0
    prec: int = 0
    seed: int = 0

    @property
    def N(self) -> int:
        return int(np.asarray(self.charges).size)

    @property
    def M(self) -> int:
        return int(np.asarray(self.ref_idx).size)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint))


def sha256(data: bytes) -> bytes:
    """SHA-256 digest computed by the library."""
    buf = (C.c_ubyte * max(1, len(data))).from_buffer_copy(data if data else b"\0")
    out = (C.c_ubyte * 32)()
    if lib().pdmk_campaign_sha256(buf, len(data), out) != 0:
        raise RuntimeError("pdmk_campaign_sha256 failed")
    return bytes(out)


def write_ref(path: str, d: RefData) -> None:
    """Write ``d`` to ``path`` (magic, version, lengths and digest are filled in)."""
    cell = np.ascontiguousarray(np.asarray(d.cell, dtype=np.float64).reshape(3, 3).T)
    pos = np.ascontiguousarray(np.asarray(d.positions, dtype=np.float64).reshape(-1, 3))
    chg = np.ascontiguousarray(np.asarray(d.charges, dtype=np.float64).ravel())
    idx = np.ascontiguousarray(np.asarray(d.ref_idx, dtype=np.uint32).ravel())
    u = np.ascontiguousarray(np.asarray(d.ref_u, dtype=np.float64).ravel())
    if pos.shape[0] != chg.size:
        raise ValueError("positions and charges disagree on N")
    if idx.size != u.size:
        raise ValueError("ref_idx and ref_u disagree on M")
    rc = lib().pdmk_campaign_write(str(path).encode(), int(d.config_id), int(chg.size), int(d.digits), int(d.prec),
                                   int(d.seed), _dp(cell), int(idx.size), _up(idx), _dp(u), _dp(pos), _dp(chg))
    if rc != 0:
        raise RuntimeError(f"pdmk_campaign_write failed for {path}")


def read_ref(path: str) -> RefData:
    """Read and verify ``path``; raises RuntimeError on a bad header, length or digest."""
    L = lib()
    cfg, n, m = C.c_uint(), C.c_uint(), C.c_uint()
    dg, pr = C.c_int(), C.c_int()
    seed = C.c_ulonglong()
    p = str(path).encode()
    if L.pdmk_campaign_read_header(p, C.byref(cfg), C.byref(n), C.byref(dg), C.byref(pr), C.byref(seed),
                                   C.byref(m)) != 0:
        raise RuntimeError(f"not a valid reference-data file: {path}")
    cell = np.zeros(9)
    idx = np.zeros(m.value, dtype=np.uint32)
    u = np.zeros(m.value)
    pos = np.zeros((n.value, 3))
    chg = np.zeros(n.value)
    if L.pdmk_campaign_read(p, _dp(cell), _up(idx), _dp(u), _dp(pos), _dp(chg)) != 0:
        raise RuntimeError(f"not a valid reference-data file: {path}")
    return RefData(cell=cell.reshape(3, 3).T.copy(), positions=pos, charges=chg, ref_idx=idx, ref_u=u,
                   config_id=cfg.value, digits=dg.value, prec=pr.value, seed=seed.value)
