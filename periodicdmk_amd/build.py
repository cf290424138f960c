"""Build libpdmk.so (host C++ + gfx950 HIP kernels) in-tree with hipcc.

Usage: ``python -m periodicdmk_amd.build``.  Objects go to ``periodicdmk_amd/_build``,
the library to ``periodicdmk_amd/lib/libpdmk.so``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libpdmk.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(REPO, "include"), "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable"]
# Device code: gfx950 only, no packed FP32 instructions.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

SOURCES = [
    ("host/tree_build.cpp", False),
    ("host/tables.cpp", False),
    ("host/nufft.cpp", False),
    ("host/host_capi.cpp", False),
    ("host/campaign_capi.cpp", False),
    ("hip/plan.hip", True),
    ("hip/ewald.hip", True),
    ("hip/tree_gpu.hip", True),
    ("hip/api.hip", True),
    ("hip/costmodel.hip", True),
]


def _compile(src: str, device: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    if device:
        cmd = [HIPCC] + COMMON + DEVICE + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + COMMON + ["-march=x86-64-v3", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        # hipcc's host pass does not know the packed-fp32 feature; keep other warnings visible
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
        if any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
