"""Run selected stages of evaluate() on the benchmark workload, for per-kernel profiles of a
stage on its own (no other stream competing for the GPU):

    python -m periodicdmk_amd.stages --paths mixed6,fp64_9 --stages 1 --calls 3

Stage bits: 1 proxy / plane-wave chain, 2 near field, 4 self correction, 8 far field.
Workload: cubic L=1, N uniform positions (splitmix64 seed 11), neutral U[-1,1] charges
(seed 100), n_per_leaf 300, nz 4 -- the same points as bench.py.
"""
import argparse
import sys

import periodicdmk_amd as P
from periodicdmk_amd import workloads as W

PATHS = {"mixed6": ("float", 6), "fp64_6": ("double", 6), "fp64_9": ("double", 9)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--paths", default="mixed6,fp64_9")
    ap.add_argument("--stages", type=int, default=1)
    ap.add_argument("--calls", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    args = ap.parse_args()
    if P.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    cell = W.cubic_cell()
    pos = W.uniform_positions(args.n, 11, cell)
    q = P.DeviceArray.from_numpy(W.neutral_uniform_charges(args.n, 100))
    for name in [p for p in args.paths.split(",") if p]:
        prec, nd = PATHS[name]
        s = P.PDMKSolver(cell, pos, n_digits=nd, n_per_leaf=300, nz=4, precision=prec)
        out = P.DeviceArray(args.n, s.dtype)
        for _ in range(args.calls):
            s.evaluate_stages(q.ptr, out.ptr, args.stages)
        P.synchronize()
        print(f"[stages] {name}: {args.calls} calls of stages {args.stages}", flush=True)
        out.free()
        s.close()
    q.free()


if __name__ == "__main__":
    main()
