"""Run selected stages of evaluate() on the benchmark workload, for per-kernel profiles of a
stage on its own (no other stream competing for the GPU):

    python -m periodicdmk_amd.stages --paths mixed6,fp64_9 --stages 1 --calls 3
    python -m periodicdmk_amd.stages --paths fp64_6 --stages 15 --calls 4 --bloch 2,-1,0.5
    python -m periodicdmk_amd.stages --workload clustered --n 4000000 --paths mixed6,fp64_9,fp64_12 --stages 15
    python -m periodicdmk_amd.stages --workload clustered --n 4000000 --paths fp64_9 --stages 15 --autotune

Stage bits: 1 proxy / plane-wave chain, 2 near field, 4 self correction, 8 far field.
--stages 15 calls evaluate() itself (device pointers; far field, near field and chain side
by side); other selections run the selected stages one after another.  With --bloch the
trees are quasi-periodic and the calls are evaluate_complex (device pointers).  Prints the
mean wall time per call after the first (synchronized).
Workloads (n_per_leaf 300, nz 4):
  uniform   -- cubic L=1, N uniform positions (splitmix64 seed 11), neutral U[-1,1] charges
               (seed 100): the same points as bench.py;
  clustered -- cell columns (1,0,0), (0.3,1,0), (0.2,0.25,1), N points in 64 Gaussian
               clusters (standard deviation 0.02, seed 21), +/-1 charges (seed 22).
"""
import argparse
import sys
import time

import periodicdmk_amd as P
from periodicdmk_amd import workloads as W

PATHS = {"mixed6": ("float", 6), "fp64_6": ("double", 6), "fp64_9": ("double", 9), "fp64_12": ("double", 12)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--paths", default="mixed6,fp64_9")
    ap.add_argument("--stages", type=int, default=1)
    ap.add_argument("--calls", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--bloch", default="", help="bx,by,bz: quasi-periodic trees")
    ap.add_argument("--workload", default="uniform", choices=("uniform", "clustered"))
    ap.add_argument("--autotune", action="store_true", help="run the launch-shape autotuner on each tree first")
    args = ap.parse_args()
    if P.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "uniform":
        cell = W.cubic_cell()
        pos = W.uniform_positions(args.n, 11, cell)
        q = P.DeviceArray.from_numpy(W.neutral_uniform_charges(args.n, 100))
    else:
        cell = W.triclinic_cell()
        pos = W.gaussian_mixture(args.n, 64, 0.02, 21, cell)
        q = P.DeviceArray.from_numpy(W.pm1_charges(args.n, 22))
    for name in [p for p in args.paths.split(",") if p]:
        prec, nd = PATHS[name]
        bloch = [float(v) for v in args.bloch.split(",")] if args.bloch else None
        s = P.PDMKSolver(cell, pos, n_digits=nd, n_per_leaf=300, nz=4, precision=prec, bloch=bloch)
        out = P.DeviceArray(2 * args.n if bloch else args.n, s.dtype)
        if args.autotune:
            r = s.autotune(2)
            print(f"[stages] {name}: autotune {r['ms_default']:.3f} -> {r['ms_tuned']:.3f} ms, knobs {r['knobs']}",
                  flush=True)
        times = []
        for _ in range(args.calls):
            t0 = time.perf_counter()
            if bloch:
                if P.lib().pdmk_gpu_evaluate_complex(s._h, P.C.c_void_p(q.ptr), P.C.c_void_p(out.ptr), args.stages) != 0:
                    raise RuntimeError("pdmk_gpu_evaluate_complex failed")
            elif args.stages == 15:
                s.evaluate_device(q.ptr, out.ptr)  # evaluate() itself, its stages side by side
            else:
                s.evaluate_stages(q.ptr, out.ptr, args.stages)  # the selected stages one after another
            P.synchronize()
            times.append(time.perf_counter() - t0)
        ms = 1e3 * sum(times[1:]) / max(1, len(times) - 1)
        kind = "evaluate_complex" if bloch else ("evaluate" if args.stages == 15 else "evaluate_stages")
        info = s.build_info()
        print(f"[stages] {name}: {args.calls} calls of {kind} stages {args.stages}, {ms:.3f} ms per call after the first"
              f" (build {info['build_ms']:.0f} ms, {info['proxy_slots']} proxy slots, plane waves by level: "
              f"{info['pw_by_level']})", flush=True)
        out.free()
        s.close()
    q.free()


if __name__ == "__main__":
    main()
