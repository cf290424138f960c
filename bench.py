#!/usr/bin/env python3
"""Benchmark: evaluate() throughput of the MI355X periodic DMK solver.

Workload (the reference metric's configuration): cubic cell L = 1, N = 1,000,000
positions uniform in the cell (splitmix64 seed 11), neutral U[-1,1] charges (seeds
100..107, one fresh vector per step), n_per_leaf = 300, nz = 4.  The tree is built once
(not timed); each step is one evaluate() with charges and potentials resident in HBM.

Three paths are timed in one run:
  * mixed6  -- n_digits = 6, FP32 kernels, FP64 far field and final accumulation (the value);
  * fp64_6  -- n_digits = 6, plain FP64 path;
  * fp64_9  -- n_digits = 9, plain FP64 path.
The CPU baseline is the CPU oracle (a C++/OpenMP restatement of the reference
evaluate()) timed on this host on the same workload.  Errors are reported against that
oracle (all N targets) and against the FP64 Ewald sum at 1000 sampled targets.

Prints ONE JSON line.  ``--gpus`` above 1 is not supported (single-GPU build).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "evaluate() Mpts/s @N=1e6, 6 & 9 digits: tuned mixed vs plain FP64 GPU"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
EWALD_TARGETS = 1000


def path_roofline(res):
    """Composite roofline of one evaluate(): t_roof = F_main / P_main + F_far / P_fp64 + B / BW
    over the plan's exact algorithmic counts (near-field flops of in-support pairs, transform
    flops, far-field flops, O(N) and near-field bytes), with the published peaks (vector or
    matrix FP32 157.3, FP64 78.6 TFLOP/s; HBM 8 TB/s); frac = t_roof / measured time."""
    w = res["work"]
    main_peak = PEAK_FP32_TFLOPS if res["precision"] == "float" else PEAK_FP64_TFLOPS
    f_main = w["nf_flops"] + w["transform_flops"]
    nbytes = w["on_bytes"] + w["nf_bytes"]
    t_ms = 1e3 * (f_main / (main_peak * 1e12) + w["ff_flops"] / (PEAK_FP64_TFLOPS * 1e12) +
                  nbytes / (PEAK_HBM_GBS * 1e9))
    return {"t_roof_ms": round(t_ms, 4), "ms_measured": round(res["ms_per_step"], 4),
            "frac": round(t_ms / res["ms_per_step"], 4), "flops_main": f_main, "main_peak_tflops": main_peak,
            "flops_far_fp64": w["ff_flops"], "bytes": nbytes}


def _run_path(P, cell, pos, n, precision, n_digits, steps, warmup, npl, nz, charges_dev, n_vec):
    t0 = time.time()
    s = P.PDMKSolver(cell, pos, n_digits=n_digits, n_per_leaf=npl, nz=nz, precision=precision)
    build_s = time.time() - t0
    out = P.DeviceArray(n, s.dtype)
    for k in range(warmup):
        s.evaluate_device(charges_dev.offset((k % n_vec) * n), out.ptr, 0)
    s.enable_kernel_timing(max(1, steps))
    P.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        s.evaluate_device(charges_dev.offset((k % n_vec) * n), out.ptr, 0)
    P.synchronize()
    wall = time.perf_counter() - t0
    ktimes, ncalls = s.kernel_times()
    ms = 1000.0 * wall / steps
    # potentials of the last step, for the error check against the CPU oracle (read before
    # the pass below overwrites them)
    last_q = (steps - 1) % n_vec
    pot = out.numpy().astype("float64")
    # the near-field kernel alone (stage bit 2 only), for its duration without the other
    # streams of evaluate() competing for the CUs
    s.enable_kernel_timing(max(1, steps))
    for k in range(steps):
        s.evaluate_stages(charges_dev.offset((k % n_vec) * n), out.ptr, 2)
    P.synchronize()
    near_alone_ms = s.kernel_times()[0]["near_field"]
    out.free()
    res = {
        "precision": precision, "n_digits": n_digits, "ms_per_step": ms, "mpts_per_s": n / (ms / 1000.0) / 1e6,
        "build_s": build_s, "build": s.build_info(), "kernel_ms": ktimes, "near_field_alone_ms": near_alone_ms,
        "stats": s.stats(), "work": s.work(), "nf_steps": s.stats()["n_nf_steps"],
        "last_charge_index": last_q,
    }
    s.close()
    return res, pot


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--npl", type=int, default=300)
    ap.add_argument("--nz", type=int, default=4)
    ap.add_argument("--paths", default="mixed6,fp64_6,fp64_9", help="comma list of mixed6, fp64_6, fp64_9")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="only run 4 device-to-device copies of MB MiB (counter calibration for the traffic figure)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this solver runs on one GPU; --gpus must be 1 (multi-GPU sharding is not implemented)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import periodicdmk_amd as P
    from periodicdmk_amd import workloads as W

    if P.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        a = P.DeviceArray(nbytes // 8, np.float64)
        b = P.DeviceArray(nbytes // 8, np.float64)
        a.upload(np.ones(nbytes // 8))
        for _ in range(4):
            P.memcpy_d2d(b.ptr, a.ptr, nbytes)
        P.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 4}))
        return

    n = args.n
    cell = W.cubic_cell()
    pos = W.uniform_positions(n, 11, cell)
    qs = np.stack([W.neutral_uniform_charges(n, 100 + k) for k in range(8)])
    charges_dev = P.DeviceArray.from_numpy(qs.reshape(-1))  # 8 vectors, resident in HBM

    paths = {"mixed6": ("float", 6), "fp64_6": ("double", 6), "fp64_9": ("double", 9)}
    wanted = [p for p in args.paths.split(",") if p]
    results, pots = {}, {}
    for name in wanted:
        prec, nd = paths[name]
        print(f"[bench] {name}: building and timing", file=sys.stderr, flush=True)
        results[name], pots[name] = _run_path(P, cell, pos, n, prec, nd, args.steps, args.warmup, args.npl, args.nz,
                                              charges_dev, qs.shape[0])
        print(f"[bench] {name}: {results[name]['ms_per_step']:.3f} ms/step "
              f"({results[name]['mpts_per_s']:.1f} Mpts/s) kernels {results[name]['kernel_ms']}",
              file=sys.stderr, flush=True)

    # tree construction: per-point work on the GPU vs the same construction on the host
    tree_build = None
    if wanted:
        t0 = time.time()
        P.host_tree(cell, pos, args.npl, args.nz)
        host_ms = (time.time() - t0) * 1e3
        b = results[wanted[0]]["build"]
        tree_build = {"tree_ms_gpu_points": round(b["tree_ms"], 1), "tree_ms_host_points": round(host_ms, 1),
                      "tree_on_gpu": b["tree_on_gpu"], "build_ms_total": round(b["build_ms"], 1),
                      "far_plan_ms": round(b["far_plan_ms"], 1)}
        print(f"[bench] tree build {tree_build}", file=sys.stderr, flush=True)

    # CPU oracle on the same workload: timed baseline and error reference
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ref = {}
        for nd in sorted({paths[nm][1] for nm in wanted}):
            print(f"[bench] CPU oracle, n_digits={nd}", file=sys.stderr, flush=True)
            t0 = time.time()
            o = O.OracleTree(cell, pos, nd, args.npl, args.nz, precision="double")
            tb = time.time() - t0
            qi = (args.steps - 1) % 8
            t0 = time.perf_counter()
            ref[nd] = o.evaluate(qs[qi])
            te = time.perf_counter() - t0
            if nd == 6:
                cpu = {"value": n / te / 1e6, "unit": "Mpts/s", "cores": threads, "kind": "port",
                       "sample": f"full workload (N={n}, n_digits=6, FP64), one evaluate() after a "
                                 f"{tb:.1f} s tree build; CPU oracle restating the reference evaluate() "
                                 f"with an exact direct-DFT far field", "seconds": te}
                # the same evaluate() on the same tree on one core
                print("[bench] CPU oracle on one thread", file=sys.stderr, flush=True)
                prev = O.set_threads(1)
                t0 = time.perf_counter()
                o.evaluate(qs[qi])
                t1 = time.perf_counter() - t0
                O.set_threads(prev)
                cpu["single_core"] = {"value": n / t1 / 1e6, "unit": "Mpts/s", "cores": 1, "seconds": t1}
            o.close()
        for nm in wanted:
            r = ref[paths[nm][1]]
            d = pots[nm] - r
            results[nm]["err_vs_cpu_oracle_rel_l2"] = float(np.linalg.norm(d) / np.linalg.norm(r))
            results[nm]["err_vs_cpu_oracle_max_rel"] = float(np.max(np.abs(d)) / np.max(np.abs(r)))
        # FP64 Ewald sum at EWALD_TARGETS sampled targets (splitting parameter at its lower
        # bound, which keeps the reciprocal sum small; s = 5.5 truncates below 1e-12)
        print(f"[bench] CPU Ewald oracle at {EWALD_TARGETS} sampled targets", file=sys.stderr, flush=True)
        tg = np.sort(np.random.default_rng(5).choice(n, min(n, EWALD_TARGETS), replace=False)).astype(np.int32)
        qi = (args.steps - 1) % 8
        t0 = time.perf_counter()
        ew = O.ewald_cells(cell, pos, qs[qi], s=5.5, targets=tg, alpha=1e-9)
        ew_s = time.perf_counter() - t0
        for nm in wanted:
            d = pots[nm][tg] - ew
            results[nm]["err_vs_ewald_rel_l2"] = float(np.linalg.norm(d) / np.linalg.norm(ew))
            results[nm]["err_vs_ewald_max_rel"] = float(np.max(np.abs(d)) / np.max(np.abs(ew)))
        if cpu is not None:
            cpu["ewald_reference"] = f"FP64 Ewald (s=5.5) at {len(tg)} sampled targets, {ew_s:.1f} s"

    head = results.get("mixed6") or next(iter(results.values()))
    w = head["work"]
    nf_ms = head["kernel_ms"]["near_field"]
    peak = PEAK_FP32_TFLOPS if head["precision"] == "float" else PEAK_FP64_TFLOPS
    achieved = w["nf_flops"] / (nf_ms / 1000.0) / 1e12 if nf_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(REPO, "profiles", "near_field_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": round(head["mpts_per_s"], 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64" if head["precision"] == "float" else "f64",
        "data": "synthetic",
        "config": {"workload": f"cubic L=1, N={n} uniform positions (splitmix64 seed 11), neutral U[-1,1] charges "
                               f"(seeds 100-107, fresh vector per step), n_per_leaf={args.npl}, nz={args.nz}; "
                               f"value = n_digits=6 mixed path",
                   "n_points": n, "n_digits": head["n_digits"], "n_per_leaf": args.npl, "nz": args.nz},
        "roofline": {"bound": "mfma", "engine": "VALU (near-field residual kernel, vector FP32)" if
                     head["precision"] == "float" else "VALU (near-field residual kernel, vector FP64)",
                     "kernel": "k_near_field", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "flops_per_launch": w["nf_flops"], "pairs_in_support": w["nf_pairs_support"],
                     "pairs_candidate": w["nf_pairs_candidate"], "kernel_ms": round(nf_ms, 4),
                     "kernel_ms_alone": round(head["near_field_alone_ms"], 4),
                     "achieved_alone": round(w["nf_flops"] / (head["near_field_alone_ms"] / 1000.0) / 1e12, 3),
                     "note": "kernel_ms is the near-field launch inside evaluate(), which runs it concurrently with "
                             "the transform chain and the far field; kernel_ms_alone is the same kernel run alone"},
        "cpu_baseline": cpu,
        "tree_build": tree_build,
        "paths": {k: {kk: vv for kk, vv in v.items() if kk not in ("stats",)} for k, v in results.items()},
    }
    for nm in results:
        results[nm]["path_roofline"] = path_roofline(results[nm])
    line["path_roofline"] = results[next(iter(results)) if "mixed6" not in results else "mixed6"]["path_roofline"]
    line["paths"] = {k: {kk: vv for kk, vv in v.items() if kk not in ("stats",)} for k, v in results.items()}
    if "mixed6" in results and "fp64_6" in results:
        line["mixed_vs_fp64_6"] = round(results["fp64_6"]["ms_per_step"] / results["mixed6"]["ms_per_step"], 3)
    # a throughput of wrong potentials is no measurement: flag any path whose potentials are
    # off the CPU oracle by more than 100x its digit target, and exit non-zero
    bad = [nm for nm in wanted if results[nm].get("err_vs_cpu_oracle_rel_l2", 0.0) > 10.0 ** (2 - paths[nm][1])]
    line["accuracy_ok"] = None if args.no_cpu_baseline else not bad
    print(json.dumps(line))
    if bad:
        print(f"[bench] ERROR: potentials of {bad} disagree with the CPU oracle", file=sys.stderr, flush=True)
        sys.exit(1)


if __name__ == "__main__":
    main()
