#!/usr/bin/env python3
"""Benchmark: evaluate() throughput of the MI355X periodic DMK solver.

Workload (the reference metric's configuration): cubic cell L = 1, N = 1,000,000
positions uniform in the cell (splitmix64 seed 11), neutral U[-1,1] charges (seeds
100..107, one fresh vector per step), n_per_leaf = 300, nz = 4.  The tree is built once
(not timed); each step is one evaluate() with charges and potentials resident in HBM.

Four paths are timed in one run:
  * mixed6_tuned -- n_digits = 6, FP32 kernels, FP64 far field and final accumulation, with
    n_per_leaf and nz chosen by the learned cost model (the defaults when it has no model) and
    the launch shapes by the autotuner (the value);
  * mixed6  -- the same at n_per_leaf = 300, nz = 4 and the built-in launch shapes;
  * fp64_6  -- n_digits = 6, plain FP64 path;
  * fp64_9  -- n_digits = 9, plain FP64 path.
The CPU baseline is the CPU oracle (a C++/OpenMP restatement of the reference
evaluate()) timed on this host on the same workload.  Errors are reported against that
oracle (all N targets) and against the FP64 Ewald sum at 1000 sampled targets; every path
is gated (accuracy_gates) and the run exits non-zero when one misses.

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
CPU_CALLS = 8


def accuracy_gates(name, n_digits):
    """Relative-L2 limits of a path: against the FP64 Ewald sum, 1e-6 for the mixed 6-digit
    path (the target of the mixed mode) and otherwise the reference's own end-to-end
    tolerances (1e-4 up to 6 digits, 1e-7 at 9: cpp/tests/test_nonuniform.cpp:282-304,
    10^(2-d) above); against the CPU oracle, the requested 10^-d."""
    ewald = 1e-6 if name.startswith("mixed") else (1e-4 if n_digits <= 6 else (1e-7 if n_digits <= 9 else 10.0 ** (2 - n_digits)))
    return {"ewald": ewald, "oracle": 10.0 ** (-n_digits)}


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


def _run_path(P, cell, pos, n, precision, n_digits, steps, warmup, npl, nz, charges_dev, n_vec, q_host, host_calls=3,
              tune=False):
    tuning = None
    if tune:
        # the learned cost model's n_per_leaf and nz for this problem (FP32 kernels: the mixed
        # mode), then the launch-shape autotuner on the built tree
        from periodicdmk_amd import costmodel
        t0 = time.time()
        ch = costmodel.choose(cell, pos, n_digits, precision=precision)
        npl, nz = ch["n_per_leaf"], ch["nz"]
        tuning = {"n_per_leaf": npl, "nz": nz, "cost_model": ch["model"], "choose_s": round(time.time() - t0, 3)}
        if ch["model"]:
            tuning.update({k: ch[k] for k in ("pred_ms", "pred_default_ms", "pred_err", "pred_default_err")})
    t0 = time.time()
    s = P.PDMKSolver(cell, pos, n_digits=n_digits, n_per_leaf=npl, nz=nz, precision=precision)
    build_s = time.time() - t0
    if tune:
        tuning["autotune"] = s.autotune(repeats=3)
        # the other paths keep the built-in launch shapes
        P.lib().pdmk_gpu_tune_cache_clear()
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
    # the drop-in rate: host charges in, host potentials out (pdmk_evaluate_*, the reference's
    # C ABI), PCIe copies included
    host_ms = None
    if host_calls:
        s.evaluate(q_host)
        t0 = time.perf_counter()
        for _ in range(host_calls):
            s.evaluate(q_host)
        host_ms = 1000.0 * (time.perf_counter() - t0) / host_calls
    out.free()
    res = {
        "precision": precision, "n_digits": n_digits, "ms_per_step": ms, "mpts_per_s": n / (ms / 1000.0) / 1e6,
        "build_s": build_s, "build": s.build_info(), "kernel_ms": ktimes, "near_field_alone_ms": near_alone_ms,
        "stats": s.stats(), "work": s.work(), "nf_steps": s.stats()["n_nf_steps"],
        "last_charge_index": last_q,
        "host_buffer_ms": host_ms, "host_buffer_mpts_per_s": (n / (host_ms / 1000.0) / 1e6) if host_ms else None,
        "n_per_leaf": npl, "nz": nz, "tuning": tuning, "knobs": s.knobs(),
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
    ap.add_argument("--paths", default="mixed6_tuned,mixed6,fp64_6,fp64_9",
                    help="comma list of mixed6_tuned, mixed6, fp64_6, fp64_9")
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

    paths = {"mixed6_tuned": ("float", 6), "mixed6": ("float", 6), "fp64_6": ("double", 6), "fp64_9": ("double", 9)}
    wanted = [p for p in args.paths.split(",") if p]
    results, pots = {}, {}
    for name in wanted:
        prec, nd = paths[name]
        print(f"[bench] {name}: building and timing", file=sys.stderr, flush=True)
        results[name], pots[name] = _run_path(P, cell, pos, n, prec, nd, args.steps, args.warmup, args.npl, args.nz,
                                              charges_dev, qs.shape[0], qs[0], tune=name.endswith("_tuned"))
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

    # FP64 Ewald sum at EWALD_TARGETS sampled targets on the GPU (pdmk_gpu_ewald_targets,
    # checked against the CPU oracle's Ewald sum by the tests): the accuracy reference of
    # every path, computed in every run
    tg = np.sort(np.random.default_rng(5).choice(n, min(n, EWALD_TARGETS), replace=False)).astype(np.int32)
    qi = (args.steps - 1) % 8
    t0 = time.perf_counter()
    ew = P.ewald_targets(cell, pos, qs[qi], tg, s=5.5)
    ew_s = time.perf_counter() - t0
    for nm in wanted:
        d = pots[nm][tg] - ew
        results[nm]["err_vs_ewald_rel_l2"] = float(np.linalg.norm(d) / np.linalg.norm(ew))
        results[nm]["err_vs_ewald_max_rel"] = float(np.max(np.abs(d)) / np.max(np.abs(ew)))
    ewald_info = {"targets": int(len(tg)), "seconds": round(ew_s, 2), "s": 5.5, "where": "GPU FP64"}

    # CPU oracle on the same workload: timed baseline (built -march=native for this host; one
    # warm-up, then the median of CPU_CALLS evaluate() calls, with a per-stage split) and the
    # all-N error reference
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.use_native()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ref = {}
        for nd in sorted({paths[nm][1] for nm in wanted}):
            print(f"[bench] CPU oracle, n_digits={nd}", file=sys.stderr, flush=True)
            t0 = time.time()
            o = O.OracleTree(cell, pos, nd, args.npl, args.nz, precision="double")
            tb = time.time() - t0
            ref[nd] = o.evaluate(qs[qi])  # also the warm-up
            if nd == 6:
                times = []
                for k in range(CPU_CALLS):
                    t0 = time.perf_counter()
                    o.evaluate(qs[k % 8])
                    times.append(time.perf_counter() - t0)
                te = float(np.median(times))
                split = {}
                for name, bit in (("proxy_planewave", 1), ("near_field", 2), ("self_correction", 4), ("far_field", 8)):
                    t0 = time.perf_counter()
                    o.evaluate(qs[qi], stages=bit)
                    split[name] = round(time.perf_counter() - t0, 3)
                cpu = {"value": n / te / 1e6, "unit": "Mpts/s", "cores": threads, "kind": "port",
                       "sample": f"full workload (N={n}, n_digits=6, FP64): median of {CPU_CALLS} evaluate() calls "
                                 f"after one warm-up, on a tree built in {tb:.1f} s; CPU oracle (C++/OpenMP "
                                 f"restatement of the reference evaluate(), -O3 -march=native on this host, exact "
                                 f"direct-DFT far field)",
                       "seconds": te, "seconds_min": float(np.min(times)), "stage_seconds": split}
                # the same evaluate() on the same tree on one core
                print("[bench] CPU oracle on one thread", file=sys.stderr, flush=True)
                prev = O.set_threads(1)
                t0 = time.perf_counter()
                o.evaluate(qs[qi])
                t1 = time.perf_counter() - t0
                O.set_threads(prev)
                cpu["single_core"] = {"value": n / t1 / 1e6, "unit": "Mpts/s", "cores": 1, "seconds": t1,
                                      "sample": "one evaluate() of the same workload"}
            o.close()
        for nm in wanted:
            r = ref[paths[nm][1]]
            d = pots[nm] - r
            results[nm]["err_vs_cpu_oracle_rel_l2"] = float(np.linalg.norm(d) / np.linalg.norm(r))
            results[nm]["err_vs_cpu_oracle_max_rel"] = float(np.max(np.abs(d)) / np.max(np.abs(r)))
            k = int(np.argmax(np.abs(d)))
            results[nm]["err_vs_cpu_oracle_worst"] = {"index": k, "abs": float(abs(d[k])), "pot": float(r[k])}
        # the CPU oracle's Ewald sum at the same targets (the GPU reference's own check)
        print(f"[bench] CPU Ewald oracle at {EWALD_TARGETS} sampled targets", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        ew_cpu = O.ewald_cells(cell, pos, qs[qi], s=5.5, targets=tg, alpha=1e-9)
        ewald_info["cpu_seconds"] = round(time.perf_counter() - t0, 1)
        ewald_info["gpu_vs_cpu_rel_l2"] = float(np.linalg.norm(ew - ew_cpu) / np.linalg.norm(ew_cpu))

    head_name = next((k for k in ("mixed6_tuned", "mixed6") if k in results), next(iter(results)))
    head = results[head_name]
    w = head["work"]
    nf_ms = head["kernel_ms"]["near_field"]
    peak = PEAK_FP32_TFLOPS if head["precision"] == "float" else PEAK_FP64_TFLOPS
    achieved = w["nf_flops"] / (nf_ms / 1000.0) / 1e12 if nf_ms > 0 else 0.0
    # HBM bytes per near-field launch from the FETCH_SIZE / WRITE_SIZE counter passes recorded
    # in profiles/ (counters need their own profiler runs)
    traffic, traffic_src = None, None
    tpath = os.path.join(REPO, "profiles", "near_field_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        traffic, traffic_src = tr.get("bytes_per_launch"), f"{tr.get('source')} ({tr.get('kernel')})"
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
                               f"value = n_digits=6 mixed path ({head_name}: n_per_leaf "
                               f"{head['n_per_leaf']}, nz {head['nz']})",
                   "n_points": n, "n_digits": head["n_digits"], "n_per_leaf": head["n_per_leaf"], "nz": head["nz"]},
        "roofline": {"bound": "valu", "engine": "VALU (near-field residual kernel, vector FP32)" if
                     head["precision"] == "float" else "VALU (near-field residual kernel, vector FP64)",
                     "kernel": (("k_near_rows" if head["precision"] == "float" else "k_near_rows64")
                                if head["build"].get("near_rows") else "k_near_field"), "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic, "traffic_source": traffic_src,
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
    line["path_roofline"] = results[head_name]["path_roofline"]
    line["paths"] = {k: {kk: vv for kk, vv in v.items() if kk not in ("stats",)} for k, v in results.items()}
    for nm in ("mixed6", "mixed6_tuned"):
        if nm in results and "fp64_6" in results:
            line[f"{nm}_vs_fp64_6"] = round(results["fp64_6"]["ms_per_step"] / results[nm]["ms_per_step"], 3)
    # a throughput of wrong potentials is no measurement: every path is gated against the FP64
    # Ewald sum and, when it ran, the CPU oracle, and the run exits non-zero on a failure
    gates, bad = {}, []
    for nm in wanted:
        g = accuracy_gates(nm, paths[nm][1])
        got = {"ewald": results[nm]["err_vs_ewald_rel_l2"], "oracle": results[nm].get("err_vs_cpu_oracle_rel_l2")}
        if (results[nm]["n_per_leaf"], results[nm]["nz"]) != (args.npl, args.nz):
            got["oracle"] = None  # another tree: its own DMK approximation, gated against Ewald only
        gates[nm] = {k: {"rel_l2": got[k], "limit": g[k]} for k in g if got[k] is not None}
        if any(v["rel_l2"] > v["limit"] for v in gates[nm].values()):
            bad.append(nm)
    line["ewald_reference"] = ewald_info
    line["accuracy_gates"] = gates
    line["accuracy_ok"] = not bad
    print(json.dumps(line))
    if bad:
        print(f"[bench] ERROR: potentials of {bad} miss their accuracy gates {gates}", file=sys.stderr, flush=True)
        sys.exit(1)


if __name__ == "__main__":
    main()
