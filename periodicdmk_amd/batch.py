"""The batched workload: many small trees, several charge vectors each, evaluated eagerly
(every call enqueued on its tree's own stream) and as HIP graph replays
(pdmk_gpu_graph_create):

    python -m periodicdmk_amd.batch --cells 256 --vectors 32 --per-graph 32 --branches 8

Cells and point counts from workloads.batch_cells (cubic or triclinic, N log-uniform in
[2e3, 2e5]); 6 digits in FP32, n_per_leaf 300, nz 4; charges resident in device memory.
Many trees run side by side only on as many hardware queues as the process has
(GPU_MAX_HW_QUEUES, 4 by default): measured on 256 cells x 8 vectors, eager calls reach
97 Mpts/s with 4 queues and 172 with 16, graphs 98 and 110; with --single-stream (each
tree's stages one after another on its own stream) eager calls reach 141 with 4 queues.
Each graph holds every tree's calls for --per-graph of the vectors; a pass launches
vectors / per-graph graphs back to back.  A third pass evaluates all of a tree's vectors in one
call (pdmk_gpu_evaluate_many: the near field of 16 vectors per pass).  The graph's and the
many-vector potentials are checked bit for bit against the eager pass.  Prints one JSON line:
points x vectors per second for each.
"""
import argparse
import json
import sys
import time

import numpy as np

import periodicdmk_amd as P
from periodicdmk_amd import workloads as W


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--vectors", type=int, default=32)
    ap.add_argument("--per-graph", type=int, default=32)
    ap.add_argument("--branches", type=int, default=8)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--n-max", type=int, default=200_000)
    ap.add_argument("--single-stream", action="store_true",
                    help="each tree's evaluate() on its own stream only (pdmk_gpu_set_single_stream)")
    args = ap.parse_args()
    if P.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.vectors % args.per_graph:
        sys.exit("--vectors must be a multiple of --per-graph")
    cells = W.batch_cells(args.cells, n_max=args.n_max)
    t0 = time.perf_counter()
    trees = []
    for i, (cell, n) in enumerate(cells):
        s = P.PDMKSolver(cell, W.uniform_positions(n, 1000 + i, cell), n_digits=6, n_per_leaf=300, nz=4,
                         precision="float")
        if args.single_stream:
            s.set_single_stream(True)
        qs = [P.DeviceArray.from_numpy(W.neutral_uniform_charges(n, 100000 + 32 * i + v)) for v in range(args.vectors)]
        outs = [P.DeviceArray(n, np.float32) for _ in range(args.vectors)]
        trees.append((s, qs, outs))
        if (i + 1) % 32 == 0:
            print(f"[batch] {i + 1} trees built, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    build_s = time.perf_counter() - t0
    total = sum(n for _, n in cells) * args.vectors

    def eager_pass():
        for s, qs, outs in trees:
            for q, o in zip(qs, outs):
                s.evaluate_device(q.ptr, o.ptr)
        P.synchronize()

    eager_pass()
    ref = [[o.numpy() for o in outs] for _, _, outs in trees]
    t_eager = []
    for _ in range(args.reps):
        t0 = time.perf_counter()
        eager_pass()
        t_eager.append(time.perf_counter() - t0)
    print(f"[batch] eager pass {1e3 * min(t_eager):.2f} ms", file=sys.stderr, flush=True)

    # every vector of a tree in one call (pdmk_gpu_evaluate_many): contiguous [V][N] buffers
    many = []
    for i, (s, qs, outs) in enumerate(trees):
        n = s.n
        qall = P.DeviceArray.from_numpy(
            np.concatenate([W.neutral_uniform_charges(n, 100000 + 32 * i + v) for v in range(args.vectors)]))
        many.append((s, qall, P.DeviceArray(n * args.vectors, np.float32)))

    def many_pass():
        for s, qall, oall in many:
            s.evaluate_many_device(args.vectors, qall.ptr, oall.ptr)
        P.synchronize()

    many_pass()
    many_same = all(np.array_equal(oall.numpy().reshape(args.vectors, -1)[v], ref[t][v])
                    for t, (_, _, oall) in enumerate(many) for v in range(args.vectors))
    t_many = []
    for _ in range(args.reps):
        t0 = time.perf_counter()
        many_pass()
        t_many.append(time.perf_counter() - t0)
    print(f"[batch] evaluate_many pass {1e3 * min(t_many):.2f} ms (equal to eager: {many_same})", file=sys.stderr,
          flush=True)
    for _, qall, oall in many:
        qall.free()
        oall.free()

    t0 = time.perf_counter()
    graphs = []
    for v0 in range(0, args.vectors, args.per_graph):
        calls = [(s, qs[v], outs[v]) for s, qs, outs in trees for v in range(v0, v0 + args.per_graph)]
        graphs.append(P.BatchGraph(calls, n_branches=args.branches))
        print(f"[batch] graph of {len(calls)} calls captured and instantiated, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    graph_build_s = time.perf_counter() - t0
    for _, _, outs in trees:
        for o in outs:
            o.upload(np.zeros(o.n, dtype=o.dtype))
    for g in graphs:
        g.launch()
    P.synchronize()
    same = all(np.array_equal(o.numpy(), ref[t][v]) for t, (_, _, outs) in enumerate(trees) for v, o in enumerate(outs))
    t_graph = []
    for _ in range(args.reps):
        t0 = time.perf_counter()
        for g in graphs:
            g.launch()
        P.synchronize()
        t_graph.append(time.perf_counter() - t0)
    print(f"[batch] graph pass {1e3 * min(t_graph):.2f} ms", file=sys.stderr, flush=True)
    line = {
        "workload": f"{args.cells} cells (workloads.batch_cells: cubic / triclinic, N log-uniform in "
                    f"[2e3, {args.n_max:.0e}]), {args.vectors} neutral U[-1,1] charge vectors per tree, 6 digits FP32, "
                    f"n_per_leaf 300, nz 4",
        "points_x_vectors": total,
        "eager_ms": round(1e3 * min(t_eager), 3),
        "graph_ms": round(1e3 * min(t_graph), 3),
        "eager_mpts_per_s": round(total / min(t_eager) / 1e6, 2),
        "graph_mpts_per_s": round(total / min(t_graph) / 1e6, 2),
        "many_ms": round(1e3 * min(t_many), 3),
        "many_mpts_per_s": round(total / min(t_many) / 1e6, 2),
        "many_equals_eager": bool(many_same),
        "graphs": len(graphs),
        "branches": args.branches,
        "single_stream": bool(args.single_stream),
        "graph_equals_eager": bool(same),
        "tree_build_s": round(build_s, 1),
        "graph_build_s": round(graph_build_s, 1),
    }
    print(json.dumps(line), flush=True)
    for g in graphs:
        g.close()
    for s, qs, outs in trees:
        for a in qs + outs:
            a.free()
        s.close()
    if not (same and many_same):
        sys.exit(1)


if __name__ == "__main__":
    main()
