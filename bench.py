#!/usr/bin/env python3
"""Benchmark of the HLBVH build + self-kNN hot path on one MI355X.

Workload: the clustered production snapshot -- N = 10M particles, int64
fixed-point coordinates (21 bits/dim) in a periodic box L = 100, 70% in 2000
NFW halos + 30% uniform background, 63-bit Peano-Hilbert keys; build_bvh then
self_knn_query with k = 32, exclude_self.  A step is one build plus one
self-kNN pass over all N particles, inputs resident in HBM.

Prints ONE JSON line.  Run: python bench.py [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-ladder", action="store_true", help="skip the plain/grouped kernels")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--bucket", type=int, default=0, help="packet leaf bucket (0 = auto)")
    ap.add_argument("--sweep-buckets", action="store_true")
    ap.add_argument("--model", default=os.path.join(ROOT, "fasttree_amd", "models",
                                                    "nfw10m_int64_ph21_k32.npz"),
                    help="trained learned-radius model (npz); the learned step is skipped "
                         "when it is missing")
    ap.add_argument("--seed-probe", action="store_true",
                    help="time the packet kernel seeded with scaled exact k-th distances")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import fasttree_amd as ft
    from fasttree_amd import synth

    n, k = args.n, args.k
    t0 = time.time()
    x, y, z = synth.clustered_fixed(n, args.seed, bits=21, box=100.0, dtype=np.uint64)
    log(f"snapshot N={n} generated in {time.time() - t0:.1f}s")

    stream = ft.Stream()
    cfg = ft.Config(sfc=ft.SFC_PEANO_HILBERT, bits_per_dim=21, periodic=True)
    tree = ft.TreeSoA(n, ft.COORD_U64, cfg)
    scratch = ft.BvhScratch()
    p = ft.particles(*(ft.DeviceArray.from_host(a) for a in (x, y, z)))
    res = ft.DeviceArray(n * k, np.uint64)
    dst = ft.DeviceArray(n * k, np.uint64)
    ev = [ft.Event() for _ in range(4)]

    def step(algorithm=ft.ALGO_PACKET):
        ev[0].record(stream)
        ft.build_bvh(p, tree, scratch=scratch, stream=stream.handle)
        ev[1].record(stream)
        ft.self_knn_query(tree, k, results=res, dists=dst, exclude_self=True,
                          stream=stream.handle, algorithm=algorithm, leaf_bucket=args.bucket)
        ev[2].record(stream)
        stream.synchronize()
        return ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])

    for w in range(args.warmup):
        b, q = step()
        log(f"warmup {w}: build {b:.3f} ms, self-kNN {q:.3f} ms")

    stream.synchronize()
    builds, queries = [], []
    t_start = time.perf_counter()
    for s in range(args.steps):
        b, q = step()
        builds.append(b)
        queries.append(q)
        log(f"step {s}: build {b:.3f} ms, self-kNN {q:.3f} ms")
    stream.synchronize()
    wall_ms = (time.perf_counter() - t_start) * 1e3 / max(args.steps, 1)

    build_ms = float(np.median(builds))
    query_ms = float(np.median(queries))
    qps = n / (query_ms * 1e-3)

    # Per-stage build times (events between stages on the same stream).
    scratch.set_profiling(True)
    ft.build_bvh(p, tree, scratch=scratch, stream=stream.handle)
    stream.synchronize()
    stages = scratch.stage_ms()
    scratch.set_profiling(False)

    # Traversal statistics of the measured kernel.
    stats = ft.DeviceArray(8, np.uint64).zero()
    ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle, stats=stats,
                      algorithm=ft.ALGO_PACKET)
    stream.synchronize()
    st = stats.to_host()

    ladder = {"packet_qps": qps}
    learned = None
    if os.path.exists(args.model):
        from fasttree_amd import learned as L
        model = L.MLP.load(args.model)
        seeds = ft.DeviceArray(n, np.uint64)
        feats = ft.DeviceArray(n * 16, np.float32)
        res2 = ft.DeviceArray(n * k, np.uint64)
        dst2 = ft.DeviceArray(n * k, np.uint64)
        lst = ft.DeviceArray(8, np.uint64)
        times = []
        for rep in range(max(args.steps, 1) + 1):
            lst.zero()
            ev[0].record(stream)
            model.seed(tree, k, seeds, feats, stream=stream.handle)
            ev[1].record(stream)
            ft.self_knn_query(tree, k, results=res2, dists=dst2, stream=stream.handle,
                              seed_bound=seeds, stats=lst, leaf_bucket=args.bucket)
            ev[2].record(stream)
            stream.synchronize()
            if rep:
                times.append((ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])))
        seed_ms = float(np.median([a for a, b in times]))
        trav_ms = float(np.median([b for a, b in times]))
        ls = lst.to_host()
        rows = np.linspace(0, n - 1, 2000).astype(np.int64)
        same = all(np.array_equal(res.to_host_range(r * k, k), res2.to_host_range(r * k, k)) and
                   np.array_equal(dst.to_host_range(r * k, k), dst2.to_host_range(r * k, k))
                   for r in rows)
        learned = {"learned_ms": round(seed_ms + trav_ms, 4), "seed_ms": round(seed_ms, 4),
                   "seeded_traversal_ms": round(trav_ms, 4),
                   "rerun_fraction": round(float(ls[3]) / n, 5),
                   "identical_on_2000_sampled_rows": bool(same),
                   "model": os.path.relpath(args.model, ROOT)}
        ladder["learned_qps"] = n / ((seed_ms + trav_ms) * 1e-3)
        log(f"learned: seed {seed_ms:.2f} ms + traversal {trav_ms:.2f} ms, reruns {ls[3]}, "
            f"identical sample {same}")
        del res2, dst2, feats, seeds
    if args.sweep_buckets:
        for bk in (8, 16, 24, 32, 48, 64):
            ev[0].record(stream)
            ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle,
                              algorithm=ft.ALGO_PACKET, leaf_bucket=bk)
            ev[1].record(stream)
            stream.synchronize()
            log(f"bucket {bk}: {ev[0].elapsed_ms(ev[1]):.2f} ms")
    if args.seed_probe:
        dk = dst.to_host().reshape(n, k)[:, k - 1].astype(np.float64)
        for f in (1.0, 1.2, 1.5, 2.0, 0.9):
            sd = ft.DeviceArray.from_host(np.floor(dk * f).astype(np.uint64))
            st2 = ft.DeviceArray(8, np.uint64).zero()
            ev[0].record(stream)
            ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle,
                              seed_bound=sd, stats=st2, leaf_bucket=args.bucket)
            ev[1].record(stream)
            stream.synchronize()
            s2 = st2.to_host()
            log(f"seed x{f}: {ev[0].elapsed_ms(ev[1]):.2f} ms, nodes/wave "
                f"{s2[0] / ((n + 63) // 64):.1f}, leaves/wave {s2[1] / ((n + 63) // 64):.1f}, "
                f"inserts/q {s2[4] / n:.1f}, reruns {int(s2[3])}")
    if not args.no_ladder:
        for name, algo in (("grouped_qps", ft.ALGO_GROUPED), ("plain_qps", ft.ALGO_PLAIN)):
            ev[0].record(stream)
            ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle,
                              algorithm=algo)
            ev[1].record(stream)
            stream.synchronize()
            ms = ev[0].elapsed_ms(ev[1])
            ladder[name] = n / (ms * 1e-3)
            log(f"{name}: {ms:.1f} ms")

    # Roofline of the dominant kernel (the self-kNN traversal): algorithmic bytes
    # N*(3s + 16k) + (12s + 8)*N with s = 8 bytes per coordinate.
    s_b = 8
    knn_bytes = n * (3 * s_b + 16 * k) + (12 * s_b + 8) * n
    build_bytes = 492 * n
    achieved = knn_bytes / (query_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": None,
            "kernel": "packet_knn_kernel<uint64_t,32>", "algorithmic_bytes": knn_bytes}
    build_gbs = build_bytes / (build_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ft, tree, n, k, args.cpu_seconds)

    out = {
        "metric": "self-kNN queries/s, N=10M k=32, learned-bound vs plain; BVH build ms",
        "value": round(ladder.get("learned_qps", qps), 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(build_ms + query_ms, 4),
        "wall_ms_per_step": round(wall_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded NFW-clustered snapshot)",
        "config": {"workload": "clustered NFW snapshot N=%d, int64 21-bit PH keys, periodic "
                               "L=100; build_bvh + self_knn_query k=%d exclude_self" % (n, k),
                   "n": n, "k": k, "global_batch": n, "parallelism": "single GPU"},
        "build_ms": round(build_ms, 4),
        "build_stage_ms": {a: round(b, 4) for a, b in stages.items()},
        "build_hbm_frac": round(build_gbs / PEAK_HBM_GBS, 4),
        "self_knn_ms": round(query_ms, 4),
        "ladder": {a: round(b, 1) for a, b in ladder.items()},
        "learned": learned,
        "traversal": {"nodes_per_wave": float(st[0]) / ((n + 63) // 64),
                      "leaves_per_wave": float(st[1]) / ((n + 63) // 64),
                      "inserts_per_query": float(st[4]) / n, "stack_overflows": int(st[2])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(ft, tree, n, k, seconds):
    """The oracle's restatement of the reference's self-kNN (64-entry stack,
    max-heap, `>=` prune) on the host cores over a strided sample of queries."""
    import oracle

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    g = tree.to_host()
    t = oracle.Tree(ct=oracle.CT_U64, n=n, sfc=oracle.SFC_PH, bits=21, periodic=1)
    t.arrays = {"min_x": g["min_x"], "max_x": g["max_x"], "min_y": g["min_y"],
                "max_y": g["max_y"], "min_z": g["min_z"], "max_z": g["max_z"],
                "left": g["left_child"], "right": g["right_child"], "parent": g["parent"],
                "id": g["id"], "ghost": g["is_ghost"], "orig": g["orig_idx"]}
    probe = np.arange(0, n, max(1, n // 20000), dtype=np.int32)
    t0 = time.perf_counter()
    oracle.knn_reference(t, k, True, leaf_ids=probe, max_k=32)
    rate = len(probe) / (time.perf_counter() - t0)
    m = int(min(n, max(len(probe), rate * seconds)))
    ids = np.linspace(0, n - 1, m).astype(np.int32)
    t0 = time.perf_counter()
    oracle.knn_reference(t, k, True, leaf_ids=ids, max_k=32)
    dt = time.perf_counter() - t0
    log(f"cpu baseline: {m} queries in {dt:.1f}s on {threads} threads")
    return {"value": round(m / dt, 1), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{m} of {n} self-kNN queries (evenly spaced leaf ranks), k={k}, "
                      f"reference traversal restated in oracle/oracle.cpp, OpenMP"}


if __name__ == "__main__":
    main()
