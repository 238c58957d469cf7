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
    ap.add_argument("--margin-probe", action="store_true",
                    help="time the learned path at other seed margins (log d2 units)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--bucket", type=int, default=0, help="packet leaf bucket (0 = auto)")
    ap.add_argument("--sweep-buckets", action="store_true")
    ap.add_argument("--topk", type=int, default=0,
                    help="packet top-k container: 0 auto, 1 registers, 2 LDS heap")
    ap.add_argument("--window", type=int, default=0,
                    help="packet window pre-scan (leaves on each side; 0 auto, -1 none)")
    ap.add_argument("--model", default=os.path.join(ROOT, "fasttree_amd", "models",
                                                    "nfw10m_int64_ph21_k32.npz"),
                    help="trained learned-radius model (npz); the learned step is skipped "
                         "when it is missing")
    ap.add_argument("--sph-model", default=os.path.join(ROOT, "fasttree_amd", "models",
                                                        "nfw20m_f64_ph21_k32k64.npz"),
                    help="learned-radius model for the SPH mixed query (double coordinates, "
                         "k = 32 and 64); its seeded legs are skipped when it is missing")
    ap.add_argument("--seed-probe", action="store_true",
                    help="time the packet kernel seeded with scaled exact k-th distances")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one build, one self-kNN and one 1 GiB copy kernel, then exit (for "
                         "hardware counter runs); prints no JSON")
    ap.add_argument("--wave-probe", action="store_true",
                    help="one build and 1M external k=128 queries through the wave-per-query "
                         "tier (half near the particles, half uniform in the box), twice, "
                         "then exit (for hardware counter runs); prints no JSON")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched multi-tree workload (development)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched multi-tree workload (64 trees, 4M queries)")
    ap.add_argument("--no-large-k", action="store_true",
                    help="skip the wave-per-query tier's large-k comparison on the C2 tree")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the small plain and the SPH mixed-query configurations")
    ap.add_argument("--configs-only", action="store_true",
                    help="run only the small plain and SPH mixed-query configurations")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import fasttree_amd as ft
    from fasttree_amd import synth

    if args.configs_only:
        print(json.dumps({"configs": configs_bench(ft, args.steps, args.sph_model)}), flush=True)
        return

    if args.batched_only:
        print(json.dumps({"batched": batched_bench(ft, args.steps, args.warmup, args.seed)}),
              flush=True)
        return

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
                          stream=stream.handle, algorithm=algorithm, leaf_bucket=args.bucket,
                          topk=args.topk, window=args.window)
        ev[2].record(stream)
        stream.synchronize()
        return ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])

    if args.wave_probe:
        ft.build_bvh(p, tree, scratch=scratch, stream=stream.handle)
        rng = np.random.default_rng(3)
        nq, kw = 1_000_000, 128
        sel = rng.integers(0, n, nq)
        mask = np.uint64((1 << 21) - 1)
        qs = []
        for a in (x, y, z):
            v = (a[sel] + rng.integers(0, 64, nq).astype(np.uint64)) & mask
            v[nq // 2:] = rng.integers(0, 1 << 21, nq - nq // 2).astype(np.uint64)
            qs.append(ft.DeviceArray.from_host(v))
        rw, dw = ft.DeviceArray(nq * kw, np.uint64), ft.DeviceArray(nq * kw, np.uint64)
        stw = ft.DeviceArray(8, np.uint64).zero()
        # the first call sizes the query-order buffers and counts the work;
        # the second (timed) runs without counters
        ft.knn_query(tree, *qs, kw, results=rw, dists=dw, algorithm=ft.ALGO_WAVE,
                     stream=stream.handle, stats=stw)
        ev[0].record(stream)
        ft.knn_query(tree, *qs, kw, results=rw, dists=dw, algorithm=ft.ALGO_WAVE,
                     stream=stream.handle)
        ev[1].record(stream)
        stream.synchronize()
        sv = stw.to_host() / nq
        log(f"wave probe: 1M external k=128 in {ev[0].elapsed_ms(ev[1]):.2f} ms; per query "
            f"{sv[0]:.1f} nodes, {sv[1]:.1f} leaves, {sv[4]:.1f} candidates merged")
        return

    if args.pmc_probe:
        b, q = step()
        cb = 1 << 30
        src, dstc = ft.DeviceArray(cb // 8, np.uint64), ft.DeviceArray(cb // 8, np.uint64)
        ft._check(ft.lib().ft_copy_kernel(dstc.ptr, src.ptr, cb, stream.handle))
        stream.synchronize()
        log(f"pmc probe: build {b:.3f} ms, self-kNN {q:.3f} ms, copy kernel {cb} bytes")
        return

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

    # The same step (build + packet self-kNN) captured once in a hipGraph and
    # replayed: what a caller pays per rebuild of moved particles.
    with ft.Graph(stream) as graph:
        ft.build_bvh(p, tree, scratch=scratch, stream=stream.handle)
        ft.self_knn_query(tree, k, results=res, dists=dst, exclude_self=True,
                          stream=stream.handle, algorithm=ft.ALGO_PACKET, leaf_bucket=args.bucket,
                          topk=args.topk, window=args.window)
    gts = []
    for _ in range(max(args.steps, 2)):
        ev[0].record(stream)
        graph.launch()
        ev[3].record(stream)
        stream.synchronize()
        gts.append(ev[0].elapsed_ms(ev[3]))
    graph_step_ms = float(np.median(gts))
    del graph
    log(f"graph-captured step (build + self-kNN): {graph_step_ms:.3f} ms")

    # Per-stage build times (events between stages on the same stream).
    scratch.set_profiling(True)
    ft.build_bvh(p, tree, scratch=scratch, stream=stream.handle)
    stream.synchronize()
    stages = scratch.stage_ms()
    scratch.set_profiling(False)

    # Traversal statistics of the measured kernel.
    stats = ft.DeviceArray(8, np.uint64).zero()
    ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle, stats=stats,
                      algorithm=ft.ALGO_PACKET, leaf_bucket=args.bucket, topk=args.topk,
                      window=args.window)
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
                              seed_bound=seeds, stats=lst, leaf_bucket=args.bucket,
                              topk=args.topk, window=args.window)
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
        nodes_seeded = float(ls[0])  # the stats of the last repetition
        learned = {"learned_ms": round(seed_ms + trav_ms, 4), "seed_ms": round(seed_ms, 4),
                   "seeded_traversal_ms": round(trav_ms, 4),
                   "rerun_fraction": round(float(ls[3]) / n, 5),
                   "inserts_per_query": round(float(ls[4]) / n, 2),
                   "nodes_per_wave": round(nodes_seeded / ((n + 63) // 64), 2),
                   "leaves_per_wave": round(float(ls[1]) / ((n + 63) // 64), 2),
                   "pruned_node_fraction": round(1.0 - nodes_seeded / max(float(st[0]), 1.0), 4),
                   "identical_on_2000_sampled_rows": bool(same),
                   "model": os.path.relpath(args.model, ROOT)}
        ladder["learned_qps"] = n / ((seed_ms + trav_ms) * 1e-3)
        log(f"learned: seed {seed_ms:.2f} ms + traversal {trav_ms:.2f} ms, reruns "
            f"{learned['rerun_fraction']:.4f}, inserts/q {learned['inserts_per_query']}, "
            f"nodes/wave {learned['nodes_per_wave']}, identical sample {same}")
        if args.margin_probe:
            m0 = model.get()[2]
            for mg in (0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 1.0):
                model.set(margin=mg)
                lst.zero()
                ev[0].record(stream)
                model.seed(tree, k, seeds, feats, stream=stream.handle)
                ft.self_knn_query(tree, k, results=res2, dists=dst2, stream=stream.handle,
                                  seed_bound=seeds, stats=lst, leaf_bucket=args.bucket,
                                  topk=args.topk, window=args.window)
                ev[1].record(stream)
                stream.synchronize()
                lm = lst.to_host()
                log(f"margin {mg}: learned {ev[0].elapsed_ms(ev[1]):.2f} ms, reruns "
                    f"{lm[3] / n:.4f}, inserts/q {lm[4] / n:.1f}")
            model.set(margin=m0)
        del res2, dst2, feats, seeds
    if args.sweep_buckets:
        for tk in (ft.TOPK_REGISTERS, ft.TOPK_LDS_HEAP):
            for bk in (16, 32, 48, 64):
                ev[0].record(stream)
                ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle,
                                  algorithm=ft.ALGO_PACKET, leaf_bucket=bk, topk=tk)
                ev[1].record(stream)
                stream.synchronize()
                log(f"topk {tk} bucket {bk}: {ev[0].elapsed_ms(ev[1]):.2f} ms")
        for gw in (16, 32, 64):
            ev[0].record(stream)
            ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle,
                              algorithm=ft.ALGO_PACKET, group_width=gw)
            ev[1].record(stream)
            stream.synchronize()
            log(f"group width {gw}: {ev[0].elapsed_ms(ev[1]):.2f} ms")
        for wn in (-1, 32, 64, 128, 256):
            ev[0].record(stream)
            ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle,
                              algorithm=ft.ALGO_PACKET, leaf_bucket=args.bucket or 64, window=wn)
            ev[1].record(stream)
            stream.synchronize()
            log(f"window {wn}: {ev[0].elapsed_ms(ev[1]):.2f} ms")
    if args.seed_probe:
        dk = dst.to_host().reshape(n, k)[:, k - 1].astype(np.float64)
        for f in (1.0, 1.2, 1.5, 2.0):
            sd = ft.DeviceArray.from_host(np.floor(dk * f).astype(np.uint64))
            st2 = ft.DeviceArray(8, np.uint64).zero()
            ev[0].record(stream)
            ft.self_knn_query(tree, k, results=res, dists=dst, stream=stream.handle,
                              seed_bound=sd, stats=st2, leaf_bucket=args.bucket,
                              window=args.window)
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

    large_k = None if args.no_large_k else large_k_bench(ft, tree, x, y, z, n, stream, ev)

    # Roofline of the dominant kernel (the self-kNN traversal): algorithmic bytes
    # N*(3s + 16k) + (12s + 8)*N with s = 8 bytes per coordinate.
    s_b = 8
    knn_bytes = n * (3 * s_b + 16 * k) + (12 * s_b + 8) * n
    build_bytes = 492 * n
    achieved = knn_bytes / (query_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": traffic,
            "kernel": "packet_knn_kernel<uint64_t,32> (unseeded, packed keys)",
            "algorithmic_bytes": knn_bytes,
            "traffic_source": TRAFFIC_PROFILE if traffic is not None else None}
    build_gbs = build_bytes / (build_ms * 1e-3) / 1e9
    issue = valu_issue(query_ms)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ft, tree, x, y, z, n, k, args.cpu_seconds)

    batched = None
    if not args.no_batched or not args.no_configs:
        del res, dst, tree, scratch, p
    if not args.no_batched:
        batched = batched_bench(ft, args.steps, args.warmup, args.seed)
    configs = None if args.no_configs else configs_bench(ft, args.steps, args.sph_model)

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
        "graph_step_ms": round(graph_step_ms, 4),
        "build_hbm_frac": round(build_gbs / PEAK_HBM_GBS, 4),
        "self_knn_ms": round(query_ms, 4),
        "ladder": {a: round(b, 1) for a, b in ladder.items()},
        "learned": learned,
        "traversal": {"nodes_per_wave": float(st[0]) / ((n + 63) // 64),
                      "leaves_per_wave": float(st[1]) / ((n + 63) // 64),
                      "inserts_per_query": float(st[4]) / n, "stack_overflows": int(st[2]),
                      "insert_rounds_per_wave": float(st[5]) / ((n + 63) // 64),
                      "chains_per_wave": float(st[6]) / ((n + 63) // 64)},
        "roofline": roof,
        "valu_issue": issue,
        "cpu_baseline": cpu,
        "large_k": large_k,
        "batched": batched,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)


def large_k_bench(ft, tree, x, y, z, n, stream, ev):
    """The wave-per-query tier (one wavefront per query, a sorted list spread
    over its lanes) against the tier the other algorithms give, on the C2 tree:
    1M external queries near the particles at k = 128 (against the packet tier
    in passes of 32) and 2.5M self queries (every fourth leaf) at k = 256
    (against the per-lane LDS heaps); rows compared on a sample."""
    rng = np.random.default_rng(3)

    def timed(fn, reps=2):
        fn()
        stream.synchronize()
        ts = []
        for _ in range(reps):
            ev[0].record(stream)
            fn()
            ev[1].record(stream)
            stream.synchronize()
            ts.append(ev[0].elapsed_ms(ev[1]))
        return min(ts)

    def same_rows(ra, da, rb, db, nq, k):
        rows = np.linspace(0, nq - 1, 1000).astype(np.int64)
        return all(np.array_equal(ra.to_host_range(r * k, k), rb.to_host_range(r * k, k)) and
                   np.array_equal(da.to_host_range(r * k, k), db.to_host_range(r * k, k))
                   for r in rows)

    out = {}
    nq, k = 1_000_000, 128
    sel = rng.integers(0, n, nq)
    mask = np.uint64((1 << 21) - 1)
    q = [ft.DeviceArray.from_host((a[sel] + rng.integers(0, 64, nq).astype(np.uint64)) & mask)
         for a in (x, y, z)]
    ra, da, rb, db = (ft.DeviceArray(nq * k, np.uint64) for _ in range(4))
    t_other = timed(lambda: ft.knn_query(tree, *q, k, results=ra, dists=da, stream=stream.handle,
                                         algorithm=ft.ALGO_PACKET))
    t_wave = timed(lambda: ft.knn_query(tree, *q, k, results=rb, dists=db, stream=stream.handle,
                                        algorithm=ft.ALGO_WAVE))
    out["external_k128"] = {"queries": nq, "wave_ms": round(t_wave, 3),
                            "packet_passes_ms": round(t_other, 3),
                            "wave_queries_per_s": round(nq / (t_wave * 1e-3), 1),
                            "identical_on_1000_sampled_rows": bool(same_rows(ra, da, rb, db, nq, k))}
    log(f"large k: external k=128 wave {t_wave:.2f} ms, packet passes {t_other:.2f} ms")
    del ra, da, rb, db, q
    nq, k = n // 4, 256
    ids = ft.DeviceArray.from_host(np.arange(nq, dtype=np.int32) * 4)
    ra, da, rb, db = (ft.DeviceArray(nq * k, np.uint64) for _ in range(4))
    t_other = timed(lambda: ft.self_knn_query_subset(tree, ids, k, results=ra, dists=da,
                                                     stream=stream.handle,
                                                     algorithm=ft.ALGO_GROUPED), reps=1)
    t_wave = timed(lambda: ft.self_knn_query_subset(tree, ids, k, results=rb, dists=db,
                                                    stream=stream.handle,
                                                    algorithm=ft.ALGO_WAVE))
    out["self_subset_k256"] = {"queries": nq, "wave_ms": round(t_wave, 3),
                               "lane_heap_ms": round(t_other, 3),
                               "wave_queries_per_s": round(nq / (t_wave * 1e-3), 1),
                               "identical_on_1000_sampled_rows": bool(same_rows(ra, da, rb, db, nq, k))}
    log(f"large k: self subset k=256 wave {t_wave:.2f} ms, per-lane heaps {t_other:.2f} ms")
    return out


TRAFFIC_PROFILE = "profiles/c2_hbm_traffic_counters.md"
SQ_PROFILE = "profiles/c2_sq_instruction_counters.md"
# MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock; a wave64 VALU
# instruction occupies its SIMD for 4 cycles (FP64 included: full-rate FP64).
SIMDS, CLOCK_HZ, CYCLES_PER_VALU = 1024, 2.4e9, 4


def valu_issue(kernel_ms):
    """The self-kNN kernel against its VALU-issue bound: the committed SQ counter
    summary's VALU instructions per wave times waves, at 4 cycles each spread
    over every SIMD, against the kernel's measured time; or None."""
    import re
    try:
        text = open(os.path.join(ROOT, SQ_PROFILE)).read()
    except OSError:
        return None
    m = re.search(r"Self-kNN VALU: ([0-9,]+) instructions per wave over ([0-9,]+) waves", text)
    if not m:
        return None
    per_wave, waves = (int(g.replace(",", "")) for g in m.groups())
    issue_ms = per_wave * waves * CYCLES_PER_VALU / (SIMDS * CLOCK_HZ) * 1e3
    return {"valu_per_wave": per_wave, "waves": waves, "issue_ms": round(issue_ms, 3),
            "kernel_ms": round(kernel_ms, 3), "frac": round(issue_ms / kernel_ms, 4),
            "model": "VALU instructions x 4 cycles / (1024 SIMDs x 2.4 GHz)",
            "source": SQ_PROFILE}


def measured_traffic():
    """HBM bytes per self-kNN launch from the committed FETCH_SIZE / WRITE_SIZE
    counter summary (scaled by the copy-kernel calibration), or None."""
    import re
    try:
        text = open(os.path.join(ROOT, TRAFFIC_PROFILE)).read()
    except OSError:
        return None
    m = re.search(r"Self-kNN traffic: ([0-9.]+) GB per launch", text)
    return float(m.group(1)) * 1e9 if m else None


def configs_bench(ft, steps, sph_model=None):
    """The small plain configuration and the SPH/Voronoi mixed query, timed with
    HIP events on one stream, inputs resident in HBM.

    Small plain: N = 100,000 uniform double points (seed 1), 30-bit Morton,
    non-periodic; build_bvh + self_knn_query k = 16.  SPH mixed query: N = 20M
    clustered double points in a periodic box L = 100, 63-bit Peano-Hilbert;
    (a) self_knn_query_subset over a seeded 20% of the leaves at k = 64 (FP
    k > 32: the grouped and plain per-lane walks, which must agree), (b) the
    k = 32 self query, h_i = 1.2 x its 32nd neighbour distance, then the CSR
    range query over every particle with radius h_i (count + scan pass, fill
    pass).  Both queries also run seeded by the learned radius (a model trained
    on double-coordinate snapshots of other seeds at k = 32 and 64): features +
    MLP seeds, the seeded walk, the exact re-run of queries whose seed held
    fewer than k; their rows must equal the unseeded ones, so the radii h_i
    and the range query are the same either way."""
    from fasttree_amd import synth
    stream = ft.Stream()
    ev = [ft.Event() for _ in range(3)]
    out = {}

    def timed(fn, reps):
        ts = []
        for _ in range(reps + 1):
            ev[0].record(stream)
            fn()
            ev[1].record(stream)
            stream.synchronize()
            ts.append(ev[0].elapsed_ms(ev[1]))
        return float(np.median(ts[1:]))

    # small plain
    n1, k1 = 100_000, 16
    x, y, z = synth.uniform(n1, 1, 1.0)
    cfg = ft.Config(sfc=ft.SFC_MORTON, bits_per_dim=10, periodic=False)
    t1 = ft.TreeSoA(n1, ft.COORD_F64, cfg)
    sc1 = ft.BvhScratch()
    p1 = ft.particles(*(ft.DeviceArray.from_host(a) for a in (x, y, z)))
    r1, d1 = ft.DeviceArray(n1 * k1, np.uint64), ft.DeviceArray(n1 * k1, np.float64)

    def step1():
        ft.build_bvh(p1, t1, scratch=sc1, stream=stream.handle)
        ft.self_knn_query(t1, k1, results=r1, dists=d1, stream=stream.handle)
    ms1 = timed(step1, max(steps, 3))
    out["small_plain"] = {"workload": "N=100k uniform f64, Morton 10 bits/dim, non-periodic; "
                                      "build_bvh + self_knn_query k=16",
                          "step_ms": round(ms1, 4), "queries_per_s": round(n1 / (ms1 * 1e-3), 1)}
    log(f"small plain: {ms1:.3f} ms per build + self-kNN")
    del t1, sc1, p1, r1, d1

    # SPH mixed query
    n5, L = 20_000_000, 100.0
    t0 = time.time()
    x, y, z = synth.nfw_clustered(n5, 21, box=L, n_halos=4000)
    cfg = ft.Config(sfc=ft.SFC_PEANO_HILBERT, bits_per_dim=21, periodic=True, box_length=L)
    t5 = ft.TreeSoA(n5, ft.COORD_F64, cfg)
    sc5 = ft.BvhScratch()
    p5 = ft.particles(*(ft.DeviceArray.from_host(a) for a in (x, y, z)))
    del x, y, z
    log(f"SPH snapshot N={n5} generated in {time.time() - t0:.1f}s")
    build_ms = timed(lambda: ft.build_bvh(p5, t5, scratch=sc5, stream=stream.handle), steps)
    del p5
    rng = np.random.default_rng(5)
    sub = np.sort(rng.permutation(n5)[: n5 // 5]).astype(np.int32)
    rng.shuffle(sub)
    dsub = ft.DeviceArray.from_host(sub)
    ns, ks = len(sub), 64
    ra, da = ft.DeviceArray(ns * ks, np.uint64), ft.DeviceArray(ns * ks, np.float64)
    rb, db = ft.DeviceArray(ns * ks, np.uint64), ft.DeviceArray(ns * ks, np.float64)
    g_ms = timed(lambda: ft.self_knn_query_subset(t5, dsub, ks, results=ra, dists=da,
                                                  stream=stream.handle,
                                                  algorithm=ft.ALGO_GROUPED), steps)
    p_ms = timed(lambda: ft.self_knn_query_subset(t5, dsub, ks, results=rb, dists=db,
                                                  stream=stream.handle,
                                                  algorithm=ft.ALGO_PLAIN), steps)
    rows = rng.choice(ns, 2000, replace=False)
    same = all(np.array_equal(ra.to_host_range(int(r) * ks, ks), rb.to_host_range(int(r) * ks, ks))
               for r in rows)
    k_ms = timed(lambda: ft.self_knn_query_subset(t5, dsub, ks, results=rb, dists=db,
                                                  stream=stream.handle,
                                                  algorithm=ft.ALGO_PACKET), steps)
    same_pk = all(np.array_equal(ra.to_host_range(int(r) * ks, ks),
                                 rb.to_host_range(int(r) * ks, ks)) and
                  np.array_equal(da.to_host_range(int(r) * ks, ks),
                                 db.to_host_range(int(r) * ks, ks)) for r in rows)
    # the default for a subset above k = 32: the wave-per-query tier
    w_ms = timed(lambda: ft.self_knn_query_subset(t5, dsub, ks, results=rb, dists=db,
                                                  stream=stream.handle), steps)
    same_w = all(np.array_equal(ra.to_host_range(int(r) * ks, ks),
                                rb.to_host_range(int(r) * ks, ks)) and
                 np.array_equal(da.to_host_range(int(r) * ks, ks),
                                db.to_host_range(int(r) * ks, ks)) for r in rows)
    model = None
    if sph_model and os.path.exists(sph_model):
        from fasttree_amd import learned as L
        model = L.MLP.load(sph_model)
    sub_learned = None
    if model is not None:
        sd = ft.DeviceArray(ns, np.float64)
        sf = ft.DeviceArray(ns * L.NF, np.float32)
        st = ft.DeviceArray(8, np.uint64)

        def learned_subset():
            model.seed(t5, ks, sd, sf, leaf_ids=dsub, stream=stream.handle)
            ft.self_knn_query_subset(t5, dsub, ks, results=ra, dists=da, stream=stream.handle,
                                     seed_bound=sd, stats=st)
        seed_ms = timed(lambda: model.seed(t5, ks, sd, sf, leaf_ids=dsub, stream=stream.handle),
                        steps)
        st.zero()
        l_ms = timed(learned_subset, steps)
        reruns = int(st.to_host()[3]) / ((steps + 1) * ns)
        same_l = all(np.array_equal(ra.to_host_range(int(r) * ks, ks),
                                    rb.to_host_range(int(r) * ks, ks)) and
                     np.array_equal(da.to_host_range(int(r) * ks, ks),
                                    db.to_host_range(int(r) * ks, ks)) for r in rows)
        sub_learned = {"learned_ms": round(l_ms, 3), "seed_ms": round(seed_ms, 3),
                       "rerun_fraction": round(reruns, 5), "vs_packet": round(k_ms / l_ms, 3),
                       "equals_unseeded_on_2000_sampled_rows": bool(same_l)}
        log(f"SPH subset k=64 learned: {l_ms:.2f} ms (seed {seed_ms:.2f} ms), reruns "
            f"{reruns:.4f}, same {same_l}")
        del sd, sf
    del ra, da, rb, db, dsub
    log(f"SPH subset k=64: wave (default) {w_ms:.2f} ms, packet (passes of 32) {k_ms:.2f} ms, "
        f"grouped {g_ms:.2f} ms, plain {p_ms:.2f} ms, same {same} {same_pk} {same_w}")
    # smoothing lengths from the k = 32 self query, then the CSR range query
    k32 = 32
    r32, d32 = ft.DeviceArray(n5 * k32, np.uint64), ft.DeviceArray(n5 * k32, np.float64)
    s_ms = timed(lambda: ft.self_knn_query(t5, k32, results=r32, dists=d32,
                                           stream=stream.handle), 1)
    h = 1.2 * np.sqrt(d32.to_host().reshape(n5, k32)[:, k32 - 1])
    self_learned = None
    if model is not None:
        sd = ft.DeviceArray(n5, np.float64)
        sf = ft.DeviceArray(n5 * L.NF, np.float32)
        st = ft.DeviceArray(8, np.uint64).zero()
        r32b, d32b = ft.DeviceArray(n5 * k32, np.uint64), ft.DeviceArray(n5 * k32, np.float64)

        def learned_self():
            model.seed(t5, k32, sd, sf, stream=stream.handle)
            ft.self_knn_query(t5, k32, results=r32b, dists=d32b, stream=stream.handle,
                              seed_bound=sd, stats=st)
        ls_ms = timed(learned_self, 1)
        reruns = int(st.to_host()[3]) / (2 * n5)
        rows5 = np.random.default_rng(6).choice(n5, 2000, replace=False)
        same_s = all(np.array_equal(r32.to_host_range(int(r) * k32, k32),
                                    r32b.to_host_range(int(r) * k32, k32)) and
                     np.array_equal(d32.to_host_range(int(r) * k32, k32),
                                    d32b.to_host_range(int(r) * k32, k32)) for r in rows5)
        hb = 1.2 * np.sqrt(d32b.to_host().reshape(n5, k32)[:, k32 - 1])
        self_learned = {"learned_ms": round(ls_ms, 3), "rerun_fraction": round(reruns, 5),
                        "vs_packet": round(s_ms / ls_ms, 3),
                        "equals_unseeded_on_2000_sampled_rows": bool(same_s),
                        "smoothing_lengths_identical": bool(np.array_equal(h, hb))}
        log(f"SPH self k=32 learned: {ls_ms:.2f} ms vs {s_ms:.2f} ms, reruns {reruns:.4f}, "
            f"same {same_s}, h identical {self_learned['smoothing_lengths_identical']}")
        del sd, sf, r32b, d32b, hb
    del r32, d32
    g = t5.to_host()
    qx, qy, qz = (ft.DeviceArray.from_host(np.ascontiguousarray(g[c][n5 - 1:]))
                  for c in ("min_x", "min_y", "min_z"))
    del g
    rmin = ft.DeviceArray.from_host(np.zeros(n5))
    rmax = ft.DeviceArray.from_host(h)
    rp = ft.DeviceArray(n5 + 1, np.int64)
    lib = ft.lib()

    def count_pass():
        ft._check(lib.ft_range_query_csr(stream.handle, ft.C.byref(t5.c), qx.ptr, qy.ptr, qz.ptr,
                                         rmin.ptr, rmax.ptr, n5, rp.ptr, None, 0))
    c_ms = timed(count_pass, 1)
    total = int(rp.to_host_range(n5, 1)[0])
    cols = ft.DeviceArray(max(total, 1), np.int32)

    def fill_pass():
        ft._check(lib.ft_range_query_csr(stream.handle, ft.C.byref(t5.c), qx.ptr, qy.ptr, qz.ptr,
                                         rmin.ptr, rmax.ptr, n5, rp.ptr, cols.ptr, total))
    f_ms = timed(fill_pass, 1)
    min_row = int(np.diff(rp.to_host()).min())
    log(f"SPH range CSR: {total} hits ({total / n5:.1f} per query), count+scan {c_ms:.1f} ms, "
        f"count+scan+fill {f_ms:.1f} ms")
    out["sph_mixed"] = {
        "workload": "N=20M NFW-clustered f64, periodic L=100, PH 21 bits/dim",
        "build_ms": round(build_ms, 3),
        "subset_k64": {"queries": ns, "wave_ms": round(w_ms, 3), "packet_ms": round(k_ms, 3),
                       "grouped_ms": round(g_ms, 3), "plain_ms": round(p_ms, 3),
                       "queries_per_s": round(ns / (min(w_ms, k_ms, g_ms, p_ms) * 1e-3), 1),
                       "grouped_equals_plain_on_2000_sampled_rows": bool(same),
                       "packet_equals_grouped_on_2000_sampled_rows": bool(same_pk),
                       "wave_equals_grouped_on_2000_sampled_rows": bool(same_w),
                       "learned": sub_learned},
        "self_knn_k32_ms": round(s_ms, 3),
        "self_knn_k32_learned": self_learned,
        "range_csr": {"queries": n5, "hits": total, "hits_per_query": round(total / n5, 2),
                      "count_scan_ms": round(c_ms, 3), "count_scan_fill_ms": round(f_ms, 3),
                      "queries_per_s": round(n5 / (f_ms * 1e-3), 1),
                      "fewest_hits_in_a_row": min_row,
                      "every_row_holds_itself_and_32_neighbours": bool(min_row >= 33)}}
    del t5, sc5, qx, qy, qz, rmin, rmax, rp, cols
    return out


def batched_bench(ft, steps, warmup, seed):
    """64 trees (N log-uniform in [10k, 2M], clustered, int32 21-bit Morton,
    non-periodic), 4M external queries in 256 requests with k in {1, 8, 32,
    128}: one step = build every tree + answer every request, recorded in one
    hipGraph after autotuning.  Beside it: the same step launched eagerly with
    the untuned defaults, and a plain loop of one build_bvh per tree and one
    knn_query per request on one stream."""
    import ctypes as C
    from fasttree_amd import synth
    from fasttree_amd.batch import Batch

    t0 = time.time()
    sizes, pts, reqs, q = synth.batched_workload(seed)
    log(f"batched workload: {len(sizes)} trees, {int(sizes.sum())} points, {len(q[0])} queries "
        f"in {len(reqs)} requests, generated in {time.time() - t0:.1f}s")
    cfg = ft.Config(sfc=ft.SFC_MORTON, bits_per_dim=21, periodic=False)
    trees = [ft.TreeSoA(int(m), ft.COORD_U32, cfg) for m in sizes]
    parts = [tuple(ft.DeviceArray.from_host(a) for a in pp) for pp in pts]
    dq = [ft.DeviceArray.from_host(a) for a in q]
    total_r = sum(r[1] * r[3] for r in reqs)
    res = ft.DeviceArray(total_r, np.uint64)
    dst = ft.DeviceArray(total_r, np.uint64)
    stream = ft.Stream()
    e0, e1 = ft.Event(), ft.Event()

    def timed(fn):
        e0.record(stream)
        fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1)

    b = Batch(ft.COORD_U32)
    b.set_trees(trees, parts)
    b.set_queries(reqs, *dq, res, dst)
    b.run(stream)
    stream.synchronize()
    eager_ms = min(timed(lambda: b.run(stream)) for _ in range(2))
    t0 = time.time()
    b.tune(stream)
    tune_s = time.time() - t0
    b.capture(stream)
    for _ in range(max(warmup, 1)):
        b.launch(stream)
    stream.synchronize()
    graph = [timed(lambda: b.launch(stream)) for _ in range(max(steps, 1))]
    graph_ms = float(np.median(graph))
    log(f"batched: graph step {graph_ms:.2f} ms (eager untuned {eager_ms:.2f} ms, "
        f"tuning {tune_s:.1f}s)")

    # plain loop: one build_bvh per tree, one knn_query per request
    L = ft.lib()
    res2 = ft.DeviceArray(total_r, np.uint64)
    dst2 = ft.DeviceArray(total_r, np.uint64)
    scratch = ft.BvhScratch()

    def loop(verbose=False):
        t0 = time.time()
        for t, tr in enumerate(trees):
            ft.build_bvh(ft.particles(*parts[t]), tr, scratch=scratch, stream=stream.handle)
        for i, (t, kk, qf, m, rf) in enumerate(reqs):
            if m == 0:
                continue
            o = ft.QueryOpts(256, 0, 0, 0, 0, 0, None, None, 0, 0)
            ft._check(L.ft_knn_query(stream.handle, C.byref(trees[t].c), dq[0].ptr + 4 * qf,
                                     dq[1].ptr + 4 * qf, dq[2].ptr + 4 * qf, int(kk), int(m),
                                     res2.ptr + 8 * rf, dst2.ptr + 8 * rf, C.byref(o)))
            if verbose and i % 32 == 31:
                stream.synchronize()
                log(f"per-tree loop: {i + 1} requests, {time.time() - t0:.1f}s")

    loop(verbose=True)
    stream.synchronize()
    loop_ms = timed(loop)
    same = True
    rng = np.random.default_rng(seed)
    for i in rng.choice(len(reqs), 24, replace=False):
        t, kk, qf, m, rf = reqs[i]
        if m == 0:
            continue
        same &= bool(np.array_equal(res.to_host_range(rf, m * kk), res2.to_host_range(rf, m * kk)))
        same &= bool(np.array_equal(dst.to_host_range(rf, m * kk), dst2.to_host_range(rf, m * kk)))
    log(f"batched: per-tree loop {loop_ms:.2f} ms, identical sample {same}")
    nq = int(len(q[0]))
    return {"workload": "64 clustered trees, N log-uniform in [10k, 2M], int32 21-bit Morton, "
                        "non-periodic; 4M external queries in 256 requests, k in {1,8,32,128}; "
                        "step = build every tree + answer every request",
            "trees": len(sizes), "points": int(sizes.sum()), "queries": nq,
            "requests": len(reqs), "graph_step_ms": round(graph_ms, 3),
            "queries_per_s_incl_builds": round(nq / (graph_ms * 1e-3), 1),
            "eager_untuned_step_ms": round(eager_ms, 3),
            "per_tree_loop_step_ms": round(loop_ms, 3),
            "speedup_vs_loop": round(loop_ms / graph_ms, 3),
            "identical_to_loop_on_24_sampled_requests": same,
            "tuning_s": round(tune_s, 2), "tuning": b.tuning()}


def host_cpus():
    """What the host grants this process: nproc, the affinity mask, the cgroup
    CPU quota (cpu.max), and the CPU model."""
    info = {"nproc": os.cpu_count() or 1}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity"] = info["nproc"]
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    info["cgroup_quota"] = quota
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info["model"] = model
    info["threads"] = min(info["affinity"], quota or info["affinity"])
    return info


def cpu_baseline(ft, tree, x, y, z, n, k, seconds):
    """The oracle's restatement of the reference's CPU path on every core this
    process may use: the build (stable key sort, Karras topology, refit) over
    the whole snapshot, and the self-kNN walk (64-entry stack, max-heap, `>=`
    prune) over a sample of queries, both as the grouped lockstep walk of the
    reference kernel (8 lanes: the native vector width for 64-bit coordinates)
    and one query at a time.  `value` is the faster of the two walks."""
    import oracle

    cpus = host_cpus()
    threads = cpus["threads"]
    oracle.set_threads(threads)
    builds = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.build_bvh(oracle.CT_U64, x, y, z, sfc=oracle.SFC_PH, bits=21, periodic=1)
        builds.append((time.perf_counter() - t0) * 1e3)
    g = tree.to_host()
    t = oracle.Tree(ct=oracle.CT_U64, n=n, sfc=oracle.SFC_PH, bits=21, periodic=1)
    t.arrays = {"min_x": g["min_x"], "max_x": g["max_x"], "min_y": g["min_y"],
                "max_y": g["max_y"], "min_z": g["min_z"], "max_z": g["max_z"],
                "left": g["left_child"], "right": g["right_child"], "parent": g["parent"],
                "id": g["id"], "ghost": g["is_ghost"], "orig": g["orig_idx"]}

    def blocks(m):  # m queries in 64 evenly spaced runs of consecutive leaf ranks
        run = max(8, -(-m // 64))
        starts = np.linspace(0, n - run, 64).astype(np.int64)
        return np.unique(np.concatenate([np.arange(a, a + run) for a in starts])).astype(np.int32)

    rates = {}
    for name, walk in (("per_query", lambda ids: oracle.knn_reference(t, k, True, leaf_ids=ids,
                                                                        max_k=32)),
                       ("grouped_w8", lambda ids: oracle.knn_reference_grouped(
                           t, k, True, leaf_ids=ids, width=8))):
        probe = blocks(20000)
        t0 = time.perf_counter()
        walk(probe)
        rate = len(probe) / (time.perf_counter() - t0)
        ids = blocks(int(min(n, max(len(probe), rate * seconds / 2))))
        t0 = time.perf_counter()
        walk(ids)
        dt = time.perf_counter() - t0
        rates[name] = (len(ids), len(ids) / dt)
        log(f"cpu baseline {name}: {len(ids)} queries in {dt:.1f}s on {threads} threads")
    best = max(rates, key=lambda r: rates[r][1])
    return {"value": round(rates[best][1], 1), "unit": "queries/s", "cores": threads,
            "kind": "port",
            "sample": f"{rates[best][0]} of {n} self-kNN queries (64 evenly spaced runs of "
                      f"consecutive leaf ranks), k={k}; reference traversal restated in "
                      f"oracle/oracle.cpp, OpenMP; value = the {best} walk",
            "per_query_qps": round(rates["per_query"][1], 1),
            "grouped_w8_qps": round(rates["grouped_w8"][1], 1),
            "build_ms": round(float(np.median(builds)), 1),
            "build_sample": f"oracle build_bvh over all {n} points, median of 3",
            "host": cpus}


if __name__ == "__main__":
    main()
