"""GPU parity at full size on the benchmark configurations.

Clustered production shape: N = 10M NFW-clustered int64 points (21 bits per
dimension, periodic, 63-bit Peano-Hilbert keys), snapshot seed 0 -- held out
from the learned model's training seeds.  The build equals the oracle's build;
self-kNN at k = 1, 32 and 64 gives the same arrays from the packet walk and the
one-lane-per-query walk on every row, and 1,000 sampled rows equal the oracle's
brute force.  Learned bound on the same snapshot at k = 32 with the committed
model: the seeded run returns the unseeded arrays on every row, and exactly
the queries whose seed lies below their k-th distance are re-run.

SPH/Voronoi mixed query shape: N = 20M clustered double points, periodic
L = 100: the build equals the oracle's; the packet walk (the benchmarked one,
with its float32 leaf filter) equals the one-lane-per-query walk on every row
of the 20% subset at k = 64 and of the self query at k = 32.

The acceptance rule is the reference's (test/test_self_knn.cpp:153-266),
tightened to identical indices at tied ranks as well (canonical order).
"""
import os

import numpy as np
import pytest

import oracle as O
from fasttree_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "fasttree_amd", "models", "nfw10m_int64_ph21_k32.npz")


def say(msg):
    print(msg, flush=True)


def same_rows(a, b, rows, k, chunk=1 << 20):
    """Device arrays a and b (rows x k) are equal, compared chunk by chunk."""
    for r0 in range(0, rows, chunk):
        c = min(chunk, rows - r0)
        if not np.array_equal(a.to_host_range(r0 * k, c * k), b.to_host_range(r0 * k, c * k)):
            return False
    return True


def rows_of(a, ids, k):
    return np.stack([a.to_host_range(int(i) * k, k) for i in ids])


def compare_trees(g, t, n):
    m, o = 2 * n - 1, n - 1
    assert np.array_equal(g["sorted_keys"], t.sorted_keys[:n]), "sorted keys"
    assert np.array_equal(g["orig_idx"][o:m], t.orig[o:m]), "permutation"
    for a in ("min_x", "max_x", "min_y", "max_y", "min_z", "max_z"):
        assert np.array_equal(g[a][:m].view(np.uint8), t.arrays[a][:m].view(np.uint8)), a
    assert np.array_equal(g["left_child"], t.left), "left"
    assert np.array_equal(g["right_child"], t.right), "right"
    assert np.array_equal(g["parent"][:m], t.parent[:m]), "parent"
    assert np.array_equal(g["range_first"], t.range_first)
    assert np.array_equal(g["range_last"], t.range_last)


def test_clustered_10m_int64_selfknn_and_learned_bound(ft, orc):
    n = 10_000_000
    x, y, z = synth.clustered_fixed(n, 0, bits=21, box=100.0, dtype=np.uint64)
    cfg = ft.Config(sfc=ft.SFC_PEANO_HILBERT, bits_per_dim=21, periodic=True)
    tree = ft.TreeSoA(n, ft.COORD_U64, cfg)
    ft.build_bvh(ft.particles(*(ft.DeviceArray.from_host(a) for a in (x, y, z))), tree)
    t = O.build_bvh(O.CT_U64, x, y, z, sfc=O.SFC_PH, bits=21, periodic=1)
    compare_trees(tree.to_host(), t, n)
    say("10M build equals the oracle's")
    del x, y, z
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(n, 1000, replace=False)).astype(np.int32)
    keep = {}
    for k in (1, 32, 64):
        out = {}
        for name, algo in (("packet", ft.ALGO_AUTO), ("plain", ft.ALGO_PLAIN)):
            r, d = ft.DeviceArray(n * k, np.uint64), ft.DeviceArray(n * k, np.uint64)
            ft.self_knn_query(tree, k, results=r, dists=d, algorithm=algo, max_k=64)
            out[name] = (r, d)
        assert same_rows(out["packet"][1], out["plain"][1], n, k), f"d2 k={k}"
        assert same_rows(out["packet"][0], out["plain"][0], n, k), f"indices k={k}"
        rb, db = O.knn_bruteforce(t, k, True, leaf_ids=sample)
        assert np.array_equal(rows_of(out["packet"][1], sample, k), db), k
        assert np.array_equal(rows_of(out["packet"][0], sample, k), rb), k
        say(f"k={k}: packet == plain on all {n} rows, 1000 rows == brute force")
        if k == 32:
            keep = {"r": out["packet"][0], "d": out["packet"][1]}
        del out

    # learned bound at k = 32 with the committed model
    from fasttree_amd import learned as L
    k = 32
    model = L.MLP.load(MODEL)
    seeds = ft.DeviceArray(n, np.uint64)
    feats = ft.DeviceArray(n * 16, np.float32)
    model.seed(tree, k, seeds, feats)
    r, d = ft.DeviceArray(n * k, np.uint64), ft.DeviceArray(n * k, np.uint64)
    stats = ft.DeviceArray(8, np.uint64).zero()
    ft.self_knn_query(tree, k, results=r, dists=d, seed_bound=seeds, stats=stats)
    assert same_rows(d, keep["d"], n, k) and same_rows(r, keep["r"], n, k)
    kth = np.concatenate([keep["d"].to_host_range(r0 * k, min(1 << 20, n - r0) * k)
                          .reshape(-1, k)[:, k - 1] for r0 in range(0, n, 1 << 20)])
    too_small = int(np.count_nonzero(seeds.to_host() < kth))
    reruns = int(stats.to_host()[3])
    assert reruns == too_small, (reruns, too_small)
    assert 0 < reruns < n // 100
    say(f"learned bound: identical on all rows, {reruns} re-runs == seeds below the k-th")


def test_sph_20m_double_build_and_packet_walks(ft, orc):
    n, L = 20_000_000, 100.0
    x, y, z = synth.nfw_clustered(n, 21, box=L, n_halos=4000)
    cfg = ft.Config(sfc=ft.SFC_PEANO_HILBERT, bits_per_dim=21, periodic=True, box_length=L)
    tree = ft.TreeSoA(n, ft.COORD_F64, cfg)
    ft.build_bvh(ft.particles(*(ft.DeviceArray.from_host(a) for a in (x, y, z))), tree)
    t = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_PH, bits=21, periodic=1, box=L)
    compare_trees(tree.to_host(), t, n)
    say("20M double build equals the oracle's")
    del x, y, z, t
    rng = np.random.default_rng(5)
    sub = np.sort(rng.permutation(n)[: n // 5]).astype(np.int32)
    rng.shuffle(sub)
    dsub = ft.DeviceArray.from_host(sub)
    nq = len(sub)
    for name, nrows, k, run in (
            ("subset k=64", nq, 64, lambda a, r, d: ft.self_knn_query_subset(
                tree, dsub, 64, results=r, dists=d, algorithm=a)),
            ("self k=32", n, 32, lambda a, r, d: ft.self_knn_query(
                tree, 32, results=r, dists=d, algorithm=a))):
        out = {}
        for algo in (ft.ALGO_AUTO, ft.ALGO_PLAIN):
            r, d = ft.DeviceArray(nrows * k, np.uint64), ft.DeviceArray(nrows * k, np.float64)
            run(algo, r, d)
            out[algo] = (r, d)
        assert same_rows(out[ft.ALGO_AUTO][1], out[ft.ALGO_PLAIN][1], nrows, k), name
        assert same_rows(out[ft.ALGO_AUTO][0], out[ft.ALGO_PLAIN][0], nrows, k), name
        say(f"{name}: packet == plain on all {nrows} rows")
        del out
