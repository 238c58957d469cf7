"""GPU parity of the wave-per-query tier (algorithm=ALGO_WAVE): one wavefront
per query, a sorted list spread over its lanes and leaf batches merged by
cross-lane bitonic networks -- the canonical counterpart of the reference's
SortedMergeHeap (src/maxheap.hpp:226-324, knn_query_large_k
src/hlbvh.hpp:1232-1387).  Every result equals the oracle's brute force
ordered by (d2, index), ties included, bit for bit."""
import numpy as np
import pytest

import oracle as O
from test_gpu_knn import DATASETS, comb_points, make, nfw_points, ref_points

pytestmark = pytest.mark.gpu

WAVE = 4


@pytest.mark.parametrize("ds", DATASETS, ids=[d[0] for d in DATASETS])
def test_wave_self_knn(ft, orc, ds):
    """Self queries at every list size (64 to 1024 entries: k up to 64, 128,
    256, 512, 1024), with and without the query itself."""
    name, ct, sfc, bits, periodic, gen = ds
    tree, t = make(ft, ct, sfc, bits, periodic, gen())
    for k in (1, 5, 32, 33, 64, 65, 128, 200, 256, 300):
        for excl in (True, False):
            print(f"[wave] {name} k={k} excl={excl}", flush=True)
            stats = ft.DeviceArray(8, np.uint64).zero()
            r, d = ft.self_knn_query(tree, k, exclude_self=excl, max_k=1024, algorithm=WAVE,
                                     stats=stats)
            rb, db = O.knn_bruteforce(t, k, excl)
            assert np.array_equal(d, db), f"d2 k={k} excl={excl}"
            assert np.array_equal(r, rb), f"idx k={k} excl={excl}"
            assert stats.to_host()[3] == 0  # exact in one walk: nothing re-run


def test_wave_large_k(ft, orc):
    """k = 700 and 1024 (lists of 16 slots per lane) on clustered points."""
    pts = nfw_points(6000, 31, O.CT_U64)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, pts)
    ids = np.random.default_rng(1).choice(6000, 700, replace=False).astype(np.int32)
    for k in (700, 1024):
        r, d = ft.self_knn_query_subset(tree, ids, k, max_k=1024, algorithm=WAVE)
        rb, db = O.knn_bruteforce(t, k, True, leaf_ids=ids)
        assert np.array_equal(d, db) and np.array_equal(r, rb), k


def test_wave_subset_and_external(ft, orc):
    pts = nfw_points(20000, 7, O.CT_U64)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, pts)
    rng = np.random.default_rng(11)
    ids = rng.permutation(20000)[:3001].astype(np.int32)
    q = tuple(rng.integers(0, 1 << 21, 777, dtype=np.uint64) for _ in range(3))
    for k in (1, 8, 33, 128, 256):
        r, d = ft.self_knn_query_subset(tree, ids, k, algorithm=WAVE)
        rb, db = O.knn_bruteforce(t, k, True, leaf_ids=ids)
        assert np.array_equal(d, db) and np.array_equal(r, rb), ("subset", k)
        r, d = ft.knn_query(tree, *q, k, algorithm=WAVE)
        rb, db = O.knn_bruteforce(t, k, queries=q)
        assert np.array_equal(d, db) and np.array_equal(r, rb), ("external", k)
    # double coordinates, periodic box
    x, y, z = nfw_points(8000, 9, O.CT_F64)
    tree, t = make(ft, O.CT_F64, O.SFC_PH, 21, 1, (x, y, z), box=100.0)
    q = tuple(rng.uniform(0, 100.0, 500) for _ in range(3))
    for k in (40, 150):
        r, d = ft.knn_query(tree, *q, k, algorithm=WAVE)
        rb, db = O.knn_bruteforce(t, k, queries=q)
        assert np.array_equal(d, db) and np.array_equal(r, rb), ("external f64", k)


@pytest.mark.parametrize("n", [1, 2, 7, 63, 64, 65, 130, 200, 400])
def test_wave_edge_sizes_and_sentinels(ft, orc, n):
    """Trees smaller than, equal to and just above one leaf batch or the
    anchor window; k above n - 1 leaves missing entries (index UINT64_MAX,
    distance -DBL_MAX / 0)."""
    for ct in (O.CT_F64, O.CT_U64):
        x, y, z = ref_points(O.GEN_UNIFORM, n, 777, ct)
        tree, t = make(ft, ct, O.SFC_PH, 21, 0, (x, y, z))
        for k in (1, 4, 33, 64, 129, 300):
            for excl in (True, False):
                r, d = ft.self_knn_query(tree, k, exclude_self=excl, max_k=1024, algorithm=WAVE)
                rb, db = O.knn_bruteforce(t, k, excl)
                assert np.array_equal(d, db) and np.array_equal(r, rb), (ct, k, excl)


@pytest.mark.parametrize("ct", [O.CT_U64, O.CT_U32])
def test_wave_deep_comb_tree(ft, orc, ct):
    """A 63-level comb: the wave's stack (128 entries) holds the whole descent."""
    pts = comb_points(ct)
    tree, t = make(ft, ct, O.SFC_MORTON, 21, 0, pts)
    for k in (1, 4, 33, 64):
        rb, db = O.knn_bruteforce(t, k, True)
        r, d = ft.self_knn_query(tree, k, max_k=256, algorithm=WAVE)
        assert np.array_equal(d, db) and np.array_equal(r, rb), k


def test_wave_orig_indices_and_ties(ft, orc):
    """RETURN_ORIG_INDICES (ties ordered by the original index) and a lattice
    with many equal distances."""
    x, y, z = nfw_points(5000, 8, O.CT_F64)
    cfg = ft.Config(sfc=O.SFC_PH, bits_per_dim=21, return_orig_indices=True)
    tree = ft.TreeSoA(len(x), O.CT_F64, cfg)
    ft.build_bvh(ft.particles(x, y, z), tree)
    t = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_PH, bits=21, orig_indices=1)
    for k in (12, 100):
        r, d = ft.self_knn_query(tree, k, max_k=256, algorithm=WAVE)
        rb, db = O.knn_bruteforce(t, k, True)
        assert np.array_equal(d, db) and np.array_equal(r, rb), k
    g = np.arange(12, dtype=np.uint64)
    gx, gy, gz = (a.ravel().copy() for a in np.meshgrid(g, g, g, indexing="ij"))
    tree, t = make(ft, O.CT_U64, O.SFC_MORTON, 21, 0, (gx, gy, gz))
    for k in (6, 18, 26, 100):
        r, d = ft.self_knn_query(tree, k, max_k=256, algorithm=WAVE)
        rb, db = O.knn_bruteforce(t, k, True)
        assert np.array_equal(d, db) and np.array_equal(r, rb), ("lattice", k)
