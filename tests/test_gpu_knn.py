"""GPU parity of the k-nearest-neighbour queries.

Canonical mode (the default): indices and squared distances equal the oracle's
brute force ordered by (d2, index) -- bit for bit, ties included.
Reference mode: equal to the oracle's restatement of the reference's traversal
and heap mechanics -- the same tie picks the reference makes.
"""
import numpy as np
import pytest

import oracle as O
from fasttree_amd import synth

pytestmark = pytest.mark.gpu


def make(ft, ct, sfc, bits, periodic, pts, box=0.0):
    x, y, z = pts
    cfg = ft.Config(sfc=sfc, bits_per_dim=bits, periodic=periodic, box_length=box)
    tree = ft.TreeSoA(len(x), ct, cfg)
    ft.build_bvh(ft.particles(*(ft.DeviceArray.from_host(a) for a in (x, y, z))), tree)
    t = O.build_bvh(ct, x, y, z, sfc=sfc, bits=bits, periodic=periodic, box=box)
    return tree, t


def ref_points(kind, n, seed, ct):
    return O.gen_reference(kind, n, seed, ct, bits=21)


def nfw_points(n, seed, ct, halos=200):
    x, y, z = synth.nfw_clustered(n, seed, box=100.0, n_halos=halos)
    if ct in (O.CT_U64, O.CT_U32):
        dt = O.COORD_DTYPES[ct]
        return tuple(synth.to_fixed(a, 100.0, 21, dt) for a in (x, y, z))
    return tuple(a.astype(O.COORD_DTYPES[ct]) for a in (x, y, z))


ALGOS = {"packet": 3, "grouped": 2, "plain": 1, "packet-heap": 3}
TOPK = {"packet-heap": 2}  # LDS heap container also for k <= 32 (integer metrics)

DATASETS = [
    ("ref-uniform-u64-periodic", O.CT_U64, O.SFC_PH, 21, 1, lambda: ref_points(O.GEN_UNIFORM, 1000, 42, O.CT_U64)),
    ("ref-clustered-f64", O.CT_F64, O.SFC_PH, 21, 0, lambda: ref_points(O.GEN_CLUSTERED, 1000, 12345, O.CT_F64)),
    ("ref-degenerate-u64", O.CT_U64, O.SFC_PH, 21, 0, lambda: ref_points(O.GEN_DEGENERATE, 1000, 999, O.CT_U64)),
    ("ref-clustered10k-u32-morton", O.CT_U32, O.SFC_MORTON, 21, 0, lambda: ref_points(O.GEN_CLUSTERED, 10000, 54321, O.CT_U32)),
    ("nfw-u64-periodic", O.CT_U64, O.SFC_PH, 21, 1, lambda: nfw_points(12000, 5, O.CT_U64)),
    ("uniform-f64-morton10", O.CT_F64, O.SFC_MORTON, 10, 0, lambda: tuple(synth.uniform(8000, 1, 1.0))),
    ("nfw-f32", O.CT_F32, O.SFC_PH, 21, 0, lambda: nfw_points(6000, 6, O.CT_F32)),
]


@pytest.mark.parametrize("ds", DATASETS, ids=[d[0] for d in DATASETS])
@pytest.mark.parametrize("algo", list(ALGOS))
def test_self_knn_canonical(ft, orc, ds, algo):
    name, ct, sfc, bits, periodic, gen = ds
    tree, t = make(ft, ct, sfc, bits, periodic, gen())
    # packet: register tiers (k <= 32) and k in passes of 32 (32 < k <= 128,
    # every metric; "packet-heap": the LDS heap tiers of integer metrics).
    ks = (1, 5, 16, 32, 33, 64, 128) if algo.startswith("packet") else (1, 16, 32, 33, 64)
    for k in ks:
        for excl in (True, False):
            print(f"[knn] {name} {algo} k={k} excl={excl}", flush=True)
            r, d = ft.self_knn_query(tree, k, exclude_self=excl, max_k=256,
                                     algorithm=ALGOS[algo], topk=TOPK.get(algo, 0))
            rb, db = O.knn_bruteforce(t, k, excl)
            assert np.array_equal(d, db), f"d2 k={k} excl={excl}"
            assert np.array_equal(r, rb), f"idx k={k} excl={excl}"


@pytest.mark.parametrize("ds", DATASETS[:4], ids=[d[0] for d in DATASETS[:4]])
def test_self_knn_reference_mode(ft, orc, ds):
    """Reference heap mechanics on the GPU reproduce the reference's tie picks."""
    name, ct, sfc, bits, periodic, gen = ds
    tree, t = make(ft, ct, sfc, bits, periodic, gen())
    for k in (16, 32, 33, 64, 128, 256):
        for excl in (True, False):
            print(f"[knn-ref] {name} k={k} excl={excl}", flush=True)
            r, d = ft.self_knn_query(tree, k, exclude_self=excl, max_k=256,
                                     mode=ft.MODE_REFERENCE)
            rr, dr = O.knn_reference(t, k, excl, max_k=256)
            assert np.array_equal(d, dr)
            assert np.array_equal(r, rr)


@pytest.mark.parametrize("algo", ["packet", "grouped"])
def test_self_knn_subset(ft, orc, algo):
    """self_knn_query_subset: leaf ranks in any order, results in query order,
    including the reference's adversarial interleaved-cluster subset
    (test/test_self_knn.cpp:688-782)."""
    pts = ref_points(O.GEN_CLUSTERED, 10000, 54321, O.CT_U64)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 0, pts)
    lx = t.leaf_coords()[0]
    mid = np.uint64(1 << 20)
    c1 = np.nonzero(lx < mid)[0]
    c2 = np.nonzero(lx >= mid)[0]
    inter = np.array([c1[(i // 2) % len(c1)] if i % 2 == 0 else c2[(i // 2) % len(c2)]
                      for i in range(64)], np.int32)
    rng = np.random.default_rng(3)
    rand = rng.permutation(10000)[:1999].astype(np.int32)
    for ids in (inter, rand, np.array([0, 9999, 5000], np.int32)):
        for k in (8, 32):
            r, d = ft.self_knn_query_subset(tree, ids, k, algorithm=ALGOS[algo])
            rb, db = O.knn_bruteforce(t, k, True, leaf_ids=ids)
            assert np.array_equal(d, db) and np.array_equal(r, rb)


@pytest.mark.parametrize("width", [16, 32])
def test_packet_group_width(ft, orc, width):
    """Packets of 16 or 32 queries (group_width; the reference's sub-group
    width W) return the brute-force rows for self, subset and external queries,
    k in the register tier, in passes of 32 and with a seed."""
    pts = nfw_points(20000, 17, O.CT_U64)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, pts)
    for k in (1, 16, 32, 64):
        r, d = ft.self_knn_query(tree, k, group_width=width, max_k=64)
        rb, db = O.knn_bruteforce(t, k, True)
        assert np.array_equal(d, db) and np.array_equal(r, rb), k
    ids = np.random.default_rng(3).choice(20000, 3001, replace=False).astype(np.int32)
    r, d = ft.self_knn_query_subset(tree, ids, 32, group_width=width)
    rb, db = O.knn_bruteforce(t, 32, True, leaf_ids=ids)
    assert np.array_equal(d, db) and np.array_equal(r, rb)
    rng = np.random.default_rng(5)
    q = tuple(rng.integers(0, 1 << 21, 999, dtype=np.uint64) for _ in range(3))
    r, d = ft.knn_query(tree, *q, 8, group_width=width)
    rb, db = O.knn_bruteforce(t, 8, queries=q)
    assert np.array_equal(d, db) and np.array_equal(r, rb)
    _, d32 = O.knn_bruteforce(t, 16, True)
    seed = ft.DeviceArray.from_host((d32[:, -1] * np.where(np.arange(20000) % 4 == 0, 0.5, 1.5))
                                    .astype(np.uint64))
    r, d = ft.self_knn_query(tree, 16, group_width=width, seed_bound=seed)
    rb, db = O.knn_bruteforce(t, 16, True)
    assert np.array_equal(d, db) and np.array_equal(r, rb)
    with pytest.raises(ValueError):
        ft.self_knn_query(tree, 4, group_width=48)


def test_external_knn(ft, orc):
    pts = nfw_points(20000, 7, O.CT_U64)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, pts)
    rng = np.random.default_rng(11)
    q = tuple(rng.integers(0, 1 << 21, 777, dtype=np.uint64) for _ in range(3))
    for k in (1, 8, 32, 33, 128):
        r, d = ft.knn_query(tree, *q, k)
        rb, db = O.knn_bruteforce(t, k, queries=q)
        assert np.array_equal(d, db) and np.array_equal(r, rb), f"k={k}"


def test_queries_on_two_streams_run_independently(ft, orc):
    """Seeded self-kNN (re-run lists) and external kNN (query ordering) issued on
    two streams with no synchronisation in between: each stream's results equal
    the brute force of its own tree (every stream has its own work buffers)."""
    k = 16
    cases = []
    for seed, n, nq in ((31, 30000, 9000), (32, 45000, 5000)):
        pts = nfw_points(n, seed, O.CT_U64)
        tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, pts)
        rb, db = O.knn_bruteforce(t, k, True)
        scale = np.where(np.arange(n) % 3 == 0, 0.5, 1.5)  # a third of the seeds too small
        sb = np.floor(db[:, -1].astype(np.float64) * scale).astype(np.uint64)
        rng = np.random.default_rng(seed)
        q = tuple(rng.integers(0, 1 << 21, nq, dtype=np.uint64) for _ in range(3))
        qb = O.knn_bruteforce(t, k, queries=q)
        cases.append(dict(tree=tree, n=n, nq=nq, self_ref=(rb, db), ext_ref=qb,
                          seed=ft.DeviceArray.from_host(sb),
                          q=tuple(ft.DeviceArray.from_host(a) for a in q),
                          stats=ft.DeviceArray(8, np.uint64).zero(),
                          out=[ft.DeviceArray(m * k, np.uint64) for m in (n, n, nq, nq)],
                          stream=ft.Stream()))
    for rep in range(2):  # the second round reuses each stream's grown buffers
        for c in cases:  # interleaved: both streams' work overlaps
            ft.self_knn_query(c["tree"], k, results=c["out"][0], dists=c["out"][1],
                              seed_bound=c["seed"], stats=c["stats"], max_k=256,
                              stream=c["stream"])
        for c in cases:
            ft.knn_query(c["tree"], *c["q"], k, results=c["out"][2], dists=c["out"][3],
                         stream=c["stream"])
        for c in cases:
            c["stream"].synchronize()
            n, nq = c["n"], c["nq"]
            r, d, qr, qd = (a.to_host() for a in c["out"])
            assert np.array_equal(r.reshape(n, k), c["self_ref"][0]), (rep, n)
            assert np.array_equal(d.reshape(n, k), c["self_ref"][1]), (rep, n)
            assert np.array_equal(qr.reshape(nq, k), c["ext_ref"][0]), (rep, n)
            assert np.array_equal(qd.reshape(nq, k), c["ext_ref"][1]), (rep, n)
            assert c["stats"].to_host()[3] > 0  # re-runs happened on this stream
    # after ft_stream_release the next query on the stream makes its buffers again
    c = cases[0]
    ft._check(ft.lib().ft_stream_release(c["stream"].handle))
    ft.self_knn_query(c["tree"], k, results=c["out"][0], dists=c["out"][1], seed_bound=c["seed"],
                      max_k=256, stream=c["stream"])
    ft.knn_query(c["tree"], *c["q"], k, results=c["out"][2], dists=c["out"][3],
                 stream=c["stream"])
    c["stream"].synchronize()
    assert np.array_equal(c["out"][1].to_host().reshape(c["n"], k), c["self_ref"][1])
    assert np.array_equal(c["out"][2].to_host().reshape(c["nq"], k), c["ext_ref"][0])


def test_graph_replay_after_stream_buffers_grow(ft, orc):
    """A seeded self-kNN (re-run buffers) + external kNN (query-order buffers)
    step captured on a stream at small sizes; then larger eager queries on the
    same stream grow both buffer sets, and ft_stream_release drops them.  Every
    replay still equals the brute force: the graph keeps the blocks it was
    captured with until it is destroyed."""
    k = 16
    pts = nfw_points(30000, 41, O.CT_U64)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, pts)
    n = 30000
    rb, db = O.knn_bruteforce(t, k, True)
    scale = np.where(np.arange(n) % 3 == 0, 0.5, 1.5)  # a third of the seeds too small
    seed = ft.DeviceArray.from_host(np.floor(db[:, -1].astype(np.float64) * scale).astype(np.uint64))
    rng = np.random.default_rng(41)
    nq = 2000
    q = tuple(rng.integers(0, 1 << 21, nq, dtype=np.uint64) for _ in range(3))
    qr_b, qd_b = O.knn_bruteforce(t, k, queries=q)
    dq = tuple(ft.DeviceArray.from_host(a) for a in q)
    out = [ft.DeviceArray(m * k, np.uint64) for m in (n, n, nq, nq)]
    s = ft.Stream()

    def step():
        ft.self_knn_query(tree, k, results=out[0], dists=out[1], seed_bound=seed, max_k=256,
                          stream=s)
        ft.knn_query(tree, *dq, k, results=out[2], dists=out[3], stream=s)

    def check(tag):
        for a in out:
            a.zero()
        g.launch()
        s.synchronize()
        r, d, qr, qd = (a.to_host() for a in out)
        assert np.array_equal(r.reshape(n, k), rb) and np.array_equal(d.reshape(n, k), db), tag
        assert np.array_equal(qr.reshape(nq, k), qr_b) and np.array_equal(qd.reshape(nq, k), qd_b), tag

    step()  # eager: sizes the stream's buffers
    s.synchronize()
    with ft.Graph(s) as g:
        step()
    check("first replay")
    # larger eager queries on the same stream: both buffer sets grow
    big = nfw_points(120000, 42, O.CT_U64)
    tree2, t2 = make(ft, O.CT_U64, O.SFC_PH, 21, 1, big)
    r2, d2 = ft.self_knn_query(tree2, k, stream=s)
    bq = tuple(rng.integers(0, 1 << 21, 90000, dtype=np.uint64) for _ in range(3))
    ft.knn_query(tree2, *(ft.DeviceArray.from_host(a) for a in bq), k, stream=s)
    s.synchronize()
    check("replay after growth")
    ft._check(ft.lib().ft_stream_release(s.handle))
    check("replay after ft_stream_release")
    sample = np.arange(0, 120000, 997)
    rb2, db2 = O.knn_bruteforce(t2, k, True, leaf_ids=sample.astype(np.int32))
    assert np.array_equal(d2[sample], db2) and np.array_equal(r2[sample], rb2)
    del g  # frees the retired blocks; the stream keeps working
    step()
    s.synchronize()
    assert np.array_equal(out[1].to_host().reshape(n, k), db)
    assert np.array_equal(out[2].to_host().reshape(nq, k), qr_b)


def test_reference_mode_tie_picks_by_hand(ft, orc):
    """Reference mode on the GPU against literal arrays derived by hand from the
    reference source (the derivations are in tests/test_oracle.py): the
    per-query walk with the max-heap tier (8 points on a line, query x = 7,
    k = 4: [4, 3, 2, 5], canonical [3, 4, 2, 5]) and the SortedMergeHeap tier
    of knn_query k > 32 (34 points, query x = 50, k = 33)."""
    import test_oracle as TO
    x, y, z = TO._line_tree(range(0, 16, 2))
    tree, _ = make(ft, O.CT_U64, O.SFC_MORTON, 21, 0, (x, y, z))
    q = (np.array([7], np.uint64), np.zeros(1, np.uint64), np.zeros(1, np.uint64))
    r, d = ft.knn_query(tree, *q, 4, mode=ft.MODE_REFERENCE)
    assert list(r[0]) == [4, 3, 2, 5] and list(d[0]) == [1, 1, 9, 9]
    r, d = ft.knn_query(tree, *q, 4)
    assert list(r[0]) == [3, 4, 2, 5] and list(d[0]) == [1, 1, 9, 9]
    pts, q, _ = TO.merge_tier_case()
    tree, _ = make(ft, O.CT_U64, O.SFC_MORTON, 21, 0, pts)
    r, d = ft.knn_query(tree, *q, 33, mode=ft.MODE_REFERENCE, max_k=256)
    assert list(r[0]) == TO.MERGE_TIER_IDX
    r, d = ft.knn_query(tree, *q, 33)  # canonical: ties by index
    assert np.array_equal(np.lexsort((r[0], d[0])), np.arange(33))


def test_external_knn_reference_merge_tier(ft, orc):
    """knn_query k > 32 in reference mode: SortedMergeHeap tie order."""
    pts = ref_points(O.GEN_UNIFORM, 1000, 42, O.CT_U64)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 0, pts)
    x, y, z = pts
    q = (x[::31].copy(), y[::31].copy(), z[::31].copy())
    for k in (33, 64, 128):
        r, d = ft.knn_query(tree, *q, k, mode=ft.MODE_REFERENCE)
        rr, dr = O.knn_reference(t, k, queries=q, kind=1, max_k=256)
        assert np.array_equal(d, dr) and np.array_equal(r, rr), f"k={k}"


def test_return_orig_indices(ft, orc):
    x, y, z = nfw_points(5000, 8, O.CT_F64)
    cfg = ft.Config(sfc=O.SFC_PH, bits_per_dim=21, return_orig_indices=True)
    tree = ft.TreeSoA(len(x), O.CT_F64, cfg)
    ft.build_bvh(ft.particles(x, y, z), tree)
    t = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_PH, bits=21, orig_indices=1)
    for algo in ("packet", "grouped"):
        r, d = ft.self_knn_query(tree, 12, algorithm=ALGOS[algo])
        rb, db = O.knn_bruteforce(t, 12, True)
        assert np.array_equal(d, db) and np.array_equal(r, rb)


@pytest.mark.parametrize("n", [1, 2, 7, 31, 33, 63, 65, 99])
def test_edge_sizes_and_sentinels(ft, orc, n):
    """Edge cases of test/test_self_knn.cpp:876-1003: n=1, n=2, partial wavefronts;
    k > n-1 leaves missing neighbours (index UINT64_MAX, -DBL_MAX / 0)."""
    for ct in (O.CT_F64, O.CT_U64):
        x, y, z = ref_points(O.GEN_UNIFORM, n, 777, ct)
        tree, t = make(ft, ct, O.SFC_PH, 21, 0, (x, y, z))
        for k in (1, 2, 4, 32, 33, 64):
            for excl in (True, False):
                for algo in ("packet", "grouped"):
                    r, d = ft.self_knn_query(tree, k, exclude_self=excl, max_k=256,
                                             algorithm=ALGOS[algo])
                    rb, db = O.knn_bruteforce(t, k, excl)
                    assert np.array_equal(d, db) and np.array_equal(r, rb), (ct, k, excl, algo)
                r, d = ft.self_knn_query(tree, k, exclude_self=excl, max_k=256,
                                         mode=ft.MODE_REFERENCE)
                rr, dr = O.knn_reference(t, k, excl, max_k=256)
                assert np.array_equal(d, dr) and np.array_equal(r, rr)


@pytest.mark.parametrize("n", [1, 2, 3, 50, 65, 130, 1000])
def test_windows_on_small_trees_subset_and_external(ft, orc, n):
    """The packet walk's window pre-scan on trees smaller than or near one
    64-leaf run: subset queries (window around the middle lane's rank, widened
    for sparse subsets) and external queries (window around the leaf a greedy
    descent reaches), with full, partial and single-query packets."""
    rng = np.random.default_rng(n)
    for ct in (O.CT_U64, O.CT_F64):
        x, y, z = ref_points(O.GEN_UNIFORM, n, 4242 + n, ct)
        periodic = 1 if ct == O.CT_U64 else 0
        tree, t = make(ft, ct, O.SFC_PH, 21, periodic, (x, y, z))
        for m in sorted({1, max(1, n // 3), n, 150}):
            ids = rng.integers(0, n, m).astype(np.int32)
            for k in (1, 8, 40):
                r, d = ft.self_knn_query_subset(tree, ids, k, algorithm=ALGOS["packet"])
                rb, db = O.knn_bruteforce(t, k, True, leaf_ids=ids)
                assert np.array_equal(d, db) and np.array_equal(r, rb), (ct, m, k, "subset")
        q = tuple(a[rng.integers(0, n, 97)].copy() for a in (x, y, z))
        for k in (1, 8, 40):
            r, d = ft.knn_query(tree, *q, k)
            rb, db = O.knn_bruteforce(t, k, queries=q)
            assert np.array_equal(d, db) and np.array_equal(r, rb), (ct, k, "external")


def test_large_sample_against_bruteforce(ft, orc):
    """300k clustered int64 periodic particles, k=32: all queries on the GPU,
    a 3000-query sample checked against brute force."""
    pts = nfw_points(300_000, 13, O.CT_U64, halos=2000)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, pts)
    ids = np.random.default_rng(5).choice(300_000, 3000, replace=False).astype(np.int32)
    rb, db = O.knn_bruteforce(t, 32, True, leaf_ids=ids)
    for algo in ("packet", "packet-heap", "grouped", "plain"):
        r, d = ft.self_knn_query(tree, 32, algorithm=ALGOS[algo], topk=TOPK.get(algo, 0))
        assert np.array_equal(d[ids], db) and np.array_equal(r[ids], rb), algo


def test_seeded_bound_never_changes_results(ft, orc):
    """A per-query seed bound (too small for some queries, loose for others)
    gives exactly the unbounded results; too-small seeds are re-run."""
    pts = nfw_points(40000, 17, O.CT_U64)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, pts)
    for k in (16, 64):  # register tier, LDS heap tier
        rb, db = O.knn_bruteforce(t, k, True)
        true_r2 = db[:, -1].astype(np.float64)
        rng = np.random.default_rng(2)
        scale = rng.choice([0.25, 0.9, 1.0, 1.5, 4.0], size=len(true_r2))
        seed = np.floor(true_r2 * scale).astype(np.uint64)
        stats = ft.DeviceArray(8, np.uint64).zero()
        r, d = ft.self_knn_query(tree, k, seed_bound=ft.DeviceArray.from_host(seed), stats=stats,
                                 max_k=256)
        assert np.array_equal(d, db) and np.array_equal(r, rb), k
        st = stats.to_host()
        assert st[3] == int(np.sum(seed < db[:, -1]))  # exactly the too-small seeds re-ran


def test_seeded_bound_fp_periodic_subset_and_self(ft, orc):
    """Seeds on the SPH shape at test size (double coordinates, periodic box,
    Peano-Hilbert): a random 20% subset at k = 64 (passes of 32) and every leaf
    at k = 32, each query seeded too small, exactly right or loose: results
    equal the brute force and exactly the queries whose seed is below their
    k-th distance re-run."""
    n = 20000
    pts = nfw_points(n, 23, O.CT_F64)
    tree, t = make(ft, O.CT_F64, O.SFC_PH, 21, 1, pts, box=100.0)
    rng = np.random.default_rng(4)
    ids = rng.permutation(n)[: n // 5].astype(np.int32)
    for k, leaf_ids in ((64, ids), (32, None)):
        rb, db = O.knn_bruteforce(t, k, True, leaf_ids=leaf_ids)
        scale = rng.choice([0.25, 0.9, 1.0, 1.5, 4.0], size=len(db))
        seed = db[:, -1] * scale
        stats = ft.DeviceArray(8, np.uint64).zero()
        sd = ft.DeviceArray.from_host(seed)
        if leaf_ids is None:
            r, d = ft.self_knn_query(tree, k, seed_bound=sd, stats=stats)
        else:
            r, d = ft.self_knn_query_subset(tree, ids, k, seed_bound=sd, stats=stats)
        assert np.array_equal(d, db) and np.array_equal(r, rb), k
        assert stats.to_host()[3] == int(np.sum(seed < db[:, -1])), k


def comb_points(ct):
    """63 points whose Morton keys are the 63 powers of two below 2^63, plus the
    origin: the radix tree is a comb 63 levels deep, deeper than the
    reference's 64-entry per-lane stack can hold."""
    pts = [(0, 0, 0)]
    for i in range(21):
        pts += [(1 << i, 0, 0), (0, 1 << i, 0), (0, 0, 1 << i)]
    a = np.array(pts, dtype=O.COORD_DTYPES[ct])
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


@pytest.mark.parametrize("ct", [O.CT_U64, O.CT_U32])
def test_deep_comb_tree(ft, orc, ct):
    """Canonical walks stay exact on a tree deeper than the 64-entry stack (the
    overflowing queries are re-run with a deeper stack); reference mode drops
    the deep children exactly as the oracle's restatement of the reference."""
    pts = comb_points(ct)
    tree, t = make(ft, ct, O.SFC_MORTON, 21, 0, pts)
    for k in (1, 4, 33):
        rb, db = O.knn_bruteforce(t, k, True)
        for algo in ("plain", "grouped", "packet"):
            r, d = ft.self_knn_query(tree, k, max_k=256, algorithm=ALGOS[algo])
            assert np.array_equal(d, db) and np.array_equal(r, rb), (algo, k)
        rr, dr = O.knn_reference(t, k, True, max_k=256)
        r, d = ft.self_knn_query(tree, k, max_k=256, mode=ft.MODE_REFERENCE)
        assert np.array_equal(d, dr) and np.array_equal(r, rr), ("reference", k)
