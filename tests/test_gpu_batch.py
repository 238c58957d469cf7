"""GPU checks of the batched multi-tree path (ft_batch_*): every request's
results equal knn_query's canonical results on its own tree (which the kNN
tests pin to the oracle's brute force), a sample is checked against the oracle
directly, and graph replays -- after retuning, with other GPU work between
them, and after the query points change -- keep giving those results."""
import numpy as np
import pytest

import oracle as O
from fasttree_amd import synth

pytestmark = pytest.mark.gpu


def make_batch(ft, ct, sizes, ks, seed, periodic=False):
    from fasttree_amd.batch import Batch, plan_requests
    rng = np.random.default_rng(seed)
    dt = O.COORD_DTYPES[ct]
    fp = ct in (O.CT_F32, O.CT_F64)
    cfg = ft.Config(sfc=ft.SFC_MORTON, bits_per_dim=21, periodic=periodic,
                    box_length=100.0 if (fp and periodic) else 0.0)
    trees, parts, host = [], [], []
    for i, n in enumerate(sizes):
        if n >= 64:
            x, y, z = synth.nfw_clustered(n, seed * 100 + i, box=100.0, n_halos=max(2, n // 2000))
        else:
            x, y, z = synth.uniform(n, seed * 100 + i, 100.0)
        if fp:
            pts = tuple(a.astype(dt) for a in (x, y, z))
        else:
            pts = tuple(synth.to_fixed(a, 100.0, 21, dt) for a in (x, y, z))
        host.append(pts)
        trees.append(ft.TreeSoA(max(n, 1), ct, cfg))
        parts.append(tuple(ft.DeviceArray.from_host(a) for a in pts))
    # two requests per tree (one may be empty), k from the mixed set
    nq, kk, tt = [], [], []
    for i in range(len(sizes)):
        for _ in range(2):
            nq.append(int(rng.choice([0, 1, 63, 64, 65, 700, 2500])))
            kk.append(int(rng.choice(ks)))
            tt.append(i)
    reqs, total_q, total_r = plan_requests(nq, kk, tt)
    # queries: half near tree points, half uniform in the box (random order)
    qs = []
    for (t, k, qf, n, rf) in reqs:
        src = host[t]
        m = len(src[0])
        pick = rng.integers(0, m, n)
        near = [src[a][pick].astype(np.float64) + rng.normal(0, 0.3 if fp else 2000.0, n)
                for a in range(3)]
        far = [rng.uniform(0, 100.0 if fp else float(1 << 21), n) for _ in range(3)]
        sel = rng.random(n) < 0.5
        q = [np.where(sel, near[a], far[a]) for a in range(3)]
        if fp:
            q = [a.astype(dt) for a in q]
        else:
            q = [np.clip(a, 0, (1 << 21) - 1).astype(dt) for a in q]
        qs.append(q)
    qx = np.concatenate([q[0] for q in qs]) if qs else np.zeros(0, dt)
    qy = np.concatenate([q[1] for q in qs]) if qs else np.zeros(0, dt)
    qz = np.concatenate([q[2] for q in qs]) if qs else np.zeros(0, dt)
    dq = [ft.DeviceArray.from_host(a) for a in (qx, qy, qz)]
    res = ft.DeviceArray(max(total_r, 1), np.uint64)
    dst = ft.DeviceArray(max(total_r, 1), ft.dist_dtype(ct))
    b = Batch(ct)
    b.set_trees(trees, parts)
    b.set_queries(reqs, *dq, res, dst)
    return b, trees, reqs, (qx, qy, qz), dq, res, dst, host


def expected(ft, trees, reqs, q):
    out = []
    for (t, k, qf, n, rf) in reqs:
        if n == 0:
            out.append(None)
            continue
        sl = slice(qf, qf + n)
        r, d = ft.knn_query(trees[t], q[0][sl], q[1][sl], q[2][sl], k)
        out.append((r, d))
    return out


def check(reqs, exp, res, dst, tag):
    r_all = res.to_host()
    d_all = dst.to_host()
    for (t, k, qf, n, rf), e in zip(reqs, exp):
        if e is None:
            continue
        r = r_all[rf:rf + n * k].reshape(n, k)
        d = d_all[rf:rf + n * k].reshape(n, k)
        assert np.array_equal(d, e[1]), f"{tag}: d2 tree={t} k={k} n={n}"
        assert np.array_equal(r, e[0]), f"{tag}: idx tree={t} k={k} n={n}"


@pytest.mark.parametrize("ct", [O.CT_U32, O.CT_F64])
def test_batch_run_tune_and_graph_replays(ft, orc, ct):
    sizes = [3000, 17000, 1, 2, 45000, 9000, 130]
    b, trees, reqs, q, dq, res, dst, host = make_batch(ft, ct, sizes, (1, 8, 32, 128), 3)
    print(f"[batch] ct={ct} eager run", flush=True)
    b.run()
    exp = expected(ft, trees, reqs, q)
    check(reqs, exp, res, dst, "eager")

    # oracle brute force on a sample of each non-empty request
    for (t, k, qf, n, rf), e in zip(reqs, exp):
        if e is None:
            continue
        ot = O.build_bvh(ct, *host[t], sfc=O.SFC_MORTON, bits=21)
        m = min(n, 40)
        sl = slice(qf, qf + m)
        rb, db = O.knn_bruteforce(ot, k, queries=(q[0][sl], q[1][sl], q[2][sl]))
        assert np.array_equal(e[1][:m], db) and np.array_equal(e[0][:m], rb), (t, k)

    print("[batch] tune", flush=True)
    b.tune()
    tun = b.tuning()
    assert tun["tuned"] and tun["groups"]
    assert set(tun["radix_bits"]) == {"<40k", "<160k", "<640k", ">=640k"}
    # 8 or 9 for the size classes the batch's trees fall in; 0 (automatic)
    # for classes with no tree
    assert all(v in (0, 8, 9) for v in tun["radix_bits"].values())
    assert any(v in (8, 9) for v in tun["radix_bits"].values())
    print(f"[batch] tuning {tun}", flush=True)
    for rep in range(3):
        res.zero()
        dst.zero()
        print(f"[batch] graph replay {rep}", flush=True)
        b.launch()
        # unrelated eager GPU work between replays
        ft.self_knn_query(trees[0], 4)
        check(reqs, exp, res, dst, f"replay {rep}")

    # new query coordinates, same layout: the graph reads the new points
    rng = np.random.default_rng(99)
    perm = rng.permutation(len(q[0]))
    q2 = tuple(np.ascontiguousarray(a[perm]) for a in q)
    for dev, host_new in zip(dq, q2):
        dev.copy_from(host_new)
    b.launch()
    check(reqs, expected(ft, trees, reqs, q2), res, dst, "moved queries")


def test_batch_queries_only_on_prebuilt_trees(ft, orc):
    """Trees built beforehand (no build inputs): the step only queries."""
    from fasttree_amd.batch import Batch, plan_requests
    ct = O.CT_U64
    cfg = ft.Config(sfc=ft.SFC_PEANO_HILBERT, bits_per_dim=21, periodic=True)
    trees, hosts = [], []
    for i, n in enumerate((5000, 20000)):
        pts = synth.clustered_fixed(n, 40 + i, n_halos=20)
        t = ft.TreeSoA(n, ct, cfg)
        ft.build_bvh(ft.particles(*pts), t)
        trees.append(t)
        hosts.append(pts)
    reqs, total_q, total_r = plan_requests([300, 500, 200], [16, 64, 1], [0, 1, 1])
    rng = np.random.default_rng(5)
    q = tuple(rng.integers(0, 1 << 21, total_q, dtype=np.uint64) for _ in range(3))
    dq = [ft.DeviceArray.from_host(a) for a in q]
    res = ft.DeviceArray(total_r, np.uint64)
    dst = ft.DeviceArray(total_r, np.uint64)
    stats = ft.DeviceArray(8, np.uint64).zero()
    b = Batch(ct)
    b.set_trees(trees)
    b.set_queries(reqs, *dq, res, dst)
    b.set_stats(stats)
    b.launch()
    check(reqs, expected(ft, trees, reqs, q), res, dst, "prebuilt")
    assert stats.to_host()[0] > 0


def _tiers_for(kt, fp):
    """The tier kernels that serve k tier kt (k up to 1, 4, 8, 16, 32, 64, 128,
    256), as the batch's tuner offers them."""
    cap = (1, 4, 8, 16, 32, 64, 128, 256)[kt]
    t = []
    if cap <= 32:
        t.append(f"packet-reg{cap}")
    elif cap <= 128:
        t.append("packet-reg32-passes")
    if not fp and cap <= 128:
        t.append(f"packet-heap{max(cap, 32)}")
    t.append("lane-heap")
    if cap > 32:
        t.append(f"wave-merge{cap}")
    return t


@pytest.mark.parametrize("ct", [O.CT_U32, O.CT_F64])
def test_batch_pinned_tiers_every_size_class(ft, orc, ct):
    """Every tier kernel, pinned (ft_batch_set_tier) for every (size class, k
    tier) group it serves, at several leaf buckets and queries per item,
    returns knn_query's results -- on trees in all four size classes (below
    40k, 160k, 640k leaves and above), which the tuner alone may never run."""
    from fasttree_amd.batch import Batch, plan_requests
    fp = ct == O.CT_F64
    dt = O.COORD_DTYPES[ct]
    cfg = ft.Config(sfc=ft.SFC_MORTON, bits_per_dim=21)
    rng = np.random.default_rng(8)
    sizes = (20000, 100000, 300000, 700000)
    ks = (1, 3, 8, 13, 32, 50, 128, 200)  # one per k tier
    trees, parts = [], []
    for i, n in enumerate(sizes):
        x, y, z = synth.nfw_clustered(n, 500 + i, box=100.0, n_halos=max(2, n // 2000))
        pts = tuple(a.astype(dt) if fp else synth.to_fixed(a, 100.0, 21, dt) for a in (x, y, z))
        trees.append(ft.TreeSoA(n, ct, cfg))
        parts.append(tuple(ft.DeviceArray.from_host(a) for a in pts))
    nq, kk, tt = [], [], []
    for i in range(len(sizes)):
        for k in ks:
            nq.append(int(rng.choice([1, 65, 300])))
            kk.append(k)
            tt.append(i)
    reqs, total_q, total_r = plan_requests(nq, kk, tt)
    hi = 100.0 if fp else float(1 << 21)
    q = tuple(rng.uniform(0, hi, total_q).astype(dt) for _ in range(3))
    dq = [ft.DeviceArray.from_host(a) for a in q]
    res = ft.DeviceArray(total_r, np.uint64)
    dst = ft.DeviceArray(total_r, ft.dist_dtype(ct))
    b = Batch(ct)
    b.set_trees(trees, parts)
    b.set_queries(reqs, *dq, res, dst)
    b.run()  # builds the trees
    exp = expected(ft, trees, reqs, q)
    tiers = [_tiers_for(kt, fp) for kt in range(8)]
    for rnd in range(max(len(t) for t in tiers)):
        for c in range(4):
            for kt in range(8):
                name = tiers[kt][(rnd + c) % len(tiers[kt])]
                if name.startswith("packet"):
                    b.set_tier(c, kt, name, bucket=(16, 32, 64)[(rnd + kt) % 3],
                               queries_per_item=(64, 32, 16)[(rnd + c) % 3])
                elif name.startswith("wave"):
                    b.set_tier(c, kt, name, queries_per_item=(8, 1, 64)[(rnd + c) % 3])
                else:
                    b.set_tier(c, kt, name)
        res.zero()
        dst.zero()
        print(f"[batch] pinned round {rnd}: {b.tuning()['groups'][:3]}", flush=True)
        b.run()
        check(reqs, exp, res, dst, f"pinned round {rnd}")
    with pytest.raises(ValueError):
        b.set_tier(0, 2, "packet-reg32-passes")  # does not serve k <= 8
    with pytest.raises(ValueError):
        b.set_tier(0, 6, "wave-merge128", queries_per_item=65)
    # the tuner keeps pinned groups
    b.set_tier(3, 6, "wave-merge128")
    b.tune()
    g = [x for x in b.tuning()["groups"] if x["n_class"] == ">=640k" and x["k_max"] == 128]
    assert g and g[0]["tier"] == "wave-merge128"
    res.zero()
    b.launch()
    check(reqs, exp, res, dst, "tuned with a pinned group")


SEG_CASES = [
    ("u32-morton", O.CT_U32, 0, 21, 0),
    ("u64-ph-periodic", O.CT_U64, 1, 21, 1),
    ("f64-ph", O.CT_F64, 1, 21, 0),
    ("f32-morton10", O.CT_F32, 0, 10, 0),
]


@pytest.mark.parametrize("case", SEG_CASES, ids=[c[0] for c in SEG_CASES])
def test_segmented_build_equals_build_bvh(ft, orc, case):
    """The batch's segmented build (every tree in one launch of each build
    kernel: per-tree sort tiles, digit bases and look-back chains, per-tree
    assembly chunks) gives every tree exactly the arrays build_bvh gives it
    alone: keys, permutation, topology, boxes, leaf ranges -- for trees of 0,
    1 and 2 points, at and around the 512-leaf chunk and 6,144-key tile
    boundaries, and large ones."""
    from fasttree_amd.batch import Batch, plan_requests
    name, ct, sfc, bits, periodic = case
    dt = O.COORD_DTYPES[ct]
    fp = ct in (O.CT_F32, O.CT_F64)
    cfg = ft.Config(sfc=sfc, bits_per_dim=bits, periodic=periodic)
    sizes = [0, 1, 2, 511, 512, 513, 6143, 6144, 6145, 20000, 130000, 3]
    trees, parts, hosts = [], [], []
    for i, n in enumerate(sizes):
        if n >= 64:
            x, y, z = synth.nfw_clustered(n, 900 + i, box=100.0, n_halos=max(2, n // 3000))
        else:
            x, y, z = synth.uniform(max(n, 1), 900 + i, 100.0)
            x, y, z = x[:n], y[:n], z[:n]
        pts = tuple(a.astype(dt) if fp else synth.to_fixed(a, 100.0, bits, dt) for a in (x, y, z))
        if n and not fp:  # duplicate keys: equal points
            for a in pts:
                a[n // 2] = a[0]
        hosts.append(pts)
        trees.append(ft.TreeSoA(max(n, 1), ct, cfg))
        parts.append(tuple(ft.DeviceArray.from_host(a) for a in pts))
    reqs, total_q, total_r = plan_requests([5 if n else 0 for n in sizes], [4] * len(sizes),
                                           list(range(len(sizes))))
    rng = np.random.default_rng(1)
    q = tuple((rng.uniform(0, 100.0, total_q) if fp else
               rng.integers(0, 1 << bits, total_q)).astype(dt) for _ in range(3))
    dq = [ft.DeviceArray.from_host(a) for a in q]
    res = ft.DeviceArray(total_r, np.uint64)
    dst = ft.DeviceArray(total_r, ft.dist_dtype(ct))
    b = Batch(ct)
    b.set_trees(trees, parts)
    b.set_queries(reqs, *dq, res, dst)
    b.run()
    assert b.tuning()["build"] == "segmented"
    for i, n in enumerate(sizes):
        if n == 0:
            continue
        ref = ft.TreeSoA(n, ct, cfg)
        ft.build_bvh(ft.particles(*(ft.DeviceArray.from_host(a) for a in hosts[i])), ref)
        got, want = trees[i].to_host(), ref.to_host()
        leaves = slice(n - 1, 2 * n - 1)
        for key in want:
            g, w = got[key], want[key]
            if key in ("id", "is_ghost", "orig_idx"):
                g, w = g[leaves], w[leaves]
            assert np.array_equal(g, w), f"{name} tree {i} (n={n}): {key}"
    check(reqs, expected(ft, trees, reqs, q), res, dst, f"segmented {name}")
