"""GPU checks of the learned search radius: MFMA forward and backward against
the fp64 oracle (oracle/mlp.py, itself pinned by finite differences), the
feature definition against a numpy recomputation, training on exact kNN
distances, and exact results with the learned seed."""
import numpy as np
import pytest

import oracle as O
from oracle import mlp as M
from fasttree_amd import synth

pytestmark = pytest.mark.gpu


def random_model(rng):
    p = np.zeros(M.NPARAM, np.float32)
    p[M.P_W1:M.P_B1] = (rng.standard_normal((16, 64)) * 0.3).ravel()
    p[M.P_W1 + 12 * 64:M.P_B1] = 0.0
    p[M.P_B1:M.P_W2] = rng.standard_normal(64) * 0.1
    p[M.P_W2:M.P_B2] = (rng.standard_normal((64, 64)) * 0.15).ravel()
    p[M.P_B2:M.P_W3] = rng.standard_normal(64) * 0.1
    p[M.P_W3:M.P_B3] = (rng.standard_normal((64, 64)) * 0.15).ravel()
    p[M.P_B3:M.P_W4] = rng.standard_normal(64) * 0.1
    p[M.P_W4:M.P_B4] = rng.standard_normal(64) * 0.2
    p[M.P_B4] = 0.3
    norm = np.zeros(34, np.float32)
    norm[:12] = rng.standard_normal(12)
    norm[16:28] = 1.0 / (0.5 + rng.random(12))
    norm[32], norm[33] = 2.0, 1.5
    return p, norm


def test_mlp_forward_matches_oracle(ft):
    from fasttree_amd import learned as L
    rng = np.random.default_rng(1)
    p, norm = random_model(rng)
    f = np.zeros((1000, 16), np.float32)
    f[:, :12] = rng.standard_normal((1000, 12))
    m = L.MLP(p, norm, 0.0)
    got = m.forward(ft.DeviceArray.from_host(f.ravel()), 1000)
    out, _ = M.forward(p, norm.astype(np.float64), f, rounding=True)
    exp = norm[32] + norm[33] * out
    assert np.max(np.abs(got - exp)) <= 2e-3 * (1 + np.max(np.abs(exp)))


def test_mlp_gradient_matches_oracle(ft):
    """GPU backward (MFMA, bf16 operands, fp32 accumulation) vs the fp64 oracle
    with the same bf16 weights and activations.  The GPU also rounds the
    backpropagated error to bf16 before each MFMA, hence the tolerance:
    max |difference| <= 3% of max |gradient|, cosine >= 0.999."""
    from fasttree_amd import learned as L
    rng = np.random.default_rng(2)
    p, norm = random_model(rng)
    n = 3000
    f = np.zeros((n, 16), np.float32)
    f[:, :12] = rng.standard_normal((n, 12))
    y = (rng.standard_normal(n) * 2 + 1).astype(np.float32)
    m = L.MLP(p, norm, 0.0)
    g, loss = m.grad(ft.DeviceArray.from_host(f.ravel()), ft.DeviceArray.from_host(y), n)
    el, eg = M.loss_and_grad(p, norm.astype(np.float64), f, y, rounding=True)
    assert abs(loss - el) <= 1e-3 * el
    live = np.ones(M.NPARAM, bool)
    live[M.P_W1 + 12 * 64:M.P_B1] = False
    live[M.P_B4 + 1:] = False
    gd, ed = g[live].astype(np.float64), eg[live]
    cos = gd @ ed / (np.linalg.norm(gd) * np.linalg.norm(ed))
    assert cos >= 0.999, cos
    assert np.max(np.abs(gd - ed)) <= 0.03 * np.max(np.abs(ed))


def test_features_definition(ft):
    from fasttree_amd import learned as L
    x, y, z = synth.clustered_fixed(20000, 3, n_halos=40)
    cfg = ft.Config(sfc=ft.SFC_PEANO_HILBERT, bits_per_dim=21, periodic=True)
    tree = ft.TreeSoA(len(x), ft.COORD_U64, cfg)
    ft.build_bvh(ft.particles(x, y, z), tree)
    k = 16
    ids = np.array([0, 1, 777, 12345, 19999], np.int32)
    f = L.features(tree, k, leaf_ids=ids).to_host().reshape(-1, 16)
    g = tree.to_host()
    n = len(x)
    off = n - 1
    t = O.Tree(ct=O.CT_U64, n=n, bits=21, periodic=1)
    for row, r in zip(f, ids):
        anc, lvl = r + off, 0
        for a in range(4):
            while lvl < 3 * (a + 1) and g["parent"][anc] >= 0:
                anc = g["parent"][anc]
                lvl += 1
            e = [float(g[h][anc]) - float(g[l][anc]) + 1.0 for l, h in
                 (("min_x", "max_x"), ("min_y", "max_y"), ("min_z", "max_z"))]
            cnt = g["range_last"][anc] - g["range_first"][anc] + 1 if anc < off else 1
            assert np.isclose(row[a], np.log(e[0] * e[1] * e[2]), rtol=1e-5)
            assert np.isclose(row[4 + a], np.log(cnt), rtol=1e-6)
        assert np.isclose(row[8], np.log(k))

        def d2(rr):
            rr = min(max(rr, 0), n - 1)
            a_ = np.array([g["min_x"][r + off], g["min_y"][r + off], g["min_z"][r + off]])
            b_ = np.array([g["min_x"][rr + off], g["min_y"][rr + off], g["min_z"][rr + off]])
            return float(O.point_d2(t, a_, b_))
        assert np.isclose(row[9], np.log(d2(r - 1 if r > 0 else r + 1) + 1.0), rtol=1e-5)
        assert np.isclose(row[10], np.log(d2(r + 1 if r + 1 < n else r - 1) + 1.0), rtol=1e-5)
        assert np.isclose(row[11], np.log(max(d2(r - 8), d2(r + 8)) + 1.0), rtol=1e-5)
        assert np.all(row[12:] == 0)


def test_training_learns_and_seeds_are_exact(ft):
    """A short GPU training run on exact k=16 distances of one snapshot lowers
    the loss well below the constant predictor's; seeding the traversal with
    the trained model leaves every index and distance unchanged, and what the
    network gains is measured on the same tree: with the seed as each lane's
    only prior bound, the seeded walk performs fewer top-k insertions than the
    walk without it."""
    from fasttree_amd import learned as L
    k = 16
    tree = L.snapshot_tree(200_000, 11)
    rng = np.random.default_rng(0)
    F, Y = L.examples(tree, k, 100_000, rng)
    m = L.MLP(L.init_params(3), L.norm_from(F, Y), 0.0)
    dF = ft.DeviceArray.from_host(F.ravel(), np.float32)
    dY = ft.DeviceArray.from_host(Y, np.float32)
    first = None
    for step in range(400):
        b = step % 6
        m.train_step(dF.ptr + b * 16384 * 16 * 4, dY.ptr + b * 16384 * 4, 16384, 3e-3)
        if step == 0:
            first = m.last_loss()
    last = m.last_loss()
    print(f"[learned] loss {first:.3f} -> {last:.3f}")
    assert last < 0.5 * first and last < 0.6  # standardised: constant predictor = 1.0
    pred = m.forward(dF, len(Y))
    resid = Y - pred
    m.set(margin=float(np.quantile(resid, 0.98)))
    n = tree.num_leaves
    seeds = ft.DeviceArray(n, np.uint64)
    scratch = ft.DeviceArray(n * 16, np.float32)
    m.seed(tree, k, seeds, scratch)
    # The network's gain, measured where the seed is each lane's only prior
    # bound (no window pre-scan or curve-neighbour fill): fewer insertions.
    stats = ft.DeviceArray(8, np.uint64).zero()
    r1, d1 = ft.self_knn_query(tree, k, seed_bound=seeds, stats=stats, window=-1)
    stats0 = ft.DeviceArray(8, np.uint64).zero()
    r0, d0 = ft.self_knn_query(tree, k, algorithm=ft.ALGO_PACKET, stats=stats0, window=-1)
    assert np.array_equal(d1, d0) and np.array_equal(r1, r0)
    s1, s0 = stats.to_host(), stats0.to_host()
    reruns = int(s1[3])
    print(f"[learned] re-run fraction {reruns / n:.4f}; insertions per query "
          f"{s1[4] / n:.1f} seeded vs {s0[4] / n:.1f} without the network")
    assert reruns < 0.2 * n
    assert s1[4] < 0.9 * s0[4]
    # with the default pre-scan and fill the results are the same again
    stats2 = ft.DeviceArray(8, np.uint64).zero()
    r2, d2 = ft.self_knn_query(tree, k, seed_bound=seeds, stats=stats2)
    assert np.array_equal(d2, d0) and np.array_equal(r2, r0)
    assert int(stats2.to_host()[3]) == reruns


def test_fp_model_seeds_subset_and_self_queries_exactly(ft):
    """The learned bound on double coordinates (the SPH shape at test size):
    a network trained on the GPU on exact k = 32 and k = 64 distances of one
    periodic f64 snapshot learns (loss well below the constant predictor's);
    its seeds leave the k = 64 subset query and the k = 32 self query
    bit-identical to the unseeded packet walk, with few re-runs."""
    from fasttree_amd import learned as L
    tree = L.snapshot_tree(200_000, 12, coord="f64", n_halos=40)
    assert tree.coord_type == ft.COORD_F64
    eps = L.target_eps(tree)
    assert 0 < eps < 1e-6  # 1e-12 * (root extent ~ 100)^2
    rng = np.random.default_rng(0)
    F = []
    Y = []
    for k in (32, 64):
        f, y = L.examples(tree, k, 49_152, rng)
        F.append(f)
        Y.append(y)
    F, Y = np.concatenate(F), np.concatenate(Y)
    perm = rng.permutation(len(Y))
    F, Y = F[perm], Y[perm]
    m = L.MLP(L.init_params(5), L.norm_from(F, Y), 0.0)
    dF = ft.DeviceArray.from_host(F.ravel(), np.float32)
    dY = ft.DeviceArray.from_host(Y, np.float32)
    first = None
    for step in range(400):
        b = step % 6
        m.train_step(dF.ptr + b * 16384 * 16 * 4, dY.ptr + b * 16384 * 4, 16384, 3e-3)
        if step == 0:
            first = m.last_loss()
    last = m.last_loss()
    print(f"[learned f64] loss {first:.3f} -> {last:.3f}")
    assert last < 0.5 * first and last < 0.6
    resid = Y - m.forward(dF, len(Y))
    m.set(margin=float(np.quantile(resid, 0.98)))
    n = tree.num_leaves
    ids = rng.permutation(n)[: n // 5].astype(np.int32)
    dids = ft.DeviceArray.from_host(ids)
    for k, sub in ((64, dids), (32, None)):
        nq = n if sub is None else len(ids)
        seeds = ft.DeviceArray(nq, np.float64)
        scratch = ft.DeviceArray(nq * 16, np.float32)
        m.seed(tree, k, seeds, scratch, leaf_ids=sub)
        stats = ft.DeviceArray(8, np.uint64).zero()
        if sub is None:
            r1, d1 = ft.self_knn_query(tree, k, seed_bound=seeds, stats=stats)
            r0, d0 = ft.self_knn_query(tree, k, algorithm=ft.ALGO_PACKET)
        else:
            r1, d1 = ft.self_knn_query_subset(tree, sub, k, seed_bound=seeds, stats=stats)
            r0, d0 = ft.self_knn_query_subset(tree, sub, k, algorithm=ft.ALGO_PACKET)
        assert np.array_equal(d1, d0) and np.array_equal(r1, r0), k
        reruns = int(stats.to_host()[3])
        print(f"[learned f64] k={k}: re-run fraction {reruns / nq:.4f}")
        assert reruns < 0.2 * nq
