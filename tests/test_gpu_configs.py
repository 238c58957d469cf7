"""GPU parity at the benchmark configurations' own shapes.

Plain method, small: N = 100,000 uniform double points in [0,1)^3 (seeded),
non-periodic, 30-bit Morton keys, self_knn_query k = 16 exclude_self.

SPH/Voronoi mixed query: N = 20M clustered double points in a periodic box
L = 100, 63-bit Peano-Hilbert keys; (a) self_knn_query_subset over a seeded
random 20% of the leaves at k = 64, (b) variable-radius range query with
h_i = 1.2 x the k = 32 neighbour distance, CSR output.

Full-size results are checked through size-independent properties (two
different kernels agree, CSR equals the fixed-stride output) and against the
oracle's brute force on a seeded sample of queries.
"""
import numpy as np
import pytest

import oracle as O
from fasttree_amd import synth

pytestmark = pytest.mark.gpu


def oracle_view(ft, tree, ct, sfc, bits, periodic, box=0.0):
    """An oracle Tree over the GPU tree's arrays (brute force needs only the
    leaves and the metric)."""
    g = tree.to_host()
    t = O.Tree(ct=ct, n=tree.num_leaves, sfc=sfc, bits=bits, periodic=periodic, box=box)
    t.arrays = {"min_x": g["min_x"], "max_x": g["max_x"], "min_y": g["min_y"],
                "max_y": g["max_y"], "min_z": g["min_z"], "max_z": g["max_z"],
                "left": g["left_child"], "right": g["right_child"], "parent": g["parent"],
                "id": g["id"], "ghost": g["is_ghost"], "orig": g["orig_idx"]}
    return t


def test_plain_small_uniform_double_morton10(ft, orc):
    n, k = 100_000, 16
    x, y, z = synth.uniform(n, 1, 1.0)
    cfg = ft.Config(sfc=ft.SFC_MORTON, bits_per_dim=10, periodic=False)
    tree = ft.TreeSoA(n, ft.COORD_F64, cfg)
    ft.build_bvh(ft.particles(x, y, z), tree)
    t = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_MORTON, bits=10)
    g = tree.to_host()
    assert np.array_equal(g["min_x"], t.min_x) and np.array_equal(g["left_child"], t.left)
    r, d = ft.self_knn_query(tree, k)
    ids = np.random.default_rng(7).choice(n, 4000, replace=False).astype(np.int32)
    rb, db = O.knn_bruteforce(t, k, True, leaf_ids=ids)
    assert np.array_equal(d[ids], db) and np.array_equal(r[ids], rb)
    rr, dr = ft.self_knn_query(tree, k, mode=ft.MODE_REFERENCE)
    er, ed = O.knn_reference(t, k, True, leaf_ids=ids, max_k=32)
    assert np.array_equal(dr[ids], ed) and np.array_equal(rr[ids], er)
    assert np.array_equal(dr, d)  # the d2 arrays are unique whatever the tie picks


def test_sph_mixed_query_20m_periodic_double(ft, orc):
    n, L = 20_000_000, 100.0
    x, y, z = synth.nfw_clustered(n, 21, box=L, n_halos=4000)
    cfg = ft.Config(sfc=ft.SFC_PEANO_HILBERT, bits_per_dim=21, periodic=True, box_length=L)
    tree = ft.TreeSoA(n, ft.COORD_F64, cfg)
    ft.build_bvh(ft.particles(*(ft.DeviceArray.from_host(a) for a in (x, y, z))), tree)
    del x, y, z
    t = oracle_view(ft, tree, O.CT_F64, O.SFC_PH, 21, 1, L)
    rng = np.random.default_rng(5)

    # (a) 20% subset at k = 64: the grouped and plain walks agree everywhere,
    # and a sample matches brute force.
    sub = np.sort(rng.permutation(n)[: n // 5]).astype(np.int32)
    rng.shuffle(sub)
    dsub = ft.DeviceArray.from_host(sub)
    k = 64
    nq = len(sub)
    ra, da = (ft.DeviceArray(nq * k, np.uint64), ft.DeviceArray(nq * k, np.float64))
    rp, dp = (ft.DeviceArray(nq * k, np.uint64), ft.DeviceArray(nq * k, np.float64))
    ft.self_knn_query_subset(tree, dsub, k, results=ra, dists=da, algorithm=ft.ALGO_GROUPED)
    ft.self_knn_query_subset(tree, dsub, k, results=rp, dists=dp, algorithm=ft.ALGO_PLAIN)
    ha, hp = ra.to_host(), rp.to_host()
    assert np.array_equal(ha, hp)
    assert np.array_equal(da.to_host(), dp.to_host())
    pick = rng.choice(nq, 60, replace=False)
    rb, db = O.knn_bruteforce(t, k, True, leaf_ids=sub[pick])
    ha = ha.reshape(nq, k)
    hd = da.to_host().reshape(nq, k)
    assert np.array_equal(hd[pick], db) and np.array_equal(ha[pick], rb)
    del ra, da, rp, dp, ha, hp, hd

    # (b) h_i = 1.2 x the k = 32 neighbour distance; CSR equals fixed stride,
    # a sample matches brute force.
    q_ids = np.arange(0, n, 20, dtype=np.int32)  # 1M query particles
    _, d32 = ft.self_knn_query_subset(tree, q_ids, 32)
    d32h = d32[:, 31]
    lx, ly, lz = t.leaf_coords()
    qx, qy, qz = lx[q_ids].copy(), ly[q_ids].copy(), lz[q_ids].copy()
    h = 1.2 * np.sqrt(d32h)
    rp_csr, cols = ft.range_query_csr(tree, qx, qy, qz, np.zeros(len(h)), h)
    counts = np.diff(rp_csr)
    assert np.all(counts >= 33)  # itself and at least its 32 nearest neighbours
    m = 256
    res, cnt = ft.range_query(tree, qx, qy, qz, np.zeros(len(h)), h, m)
    assert np.array_equal(cnt, counts)
    for i in rng.choice(len(h), 200, replace=False):
        c = min(cnt[i], m)
        assert np.array_equal(res[i, :c], cols[rp_csr[i]:rp_csr[i] + c])
    pick = rng.choice(len(h), 30, replace=False)
    erp, ecols = O.range_csr(t, qx[pick], qy[pick], qz[pick], np.zeros(30), h[pick])
    for j, i in enumerate(pick):
        assert np.array_equal(cols[rp_csr[i]:rp_csr[i + 1]], ecols[erp[j]:erp[j + 1]])
