"""GPU parity of range_query (fixed stride + counts) and the CSR range query."""
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


@pytest.mark.parametrize("ct,periodic", [(O.CT_U64, 1), (O.CT_U64, 0), (O.CT_F64, 0),
                                         (O.CT_U32, 0), (O.CT_F32, 0)])
def test_range_matches_reference(ft, orc, ct, periodic):
    x, y, z = synth.nfw_clustered(20000, 3, box=100.0, n_halos=100)
    dt = O.COORD_DTYPES[ct]
    if np.issubdtype(dt, np.integer):
        pts = tuple(synth.to_fixed(a, 100.0, 21, dt) for a in (x, y, z))
        rmax = np.full(500, 1 << 14, dt)
        rmax[::3] = 1 << 16
    else:
        pts = tuple(a.astype(dt) for a in (x, y, z))
        rmax = np.full(500, 0.8, dt)
        rmax[::3] = 3.0
    tree, t = make(ft, ct, O.SFC_PH, 21, periodic, pts)
    q = tuple(a[::40].copy() for a in pts)
    rmin = np.zeros(500, dt)
    rmin[::5] = rmax[::5] // 2 if np.issubdtype(dt, np.integer) else rmax[::5] / 2
    for maxr in (0, 7, 64, 4000):
        res, cnt = ft.range_query(tree, *q, rmin, rmax, maxr)
        rr, rc = O.range_reference(t, *q, rmin, rmax, maxr)
        assert np.array_equal(cnt, rc)
        for i in range(500):
            m = min(cnt[i], maxr)
            assert np.array_equal(res[i, :m], rr[i, :m])
    rp, cols = ft.range_query_csr(tree, *q, rmin, rmax)
    erp, ecols = O.range_csr(t, *q, rmin, rmax)
    assert np.array_equal(rp, erp) and np.array_equal(cols, ecols)
    # the CSR fill copies rows of at most 128 hits from its first walk and walks
    # the longer rows again: both kinds occur here
    rows = np.diff(rp)
    assert (rows > 128).any() and (rows <= 128).any()


def test_csr_column_capacity_too_small(ft, orc):
    """A column buffer smaller than the number of hits is an error (FT_ERANGE)
    and leaves the row offsets filled."""
    import ctypes as C
    x, y, z = synth.nfw_clustered(5000, 8, box=100.0, n_halos=20)
    tree, t = make(ft, O.CT_F64, O.SFC_PH, 21, 0, (x, y, z))
    q = tuple(np.ascontiguousarray(a[::10]) for a in (x, y, z))
    nq = len(q[0])
    rmin, rmax = np.zeros(nq), np.full(nq, 2.0)
    erp, _ = O.range_csr(t, *q, rmin, rmax)
    assert erp[-1] > 10
    rp = np.zeros(nq + 1, np.int64)
    cols = np.zeros(10, np.int32)
    rc = ft.lib().ft_range_query_csr(None, C.byref(tree.c), *(a.ctypes.data for a in q),
                                     rmin.ctypes.data, rmax.ctypes.data, nq, rp.ctypes.data,
                                     cols.ctypes.data, 10)
    assert rc == ft.FT_ERANGE
    assert np.array_equal(rp, erp)


def test_range_grid_kats(ft, orc):
    """test/main.cpp TEST 2 and TEST 4 (10^3 grid; non-periodic at (52,52,52),
    r <= 20; periodic int64 at (2,50,50), r <= 10)."""
    x, y, z = O.gen_reference(O.GEN_GRID, 1000, 0, O.CT_F64)
    tree, t = make(ft, O.CT_F64, O.SFC_MORTON, 10, 0, (x, y, z))
    res, cnt = ft.range_query(tree, [52.0], [52.0], [52.0], [0.0], [20.0], 1000)
    rr, rc = O.range_reference(t, [52.0], [52.0], [52.0], [0.0], [20.0], 1000)
    assert cnt[0] == rc[0] and np.array_equal(res[0, :cnt[0]], rr[0, :rc[0]])
    xi, yi, zi = O.gen_reference(O.GEN_GRID, 1000, 0, O.CT_U64, bits=21)
    tree, t = make(ft, O.CT_U64, O.SFC_PH, 21, 1, (xi, yi, zi))
    s = float(1 << 21) / 100.0
    q = ([np.uint64(int(2 * s))], [np.uint64(int(50 * s))], [np.uint64(int(50 * s))])
    r = [np.uint64(int(10 * s))]
    res, cnt = ft.range_query(tree, *q, [np.uint64(0)], r, 1000)
    rr, rc = O.range_reference(t, *q, [np.uint64(0)], r, 1000)
    assert cnt[0] == rc[0] and cnt[0] > 0
    assert np.array_equal(res[0, :cnt[0]], rr[0, :rc[0]])


def test_fp_periodic_range_and_knn(ft, orc):
    """Double coordinates in a periodic box (no reference counterpart: the
    reference ignores PERIODIC_BC for FP).  Pinned by brute force."""
    L = 100.0
    x, y, z = synth.nfw_clustered(15000, 4, box=L, n_halos=60)
    tree, t = make(ft, O.CT_F64, O.SFC_PH, 21, 1, (x, y, z), box=L)
    r, d = ft.self_knn_query(tree, 24)
    rb, db = O.knn_bruteforce(t, 24, True)
    assert np.array_equal(d, db) and np.array_equal(r, rb)
    lx, ly, lz = t.leaf_coords()
    q = (lx[::50].copy(), ly[::50].copy(), lz[::50].copy())
    rad = np.sqrt(db[::50, -1]) * 1.2
    rp, cols = ft.range_query_csr(tree, *q, np.zeros(len(rad)), rad)
    erp, ecols = O.range_csr(t, *q, np.zeros(len(rad)), rad)
    assert np.array_equal(rp, erp) and np.array_equal(cols, ecols)
    # each query finds itself and at least its 24 nearest neighbours
    assert np.all(np.diff(rp) >= 25)


def comb_points(ct):
    """63 points whose Morton keys are the 63 powers of two below 2^63, plus the
    origin: the radix tree is a comb 63 levels deep whose every step goes left,
    so the reference's per-query stack is full at its deepest nodes and it drops
    their children."""
    pts = [(0, 0, 0)]
    for i in range(21):
        pts += [(1 << i, 0, 0), (0, 1 << i, 0), (0, 0, 1 << i)]
    a = np.array(pts, dtype=O.COORD_DTYPES[ct])
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


@pytest.mark.parametrize("ct", [O.CT_U64, O.CT_U32])
def test_range_deep_comb_tree(ft, orc, ct):
    """The shared-walk range query reproduces the reference's stack limit: on a
    tree whose deepest nodes sit below 62 left turns, the leaves the reference
    never reaches are missing from the GPU output too, in the same order."""
    pts = comb_points(ct)
    tree, t = make(ft, ct, O.SFC_MORTON, 21, 0, pts)
    dt = O.COORD_DTYPES[ct]
    x, y, z = pts
    qx = np.concatenate([x, np.array([3, 0, 1 << 20], dt)])
    qy = np.concatenate([y, np.array([0, 5, 0], dt)])
    qz = np.concatenate([z, np.array([1, 0, 0], dt)])
    nq = len(qx)
    rmax = np.full(nq, (1 << 21) - 1, dt)  # every point inside
    rmax[::4] = 4
    rmin = np.zeros(nq, dt)
    rmin[::7] = 1
    res, cnt = ft.range_query(tree, qx, qy, qz, rmin, rmax, 64)
    rr, rc = O.range_reference(t, qx, qy, qz, rmin, rmax, 64)
    assert np.array_equal(cnt, rc)
    for i in range(nq):
        m = min(cnt[i], 64)
        assert np.array_equal(res[i, :m], rr[i, :m])
    assert rc.max() == len(x) - 2  # the reference never reaches the two deepest leaves
    rp, cols = ft.range_query_csr(tree, qx, qy, qz, rmin, rmax)
    erp, ecols = O.range_csr(t, qx, qy, qz, rmin, rmax)
    assert np.array_equal(rp, erp) and np.array_equal(cols, ecols)
