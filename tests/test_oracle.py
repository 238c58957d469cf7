"""CPU tests that pin the oracle (oracle/oracle.cpp) independently of the GPU.

Each check is chosen so a plausible slip (a dropped term, a swapped axis, a
wrong tie rule, an off-by-one) fails: closed forms computed here in Python,
invariants of the curves and of radix trees, the reference's own heap
known-answer tests and test generators, and brute force on small inputs.
"""
import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- SFC keys ----
def interleave(ix, iy, iz, bits):
    k = 0
    for b in range(bits):
        k |= ((ix >> b) & 1) << (3 * b)
        k |= ((iy >> b) & 1) << (3 * b + 1)
        k |= ((iz >> b) & 1) << (3 * b + 2)
    return k


def test_morton_integer_interleave(orc):
    for bits in (10, 21):
        x = RNG.integers(0, 1 << 32, 500, dtype=np.uint64)
        y = RNG.integers(0, 1 << 32, 500, dtype=np.uint64)
        z = RNG.integers(0, 1 << 32, 500, dtype=np.uint64)
        keys = O.encode(O.CT_U64, O.SFC_MORTON, bits, x, y, z)
        m = (1 << bits) - 1
        exp = [interleave(int(a) & m, int(b) & m, int(c) & m, bits) for a, b, c in zip(x, y, z)]
        assert [int(k) for k in keys] == exp


def test_morton_fp_quantisation_and_max_quirk(orc):
    # nx = (x - min) / (max - min) in double; ix = top-B mantissa bits of fl(1 + nx).
    x = np.array([0.0, 0.25, 0.5, 1.0 - 2.0 ** -11, 1.0], np.float64)
    y = np.zeros(5)
    z = np.zeros(5)
    bb = np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    keys = O.encode(O.CT_F64, O.SFC_MORTON, 10, x, y, z, bb)
    ix = [interleave(i, 0, 0, 10) for i in (0, 256, 512, 1023, 0)]
    # the maximum clamps to nextafter(1, 0) and 1 + (1 - 2^-53) rounds to 2.0: ix = 0
    assert [int(k) for k in keys] == ix


def test_morton_fp_matches_python_double_arithmetic(orc):
    x, y, z = (RNG.random(2000) * 37.0 - 5.0 for _ in range(3))
    bb = O.bbox(O.CT_F64, x, y, z)
    keys = O.encode(O.CT_F64, O.SFC_MORTON, 10, x, y, z)
    exp = []
    for p in zip(x, y, z):
        q = []
        for a in range(3):
            lo, hi = float(bb[2 * a]), float(bb[2 * a + 1])
            inv = 1.0 / (hi - lo)
            nx = min(max((float(p[a]) - lo) * inv, 0.0), np.nextafter(1.0, 0.0))
            v = np.float64(1.0) + np.float64(nx)
            bits = int(np.array([v]).view(np.uint64)[0])
            q.append((bits & ((1 << 52) - 1)) >> (52 - 10))
        exp.append(interleave(q[0], q[1], q[2], 10))
    assert [int(k) for k in keys] == exp


def test_bbox_matches_numpy(orc):
    for ct, dt in ((O.CT_F64, np.float64), (O.CT_F32, np.float32), (O.CT_U64, np.uint64),
                   (O.CT_U32, np.uint32)):
        if np.issubdtype(dt, np.integer):
            a = [RNG.integers(0, 1 << 30, 999).astype(dt) for _ in range(3)]
        else:
            a = [(RNG.standard_normal(999) * 50).astype(dt) for _ in range(3)]
        bb = O.bbox(ct, *a)
        exp = np.array([a[0].min(), a[0].max(), a[1].min(), a[1].max(), a[2].min(), a[2].max()],
                       dt)
        assert np.array_equal(bb, exp)


def _ph_all_cells(bits):
    side = 1 << bits
    g = np.arange(side, dtype=np.uint64)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    return x, y, z, O.encode(O.CT_U64, O.SFC_PH, bits, x, y, z)


@pytest.mark.parametrize("bits", [1, 2, 3, 4])
def test_peano_hilbert_bijection_and_adjacency(orc, bits):
    x, y, z, k = _ph_all_cells(bits)
    # a bijection onto [0, 8^B)
    assert np.array_equal(np.sort(k), np.arange(1 << (3 * bits), dtype=np.uint64))
    order = np.argsort(k)
    pts = np.stack([x[order], y[order], z[order]], 1).astype(np.int64)
    # consecutive keys are face-adjacent cells (a Hilbert curve, not a Z-curve)
    step = np.abs(np.diff(pts, axis=0)).sum(1)
    assert np.all(step == 1)
    assert int(k[0]) == 0  # origin cell starts the curve


def test_peano_hilbert_levels_nest(orc):
    # The key of a cell at B bits extends the key of its parent cell at B-1 bits.
    x, y, z = (RNG.integers(0, 1 << 21, 400, dtype=np.uint64) for _ in range(3))
    k21 = O.encode(O.CT_U64, O.SFC_PH, 21, x, y, z)
    k20 = O.encode(O.CT_U64, O.SFC_PH, 20, x >> np.uint64(1), y >> np.uint64(1),
                   z >> np.uint64(1))
    assert np.array_equal(k21 >> np.uint64(3), k20)


REF_PH_HEADER = "/root/reference/src/sfc.peano_hilbert.hpp"


def _parse_table(text, name):
    import re
    m = re.search(name + r"\[48\]\[8\]\s*=\s*\{(.*?)\};", text, re.S)
    vals = [int(v) for v in re.findall(r"\d+", m.group(1))]
    assert len(vals) == 384
    return np.array(vals, np.int64).reshape(48, 8)


@pytest.mark.skipif(not __import__("os").path.exists(REF_PH_HEADER),
                    reason="reference sources not present")
def test_peano_hilbert_tables_match_reference_source(orc):
    """Orientation pin: bijection, adjacency and nesting hold for every one of
    the 48 Hilbert-curve orientations, so they cannot tell a transposed axis or a
    mis-typed row.  This parses the rotation/digit tables from the reference
    header's text (src/sfc.peano_hilbert.hpp:332-376) and runs the encode loop of
    :569-600 on them in Python: pix = bx<<2 | by<<1 | bz, most significant level
    first."""
    text = open(REF_PH_HEADER).read()
    rot, sub = _parse_table(text, "rottable3"), _parse_table(text, "subpix3")
    B = 21
    x, y, z = (RNG.integers(0, 1 << B, 300, dtype=np.uint64) for _ in range(3))
    got = O.encode(O.CT_U64, O.SFC_PH, B, x, y, z)
    for i in range(len(x)):
        key, r = 0, 0
        for b in range(B - 1, -1, -1):
            pix = (((int(x[i]) >> b) & 1) << 2) | (((int(y[i]) >> b) & 1) << 1) | \
                  ((int(z[i]) >> b) & 1)
            key = (key << 3) | int(sub[r][pix])
            r = int(rot[r][pix])
        assert int(got[i]) == key


def test_peano_hilbert_fp_quantisation(orc):
    # PH for FP coords quantises min(1 + nx, 0x1.fffffffffffffp+0): the maximum maps
    # to the last cell (no Morton-style wrap to 0).
    x = np.array([0.0, 1.0]); y = np.array([0.0, 1.0]); z = np.array([0.0, 1.0])
    bb = np.array([0.0, 1.0] * 3)
    k = O.encode(O.CT_F64, O.SFC_PH, 3, x, y, z, bb)
    ki = O.encode(O.CT_U64, O.SFC_PH, 3, np.array([0, 7], np.uint64),
                  np.array([0, 7], np.uint64), np.array([0, 7], np.uint64))
    assert np.array_equal(k, ki)


# ---------------------------------------------------------------- sort ---------
def test_sort_is_stable_by_key(orc):
    keys = RNG.integers(0, 50, 5000).astype(np.uint64)  # many equal keys
    perm, srt = O.sort(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable").astype(np.uint32))
    assert np.array_equal(srt, keys[perm])


# ---------------------------------------------------------------- Karras -------
def check_radix_tree(sorted_keys, left, right, parent, rf, rl):
    """Checks topology and labelling against the radix-tree definition over the
    (key, index)-augmented sequence -- independent of the Karras search."""
    n = len(sorted_keys)
    aug = [(int(k) << 32) | i for i, k in enumerate(sorted_keys)]
    assert parent[0] == -1
    seen = np.zeros(2 * n - 1, np.int64)
    stack = [(0, 0, n - 1)]
    while stack:
        node, f, l = stack.pop()
        if f == l:
            assert node == f + n - 1  # a leaf
            continue
        assert rf[node] == f and rl[node] == l
        # Karras labelling: internal node label is an end of its range
        assert node in (f, l) or node == 0
        hb = (aug[f] ^ aug[l]).bit_length() - 1
        split = max(j for j in range(f, l + 1) if not (aug[j] >> hb) & 1)
        assert all((aug[j] >> hb) & 1 for j in range(split + 1, l + 1))
        lc, rc = int(left[node]), int(right[node])
        exp_l = split + n - 1 if split == f else split
        exp_r = split + 1 + n - 1 if split + 1 == l else split + 1
        assert (lc, rc) == (exp_l, exp_r)
        assert parent[lc] == node and parent[rc] == node
        seen[lc] += 1
        seen[rc] += 1
        stack.append((lc, f, split))
        stack.append((rc, split + 1, l))
    assert seen[0] == 0 and np.all(seen[1:] == 1)


@pytest.mark.parametrize("n,vals", [(10, None), (1000, None), (777, 40), (513, 3), (2, 1)])
def test_karras_is_the_radix_tree(orc, n, vals):
    if vals is None:
        keys = np.arange(n, dtype=np.uint64)  # the collinear KATs of test/test_query.cpp
    else:
        keys = np.sort(RNG.integers(0, vals, n).astype(np.uint64))
    left, right, parent, rf, rl = O.karras(keys)
    check_radix_tree(keys, left, right, parent, rf, rl)


def test_karras_random_63bit_keys(orc):
    keys = np.sort(RNG.integers(0, 1 << 63, 3000, dtype=np.uint64))
    keys[100:140] = keys[100]  # a run of duplicates
    keys = np.sort(keys)
    check_radix_tree(keys, *O.karras(keys))


def test_depth_kat(orc):
    # test/test_depth.cpp: 1000 collinear leaves keyed 0..999; the radix tree over
    # distinct 10-bit keys has depth 10.
    keys = np.arange(1000, dtype=np.uint64)
    left, right, parent, rf, rl = O.karras(keys)
    depth = 0
    for leaf in range(999, 1999):
        d, c = 0, leaf
        while c != 0:
            c = parent[c]
            d += 1
        depth = max(depth, d)
    assert depth == 10


# ---------------------------------------------------------------- build --------
def leaf_box_closed_form(t):
    o = t.n - 1
    for node in range(t.n - 1):
        f, l = t.range_first[node], t.range_last[node]
        for lo, hi in (("min_x", "max_x"), ("min_y", "max_y"), ("min_z", "max_z")):
            seg = t.arrays[lo][o + f:o + l + 1]
            assert t.arrays[lo][node] == seg.min()
            assert t.arrays[hi][node] == t.arrays[hi][o + f:o + l + 1].max()


@pytest.mark.parametrize("ct,sfc,bits", [(O.CT_F64, O.SFC_MORTON, 10), (O.CT_U64, O.SFC_PH, 21),
                                         (O.CT_U32, O.SFC_MORTON, 21), (O.CT_F32, O.SFC_PH, 21)])
def test_build_closed_forms(orc, ct, sfc, bits):
    dt = O.COORD_DTYPES[ct]
    n = 3000
    if np.issubdtype(dt, np.integer):
        a = [RNG.integers(0, 1 << bits, n).astype(dt) for _ in range(3)]
        a[0][:200] = a[0][0]; a[1][:200] = a[1][0]; a[2][:200] = a[2][0]  # duplicates
    else:
        a = [RNG.random(n).astype(dt) for _ in range(3)]
    t = O.build_bvh(ct, *a, sfc=sfc, bits=bits)
    keys = O.encode(ct, sfc, bits, *a)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(t.sorted_keys, keys[perm])
    o = n - 1
    assert np.array_equal(t.orig[o:], perm.astype(np.int32))
    assert np.array_equal(t.id[o:], perm.astype(np.uint32))
    for name, src in (("min_x", a[0]), ("max_y", a[1]), ("min_z", a[2])):
        assert np.array_equal(t.arrays[name][o:], src[perm])
    check_radix_tree(t.sorted_keys, t.left, t.right, t.parent, t.range_first, t.range_last)
    leaf_box_closed_form(t)


# ---------------------------------------------------------------- distances ----
def test_integer_periodic_distance_closed_form(orc):
    B = 21
    P = 1 << B
    t = O.Tree.empty(O.CT_U64, 1, bits=B, periodic=1)
    for _ in range(300):
        a = RNG.integers(0, P, 3, dtype=np.uint64)
        b = RNG.integers(0, P, 3, dtype=np.uint64)
        exp = 0
        for i in range(3):
            d = (int(a[i]) - int(b[i])) % P
            d = min(d, P - d)
            exp += d * d
        assert int(O.point_d2(t, a, b)) == exp


def test_integer_nonperiodic_distance_closed_form(orc):
    t = O.Tree.empty(O.CT_U32, 1, bits=21, periodic=0)
    for _ in range(300):
        a = RNG.integers(0, 1 << 21, 3, dtype=np.uint32)
        b = RNG.integers(0, 1 << 21, 3, dtype=np.uint32)
        exp = sum((int(a[i]) - int(b[i])) ** 2 for i in range(3))
        assert int(O.point_d2(t, a, b)) == exp


def test_fp_distance_operation_order(orc):
    # (dx*dx + dy*dy) + dz*dz, each product and sum rounded in double.
    t = O.Tree.empty(O.CT_F64, 1)
    for _ in range(300):
        a = RNG.random(3) * 1e3
        b = RNG.random(3) * 1e3
        dx, dy, dz = (float(a[i]) - float(b[i]) for i in range(3))
        s = dx * dx
        s = s + dy * dy
        s = s + dz * dz
        assert O.point_d2(t, a, b) == s
    # float coordinates: arithmetic in float, widened to double
    t32 = O.Tree.empty(O.CT_F32, 1)
    a = np.array([1.1, 2.2, 3.3], np.float32)
    b = np.array([0.7, -4.1, 9.9], np.float32)
    d = a - b
    s = np.float32(d[0] * d[0])
    s = np.float32(s + np.float32(d[1] * d[1]))
    s = np.float32(s + np.float32(d[2] * d[2]))
    assert O.point_d2(t32, a, b) == float(s)


def test_fp_periodic_distance(orc):
    L = 100.0
    t = O.Tree.empty(O.CT_F64, 1, periodic=1, box=L)
    a = np.array([1.0, 50.0, 99.0])
    b = np.array([99.0, 50.0, 2.0])
    assert O.point_d2(t, a, b) == 2.0 * 2.0 + 0.0 + 3.0 * 3.0


# ---------------------------------------------------------------- heaps --------
def test_local_max_heap_kat(orc):
    # test/test_self_knn.cpp:381-451: per lane, d = (val*17 + lane*13) % 73 for
    # val = 100..1 with a prune check, k = 48, cap = 64; sorted output = the 48
    # smallest.
    for lane in range(32):
        vals = np.arange(100, 0, -1)
        d = ((vals * 17 + lane * 13) % 73).astype(np.uint64)
        od, oi, cnt = O.heap_u64(0, 48, 64, d, vals.astype(np.int32))
        assert cnt == 48
        assert np.array_equal(od[:48], np.sort(d)[:48])
        # each returned index carries its own distance
        assert all(int((oi[j] * 17 + lane * 13) % 73) == int(od[j]) for j in range(48))


def test_sorted_merge_heap_kat(orc):
    # test/test_self_knn.cpp:453-559: three 32-lane merge rounds, k = 48, cap = 64.
    lid = np.arange(32)
    d1 = ((31 - lid) * 10).astype(np.uint64)
    d2 = ((lid % 4) * 5).astype(np.uint64)
    d3 = np.where(lid < 16, lid + 1, np.iinfo(np.uint64).max).astype(np.uint64)
    i3 = np.where(lid < 16, 200 + lid, -1)
    d = np.concatenate([d1, d2, d3])
    idx = np.concatenate([lid, 100 + lid, i3]).astype(np.int32)
    od, oi, _ = O.heap_u64(1, 48, 64, d, idx)
    exp = np.sort(np.concatenate([d1, d2, d3[:16]]))[:48]
    assert np.array_equal(od[:48], exp)


# ---------------------------------------------------------------- generators ---
def _libstdcxx_uniform(seed, n, a, b):
    """std::uniform_real_distribution<double>(a, b) over std::mt19937(seed) as
    libstdc++ implements it: generate_canonical<double, 53> from two 32-bit
    draws, (x1 + x2 * 2^32) / 2^64."""
    raw = np.random.RandomState(seed).randint(0, 2 ** 32, size=2 * n, dtype=np.uint64)
    x1, x2 = raw[0::2].astype(np.float64), raw[1::2].astype(np.float64)
    u = (x1 + x2 * 4294967296.0) / 18446744073709551616.0
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return a + (b - a) * u


def test_reference_uniform_generator_is_mt19937(orc):
    x, y, z = O.gen_reference(O.GEN_UNIFORM, 50, 42, O.CT_F64)
    u = _libstdcxx_uniform(42, 150, 0.0, 100.0)
    assert np.array_equal(x, u[0::3]) and np.array_equal(y, u[1::3]) and np.array_equal(z, u[2::3])


def test_reference_fixed_point_conversion(orc):
    # float_to_int_rep(v, 0, 1/100): top-21 mantissa bits of min(1 + v/100, 2 - 2^-52)
    x, y, z = O.gen_reference(O.GEN_UNIFORM, 50, 42, O.CT_U64, bits=21)
    u = _libstdcxx_uniform(42, 150, 0.0, 100.0)[0::3]
    exp = []
    for v in u:
        nx = min(max(v * (1.0 / 100.0), 0.0), 2.0 - 2.0 ** -52)
        bb = min(1.0 + nx, 2.0 - 2.0 ** -52)
        bits = int(np.array([bb]).view(np.uint64)[0])
        exp.append((bits & ((1 << 52) - 1)) >> 31)
    assert [int(v) for v in x] == exp


# ---------------------------------------------------------------- kNN ----------
def numpy_knn(t, k, exclude_self, leaf_ids=None, queries=None):
    lx, ly, lz = t.leaf_coords()
    n = t.n
    ids = np.arange(n) if leaf_ids is None else np.asarray(leaf_ids)
    outi, outd = [], []
    rng = range(len(ids)) if queries is None else range(len(queries[0]))
    for qi in rng:
        if queries is None:
            s = ids[qi]
            p = (lx[s], ly[s], lz[s])
        else:
            s = -1
            p = (queries[0][qi], queries[1][qi], queries[2][qi])
        d = np.array([O.point_d2(t, np.array(p), np.array([lx[j], ly[j], lz[j]]))
                      for j in range(n)])
        idx = np.arange(n)
        keep = idx != s if exclude_self else np.ones(n, bool)
        d, idx = d[keep], idx[keep]
        o = np.lexsort((idx, d))[:k]
        outi.append(idx[o])
        outd.append(d[o])
    return outi, outd


def test_bruteforce_matches_numpy(orc):
    x, y, z = (RNG.integers(0, 64, 300, dtype=np.uint64) for _ in range(3))  # many ties
    t = O.build_bvh(O.CT_U64, x, y, z, sfc=O.SFC_PH, bits=6, periodic=1)
    r, d = O.knn_bruteforce(t, 12, True)
    ei, ed = numpy_knn(t, 12, True)
    for q in range(0, 300, 7):
        assert np.array_equal(r[q], ei[q].astype(np.uint64))
        assert np.array_equal(d[q], ed[q].astype(np.uint64))


def untied(d_ext, k):
    """Ranks < k whose distance equals no neighbouring rank's distance, from a
    brute-force row of k+1 entries (a tie with the first excluded candidate
    counts, as in test/test_self_knn.cpp:122-141)."""
    out = np.ones(d_ext.shape, bool)
    out[:, 1:] &= d_ext[:, 1:] != d_ext[:, :-1]
    out[:, :-1] &= d_ext[:, :-1] != d_ext[:, 1:]
    return out[:, :k]


@pytest.mark.parametrize("kind,seed", [(O.GEN_UNIFORM, 42), (O.GEN_CLUSTERED, 12345),
                                       (O.GEN_DEGENERATE, 999)])
@pytest.mark.parametrize("ct", [O.CT_U64, O.CT_F64])
def test_reference_traversal_vs_bruteforce(orc, kind, seed, ct):
    """The reference suite's acceptance rule (test/test_self_knn.cpp:153-266) on its
    own datasets: identical d2, identical indices at untied ranks."""
    x, y, z = O.gen_reference(kind, 1000, seed, ct, bits=21)
    t = O.build_bvh(ct, x, y, z, sfc=O.SFC_PH, bits=21)
    for k in (16, 32, 33, 64, 128, 256):
        for excl in (True, False):
            rb, db = O.knn_bruteforce(t, k + 1, excl)
            rr, dr = O.knn_reference(t, k, excl, max_k=256)
            assert np.array_equal(db[:, :k], dr)
            m = untied(db, k)
            assert np.array_equal(rb[:, :k][m], rr[m])


def test_reference_sorted_merge_vs_bruteforce(orc):
    x, y, z = O.gen_reference(O.GEN_UNIFORM, 1000, 42, O.CT_U64)
    t = O.build_bvh(O.CT_U64, x, y, z, sfc=O.SFC_PH, bits=21)
    q = (x[::31], y[::31], z[::31])
    for k in (33, 64, 128):
        rb, db = O.knn_bruteforce(t, k, queries=q)
        rr, dr = O.knn_reference(t, k, queries=q, kind=1, max_k=256)
        assert np.array_equal(db, dr)


def test_sentinels_when_k_exceeds_available(orc):
    x = np.array([1.0, 2.0]); y = np.zeros(2); z = np.zeros(2)
    t = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_MORTON, bits=10)
    r, d = O.knn_reference(t, 5, True)
    assert list(r[0]) == [1] + [2 ** 64 - 1] * 4
    assert d[0, 1] == -np.finfo(np.float64).max


# ---------------------------------------------------------------- range --------
def test_range_grid_kat(orc):
    """test/main.cpp:176-236: 10^3 grid, query (52,52,52), r in [0, 20] -> brute force."""
    x, y, z = O.gen_reference(O.GEN_GRID, 1000, 0, O.CT_F64)
    t = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_MORTON, bits=10)
    res, cnt = O.range_reference(t, [52.0], [52.0], [52.0], [0.0], [20.0], 1000)
    lx, ly, lz = t.leaf_coords()
    dist = np.sqrt((lx - 52) ** 2 + (ly - 52) ** 2 + (lz - 52) ** 2)
    assert cnt[0] == int(np.sum(dist <= 20.0))
    assert np.array_equal(res[0, :cnt[0]], np.nonzero(dist <= 20.0)[0])  # ascending rank


def test_range_collinear_kat(orc):
    """test/test_mpi.cpp:186-217: 1000 collinear points, query x=500.5, r=10 -> 20 hits."""
    x = np.arange(1000, dtype=np.float64)
    t = O.build_bvh(O.CT_F64, x, np.zeros(1000), np.zeros(1000), sfc=O.SFC_MORTON, bits=10)
    _, cnt = O.range_reference(t, [500.5], [0.0], [0.0], [0.0], [10.0], 1000)
    assert cnt[0] == 20


def test_range_periodic_integer(orc):
    """test/main.cpp:326-400: a query near the wrap finds images across it."""
    x, y, z = O.gen_reference(O.GEN_GRID, 1000, 0, O.CT_U64, bits=21)
    t = O.build_bvh(O.CT_U64, x, y, z, sfc=O.SFC_PH, bits=21, periodic=1)
    s = float(1 << 21) / 100.0
    q = [np.uint64(int(2.0 * s))], [np.uint64(int(50.0 * s))], [np.uint64(int(50.0 * s))]
    r = np.uint64(int(10.0 * s))
    res, cnt = O.range_reference(t, *q, [np.uint64(0)], [r], 1000)
    lx, ly, lz = t.leaf_coords()
    P = 1 << 21
    d2 = 0
    for a, b in ((lx, q[0][0]), (ly, q[1][0]), (lz, q[2][0])):
        dd = (a.astype(np.int64) - int(b)) % P
        dd = np.minimum(dd, P - dd)
        d2 = d2 + dd * dd
    assert cnt[0] == int(np.sum(d2 <= int(r) ** 2))
    # some hits are images across the wrap (x near the top of the box)
    hits = res[0, :cnt[0]]
    assert np.any(lx[hits] > P // 2)
