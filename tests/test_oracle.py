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


# ------------------------------------------ reference tie picks, by hand -------
# Expected arrays below are derived by hand from the reference source (sift and
# heapsort rules src/maxheap.hpp:94-134, the bitonic network :193-215 and
# merge_batch :263-323, the DFS child order src/hlbvh.hpp:1618-1626), in the
# comments; the oracle restatement must reproduce them.
S = np.iinfo(np.uint64).max


def test_max_heap_tie_picks_by_hand(orc):
    """k = 4, pushes (5,10) (5,11) (3,12) (5,13) (2,14) (5,15) with the prune check.
    sift_up stops on <=, so the first four stay in arrival order: [5/10, 5/11,
    3/12, 5/13].  (2,14) replaces the root; sift_down takes the larger child,
    the left one on a tie: 2/14 swaps with 5/11 (slot 1), then with 5/13 (slot
    3): [5/11, 5/13, 3/12, 2/14].  (5,15) is pruned (5 >= root 5).  Heapsort:
    swap root and last, sift down over the rest, three times: [2/14, 3/12, 5/13,
    5/11].  The (d, index) order would keep 10 and 11 instead."""
    od, oi, cnt = O.heap_u64(0, 4, 4, [5, 5, 3, 5, 2, 5], [10, 11, 12, 13, 14, 15])
    assert cnt == 4
    assert list(od) == [2, 3, 5, 5] and list(oi) == [14, 12, 13, 11]


def test_sorted_merge_tie_order_by_hand(orc):
    """merge_batch with one candidate per batch, k = 2, cap = 64: batch 1 stages
    (7,2) (the network keeps a lone entry at slot 0 of every ascending block),
    count = min(32, 2) = 2.  Batch 2 appends (7,1) at slot 2; in the 64-wide
    sort, k = 2 / j = 1 is descending at slots 2-3, so (7,1) moves to slot 3,
    then k = 4 / j = 2 ascending moves it to slot 1 (a sentinel sat there); the
    tie (7,2)/(7,1) at slots 0-1 is only ever compared ascending, so it stays:
    [(7,2), (7,1)], arrival order, where (d, index) order is [1, 2].  One batch
    of lanes [(5,3), (5,1), sentinels], k = 2, cap = 32: slots 0 and 1 are
    compared ascending at every stage, so lane order survives: [3, 1]."""
    sent = [S] * 31
    d = [7] + sent + [7] + sent
    i = [2] + [-1] * 31 + [1] + [-1] * 31
    od, oi, cnt = O.heap_u64(1, 2, 64, d, i)
    assert cnt == 2 and list(od[:2]) == [7, 7] and list(oi[:2]) == [2, 1]
    od, oi, cnt = O.heap_u64(1, 2, 32, [5, 5] + [S] * 30, [3, 1] + [-1] * 30)
    assert cnt == 2 and list(oi[:2]) == [3, 1]


def _line_tree(xs):
    x = np.array(xs, np.uint64)
    zero = np.zeros(len(xs), np.uint64)
    return x, zero, zero.copy()


def test_reference_dfs_tie_picks_by_hand(orc):
    """Eight points at x = 0, 2, ..., 14 (y = z = 0, Morton keys: a perfect
    tree; leaf rank r at x = 2r), query x = 7, k = 4, knn_query's per-query walk.
    DFS (farther child pushed first, left on a tie): root -> [0,6] -> [4,6] ->
    leaf r3 (d2 1), r2 (9) -> [0,2] -> r1 (25), r0 (49) -> [8,14] -> [8,10] ->
    r4 (1, replaces the root 49), r5 (9, replaces the root 25) -> [12,14]
    pruned.  Heap [9/5, 1/4, 9/2, 1/3]; heapsort gives [1/4, 1/3, 9/2, 9/5]:
    indices [4, 3, 2, 5], where the (d2, index) order is [3, 4, 2, 5]."""
    x, y, z = _line_tree(range(0, 16, 2))
    t = O.build_bvh(O.CT_U64, x, y, z, sfc=O.SFC_MORTON, bits=21, periodic=0)
    q = (np.array([7], np.uint64), np.zeros(1, np.uint64), np.zeros(1, np.uint64))
    r, d = O.knn_reference(t, 4, queries=q, max_k=32)
    assert list(r[0]) == [4, 3, 2, 5] and list(d[0]) == [1, 1, 9, 9]
    rb, db = O.knn_bruteforce(t, 4, queries=q)
    assert list(rb[0]) == [3, 4, 2, 5] and list(db[0]) == [1, 1, 9, 9]


MERGE_TIER_IDX = [17, 18, 16, 19, 20, 15, 21, 14, 22, 13, 23, 12, 24, 11, 10, 25, 26, 9, 27,
                  8, 28, 7, 6, 29, 30, 5, 4, 31, 32, 3, 33, 2, 1]


def merge_tier_case():
    """34 points on a line: ranks 0-16 at x = 2r, ranks 17-33 at x = 64 + 2j
    (the root splits at x bit 6); query x = 50, k = 33 (the SortedMergeHeap
    tier of knn_query_large_k, cap 256).  The right subtree is nearer (gap 14
    against 18), and inside each group the side facing the query comes first,
    so leaves arrive as ranks 17, 18, ..., 33 then 16, 15, ..., 1; rank 0
    (d2 2500) meets a full list whose worst is 2304 and is pruned."""
    x, y, z = _line_tree([2 * r for r in range(17)] + [64 + 2 * j for j in range(17)])
    q = (np.array([50], np.uint64), np.zeros(1, np.uint64), np.zeros(1, np.uint64))
    arrivals = [((14 + 2 * j) ** 2, 17 + j) for j in range(17)] + \
               [((18 + 2 * i) ** 2, 16 - i) for i in range(16)]
    return (x, y, z), q, arrivals


def _bitonic_textbook(d, ix, lo, n, asc):
    """Batcher's recursive bitonic sort (the textbook form of the iterative
    network of src/maxheap.hpp:193-215: the same comparators in the same
    directions, swap when (d[i] > d[i + h]) == ascending)."""
    if n <= 1:
        return
    h = n // 2
    _bitonic_textbook(d, ix, lo, h, True)
    _bitonic_textbook(d, ix, lo + h, h, False)
    _bitonic_merge(d, ix, lo, n, asc)


def _bitonic_merge(d, ix, lo, n, asc):
    if n <= 1:
        return
    h = n // 2
    for i in range(lo, lo + h):
        if (d[i] > d[i + h]) == asc:
            d[i], d[i + h] = d[i + h], d[i]
            ix[i], ix[i + h] = ix[i + h], ix[i]
    _bitonic_merge(d, ix, lo, h, asc)
    _bitonic_merge(d, ix, lo + h, h, asc)


def test_reference_merge_tier_tie_order(orc):
    """The merge tier's final order for the hand-derived arrival sequence,
    computed with the textbook network (one candidate + 31 sentinels staged and
    sorted, appended at `count`, padded to a power of two, sorted, count
    clamped to k): ties come out in a network-defined order (e.g. d2 324:
    ranks [16, 19], though 19 arrived first; d2 900: [10, 25]).  The oracle's
    restatement of knn_query_large_k gives the same literal array."""
    pts, q, arrivals = merge_tier_case()
    cap, k = 256, 33
    d, ix, count = [int(S)] * cap, [-1] * cap, 0
    for dd, ii in arrivals:
        sd, si = [dd] + [int(S)] * 31, [ii] + [-1] * 31
        _bitonic_textbook(sd, si, 0, 32, True)
        for lane in range(32):
            if count + lane < cap:
                d[count + lane], ix[count + lane] = sd[lane], si[lane]
        nc = min(count + 32, cap)
        p2 = 1
        while p2 < nc:
            p2 <<= 1
        for lane in range(nc, min(p2, cap)):
            d[lane], ix[lane] = int(S), -1
        _bitonic_textbook(d, ix, 0, p2, True)
        count = min(nc, k)
    assert ix[:k] == MERGE_TIER_IDX
    t = O.build_bvh(O.CT_U64, *pts, sfc=O.SFC_MORTON, bits=21, periodic=0)
    r, dr = O.knn_reference(t, k, queries=q, kind=1, max_k=256)
    assert list(r[0]) == MERGE_TIER_IDX and list(dr[0]) == d[:k]


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


# ---------------------------------------- snapshot positions -> coordinates ----
def test_fill_particle_coords_kat(orc):
    """fill_particle_coords (test/benchmark/BenchmarkUtils.hpp:164-205) on a hand
    case: one cube over all axes, min 0, max 100, size 100 * 1.0001, so a value v
    maps to floor(v / 100.01 * 2^21), computed here in exact rationals: 0, 50,
    100 -> 0, 1048471, 2096942; y = 10 -> 209694, z = 20 -> 419388."""
    px = np.array([0.0, 50.0, 100.0], np.float32)
    py = np.full(3, 10.0, np.float32)
    pz = np.full(3, 20.0, np.float32)
    for ct in (O.CT_U64, O.CT_U32):
        x, y, z, bmin, bsize = O.fill_particle_coords(px, py, pz, ct, 21)
        assert list(x) == [0, 1048471, 2096942]
        assert list(y) == [209694] * 3 and list(z) == [419388] * 3
        assert bmin == 0.0 and bsize == 100.0 * 1.0001


def test_fill_particle_coords_invariants(orc):
    """The largest coordinate always maps to floor(2^B / 1.0001) (2096942 at
    B = 21, 1023 at B = 10) and the smallest to 0, whichever axis holds them; the
    box is shared by the axes (moving y's minimum shifts every x); equal
    positions give box (v, 1.0) and zeros; the map is monotone; FP coordinate
    types are a plain cast with box (0, 1); empty input gives box (0, 1)."""
    rng = np.random.default_rng(7)
    for bits, top in ((21, 2096942), (10, 1023)):
        px = rng.uniform(-3.0, 40.0, 5000).astype(np.float32)
        py = rng.uniform(-7.5, 12.0, 5000).astype(np.float32)
        pz = rng.uniform(0.0, 55.25, 5000).astype(np.float32)
        x, y, z, bmin, bsize = O.fill_particle_coords(px, py, pz, O.CT_U64, bits)
        allv = np.concatenate([px, py, pz]).astype(np.float64)
        allc = np.concatenate([x, y, z])
        assert bmin == allv.min() and allc[allv.argmax()] == top and allc[allv.argmin()] == 0
        o = np.argsort(allv, kind="stable")
        assert np.all(np.diff(allc[o].astype(np.int64)) >= 0)
        # floor(nx * 2^B) is exact for nx away from the cell edges (nx is
        # within 2^-52 of an edge only for a handful of values)
        nx = (allv - bmin) / bsize
        lo = np.floor(nx * 2.0 ** bits).astype(np.int64)
        d = allc.astype(np.int64) - lo
        assert np.all((d == 0) | (d == 1)) and np.count_nonzero(d) <= 3
        # a lower y minimum moves the shared box, so every x changes
        py2 = py.copy()
        py2[17] = -20.0
        x2, _, _, bmin2, _ = O.fill_particle_coords(px, py2, pz, O.CT_U64, bits)
        assert bmin2 == -20.0 and np.all(x2 >= x) and np.count_nonzero(x2 != x) > 4900
    same = np.full(4, 3.5, np.float32)
    x, y, z, bmin, bsize = O.fill_particle_coords(same, same, same, O.CT_U64, 21)
    assert bmin == 3.5 and bsize == 1.0 and not x.any() and not y.any() and not z.any()
    f = rng.uniform(0, 1, 10).astype(np.float32)
    x, y, z, bmin, bsize = O.fill_particle_coords(f, f, f, O.CT_F64, 21)
    assert np.array_equal(x, f.astype(np.float64)) and (bmin, bsize) == (0.0, 1.0)
    e = np.zeros(0, np.float32)
    assert O.fill_particle_coords(e, e, e, O.CT_U64, 21)[3:] == (0.0, 1.0)


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


@pytest.mark.parametrize("kind,seed", [(O.GEN_UNIFORM, 42), (O.GEN_CLUSTERED, 12345),
                                       (O.GEN_DEGENERATE, 999)])
def test_reference_grouped_traversal(orc, kind, seed):
    """The grouped lockstep walk (8 and 16 lanes; the CPU baseline's form of the
    reference kernel) returns the per-query walk's arrays exactly -- a lane pushes
    a leaf only when its own test passes -- on trees too shallow for the stack
    limit, and makes at least as many visits (a lane expands nodes it pruned
    while another lane needs its step); with a subset whose size is not a
    multiple of the width, the shadow lanes change nothing."""
    x, y, z = O.gen_reference(kind, 1000, seed, O.CT_U64, bits=21)
    t = O.build_bvh(O.CT_U64, x, y, z, sfc=O.SFC_PH, bits=21)
    ids = np.arange(3, 1000, 7, dtype=np.int32)
    for k in (1, 16, 32, 64):
        rr, dr, cnt = O.knn_reference(t, k, True, leaf_ids=ids, max_k=256, counters=True)
        rb, db = O.knn_bruteforce(t, k + 1, True, leaf_ids=ids)
        assert np.array_equal(db[:, :k], dr)
        for w in (8, 16):
            rg, dg, vis = O.knn_reference_grouped(t, k, True, leaf_ids=ids, width=w, visits=True)
            assert np.array_equal(rg, rr) and np.array_equal(dg, dr), (k, w)
            assert np.all(vis >= cnt[:, 0]) and vis.sum() > cnt[:, 0].sum()


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
