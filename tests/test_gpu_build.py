"""GPU parity of the build: every TreeSoA array, the sorted keys and the node
ranges must equal the oracle's, bit for bit."""
import numpy as np
import pytest

import oracle as O
from fasttree_amd import synth

pytestmark = pytest.mark.gpu

ARRAYS = ["min_x", "max_x", "min_y", "max_y", "min_z", "max_z"]


def gpu_build(ft, ct, x, y, z, sfc, bits, periodic=0, box=0.0, ids=None, ghost=None,
              bbox=None, keys=None, scratch=None):
    cfg = ft.Config(sfc=sfc, bits_per_dim=bits, periodic=periodic, box_length=box)
    tree = ft.TreeSoA(len(x), ct, cfg)
    dt = ft.COORD_DTYPES[ct]
    p = ft.particles(*(ft.DeviceArray.from_host(a, dt) for a in (x, y, z)),
                     id=None if ids is None else ft.DeviceArray.from_host(ids, np.uint32),
                     is_ghost=None if ghost is None else ft.DeviceArray.from_host(ghost, np.int8))
    ft.build_bvh(p, tree, scratch=scratch, bbox=bbox, key=keys)
    return tree


def compare_trees(g, t, n):
    m = 2 * n - 1
    o = n - 1
    assert np.array_equal(g["sorted_keys"], t.sorted_keys[:n]), "sorted keys"
    assert np.array_equal(g["orig_idx"][o:m], t.orig[o:m]), "permutation"
    assert np.array_equal(g["id"][o:m], t.id[o:m]), "ids"
    assert np.array_equal(g["is_ghost"][o:m], t.ghost[o:m]), "ghost flags"
    for a in ARRAYS:
        assert np.array_equal(g[a][:m].view(np.uint8), t.arrays[a][:m].view(np.uint8)), a
    if n > 1:
        assert np.array_equal(g["left_child"], t.left), "left"
        assert np.array_equal(g["right_child"], t.right), "right"
        assert np.array_equal(g["range_first"], t.range_first)
        assert np.array_equal(g["range_last"], t.range_last)
    assert np.array_equal(g["parent"][:m], t.parent[:m]), "parent"


CASES = [
    # (name, coord type, sfc, bits, periodic, generator)
    ("ref-uniform-u64-ph", O.CT_U64, O.SFC_PH, 21, 1, ("ref", O.GEN_UNIFORM, 1000, 42)),
    ("ref-clustered-f64-ph", O.CT_F64, O.SFC_PH, 21, 0, ("ref", O.GEN_CLUSTERED, 10000, 54321)),
    ("ref-degenerate-u32-morton", O.CT_U32, O.SFC_MORTON, 21, 0, ("ref", O.GEN_DEGENERATE, 1000, 999)),
    ("ref-grid-f32-morton", O.CT_F32, O.SFC_MORTON, 10, 0, ("ref", O.GEN_GRID, 1000, 0)),
    ("uniform-f64-morton10", O.CT_F64, O.SFC_MORTON, 10, 0, ("uni", 100_000, 1)),
    ("nfw-u64-ph21", O.CT_U64, O.SFC_PH, 21, 1, ("nfw", 300_000, 3)),
    ("nfw-u32-morton21", O.CT_U32, O.SFC_MORTON, 21, 0, ("nfw", 150_000, 5)),
    ("nfw-f64-ph21", O.CT_F64, O.SFC_PH, 21, 0, ("nfw", 120_000, 9)),
    ("tiny-3", O.CT_F64, O.SFC_MORTON, 10, 0, ("uni", 3, 2)),
    ("tiny-2", O.CT_U64, O.SFC_PH, 21, 0, ("uni", 2, 4)),
]


def make_points(gen, ct):
    dt = O.COORD_DTYPES[ct]
    if gen[0] == "ref":
        return O.gen_reference(gen[1], gen[2], gen[3], ct, bits=21)
    if gen[0] == "uni":
        x, y, z = synth.uniform(gen[1], gen[2], 1.0)
    else:
        x, y, z = synth.nfw_clustered(gen[1], gen[2], box=100.0, n_halos=200)
    if np.issubdtype(dt, np.integer):
        return tuple(synth.to_fixed(a, 100.0 if gen[0] == "nfw" else 1.0, 21, dt)
                     for a in (x, y, z))
    return tuple(a.astype(dt) for a in (x, y, z))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_build_matches_oracle(ft, orc, case):
    name, ct, sfc, bits, periodic, gen = case
    print(f"[build] {name}", flush=True)
    x, y, z = make_points(gen, ct)
    n = len(x)
    ghost = (np.arange(n) % 7 == 0).astype(np.int8)
    ids = (np.arange(n, dtype=np.uint32) * 3 + 11)
    tree = gpu_build(ft, ct, x, y, z, sfc, bits, periodic, ids=ids, ghost=ghost)
    g = tree.to_host()
    t = O.build_bvh(ct, x, y, z, sfc=sfc, bits=bits, periodic=periodic, ids=ids, ghost=ghost)
    compare_trees(g, t, n)


RADIX_CASES = [c for c in CASES if c[0] in ("ref-clustered-f64-ph", "ref-degenerate-u32-morton",
                                             "uniform-f64-morton10", "nfw-u64-ph21", "tiny-2")]


@pytest.mark.parametrize("case", RADIX_CASES, ids=[c[0] for c in RADIX_CASES])
def test_build_with_9_bit_digits_matches_oracle(ft, orc, case):
    """The radix sort forced to 9-bit digits (7 passes for 63-bit keys, 4 for
    30-bit keys, where the default is 8-bit) gives the same stable order: every array equals the oracle's,
    including the degenerate set's runs of equal keys; so do keys given by the
    caller with many duplicates."""
    name, ct, sfc, bits, periodic, gen = case
    print(f"[build 9-bit] {name}", flush=True)
    x, y, z = make_points(gen, ct)
    sc = ft.BvhScratch()
    sc.set_radix_bits(9)
    tree = gpu_build(ft, ct, x, y, z, sfc, bits, periodic, scratch=sc)
    t = O.build_bvh(ct, x, y, z, sfc=sfc, bits=bits, periodic=periodic)
    compare_trees(tree.to_host(), t, len(x))
    if len(x) > 1000:
        keys = (np.arange(len(x), dtype=np.uint64) * 2654435761) % 997
        tree2 = gpu_build(ft, ct, x, y, z, sfc, bits, periodic, scratch=sc,
                          keys=ft.DeviceArray.from_host(keys, np.uint64))
        t2 = O.build_bvh(ct, x, y, z, sfc=sfc, bits=bits, periodic=periodic, keys=keys)
        compare_trees(tree2.to_host(), t2, len(x))


def test_radix_bits_argument(ft):
    sc = ft.BvhScratch()
    for bad in (-1, 7, 10, 16):
        with pytest.raises(Exception):
            sc.set_radix_bits(bad)
    sc.set_radix_bits(9)
    sc.set_radix_bits(8)
    sc.set_radix_bits(0)


def test_single_particle(ft, orc):
    x = np.array([5.0]); y = np.array([6.0]); z = np.array([7.0])
    tree = gpu_build(ft, O.CT_F64, x, y, z, O.SFC_MORTON, 10)
    g = tree.to_host()
    assert g["min_x"][0] == 5.0 and g["max_z"][0] == 7.0 and g["parent"][0] == -1
    assert g["orig_idx"][0] == 0 and g["id"][0] == 0


def test_default_ids_and_ghosts(ft, orc):
    x, y, z = make_points(("uni", 5000, 8), O.CT_F64)
    tree = gpu_build(ft, O.CT_F64, x, y, z, O.SFC_MORTON, 10)
    g = tree.to_host()
    o = 4999
    assert np.array_equal(g["id"][o:], g["orig_idx"][o:].astype(np.uint32))
    assert not g["is_ghost"][o:].any()


def test_precomputed_bbox_and_keys(ft, orc):
    x, y, z = make_points(("uni", 20000, 12), O.CT_F64)
    bb = np.array([-1.0, 2.0, -1.0, 2.0, -1.0, 2.0])
    tree = gpu_build(ft, O.CT_F64, x, y, z, O.SFC_PH, 21, bbox=bb)
    t = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_PH, bits=21, bbox6=bb)
    compare_trees(tree.to_host(), t, len(x))
    keys = (np.arange(len(x), dtype=np.uint64) * 2654435761) % 1000
    tree2 = gpu_build(ft, O.CT_F64, x, y, z, O.SFC_PH, 21,
                      keys=ft.DeviceArray.from_host(keys, np.uint64))
    t2 = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_PH, bits=21, keys=keys)
    compare_trees(tree2.to_host(), t2, len(x))


def test_scratch_reuse_and_tree_pool(ft, orc):
    """BvhScratch / TreePool semantics: one scratch and one pooled tree across
    builds of different sizes (src/hlbvh.hpp:682-745)."""
    scratch = ft.BvhScratch()
    cfg = ft.Config(sfc=O.SFC_PH, bits_per_dim=21, periodic=1)
    tree = ft.TreeSoA(10, O.CT_U64, cfg)
    for n, seed in ((50_000, 21), (7_000, 22), (80_000, 23)):
        x, y, z = make_points(("nfw", n, seed), O.CT_U64)
        tree.ensure(n)
        p = ft.particles(*(ft.DeviceArray.from_host(a) for a in (x, y, z)))
        ft.build_bvh(p, tree, scratch=scratch)
        t = O.build_bvh(O.CT_U64, x, y, z, sfc=O.SFC_PH, bits=21, periodic=1)
        compare_trees(tree.to_host(), t, n)


def test_host_pointer_staging(ft, orc):
    x, y, z = make_points(("uni", 4096, 31), O.CT_F64)
    cfg = ft.Config(sfc=O.SFC_MORTON, bits_per_dim=10)
    tree = ft.TreeSoA(len(x), O.CT_F64, cfg)
    ft.build_bvh(ft.particles(x, y, z), tree)  # host numpy arrays: staged
    t = O.build_bvh(O.CT_F64, x, y, z, sfc=O.SFC_MORTON, bits=10)
    compare_trees(tree.to_host(), t, len(x))


@pytest.mark.parametrize("ct", [O.CT_F64, O.CT_F32, O.CT_U32, O.CT_U64])
@pytest.mark.parametrize("sfc", [O.SFC_MORTON, O.SFC_PH])
def test_encode_and_bbox_match_oracle(ft, orc, ct, sfc):
    x, y, z = make_points(("nfw", 30000, 40 + ct), ct)
    if ct in (O.CT_F64, O.CT_F32):
        # include exact extremes so the Morton maximum quirk is exercised
        x = x.copy(); x[5] = x.max(); x[6] = x.min()
    bb = ft.compute_bbox(x, y, z)
    assert np.array_equal(bb, O.bbox(ct, x, y, z))
    bits = 10 if (ct == O.CT_F32 and sfc == O.SFC_MORTON) else 21
    cfg = ft.Config(sfc=sfc, bits_per_dim=bits)
    keys = ft.sfc_encode(x, y, z, cfg, bbox=bb)
    assert np.array_equal(keys, O.encode(ct, sfc, bits, x, y, z, bb))


def test_build_tree_collinear_kat(ft, orc):
    """test/test_query.cpp: build_tree over 10 collinear points with keys = i, then
    the range search [5.5 +- 2] finds leaves 4..7."""
    n = 10
    cfg = ft.Config(sfc=O.SFC_MORTON, bits_per_dim=10)
    tree = ft.TreeSoA(n, O.CT_F64, cfg)
    keys = np.arange(n, dtype=np.uint64)
    sx = np.arange(n, dtype=np.float64)
    ft.build_tree(tree, keys, sx, np.zeros(n), np.zeros(n))
    g = tree.to_host()
    left, right, parent, rf, rl = O.karras(keys)
    assert np.array_equal(g["left_child"], left) and np.array_equal(g["right_child"], right)
    res, cnt = ft.range_query(tree, np.array([5.5]), np.zeros(1), np.zeros(1), np.zeros(1),
                              np.array([2.0]), 10)
    assert cnt[0] == 4 and list(res[0, :4]) == [4, 5, 6, 7]


@pytest.mark.parametrize("n", [10_000_000])
def test_full_size_build_properties(ft, orc, n):
    """The 10M clustered int64 snapshot: size-independent properties on the GPU
    tree, then a full comparison with the oracle build."""
    print(f"[build] full size n={n}", flush=True)
    x, y, z = synth.clustered_fixed(n, 0)
    tree = gpu_build(ft, O.CT_U64, x, y, z, O.SFC_PH, 21, periodic=1)
    g = tree.to_host()
    k = g["sorted_keys"]
    assert np.all(k[1:] >= k[:-1])
    o = n - 1
    perm = g["orig_idx"][o:]
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.int32))
    assert np.array_equal(g["min_x"][o:], x[perm])
    t = O.build_bvh(O.CT_U64, x, y, z, sfc=O.SFC_PH, bits=21, periodic=1)
    compare_trees(g, t, n)
    # the same snapshot sorted in 8 passes of 8-bit digits (the default for
    # 63-bit keys is 7 passes of 9 bits): the same tree
    sc = ft.BvhScratch()
    sc.set_radix_bits(8)
    g9 = gpu_build(ft, O.CT_U64, x, y, z, O.SFC_PH, 21, periodic=1, scratch=sc).to_host()
    for a in ("sorted_keys", "orig_idx", "left_child", "right_child", "parent", "range_first",
              "range_last", *ARRAYS):
        assert np.array_equal(g9[a], g[a]), a


@pytest.mark.parametrize("ct", [O.CT_U64, O.CT_F64, O.CT_U32])
def test_serialized_tree_round_trip(ft, orc, ct):
    """A serialised tree carries the reference's MPI transfer sequence after its
    64-byte header (test/test_mpi.cpp:76-100: the u64 dims [num_leaves,
    num_internal], then min_x, max_x, min_y, max_y, min_z, max_z, left_child,
    right_child, parent), then id, is_ghost, orig_idx, the sorted keys and the
    leaf ranges.  Every byte range is checked against the arrays of the tree the
    oracle builds from the same points, so a field-order or size mistake shared
    by serialise and deserialise fails.  The buffer restores to an identical tree
    that answers queries identically; damaged, truncated and topologically
    corrupt buffers are errors, not trees."""
    import ctypes as C
    periodic = 1 if ct == O.CT_U64 else 0
    x, y, z = make_points(("nfw", 30000, 41), ct)
    tree = gpu_build(ft, ct, x, y, z, O.SFC_PH, 21, periodic)
    t = O.build_bvh(ct, x, y, z, sfc=O.SFC_PH, bits=21, periodic=periodic)
    g = tree.to_host()
    data = tree.serialize()
    n = len(x)
    m = 2 * n - 1
    assert bytes(data[:8]) == b"FTTREE01"
    dt = np.dtype(O.COORD_DTYPES[ct])
    # the header's coordinate type and configuration (ft_config after 16 bytes)
    assert np.frombuffer(data, np.int32, 1, 8)[0] == ct
    off = 64
    assert np.array_equal(np.frombuffer(data, np.uint64, 2, off), [n, n - 1])
    off += 16
    seq = [(a, dt, m, t.arrays[a]) for a in ARRAYS]
    seq += [("left_child", np.int32, n - 1, t.left), ("right_child", np.int32, n - 1, t.right),
            ("parent", np.int32, m, t.parent)]
    seq += [("id", np.uint32, m, None), ("is_ghost", np.int8, m, None),
            ("orig_idx", np.int32, m, None), ("sorted_keys", np.uint64, n, t.sorted_keys),
            ("range_first", np.int32, n - 1, t.range_first),
            ("range_last", np.int32, n - 1, t.range_last)]
    oracle_leaf = {"id": t.id, "is_ghost": t.ghost, "orig_idx": t.orig}
    for name, adt, cnt, ref in seq:
        part = np.frombuffer(data, adt, cnt, off)
        if ref is not None:
            assert np.array_equal(part.view(np.uint8), np.ascontiguousarray(ref[:cnt], adt).view(np.uint8)), name
        else:  # only leaves carry id, is_ghost and orig_idx
            assert np.array_equal(part[n - 1:m], oracle_leaf[name][n - 1:m].astype(adt)), name
        off += cnt * np.dtype(adt).itemsize
    assert off == len(data)
    t2 = ft.TreeSoA.deserialize(data)
    assert t2.num_leaves == n and t2.coord_type == ct
    g2 = t2.to_host()
    for name in g:
        assert np.array_equal(g2[name], g[name]), name
    r1, d1 = ft.self_knn_query(tree, 8)
    r2, d2 = ft.self_knn_query(t2, 8)
    assert np.array_equal(r1, r2) and np.array_equal(d1, d2)
    # a damaged header and a truncated buffer are errors
    bad = bytearray(data)
    bad[0:8] = b"NOTATREE"
    t3 = ft.TreeSoA(1, ct)

    def load(b):
        buf = np.frombuffer(bytes(b), np.uint8)
        return ft.lib().ft_tree_deserialize(None, C.byref(t3.c), buf.ctypes.data, buf.nbytes)

    assert load(bad) == ft.FT_EINVAL
    assert load(data[:len(data) // 2]) == ft.FT_ERANGE
    # corrupt topology: a child out of range, a child listed twice (a cycle a
    # walk would never leave), a wrong parent, a leaf range past the leaves
    o_left = 80 + 6 * m * dt.itemsize
    o_right = o_left + 4 * (n - 1)
    o_parent = o_right + 4 * (n - 1)
    o_rf = o_parent + 4 * m + 4 * m + m + 4 * m + 8 * n
    o_rl = o_rf + 4 * (n - 1)
    for at, val in ((o_left + 4 * 5, 2 * n - 1), (o_left + 4 * 7, -3), (o_right + 4 * 9, 0),
                    (o_right + 4 * 11, int(t.left[11])), (o_parent + 4 * int(t.left[3]), 4),
                    (o_rf + 4 * 2, -1), (o_rl + 4 * 6, n), (o_parent, 0)):
        b = bytearray(data)
        b[at:at + 4] = np.int32(val).tobytes()
        assert load(b) == ft.FT_EINVAL, (at, val)
    assert load(data) == ft.FT_OK


def test_graph_captured_rebuild_and_query(ft, orc):
    """A build_bvh + self_knn_query step captured in a hipGraph and replayed on
    kicked particles (new positions written into the same device arrays,
    test/benchmark/rebuild_scaling.cpp's setting): every replay equals an eager
    build and query of the same positions, with unrelated eager work and
    synchronisations between the replays."""
    n, k = 100_000, 16
    x, y, z = make_points(("nfw", n, 51), O.CT_U64)
    cfg = ft.Config(sfc=O.SFC_PH, bits_per_dim=21, periodic=1)
    s = ft.Stream()
    dx, dy, dz = (ft.DeviceArray.from_host(a) for a in (x, y, z))
    p = ft.particles(dx, dy, dz)
    tree, sc = ft.TreeSoA(n, O.CT_U64, cfg), ft.BvhScratch()
    res, dst = ft.DeviceArray(n * k, np.uint64), ft.DeviceArray(n * k, np.uint64)

    def step():
        ft.build_bvh(p, tree, scratch=sc, stream=s.handle)
        ft.self_knn_query(tree, k, results=res, dists=dst, stream=s.handle)

    step()  # eager: sizes every buffer
    s.synchronize()
    with ft.Graph(s) as g:
        step()
    ref_tree, ref_sc = ft.TreeSoA(n, O.CT_U64, cfg), ft.BvhScratch()
    rng = np.random.default_rng(3)
    mask = np.uint64((1 << 21) - 1)
    for rep in range(3):
        kick = [rng.integers(-40, 41, n).astype(np.int64) for _ in range(3)]
        new = [((a.astype(np.int64) + d) & int(mask)).astype(np.uint64) for a, d in zip((x, y, z), kick)]
        for dev, a in zip((dx, dy, dz), new):
            dev.copy_from(a)
        g.launch()
        s.synchronize()
        # eager reference on separate buffers, and unrelated work in between
        ft.build_bvh(ft.particles(*(ft.DeviceArray.from_host(a) for a in new)), ref_tree,
                     scratch=ref_sc)
        r_ref, d_ref = ft.self_knn_query(ref_tree, k)
        ft.self_knn_query(ref_tree, 4)
        assert np.array_equal(res.to_host().reshape(n, k), r_ref), rep
        assert np.array_equal(dst.to_host().reshape(n, k), d_ref), rep
        assert np.array_equal(tree.to_host()["parent"], ref_tree.to_host()["parent"]), rep
