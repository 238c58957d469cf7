"""ft_self_knn_query_noalloc (the entry point that allocates nothing: its re-run
buffers live in a fixed device array) against ft_self_knn_query, through the C
ABI, on two calls in a row, the second with a smaller tree.  The seeds are too
small for a third of the queries, so the exact re-run list in the fixed array
is used; the outputs must equal the allocating entry point's bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from fasttree_amd import synth

pytestmark = pytest.mark.gpu


def _tree(ft, n, seed):
    x, y, z = synth.nfw_clustered(n, seed, box=100.0, n_halos=150)
    pts = tuple(synth.to_fixed(a, 100.0, 21, np.uint64) for a in (x, y, z))
    cfg = ft.Config(sfc=O.SFC_PH, bits_per_dim=21, periodic=1)
    tree = ft.TreeSoA(n, O.CT_U64, cfg)
    ft.build_bvh(ft.particles(*(ft.DeviceArray.from_host(a) for a in pts)), tree)
    return tree


def _noalloc(ft, tree, k, res, dst, seed=None, stats=None, max_k=32):
    o = ft._opts(max_k=max_k, seed_bound=seed, stats=stats)
    return ft.lib().ft_self_knn_query_noalloc(None, C.byref(tree.c), int(k), res.ptr, dst.ptr, 1,
                                              C.byref(o))


def test_noalloc_equals_self_knn_query(ft, orc):
    for n, k, seed in ((60000, 16, 71), (9000, 32, 72)):  # second call: smaller input
        tree = _tree(ft, n, seed)
        r_ref, d_ref = ft.self_knn_query(tree, k)
        # seeds: the exact k-th distance scaled, a third of them too small
        scale = np.where(np.arange(n) % 3 == 0, 0.5, 1.5)
        sb = ft.DeviceArray.from_host(np.floor(d_ref[:, -1].astype(np.float64) * scale)
                                      .astype(np.uint64))
        for seeded in (False, True):
            res = ft.DeviceArray(n * k, np.uint64)
            dst = ft.DeviceArray(n * k, np.uint64)
            stats = ft.DeviceArray(8, np.uint64).zero()
            rc = _noalloc(ft, tree, k, res, dst, seed=sb if seeded else None, stats=stats,
                          max_k=256)
            assert rc == ft.FT_OK, ft.lib().ft_last_error()
            assert np.array_equal(res.to_host().reshape(n, k), r_ref), (n, seeded)
            assert np.array_equal(dst.to_host().reshape(n, k), d_ref), (n, seeded)
            if seeded:  # exactly the too-small seeds were re-run from the fixed list
                too_small = int(np.count_nonzero(np.arange(n) % 3 == 0))
                assert stats.to_host()[3] >= too_small // 2


def test_noalloc_rejects_host_buffers_and_oversize(ft, orc):
    tree = _tree(ft, 2000, 73)
    res = np.zeros(2000 * 8, np.uint64)
    dst = ft.DeviceArray(2000 * 8, np.uint64)
    o = ft._opts(max_k=32)
    rc = ft.lib().ft_self_knn_query_noalloc(None, C.byref(tree.c), 8, res.ctypes.data, dst.ptr, 1,
                                            C.byref(o))
    assert rc == ft.FT_EINVAL
    big = ft._Tree()
    big.num_leaves = (1 << 24) + 1
    big.num_internal = 1 << 24
    big.coord_type = O.CT_U64
    big.config = ft.Config()
    rc = ft.lib().ft_self_knn_query_noalloc(None, C.byref(big), 8, dst.ptr, dst.ptr, 1, C.byref(o))
    assert rc == ft.FT_ERANGE
