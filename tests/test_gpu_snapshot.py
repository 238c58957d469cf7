"""GPU parity of fill_particle_coords (test/benchmark/BenchmarkUtils.hpp:164-205):
snapshot positions (float32 as the HDF5 read gives them, or float64) to tree
coordinates, bit for bit against the oracle, and the converted snapshot fed to
build_bvh + self_knn_query as the reference's benchmarks do
(test/benchmark/rebuild_scaling.cpp:14-57)."""
import numpy as np
import pytest

import oracle as O
from fasttree_amd import synth

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(5)
    x, y, z = synth.nfw_clustered(400_000, 3, box=100.0, n_halos=300)
    yield "nfw-f32", tuple(a.astype(np.float32) for a in (x, y, z))
    yield "nfw-f64", (x, y, z)
    yield "negative-f32", tuple(rng.uniform(-1e4, -2.5, 100_000).astype(np.float32) for _ in range(3))
    yield "mixed-axes-f64", (rng.normal(0, 1, 70_000), rng.uniform(-3, 9, 70_000),
                             rng.uniform(5, 6, 70_000))
    yield "constant", tuple(np.full(1000, 7.25, np.float32) for _ in range(3))
    yield "one", tuple(np.array([1.5], np.float32) for _ in range(3))
    yield "two", (np.array([0.0, 1.0], np.float32), np.array([2.0, 2.0], np.float32),
                  np.array([-1.0, 3.0], np.float32))


@pytest.mark.parametrize("ct,bits", [(O.CT_U64, 21), (O.CT_U32, 21), (O.CT_U64, 42),
                                     (O.CT_U32, 32), (O.CT_U64, 10), (O.CT_F32, 21),
                                     (O.CT_F64, 21)])
def test_fill_particle_coords_matches_oracle(ft, orc, ct, bits):
    for name, (px, py, pz) in _cases():
        want = O.fill_particle_coords(px, py, pz, ct, bits)
        got = ft.fill_particle_coords(px, py, pz, ct, bits)  # host arrays: staged
        for a, b in zip(got[:3], want[:3]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
        assert got[3:] == want[3:], name
        # device-resident inputs give device outputs with the same bits
        dev = [ft.DeviceArray.from_host(a) for a in (px, py, pz)]
        gd = ft.fill_particle_coords(*dev, ct, bits)
        for a, b in zip(gd[:3], want[:3]):
            assert np.array_equal(a.to_host().view(np.uint8), b.view(np.uint8)), name
        assert gd[3:] == want[3:], name


def test_fill_particle_coords_full_size_then_build_and_query(ft, orc):
    """A 10M float32 snapshot converted on the GPU equals the oracle's
    conversion; 200k of it built and queried equal the oracle's build and
    brute force (int64 coordinates, 21-bit Peano-Hilbert keys)."""
    x, y, z = synth.nfw_clustered(10_000_000, 9, box=100.0, n_halos=2000)
    px, py, pz = (a.astype(np.float32) for a in (x, y, z))
    want = O.fill_particle_coords(px, py, pz, O.CT_U64, 21)
    got = ft.fill_particle_coords(*(ft.DeviceArray.from_host(a) for a in (px, py, pz)),
                                  O.CT_U64, 21)
    for a, b in zip(got[:3], want[:3]):
        assert np.array_equal(a.to_host(), b)
    assert got[3:] == want[3:]
    n, k = 200_000, 16
    sub = [a[:n] for a in (px, py, pz)]
    cx, cy, cz, _, _ = ft.fill_particle_coords(*sub, O.CT_U64, 21)
    cfg = ft.Config(sfc=O.SFC_PH, bits_per_dim=21, periodic=0)
    tree = ft.TreeSoA(n, O.CT_U64, cfg)
    ft.build_bvh(ft.particles(*(ft.DeviceArray.from_host(a) for a in (cx, cy, cz))), tree)
    t = O.build_bvh(O.CT_U64, cx, cy, cz, sfc=O.SFC_PH, bits=21, periodic=0)
    assert np.array_equal(tree.to_host()["sorted_keys"], t.sorted_keys)
    r, d = ft.self_knn_query(tree, k)
    ids = np.arange(0, n, 101, dtype=np.int32)
    rb, db = O.knn_bruteforce(t, k, True, leaf_ids=ids)
    assert np.array_equal(r[ids], rb) and np.array_equal(d[ids], db)


def test_fill_particle_coords_errors(ft, orc):
    a = np.zeros(4, np.float32)
    with pytest.raises(Exception):
        ft.fill_particle_coords(a, a, a, O.CT_U32, 33)
    with pytest.raises(Exception):
        ft.fill_particle_coords(a, a, a, O.CT_U64, 53)
    e = np.zeros(0, np.float32)
    assert ft.fill_particle_coords(e, e, e, O.CT_U64, 21)[3:] == (0.0, 1.0)
