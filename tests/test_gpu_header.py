"""The C++ drop-in header on the GPU: the reference-shaped caller
tests/host/selfknn_runner.cpp (built from source with hipcc by
tests/host/Makefile, in two of the reference's compile-time configurations)
runs build_bvh, self_knn_query, self_knn_query_subset and knn_query on the
reference suite's own datasets (test/test_self_knn.cpp: uniform seed 42,
clustered seed 12345, degenerate seed 999, N = 1000), k = 16 and 32,
exclude_self on and off.  Every output array equals the oracle's brute force
ordered by (d2, index) -- the canonical semantics, stricter than the
reference's own acceptance rule (test/test_self_knn.cpp:153-266), which
tolerates tie swaps."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "bin")
CONFIGS = [("selfknn_runner_i64_periodic", O.CT_U64, 1), ("selfknn_runner_f64", O.CT_F64, 0)]
DATASETS = [(O.GEN_UNIFORM, 42), (O.GEN_CLUSTERED, 12345), (O.GEN_DEGENERATE, 999)]


@pytest.mark.parametrize("exe,ct,periodic", CONFIGS, ids=[c[0] for c in CONFIGS])
@pytest.mark.parametrize("kind,seed", DATASETS, ids=["uniform42", "clustered12345",
                                                     "degenerate999"])
def test_header_runner_matches_oracle(orc, tmp_path, exe, ct, periodic, kind, seed):
    path = os.path.join(BIN, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built: run `make -C tests/host` (build() does)")
    x, y, z = O.gen_reference(kind, 1000, seed, ct, bits=21)
    n = len(x)
    pts = tmp_path / "points.bin"
    with open(pts, "wb") as f:
        f.write(np.uint64(n).tobytes())
        for a in (x, y, z):
            f.write(np.ascontiguousarray(a).tobytes())
    out = subprocess.run([path, str(pts), str(tmp_path)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "Instantiate self_knn_query with larger _MAX_K_." in out.stdout
    t = O.build_bvh(ct, x, y, z, sfc=O.SFC_PH, bits=21, periodic=periodic)
    dd = O.dist_dtype(ct)
    ids = np.fromfile(tmp_path / "subset_ids.bin", np.int32)
    assert ids[0] == 0 and ids[-1] == n - 1 and n // 2 in ids

    def load(name, rows, k):
        r = np.fromfile(tmp_path / f"{name}_idx.bin", np.uint64).reshape(rows, k)
        d = np.fromfile(tmp_path / f"{name}_d2.bin", dd).reshape(rows, k)
        return r, d

    q = (x[ids], y[ids], z[ids])
    for k in (16, 32):
        for excl in (True, False):
            tag = f"k{k}_{'excl' if excl else 'incl'}"
            r, d = load(f"self_{tag}", n, k)
            rb, db = O.knn_bruteforce(t, k, excl)
            assert np.array_equal(d.view(np.uint8), db.view(np.uint8)), tag
            assert np.array_equal(r, rb), tag
            r, d = load(f"subset_{tag}", len(ids), k)
            rb, db = O.knn_bruteforce(t, k, excl, leaf_ids=ids)
            assert np.array_equal(d.view(np.uint8), db.view(np.uint8)) and np.array_equal(r, rb), tag
        r, d = load(f"ext_k{k}", len(ids), k)
        rb, db = O.knn_bruteforce(t, k, queries=q)
        assert np.array_equal(d.view(np.uint8), db.view(np.uint8)) and np.array_equal(r, rb), k
