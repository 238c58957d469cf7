"""CPU oracle for the HLBVH build and neighbour queries -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this package; the product (``fasttree_amd``) never does.  It wraps
``liboracle.so`` (built from ``oracle.cpp`` by ``oracle/Makefile``), a C++
restatement of the reference's algorithm in which every function cites the
reference file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CT_F32, CT_F64, CT_U32, CT_U64 = 0, 1, 2, 3
SFC_MORTON, SFC_PH = 0, 1
COORD_DTYPES = {CT_F32: np.float32, CT_F64: np.float64, CT_U32: np.uint32, CT_U64: np.uint64}
GEN_UNIFORM, GEN_CLUSTERED, GEN_DEGENERATE, GEN_GRID = 0, 1, 2, 3


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []))
    return LIB_PATH


class _Tree(C.Structure):
    _fields_ = [
        ("min_x", C.c_void_p), ("max_x", C.c_void_p), ("min_y", C.c_void_p),
        ("max_y", C.c_void_p), ("min_z", C.c_void_p), ("max_z", C.c_void_p),
        ("left", C.c_void_p), ("right", C.c_void_p), ("parent", C.c_void_p),
        ("id", C.c_void_p), ("ghost", C.c_void_p), ("orig", C.c_void_p),
        ("n", C.c_uint64), ("ct", C.c_int32), ("sfc", C.c_int32), ("bits", C.c_int32),
        ("periodic", C.c_int32), ("orig_indices", C.c_int32), ("pad", C.c_int32),
        ("box", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dist_dtype(ct):
    return np.float64 if ct in (CT_F32, CT_F64) else np.uint64


@dataclass
class Tree:
    """Host TreeSoA (src/hlbvh.hpp:609-673) plus the sorted keys and node ranges."""
    ct: int
    n: int
    sfc: int = SFC_MORTON
    bits: int = 21
    periodic: int = 0
    orig_indices: int = 0
    box: float = 0.0
    arrays: dict = field(default_factory=dict)

    @classmethod
    def empty(cls, ct, n, **cfg):
        t = cls(ct=ct, n=n, **cfg)
        dt = COORD_DTYPES[ct]
        m = max(2 * n - 1, 1)
        for name in ("min_x", "max_x", "min_y", "max_y", "min_z", "max_z"):
            t.arrays[name] = np.zeros(m, dt)
        t.arrays["left"] = np.full(max(n - 1, 1), -1, np.int32)
        t.arrays["right"] = np.full(max(n - 1, 1), -1, np.int32)
        t.arrays["parent"] = np.full(m, -1, np.int32)
        t.arrays["id"] = np.zeros(m, np.uint32)
        t.arrays["ghost"] = np.zeros(m, np.int8)
        t.arrays["orig"] = np.zeros(m, np.int32)
        t.arrays["sorted_keys"] = np.zeros(max(n, 1), np.uint64)
        t.arrays["range_first"] = np.zeros(max(n - 1, 1), np.int32)
        t.arrays["range_last"] = np.zeros(max(n - 1, 1), np.int32)
        return t

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays", {})
        if name in arrays:
            return arrays[name]
        raise AttributeError(name)

    def cstruct(self):
        a = self.arrays
        s = _Tree()
        for name in ("min_x", "max_x", "min_y", "max_y", "min_z", "max_z", "left", "right",
                     "parent", "id", "ghost", "orig"):
            # a metric-only tree (no arrays) serves point_d2
            setattr(s, name, a[name].ctypes.data if name in a else None)
        s.n, s.ct, s.sfc, s.bits = self.n, self.ct, self.sfc, self.bits
        s.periodic, s.orig_indices, s.box = self.periodic, self.orig_indices, self.box
        return s

    def leaf_coords(self):
        o = self.n - 1
        return self.min_x[o:], self.min_y[o:], self.min_z[o:]


def set_threads(t: int) -> int:
    return lib().or_set_threads(int(t))


def bbox(ct, x, y, z):
    out = np.zeros(6, COORD_DTYPES[ct])
    rc = lib().or_bbox(ct, _p(x), _p(y), _p(z), C.c_size_t(len(x)), _p(out))
    assert rc == 0
    return out


def encode(ct, sfc, bits, x, y, z, bbox6=None):
    keys = np.zeros(len(x), np.uint64)
    b = None if bbox6 is None else np.ascontiguousarray(bbox6, COORD_DTYPES[ct])
    rc = lib().or_encode(ct, sfc, bits, _p(x), _p(y), _p(z), C.c_size_t(len(x)), _p(b), _p(keys))
    assert rc == 0
    return keys


def sort(keys):
    n = len(keys)
    perm = np.zeros(n, np.uint32)
    srt = np.zeros(n, np.uint64)
    lib().or_sort(_p(keys), C.c_size_t(n), _p(perm), _p(srt))
    return perm, srt


def karras(sorted_keys):
    n = len(sorted_keys)
    left = np.zeros(n - 1, np.int32)
    right = np.zeros(n - 1, np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    rf = np.zeros(n - 1, np.int32)
    rl = np.zeros(n - 1, np.int32)
    rc = lib().or_karras(_p(sorted_keys), C.c_size_t(n), _p(left), _p(right), _p(parent),
                         _p(rf), _p(rl))
    assert rc == 0
    return left, right, parent, rf, rl


def build_bvh(ct, x, y, z, sfc=SFC_MORTON, bits=21, periodic=0, orig_indices=0, box=0.0,
              ids=None, ghost=None, bbox6=None, keys=None) -> Tree:
    n = len(x)
    t = Tree.empty(ct, n, sfc=sfc, bits=bits, periodic=periodic, orig_indices=orig_indices,
                   box=box)
    dt = COORD_DTYPES[ct]
    x, y, z = (np.ascontiguousarray(a, dt) for a in (x, y, z))
    s = t.cstruct()
    b = None if bbox6 is None else np.ascontiguousarray(bbox6, dt)
    rc = lib().or_build_bvh(C.byref(s), _p(x), _p(y), _p(z), _p(ids), _p(ghost), C.c_size_t(n),
                            _p(b), _p(keys), _p(t.sorted_keys), _p(t.range_first),
                            _p(t.range_last))
    assert rc == 0
    return t


def knn_bruteforce(t: Tree, k, exclude_self=True, queries=None, leaf_ids=None):
    """Canonical k nearest: ordered by (d2, returned index)."""
    if queries is not None:
        qx, qy, qz = (np.ascontiguousarray(a, COORD_DTYPES[t.ct]) for a in queries)
        nq = len(qx)
        ex = 0
    else:
        qx = qy = qz = None
        nq = t.n if leaf_ids is None else len(leaf_ids)
        ex = int(bool(exclude_self))
    lid = None if leaf_ids is None else np.ascontiguousarray(leaf_ids, np.int32)
    res = np.zeros(nq * k, np.uint64)
    d = np.zeros(nq * k, dist_dtype(t.ct))
    s = t.cstruct()
    rc = lib().or_knn_bruteforce(C.byref(s), _p(qx), _p(qy), _p(qz), _p(lid), nq, k, ex,
                                 _p(res), _p(d))
    assert rc == 0
    return res.reshape(nq, k), d.reshape(nq, k)


def knn_reference(t: Tree, k, exclude_self=True, queries=None, leaf_ids=None, kind=0,
                  max_k=32, counters=False):
    """The reference's traversal and heap mechanics (tie picks, 64-entry stack)."""
    if queries is not None:
        qx, qy, qz = (np.ascontiguousarray(a, COORD_DTYPES[t.ct]) for a in queries)
        nq = len(qx)
        ex = 0
    else:
        qx = qy = qz = None
        nq = t.n if leaf_ids is None else len(leaf_ids)
        ex = int(bool(exclude_self))
    lid = None if leaf_ids is None else np.ascontiguousarray(leaf_ids, np.int32)
    res = np.zeros(nq * k, np.uint64)
    d = np.zeros(nq * k, dist_dtype(t.ct))
    cnt = np.zeros(2 * nq, np.uint64) if counters else None
    s = t.cstruct()
    rc = lib().or_knn_reference(C.byref(s), _p(qx), _p(qy), _p(qz), _p(lid), nq, k, ex, kind,
                                max_k, _p(res), _p(d), _p(cnt))
    assert rc == 0
    if counters:
        return res.reshape(nq, k), d.reshape(nq, k), cnt.reshape(nq, 2)
    return res.reshape(nq, k), d.reshape(nq, k)


def knn_reference_grouped(t: Tree, k, exclude_self=True, leaf_ids=None, width=8,
                          visits=False):
    """The reference's grouped self-kNN as a CPU runs it: `width` lanes in
    lockstep (native vector width: 8 for 64-bit, 16 for 32-bit coordinates)."""
    nq = t.n if leaf_ids is None else len(leaf_ids)
    lid = None if leaf_ids is None else np.ascontiguousarray(leaf_ids, np.int32)
    res = np.zeros(nq * k, np.uint64)
    d = np.zeros(nq * k, dist_dtype(t.ct))
    cnt = np.zeros(nq, np.uint64)
    s = t.cstruct()
    rc = lib().or_knn_reference_grouped(C.byref(s), _p(lid), nq, k, int(bool(exclude_self)),
                                        width, _p(res), _p(d), _p(cnt))
    assert rc == 0
    if visits:
        return res.reshape(nq, k), d.reshape(nq, k), cnt
    return res.reshape(nq, k), d.reshape(nq, k)


def range_reference(t: Tree, qx, qy, qz, rmin, rmax, max_res):
    dt = COORD_DTYPES[t.ct]
    qx, qy, qz, rmin, rmax = (np.ascontiguousarray(a, dt) for a in (qx, qy, qz, rmin, rmax))
    nq = len(qx)
    res = np.full(nq * max(max_res, 1), -7, np.int32)
    cnt = np.zeros(nq, np.int32)
    s = t.cstruct()
    rc = lib().or_range_reference(C.byref(s), _p(qx), _p(qy), _p(qz), _p(rmin), _p(rmax), nq,
                                  _p(res), _p(cnt), max_res, None, None)
    assert rc == 0
    return res.reshape(nq, max(max_res, 1)), cnt


def range_csr(t: Tree, qx, qy, qz, rmin, rmax):
    """Counts plus CSR neighbour lists in ascending leaf-rank order."""
    dt = COORD_DTYPES[t.ct]
    qx, qy, qz, rmin, rmax = (np.ascontiguousarray(a, dt) for a in (qx, qy, qz, rmin, rmax))
    nq = len(qx)
    cnt = np.zeros(nq, np.int32)
    s = t.cstruct()
    lib().or_range_reference(C.byref(s), _p(qx), _p(qy), _p(qz), _p(rmin), _p(rmax), nq, None,
                             _p(cnt), 0, None, None)
    rp = np.zeros(nq + 1, np.int64)
    np.cumsum(cnt, out=rp[1:])
    cols = np.zeros(max(int(rp[-1]), 1), np.int32)
    lib().or_range_reference(C.byref(s), _p(qx), _p(qy), _p(qz), _p(rmin), _p(rmax), nq, None,
                             None, 0, _p(rp), _p(cols))
    return rp, cols[: rp[-1]]


def point_d2(t: Tree, a, b):
    dt = COORD_DTYPES[t.ct]
    a = np.ascontiguousarray(a, dt)
    b = np.ascontiguousarray(b, dt)
    out = np.zeros(1, dist_dtype(t.ct))
    s = t.cstruct()
    lib().or_point_d2(C.byref(s), _p(a), _p(b), _p(out))
    return out[0]


def heap_u64(kind, k, cap, d, idx):
    d = np.ascontiguousarray(d, np.uint64)
    idx = np.ascontiguousarray(idx, np.int32)
    od = np.zeros(cap, np.uint64)
    oi = np.zeros(cap, np.int32)
    cnt = lib().or_heap_u64(kind, k, cap, _p(d), _p(idx), len(d), _p(od), _p(oi))
    return od, oi, cnt


def gen_reference(kind, n, seed, ct, bits=21, box_min=0.0, box_size=100.0):
    """The reference tests' seeded particle sets (std::mt19937 + libstdc++ distributions)."""
    if kind == GEN_GRID:
        n = 1000
    dt = COORD_DTYPES[ct]
    x, y, z = np.zeros(n, dt), np.zeros(n, dt), np.zeros(n, dt)
    lib().or_gen_reference(kind, C.c_size_t(n), C.c_uint32(seed), C.c_double(box_min),
                           C.c_double(box_size), ct, bits, _p(x), _p(y), _p(z))
    return x, y, z


def fill_particle_coords(px, py, pz, ct, bits=21):
    """fill_particle_coords (test/benchmark/BenchmarkUtils.hpp:164-205): float32 or
    float64 positions -> coordinates of type ct; returns (x, y, z, box_min, box_size)."""
    in_dt = np.float64 if np.asarray(px).dtype == np.float64 else np.float32
    px, py, pz = (np.ascontiguousarray(a, in_dt) for a in (px, py, pz))
    n = len(px)
    dt = COORD_DTYPES[ct]
    x, y, z = np.zeros(n, dt), np.zeros(n, dt), np.zeros(n, dt)
    box = np.zeros(2, np.float64)
    rc = lib().or_fill_particle_coords(CT_F64 if in_dt == np.float64 else CT_F32, _p(px), _p(py),
                                       _p(pz), C.c_size_t(n), ct, bits, _p(x), _p(y), _p(z),
                                       _p(box))
    assert rc == 0
    return x, y, z, float(box[0]), float(box[1])
