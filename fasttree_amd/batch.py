"""Batched trees and mixed-k queries in one hipGraph (ft_batch_* of the C ABI).

Many trees of uneven size are built, and query requests of mixed k over them
are answered, as one step: per-tree builds on parallel graph branches, one
curve-order sort of all queries, persistent kernels per heap tier draining a
work list over all trees, and an exact re-run pass.  The autotuner picks the
tier kernel and leaf bucket per (tree-size class, k tier) and the build branch
count by timing them on the bound data.  Results equal knn_query's canonical
results on each tree.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import DeviceArray, TreeSoA, _Tree, _check, _ptr, _sh, dist_dtype, lib


class BatchRequest(C.Structure):
    _fields_ = [("tree", C.c_int32), ("k", C.c_int32), ("query_first", C.c_int64),
                ("num_queries", C.c_int64), ("result_first", C.c_int64)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    vp, i32 = C.c_void_p, C.c_int32
    L.ft_batch_create.argtypes = [C.POINTER(vp), i32]
    L.ft_batch_destroy.argtypes = [vp]
    L.ft_batch_set_trees.argtypes = [vp, i32, C.POINTER(C.POINTER(_Tree)), C.POINTER(vp),
                                     C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_size_t)]
    L.ft_batch_set_queries.argtypes = [vp, i32, C.POINTER(BatchRequest), vp, vp, vp, vp, vp]
    L.ft_batch_set_stats.argtypes = [vp, vp]
    L.ft_batch_tune.argtypes = [vp, vp]
    L.ft_batch_run.argtypes = [vp, vp]
    L.ft_batch_capture.argtypes = [vp, vp]
    L.ft_batch_launch.argtypes = [vp, vp]
    L.ft_batch_tuning.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.ft_batch_set_tier.argtypes = [vp, i32, i32, C.c_char_p, i32, i32]
    _bound = True


class Batch:
    """A set of trees (optionally rebuilt every step from device particles) and
    query requests over them.  Device arrays given here must outlive the batch."""

    def __init__(self, coord_type: int):
        _bind()
        h = C.c_void_p()
        _check(lib().ft_batch_create(C.byref(h), int(coord_type)))
        self.h = h.value
        self.coord_type = coord_type
        self._keep = []

    def set_trees(self, trees, particles=None):
        """trees: list of TreeSoA; particles: list of (x, y, z) DeviceArrays
        (the step builds every tree from them), or None (trees already built)."""
        nt = len(trees)
        arr = (C.POINTER(_Tree) * nt)(*[C.pointer(t.c) for t in trees])
        self._trees = list(trees)
        if particles is None:
            _check(lib().ft_batch_set_trees(self.h, nt, arr, None, None, None, None))
        else:
            xs = (C.c_void_p * nt)(*[p[0].ptr for p in particles])
            ys = (C.c_void_p * nt)(*[p[1].ptr for p in particles])
            zs = (C.c_void_p * nt)(*[p[2].ptr for p in particles])
            ns = (C.c_size_t * nt)(*[p[0].n for p in particles])
            self._keep.append(list(particles))
            _check(lib().ft_batch_set_trees(self.h, nt, arr, xs, ys, zs, ns))
        self._tree_arr = arr

    def set_queries(self, requests, qx: DeviceArray, qy: DeviceArray, qz: DeviceArray,
                    results: DeviceArray, dists: DeviceArray):
        """requests: list of (tree, k, query_first, num_queries, result_first)."""
        nr = len(requests)
        arr = (BatchRequest * max(nr, 1))(*[BatchRequest(*map(int, r)) for r in requests])
        self._req = arr
        self._keep.append((qx, qy, qz, results, dists))
        _check(lib().ft_batch_set_queries(self.h, nr, arr, qx.ptr, qy.ptr, qz.ptr, results.ptr,
                                          dists.ptr))

    def set_stats(self, stats: DeviceArray | None):
        _check(lib().ft_batch_set_stats(self.h, _ptr(stats)))

    def set_tier(self, n_class: int, k_tier: int, tier: str, bucket: int = 0,
                 queries_per_item: int = 0):
        """Pins one (tree-size class, k tier) group's tier kernel (a name as
        tuning() reports it), leaf bucket and queries per item; tune() keeps it."""
        _check(lib().ft_batch_set_tier(self.h, int(n_class), int(k_tier), tier.encode(),
                                       int(bucket), int(queries_per_item)))

    def tune(self, stream=None):
        _check(lib().ft_batch_tune(self.h, _sh(stream)))

    def run(self, stream=None):
        _check(lib().ft_batch_run(self.h, _sh(stream)))

    def capture(self, stream=None):
        _check(lib().ft_batch_capture(self.h, _sh(stream)))

    def launch(self, stream=None):
        _check(lib().ft_batch_launch(self.h, _sh(stream)))

    def tuning(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        _check(lib().ft_batch_tuning(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def __del__(self):
        try:
            if self.h:
                lib().ft_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


def plan_requests(sizes_q, ks, trees_of):
    """Lays out requests contiguously: query and result offsets from counts.
    sizes_q[i] queries with k = ks[i] against tree trees_of[i]."""
    reqs, qf, rf = [], 0, 0
    for nq, k, t in zip(sizes_q, ks, trees_of):
        reqs.append((int(t), int(k), qf, int(nq), rf))
        qf += int(nq)
        rf += int(nq) * int(k)
    return reqs, qf, rf


__all__ = ["Batch", "BatchRequest", "plan_requests", "TreeSoA", "dist_dtype", "np"]
