"""Python host mirror of FastTree's HLBVH build / neighbour-query API on MI355X.

The functions keep the reference's names, argument meaning and error behaviour
(src/hlbvh.hpp: build_bvh, self_knn_query, self_knn_query_subset, knn_query,
range_query, compute_bbox, sfc_encode, build_tree; TreeSoA, BvhScratch) and
call the gfx950 kernels through the C ABI of ``libfasttree_amd.so``
(include/fasttree_c.h).  A stream handle replaces the in-order SYCL queue.

The library must be built (``python -m fasttree_amd.build``); importing the
kernels without it raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfasttree_amd.so")

FT_OK, FT_EINVAL, FT_ERANGE, FT_ELDS, FT_ENOMEM, FT_EHIP = 0, 1, 2, 3, 4, 5
COORD_F32, COORD_F64, COORD_U32, COORD_U64 = 0, 1, 2, 3
SFC_MORTON, SFC_PEANO_HILBERT = 0, 1
MODE_CANONICAL, MODE_REFERENCE = 0, 1
ALGO_AUTO, ALGO_PLAIN, ALGO_GROUPED, ALGO_PACKET, ALGO_WAVE = 0, 1, 2, 3, 4
TOPK_AUTO, TOPK_REGISTERS, TOPK_LDS_HEAP = 0, 1, 2

COORD_DTYPES = {COORD_F32: np.float32, COORD_F64: np.float64, COORD_U32: np.uint32,
                COORD_U64: np.uint64}
_CT_BY_DTYPE = {np.dtype(v): k for k, v in COORD_DTYPES.items()}


def coord_type_of(dtype) -> int:
    return _CT_BY_DTYPE[np.dtype(dtype)]


def dist_dtype(ct: int):
    """dist_t: double for FP coordinates, uint64 for integer coordinates."""
    return np.float64 if ct in (COORD_F32, COORD_F64) else np.uint64


def missing_dist(ct: int):
    """type_identity_max<dist_t>() written for a missing neighbour."""
    return -np.finfo(np.float64).max if ct in (COORD_F32, COORD_F64) else np.uint64(0)


MISSING_INDEX = np.uint64(0xFFFFFFFFFFFFFFFF)


class Config(C.Structure):
    """Runtime form of the reference's compile-time switches."""
    _fields_ = [("sfc", C.c_int32), ("bits_per_dim", C.c_int32), ("periodic", C.c_int32),
                ("return_orig_indices", C.c_int32), ("box_length", C.c_double)]

    def __init__(self, sfc=SFC_PEANO_HILBERT, bits_per_dim=21, periodic=False,
                 return_orig_indices=False, box_length=0.0):
        super().__init__(int(sfc), int(bits_per_dim), int(bool(periodic)),
                         int(bool(return_orig_indices)), float(box_length))


class _Tree(C.Structure):
    _fields_ = [
        ("min_x", C.c_void_p), ("max_x", C.c_void_p), ("min_y", C.c_void_p),
        ("max_y", C.c_void_p), ("min_z", C.c_void_p), ("max_z", C.c_void_p),
        ("left_child", C.c_void_p), ("right_child", C.c_void_p), ("parent", C.c_void_p),
        ("id", C.c_void_p), ("is_ghost", C.c_void_p), ("orig_idx", C.c_void_p),
        ("num_leaves", C.c_size_t), ("num_internal", C.c_size_t),
        ("coord_type", C.c_int32), ("reserved0", C.c_int32), ("config", Config),
        ("sorted_keys", C.c_void_p), ("range_first", C.c_void_p), ("range_last", C.c_void_p),
        ("capacity", C.c_size_t), ("block", C.c_void_p),
    ]


class QueryOpts(C.Structure):
    _fields_ = [("max_k", C.c_int32), ("mode", C.c_int32), ("algorithm", C.c_int32),
                ("group_width", C.c_int32), ("leaf_bucket", C.c_int32), ("topk", C.c_int32),
                ("seed_bound", C.c_void_p), ("stats", C.c_void_p), ("window", C.c_int32),
                ("reserved2", C.c_int32)]


EXPORTS = [
    "ft_last_error", "ft_version", "ft_device_count", "ft_malloc", "ft_free", "ft_memcpy_h2d",
    "ft_memcpy_d2h", "ft_memcpy_d2d", "ft_memset", "ft_copy_kernel", "ft_stream_create",
    "ft_stream_destroy", "ft_stream_release", "ft_stream_synchronize", "ft_graph_begin", "ft_graph_end", "ft_graph_launch",
    "ft_graph_destroy", "ft_event_create", "ft_event_destroy", "ft_event_record",
    "ft_event_elapsed_ms", "ft_tree_alloc", "ft_tree_ensure", "ft_tree_free",
    "ft_tree_serialized_size", "ft_tree_serialize", "ft_tree_deserialize",
    "ft_scratch_create", "ft_scratch_destroy", "ft_scratch_set_profiling", "ft_scratch_stage_ms",
    "ft_scratch_set_radix_bits",
    "ft_compute_bbox", "ft_sfc_encode", "ft_build_tree", "ft_fill_particle_coords",
    "ft_build_bvh", "ft_self_knn_query",
    "ft_self_knn_query_subset", "ft_self_knn_query_noalloc", "ft_knn_query", "ft_range_query", "ft_range_query_csr",
    "ft_mlp_num_params", "ft_mlp_create", "ft_mlp_destroy", "ft_mlp_set", "ft_mlp_get",
    "ft_knn_features", "ft_mlp_forward", "ft_mlp_grad", "ft_mlp_train_step",
    "ft_mlp_last_loss", "ft_mlp_seed", "ft_batch_create", "ft_batch_destroy",
    "ft_batch_set_trees", "ft_batch_set_queries", "ft_batch_set_stats", "ft_batch_tune",
    "ft_batch_run", "ft_batch_capture", "ft_batch_launch", "ft_batch_tuning", "ft_batch_set_tier",
]

_lib = None


def lib():
    """Loads libfasttree_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m fasttree_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, sz, i32 = C.c_void_p, C.c_size_t, C.c_int
        L.ft_last_error.restype = C.c_char_p
        L.ft_version.restype = C.c_char_p
        L.ft_malloc.argtypes = [C.POINTER(C.c_void_p), sz]
        L.ft_free.argtypes = [vp]
        for f in ("ft_memcpy_h2d", "ft_memcpy_d2h", "ft_memcpy_d2d"):
            getattr(L, f).argtypes = [vp, vp, sz, vp]
        L.ft_memset.argtypes = [vp, i32, sz, vp]
        L.ft_copy_kernel.argtypes = [vp, vp, sz, vp]
        L.ft_stream_create.argtypes = [C.POINTER(C.c_void_p)]
        L.ft_stream_destroy.argtypes = [vp]
        L.ft_stream_release.argtypes = [vp]
        L.ft_stream_synchronize.argtypes = [vp]
        L.ft_graph_begin.argtypes = [vp]
        L.ft_graph_end.argtypes = [vp, C.POINTER(C.c_void_p)]
        L.ft_graph_launch.argtypes = [vp, vp]
        L.ft_graph_destroy.argtypes = [vp]
        L.ft_event_create.argtypes = [C.POINTER(C.c_void_p)]
        L.ft_event_destroy.argtypes = [vp]
        L.ft_event_record.argtypes = [vp, vp]
        L.ft_event_elapsed_ms.argtypes = [vp, vp, C.POINTER(C.c_float)]
        L.ft_tree_alloc.argtypes = [C.POINTER(_Tree), sz, C.c_int32, C.POINTER(Config), vp]
        L.ft_tree_ensure.argtypes = [C.POINTER(_Tree), sz, C.c_int32, C.POINTER(Config), vp]
        L.ft_tree_free.argtypes = [C.POINTER(_Tree), vp]
        L.ft_tree_serialized_size.argtypes = [C.POINTER(_Tree), C.POINTER(C.c_size_t)]
        L.ft_tree_serialize.argtypes = [vp, C.POINTER(_Tree), vp, sz]
        L.ft_tree_deserialize.argtypes = [vp, C.POINTER(_Tree), vp, sz]
        L.ft_scratch_create.argtypes = [C.POINTER(C.c_void_p)]
        L.ft_scratch_destroy.argtypes = [vp]
        L.ft_scratch_set_profiling.argtypes = [vp, i32]
        L.ft_scratch_set_radix_bits.argtypes = [vp, i32]
        L.ft_scratch_stage_ms.argtypes = [vp, C.POINTER(C.c_double), i32, C.POINTER(C.c_int)]
        L.ft_compute_bbox.argtypes = [vp, C.c_int32, vp, vp, vp, sz, vp]
        L.ft_sfc_encode.argtypes = [vp, C.c_int32, C.POINTER(Config), vp, vp, vp, sz, vp, vp]
        L.ft_build_tree.argtypes = [vp, C.POINTER(_Tree), vp, vp, vp, vp, vp, vp, vp]
        L.ft_build_bvh.argtypes = [vp, C.POINTER(_Tree), vp, vp, vp, vp, vp, vp, sz, vp, vp]
        L.ft_fill_particle_coords.argtypes = [vp, C.c_int32, vp, vp, vp, sz, C.c_int32, i32, vp, vp,
                                              vp, vp]
        L.ft_self_knn_query.argtypes = [vp, C.POINTER(_Tree), i32, vp, vp, i32,
                                        C.POINTER(QueryOpts)]
        L.ft_self_knn_query_noalloc.argtypes = [vp, C.POINTER(_Tree), i32, vp, vp, i32,
                                                C.POINTER(QueryOpts)]
        L.ft_self_knn_query_subset.argtypes = [vp, C.POINTER(_Tree), vp, i32, i32, vp, vp, i32,
                                               C.POINTER(QueryOpts)]
        L.ft_knn_query.argtypes = [vp, C.POINTER(_Tree), vp, vp, vp, i32, i32, vp, vp,
                                   C.POINTER(QueryOpts)]
        L.ft_range_query.argtypes = [vp, C.POINTER(_Tree), vp, vp, vp, vp, vp, i32, vp, vp, i32]
        L.ft_range_query_csr.argtypes = [vp, C.POINTER(_Tree), vp, vp, vp, vp, vp, i32, vp, vp,
                                         C.c_int64]
        _lib = L
    return _lib


def _check(rc: int):
    if rc == FT_OK:
        return
    msg = lib().ft_last_error().decode(errors="replace")
    if rc == FT_EINVAL:
        raise ValueError(msg)          # std::invalid_argument
    if rc == FT_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)            # std::runtime_error / HIP errors


# ---------------------------------------------------------------------------
# Device memory and streams (plumbing for tests and benchmarks)
# ---------------------------------------------------------------------------
class DeviceArray:
    """A device allocation with a dtype and length; frees itself."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        _check(lib().ft_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @classmethod
    def from_host(cls, a, dtype=None, stream=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, a.dtype)
        if a.size:
            _check(lib().ft_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes, stream))
        return d

    def to_host(self, stream=None):
        out = np.empty(self.n, self.dtype)
        if self.n:
            _check(lib().ft_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream))
        return out

    def to_host_range(self, start: int, count: int, stream=None):
        """Copies elements [start, start + count) to the host."""
        out = np.empty(int(count), self.dtype)
        if count:
            _check(lib().ft_memcpy_d2h(out.ctypes.data,
                                       int(self.ptr) + int(start) * self.dtype.itemsize,
                                       out.nbytes, stream))
        return out

    def zero(self, stream=None):
        _check(lib().ft_memset(self.ptr, 0, self.nbytes, stream))
        return self

    def copy_from(self, a, stream=None):
        """Overwrites the array with host data of the same length."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError("copy_from: length differs")
        if a.size:
            _check(lib().ft_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream))
        return self

    def free(self):
        if getattr(self, "ptr", None):
            lib().ft_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a):
    """Pointer of a DeviceArray or a contiguous numpy array (host: staged by the library)."""
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        assert a.flags.c_contiguous
        return a.ctypes.data
    if isinstance(a, int):
        return a
    raise TypeError(type(a))


class Stream:
    def __init__(self):
        h = C.c_void_p()
        _check(lib().ft_stream_create(C.byref(h)))
        self.handle = h.value

    def synchronize(self):
        _check(lib().ft_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().ft_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _sh(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class Graph:
    """The library calls issued on `stream` inside ``with Graph(stream) as g:``,
    captured as a hipGraph; ``g.launch()`` replays them.  Run the step once
    eagerly first, so that no buffer grows during the capture."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.exec = None

    def __enter__(self):
        _check(lib().ft_graph_begin(_sh(self.stream)))
        return self

    def __exit__(self, et, ev, tb):
        h = C.c_void_p()
        rc = lib().ft_graph_end(_sh(self.stream), C.byref(h))
        if et is None:
            _check(rc)
            self.exec = h.value
        return False

    def launch(self, stream=None):
        _check(lib().ft_graph_launch(self.exec, _sh(stream if stream is not None else self.stream)))

    def __del__(self):
        try:
            if self.exec:
                lib().ft_graph_destroy(self.exec)
                self.exec = None
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        _check(lib().ft_event_create(C.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        _check(lib().ft_event_record(self.handle, _sh(stream)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _check(lib().ft_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                lib().ft_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def device_count() -> int:
    c = C.c_int()
    _check(lib().ft_device_count(C.byref(c)))
    return c.value


# ---------------------------------------------------------------------------
# particles / TreeSoA / BvhScratch
# ---------------------------------------------------------------------------
@dataclass
class particles:
    """particles<coord_t> (src/hlbvh.hpp:26-34): SoA positions, optional id / is_ghost."""
    pos_x: object
    pos_y: object
    pos_z: object
    id: object = None
    is_ghost: object = None

    def __len__(self):
        return self.pos_x.n if isinstance(self.pos_x, DeviceArray) else len(self.pos_x)

    @property
    def coord_type(self):
        dt = self.pos_x.dtype
        return coord_type_of(dt)


class TreeSoA:
    """TreeSoA (src/hlbvh.hpp:609-673) in device memory.

    ``TreeSoA(n, coord_type, config)`` allocates for n leaves; ``free()``
    releases it.  ``ensure(n)`` follows TreePool::ensure (grow by n/2 + 1024).
    """

    def __init__(self, n: int, coord_type: int, config: Config | None = None, stream=None):
        self.c = _Tree()
        self.config = config if config is not None else Config()
        _check(lib().ft_tree_alloc(C.byref(self.c), int(n), int(coord_type),
                                   C.byref(self.config), _sh(stream)))

    @property
    def num_leaves(self):
        return self.c.num_leaves

    @property
    def num_internal(self):
        return self.c.num_internal

    @property
    def coord_type(self):
        return self.c.coord_type

    def ensure(self, n: int, stream=None):
        _check(lib().ft_tree_ensure(C.byref(self.c), int(n), self.c.coord_type,
                                    C.byref(self.config), _sh(stream)))
        return self

    def free(self, stream=None):
        if self.c.block:
            _check(lib().ft_tree_free(C.byref(self.c), _sh(stream)))

    def serialize(self, stream=None) -> bytes:
        """The tree as bytes: header, the reference's MPI transfer sequence
        (dims, boxes, children, parents), then ids, ghost flags, original
        indices, sorted keys and node ranges."""
        nb = C.c_size_t()
        _check(lib().ft_tree_serialized_size(C.byref(self.c), C.byref(nb)))
        buf = np.empty(nb.value, np.uint8)
        _check(lib().ft_tree_serialize(_sh(stream), C.byref(self.c), buf.ctypes.data,
                                       buf.nbytes))
        return buf.tobytes()

    @classmethod
    def deserialize(cls, data, stream=None) -> "TreeSoA":
        """A tree on the device from `serialize` output."""
        buf = np.frombuffer(data, np.uint8)
        t = cls.__new__(cls)
        t.c = _Tree()
        _check(lib().ft_tree_deserialize(_sh(stream), C.byref(t.c), buf.ctypes.data, buf.nbytes))
        t.config = Config()
        C.memmove(C.byref(t.config), C.byref(t.c.config), C.sizeof(Config))
        return t

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def to_host(self) -> dict:
        """Copies every array to host numpy arrays (sizes per the reference layout)."""
        n = self.c.num_leaves
        if n == 0:
            return {}
        ct = self.c.coord_type
        dt = np.dtype(COORD_DTYPES[ct])
        m, inner = 2 * n - 1, max(n - 1, 0)
        out = {}

        def get(ptr, count, dtype):
            a = np.empty(count, dtype)
            if count:
                _check(lib().ft_memcpy_d2h(a.ctypes.data, ptr, a.nbytes, None))
            return a

        for name in ("min_x", "max_x", "min_y", "max_y", "min_z", "max_z"):
            out[name] = get(getattr(self.c, name), m, dt)
        out["left_child"] = get(self.c.left_child, inner, np.int32)
        out["right_child"] = get(self.c.right_child, inner, np.int32)
        out["parent"] = get(self.c.parent, m, np.int32)
        out["id"] = get(self.c.id, m, np.uint32)
        out["is_ghost"] = get(self.c.is_ghost, m, np.int8)
        out["orig_idx"] = get(self.c.orig_idx, m, np.int32)
        out["sorted_keys"] = get(self.c.sorted_keys, n, np.uint64)
        out["range_first"] = get(self.c.range_first, inner, np.int32)
        out["range_last"] = get(self.c.range_last, inner, np.int32)
        return out


class BvhScratch:
    """BvhScratch (src/hlbvh.hpp:720-745): persistent build buffers."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().ft_scratch_create(C.byref(h)))
        self.handle = h.value

    def set_profiling(self, on: bool):
        _check(lib().ft_scratch_set_profiling(self.handle, int(bool(on))))

    def set_radix_bits(self, bits: int):
        """Digit width of the build's radix sort: 8 or 9 bits per pass, or 0 (the
        default) for the width with fewer passes."""
        _check(lib().ft_scratch_set_radix_bits(self.handle, int(bits)))

    def stage_ms(self):
        ms = (C.c_double * 8)()
        n = C.c_int()
        _check(lib().ft_scratch_stage_ms(self.handle, ms, 8, C.byref(n)))
        names = ["bbox", "encode", "sort", "assemble", "assemble_top"]
        return {names[i]: ms[i] for i in range(n.value)}

    def free_all(self):
        if self.handle:
            lib().ft_scratch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free_all()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Build
# ---------------------------------------------------------------------------
def fill_particle_coords(pos_x, pos_y, pos_z, coord_type: int, bits_per_dim: int = 21,
                         stream=None):
    """fill_particle_coords (test/benchmark/BenchmarkUtils.hpp:164-205): snapshot
    positions (float32 as the HDF5 read gives them, or float64; numpy arrays or
    DeviceArrays) to tree coordinates of `coord_type`.  Integer types: one cube
    over all three axes, padded by 1.0001, then float_to_int_rep with
    `bits_per_dim` bits.  Returns (x, y, z, box_min, box_size), the arrays in the
    inputs' kind (host or device)."""
    dev = isinstance(pos_x, DeviceArray)
    if dev:
        in_dt = pos_x.dtype
        n = pos_x.n
    else:
        in_dt = np.dtype(np.float64 if np.asarray(pos_x).dtype == np.float64 else np.float32)
        pos_x, pos_y, pos_z = (np.ascontiguousarray(a, in_dt) for a in (pos_x, pos_y, pos_z))
        n = len(pos_x)
    in_type = COORD_F64 if in_dt == np.float64 else COORD_F32
    dt = COORD_DTYPES[coord_type]
    if dev:
        outs = [DeviceArray(n, dt) for _ in range(3)]
    else:
        outs = [np.zeros(n, dt) for _ in range(3)]
    box = np.zeros(2, np.float64)
    _check(lib().ft_fill_particle_coords(_sh(stream), in_type, _ptr(pos_x), _ptr(pos_y),
                                         _ptr(pos_z), int(n), int(coord_type), int(bits_per_dim),
                                         _ptr(outs[0]), _ptr(outs[1]), _ptr(outs[2]),
                                         box.ctypes.data))
    return outs[0], outs[1], outs[2], float(box[0]), float(box[1])


def compute_bbox(pos_x, pos_y, pos_z, n=None, coord_type=None, stream=None):
    """compute_bbox (src/hlbvh.hpp:462-578) -> (min_x, max_x, min_y, max_y, min_z, max_z)."""
    ct = coord_type if coord_type is not None else coord_type_of(pos_x.dtype)
    n = len(pos_x) if n is None else n
    if isinstance(pos_x, DeviceArray):
        n = pos_x.n if n is None else n
    out = np.zeros(6, COORD_DTYPES[ct])
    _check(lib().ft_compute_bbox(_sh(stream), ct, _ptr(pos_x), _ptr(pos_y), _ptr(pos_z), int(n),
                                 out.ctypes.data))
    return out


def sfc_encode(pos_x, pos_y, pos_z, config: Config, bbox=None, keys=None, stream=None):
    """sfc_encode (src/hlbvh.hpp:357-423): one 64-bit SFC key per particle."""
    ct = coord_type_of(pos_x.dtype)
    n = pos_x.n if isinstance(pos_x, DeviceArray) else len(pos_x)
    out = keys if keys is not None else np.zeros(n, np.uint64)
    b = None if bbox is None else np.ascontiguousarray(bbox, COORD_DTYPES[ct])
    _check(lib().ft_sfc_encode(_sh(stream), ct, C.byref(config), _ptr(pos_x), _ptr(pos_y),
                               _ptr(pos_z), int(n), _ptr(out), _ptr(b)))
    return out


def build_tree(tree: TreeSoA, sorted_keys, sorted_x=None, sorted_y=None, sorted_z=None,
               sorted_id=None, sorted_is_ghost=None, sorted_orig_idx=None, stream=None):
    """build_tree (src/hlbvh.hpp:773-982): Karras topology + refit over sorted keys."""
    sk = np.ascontiguousarray(sorted_keys, np.uint64) if isinstance(sorted_keys, np.ndarray) \
        else sorted_keys
    _check(lib().ft_build_tree(_sh(stream), C.byref(tree.c), _ptr(sk), _ptr(sorted_x),
                               _ptr(sorted_y), _ptr(sorted_z), _ptr(sorted_id),
                               _ptr(sorted_is_ghost), _ptr(sorted_orig_idx)))


def build_bvh(p: particles, tree: TreeSoA, scratch: BvhScratch | None = None, bbox=None,
              key=None, stream=None):
    """build_bvh (src/hlbvh.hpp:2165-2302)."""
    n = len(p)
    b = None if bbox is None else np.ascontiguousarray(bbox, COORD_DTYPES[tree.coord_type])
    _check(lib().ft_build_bvh(_sh(stream), C.byref(tree.c),
                              scratch.handle if scratch is not None else None,
                              _ptr(p.pos_x), _ptr(p.pos_y), _ptr(p.pos_z), _ptr(p.id),
                              _ptr(p.is_ghost), int(n), _ptr(b), _ptr(key)))


# ---------------------------------------------------------------------------
# Queries
# ---------------------------------------------------------------------------
def _opts(max_k=0, mode=MODE_CANONICAL, algorithm=ALGO_AUTO, group_width=0, leaf_bucket=0,
          seed_bound=None, stats=None, topk=0, window=0):
    return QueryOpts(int(max_k), int(mode), int(algorithm), int(group_width), int(leaf_bucket),
                     int(topk), _ptr(seed_bound), _ptr(stats), int(window), 0)


def _outputs(tree, nq, k, results, dists):
    if results is None:
        results = np.zeros(nq * k, np.uint64)
    if dists is None:
        dists = np.zeros(nq * k, dist_dtype(tree.coord_type))
    return results, dists


def _shape(x, nq, k):
    return x.reshape(nq, k) if isinstance(x, np.ndarray) else x


def self_knn_query(tree: TreeSoA, k: int, results=None, dists=None, exclude_self=True,
                   max_k=32, stream=None, **opts):
    """self_knn_query<_MAX_K_=32> (src/hlbvh.hpp:1929-1954).  Returns (indices, d2)."""
    n = tree.num_leaves
    results, dists = _outputs(tree, n, k, results, dists)
    o = _opts(max_k=max_k, **opts)
    _check(lib().ft_self_knn_query(_sh(stream), C.byref(tree.c), int(k), _ptr(results),
                                   _ptr(dists), int(bool(exclude_self)), C.byref(o)))
    return _shape(results, n, k), _shape(dists, n, k)


def self_knn_query_subset(tree: TreeSoA, leaf_ids, k: int, results=None, dists=None,
                          exclude_self=True, max_k=256, stream=None, **opts):
    """self_knn_query_subset<_MAX_K_=256> (src/hlbvh.hpp:1956-1983)."""
    ids = np.ascontiguousarray(leaf_ids, np.int32) if isinstance(leaf_ids, np.ndarray) \
        else leaf_ids
    nq = ids.n if isinstance(ids, DeviceArray) else len(ids)
    results, dists = _outputs(tree, nq, k, results, dists)
    o = _opts(max_k=max_k, **opts)
    _check(lib().ft_self_knn_query_subset(_sh(stream), C.byref(tree.c), _ptr(ids), int(nq),
                                          int(k), _ptr(results), _ptr(dists),
                                          int(bool(exclude_self)), C.byref(o)))
    return _shape(results, nq, k), _shape(dists, nq, k)


def knn_query(tree: TreeSoA, qx, qy, qz, k: int, results=None, dists=None, max_k=256,
              stream=None, **opts):
    """knn_query<_MAX_K_=256> (src/hlbvh.hpp:1405-1472) over external points."""
    dt = COORD_DTYPES[tree.coord_type]
    if not isinstance(qx, DeviceArray):
        qx, qy, qz = (np.ascontiguousarray(a, dt) for a in (qx, qy, qz))
    nq = qx.n if isinstance(qx, DeviceArray) else len(qx)
    results, dists = _outputs(tree, nq, k, results, dists)
    o = _opts(max_k=max_k, **opts)
    _check(lib().ft_knn_query(_sh(stream), C.byref(tree.c), _ptr(qx), _ptr(qy), _ptr(qz),
                              int(k), int(nq), _ptr(results), _ptr(dists), C.byref(o)))
    return _shape(results, nq, k), _shape(dists, nq, k)


def range_query(tree: TreeSoA, qx, qy, qz, r_min, r_max, max_results_per_query: int,
                results=None, result_counts=None, stream=None):
    """range_query (src/hlbvh.hpp:2006-2136): fixed-stride indices, counts = total hits."""
    dt = COORD_DTYPES[tree.coord_type]
    if not isinstance(qx, DeviceArray):
        qx, qy, qz, r_min, r_max = (np.ascontiguousarray(a, dt) for a in (qx, qy, qz, r_min,
                                                                           r_max))
    nq = qx.n if isinstance(qx, DeviceArray) else len(qx)
    if results is None:
        results = np.full(nq * max(max_results_per_query, 1), -7, np.int32)
    if result_counts is None:
        result_counts = np.zeros(nq, np.int32)
    _check(lib().ft_range_query(_sh(stream), C.byref(tree.c), _ptr(qx), _ptr(qy), _ptr(qz),
                                _ptr(r_min), _ptr(r_max), int(nq), _ptr(results),
                                _ptr(result_counts), int(max_results_per_query)))
    if isinstance(results, np.ndarray):
        results = results.reshape(nq, max(max_results_per_query, 1))
    return results, result_counts


def range_query_csr(tree: TreeSoA, qx, qy, qz, r_min, r_max, stream=None):
    """Range query with counts and a CSR neighbour list (ascending leaf rank per row)."""
    dt = COORD_DTYPES[tree.coord_type]
    if not isinstance(qx, DeviceArray):
        qx, qy, qz, r_min, r_max = (np.ascontiguousarray(a, dt) for a in (qx, qy, qz, r_min,
                                                                           r_max))
    nq = qx.n if isinstance(qx, DeviceArray) else len(qx)
    rp = np.zeros(nq + 1, np.int64)
    _check(lib().ft_range_query_csr(_sh(stream), C.byref(tree.c), _ptr(qx), _ptr(qy), _ptr(qz),
                                    _ptr(r_min), _ptr(r_max), int(nq), rp.ctypes.data, None, 0))
    total = int(rp[-1])
    cols = np.zeros(max(total, 1), np.int32)
    _check(lib().ft_range_query_csr(_sh(stream), C.byref(tree.c), _ptr(qx), _ptr(qy), _ptr(qz),
                                    _ptr(r_min), _ptr(r_max), int(nq), rp.ctypes.data,
                                    cols.ctypes.data, int(cols.size)))
    return rp, cols[:total]


def version() -> str:
    return lib().ft_version().decode()
