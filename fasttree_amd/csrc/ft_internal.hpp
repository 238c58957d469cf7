// ft_internal.hpp -- shared definitions of the gfx950 HLBVH library.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>

#include "../../include/fasttree_c.h"

namespace ft {

// ---------------------------------------------------------------------------
// Errors: every C entry point returns a status; the message is thread-local.
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define FT_HIP(expr)                                                                      \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return ::ft::fail(FT_EHIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

#define FT_CHECK_LAUNCH()                                                                 \
    do {                                                                                  \
        hipError_t _e = hipGetLastError();                                                \
        if (_e != hipSuccess)                                                             \
            return ::ft::fail(FT_EHIP, std::string("kernel launch: ") + hipGetErrorString(_e)); \
    } while (0)

#define FT_TRY(...)                    \
    do {                               \
        int _rc = (__VA_ARGS__);       \
        if (_rc != FT_OK) return _rc;  \
    } while (0)

// ---------------------------------------------------------------------------
// Coordinate / distance types.  dist_t is double for FP coordinates and
// uint64_t for integer coordinates (src/sfc.peano_hilbert.hpp:185-193).
// ---------------------------------------------------------------------------
template <typename T> struct CoordTraits;
template <> struct CoordTraits<float> { using D = double; static constexpr bool fp = true; static constexpr int ct = FT_COORD_F32; };
template <> struct CoordTraits<double> { using D = double; static constexpr bool fp = true; static constexpr int ct = FT_COORD_F64; };
template <> struct CoordTraits<uint32_t> { using D = uint64_t; static constexpr bool fp = false; static constexpr int ct = FT_COORD_U32; };
template <> struct CoordTraits<uint64_t> { using D = uint64_t; static constexpr bool fp = false; static constexpr int ct = FT_COORD_U64; };

inline size_t coord_size(int ct) { return (ct == FT_COORD_F32 || ct == FT_COORD_U32) ? 4 : 8; }

// Runtime dispatch over the coordinate type recorded in the tree.
#define FT_DISPATCH_COORD(ct, ...)                              \
    switch (ct) {                                               \
    case FT_COORD_F32: { using T = float; __VA_ARGS__; } break;    \
    case FT_COORD_F64: { using T = double; __VA_ARGS__; } break;   \
    case FT_COORD_U32: { using T = uint32_t; __VA_ARGS__; } break; \
    case FT_COORD_U64: { using T = uint64_t; __VA_ARGS__; } break; \
    default: return ::ft::fail(FT_EINVAL, "unknown coordinate type"); \
    }

// Metric parameters handed to every query kernel by value.
struct Metric {
    int bits;      // B
    int periodic;  // PERIODIC_BC
    int dshift;    // _DIST_SHIFT = max(B - 31, 0)
    int pad;
    double box;    // FP periodic period L
};

inline Metric make_metric(const ft_config &c) {
    Metric m;
    m.bits = c.bits_per_dim;
    m.periodic = c.periodic;
    m.dshift = c.bits_per_dim > 31 ? c.bits_per_dim - 31 : 0;
    m.pad = 0;
    m.box = c.box_length;
    return m;
}

__host__ __device__ inline float ffmax(float a, float b) { return fmaxf(a, b); }
__host__ __device__ inline double ffmax(double a, double b) { return fmax(a, b); }
__host__ __device__ inline float ffmin(float a, float b) { return fminf(a, b); }
__host__ __device__ inline double ffmin(double a, double b) { return fmin(a, b); }
__host__ __device__ inline float ffabs(float a) { return fabsf(a); }
__host__ __device__ inline double ffabs(double a) { return fabs(a); }

// Periodic gap of one axis for integer coordinates: the signed difference
// wrapped to B bits (shortest_periodic_distance, src/sfc.peano_hilbert.hpp:532-538,
// with SHIFT_BITS = bits(T) - B), absolute value.
template <typename T> __host__ __device__ inline uint64_t wrap_gap(T a, T b, int B) {
    using S = typename std::make_signed<T>::type;
    const int sh = (int)(sizeof(T) * 8) - B;
    T diff = (T)(a - b);
    S v = (S)(T)(diff << sh);
    v = v >> sh;
    return (uint64_t)(v < 0 ? -v : v);
}

// Squared distance from a point to a box (node_distance_sq, src/hlbvh.hpp:998-1055).
// FP: (dx*dx + dy*dy) + dz*dz in T, no contraction (built with -ffp-contract=off),
// widened to double.  Integer: per-axis gap >> _DIST_SHIFT, summed squares in u64.
template <typename T>
__host__ __device__ inline typename CoordTraits<T>::D box_d2(const Metric &m, T px, T py, T pz,
                                                             T x0, T x1, T y0, T y1, T z0, T z1) {
    if constexpr (CoordTraits<T>::fp) {
        T dx, dy, dz;
        if (!m.periodic) {
            dx = ffmax(x0 - px, ffmax(T(0), px - x1));
            dy = ffmax(y0 - py, ffmax(T(0), py - y1));
            dz = ffmax(z0 - pz, ffmax(T(0), pz - z1));
        } else {
            // No reference counterpart (PERIODIC_BC is a no-op for FP coords): 0 inside
            // [lo, hi], else the smaller minimum-image distance to either face.
            const T L = (T)m.box;
            auto ax = [L](T p, T lo, T hi) -> T {
                if (p >= lo && p <= hi) return T(0);
                T a = ffabs(p - lo);
                a = ffmin(a, L - a);
                T b = ffabs(p - hi);
                b = ffmin(b, L - b);
                return ffmin(a, b);
            };
            dx = ax(px, x0, x1);
            dy = ax(py, y0, y1);
            dz = ax(pz, z0, z1);
        }
        T s = dx * dx;
        s = s + dy * dy;
        s = s + dz * dz;
        return (double)s;
    } else {
        uint64_t g[3];
        const T p[3] = {px, py, pz}, lo[3] = {x0, y0, z0}, hi[3] = {x1, y1, z1};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T v = p[a];
            if (m.periodic) {
                if (sizeof(T) >= 8) v &= (T)((T(1) << m.bits) - T(1));
                if (v >= lo[a] && v <= hi[a]) {
                    g[a] = 0;
                } else {
                    uint64_t al = wrap_gap<T>(v, lo[a], m.bits);
                    uint64_t ah = wrap_gap<T>(v, hi[a], m.bits);
                    g[a] = (al < ah ? al : ah) >> m.dshift;
                }
            } else {
                if (v < lo[a]) g[a] = (uint64_t)(T)(lo[a] - v) >> m.dshift;
                else if (v > hi[a]) g[a] = (uint64_t)(T)(v - hi[a]) >> m.dshift;
                else g[a] = 0;
            }
        }
        return g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    }
}

// Point-to-point distance: the box degenerates to the point.
template <typename T>
__host__ __device__ inline typename CoordTraits<T>::D point_d2(const Metric &m, T px, T py, T pz,
                                                               T x, T y, T z) {
    return box_d2<T>(m, px, py, pz, x, x, y, y, z, z);
}

// The same value as point_d2, computed with fewer operations:
//  * FP: with lo == hi the face gap fmax(lo - p, fmax(0, p - hi)) is |p - x|
//    exactly (negation is exact), periodic: min(|p - x|, L - |p - x|).
//  * integer periodic: only the low B bits of (p - x) survive the wrap
//    (shortest_periodic_distance shifts everything above them out), so the
//    signed B-bit difference is formed in 32 bits; the coordinate mask of the
//    query does not change those bits either.  Valid for B <= 31.
//  * integer non-periodic: the general formula.
template <typename T>
__host__ __device__ inline typename CoordTraits<T>::D leaf_d2(const Metric &m, T px, T py, T pz,
                                                              T x, T y, T z) {
    if constexpr (CoordTraits<T>::fp) {
        T dx = ffabs(px - x), dy = ffabs(py - y), dz = ffabs(pz - z);
        if (m.periodic) {
            const T L = (T)m.box;
            dx = ffmin(dx, L - dx);
            dy = ffmin(dy, L - dy);
            dz = ffmin(dz, L - dz);
        }
        T s = dx * dx;
        s = s + dy * dy;
        s = s + dz * dz;
        return (double)s;
    } else {
        if (m.periodic && m.bits <= 31) {
            const int sh = 32 - m.bits;
            auto g = [sh](T a, T b) -> uint64_t {
                int32_t v = (int32_t)(((uint32_t)a - (uint32_t)b) << sh) >> sh;
                return (uint64_t)(uint32_t)(v < 0 ? -v : v);
            };
            const uint64_t gx = g(px, x) >> m.dshift, gy = g(py, y) >> m.dshift,
                           gz = g(pz, z) >> m.dshift;
            return gx * gx + gy * gy + gz * gz;
        }
        return box_d2<T>(m, px, py, pz, x, x, y, y, z, z);
    }
}

// Sentinel written for a missing neighbour: type_identity_max<dist_t>()
// (src/hlbvh.hpp:59-69): -DBL_MAX for FP, 0 for integers.
template <typename D> __host__ __device__ inline D missing_dist() {
    if constexpr (std::is_floating_point<D>::value) return -DBL_MAX;
    else return 0;
}

// ---------------------------------------------------------------------------
// XCD-aware block order.  MI355X hands workgroup b of a launch to XCD b % 8,
// and each XCD has its own L2.  With the grid padded to a multiple of 8,
// xcd_block() maps the workgroups of XCD x to the contiguous run of logical
// blocks [x * G/8, (x+1) * G/8), so neighbouring data (neighbouring packets
// of queries, neighbouring sorted particles) is read through one L2 instead
// of being fetched into all eight.
// ---------------------------------------------------------------------------
constexpr int FT_XCDS = 8;
__device__ inline int xcd_block() {
    const int per = (int)gridDim.x / FT_XCDS;
    return ((int)blockIdx.x % FT_XCDS) * per + (int)blockIdx.x / FT_XCDS;
}
inline unsigned xcd_grid(size_t blocks) {
    const size_t g = (blocks + FT_XCDS - 1) / FT_XCDS * FT_XCDS;
    return (unsigned)(g ? g : FT_XCDS);
}

// Device view of a built tree.
template <typename T> struct DTree {
    const T *mnx, *mxx, *mny, *mxy, *mnz, *mxz;
    const int32_t *left, *right, *parent, *orig;
    const int32_t *rfirst, *rlast;
    int n;
    int orig_indices;
    Metric m;
};

template <typename T> inline DTree<T> make_dtree(const ft_tree &t) {
    DTree<T> d;
    d.mnx = (const T *)t.min_x; d.mxx = (const T *)t.max_x;
    d.mny = (const T *)t.min_y; d.mxy = (const T *)t.max_y;
    d.mnz = (const T *)t.min_z; d.mxz = (const T *)t.max_z;
    d.left = t.left_child; d.right = t.right_child; d.parent = t.parent; d.orig = t.orig_idx;
    d.rfirst = t.range_first; d.rlast = t.range_last;
    d.n = (int)t.num_leaves;
    d.orig_indices = t.config.return_orig_indices;
    d.m = make_metric(t.config);
    return d;
}

// Output arrays of a tree build.
template <typename T> struct TreeOut {
    T *mnx, *mxx, *mny, *mxy, *mnz, *mxz;
    int32_t *left, *right, *parent, *rfirst, *rlast;
    uint32_t *id;
    int8_t *ghost;
    int32_t *orig;
};
template <typename T> inline TreeOut<T> make_tree_out(const ft_tree &t) {
    return TreeOut<T>{(T *)t.min_x,   (T *)t.max_x,   (T *)t.min_y,  (T *)t.max_y,
                      (T *)t.min_z,   (T *)t.max_z,   t.left_child,  t.right_child,
                      t.parent,       t.range_first,  t.range_last,  t.id,
                      t.is_ghost,     t.orig_idx};
}

// ---------------------------------------------------------------------------
// Segmented build: the trees of a batch built together, one launch of each
// build kernel for all of them.  Their points are numbered one after another
// (tree t owns indices [off, off + n), all below 2^31); sort tiles
// (SEG_TILE keys) and assembly chunks (SEG_CHUNK leaves) never cross a tree.
// ---------------------------------------------------------------------------
template <typename T> struct SegTree {
    const T *x, *y, *z;     // the tree's input points
    TreeOut<T> o;           // its TreeSoA arrays
    uint64_t *sorted_keys;  // its sorted_keys array
    int32_t off, n;         // first point index, points
    int32_t tile0, ntiles;  // its sort tiles (tile table entries)
};
constexpr int SEG_TILE = 6144;   // keys per sort tile
constexpr int SEG_CHUNK = 512;   // leaves per assembly chunk
constexpr int SEG_HIST_WORDS = 8 * 512;  // digit histogram words per tree (8 passes of 9 bits)
// Device tables and buffers of one segmented build (owned by the caller).
struct SegBuild {
    const void *trees;   // SegTree<T>[ntrees]
    const int4 *tiles;   // per sort tile: first key, end, tree, first tile of its tree
    const int2 *chunks;  // per assembly chunk: tree, first leaf within the tree
    int ntrees, ntiles, nchunks;
    size_t total;        // points of all trees
    uint32_t *hist;      // 2 * ntrees * SEG_HIST_WORDS: digit histograms, then digit bases
    void *bbox;          // FP coordinates: one box record (lo, hi, inverse extents) per tree
    void *partials;      // FP coordinates: 6 values per tile
};

// ---------------------------------------------------------------------------
// Host staging (ensure_device_readable / copy_back_and_free, src/hlbvh.hpp:125-200).
// ---------------------------------------------------------------------------
bool is_device_accessible(const void *p);

// A staged buffer frees its device copy (stream-ordered on the stream it was
// staged on) when it goes out of scope without unstage, so an entry point
// that returns early on an error leaks nothing.
struct Staged {
    void *dev = nullptr;
    const void *host_src = nullptr;
    void *host_dst = nullptr;
    size_t bytes = 0;
    bool owned = false;
    hipStream_t stream = nullptr;
    Staged() = default;
    Staged(const Staged &) = delete;
    Staged &operator=(const Staged &) = delete;
    ~Staged() {
        if (owned) (void)hipFreeAsync(dev, stream);
    }
};

// Returns a device-readable pointer for p (copying host data in), or p itself.
int stage_in(const void *p, size_t bytes, hipStream_t s, Staged &st);
// Returns a device-writable pointer (device scratch when p is host memory).
int stage_out(void *p, size_t bytes, hipStream_t s, Staged &st);
// Copies staged output back (synchronising the stream) and frees staging memory.
int unstage(Staged &st, hipStream_t s, bool copy_back);

inline hipStream_t as_stream(ft_stream s) { return (hipStream_t)s; }

// Device scratch owned by ft_scratch: grows as n + n/2 + 1024 (BvhScratch::ensure).
struct Scratch {
    void *block = nullptr;
    size_t block_bytes = 0;
    size_t cap_n = 0;
    // carved views
    uint64_t *keys_in = nullptr;  // encoded keys (n)
    uint64_t *keys_a = nullptr, *keys_b = nullptr;
    uint32_t *vals_a = nullptr, *vals_b = nullptr;
    uint32_t *perm = nullptr;     // final permutation (n)
    void *points = nullptr;       // encode's copy of the input points, (x, y, z, 0) per
                                  // point (4 coordinates, 32 bytes at most): the leaf
                                  // gather reads one record instead of three arrays
    uint64_t *split_words = nullptr; // tree assembly: arrival word per split (n), epoch-tagged
    int32_t *tops = nullptr;         // tree assembly: labels of the chunk tops, by chunk (n)
    uint32_t *chunk_tops = nullptr;  // tree assembly: chunk tops per chunk
    uint32_t *topo_epoch = nullptr;  // tree assembly counter: tags the arrival words
    uint32_t *hist = nullptr;     // 8 passes x 512 digit counts, then 8 x 512 digit bases
    uint64_t *lookback = nullptr; // tiles x 512 status words
    size_t lookback_words = 0;
    uint32_t *tile_ctr = nullptr; // 8 passes
    uint32_t *epoch = nullptr;    // sort counter on the device: look-back tags of a sort
                                  // are 8 * epoch + pass, so no tag is reused and a
                                  // captured graph that replays the sort stays correct
    int radix_bits = 0;           // sort digit width: 8 or 9 bits per pass, 0 = the one
                                  // with fewer passes (8 on a tie)
    void *bbox_dev = nullptr;     // 6 coords + 3 inverse extents (16 x 8 bytes)
    void *partials = nullptr;     // bbox partials (1024 x 6 x 8 bytes)
    // staging of inputs (host pointers)
    void *stage_block = nullptr;
    size_t stage_bytes = 0;
    // profiling
    int profiling = 0;
    hipEvent_t ev[8] = {};
    double stage_ms[8] = {};
    int n_stages = 0;
};

// Builds every tree of g (all of one configuration) in one launch of each
// build kernel; each tree comes out as build_bvh builds it alone.
template <typename T>
int build_segmented_t(hipStream_t s, Scratch &S, const SegBuild &g, const ft_config &cfg);

}  // namespace ft

struct ft_scratch {
    ft::Scratch s;
};
