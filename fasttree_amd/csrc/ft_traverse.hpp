// ft_traverse.hpp -- device-side pieces shared by the query kernels
// (ft_query.hip) and the batched multi-tree kernels (ft_batch.hip): lane
// shuffles, query sources, node distances, the canonical top-k containers
// (register, packed 64-bit, LDS heap) and the packet traversal walk.
#pragma once

#include "ft_internal.hpp"

namespace ft {

template <typename D> __device__ inline D dist_max() {
    if constexpr (std::is_floating_point<D>::value) return DBL_MAX;
    else return ~0ull;
}
template <typename D> __device__ inline D dist_lowest() {
    if constexpr (std::is_floating_point<D>::value) return -DBL_MAX;
    else return 0ull;
}

__device__ inline uint32_t rl32(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ inline uint64_t rl(uint64_t v, int l) {
    return ((uint64_t)rl32((uint32_t)(v >> 32), l) << 32) | rl32((uint32_t)v, l);
}
__device__ inline uint32_t rl(uint32_t v, int l) { return rl32(v, l); }
__device__ inline int32_t rl(int32_t v, int l) { return (int32_t)rl32((uint32_t)v, l); }
__device__ inline double rl(double v, int l) { return __longlong_as_double((long long)rl((uint64_t)__double_as_longlong(v), l)); }
__device__ inline float rl(float v, int l) { return __uint_as_float(rl32(__float_as_uint(v), l)); }

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Per-lane source lane (ds_bpermute): lane i receives v from lane src_i.
__device__ inline uint32_t shfl32(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}
__device__ inline uint64_t shfl_any(uint64_t v, int src) {
    return ((uint64_t)shfl32((uint32_t)(v >> 32), src) << 32) | shfl32((uint32_t)v, src);
}
__device__ inline uint32_t shfl_any(uint32_t v, int src) { return shfl32(v, src); }
__device__ inline int32_t shfl_any(int32_t v, int src) { return (int32_t)shfl32((uint32_t)v, src); }
__device__ inline double shfl_any(double v, int src) {
    return __longlong_as_double((long long)shfl_any((uint64_t)__double_as_longlong(v), src));
}
__device__ inline float shfl_any(float v, int src) { return __uint_as_float(shfl32(__float_as_uint(v), src)); }

__device__ inline void wave_add_stats(uint64_t *stats, int slot, uint64_t v) {
    if (!stats) return;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd((unsigned long long *)&stats[slot], (unsigned long long)v);
}

// Query source: external points, or leaves of the tree (self queries).
template <typename T> struct QSrc {
    const T *qx, *qy, *qz;   // external points (null for self)
    const int32_t *leaf_ids; // subset leaf ranks (null = all leaves)
};

template <typename T>
__device__ inline void load_query(const DTree<T> &t, const QSrc<T> &qs, int q, T &px, T &py, T &pz,
                                  int &self_rank) {
    if (qs.qx) {
        px = qs.qx[q]; py = qs.qy[q]; pz = qs.qz[q];
        self_rank = -1;
    } else {
        const int r = qs.leaf_ids ? qs.leaf_ids[q] : q;
        const int leaf = r + t.n - 1;
        px = t.mnx[leaf]; py = t.mny[leaf]; pz = t.mnz[leaf];
        self_rank = r;
    }
}

template <typename T>
__device__ inline typename CoordTraits<T>::D node_d2(const DTree<T> &t, T px, T py, T pz, int node) {
    return box_d2<T>(t.m, px, py, pz, t.mnx[node], t.mxx[node], t.mny[node], t.mxy[node],
                     t.mnz[node], t.mxz[node]);
}

// ---------------------------------------------------------------------------
// Compile-time metric of a walk.  MK_ANY reads the Metric at run time; MK_WRAP
// (periodic) and MK_OPEN (non-periodic) fix the flag so inner loops carry no
// metric branches.  Integer MK_WRAP needs B <= 31 (every valid config): the
// wrapped difference is then the sign-extended low B bits of a 32-bit
// difference, and _DIST_SHIFT is 0.
// ---------------------------------------------------------------------------
enum : int { MK_ANY = 0, MK_WRAP = 1, MK_OPEN = 2 };

template <typename T> inline int metric_kind(const Metric &m) {
    if (!m.periodic) return MK_OPEN;
    if (!CoordTraits<T>::fp && m.bits > 31) return MK_ANY;
    return MK_WRAP;
}

// Signed B-bit wrapped difference a - b (sh = 32 - B).
__host__ __device__ __forceinline__ int32_t wrap_diff32(uint32_t a, uint32_t b, int sh) {
    return (int32_t)((a - b) << sh) >> sh;
}
__host__ __device__ __forceinline__ uint64_t wrap_abs32(uint32_t a, uint32_t b, int sh) {
    const int32_t v = wrap_diff32(a, b, sh);
    return (uint64_t)(uint32_t)(v < 0 ? -v : v);
}

// Float32 pre-filter of the integer periodic metric (B <= 31).  The float sum
// of the three squared wrapped gaps (conversions, one product, two fmas) stays
// within 5 * 2^-24 (relative) of the exact sum, and the bound is raised by
// 2^-20 after its own conversion, so wrap_filter_near() is true for every leaf
// whose exact squared distance is at most w; leaves it passes are then tested
// exactly.
__host__ __device__ __forceinline__ float wrap_filter_bound(uint64_t w) {
    return (float)w * (1.0f + 0x1p-20f);
}
__host__ __device__ __forceinline__ bool wrap_filter_near(uint32_t qx, uint32_t qy, uint32_t qz,
                                                          uint32_t x, uint32_t y, uint32_t z,
                                                          int sh, float wf) {
    const float fx = (float)wrap_diff32(qx, x, sh);
    const float fy = (float)wrap_diff32(qy, y, sh);
    const float fz = (float)wrap_diff32(qz, z, sh);
    float d = fx * fx;
    d = __builtin_fmaf(fy, fy, d);
    d = __builtin_fmaf(fz, fz, d);
    return d <= wf;
}

// Pre-shifted form of the integer periodic metric (B <= 31): with a' = a << sh
// and b' = b << sh (32-bit, sh = 32 - B), (int32)(a' - b') is the wrapped
// difference of a and b times 2^sh, so no per-pair shifts are needed.  Its
// magnitude, shifted back, is wrap_abs32(a, b, sh) exactly (2^(B-1) included,
// read as unsigned).
__host__ __device__ __forceinline__ uint64_t wrap_gap_pre(uint32_t a, uint32_t b, int sh) {
    const int32_t x = (int32_t)(a - b);
    const uint32_t u = (uint32_t)x;
    return (uint64_t)((x < 0 ? 0u - u : u) >> sh);
}
// The float32 filter on pre-shifted coordinates: every value is the unshifted
// one times 2^sh (a power of two, so the conversions and the error bound of
// wrap_filter_near are unchanged) and the squared bound is scaled by 2^(2 sh).
__host__ __device__ __forceinline__ float wrap_filter_bound_pre(uint64_t w, int sh) {
    return wrap_filter_bound(w) * (float)(1ull << (2 * sh));
}
// Lower-bound counterpart (multi-pass queries): a leaf whose exact squared
// distance is at least lo has a pre-shifted float32 distance of at least this
// (the float sum is within 5 * 2^-24 of the exact one; the bound is lowered by
// 2^-20 after its own conversion).
__host__ __device__ __forceinline__ float wrap_filter_floor_pre(uint64_t lo, int sh) {
    return (float)lo * (1.0f - 0x1p-20f) * (float)(1ull << (2 * sh));
}
// Float32 pre-filter for double coordinates.  With u = 2^-24, coordinates of
// magnitude at most M and period L, the float gap of one axis (both points
// rounded to float, one subtraction, for a periodic metric min(g, fl(L) - g))
// is within (1 + u) g + E of the double gap g, E = 1.01 u (L + 4M) (periodic)
// or 2.02 u M (open); the float sum of the three squares (a product, two
// fmas) is within a factor 1 + 3.01u.  So a leaf with d2 <= w has a float
// squared distance at most fpf_upper(w, E), and one with d2 >= lo at least
// fpf_lower(lo, E): the filter never rejects a leaf the exact test accepts.
__host__ __device__ __forceinline__ float fpf_gap(float q, float x, float L, bool wrap) {
    const float g = fabsf(q - x);
    return wrap ? fminf(g, L - g) : g;
}
__host__ __device__ __forceinline__ float fpf_sq(float gx, float gy, float gz) {
    float d = gx * gx;
    d = __builtin_fmaf(gy, gy, d);
    d = __builtin_fmaf(gz, gz, d);
    return d;
}
__host__ __device__ __forceinline__ double fpf_err(double M, double L, bool wrap) {
    return wrap ? 1.01 * 0x1p-24 * (L + 4.0 * M) : 2.02 * 0x1p-24 * M;
}
__host__ __device__ __forceinline__ float fpf_upper(double w, double E) {
    const double b = (1.0 + 0x1p-23) * sqrt(w) + 1.75 * E;
    const double v = b * b * (1.0 + 0x1p-22);
    return v >= 3.0e38 ? __builtin_huge_valf() : (float)v * (1.0f + 0x1p-22f);
}
__host__ __device__ __forceinline__ float fpf_lower(double lo, double E) {
    const double b = (1.0 - 0x1p-23) * sqrt(lo) - 1.75 * E;
    if (!(b > 0.0)) return 0.0f;
    return (float)(b * b * (1.0 - 0x1p-22)) * (1.0f - 0x1p-22f);
}

// Float32 squared distance of a leaf, pre-shifted coordinates (scaled by
// 2^(2 sh)); compared with wrap_filter_bound_pre it never rejects a leaf whose
// exact squared distance is within the bound.
__host__ __device__ __forceinline__ float wrap_sq_pre(uint32_t qx, uint32_t qy, uint32_t qz,
                                                      uint32_t x, uint32_t y, uint32_t z) {
    const float fx = (float)(int32_t)(qx - x);
    const float fy = (float)(int32_t)(qy - y);
    const float fz = (float)(int32_t)(qz - z);
    float d = fx * fx;
    d = __builtin_fmaf(fy, fy, d);
    d = __builtin_fmaf(fz, fz, d);
    return d;
}
// One axis of a box's float32 gap (scaled by 2^sh): zero inside [lo, hi],
// tested on the full values with the query masked to B bits (qm), as in
// box_d2_k; else the smaller wrapped gap to either face from pre-shifted
// coordinates (qs = query << sh).
template <typename T>
__host__ __device__ __forceinline__ float wrap_box_gap_pre(T qm, uint32_t qs, T lo, T hi, int sh) {
    const bool in = (qm >= lo) && (qm <= hi);
    const float a = __builtin_fabsf((float)(int32_t)(qs - ((uint32_t)lo << sh)));
    const float b = __builtin_fabsf((float)(int32_t)(qs - ((uint32_t)hi << sh)));
    return in ? 0.0f : __builtin_fminf(a, b);
}

// leaf_d2 / box_d2 with the metric kind fixed (same values).
template <typename T, int MK>
__device__ __forceinline__ typename CoordTraits<T>::D leaf_d2_k(const Metric &m, T px, T py, T pz,
                                                               T x, T y, T z) {
    if constexpr (MK == MK_ANY) {
        return leaf_d2<T>(m, px, py, pz, x, y, z);
    } else if constexpr (CoordTraits<T>::fp) {
        T dx = ffabs(px - x), dy = ffabs(py - y), dz = ffabs(pz - z);
        if constexpr (MK == MK_WRAP) {
            const T L = (T)m.box;
            dx = ffmin(dx, L - dx);
            dy = ffmin(dy, L - dy);
            dz = ffmin(dz, L - dz);
        }
        T s = dx * dx;
        s = s + dy * dy;
        s = s + dz * dz;
        return (double)s;
    } else if constexpr (MK == MK_WRAP) {
        const int sh = 32 - m.bits;
        const uint64_t gx = wrap_abs32((uint32_t)px, (uint32_t)x, sh);
        const uint64_t gy = wrap_abs32((uint32_t)py, (uint32_t)y, sh);
        const uint64_t gz = wrap_abs32((uint32_t)pz, (uint32_t)z, sh);
        return gx * gx + gy * gy + gz * gz;
    } else {
        auto g = [&m](T p, T c) -> uint64_t {
            return (uint64_t)(p < c ? (T)(c - p) : (T)(p - c)) >> m.dshift;
        };
        const uint64_t gx = g(px, x), gy = g(py, y), gz = g(pz, z);
        return gx * gx + gy * gy + gz * gz;
    }
}

template <typename T, int MK>
__device__ __forceinline__ typename CoordTraits<T>::D box_d2_k(const Metric &m, T px, T py, T pz,
                                                              T x0, T x1, T y0, T y1, T z0,
                                                              T z1) {
    if constexpr (MK == MK_ANY) {
        return box_d2<T>(m, px, py, pz, x0, x1, y0, y1, z0, z1);
    } else if constexpr (CoordTraits<T>::fp) {
        Metric mm = m;
        mm.periodic = (MK == MK_WRAP);
        return box_d2<T>(mm, px, py, pz, x0, x1, y0, y1, z0, z1);
    } else if constexpr (MK == MK_WRAP) {
        const int sh = 32 - m.bits;
        auto g = [&](T p, T lo, T hi) -> uint64_t {
            T v = p;
            if constexpr (sizeof(T) >= 8) v &= (T)((T(1) << m.bits) - T(1));
            const uint64_t al = wrap_abs32((uint32_t)v, (uint32_t)lo, sh);
            const uint64_t ah = wrap_abs32((uint32_t)v, (uint32_t)hi, sh);
            const uint64_t w = al < ah ? al : ah;
            return (v >= lo && v <= hi) ? 0ull : w;
        };
        const uint64_t gx = g(px, x0, x1), gy = g(py, y0, y1), gz = g(pz, z0, z1);
        return gx * gx + gy * gy + gz * gz;
    } else {
        auto g = [&m](T p, T lo, T hi) -> uint64_t {
            const uint64_t a = (uint64_t)(T)(lo - p) >> m.dshift;
            const uint64_t b = (uint64_t)(T)(p - hi) >> m.dshift;
            return p < lo ? a : (p > hi ? b : 0ull);
        };
        const uint64_t gx = g(px, x0, x1), gy = g(py, y0, y1), gz = g(pz, z0, z1);
        return gx * gx + gy * gy + gz * gz;
    }
}

template <typename T, int MK>
__device__ __forceinline__ typename CoordTraits<T>::D node_d2_k(const DTree<T> &t, T px, T py,
                                                               T pz, int node) {
    return box_d2_k<T, MK>(t.m, px, py, pz, t.mnx[node], t.mxx[node], t.mny[node], t.mxy[node],
                           t.mnz[node], t.mxz[node]);
}

// Returned index of a leaf: its rank, or orig_idx with RETURN_ORIG_INDICES.
template <typename T> __device__ inline int32_t ret_index(const DTree<T> &t, int leaf_node) {
    return t.orig_indices ? t.orig[leaf_node] : (leaf_node - (t.n - 1));
}

// ---------------------------------------------------------------------------
// Canonical register top-k: ascending (d, idx) in M fixed registers, the live
// k entries right-aligned so the worst is always slot M-1.  Slots below M-k
// hold (-inf, -1) and never move.  Insertion is a fixed compare/shift network.
// ---------------------------------------------------------------------------
// IEEE max / min of doubles, one VALU instruction each (no NaN reaches them).
__device__ __forceinline__ double f64_max(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double f64_min(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

template <typename D, int M> struct TopK {
    static constexpr bool kPacked = false;
    static constexpr bool kSortFill = false;
    static constexpr int kRoundsPerTrip = 2;
    D d[M];
    int32_t i[M];
    __device__ inline void init(int k, D bound) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const bool live = s >= M - k;
            d[s] = live ? bound : dist_lowest<D>();
            i[s] = live ? INT32_MAX : -1;
        }
    }
    __device__ inline D worst() const { return d[M - 1]; }
    __device__ inline bool beats_worst(D cd, int32_t ci) const {
        return (cd < d[M - 1]) | ((cd == d[M - 1]) & (ci < i[M - 1]));
    }
    // (cd, ci) below the entry in slot s, in (d, index) order.
    __device__ inline bool below(int s, D cd, int32_t ci) const {
        return (cd < d[s]) | ((cd == d[s]) & (ci < i[s]));
    }
    // Predicated insertion as a compare-and-swap chain from the smallest slot:
    // the carried entry swaps into every slot it is below, so it lands in order
    // and the former worst drops off the top.  Every slot stays in its own
    // registers; the comparisons are branch-free.  A candidate not below the
    // entry in slot L-1 swaps with no slot under L (the slots are sorted), so
    // the chain can start at L.
    //
    // Double distances run the chain top-down instead: slot s takes the entry
    // of slot s-1 when the candidate is below it, the candidate when it is
    // below slot s only, else keeps its entry.  The distance of that choice is
    // the median max(d[s-1], min(c, d[s])) (ties included: the compares that
    // pick the index are lexicographic, and the three candidates' distances
    // are ordered), one v_min_f64 and one v_max_f64; a lane that takes nothing
    // offers the largest distance, which changes no slot.
    template <int L> __device__ inline void insert_from(bool take, D cd, int32_t ci) {
        if constexpr (std::is_same<D, double>::value) {
            const double c = take ? cd : dist_max<D>();
#pragma unroll
            for (int s = M - 1; s >= L; --s) {
                const bool bs = take & below(s, cd, ci);
                const bool bp = (s > L) && (take & below(s - 1, cd, ci));
                i[s] = bp ? i[s - 1] : (bs ? ci : i[s]);
                d[s] = s > 0 ? f64_max(d[s - 1], f64_min(c, d[s])) : f64_min(c, d[s]);
            }
        } else {
#pragma unroll
            for (int s = L; s < M; ++s) {
                const bool lt = take & below(s, cd, ci);
                const D td = d[s];
                const int32_t ti = i[s];
                d[s] = lt ? cd : td;
                i[s] = lt ? ci : ti;
                cd = lt ? td : cd;
                ci = lt ? ti : ci;
            }
        }
    }
    // New entries mostly land near the top: the chain starts at the lowest of
    // slots 7M/8, 3M/4, M/2 and 0 that a taking lane of the wave can reach
    // (a compare and a wave vote per level).
    __device__ inline void insert(bool take, D cd, int32_t ci) {
        if constexpr (M >= 16) {
            if (!__any(take && below(7 * M / 8 - 1, cd, ci))) { insert_from<7 * M / 8>(take, cd, ci); return; }
            if (!__any(take && below(3 * M / 4 - 1, cd, ci))) { insert_from<3 * M / 4>(take, cd, ci); return; }
            if (!__any(take && below(M / 2 - 1, cd, ci))) { insert_from<M / 2>(take, cd, ci); return; }
        }
        insert_from<0>(take, cd, ci);
    }
    __device__ inline D worst_d() const { return d[M - 1]; }
    // Multi-pass queries: a pass takes only (d, index) at or above the floor.
    struct Floor {
        D d;
        int32_t i;
    };
    __device__ static inline Floor no_floor() { return Floor{D(0), INT32_MIN}; }
    __device__ inline D floor_d(const Floor &lo) const { return lo.d; }
    __device__ inline bool above(const Floor &lo, D cd, int32_t ci) const {
        return (cd > lo.d) | ((cd == lo.d) & (ci >= lo.i));
    }
    // Above the k-th entry, or nothing when a slot is still empty.
    __device__ inline Floor next_floor() const {
        if (i[M - 1] == INT32_MAX) return Floor{dist_max<D>(), INT32_MAX};
        return Floor{d[M - 1], i[M - 1] + 1};
    }
    template <typename DOut>
    __device__ inline void store_part(int k, uint64_t *res, DOut *dists, size_t q, int k_row,
                                      int base) const {
        const size_t o = q * (size_t)k_row + (size_t)base;
#pragma unroll
        for (int s = 0; s < M; ++s) {
            if (s >= M - k) {
                const size_t j = o + (size_t)(s - (M - k));
                const bool ok = i[s] != INT32_MAX;
                res[j] = ok ? (uint64_t)(uint32_t)i[s] : ~0ull;
                dists[j] = ok ? (DOut)d[s] : missing_dist<DOut>();
            }
        }
    }
    // Number of live slots still holding the initial sentinel.
    __device__ inline int missing(int k) const {
        int c = 0;
#pragma unroll
        for (int s = 0; s < M; ++s) c += (s >= M - k && i[s] == INT32_MAX) ? 1 : 0;
        return c;
    }
    template <typename DOut>
    __device__ inline void store(int k, uint64_t *res, DOut *dists, size_t q) const {
        const size_t o = q * (size_t)k;
#pragma unroll
        for (int s = 0; s < M; ++s) {
            if (s >= M - k) {
                const size_t j = o + (size_t)(s - (M - k));
                const bool ok = i[s] != INT32_MAX;
                res[j] = ok ? (uint64_t)(uint32_t)i[s] : ~0ull;
                dists[j] = ok ? (DOut)d[s] : missing_dist<DOut>();
            }
        }
    }
};

constexpr int LANE_BLOCK = 64;
constexpr int STACK_DEPTH = 64;  // MAX_STACK_DEPTH (src/hlbvh.hpp:984)
// Canonical per-lane walks: a DFS that pushes both children holds at most one
// pending sibling per level plus the current pair, and a Karras tree over
// (64-bit key, 32-bit index) is less than 97 levels deep, so 128 entries never
// overflow (the reference's 64-entry stack silently drops deeper nodes).
constexpr int LANE_STACK_CANON = 128;

// ===========================================================================
// Per-lane DFS with a max-heap row per lane in LDS (LocalMaxHeap,
// src/maxheap.hpp:340-478).  REF = the reference's mechanics exactly (compare
// d only, `>=` prune, sift rules, heapsort), so tie picks match; otherwise the
// heap is keyed by (d, idx) and pruning is strict (canonical).  Layout
// [slot][lane] so lanes at the same slot hit distinct banks.
// ===========================================================================
template <typename D, bool REF> struct LdsHeap {
    D *dist;       // [cap][lanes]
    int32_t *idx;  // [cap][lanes]
    int lanes, lane, count;
    __device__ inline D &dd(int s) { return dist[s * lanes + lane]; }
    __device__ inline int32_t &ii(int s) { return idx[s * lanes + lane]; }
    __device__ inline bool gt(int a, int b) {  // entry a > entry b
        if (REF) return dd(a) > dd(b);
        return dd(a) > dd(b) || (dd(a) == dd(b) && ii(a) > ii(b));
    }
    __device__ inline void swp(int a, int b) {
        D td = dd(a); dd(a) = dd(b); dd(b) = td;
        int32_t ti = ii(a); ii(a) = ii(b); ii(b) = ti;
    }
    __device__ inline bool prune(D d, int k) {
        if (count < k) return false;
        return REF ? (d >= dd(0)) : (d > dd(0));
    }
    __device__ inline void sift_up(int i) {
        while (i > 0) {
            const int p = (i - 1) >> 1;
            if (!gt(i, p)) break;  // reference: dist[i] <= dist[p] -> stop
            swp(i, p);
            i = p;
        }
    }
    __device__ inline void sift_down(int i, int size) {
        while (true) {
            int lg = i;
            const int l = 2 * i + 1, r = 2 * i + 2;
            if (l < size && gt(l, lg)) lg = l;
            if (r < size && gt(r, lg)) lg = r;
            if (lg == i) break;
            swp(i, lg);
            i = lg;
        }
    }
    __device__ inline bool push(D d, int32_t i, int k) {
        if (count < k) {
            dd(count) = d;
            ii(count) = i;
            sift_up(count);
            count++;
            return true;
        }
        const bool better = REF ? (d < dd(0)) : (d < dd(0) || (d == dd(0) && i < ii(0)));
        if (better) {
            dd(0) = d;
            ii(0) = i;
            sift_down(0, k);
            return true;
        }
        return false;
    }
    __device__ inline void sort() {
        int hs = count;
        while (hs > 1) {
            swp(0, hs - 1);
            hs--;
            sift_down(0, hs);
        }
    }
};


// ---------------------------------------------------------------------------
// Packed top-k for integer distances: key = BIAS | min(d2, SAT) << ib | index,
// one 64-bit compare orders by (d2, index).  ib = bit length of n, so every
// index is below 2^ib - 1 and an all-ones index marks a sentinel.  A key whose
// d2 part saturated cannot be ordered exactly; the traversal reports such
// queries (and queries left with sentinels) for an exact re-run.
//
// SAT = 2^(61 - ib) - 1 keeps every unbiased key below 2^61, and BIAS = 2^62
// lifts it into [2^62, 2^62 + 2^61): read as an IEEE double, such a key is a
// positive, normal, finite number (exponent field 0x400 to 0x5FF), and positive
// finite doubles order exactly as their bit patterns do.  So the insertion
// chain can use the native v_max_f64 / v_min_f64 (one instruction each for a
// whole 64-bit key) instead of a 64-bit compare and four 32-bit selects.
// Dead slots (below M - k) hold 0 (+0.0), below every key.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double key_f64(uint64_t k) { return __longlong_as_double((long long)k); }
__device__ __forceinline__ uint64_t key_u64(double k) { return (uint64_t)__double_as_longlong(k); }

template <int M> struct TopKPacked {
    static constexpr bool kPacked = true;
    static constexpr bool kSortFill = true;
    // insertion rounds per trip of the leaf loop (measured: two from 32 slots)
    static constexpr int kRoundsPerTrip = M >= 32 ? 2 : 1;
    static constexpr uint64_t BIAS = 1ull << 62;
    static constexpr uint64_t NOKEY = BIAS | (1ull << 61);  // above every key, still finite
    uint64_t key[M];
    int ib;
    uint64_t sat, imask;
    __device__ inline void init(int k, uint64_t bound, int ibits) {
        ib = ibits;
        sat = (1ull << (61 - ib)) - 1;
        imask = (1ull << ib) - 1;
        const uint64_t sent = BIAS | ((bound < sat ? bound : sat) << ib) | imask;
#pragma unroll
        for (int s = 0; s < M; ++s) key[s] = (s >= M - k) ? sent : 0ull;
    }
    __device__ inline uint64_t make(uint64_t d, int32_t i) const {
        return BIAS | ((d < sat ? d : sat) << ib) | (uint64_t)(uint32_t)i;
    }
    __device__ inline uint64_t dist_of(uint64_t kb) const { return (kb & ~BIAS) >> ib; }
    __device__ inline uint64_t worst_d() const { return dist_of(key[M - 1]); }
    __device__ inline bool beats_worst(uint64_t d, int32_t i) const { return make(d, i) < key[M - 1]; }
    // Sorted insertion, in place from the top slot down: with the slots
    // sorted, the new content of slot s is the median of key[s-1], c and
    // key[s], i.e. max(key[s-1], min(c, key[s])) -- key[s-1] when c is below
    // it, c when c falls between, key[s] when c is not below it.  Slot 0 has
    // no key below it and becomes min(c, key[0]).  A lane that takes nothing
    // inserts NOKEY, which changes no slot.
    //
    // The chain runs in segments from the top: slots [A, B) with the median
    // rule.  When no taking lane's key is below key[A-1], the slots below A
    // keep their keys, and the median rule at slot A already gave
    // min(c, key[A]) (c >= key[A-1] and key[A] >= key[A-1]), so the chain can
    // stop there.
    template <int A, int B> __device__ inline void chain_seg(uint64_t c) {
#pragma unroll
        for (int s = B - 1; s >= A; --s) {
            if (s == 0) key[0] = key_u64(f64_min(key_f64(c), key_f64(key[0])));
            else key[s] = key_u64(f64_max(key_f64(key[s - 1]), f64_min(key_f64(c), key_f64(key[s]))));
        }
    }
    __device__ inline bool beats_worst_key(uint64_t kb) const { return kb < key[M - 1]; }
    __device__ inline void insert(bool take, uint64_t d, int32_t i) { insert_key(take, make(d, i)); }
    __device__ inline void insert_key(bool take, uint64_t kb) {
        const uint64_t c = take ? kb : NOKEY;
        // New keys mostly land near the top: run only the part of the chain a
        // taking lane can reach (one compare and a wave vote per level).
        if constexpr (M >= 16) {
            chain_seg<7 * M / 8, M>(c);
            if (__any(c < key[7 * M / 8 - 1])) {
                chain_seg<3 * M / 4, 7 * M / 8>(c);
                if (__any(c < key[3 * M / 4 - 1])) {
                    chain_seg<M / 2, 3 * M / 4>(c);
                    if (__any(c < key[M / 2 - 1])) chain_seg<0, M / 2>(c);
                }
            }
        } else {
            chain_seg<0, M>(c);
        }
    }
    // Bitonic sort of the M slots, ascending, in registers (each
    // compare-exchange is one v_min_f64 and one v_max_f64).
    __device__ inline void sort_slots() {
#pragma unroll
        for (int kk = 2; kk <= M; kk <<= 1) {
#pragma unroll
            for (int j = kk >> 1; j > 0; j >>= 1) {
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    const int l = i ^ j;
                    if (l > i) {
                        const double a = key_f64(key[i]), b = key_f64(key[l]);
                        const double lo = f64_min(a, b), hi = f64_max(a, b);
                        const bool asc = (i & kk) == 0;
                        key[i] = key_u64(asc ? lo : hi);
                        key[l] = key_u64(asc ? hi : lo);
                    }
                }
            }
        }
    }
    // Right after init: the k live slots take cand(0..k-1) (keys from make();
    // a key above the init sentinel, i.e. beyond the bound, becomes the
    // sentinel), then the slots are sorted.  Returns the number of keys kept.
    template <typename F> __device__ inline int fill_sorted(int k, F &&cand) {
        const uint64_t sent = key[M - 1];
        int kept = 0;
#pragma unroll
        for (int s = 0; s < M; ++s) {
            if (s >= M - k) {
                const uint64_t kb = cand(s - (M - k));
                const bool in = kb < sent;
                key[s] = in ? kb : sent;
                kept += in ? 1 : 0;
            }
        }
        sort_slots();
        return kept;
    }
    // Sentinels left in live slots, or a saturated k-th distance.
    __device__ inline bool needs_exact(int k) const {
        bool bad = dist_of(key[M - 1]) >= sat;
#pragma unroll
        for (int s = 0; s < M; ++s) bad |= (s >= M - k) & ((key[s] & imask) == imask);
        return bad;
    }
    __device__ inline bool missing(int k) const {
        bool bad = false;
#pragma unroll
        for (int s = 0; s < M; ++s) bad |= (s >= M - k) & ((key[s] & imask) == imask);
        return bad;
    }
    // Multi-pass queries: a pass takes only keys at or above the floor.
    using Floor = uint64_t;
    __device__ static inline Floor no_floor() { return 0ull; }
    __device__ inline uint64_t floor_d(Floor lo) const { return dist_of(lo); }
    __device__ inline bool above(Floor lo, uint64_t d, int32_t i) const { return make(d, i) >= lo; }
    // Lower key bound of the next pass of a multi-pass query: above the k-th
    // key, or nothing when a slot is still empty (no further neighbours).
    __device__ inline uint64_t next_floor() const {
        const uint64_t last = key[M - 1];
        return ((last & imask) == imask) ? ~0ull : last + 1;
    }
    // Entries [base, base + k) of row q (row length k_row).
    template <typename DOut>
    __device__ inline void store_part(int k, uint64_t *res, DOut *dists, size_t q, int k_row,
                                      int base) const {
        const size_t o = q * (size_t)k_row + (size_t)base;
#pragma unroll
        for (int s = 0; s < M; ++s) {
            if (s >= M - k) {
                const size_t j = o + (size_t)(s - (M - k));
                const bool ok = (key[s] & imask) != imask;
                res[j] = ok ? (key[s] & imask) : ~0ull;
                dists[j] = ok ? (DOut)dist_of(key[s]) : missing_dist<DOut>();
            }
        }
    }
    template <typename DOut>
    __device__ inline void store(int k, uint64_t *res, DOut *dists, size_t q) const {
        const size_t o = q * (size_t)k;
#pragma unroll
        for (int s = 0; s < M; ++s) {
            if (s >= M - k) {
                const size_t j = o + (size_t)(s - (M - k));
                const bool ok = (key[s] & imask) != imask;
                res[j] = ok ? (key[s] & imask) : ~0ull;
                dists[j] = ok ? (DOut)dist_of(key[s]) : missing_dist<DOut>();
            }
        }
    }
};



// ===========================================================================
// Packet walk: the 64 lanes of a wavefront share one DFS over tree t.  A node
// is expanded when some active lane's box distance does not exceed that
// lane's current k-th distance (h.worst_d()); a subtree with at most `bucket`
// leaves is scanned as one contiguous run of leaf slots.  Node data is
// wave-uniform and the stack (PK_STACK entries) lives in LDS.  Each lane's
// result depends only on its own pruning tests: a child box lies inside its
// parent's box and every bound only shrinks, so sharing the walk never drops a
// node a lane needs.  The depth of a Karras tree over (64-bit key, 32-bit
// index) is below 96, so the stack (one entry per level) cannot overflow.
//
// h must be initialised by the caller (k, bound); self = leaf rank excluded
// from the lane's results (-1: none); inactive lanes take no candidates.
// [fill_lo, fill_hi): leaf ranks the caller already put into this lane's h
// (fill_from_run); the window pre-scan skips them.
// ===========================================================================
constexpr int PK_STACK = 128;

struct PacketCounters {
    uint64_t nodes = 0, leaves = 0, inserts = 0, overflow = 0;
    uint64_t rounds = 0, chains = 0;  // per wave: insertion rounds, and those that ran the chain
};

// A 32-bit coordinate's bits, and back (LDS staging of leaf coordinates).
template <typename T> __device__ __forceinline__ uint32_t bits32(T v) {
    if constexpr (std::is_same<T, float>::value) return __float_as_uint(v);
    else return (uint32_t)v;
}
template <typename T> __device__ __forceinline__ T from_bits32(uint32_t b) {
    if constexpr (std::is_same<T, float>::value) return __uint_as_float(b);
    else return (T)b;
}

template <typename T, typename H, int MK, bool FLOOR = false>
__device__ __forceinline__ void packet_walk(const DTree<T> &t, int bucket, bool active, T px, T py,
                                            T pz, int self, H &h, int32_t *stk,
                                            PacketCounters &c, int win_lo = 0, int win_hi = 0,
                                            int win_mid = 0,
                                            typename H::Floor lo = H::no_floor(),
                                            int fill_lo = 0, int fill_hi = 0) {
    using D = typename CoordTraits<T>::D;
    const int lane = threadIdx.x & 63;
    const int n = t.n;
    const int off = n - 1;
    const T *lx = t.mnx + off;
    const T *ly = t.mny + off;
    const T *lz = t.mnz + off;
    // Multi-pass queries (FLOOR): only (d2, index) at or above the floor lo
    // are taken, so a pass finds the M nearest beyond those of the passes
    // before it.
    [[maybe_unused]] D lo_d = D(0);
    if constexpr (FLOOR) lo_d = (D)h.floor_d(lo);
    // Integer periodic metric: pre-shifted 32-bit query (wrap_gap_pre) and the
    // query masked to B bits for the exact inside-the-box tests.
    constexpr bool PRE = !CoordTraits<T>::fp && MK == MK_WRAP;
    // The leaf filter's operands (pre-shifted integer or float32 coordinates
    // of leaf j in slot j): one 16-byte LDS broadcast read per leaf serves
    // every lane, instead of three readlanes on the vector unit.  The walk
    // runs on one wave, whose LDS operations complete in order.
    __shared__ uint4 s_leaf[64];
    [[maybe_unused]] const int wsh = PRE ? 32 - t.m.bits : 0;
    [[maybe_unused]] uint32_t qxs = 0, qys = 0, qzs = 0;
    [[maybe_unused]] T qmx = px, qmy = py, qmz = pz;
    if constexpr (PRE) {
        qxs = (uint32_t)px << wsh;
        qys = (uint32_t)py << wsh;
        qzs = (uint32_t)pz << wsh;
        if constexpr (sizeof(T) >= 8) {
            const T mask = (T)((T(1) << t.m.bits) - T(1));
            qmx = px & mask;
            qmy = py & mask;
            qmz = pz & mask;
        }
    }
    // Double coordinates: the float32 leaf pre-filter (fpf_*), with E from
    // the largest coordinate magnitude of the root box and the query.  Not in
    // multi-pass walks (FLOOR): there its registers would push the kernel to
    // one wave per SIMD.
    constexpr bool FPF = std::is_same<T, double>::value && MK != MK_ANY && !FLOOR;
    [[maybe_unused]] float fqx = 0.f, fqy = 0.f, fqz = 0.f, fL = 0.f;
    [[maybe_unused]] double fE = 0.0;
    if constexpr (FPF) {
        fqx = (float)px;
        fqy = (float)py;
        fqz = (float)pz;
        const double mr = fmax(fmax(fmax(fabs(t.mnx[0]), fabs(t.mxx[0])),
                                    fmax(fabs(t.mny[0]), fabs(t.mxy[0]))),
                               fmax(fabs(t.mnz[0]), fabs(t.mxz[0])));
        const double mq = fmax(fabs(px), fmax(fabs(py), fabs(pz)));
        fE = fpf_err(fmax(mr, mq), (double)t.m.box, MK == MK_WRAP);
        fL = (float)t.m.box;
    }
    // Node test.  Integer periodic metric: a float32 squared distance that is
    // never above the exact one by more than wrap_filter_bound's margin (gaps
    // from pre-shifted coordinates, the inside test on the full values as in
    // box_d2_k), compared with the scaled, raised bound, so no node within the
    // k-th distance is skipped; a few beyond it may be visited, which changes
    // nothing in the results.  Otherwise the exact distance and bound.
    auto node_dist = [&](int nd) {
        if constexpr (PRE) {
            const float gx = wrap_box_gap_pre<T>(qmx, qxs, t.mnx[nd], t.mxx[nd], wsh);
            const float gy = wrap_box_gap_pre<T>(qmy, qys, t.mny[nd], t.mxy[nd], wsh);
            const float gz = wrap_box_gap_pre<T>(qmz, qzs, t.mnz[nd], t.mxz[nd], wsh);
            float d = gx * gx;
            d = __builtin_fmaf(gy, gy, d);
            d = __builtin_fmaf(gz, gz, d);
            return d;
        } else {
            return node_d2_k<T, MK>(t, px, py, pz, nd);
        }
    };
    auto node_bound = [&]() {
        if constexpr (PRE) return wrap_filter_bound_pre((uint64_t)h.worst_d(), wsh);
        else return (D)h.worst_d();
    };

    // Scans leaf ranks [first, first+cnt) (cnt <= 64).  One coalesced load
    // puts leaf first+j in lane j.  Phase 1 broadcasts each leaf and marks,
    // per lane, the leaves not farther than its current k-th distance;
    // phase 2 inserts the marked leaves in rounds (each lane fetches its own
    // next candidate from the holding lane), so the insertion chain runs
    // once per round instead of once per leaf.
    auto scan_range = [&](int first, int cnt, bool skip_window) {
        T bx = T(0), by = T(0), bz = T(0);
        int32_t bo = 0;
        if (lane < cnt) {
            bx = lx[first + lane];
            by = ly[first + lane];
            bz = lz[first + lane];
            bo = t.orig_indices ? t.orig[off + first + lane] : first + lane;
        }
        const D w = (D)h.worst_d();
        uint32_t clo = 0, chi = 0;
        const int c0 = cnt < 32 ? cnt : 32;
        [[maybe_unused]] uint32_t ex = 0, ey = 0, ez = 0;  // pre-shifted leaf j in lane j
        if constexpr (PRE) {
            // float32 pre-filter (wrap_filter_near on pre-shifted coordinates):
            // never rejects a leaf within w; the insertion below tests every
            // passing leaf exactly.
            ex = (uint32_t)bx << wsh;
            ey = (uint32_t)by << wsh;
            ez = (uint32_t)bz << wsh;
            s_leaf[lane] = make_uint4(ex, ey, ez, 0u);
            const float wf = wrap_filter_bound_pre((uint64_t)w, wsh);
            [[maybe_unused]] const float lf = wrap_filter_floor_pre((uint64_t)lo_d, wsh);
            auto near = [&](int j) -> uint32_t {
                const uint4 e = s_leaf[j];
                const float d = wrap_sq_pre(qxs, qys, qzs, e.x, e.y, e.z);
                if constexpr (FLOOR) return (d <= wf && d >= lf) ? 1u : 0u;
                else return d <= wf ? 1u : 0u;
            };
            // leaf j's bit shifted in from the top, leaves taken last to first;
            // unrolled so that several broadcast reads are in flight (8 for
            // lists of 32 slots and more, 4 below: measured)
            constexpr int UNROLL = H::kRoundsPerTrip == 2 ? 8 : 4;
#pragma unroll UNROLL
            for (int j = c0 - 1; j >= 0; --j) clo = (clo << 1) | near(j);
#pragma unroll UNROLL
            for (int j = cnt - 1; j >= 32; --j) chi = (chi << 1) | near(j);
        } else if constexpr (FPF) {
            // float32 pre-filter of double coordinates (fpf_*): never rejects
            // a leaf within w; the insertion below tests every passing leaf
            // exactly.
            s_leaf[lane] = make_uint4(__float_as_uint((float)bx), __float_as_uint((float)by),
                                      __float_as_uint((float)bz), 0u);
            const float wf = fpf_upper((double)w, fE);
            [[maybe_unused]] const float lf = FLOOR ? fpf_lower((double)lo_d, fE) : 0.0f;
            constexpr bool WR = MK == MK_WRAP;
            auto near = [&](int j) -> uint32_t {
                const uint4 e = s_leaf[j];
                const float d = fpf_sq(fpf_gap(fqx, __uint_as_float(e.x), fL, WR),
                                       fpf_gap(fqy, __uint_as_float(e.y), fL, WR),
                                       fpf_gap(fqz, __uint_as_float(e.z), fL, WR));
                if constexpr (FLOOR) return (d <= wf && d >= lf) ? 1u : 0u;
                else return d <= wf ? 1u : 0u;
            };
#pragma unroll 4
            for (int j = c0 - 1; j >= 0; --j) clo = (clo << 1) | near(j);
#pragma unroll 4
            for (int j = cnt - 1; j >= 32; --j) chi = (chi << 1) | near(j);
        } else if constexpr (sizeof(T) == 4) {
            // 32-bit coordinates (integer open box, float): exact test, operands
            // broadcast from LDS as above
            s_leaf[lane] = make_uint4(bits32(bx), bits32(by), bits32(bz), 0u);
            auto near = [&](int j) -> bool {
                const uint4 e = s_leaf[j];
                const D d = leaf_d2_k<T, MK>(t.m, px, py, pz, from_bits32<T>(e.x),
                                             from_bits32<T>(e.y), from_bits32<T>(e.z));
                if constexpr (FLOOR) return d <= w && d >= lo_d;
                else return d <= w;
            };
#pragma unroll 4
            for (int j = 0; j < c0; ++j) clo |= near(j) ? (1u << j) : 0u;
#pragma unroll 4
            for (int j = 32; j < cnt; ++j) chi |= near(j) ? (1u << (j - 32)) : 0u;
        } else {
            auto near = [&](int j) -> bool {
                const D d = leaf_d2_k<T, MK>(t.m, px, py, pz, rl(bx, j), rl(by, j), rl(bz, j));
                if constexpr (FLOOR) return d <= w && d >= lo_d;
                else return d <= w;
            };
            for (int j = 0; j < c0; ++j) clo |= near(j) ? (1u << j) : 0u;
            for (int j = 32; j < cnt; ++j) chi |= near(j) ? (1u << (j - 32)) : 0u;
        }
        uint64_t cand = ((uint64_t)chi << 32) | clo;
        if (!active) cand = 0;
        const int sr = self - first;
        if (sr >= 0 && sr < cnt) cand &= ~(1ull << sr);
        if (!skip_window && fill_hi > fill_lo) {  // leaves this lane's fill put in h
            const int a0 = fill_lo - first > 0 ? fill_lo - first : 0;
            const int a1 = fill_hi - first < cnt ? fill_hi - first : cnt;
            if (a1 > a0) {
                const uint64_t hi_m = a1 >= 64 ? ~0ull : ((1ull << a1) - 1);
                cand &= ~(hi_m & ~((1ull << a0) - 1));
            }
        }
        if (skip_window) {  // leaves of the pre-scanned window are already in h
            const int a0 = win_lo - first > 0 ? win_lo - first : 0;
            const int a1 = win_hi - first < cnt ? win_hi - first : cnt;
            if (a1 > a0) {
                const uint64_t hi_m = a1 >= 64 ? ~0ull : ((1ull << a1) - 1);
                cand &= ~(hi_m & ~((1ull << a0) - 1));
            }
        }
        c.leaves += (uint64_t)cnt;
        // One insertion round: each lane takes its lowest remaining candidate.
        // Integer periodic metrics need only the low 32 bits of coordinates.
        auto round = [&]() {
            const bool has = cand != 0;
            const int j = has ? (__ffsll((long long)cand) - 1) : 0;
            cand &= cand - 1;
            // index of leaf first + j: its rank, or orig_idx fetched from lane j
            const int32_t ri = t.orig_indices ? shfl_any(bo, j) : first + j;
            D d;
            if constexpr (PRE) {
                const uint64_t gx = wrap_gap_pre(qxs, shfl32(ex, j), wsh);
                const uint64_t gy = wrap_gap_pre(qys, shfl32(ey, j), wsh);
                const uint64_t gz = wrap_gap_pre(qzs, shfl32(ez, j), wsh);
                d = gx * gx + gy * gy + gz * gz;
            } else {
                const T x = shfl_any(bx, j), y = shfl_any(by, j), z = shfl_any(bz, j);
                d = leaf_d2_k<T, MK>(t.m, px, py, pz, x, y, z);
            }
            ++c.rounds;
            bool take;
            if constexpr (H::kSortFill) {  // packed keys: made once for test and insertion
                const uint64_t kb = h.make((uint64_t)d, ri);
                take = has && h.beats_worst_key(kb);
                if constexpr (FLOOR) take = take && kb >= lo;
                if (__any(take)) {  // a round no lane takes skips the chain
                    h.insert_key(take, kb);
                    ++c.chains;
                }
            } else {
                take = has && h.beats_worst(d, ri);
                if constexpr (FLOOR) take = take && h.above(lo, d, ri);
                if (__any(take)) {
                    h.insert(take, d, ri);
                    ++c.chains;
                }
            }
            c.inserts += take ? 1 : 0;
        };
        if constexpr (H::kRoundsPerTrip == 2) {
            while (__any(cand != 0)) {
                round();
                if (!__any(cand != 0)) break;
                round();
            }
        } else {
            while (__any(cand != 0)) round();
        }
    };

    // Window pre-scan: leaf ranks [win_lo, win_hi) -- the packet's own run of
    // leaves [win_mid, win_mid + 64) and its neighbours along the curve -- go
    // into h first, nearest first (own run, then 64-leaf blocks alternately
    // left and right), which gives most lanes a tight bound before the walk
    // and keeps the insertions that a tighter bound would reject few; the walk
    // then skips these leaves.
    const bool window = win_hi > win_lo;
    if (window) {
        const int own_hi = win_mid + 64 < win_hi ? win_mid + 64 : win_hi;
        scan_range(win_mid, own_hi - win_mid, false);
        int le = win_mid, rs = own_hi;
        while (le > win_lo || rs < win_hi) {
            if (le > win_lo) {
                const int lb = le - 64 > win_lo ? le - 64 : win_lo;
                scan_range(lb, le - lb, false);
                le = lb;
            }
            if (rs < win_hi) {
                const int re = rs + 64 < win_hi ? rs + 64 : win_hi;
                scan_range(rs, re - rs, false);
                rs = re;
            }
        }
    }

    int sp = 0;
    int node = 0;  // an internal node some lane needs (or the single leaf when n == 1)
    while (true) {
        ++c.nodes;
        int scan_first = 0, scan_cnt = 0, next = -1;
        if (n == 1) {
            scan_cnt = 1;
        } else {
            const int first = uni(t.rfirst[node]);
            const int last = uni(t.rlast[node]);
            if (window && first >= win_lo && last < win_hi) {
                // every leaf below is already in h
            } else if (last - first + 1 <= bucket) {
                scan_first = first;
                scan_cnt = last - first + 1;
            } else {
                const int l = uni(t.left[node]);
                const int r = uni(t.right[node]);
                const auto ld = node_dist(l);
                const auto rd = node_dist(r);
                const auto w = node_bound();
                const bool nl = active && ld <= w;
                const bool nr = active && rd <= w;
                const bool anyl = __any(nl), anyr = __any(nr);
                // With more than `bucket` >= 2 leaves below, at most one child is a leaf.
                if (l >= off) {
                    if (anyl) { scan_first = l - off; scan_cnt = 1; }
                } else if (r >= off) {
                    if (anyr) { scan_first = r - off; scan_cnt = 1; }
                }
                const bool gl = l < off && anyl;
                const bool gr = r < off && anyr;
                if (gl && gr) {
                    // Nearer child first, by vote of the lanes needing either.
                    const uint64_t needing = __ballot(nl || nr);
                    const uint64_t lfirst = __ballot((nl || nr) && ld <= rd);
                    const bool left_first = 2 * __popcll(lfirst) >= __popcll(needing);
                    next = left_first ? l : r;
                    if (sp < PK_STACK) stk[sp++] = left_first ? r : l;
                    else ++c.overflow;
                } else if (gl) {
                    next = l;
                } else if (gr) {
                    next = r;
                }
            }
        }
        if (scan_cnt > 0) scan_range(scan_first, scan_cnt, window);
        while (next < 0 && sp > 0) {
            const int cn = uni(stk[--sp]);
            if (__any(active && node_dist(cn) <= node_bound())) next = cn;
        }
        if (next < 0) break;
        node = next;
    }
}

// ---------------------------------------------------------------------------
// Curve-neighbour fill of a self-query packet (queries = leaf ranks
// [r0, r0 + 64)): each lane's k live slots take the exact keys of its k
// nearest ranks along the curve (k/2 below, the rest above, itself skipped
// when excluded), sorted in registers -- one bitonic sort instead of k
// insertion rounds, and a tight bound before the window pre-scan.  Leaves
// [r0 - 32, r0 + 96) are staged in LDS (st: 128 entries) so each lane reads
// its candidates without shuffles.  Returns this lane's filled rank range
// [flo, fhi) for the window pre-scan to skip.  k <= 32.
// ---------------------------------------------------------------------------
template <typename T> struct FillStage {
    T x[128], y[128], z[128];
    int32_t idx[128];
};

template <typename T, int MK, typename H>
__device__ __forceinline__ int fill_from_run(const DTree<T> &t, H &h, int k, bool active,
                                             bool excl, int rank, T px, T py, T pz, int r0,
                                             FillStage<T> *st, int &flo, int &fhi) {
    const int lane = threadIdx.x & 63;
    const int n = t.n, off = n - 1;
    const int base = r0 - 32;
#pragma unroll
    for (int e = lane; e < 128; e += 64) {
        const int j = base + e;
        if (j >= 0 && j < n) {
            st->x[e] = t.mnx[off + j];
            st->y[e] = t.mny[off + j];
            st->z[e] = t.mnz[off + j];
            st->idx[e] = ret_index(t, off + j);
        }
    }
    __syncthreads();  // the block is this one wave
    const int half = k >> 1;
    const int kept = h.fill_sorted(k, [&](int i) -> uint64_t {
        int o = i - half;
        if (excl && o >= 0) ++o;
        const int j = rank + o;
        if (!active || j < 0 || j >= n) return ~0ull;
        const int e = j - base;
        const auto d = leaf_d2_k<T, MK>(t.m, px, py, pz, st->x[e], st->y[e], st->z[e]);
        return h.make((uint64_t)d, st->idx[e]);
    });
    flo = active ? rank - half : 0;
    fhi = active ? rank + (k - half) + (excl ? 1 : 0) : 0;
    __syncthreads();
    return kept;
}

// ---------------------------------------------------------------------------
// Packed (d2, index) keys in a max-heap per lane in LDS, layout [slot][64
// lanes] so a wave's accesses to one slot hit 64 distinct banks: the top-k
// container for k above the register tiers (k <= KCAP).  Same key form and
// re-run rule as TopKPacked; the root (the current k-th key) is mirrored in a
// register so bound tests never touch LDS.  Insertion replaces the root and
// sifts down (log2 k levels); store() heapsorts in place.
// ---------------------------------------------------------------------------
template <int KCAP> struct TopKHeapLds {
    static constexpr bool kPacked = true;
    static constexpr bool kSortFill = false;
    static constexpr int kRoundsPerTrip = 2;
    using Floor = uint64_t;  // (single pass only)
    __device__ static inline Floor no_floor() { return 0ull; }
    uint64_t *key;  // [KCAP][64], owned by the kernel
    uint64_t root, sat, imask;
    int ib, k, lane;
    __device__ inline uint64_t &K(int s) { return key[s * 64 + lane]; }
    __device__ inline void init(int k_, uint64_t bound, int ibits) {
        k = k_;
        ib = ibits;
        lane = threadIdx.x & 63;
        sat = (1ull << (64 - ib)) - 1;
        imask = (1ull << ib) - 1;
        root = ((bound < sat ? bound : sat) << ib) | imask;
        for (int s = 0; s < k; ++s) K(s) = root;
    }
    __device__ inline uint64_t make(uint64_t d, int32_t i) const {
        return ((d < sat ? d : sat) << ib) | (uint64_t)(uint32_t)i;
    }
    __device__ inline uint64_t worst_d() const { return root >> ib; }
    __device__ inline bool beats_worst(uint64_t d, int32_t i) const { return make(d, i) < root; }
    // Sift c down from the root of a heap of `size` entries.
    __device__ inline void sift_root(uint64_t c, int size) {
        int pos = 0;
        while (true) {
            const int l = 2 * pos + 1;
            if (l >= size) break;
            uint64_t kb = K(l);
            int b = l;
            if (l + 1 < size) {
                const uint64_t kr = K(l + 1);
                if (kr > kb) {
                    kb = kr;
                    b = l + 1;
                }
            }
            if (kb <= c) break;
            K(pos) = kb;
            pos = b;
        }
        K(pos) = c;
    }
    __device__ inline void insert(bool take, uint64_t d, int32_t i) {
        if (take) {
            sift_root(make(d, i), k);
            root = K(0);
        }
    }
    __device__ inline bool needs_exact(int) const {
        return ((root >> ib) >= sat) | ((root & imask) == imask);
    }
    __device__ inline bool missing(int) const { return (root & imask) == imask; }
    template <typename DOut>
    __device__ inline void store(int, uint64_t *res, DOut *dists, size_t q) {
        for (int hs = k; hs > 1; --hs) {
            const uint64_t top = K(0);
            sift_root(K(hs - 1), hs - 1);
            K(hs - 1) = top;
        }
        const size_t o = q * (size_t)k;
        for (int s = 0; s < k; ++s) {
            const uint64_t v = K(s);
            const bool ok = (v & imask) != imask;
            res[o + s] = ok ? (v & imask) : ~0ull;
            dists[o + s] = ok ? (DOut)(v >> ib) : missing_dist<DOut>();
        }
    }
};

// Top-k container of a packet traversal: registers for M <= 32 (canonical
// (d, idx) pairs, or packed keys for integer metrics), an LDS heap above that
// (integer metrics only).
template <typename D, int M, bool PACKED> struct TopKSel { using type = TopK<D, M>; };
template <int M> struct TopKSel<uint64_t, M, true> {
    using type = typename std::conditional<(M > 32), TopKHeapLds<M>, TopKPacked<M>>::type;
};

// Leaf rank reached from the root by a greedy descent for one point: at every
// node the child whose box is nearer (ties to the left).  The point and every
// decision are uniform across the wave, so the loads are scalar.  External
// query packets centre their window pre-scan on it: it is a leaf near the
// packet's queries along the curve.
template <typename T, int MK>
__device__ inline int descend_rank(const DTree<T> &t, T cx, T cy, T cz) {
    const int off = t.n - 1;
    int v = 0;
    while (v < off) {
        const int l = uni(t.left[v]), r = uni(t.right[v]);
        const auto dl = node_d2_k<T, MK>(t, cx, cy, cz, l);
        const auto dr = node_d2_k<T, MK>(t, cx, cy, cz, r);
        v = uni(dl <= dr ? l : r);
    }
    return v - off;
}

// descend_rank with both children's boxes tested in one vector pass (even
// lanes the left child, odd lanes the right): for walks that run one query
// per wavefront, where the scalar unit is the scarcer issue slot.
template <typename T, int MK>
__device__ inline int descend_rank_v(const DTree<T> &t, T cx, T cy, T cz) {
    const int off = t.n - 1;
    const int lane = threadIdx.x & 63;
    int v = 0;
    while (v < off) {
        const int l = uni(t.left[v]), r = uni(t.right[v]);
        const auto d = node_d2_k<T, MK>(t, cx, cy, cz, (lane & 1) ? r : l);
        v = uni(rl(d, 0) <= rl(d, 1) ? l : r);
    }
    return v - off;
}

// First leaf rank of the 64-leaf run centred on rank rc, inside [0, n).
__device__ inline int run_start(int rc, int n) {
    int r0 = rc - 32 < 0 ? 0 : rc - 32;
    if (r0 > n - 64) r0 = n - 64 > 0 ? n - 64 : 0;
    return r0;
}

// ---------------------------------------------------------------------------
// Per-lane DFS walks (one query per lane): pop, node d2, prune, leaf insert or
// both children pushed, the farther first (src/hlbvh.hpp:1545-1628).  The
// stack is [DEPTH][lanes] in LDS.
//  * REF (LDS heap only): the reference's walk exactly -- `>=` prune, pushes
//    only while sp < DEPTH - 2 = 62, deeper children dropped and counted.
//  * canonical: strict prune; a push that would pass DEPTH - 2 sets the
//    returned overflow flag instead.  The caller re-runs such a query with a
//    128-entry stack (LANE_STACK_CANON), which no Karras tree over (64-bit
//    key, 32-bit index) can fill, so results never depend on the stack size.
// GROUPED: the wave skips a node unless some active lane still needs it
// (any_of_group, src/hlbvh.hpp:1563-1566); results do not depend on it.
// ---------------------------------------------------------------------------
template <typename T, int M, bool GROUPED, int DEPTH, int MK>
__device__ __forceinline__ bool lane_reg_walk(const DTree<T> &t, bool active, T px, T py, T pz,
                                              int self, TopK<typename CoordTraits<T>::D, M> &h,
                                              int32_t *stk, int lanes, int lane,
                                              uint64_t &visits, uint64_t &inserts,
                                              int wlo = 0, int whi = 0) {
    using D = typename CoordTraits<T>::D;
    const int n = t.n;
    bool overflow = false;
    int sp = 0;
    if (active) stk[(sp++) * lanes + lane] = 0;
    while (sp > 0) {
        const int node = stk[(--sp) * lanes + lane];
        const D d2 = node_d2_k<T, MK>(t, px, py, pz, node);
        ++visits;
        const bool pruned = d2 > h.worst();
        if (GROUPED) {
            if (!__any(!pruned)) continue;
        }
        if (pruned) continue;
        if (node >= n - 1 || n == 1) {
            const int leaf = (n == 1) ? 0 : node;
            const int rk = leaf - (n - 1);
            if (rk == self) continue;
            if (rk >= wlo && rk < whi) continue;  // already in h (pre-scanned)
            const int32_t ri = ret_index(t, leaf);
            const bool take = h.beats_worst(d2, ri);
            h.insert(take, d2, ri);
            inserts += take ? 1 : 0;
        } else {
            const int l = t.left[node], r = t.right[node];
            const D ld = node_d2_k<T, MK>(t, px, py, pz, l);
            const D rd = node_d2_k<T, MK>(t, px, py, pz, r);
            if (sp < DEPTH - 2) {
                if (ld <= rd) {
                    stk[(sp++) * lanes + lane] = r;
                    stk[(sp++) * lanes + lane] = l;
                } else {
                    stk[(sp++) * lanes + lane] = l;
                    stk[(sp++) * lanes + lane] = r;
                }
            } else {
                overflow = true;
            }
        }
    }
    return overflow;
}

template <typename T, bool REF, bool GROUPED, int DEPTH, int MK>
__device__ __forceinline__ bool lane_heap_walk(const DTree<T> &t, int k, bool active, T px, T py,
                                               T pz, int self,
                                               LdsHeap<typename CoordTraits<T>::D, REF> &h,
                                               int32_t *stk, int lanes, int lane,
                                               uint64_t &visits, uint64_t &drops,
                                               uint64_t &inserts) {
    using D = typename CoordTraits<T>::D;
    const int n = t.n;
    bool overflow = false;
    int sp = 0;
    if (active) stk[(sp++) * lanes + lane] = 0;
    while (sp > 0) {
        const int node = stk[(--sp) * lanes + lane];
        const D d2 = node_d2_k<T, MK>(t, px, py, pz, node);
        ++visits;
        const bool pruned = h.prune(d2, k);
        if (GROUPED) {
            if (!__any(!pruned)) continue;
        }
        if (pruned) continue;
        if (node >= n - 1 || n == 1) {
            const int leaf = (n == 1) ? 0 : node;
            if (leaf - (n - 1) == self) continue;
            if (h.push(d2, ret_index(t, leaf), k)) ++inserts;
        } else {
            const int l = t.left[node], r = t.right[node];
            const D ld = node_d2_k<T, MK>(t, px, py, pz, l);
            const D rd = node_d2_k<T, MK>(t, px, py, pz, r);
            if (sp < DEPTH - 2) {
                if (ld <= rd) {
                    stk[(sp++) * lanes + lane] = r;
                    stk[(sp++) * lanes + lane] = l;
                } else {
                    stk[(sp++) * lanes + lane] = l;
                    stk[(sp++) * lanes + lane] = r;
                }
            } else if (REF) {
                ++drops;
            } else {
                overflow = true;
            }
        }
    }
    return overflow;
}


// ===========================================================================
// Wave-per-query tier: the canonical counterpart of SortedMergeHeap
// (src/maxheap.hpp:226-324) in knn_query_large_k (src/hlbvh.hpp:1232-1387).
// One wavefront answers one query.  Its list of C = 64 S entries
// (d2, returned index), ascending, is spread over the lanes: entry
// e = 64 s + lane is slot s of that lane.  The walk is wave-uniform (scalar
// node loads; a 128-entry stack of (node, d2) held one entry per lane in
// VGPRs), and a subtree of at most 64 leaves is tested one leaf per lane.
// Leaves below the current k-th entry join the list:
//  * a few (at most WaveList::kSparse): one at a time, each by shifting the
//    entries above its place one position up (DPP wave_shr:1 inside a slot,
//    a readlane from the slot below into lane 0);
//  * more: the 64 candidates are bitonic-sorted across the lanes (DPP
//    quad_perm for lane distances 1 and 2, ds_swizzle for 4 to 16,
//    ds_bpermute for 32); the list's last 64 entries take the smaller of
//    themselves and the reversed batch, which leaves the whole list bitonic
//    and holding the C smallest of both; a bitonic merge (in-lane across
//    slots for distances of 64 and more, cross-lane below) sorts it again.
// Keys compare as (distance bits as u64, index as u32): non-negative doubles
// order as their bit patterns, so one comparison serves both distance types.
// The list is exact for every k <= C (no saturation, nothing to re-run).
// ===========================================================================
__device__ __forceinline__ bool wk_lt(uint64_t ad, uint32_t ai, uint64_t bd, uint32_t bi) {
    return (ad < bd) | ((ad == bd) & (ai < bi));
}
template <typename D> __device__ __forceinline__ uint64_t dist_bits(D d) {
    if constexpr (std::is_floating_point<D>::value) return (uint64_t)__double_as_longlong(d);
    else return (uint64_t)d;
}
template <typename D> __device__ __forceinline__ D bits_dist(uint64_t b) {
    if constexpr (std::is_floating_point<D>::value) return __longlong_as_double((long long)b);
    else return (D)b;
}

// Lane i receives v from lane i ^ J.
template <int J> __device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
    if constexpr (J == 1)
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
    else if constexpr (J == 2)
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
    else if constexpr (J < 32)
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (J << 10));  // bit mode, xor J
    else
        return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(((threadIdx.x & 63) ^ 32) << 2), (int)v);
}
// Lane i receives v from lane i - 1; lane 0 receives fill (DPP wave_shr:1).
__device__ __forceinline__ uint32_t lane_shr1(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138, 0xF, 0xF, false);
}

// Compare-exchange with lane ^ J: the lane keeps the larger key when want_max.
template <int J>
__device__ __forceinline__ void wk_cx(uint64_t &d, uint32_t &i, bool want_max) {
    const uint64_t pd = ((uint64_t)lane_xor<J>((uint32_t)(d >> 32)) << 32) | lane_xor<J>((uint32_t)d);
    const uint32_t pi = lane_xor<J>(i);
    // equal keys (two empty entries) are the same value either way
    const bool take = wk_lt(pd, pi, d, i) != want_max;
    d = take ? pd : d;
    i = take ? pi : i;
}
// Bitonic sort of one key per lane, ascending over the 64 lanes.
template <int KK, int J>
__device__ __forceinline__ void wsort_stage(uint64_t &d, uint32_t &i, int lane) {
    wk_cx<J>(d, i, ((lane & J) != 0) != ((lane & KK) != 0));
    if constexpr (J > 1) wsort_stage<KK, J / 2>(d, i, lane);
}
template <int KK> __device__ __forceinline__ void wsort_from(uint64_t &d, uint32_t &i, int lane) {
    wsort_stage<KK, KK / 2>(d, i, lane);
    if constexpr (KK < 64) wsort_from<KK * 2>(d, i, lane);
}
// Half-cleaners of lane distance J..1 (the cross-lane end of a bitonic merge).
template <int J> __device__ __forceinline__ void wmerge_lanes(uint64_t &d, uint32_t &i, int lane) {
    wk_cx<J>(d, i, (lane & J) != 0);
    if constexpr (J > 1) wmerge_lanes<J / 2>(d, i, lane);
}

constexpr int WAVE_MAX_K = 1024;

template <int S> struct WaveList {
    static constexpr int C = 64 * S;
    // Candidates of one leaf batch inserted one at a time up to this count;
    // more take the sort-and-merge path (at 10M clustered points, external
    // k = 128: 22.0 ms with 16 for every S, 21.4 ms with these).
    static constexpr int kSparse = S == 1 ? 8 : (S == 2 ? 10 : 12);
    uint64_t d[S];
    uint32_t i[S];
    uint64_t wd;  // the k-th entry (wave-uniform): pruning and admission bound
    uint32_t wi;
    int ks, kl;   // its slot and lane
    __device__ inline void init(int k) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            d[s] = ~0ull;
            i[s] = ~0u;
        }
        wd = ~0ull;
        wi = ~0u;
        ks = (k - 1) >> 6;
        kl = (k - 1) & 63;
    }
    __device__ inline void refresh() {
        uint64_t xd = d[0];
        uint32_t xi = i[0];
#pragma unroll
        for (int s = 1; s < S; ++s) {
            xd = ks == s ? d[s] : xd;
            xi = ks == s ? i[s] : xi;
        }
        wd = rl(xd, kl);
        wi = rl32(xi, kl);
    }
    __device__ inline bool admits(uint64_t cd, uint32_t ci) const { return wk_lt(cd, ci, wd, wi); }
    // One wave-uniform key into its place; the last entry drops off.  Slots
    // whose last entry is below the key keep their entries (so do all slots
    // below them): the shift stops there.
    __device__ inline void insert_one(uint64_t cd, uint32_t ci) {
#pragma unroll
        for (int s = S - 1; s >= 0; --s) {
            // lane 0 of slot s receives entry 64 s - 1 (slot s - 1, lane 63,
            // not yet shifted); entry 0 receives (0, 0), at or below every key
            uint64_t fd = 0;
            uint32_t fi = 0;
            if (s > 0) {
                fd = rl(d[s - 1], 63);
                fi = rl32(i[s - 1], 63);
            }
            if (s < S - 1) {
                const uint64_t ld = rl(d[s], 63);
                const uint32_t li = rl32(i[s], 63);
                if (wk_lt(ld, li, cd, ci)) break;
            }
            const uint64_t sd = ((uint64_t)lane_shr1((uint32_t)(d[s] >> 32), (uint32_t)(fd >> 32)) << 32) |
                                lane_shr1((uint32_t)d[s], (uint32_t)fd);
            const uint32_t si = lane_shr1(i[s], fi);
            const bool keep = wk_lt(d[s], i[s], cd, ci);
            const bool put = wk_lt(sd, si, cd, ci);
            d[s] = keep ? d[s] : (put ? cd : sd);
            i[s] = keep ? i[s] : (put ? ci : si);
        }
    }
    // Merge one key per lane (non-candidates hold (~0, ~0)) into the list.
    __device__ inline void merge_batch(uint64_t cd, uint32_t ci, int lane) {
        wsort_from<2>(cd, ci, lane);
        const int src = (63 - lane) << 2;
        const uint64_t rd =
            ((uint64_t)(uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)(cd >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)cd);
        const uint32_t ri = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)ci);
        const bool t = wk_lt(rd, ri, d[S - 1], i[S - 1]);
        d[S - 1] = t ? rd : d[S - 1];
        i[S - 1] = t ? ri : i[S - 1];
#pragma unroll
        for (int js = S / 2; js >= 1; js >>= 1) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if ((s & js) == 0) {
                    const bool sw = wk_lt(d[s + js], i[s + js], d[s], i[s]);
                    const uint64_t ad = d[s], bd = d[s + js];
                    const uint32_t ai = i[s], bi = i[s + js];
                    d[s] = sw ? bd : ad;
                    i[s] = sw ? bi : ai;
                    d[s + js] = sw ? ad : bd;
                    i[s + js] = sw ? ai : bi;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s) wmerge_lanes<32>(d[s], i[s], lane);
    }
    template <typename DOut>
    __device__ inline void store(int k, uint64_t *res, DOut *dists, size_t q, int lane) const {
        const size_t o = q * (size_t)k;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int e = 64 * s + lane;
            if (e < k) {
                const bool ok = i[s] != ~0u;
                res[o + e] = ok ? (uint64_t)i[s] : ~0ull;
                dists[o + e] = ok ? bits_dist<DOut>(d[s]) : missing_dist<DOut>();
            }
        }
    }
};

struct WaveCounters {
    uint64_t nodes = 0, leaves = 0, inserts = 0;
};

// A wave's candidate buffer in LDS: leaves below the list's k-th entry wait
// here, unsorted, and join the list 64 at a time through merge_batch.
struct WaveBuf {
    uint64_t d[64];
    uint32_t i[64];
};

// The wave's walk for one query (px, py, pz): self = leaf rank excluded from
// the results (-1: none); anchor = a leaf rank near the query (-1: none).
// The 64 (S + 2) leaves around the anchor are inserted first, so the walk
// starts with a finite bound; it then skips them (and any subtree inside
// them).  Returns true if the stack would overflow (never for a tree over
// distinct (64-bit key, 32-bit index) strings, whose depth is below 97).
template <typename T, int S, int MK>
__device__ __forceinline__ bool wave_walk(const DTree<T> &t, T px, T py, T pz, int self,
                                          int anchor, WaveList<S> &L, WaveCounters &c,
                                          WaveBuf *buf) {
    using D = typename CoordTraits<T>::D;
    const int lane = threadIdx.x & 63;
    const int n = t.n;
    const int off = n - 1;
    int w0 = 0, w1 = 0;
    // Lists of 256 entries or more (S >= 4) buffer their candidates (the
    // merge is cheap per candidate and the list's bound lags by at most 64 of
    // its k entries: at 10M clustered points, 2.5M self queries at k = 256
    // took 60.5 ms buffered, 72.1 ms inserting at once); shorter lists insert
    // at once (buffered, k = 64 external queries took 16.6 against 15.8 ms).
    constexpr bool BUFFERED = S >= 4;
    int nbuf = 0;  // candidates waiting in buf (wave-uniform)
    // The buffered candidates join the list (one merge of up to 64).
    auto flush = [&]() {
        if (!BUFFERED || nbuf == 0) return;
        const bool has = lane < nbuf;
        const uint64_t bd = has ? buf->d[lane] : ~0ull;
        const uint32_t bi = has ? buf->i[lane] : ~0u;
        L.merge_batch(bd, bi, lane);
        L.refresh();
        nbuf = 0;
    };
    // Appends the lanes' candidates that the list admits (flushing first when
    // they would not fit), or, for short lists, inserts them at once: a few
    // one at a time, more by one sort-and-merge.
    auto offer = [&](bool ok, uint64_t cd, uint32_t ci) {
        uint64_t m = __ballot(ok);
        if (!m) return;
        int cnt_in = __popcll(m);
        if constexpr (!BUFFERED) {
            c.inserts += (uint64_t)cnt_in;
            if (cnt_in <= WaveList<S>::kSparse) {
                uint64_t mm = m;
                while (mm) {
                    const int j = __builtin_ctzll(mm);
                    mm &= mm - 1;
                    const uint64_t xd = rl(cd, j);
                    const uint32_t xi = rl32(ci, j);
                    if (L.admits(xd, xi)) {
                        L.insert_one(xd, xi);
                        L.refresh();
                    }
                }
            } else {
                L.merge_batch(ok ? cd : ~0ull, ok ? ci : ~0u, lane);
                L.refresh();
            }
            return;
        }
        if (nbuf + cnt_in > 64) {
            flush();
            ok = ok && L.admits(cd, ci);  // against the merged list's k-th entry
            m = __ballot(ok);
            cnt_in = __popcll(m);
            if (!m) return;
        }
        c.inserts += (uint64_t)cnt_in;
        const int pos = nbuf + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (ok) {
            buf->d[pos] = cd;
            buf->i[pos] = ci;
        }
        nbuf += cnt_in;
    };
    // Tests leaves [r0, r0 + cnt) (cnt <= 64), one per lane, and offers the
    // ones the list admits.
    auto scan = [&](int r0, int cnt, bool skip_win) {
        const int r = r0 + lane;
        bool ok = lane < cnt && r != self;
        if (skip_win) ok = ok && (r < w0 || r >= w1);
        const int leaf = off + (lane < cnt ? r : r0);
        const D d2 = leaf_d2_k<T, MK>(t.m, px, py, pz, t.mnx[leaf], t.mny[leaf], t.mnz[leaf]);
        const uint64_t cd = dist_bits(d2);
        const uint32_t ci = (uint32_t)ret_index(t, leaf);
        ok = ok && L.admits(cd, ci);
        c.leaves += (uint64_t)cnt;
        offer(ok, cd, ci);
    };
    const int nwin = 64 * (S + 2);
    if (anchor >= 0) {
        w0 = anchor - nwin / 2;
        if (w0 > n - nwin) w0 = n - nwin;
        if (w0 < 0) w0 = 0;
        w1 = w0 + nwin < n ? w0 + nwin : n;
        for (int r0 = w0; r0 < w1; r0 += 64) scan(r0, w1 - r0 < 64 ? w1 - r0 : 64, false);
        flush();
        if (w0 == 0 && w1 == n) return false;
    }
    if (n <= 64) {
        scan(0, n, true);
        flush();
        return false;
    }
    // stack entry e: lane e & 63 of register (e >> 6)
    uint32_t sn0 = 0, sn1 = 0, sd0l = 0, sd0h = 0, sd1l = 0, sd1h = 0;
    int sp = 0;
    int node = 0;
    while (true) {
        ++c.nodes;
        if (node >= off) {
            const int r = node - off;
            if (r != self && (r < w0 || r >= w1)) {
                // on the vector unit (the node index copied into a VGPR): the
                // scalar unit, shared by the CU's four SIMDs, is this walk's
                // scarcer issue slot
                int nv;
                asm("v_mov_b32 %0, %1" : "=v"(nv) : "s"(node));
                const D d2v = leaf_d2_k<T, MK>(t.m, px, py, pz, t.mnx[nv], t.mny[nv], t.mnz[nv]);
                const uint64_t cd = dist_bits(rl(d2v, 0));
                const uint32_t ci = (uint32_t)ret_index(t, node);
                ++c.leaves;
                offer(lane == 0 && L.admits(cd, ci), cd, ci);
            }
        } else {
            const int rf = uni(t.rfirst[node]), rl_ = uni(t.rlast[node]);
            if (rf >= w0 && rl_ < w1) {
                // inside the window: already inserted
            } else if (rl_ - rf < 64) {
                scan(rf, rl_ - rf + 1, true);
            } else {
                const int l = uni(t.left[node]), r = uni(t.right[node]);
                // both children's box distances in one vector pass (even
                // lanes the left child, odd lanes the right) instead of two
                // wave-uniform computations on the scalar unit
                const int ch = (lane & 1) ? r : l;
                const D dch = node_d2_k<T, MK>(t, px, py, pz, ch);
                const uint64_t dl = dist_bits(rl(dch, 0)), dr = dist_bits(rl(dch, 1));
                const bool lfirst = dl <= dr;
                const int nn = uni(lfirst ? l : r), fn = uni(lfirst ? r : l);
                const uint64_t dn = lfirst ? dl : dr, df = lfirst ? dr : dl;
                if (df <= L.wd) {
                    if (sp >= 128) return true;
                    const uint32_t lo = (uint32_t)df, hi = (uint32_t)(df >> 32);
                    const bool at = lane == (sp & 63);
                    if (sp < 64) {
                        sn0 = at ? (uint32_t)fn : sn0;
                        sd0l = at ? lo : sd0l;
                        sd0h = at ? hi : sd0h;
                    } else {
                        sn1 = at ? (uint32_t)fn : sn1;
                        sd1l = at ? lo : sd1l;
                        sd1h = at ? hi : sd1h;
                    }
                    ++sp;
                }
                if (dn <= L.wd) {
                    node = nn;
                    continue;
                }
            }
        }
        // pop the next entry still within the bound
        bool found = false;
        while (sp > 0) {
            --sp;
            int v;
            uint64_t dv;
            if (sp < 64) {
                v = (int)rl32(sn0, sp);
                dv = ((uint64_t)rl32(sd0h, sp) << 32) | rl32(sd0l, sp);
            } else {
                v = (int)rl32(sn1, sp - 64);
                dv = ((uint64_t)rl32(sd1h, sp - 64) << 32) | rl32(sd1l, sp - 64);
            }
            if (dv <= L.wd) {
                node = v;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    flush();
    return false;
}

}  // namespace ft
