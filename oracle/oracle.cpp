// oracle.cpp -- CPU restatement of FastTree's HLBVH build and neighbour queries.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU path: only
// tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.  The
// product (fasttree_amd/) never links or calls it.
//
// Each function restates the reference's algorithm and cites the file:line it
// follows (paths relative to the reference repository root).  Pins: see
// oracle/README.md for which check pins which function; functions without an
// independent check are marked "parity unpinned" there.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <parallel/algorithm>
#include <random>
#include <tuple>
#include <type_traits>
#include <vector>

#include <omp.h>

namespace {

enum { CT_F32 = 0, CT_F64 = 1, CT_U32 = 2, CT_U64 = 3 };
enum { SFC_MORTON = 0, SFC_PH = 1 };

// ---------------------------------------------------------------------------
// Peano-Hilbert state tables (src/sfc.peano_hilbert.hpp:332-376).  They are the
// curve's definition and must be the same data: next state and 3-bit digit for
// (state, octant pix = bx<<2|by<<1|bz).
// ---------------------------------------------------------------------------
static const uint8_t kPHNext[48][8] = {
    {36,28,25,27,10,10,25,27},
    {29,11,24,24,37,11,26,26},
    { 8, 8,25,27,30,38,25,27},
    { 9,39,24,24, 9,31,26,26},
    {40,24,44,32,40, 6,44, 6},
    {25, 7,33, 7,41,41,45,45},
    { 4,42, 4,46,26,42,34,46},
    {43,43,47,47, 5,27, 5,35},
    {33,35,36,28,33,35, 2, 2},
    {32,32,29, 3,34,34,37, 3},
    {33,35, 0, 0,33,35,30,38},
    {32,32, 1,39,34,34, 1,31},
    {24,42,32,46,14,42,14,46},
    {43,43,47,47,25,15,33,15},
    {40,12,44,12,40,26,44,34},
    {13,27,13,35,41,41,45,45},
    {28,41,28,22,38,43,38,22},
    {42,40,23,23,29,39,29,39},
    {41,36,20,36,43,30,20,30},
    {37,31,37,31,42,40,21,21},
    {28,18,28,45,38,18,38,47},
    {19,19,46,44,29,39,29,39},
    {16,36,45,36,16,30,47,30},
    {37,31,37,31,17,17,46,44},
    {12, 4, 1, 3,34,34, 1, 3},
    { 5,35, 0, 0,13,35, 2, 2},
    {32,32, 1, 3, 6,14, 1, 3},
    {33,15, 0, 0,33, 7, 2, 2},
    {16, 0,20, 8,16,30,20,30},
    { 1,31, 9,31,17,17,21,21},
    {28,18,28,22, 2,18,10,22},
    {19,19,23,23,29, 3,29,11},
    { 9,11,12, 4, 9,11,26,26},
    { 8, 8, 5,27,10,10,13,27},
    { 9,11,24,24, 9,11, 6,14},
    { 8, 8,25,15,10,10,25, 7},
    { 0,18, 8,22,38,18,38,22},
    {19,19,23,23, 1,39, 9,39},
    {16,36,20,36,16, 2,20,10},
    {37, 3,37,11,17,17,21,21},
    { 4,17, 4,46,14,19,14,46},
    {18,16,47,47, 5,15, 5,15},
    {17,12,44,12,19, 6,44, 6},
    {13, 7,13, 7,18,16,45,45},
    { 4,42, 4,21,14,42,14,23},
    {43,43,22,20, 5,15, 5,15},
    {40,12,21,12,40, 6,23, 6},
    {13, 7,13, 7,41,41,22,20},
};
static const uint8_t kPHDigit[48][8] = {
    { 0, 7, 1, 6, 3, 4, 2, 5},
    { 7, 4, 6, 5, 0, 3, 1, 2},
    { 4, 3, 5, 2, 7, 0, 6, 1},
    { 3, 0, 2, 1, 4, 7, 5, 6},
    { 1, 0, 6, 7, 2, 3, 5, 4},
    { 0, 3, 7, 4, 1, 2, 6, 5},
    { 3, 2, 4, 5, 0, 1, 7, 6},
    { 2, 1, 5, 6, 3, 0, 4, 7},
    { 6, 1, 7, 0, 5, 2, 4, 3},
    { 1, 2, 0, 3, 6, 5, 7, 4},
    { 2, 5, 3, 4, 1, 6, 0, 7},
    { 5, 6, 4, 7, 2, 1, 3, 0},
    { 7, 6, 0, 1, 4, 5, 3, 2},
    { 6, 5, 1, 2, 7, 4, 0, 3},
    { 5, 4, 2, 3, 6, 7, 1, 0},
    { 4, 7, 3, 0, 5, 6, 2, 1},
    { 6, 7, 5, 4, 1, 0, 2, 3},
    { 7, 0, 4, 3, 6, 1, 5, 2},
    { 0, 1, 3, 2, 7, 6, 4, 5},
    { 1, 6, 2, 5, 0, 7, 3, 4},
    { 2, 3, 1, 0, 5, 4, 6, 7},
    { 3, 4, 0, 7, 2, 5, 1, 6},
    { 4, 5, 7, 6, 3, 2, 0, 1},
    { 5, 2, 6, 1, 4, 3, 7, 0},
    { 7, 0, 6, 1, 4, 3, 5, 2},
    { 0, 3, 1, 2, 7, 4, 6, 5},
    { 3, 4, 2, 5, 0, 7, 1, 6},
    { 4, 7, 5, 6, 3, 0, 2, 1},
    { 6, 7, 1, 0, 5, 4, 2, 3},
    { 7, 4, 0, 3, 6, 5, 1, 2},
    { 4, 5, 3, 2, 7, 6, 0, 1},
    { 5, 6, 2, 1, 4, 7, 3, 0},
    { 1, 6, 0, 7, 2, 5, 3, 4},
    { 6, 5, 7, 4, 1, 2, 0, 3},
    { 5, 2, 4, 3, 6, 1, 7, 0},
    { 2, 1, 3, 0, 5, 6, 4, 7},
    { 0, 1, 7, 6, 3, 2, 4, 5},
    { 1, 2, 6, 5, 0, 3, 7, 4},
    { 2, 3, 5, 4, 1, 0, 6, 7},
    { 3, 0, 4, 7, 2, 1, 5, 6},
    { 1, 0, 2, 3, 6, 7, 5, 4},
    { 0, 7, 3, 4, 1, 6, 2, 5},
    { 7, 6, 4, 5, 0, 1, 3, 2},
    { 6, 1, 5, 2, 7, 0, 4, 3},
    { 5, 4, 6, 7, 2, 3, 1, 0},
    { 4, 3, 7, 0, 5, 2, 6, 1},
    { 3, 2, 0, 1, 4, 5, 7, 6},
    { 2, 5, 1, 6, 3, 4, 0, 7},
};

template <typename T> constexpr bool is_fp_v = std::is_floating_point<T>::value;

// ---------------------------------------------------------------------------
// Key encoding
// ---------------------------------------------------------------------------

// spread3_u64 (src/sfc.morton.hpp:33-41): bit b of x moves to bit 3b, for b < 21.
static uint64_t spread3(uint64_t x) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((x >> b) & 1ull) << (3 * b);
    return r;
}

// Morton key: spread(ix) | spread(iy)<<1 | spread(iz)<<2 (src/hlbvh.hpp:386, :414).
static uint64_t morton_key(uint64_t ix, uint64_t iy, uint64_t iz) {
    return spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
}

// sfc_encode3D (src/sfc.peano_hilbert.hpp:569-600) with key_push_3bits
// (:396-405): digits are shifted in from the bottom, most significant level
// first, state starts at 0.  For 3B <= 63 the 64-bit sort key (to_sort_key,
// :279-300) is the key's numeric value.
static uint64_t ph_key(uint64_t ix, uint64_t iy, uint64_t iz, int B) {
    uint64_t key = 0;
    int state = 0;
    for (int bit = B - 1; bit >= 0; --bit) {
        int pix = (int)(((ix >> bit) & 1) << 2 | ((iy >> bit) & 1) << 1 | ((iz >> bit) & 1));
        key = (key << 3) | kPHDigit[state][pix];
        state = kPHNext[state][pix];
    }
    return key;
}

// Morton FP quantisation: encode_to_sfc1d<FloatT>(1 + nx) (src/hlbvh.hpp:246-271,
// :411-414): the top B mantissa bits of fl(1 + nx) in FloatT.  nx is clamped to
// [0, nextafter(1, 0)] in FloatT (:395-407); fl(1 + nextafter(1,0)) rounds to 2.0
// whose mantissa is 0, so the maximum maps to 0.
template <typename T> static uint64_t morton_quantise(T nx, int B) {
    T v = T(1) + nx;
    if (std::is_same<T, double>::value) {
        uint64_t bits;
        double dv = (double)v;
        std::memcpy(&bits, &dv, 8);
        return (bits & 0x000FFFFFFFFFFFFFull) >> (52 - B);
    } else {
        uint32_t bits;
        float fv = (float)v;
        std::memcpy(&bits, &fv, 4);
        return (uint64_t)((bits & 0x007FFFFFu) >> (23 - B));
    }
}

// quantize_coord (src/sfc.peano_hilbert.hpp:472-476) + convert_to_sfc1d_impl
// (:417-466): min(1 + (double)nx, 0x1.fffffffffffffp+0), top B mantissa bits.
static uint64_t ph_quantise(double nx, int B) {
    const double upper = 0x1.fffffffffffffp+0;
    double biased = std::fmin(1.0 + nx, upper);
    uint64_t bits;
    std::memcpy(&bits, &biased, 8);
    uint64_t mask = (B < 64) ? ((1ull << B) - 1) : ~0ull;
    return ((bits & 0x000FFFFFFFFFFFFFull) >> (52 - B)) & mask;
}

// compute_bbox (src/hlbvh.hpp:462-578): exact min/max per axis.
template <typename T>
static void bbox_t(const T *x, const T *y, const T *z, size_t n, T out[6]) {
    const T *a[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        T lo = a[d][0], hi = a[d][0];
        for (size_t i = 1; i < n; ++i) {
            if (a[d][i] < lo) lo = a[d][i];
            if (hi < a[d][i]) hi = a[d][i];
        }
        out[2 * d] = lo;
        out[2 * d + 1] = hi;
    }
}

// sfc_encode (src/hlbvh.hpp:357-423).
template <typename T>
static void encode_t(int sfc, int B, const T *x, const T *y, const T *z, size_t n,
                     const T *bbox, uint64_t *keys) {
    if (!is_fp_v<T>) {
        // Integer coords: ix = x & (2^B - 1) (:376-387).
        const uint64_t mask = (1ull << B) - 1;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            uint64_t ix = (uint64_t)x[i] & mask, iy = (uint64_t)y[i] & mask,
                     iz = (uint64_t)z[i] & mask;
            keys[i] = sfc == SFC_MORTON ? morton_key(ix, iy, iz) : ph_key(ix, iy, iz, B);
        }
        return;
    }
    // FP coords (:389-416): inv_d computed in T, 0 for a flat axis.
    T d[3], inv[3], lo[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = bbox[2 * a];
        d[a] = bbox[2 * a + 1] - bbox[2 * a];
        inv[a] = (d[a] == 0) ? T(0) : (T(1) / d[a]);
    }
    const T clamp_hi = std::nextafter(T(1), T(0));
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const T p[3] = {x[i], y[i], z[i]};
        uint64_t q[3];
        for (int a = 0; a < 3; ++a) {
            T t = (p[a] - lo[a]) * inv[a];
            // sycl::clamp(t, 0, clamp_hi) = fmin(fmax(t, 0), clamp_hi)
            T nx = std::fmin(std::fmax(t, T(0)), clamp_hi);
            q[a] = sfc == SFC_MORTON ? morton_quantise<T>(nx, B) : ph_quantise((double)nx, B);
        }
        keys[i] = sfc == SFC_MORTON ? morton_key(q[0], q[1], q[2]) : ph_key(q[0], q[1], q[2], B);
    }
}

// ---------------------------------------------------------------------------
// Sort (src/hlbvh.hpp:2202-2232): ascending key, equal keys by original index
// (stable).  The reference's oneDPL order for equal keys is not specified.
// ---------------------------------------------------------------------------
static void sort_keys(const uint64_t *keys, size_t n, uint32_t *perm, uint64_t *sorted) {
    std::vector<uint32_t> p(n);
    std::iota(p.begin(), p.end(), 0u);
    __gnu_parallel::stable_sort(p.begin(), p.end(),
                                [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (size_t i = 0; i < n; ++i) {
        perm[i] = p[i];
        sorted[i] = keys[p[i]];
    }
}

// ---------------------------------------------------------------------------
// Karras topology (src/hlbvh.hpp:773-903).  delta(i,j) = -1 outside [0,n);
// common prefix of the keys when they differ (get_common_prefix_length,
// :283-340, a constant offset from clz64 for keys below 2^63), else the key
// width plus the common prefix of the indices (:852-859).  Only comparisons of
// delta values matter, so any order-preserving offsets give the same tree.
// ---------------------------------------------------------------------------
static inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
static inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

static void karras(const uint64_t *k, size_t n_, int32_t *left, int32_t *right, int32_t *parent,
                   int32_t *rfirst, int32_t *rlast) {
    const int n = (int)n_;
    auto delta = [&](int i, int j) -> int {
        if (j < 0 || j >= n) return -1;
        if (k[i] != k[j]) return clz64(k[i] ^ k[j]);
        return 64 + clz32((uint32_t)(i ^ j));
    };
    for (int i = 0; i < 2 * n - 1; ++i) parent[i] = -1;  // q.fill(p_parent, -1) (:849)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n - 1; ++i) {
        int dl = delta(i, i + 1) - delta(i, i - 1);
        int d = (dl > 0) - (dl < 0);
        int dmin = delta(i, i - d);
        int lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = delta(i, j);
        int s = 0, ts = 1;
        while (ts <= l) ts *= 2;
        ts /= 2;
        for (int t = ts; t >= 1; t /= 2)
            if (s + t < l && delta(i, i + (s + t) * d) > dnode) s += t;
        int split = i + s * d + std::min(d, 0);
        int lo = std::min(i, j), hi = std::max(i, j);
        int li = (lo == split) ? split + n - 1 : split;
        int ri = (hi == split + 1) ? split + 1 + n - 1 : split + 1;
        left[i] = li;
        right[i] = ri;
        parent[li] = i;
        parent[ri] = i;
        if (rfirst) rfirst[i] = lo;
        if (rlast) rlast[i] = hi;
    }
}

// Refit (src/hlbvh.hpp:925-969): every leaf climbs; the first arrival at a node
// stops, the second combines the children with sfc_min/sfc_max (:580-594;
// fmin/fmax for FP).  Run with acq_rel atomics as the reference does.
template <typename T> static T smin(T a, T b) {
    if (is_fp_v<T>) return std::fmin(a, b);
    return (a < b) ? a : b;
}
template <typename T> static T smax(T a, T b) {
    if (is_fp_v<T>) return std::fmax(a, b);
    return (a < b) ? b : a;
}

template <typename T>
static void refit_t(size_t n_, const int32_t *left, const int32_t *right, const int32_t *parent,
                    T *mn[3], T *mx[3]) {
    const int n = (int)n_;
    std::vector<std::atomic<int>> counters(n > 1 ? n - 1 : 1);
    for (auto &c : counters) c.store(0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        int cur = i + n - 1;
        while (cur != 0) {
            int p = parent[cur];
            if (p < 0 || p >= n - 1) break;
            if (counters[p].fetch_add(1, std::memory_order_acq_rel) == 0) break;
            int l = left[p], r = right[p];
            for (int a = 0; a < 3; ++a) {
                mn[a][p] = smin(mn[a][l], mn[a][r]);
                mx[a][p] = smax(mx[a][l], mx[a][r]);
            }
            cur = p;
        }
    }
}

// ---------------------------------------------------------------------------
// Distances: node_distance_sq (src/hlbvh.hpp:998-1055).
// ---------------------------------------------------------------------------
struct Cfg {
    int sfc, B, periodic, orig;
    double box;
};

template <typename T> struct DistOf { using type = uint64_t; };
template <> struct DistOf<float> { using type = double; };
template <> struct DistOf<double> { using type = double; };

template <typename T> static inline int bits_of() { return (int)(sizeof(T) * 8); }

// shortest_periodic_distance (src/sfc.peano_hilbert.hpp:532-538) with
// SHIFT_BITS = bits(T) - B (the value both reference defaults satisfy).
template <typename T> static inline uint64_t periodic_gap(T a, T b, int B) {
    using S = typename std::make_signed<T>::type;
    const int sh = bits_of<T>() - B;
    T diff = (T)(a - b);
    S v = (S)(T)(diff << sh);
    v = v >> sh;
    return (uint64_t)(v < 0 ? -v : v);
}

template <typename T>
static inline typename DistOf<T>::type node_d2(const Cfg &c, T px, T py, T pz, T x0, T x1, T y0,
                                               T y1, T z0, T z1) {
    if constexpr (is_fp_v<T>) {
        if (!c.periodic) {
            // dx = fmax(bmin - p, fmax(0, p - bmax)); d2 = dx*dx + dy*dy + dz*dz in T
            T dx = std::fmax(x0 - px, std::fmax(T(0), px - x1));
            T dy = std::fmax(y0 - py, std::fmax(T(0), py - y1));
            T dz = std::fmax(z0 - pz, std::fmax(T(0), pz - z1));
            T s = dx * dx;
            s = s + dy * dy;
            s = s + dz * dz;
            return (double)s;
        }
        // FP periodic (no reference counterpart: PERIODIC_BC is ignored for FP
        // coords, src/hlbvh.hpp:1001-1006): per axis 0 inside [lo,hi], else the
        // smaller minimum-image distance to either face, period L.
        const T L = (T)c.box;
        auto ax = [&](T p, T lo, T hi) -> T {
            if (p >= lo && p <= hi) return T(0);
            T a = std::fabs(p - lo);
            a = std::fmin(a, L - a);
            T b = std::fabs(p - hi);
            b = std::fmin(b, L - b);
            return std::fmin(a, b);
        };
        T dx = ax(px, x0, x1), dy = ax(py, y0, y1), dz = ax(pz, z0, z1);
        T s = dx * dx;
        s = s + dy * dy;
        s = s + dz * dz;
        return (double)s;
    } else {
        const int dshift = c.B > 31 ? c.B - 31 : 0;  // _DIST_SHIFT (:185-193)
        uint64_t g[3];
        const T p[3] = {px, py, pz}, lo[3] = {x0, y0, z0}, hi[3] = {x1, y1, z1};
        for (int a = 0; a < 3; ++a) {
            T v = p[a];
            if (c.periodic) {
                // periodic_axis_dist (:1017-1033)
                if (sizeof(T) >= 8) v &= (T)((T(1) << c.B) - T(1));
                if (v >= lo[a] && v <= hi[a]) {
                    g[a] = 0;
                } else {
                    uint64_t al = periodic_gap<T>(v, lo[a], c.B);
                    uint64_t ah = periodic_gap<T>(v, hi[a], c.B);
                    g[a] = (al < ah ? al : ah) >> dshift;
                }
            } else {
                // axis_dist (:1043-1047)
                if (v < lo[a]) g[a] = (uint64_t)(T)(lo[a] - v) >> dshift;
                else if (v > hi[a]) g[a] = (uint64_t)(T)(v - hi[a]) >> dshift;
                else g[a] = 0;
            }
        }
        return g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    }
}

// ---------------------------------------------------------------------------
// Tree view
// ---------------------------------------------------------------------------
}  // namespace

extern "C" {

typedef struct or_tree {
    void *min_x, *max_x, *min_y, *max_y, *min_z, *max_z;
    int32_t *left, *right, *parent;
    uint32_t *id;
    int8_t *ghost;
    int32_t *orig;
    uint64_t n;
    int32_t ct, sfc, bits, periodic, orig_indices, pad;
    double box;
} or_tree;
}

namespace {

template <typename T> struct TV {
    const T *mn[3], *mx[3];
    const int32_t *left, *right, *orig;
    int n;
    Cfg c;
    explicit TV(const or_tree *t) {
        mn[0] = (const T *)t->min_x; mx[0] = (const T *)t->max_x;
        mn[1] = (const T *)t->min_y; mx[1] = (const T *)t->max_y;
        mn[2] = (const T *)t->min_z; mx[2] = (const T *)t->max_z;
        left = t->left; right = t->right; orig = t->orig;
        n = (int)t->n;
        c = Cfg{t->sfc, t->bits, t->periodic, t->orig_indices, t->box};
    }
    typename DistOf<T>::type d2(T px, T py, T pz, int node) const {
        return node_d2<T>(c, px, py, pz, mn[0][node], mx[0][node], mn[1][node], mx[1][node],
                          mn[2][node], mx[2][node]);
    }
    uint64_t ret_index(int leaf_node) const {
        // leaf rank, or orig_idx with RETURN_ORIG_INDICES (:1573-1577)
        return c.orig ? (uint64_t)(uint32_t)orig[leaf_node] : (uint64_t)(leaf_node - (n - 1));
    }
};

template <typename D> static D sentinel_dist() {
    // type_identity_max<dist_t>() (src/hlbvh.hpp:59-69)
    if (std::is_floating_point<D>::value) return (D)-DBL_MAX;
    return (D)0;
}

// Brute force (test/test_self_knn.cpp:75-147): all candidates sorted by
// (d2, returned index), truncated to k.  This is the canonical result.
template <typename T>
static void bruteforce_t(const or_tree *t, const T *qx, const T *qy, const T *qz,
                         const int32_t *leaf_ids, int nq, int k, int excl, uint64_t *res,
                         void *dists_v) {
    using D = typename DistOf<T>::type;
    TV<T> tv(t);
    D *dists = (D *)dists_v;
    const int n = tv.n;
    const int off = n - 1;
#pragma omp parallel
    {
        std::vector<std::pair<D, uint64_t>> cand;
        cand.reserve(n);
#pragma omp for schedule(dynamic, 16)
        for (int q = 0; q < nq; ++q) {
            T px, py, pz;
            int self = -1;
            if (qx) {
                px = qx[q]; py = qy[q]; pz = qz[q];
            } else {
                self = leaf_ids ? leaf_ids[q] : q;
                px = tv.mn[0][off + self]; py = tv.mn[1][off + self]; pz = tv.mn[2][off + self];
            }
            cand.clear();
            for (int j = 0; j < n; ++j) {
                if (excl && j == self) continue;
                int node = off + j;
                D d = node_d2<T>(tv.c, px, py, pz, tv.mn[0][node], tv.mn[0][node], tv.mn[1][node],
                                 tv.mn[1][node], tv.mn[2][node], tv.mn[2][node]);
                cand.emplace_back(d, tv.ret_index(node));
            }
            size_t kk = std::min<size_t>(k, cand.size());
            std::partial_sort(cand.begin(), cand.begin() + kk, cand.end());
            for (int i = 0; i < k; ++i) {
                if ((size_t)i < kk) {
                    res[(size_t)q * k + i] = cand[i].second;
                    dists[(size_t)q * k + i] = cand[i].first;
                } else {
                    res[(size_t)q * k + i] = ~0ull;
                    dists[(size_t)q * k + i] = sentinel_dist<D>();
                }
            }
        }
    }
}

// RegisterMaxHeap (src/maxheap.hpp:33-144); LocalMaxHeap (:340-478) runs the
// same push / sift / heapsort steps, so one restatement serves both.
template <typename D> struct MaxHeap {
    std::vector<D> dist;
    std::vector<int> idx;
    int count = 0;
    explicit MaxHeap(int cap) : dist(cap, std::numeric_limits<D>::max()), idx(cap, -1) {}
    void reset() {
        count = 0;
        std::fill(dist.begin(), dist.end(), std::numeric_limits<D>::max());
        std::fill(idx.begin(), idx.end(), -1);
    }
    bool should_prune(D d, int k) const { return count >= k && d >= dist[0]; }
    void swp(int a, int b) {
        std::swap(dist[a], dist[b]);
        std::swap(idx[a], idx[b]);
    }
    void sift_up(int i) {
        while (i > 0) {
            int p = (i - 1) >> 1;
            if (dist[i] <= dist[p]) break;
            swp(i, p);
            i = p;
        }
    }
    void sift_down(int i, int size) {
        while (true) {
            int lg = i, l = 2 * i + 1, r = 2 * i + 2;
            if (l < size && dist[l] > dist[lg]) lg = l;
            if (r < size && dist[r] > dist[lg]) lg = r;
            if (lg == i) break;
            swp(i, lg);
            i = lg;
        }
    }
    void push(D d, int i, int k) {
        if (count < k) {
            dist[count] = d;
            idx[count] = i;
            sift_up(count);
            count++;
        } else if (d < dist[0]) {
            dist[0] = d;
            idx[0] = i;
            sift_down(0, k);
        }
    }
    void sort_ascending() {  // extract_sorted / sort_in_place heapsort (:94-111, :424-434)
        int hs = count;
        while (hs > 1) {
            swp(0, hs - 1);
            hs--;
            sift_down(0, hs);
        }
    }
};

// SortedMergeHeap (src/maxheap.hpp:226-324) with local_bitonic_sort
// (:193-215) executed as the same data-oblivious network.
template <typename D> struct MergeList {
    std::vector<D> dist;
    std::vector<int> idx;
    D sd[32];
    int si[32];
    int count = 0, cap;
    explicit MergeList(int cap_) : dist(cap_, std::numeric_limits<D>::max()), idx(cap_, -1), cap(cap_) {}
    static void bitonic(D *d, int *ix, int n) {
        for (int k = 2; k <= n; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1)
                for (int i = 0; i < n; ++i) {
                    int l = i ^ j;
                    bool asc = (i & k) == 0;
                    if (l > i && l < n) {
                        if ((d[i] > d[l]) == asc) {
                            std::swap(d[i], d[l]);
                            std::swap(ix[i], ix[l]);
                        }
                    }
                }
    }
    bool should_prune(D d, int k) const { return count >= k && d >= dist[k - 1]; }
    // One candidate from lane 0, the other 31 lanes contribute (DIST_MAX, -1).
    void merge_one(D d, int i, int k) {
        for (int l = 0; l < 32; ++l) {
            sd[l] = std::numeric_limits<D>::max();
            si[l] = -1;
        }
        sd[0] = d;
        si[0] = i;
        bitonic(sd, si, 32);
        for (int l = 0; l < 32; ++l) {
            int dst = count + l;
            if (dst < cap) {
                dist[dst] = sd[l];
                idx[dst] = si[l];
            }
        }
        int new_count = std::min(count + 32, cap);
        int pow2 = 1;
        while (pow2 < new_count) pow2 <<= 1;
        for (int l = new_count; l < pow2 && l < cap; ++l) {
            dist[l] = std::numeric_limits<D>::max();
            idx[l] = -1;
        }
        bitonic(dist.data(), idx.data(), pow2);
        count = std::min(new_count, cap);
        count = std::min(count, k);
    }
};

// Reference traversal (knn_query_small_k src/hlbvh.hpp:1079-1182, the grouped
// self kernels :1520-1647 and :1690-1818, knn_query_large_k :1232-1387):
// 64-entry stack, children pushed only while sp < 62 (F12 drop), the farther
// child pushed first (ld2 <= rd2 -> left popped first), `>=` prune.  Grouping
// does not change a lane's result (a lane only pushes when its own test
// passes), so every query runs alone here.  kind 0 = max-heap tiers, kind 1 =
// sorted-merge tier with capacity next_pow2(max_k).
template <typename T>
static void reference_knn_t(const or_tree *t, const T *qx, const T *qy, const T *qz,
                            const int32_t *leaf_ids, int nq, int k, int excl, int kind, int max_k,
                            uint64_t *res, void *dists_v, uint64_t *counters) {
    using D = typename DistOf<T>::type;
    TV<T> tv(t);
    D *dists = (D *)dists_v;
    const int n = tv.n;
    const int off = n - 1;
    int cap = 1;
    while (cap < max_k) cap <<= 1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int q = 0; q < nq; ++q) {
        T px, py, pz;
        int self_node = -1;
        if (qx) {
            px = qx[q]; py = qy[q]; pz = qz[q];
        } else {
            int r = leaf_ids ? leaf_ids[q] : q;
            self_node = r + off;
            px = tv.mn[0][self_node]; py = tv.mn[1][self_node]; pz = tv.mn[2][self_node];
        }
        MaxHeap<D> heap(std::max(k, 1));
        MergeList<D> ml(kind == 1 ? cap : 1);
        uint64_t visits = 0, drops = 0;
        int stack[64];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            int node = stack[--sp];
            visits++;
            D d2 = tv.d2(px, py, pz, node);
            bool pruned = kind == 1 ? ml.should_prune(d2, k) : heap.should_prune(d2, k);
            if (pruned) continue;
            bool leaf = (node >= n - 1 && n > 1);
            if (leaf || n == 1) {
                int li = (n == 1) ? 0 : node;
                if (excl && (n == 1 || node == self_node)) continue;
                int ret = (int)tv.ret_index(li);
                if (kind == 1) ml.merge_one(d2, ret, k);
                else heap.push(d2, ret, k);
            } else {
                int l = tv.left[node], r = tv.right[node];
                D ld = tv.d2(px, py, pz, l), rd = tv.d2(px, py, pz, r);
                if (sp < 64 - 2) {
                    if (ld <= rd) {
                        stack[sp++] = r;
                        stack[sp++] = l;
                    } else {
                        stack[sp++] = l;
                        stack[sp++] = r;
                    }
                } else {
                    drops++;
                }
            }
        }
        size_t o = (size_t)q * k;
        if (kind == 1) {
            for (int i = 0; i < k; ++i) {
                if (i < cap && ml.idx[i] >= 0) {
                    res[o + i] = (uint64_t)ml.idx[i];
                    dists[o + i] = ml.dist[i];
                } else {
                    res[o + i] = ~0ull;
                    dists[o + i] = sentinel_dist<D>();
                }
            }
        } else {
            heap.sort_ascending();
            for (int i = 0; i < k; ++i) {
                if (i < heap.count && heap.idx[i] >= 0) {
                    res[o + i] = (uint64_t)heap.idx[i];
                    dists[o + i] = heap.dist[i];
                } else {
                    res[o + i] = ~0ull;
                    dists[o + i] = sentinel_dist<D>();
                }
            }
        }
        if (counters) {
            counters[2 * (size_t)q] = visits;
            counters[2 * (size_t)q + 1] = drops;
        }
    }
}

// The reference's grouped self-kNN (src/hlbvh.hpp:1520-1647 with the register
// heap, :1690-1818 with the local-memory heap; the same heap algorithm) as its
// CPU backend runs it: W consecutive query slots in lockstep, one step per
// loop trip.  Each lane pops its own stack and tests its own node; the group
// skips the step only when no active lane still needs its node
// (any_of_group); otherwise every lane in the loop processes its node -- a
// leaf is pushed only by an active lane that did not prune it, but an
// internal node is expanded even by a lane that pruned it.  Lanes past nq
// shadow leaf rank 0 and never count for the group.  Lanes whose stack
// emptied have left the loop.  W = the device's native vector width
// (:1828-1839): 8 for 64-bit and 16 for 32-bit coordinates on a CPU.
template <typename T>
static void reference_knn_grouped_t(const or_tree *t, const int32_t *leaf_ids, int nq, int k,
                                    int excl, int W, uint64_t *res, void *dists_v,
                                    uint64_t *counters) {
    using D = typename DistOf<T>::type;
    TV<T> tv(t);
    D *dists = (D *)dists_v;
    const int n = tv.n;
    const int off = n - 1;
    const int groups = (nq + W - 1) / W;
#pragma omp parallel
    {
        std::vector<MaxHeap<D>> heaps;
        heaps.reserve(W);
        for (int l = 0; l < W; ++l) heaps.emplace_back(std::max(k, 1));
        std::vector<int> stack((size_t)W * 64), sp(W), node(W), self_node(W);
        std::vector<uint8_t> active(W), pruned(W), in_loop(W);
        std::vector<T> px(W), py(W), pz(W);
        std::vector<D> d2(W);
        std::vector<uint64_t> visits(W);
#pragma omp for schedule(dynamic, 4)
        for (int g = 0; g < groups; ++g) {
            for (int l = 0; l < W; ++l) {
                const int q = g * W + l;
                active[l] = q < nq;
                const int r = active[l] ? (leaf_ids ? leaf_ids[q] : q) : 0;
                self_node[l] = r + off;
                px[l] = tv.mn[0][r + off];
                py[l] = tv.mn[1][r + off];
                pz[l] = tv.mn[2][r + off];
                heaps[l].reset();
                sp[l] = 0;
                stack[(size_t)l * 64 + sp[l]++] = 0;
                visits[l] = 0;
            }
            for (;;) {
                bool any_in = false, anyone = false;
                for (int l = 0; l < W; ++l) {
                    in_loop[l] = sp[l] > 0;
                    if (!in_loop[l]) continue;
                    any_in = true;
                    node[l] = stack[(size_t)l * 64 + --sp[l]];
                    visits[l]++;
                    d2[l] = tv.d2(px[l], py[l], pz[l], node[l]);
                    pruned[l] = active[l] && heaps[l].should_prune(d2[l], k);
                    anyone |= active[l] && !pruned[l];
                }
                if (!any_in) break;
                if (!anyone) continue;
                for (int l = 0; l < W; ++l) {
                    if (!in_loop[l]) continue;
                    const int nd = node[l];
                    if (nd >= n - 1 && n > 1) {
                        if (active[l] && !pruned[l] && !(excl && nd == self_node[l]))
                            heaps[l].push(d2[l], (int)tv.ret_index(nd), k);
                    } else if (n == 1) {
                        if (active[l] && !pruned[l] && !excl)
                            heaps[l].push(d2[l], (int)tv.ret_index(0), k);
                    } else {
                        const int lc = tv.left[nd], rc = tv.right[nd];
                        const D ld = tv.d2(px[l], py[l], pz[l], lc);
                        const D rd = tv.d2(px[l], py[l], pz[l], rc);
                        int *st = &stack[(size_t)l * 64];
                        if (sp[l] < 64 - 2) {
                            if (ld <= rd) {
                                st[sp[l]++] = rc;
                                st[sp[l]++] = lc;
                            } else {
                                st[sp[l]++] = lc;
                                st[sp[l]++] = rc;
                            }
                        }
                    }
                }
            }
            for (int l = 0; l < W; ++l) {
                const int q = g * W + l;
                if (!active[l]) continue;
                MaxHeap<D> &h = heaps[l];
                h.sort_ascending();
                const size_t o = (size_t)q * k;
                for (int i = 0; i < k; ++i) {
                    if (i < h.count && h.idx[i] >= 0) {
                        res[o + i] = (uint64_t)h.idx[i];
                        dists[o + i] = h.dist[i];
                    } else {
                        res[o + i] = ~0ull;
                        dists[o + i] = sentinel_dist<D>();
                    }
                }
                if (counters) counters[q] = visits[l];
            }
        }
    }
}

// range_query (src/hlbvh.hpp:2046-2119): DFS pushing right then left, so leaves
// come out in ascending rank; accept rm2 <= d2 <= RM2; store while count < max;
// count = all hits.  Radii: FP (double)r squared; integer (r >> _DIST_SHIFT)^2.
template <typename T>
static void reference_range_t(const or_tree *t, const T *qx, const T *qy, const T *qz,
                              const T *rmin, const T *rmax, int nq, int32_t *res, int32_t *counts,
                              int max_res, int64_t *row_ptr, int32_t *cols) {
    using D = typename DistOf<T>::type;
    TV<T> tv(t);
    const int n = tv.n;
    const int dshift = tv.c.B > 31 ? tv.c.B - 31 : 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int q = 0; q < nq; ++q) {
        T px = qx[q], py = qy[q], pz = qz[q];
        D rm2, RM2;
        if constexpr (is_fp_v<T>) {
            double a = (double)rmin[q], b = (double)rmax[q];
            rm2 = a * a;
            RM2 = b * b;
        } else {
            uint64_t a = (uint64_t)rmin[q] >> dshift, b = (uint64_t)rmax[q] >> dshift;
            rm2 = a * a;
            RM2 = b * b;
        }
        int stack[64];
        int sp = 0;
        stack[sp++] = 0;
        int count = 0;
        int64_t w = row_ptr ? row_ptr[q] : 0;
        while (sp > 0) {
            int node = stack[--sp];
            D d2 = tv.d2(px, py, pz, node);
            if (d2 <= RM2) {
                if ((node >= n - 1 && n > 1) || n == 1) {
                    if (d2 >= rm2) {
                        int li = (n == 1) ? 0 : node;
                        int32_t ret = (int32_t)tv.ret_index(li);
                        if (res && count < max_res) res[(size_t)q * max_res + count] = ret;
                        if (cols) cols[w + count] = ret;
                        count++;
                    }
                } else if (sp < 64 - 2) {
                    stack[sp++] = tv.right[node];
                    stack[sp++] = tv.left[node];
                }
            }
        }
        if (counts) counts[q] = count;
    }
}

// Reference test generators (test/test_self_knn.cpp:271-368, test/main.cpp:84-101)
// with phys_pos_to_int = float_to_int_rep(phys, min, 1/size) for integer coords
// (src/sfc.peano_hilbert.hpp:505-513) and a plain cast for FP coords.
// float_to_int_rep(val, min_val, inv_dx) (src/sfc.peano_hilbert.hpp:495-503):
// nx = sycl::clamp((val - min_val) * inv_dx, 0, 0x1.fffffffffffffp+0), i.e.
// fmin(fmax(., 0), upper), then quantize_coord.
static uint64_t float_to_int_rep(double val, double min_val, double inv_dx, int B) {
    const double upper = 0x1.fffffffffffffp+0;
    double nx = std::fmin(std::fmax((val - min_val) * inv_dx, 0.0), upper);
    return ph_quantise(nx, B);
}

template <typename T> static T phys_to_coord(double phys, double bmin, double bsize, int B) {
    if constexpr (is_fp_v<T>) {
        return (T)phys;
    } else {
        double inv = 1.0 / bsize;
        return (T)float_to_int_rep(phys, bmin, inv, B);
    }
}

// fill_particle_coords (test/benchmark/BenchmarkUtils.hpp:164-205).  Integer
// coordinates: per-axis std::min / std::max over the positions promoted to
// double, box_min = min of the three minima, box_size = (max of the maxima -
// box_min) * 1.0001, or 1.0 when that difference is not positive; every value
// through float_to_int_rep(v, box_min, 1 / box_size).  FP coordinates: a
// static_cast, box (0, 1).
template <typename F, typename T>
static void fill_coords_t(const F *px, const F *py, const F *pz, size_t n, int B, T *x, T *y,
                          T *z, double *box) {
    box[0] = 0.0;
    box[1] = 1.0;
    if constexpr (is_fp_v<T>) {
        for (size_t i = 0; i < n; ++i) {
            x[i] = static_cast<T>(px[i]);
            y[i] = static_cast<T>(py[i]);
            z[i] = static_cast<T>(pz[i]);
        }
    } else {
        if (n == 0) return;
        double min_x = px[0], max_x = px[0];
        double min_y = py[0], max_y = py[0];
        double min_z = pz[0], max_z = pz[0];
        for (size_t i = 1; i < n; ++i) {
            min_x = std::min(min_x, static_cast<double>(px[i]));
            max_x = std::max(max_x, static_cast<double>(px[i]));
            min_y = std::min(min_y, static_cast<double>(py[i]));
            max_y = std::max(max_y, static_cast<double>(py[i]));
            min_z = std::min(min_z, static_cast<double>(pz[i]));
            max_z = std::max(max_z, static_cast<double>(pz[i]));
        }
        const double box_min = std::min({min_x, min_y, min_z});
        const double max_val = std::max({max_x, max_y, max_z});
        const double box_size = (max_val > box_min) ? (max_val - box_min) * 1.0001 : 1.0;
        const double inv_dx = 1.0 / box_size;
        for (size_t i = 0; i < n; ++i) {
            x[i] = (T)float_to_int_rep(px[i], box_min, inv_dx, B);
            y[i] = (T)float_to_int_rep(py[i], box_min, inv_dx, B);
            z[i] = (T)float_to_int_rep(pz[i], box_min, inv_dx, B);
        }
        box[0] = box_min;
        box[1] = box_size;
    }
}

template <typename T>
static void gen_reference_t(int kind, size_t n, uint32_t seed, double bmin, double bsize, int B,
                            T *x, T *y, T *z) {
    std::mt19937 gen(seed);
    auto put = [&](size_t i, double fx, double fy, double fz) {
        x[i] = phys_to_coord<T>(fx, bmin, bsize, B);
        y[i] = phys_to_coord<T>(fy, bmin, bsize, B);
        z[i] = phys_to_coord<T>(fz, bmin, bsize, B);
    };
    if (kind == 0) {  // generate_uniform_particles (:271-289)
        std::uniform_real_distribution<double> dis(bmin, bmin + bsize);
        for (size_t i = 0; i < n; ++i) {
            double fx = dis(gen);
            double fy = dis(gen);
            double fz = dis(gen);
            put(i, fx, fy, fz);
        }
    } else if (kind == 1) {  // generate_clustered_particles (:291-330)
        std::normal_distribution<double> c1(bmin + 0.25 * bsize, 0.04 * bsize);
        std::normal_distribution<double> c2(bmin + 0.75 * bsize, 0.04 * bsize);
        std::uniform_real_distribution<double> u(bmin, bmin + bsize);
        for (size_t i = 0; i < n; ++i) {
            double fx, fy, fz;
            if (i < n * 0.45) {
                fx = c1(gen); fy = c1(gen); fz = c1(gen);
            } else if (i < n * 0.85) {
                fx = c2(gen); fy = c2(gen); fz = c2(gen);
            } else {
                fx = u(gen); fy = u(gen); fz = u(gen);
            }
            fx = std::clamp(fx, bmin + 1e-4, bmin + bsize - 1e-4);
            fy = std::clamp(fy, bmin + 1e-4, bmin + bsize - 1e-4);
            fz = std::clamp(fz, bmin + 1e-4, bmin + bsize - 1e-4);
            put(i, fx, fy, fz);
        }
    } else if (kind == 2) {  // generate_degenerate_particles (:332-368)
        std::uniform_real_distribution<double> dis(bmin + 0.1 * bsize, bmin + 0.9 * bsize);
        std::vector<std::tuple<double, double, double>> centres;
        for (int c = 0; c < 10; ++c) {
            double a = dis(gen);
            double b = dis(gen);
            double cc = dis(gen);
            centres.emplace_back(a, b, cc);
        }
        for (size_t i = 0; i < n; ++i) {
            double fx, fy, fz;
            if (i < n * 0.6) {
                std::tie(fx, fy, fz) = centres[i % centres.size()];
            } else {
                fx = dis(gen); fy = dis(gen); fz = dis(gen);
            }
            put(i, fx, fy, fz);
        }
    } else {  // kind 3: the 10^3 grid of test/main.cpp:84-101 (n ignored, 1000 points)
        const int g = 10;
        for (int ix = 0; ix < g; ++ix)
            for (int iy = 0; iy < g; ++iy)
                for (int iz = 0; iz < g; ++iz) {
                    int idx = ix * g * g + iy * g + iz;
                    put((size_t)idx, bmin + (ix + 0.5) * (bsize / g), bmin + (iy + 0.5) * (bsize / g),
                        bmin + (iz + 0.5) * (bsize / g));
                }
    }
}

template <typename T>
static void build_bvh_t(const Cfg &c, const T *x, const T *y, const T *z, const uint32_t *id,
                        const int8_t *ghost, size_t n, const T *bbox_in, const uint64_t *keys_in,
                        or_tree *out, uint64_t *sorted_keys, int32_t *rfirst, int32_t *rlast) {
    T bb[6];
    if (bbox_in) std::memcpy(bb, bbox_in, sizeof(bb));
    else bbox_t<T>(x, y, z, n, bb);
    std::vector<uint64_t> keys(n);
    if (keys_in) std::copy(keys_in, keys_in + n, keys.begin());
    else encode_t<T>(c.sfc, c.B, x, y, z, n, bb, keys.data());
    std::vector<uint32_t> perm(n);
    sort_keys(keys.data(), n, perm.data(), sorted_keys);
    // Reorder into leaf slots (src/hlbvh.hpp:2266-2279).
    T *mn[3] = {(T *)out->min_x, (T *)out->min_y, (T *)out->min_z};
    T *mx[3] = {(T *)out->max_x, (T *)out->max_y, (T *)out->max_z};
    const T *src[3] = {x, y, z};
    const size_t off = n - 1;
    for (size_t i = 0; i < n; ++i) {
        uint32_t o = perm[i];
        for (int a = 0; a < 3; ++a) mn[a][off + i] = mx[a][off + i] = src[a][o];
        out->id[off + i] = id ? id[o] : o;
        out->ghost[off + i] = ghost ? ghost[o] : 0;
        out->orig[off + i] = (int32_t)o;
    }
    if (n == 1) {
        out->parent[0] = -1;
        return;
    }
    karras(sorted_keys, n, out->left, out->right, out->parent, rfirst, rlast);
    refit_t<T>(n, out->left, out->right, out->parent, mn, mx);
}

#define OR_DISPATCH(ct, ...)                                 \
    switch (ct) {                                            \
    case CT_F32: { using T = float; __VA_ARGS__; } break;    \
    case CT_F64: { using T = double; __VA_ARGS__; } break;   \
    case CT_U32: { using T = uint32_t; __VA_ARGS__; } break; \
    case CT_U64: { using T = uint64_t; __VA_ARGS__; } break; \
    default: return 1;                                       \
    }

}  // namespace

extern "C" {

int or_set_threads(int t) {
    if (t > 0) omp_set_num_threads(t);
    return omp_get_max_threads();
}

int or_bbox(int ct, const void *x, const void *y, const void *z, size_t n, void *out6) {
    if (n == 0) return 1;
    OR_DISPATCH(ct, bbox_t<T>((const T *)x, (const T *)y, (const T *)z, n, (T *)out6));
    return 0;
}

int or_encode(int ct, int sfc, int B, const void *x, const void *y, const void *z, size_t n,
              const void *bbox6, uint64_t *keys) {
    OR_DISPATCH(ct, {
        T bb[6];
        if (bbox6) std::memcpy(bb, bbox6, sizeof(bb));
        else if (n) bbox_t<T>((const T *)x, (const T *)y, (const T *)z, n, bb);
        encode_t<T>(sfc, B, (const T *)x, (const T *)y, (const T *)z, n, bb, keys);
    });
    return 0;
}

int or_sort(const uint64_t *keys, size_t n, uint32_t *perm, uint64_t *sorted) {
    sort_keys(keys, n, perm, sorted);
    return 0;
}

int or_karras(const uint64_t *sorted, size_t n, int32_t *left, int32_t *right, int32_t *parent,
              int32_t *rfirst, int32_t *rlast) {
    if (n < 2) return 1;
    karras(sorted, n, left, right, parent, rfirst, rlast);
    return 0;
}

int or_refit(int ct, size_t n, const int32_t *left, const int32_t *right, const int32_t *parent,
             void *mnx, void *mxx, void *mny, void *mxy, void *mnz, void *mxz) {
    OR_DISPATCH(ct, {
        T *mn[3] = {(T *)mnx, (T *)mny, (T *)mnz};
        T *mx[3] = {(T *)mxx, (T *)mxy, (T *)mxz};
        refit_t<T>(n, left, right, parent, mn, mx);
    });
    return 0;
}

// Full build.  out->... arrays sized 2n-1 (n-1 for left/right) are filled.
int or_build_bvh(or_tree *out, const void *x, const void *y, const void *z, const uint32_t *id,
                 const int8_t *ghost, size_t n, const void *bbox6, const uint64_t *keys,
                 uint64_t *sorted_keys, int32_t *rfirst, int32_t *rlast) {
    if (n == 0) return 0;
    Cfg c{out->sfc, out->bits, out->periodic, out->orig_indices, out->box};
    OR_DISPATCH(out->ct, build_bvh_t<T>(c, (const T *)x, (const T *)y, (const T *)z, id, ghost, n,
                                        (const T *)bbox6, keys, out, sorted_keys, rfirst, rlast));
    return 0;
}

int or_knn_bruteforce(const or_tree *t, const void *qx, const void *qy, const void *qz,
                      const int32_t *leaf_ids, int nq, int k, int exclude_self, uint64_t *res,
                      void *dists) {
    OR_DISPATCH(t->ct, bruteforce_t<T>(t, (const T *)qx, (const T *)qy, (const T *)qz, leaf_ids,
                                       nq, k, exclude_self, res, dists));
    return 0;
}

int or_knn_reference(const or_tree *t, const void *qx, const void *qy, const void *qz,
                     const int32_t *leaf_ids, int nq, int k, int exclude_self, int kind, int max_k,
                     uint64_t *res, void *dists, uint64_t *counters) {
    OR_DISPATCH(t->ct, reference_knn_t<T>(t, (const T *)qx, (const T *)qy, (const T *)qz,
                                          leaf_ids, nq, k, exclude_self, kind, max_k, res, dists,
                                          counters));
    return 0;
}

int or_knn_reference_grouped(const or_tree *t, const int32_t *leaf_ids, int nq, int k,
                             int exclude_self, int width, uint64_t *res, void *dists,
                             uint64_t *visits) {
    if (width < 1) return 1;
    OR_DISPATCH(t->ct, reference_knn_grouped_t<T>(t, leaf_ids, nq, k, exclude_self, width, res,
                                                  dists, visits));
    return 0;
}

int or_range_reference(const or_tree *t, const void *qx, const void *qy, const void *qz,
                       const void *rmin, const void *rmax, int nq, int32_t *res, int32_t *counts,
                       int max_res, int64_t *row_ptr, int32_t *cols) {
    OR_DISPATCH(t->ct, reference_range_t<T>(t, (const T *)qx, (const T *)qy, (const T *)qz,
                                            (const T *)rmin, (const T *)rmax, nq, res, counts,
                                            max_res, row_ptr, cols));
    return 0;
}

// Squared distance between two points with the tree's metric (the brute-force
// metric of test/test_self_knn.cpp:60-63).
int or_point_d2(const or_tree *t, const void *a3, const void *b3, void *out) {
    OR_DISPATCH(t->ct, {
        Cfg c{t->sfc, t->bits, t->periodic, t->orig_indices, t->box};
        const T *a = (const T *)a3, *b = (const T *)b3;
        auto d = node_d2<T>(c, a[0], a[1], a[2], b[0], b[0], b[1], b[1], b[2], b[2]);
        std::memcpy(out, &d, 8);
    });
    return 0;
}

// Heap restatements driven directly (the heap unit tests of
// test/test_self_knn.cpp:381-559).  kind 0: max-heap, kind 1: sorted merge.
// Inputs are (d, idx) pairs pushed in order, with prune checks as the tests do
// for the max-heap; the merge list takes 32-entry batches (`batch` entries per
// row, DIST_MAX/-1 padded by the caller).  Output: the first k entries.
int or_heap_u64(int kind, int k, int cap, const uint64_t *d, const int32_t *idx, int m,
                uint64_t *out_d, int32_t *out_i) {
    if (kind == 0) {
        MaxHeap<uint64_t> h(cap);
        for (int i = 0; i < m; ++i)
            if (!h.should_prune(d[i], k)) h.push(d[i], idx[i], k);
        h.sort_ascending();
        for (int i = 0; i < cap; ++i) {
            out_d[i] = h.dist[i];
            out_i[i] = h.idx[i];
        }
        return h.count;
    }
    MergeList<uint64_t> ml(cap);
    for (int b = 0; b + 32 <= m; b += 32) {
        uint64_t sd[32];
        int si[32];
        for (int l = 0; l < 32; ++l) {
            sd[l] = d[b + l];
            si[l] = idx[b + l];
        }
        // merge_batch with a full 32-lane stage
        MergeList<uint64_t>::bitonic(sd, si, 32);
        for (int l = 0; l < 32; ++l) {
            int dst = ml.count + l;
            if (dst < cap) {
                ml.dist[dst] = sd[l];
                ml.idx[dst] = si[l];
            }
        }
        int nc = std::min(ml.count + 32, cap);
        int p2 = 1;
        while (p2 < nc) p2 <<= 1;
        for (int l = nc; l < p2 && l < cap; ++l) {
            ml.dist[l] = std::numeric_limits<uint64_t>::max();
            ml.idx[l] = -1;
        }
        MergeList<uint64_t>::bitonic(ml.dist.data(), ml.idx.data(), p2);
        ml.count = std::min(std::min(nc, cap), k);
    }
    for (int i = 0; i < cap; ++i) {
        out_d[i] = ml.dist[i];
        out_i[i] = ml.idx[i];
    }
    return ml.count;
}

int or_gen_reference(int kind, size_t n, uint32_t seed, double bmin, double bsize, int ct, int B,
                     void *x, void *y, void *z) {
    OR_DISPATCH(ct, gen_reference_t<T>(kind, n, seed, bmin, bsize, B, (T *)x, (T *)y, (T *)z));
    return 0;
}

int or_fill_particle_coords(int in_type, const void *px, const void *py, const void *pz,
                            size_t n, int ct, int B, void *x, void *y, void *z, double *box) {
    if (in_type == CT_F32) {
        using F = float;
        OR_DISPATCH(ct, fill_coords_t<F, T>((const F *)px, (const F *)py, (const F *)pz, n, B,
                                            (T *)x, (T *)y, (T *)z, box));
    } else {
        using F = double;
        OR_DISPATCH(ct, fill_coords_t<F, T>((const F *)px, (const F *)py, (const F *)pz, n, B,
                                            (T *)x, (T *)y, (T *)z, box));
    }
    return 0;
}

}  // extern "C"
