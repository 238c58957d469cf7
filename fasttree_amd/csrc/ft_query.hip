// ft_query.hip -- gfx950 neighbour-query kernels over a built TreeSoA:
//   self_knn_query / self_knn_query_subset / knn_query (src/hlbvh.hpp:1057-1983)
//   range_query (src/hlbvh.hpp:2006-2136) and a CSR range query.
//
// Two traversal families:
//  * per-lane DFS (one query per lane, 64-entry stack in LDS): the reference's
//    algorithm -- pop, node d2, prune, leaf push or two child d2 with the
//    farther child pushed first while sp < 62 (src/hlbvh.hpp:1545-1628).  Heap
//    tiers: a canonical register top-k (k <= 32) and an LDS max-heap per lane
//    (any k), the latter also in the reference's exact heap mechanics.
//  * packet traversal (one wavefront of 64 SFC-adjacent queries shares one
//    stack; node boxes are wave-uniform loads; small subtrees are scanned as
//    contiguous leaf buckets).  Canonical results only.
//
// Canonical results: the k smallest (d2, returned index) pairs in ascending
// order -- a unique answer, independent of traversal order.
#include "ft_internal.hpp"

#include <algorithm>
#include <vector>

namespace ft {

int scratch_ensure_public(Scratch &S, size_t n);

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

// Returned index of a leaf: its rank, or orig_idx with RETURN_ORIG_INDICES.
template <typename T> __device__ inline int32_t ret_index(const DTree<T> &t, int leaf_node) {
    return t.orig_indices ? t.orig[leaf_node] : (leaf_node - (t.n - 1));
}

// ---------------------------------------------------------------------------
// Canonical register top-k: ascending (d, idx) in M fixed registers, the live
// k entries right-aligned so the worst is always slot M-1.  Slots below M-k
// hold (-inf, -1) and never move.  Insertion is a fixed compare/shift network.
// ---------------------------------------------------------------------------
template <typename D, int M> struct TopK {
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
    // Predicated insertion as a compare-and-swap chain from the smallest slot:
    // the carried entry swaps into every slot it is below, so it lands in order
    // and the former worst drops off the top.  Every slot stays in its own
    // registers; the comparisons are branch-free.
    __device__ inline void insert(bool take, D cd, int32_t ci) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const bool lt = take & ((cd < d[s]) | ((cd == d[s]) & (ci < i[s])));
            const D td = d[s];
            const int32_t ti = i[s];
            d[s] = lt ? cd : td;
            i[s] = lt ? ci : ti;
            cd = lt ? td : cd;
            ci = lt ? ti : ci;
        }
    }
    __device__ inline D worst_d() const { return d[M - 1]; }
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

// ===========================================================================
// Per-lane DFS with the canonical register top-k (k <= M <= 32).
// GROUPED: the wave skips a node unless some active lane still needs it
// (any_of_group, src/hlbvh.hpp:1563-1566).  Results do not depend on it.
// ===========================================================================
template <typename T, int M, bool GROUPED>
__global__ __launch_bounds__(LANE_BLOCK) void lane_knn_reg_kernel(DTree<T> t, QSrc<T> qs, int nq,
                                                                  int k, int excl,
                                                                  uint64_t *__restrict__ res,
                                                                  typename CoordTraits<T>::D *__restrict__ dists,
                                                                  uint64_t *__restrict__ stats) {
    using D = typename CoordTraits<T>::D;
    __shared__ int32_t stk[STACK_DEPTH][LANE_BLOCK];
    const int lane = threadIdx.x;
    const int q = blockIdx.x * LANE_BLOCK + lane;
    const bool active = q < nq;
    const int n = t.n;
    T px = T(0), py = T(0), pz = T(0);
    int self = -1;
    if (active) load_query(t, qs, q, px, py, pz, self);
    TopK<D, M> h;
    h.init(k, dist_max<D>());
    uint64_t visits = 0, drops = 0, inserts = 0;
    int sp = 0;
    if (active) stk[sp++][lane] = 0;
    while (sp > 0) {
        const int node = stk[--sp][lane];
        const D d2 = node_d2(t, px, py, pz, node);
        ++visits;
        const bool pruned = d2 > h.worst();
        if (GROUPED) {
            if (!__any(!pruned)) continue;
        }
        if (pruned) continue;
        if (node >= n - 1 || n == 1) {
            const int leaf = (n == 1) ? 0 : node;
            const int rank = leaf - (n - 1);
            if (excl && rank == self) continue;
            const int32_t ri = ret_index(t, leaf);
            const bool take = h.beats_worst(d2, ri);
            h.insert(take, d2, ri);
            inserts += take ? 1 : 0;
        } else {
            const int l = t.left[node], r = t.right[node];
            const D ld = node_d2(t, px, py, pz, l);
            const D rd = node_d2(t, px, py, pz, r);
            if (sp < STACK_DEPTH - 2) {
                if (ld <= rd) {
                    stk[sp++][lane] = r;
                    stk[sp++][lane] = l;
                } else {
                    stk[sp++][lane] = l;
                    stk[sp++][lane] = r;
                }
            } else {
                ++drops;
            }
        }
    }
    if (active) h.store(k, res, dists, (size_t)q);
    wave_add_stats(stats, 0, visits);
    wave_add_stats(stats, 2, drops);
    wave_add_stats(stats, 4, inserts);
}

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

template <typename T, bool REF, bool GROUPED>
__global__ __launch_bounds__(LANE_BLOCK) void lane_knn_lds_kernel(DTree<T> t, QSrc<T> qs, int nq,
                                                                  int k, int excl, int lanes,
                                                                  uint64_t *__restrict__ res,
                                                                  typename CoordTraits<T>::D *__restrict__ dists,
                                                                  uint64_t *__restrict__ stats) {
    using D = typename CoordTraits<T>::D;
    extern __shared__ __align__(16) unsigned char lds[];
    D *hd = (D *)lds;                                   // [k][lanes]
    int32_t *hi = (int32_t *)(hd + (size_t)k * lanes);  // [k][lanes]
    int32_t *stk = hi + (size_t)k * lanes;              // [64][lanes]
    const int lane = threadIdx.x;
    const bool lane_ok = lane < lanes;
    const int q = blockIdx.x * lanes + lane;
    const bool active = lane_ok && q < nq;
    const int n = t.n;
    T px = T(0), py = T(0), pz = T(0);
    int self = -1;
    if (active) load_query(t, qs, q, px, py, pz, self);
    LdsHeap<D, REF> h{hd, hi, lanes, lane, 0};
    uint64_t visits = 0, drops = 0, inserts = 0;
    int sp = 0;
    if (active) stk[(sp++) * lanes + lane] = 0;
    while (sp > 0) {
        const int node = stk[(--sp) * lanes + lane];
        const D d2 = node_d2(t, px, py, pz, node);
        ++visits;
        const bool pruned = h.prune(d2, k);
        if (GROUPED) {
            if (!__any(!pruned)) continue;
        }
        if (pruned) continue;
        if (node >= n - 1 || n == 1) {
            const int leaf = (n == 1) ? 0 : node;
            const int rank = leaf - (n - 1);
            if (excl && rank == self) continue;
            if (h.push(d2, ret_index(t, leaf), k)) ++inserts;
        } else {
            const int l = t.left[node], r = t.right[node];
            const D ld = node_d2(t, px, py, pz, l);
            const D rd = node_d2(t, px, py, pz, r);
            if (sp < STACK_DEPTH - 2) {
                if (ld <= rd) {
                    stk[(sp++) * lanes + lane] = r;
                    stk[(sp++) * lanes + lane] = l;
                } else {
                    stk[(sp++) * lanes + lane] = l;
                    stk[(sp++) * lanes + lane] = r;
                }
            } else {
                ++drops;
            }
        }
    }
    if (active) {
        h.sort();
        const size_t o = (size_t)q * k;
        for (int s = 0; s < k; ++s) {
            const bool ok = s < h.count && h.ii(s) >= 0;
            res[o + s] = ok ? (uint64_t)(uint32_t)h.ii(s) : ~0ull;
            dists[o + s] = ok ? h.dd(s) : missing_dist<D>();
        }
    }
    wave_add_stats(stats, 0, visits);
    wave_add_stats(stats, 2, drops);
    wave_add_stats(stats, 4, inserts);
}

// ===========================================================================
// knn_query k > 32 in the reference's mechanics: one workgroup per query,
// lane 0 walks the stack, every leaf is merged through a 32-entry stage and a
// full bitonic network over the sorted list (SortedMergeHeap and
// local_bitonic_sort, src/maxheap.hpp:193-324; knn_query_large_k,
// src/hlbvh.hpp:1232-1387).  The network is data-oblivious, so running its
// compare-exchange stages across 64 lanes gives the reference's tie order.
// ===========================================================================
template <typename D>
__device__ inline void lds_bitonic(D *d, int32_t *ix, int n) {
    for (int kk = 2; kk <= n; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const int l = i ^ j;
                const bool asc = (i & kk) == 0;
                if (l > i && l < n) {
                    if ((d[i] > d[l]) == asc) {
                        D td = d[i]; d[i] = d[l]; d[l] = td;
                        int32_t ti = ix[i]; ix[i] = ix[l]; ix[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
}

template <typename T>
__global__ __launch_bounds__(64) void merge_knn_kernel(DTree<T> t, QSrc<T> qs, int nq, int k,
                                                       int cap, uint64_t *__restrict__ res,
                                                       typename CoordTraits<T>::D *__restrict__ dists,
                                                       uint64_t *__restrict__ stats) {
    using D = typename CoordTraits<T>::D;
    extern __shared__ __align__(16) unsigned char lds[];
    D *sd = (D *)lds;                                  // [cap]
    int32_t *si = (int32_t *)(sd + cap);               // [cap]
    D *gd = (D *)(si + cap + (cap & 1));               // [32] stage
    int32_t *gi = (int32_t *)(gd + 32);                // [32]
    int32_t *stk = gi + 32;                            // [64]
    __shared__ int s_count, s_sp, s_node;
    const int q = blockIdx.x;
    const int n = t.n;
    T px, py, pz;
    int self;
    load_query(t, qs, q, px, py, pz, self);
    for (int i = threadIdx.x; i < cap; i += blockDim.x) {
        sd[i] = dist_max<D>();
        si[i] = -1;
    }
    if (threadIdx.x == 0) {
        s_count = 0;
        s_sp = 1;
        stk[0] = 0;
    }
    __syncthreads();
    uint64_t visits = 0, drops = 0;
    while (true) {
        if (threadIdx.x == 0) s_node = s_sp > 0 ? stk[--s_sp] : -1;
        __syncthreads();
        const int node = s_node;
        __syncthreads();
        if (node < 0) break;
        ++visits;
        const D d2 = node_d2(t, px, py, pz, node);
        if (s_count >= k && d2 >= sd[k - 1]) continue;  // should_prune
        if (node >= n - 1 || n == 1) {
            // merge_batch: lane 0's candidate plus 31 empty entries
            if (threadIdx.x < 32) {
                const bool mine = threadIdx.x == 0;
                gd[threadIdx.x] = mine ? d2 : dist_max<D>();
                gi[threadIdx.x] = mine ? ((n == 1) ? 0 : ret_index(t, node)) : -1;
            }
            __syncthreads();
            lds_bitonic(gd, gi, 32);
            const int cnt = s_count;
            if (threadIdx.x < 32) {
                const int dst = cnt + threadIdx.x;
                if (dst < cap) {
                    sd[dst] = gd[threadIdx.x];
                    si[dst] = gi[threadIdx.x];
                }
            }
            __syncthreads();
            const int nc = min(cnt + 32, cap);
            int p2 = 1;
            while (p2 < nc) p2 <<= 1;
            for (int i = nc + threadIdx.x; i < p2 && i < cap; i += blockDim.x) {
                sd[i] = dist_max<D>();
                si[i] = -1;
            }
            __syncthreads();
            lds_bitonic(sd, si, p2);
            if (threadIdx.x == 0) s_count = min(min(nc, cap), k);
            __syncthreads();
        } else {
            if (threadIdx.x == 0) {
                const int l = t.left[node], r = t.right[node];
                const D ld = node_d2(t, px, py, pz, l);
                const D rd = node_d2(t, px, py, pz, r);
                if (s_sp < STACK_DEPTH - 2) {
                    if (ld <= rd) {
                        stk[s_sp++] = r;
                        stk[s_sp++] = l;
                    } else {
                        stk[s_sp++] = l;
                        stk[s_sp++] = r;
                    }
                } else {
                    ++drops;
                }
            }
            __syncthreads();
        }
    }
    const size_t o = (size_t)q * k;
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        const bool ok = si[i] >= 0;
        res[o + i] = ok ? (uint64_t)(uint32_t)si[i] : ~0ull;
        dists[o + i] = ok ? sd[i] : missing_dist<D>();
    }
    if (stats && threadIdx.x == 0) {
        atomicAdd((unsigned long long *)&stats[0], (unsigned long long)visits);
        if (drops) atomicAdd((unsigned long long *)&stats[2], (unsigned long long)drops);
    }
}

// ---------------------------------------------------------------------------
// Packed top-k for integer distances: key = min(d2, SAT) << ib | index, one
// 64-bit compare orders by (d2, index).  ib = bit length of n, so every index
// is below 2^ib - 1 and an all-ones index marks a sentinel.  A key whose d2
// part saturated cannot be ordered exactly; the traversal reports such queries
// (and queries left with sentinels) for an exact re-run.
// ---------------------------------------------------------------------------
template <int M> struct TopKPacked {
    uint64_t key[M];
    int ib;
    uint64_t sat, imask;
    __device__ inline void init(int k, uint64_t bound, int ibits) {
        ib = ibits;
        sat = (1ull << (64 - ib)) - 1;
        imask = (1ull << ib) - 1;
        const uint64_t sent = ((bound < sat ? bound : sat) << ib) | imask;
#pragma unroll
        for (int s = 0; s < M; ++s) key[s] = (s >= M - k) ? sent : 0ull;
    }
    __device__ inline uint64_t make(uint64_t d, int32_t i) const {
        return ((d < sat ? d : sat) << ib) | (uint64_t)(uint32_t)i;
    }
    __device__ inline uint64_t worst_d() const { return key[M - 1] >> ib; }
    __device__ inline bool beats_worst(uint64_t d, int32_t i) const { return make(d, i) < key[M - 1]; }
    __device__ inline void insert(bool take, uint64_t d, int32_t i) {
        uint64_t c = make(d, i);
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const bool lt = take & (c < key[s]);
            const uint64_t t = key[s];
            key[s] = lt ? c : t;
            c = lt ? t : c;
        }
    }
    // Sentinels left in live slots, or a saturated k-th distance.
    __device__ inline bool needs_exact(int k) const {
        bool bad = (key[M - 1] >> ib) >= sat;
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
    template <typename DOut>
    __device__ inline void store(int k, uint64_t *res, DOut *dists, size_t q) const {
        const size_t o = q * (size_t)k;
#pragma unroll
        for (int s = 0; s < M; ++s) {
            if (s >= M - k) {
                const size_t j = o + (size_t)(s - (M - k));
                const bool ok = (key[s] & imask) != imask;
                res[j] = ok ? (key[s] & imask) : ~0ull;
                dists[j] = ok ? (DOut)(key[s] >> ib) : missing_dist<DOut>();
            }
        }
    }
};

template <typename D, int M, bool PACKED> struct TopKSel { using type = TopK<D, M>; };
template <int M> struct TopKSel<uint64_t, M, true> { using type = TopKPacked<M>; };

// ===========================================================================
// Packet traversal: a wavefront of 64 queries shares one DFS.  A node is
// expanded when some lane's box distance does not exceed that lane's current
// k-th distance; leaf children are tested at once; a subtree with at most
// `bucket` leaves is scanned as a contiguous run of leaf slots.  Node data is
// wave-uniform, the stack lives in LDS, results are canonical.
//
// Query mapping: direct (lane -> query blockIdx*64+lane) or LISTED (each wave
// takes 64 entries of a device list of query positions; the list length is
// read on the device, so no host round trip is needed).
// SEEDED: each lane's k-th distance starts at its seed bound; a lane that ends
// with fewer than k neighbours inside the seed is appended to the re-run list.
// PACKED: 64-bit (d2, index) keys; queries the packed form cannot order
// exactly are appended to the re-run list too.  The re-run is LISTED,
// unseeded and lexicographic, so its results are exact.
// ===========================================================================
constexpr int PK_STACK = 128;

struct PacketArgs {
    int nq, k, excl, bucket, ibits, pad;
    const int32_t *list;        // LISTED: query positions
    const uint32_t *list_count; // LISTED: number of entries
    int32_t *rerun;             // re-run list (SEEDED or PACKED)
    uint32_t *rerun_count;
    uint64_t *stats;
};

template <typename T, int M, bool SEEDED, bool PACKED, bool LISTED>
__global__ __launch_bounds__(64) void packet_knn_kernel(
    DTree<T> t, QSrc<T> qs, PacketArgs a, const typename CoordTraits<T>::D *__restrict__ seed,
    uint64_t *__restrict__ res, typename CoordTraits<T>::D *__restrict__ dists) {
    using D = typename CoordTraits<T>::D;
    using H = typename TopKSel<D, M, PACKED>::type;
    __shared__ int32_t stk[PK_STACK];
    const int lane = threadIdx.x;
    const int n = t.n;
    const int off = n - 1;
    const T *lx = t.mnx + off;
    const T *ly = t.mny + off;
    const T *lz = t.mnz + off;
    const int k = a.k;
    const int nlist = LISTED ? (int)*a.list_count : a.nq;
    uint64_t nodes = 0, leaves = 0, inserts = 0, overflow = 0;

    for (int base = blockIdx.x * 64; base < nlist; base += (LISTED ? gridDim.x * 64 : nlist)) {
        const int slot = base + lane;
        const bool active = slot < nlist;
        const int q = active ? (LISTED ? a.list[slot] : slot) : -1;
        T px = T(0), py = T(0), pz = T(0);
        int self = -1;
        if (active) load_query(t, qs, q, px, py, pz, self);
        if (!a.excl) self = -1;
        H h;
        if constexpr (PACKED) h.init(k, (SEEDED && active) ? (uint64_t)seed[q] : ~0ull, a.ibits);
        else h.init(k, (SEEDED && active) ? seed[q] : dist_max<D>());

        // Scans leaf ranks [first, first+cnt) (cnt <= 64).  One coalesced load
        // puts leaf first+j in lane j.  Phase 1 broadcasts each leaf and marks,
        // per lane, the leaves not farther than its current k-th distance;
        // phase 2 inserts the marked leaves in rounds (each lane fetches its own
        // next candidate from the holding lane), so the insertion chain runs
        // once per round instead of once per leaf.
        auto scan_range = [&](int first, int cnt) {
            T bx = T(0), by = T(0), bz = T(0);
            int32_t bo = 0;
            if (lane < cnt) {
                bx = lx[first + lane];
                by = ly[first + lane];
                bz = lz[first + lane];
                bo = t.orig_indices ? t.orig[off + first + lane] : first + lane;
            }
            const D w = (D)h.worst_d();
            uint64_t cand = 0;
            for (int j = 0; j < cnt; ++j) {
                const D d = leaf_d2<T>(t.m, px, py, pz, rl(bx, j), rl(by, j), rl(bz, j));
                cand |= (uint64_t)(d <= w) << j;
            }
            if (!active) cand = 0;
            const int sr = self - first;
            if (sr >= 0 && sr < cnt) cand &= ~(1ull << sr);
            leaves += (uint64_t)cnt;
            while (__any(cand != 0)) {
                const bool has = cand != 0;
                const int j = has ? (__ffsll((long long)cand) - 1) : 0;
                cand &= cand - 1;
                const T x = shfl_any(bx, j), y = shfl_any(by, j), z = shfl_any(bz, j);
                const int32_t ri = shfl_any(bo, j);
                const D d = leaf_d2<T>(t.m, px, py, pz, x, y, z);
                const bool take = has && h.beats_worst(d, ri);
                h.insert(take, d, ri);
                inserts += take ? 1 : 0;
            }
        };

        int sp = 0;
        int node = 0;  // an internal node some lane needs (or the single leaf when n == 1)
        while (true) {
            ++nodes;
            int scan_first = 0, scan_cnt = 0, next = -1;
            if (n == 1) {
                scan_cnt = 1;
            } else {
                const int first = uni(t.rfirst[node]);
                const int last = uni(t.rlast[node]);
                if (last - first + 1 <= a.bucket) {
                    scan_first = first;
                    scan_cnt = last - first + 1;
                } else {
                    const int l = uni(t.left[node]);
                    const int r = uni(t.right[node]);
                    const D ld = node_d2(t, px, py, pz, l);
                    const D rd = node_d2(t, px, py, pz, r);
                    const D w = (D)h.worst_d();
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
                        else ++overflow;
                    } else if (gl) {
                        next = l;
                    } else if (gr) {
                        next = r;
                    }
                }
            }
            if (scan_cnt > 0) scan_range(scan_first, scan_cnt);
            while (next < 0 && sp > 0) {
                const int c = uni(stk[--sp]);
                const D dc = node_d2(t, px, py, pz, c);
                if (__any(active && dc <= (D)h.worst_d())) next = c;
            }
            if (next < 0) break;
            node = next;
        }
        if (active) {
            bool redo = false;
            if constexpr (PACKED) redo = h.needs_exact(k);
            else if constexpr (SEEDED) redo = h.missing(k) > 0;
            if (redo) {
                const uint32_t s2 = atomicAdd(a.rerun_count, 1u);
                a.rerun[s2] = q;
            } else {
                h.store(k, res, dists, (size_t)q);
            }
        }
        if (!LISTED) break;
    }
    if (lane == 0 && a.stats) {
        atomicAdd((unsigned long long *)&a.stats[0], (unsigned long long)nodes);
        atomicAdd((unsigned long long *)&a.stats[1], (unsigned long long)leaves);
        if (overflow) atomicAdd((unsigned long long *)&a.stats[2], (unsigned long long)overflow);
    }
    wave_add_stats(a.stats, 4, inserts);
}

// ===========================================================================
// range_query (src/hlbvh.hpp:2046-2119): per-lane DFS, right child pushed
// before left so leaves arrive in ascending rank; accept rm2 <= d2 <= RM2;
// store while count < max; count = all hits.  CSR mode writes at row offsets.
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(LANE_BLOCK) void range_kernel(DTree<T> t, const T *__restrict__ qx,
                                                           const T *__restrict__ qy,
                                                           const T *__restrict__ qz,
                                                           const T *__restrict__ rmin,
                                                           const T *__restrict__ rmax, int nq,
                                                           int32_t *__restrict__ res,
                                                           int32_t *__restrict__ counts,
                                                           int max_res,
                                                           const int64_t *__restrict__ row_ptr,
                                                           int32_t *__restrict__ cols,
                                                           int64_t col_cap) {
    using D = typename CoordTraits<T>::D;
    __shared__ int32_t stk[STACK_DEPTH][LANE_BLOCK];
    const int lane = threadIdx.x;
    const int q = blockIdx.x * LANE_BLOCK + lane;
    if (q >= nq) return;
    const int n = t.n;
    const T px = qx[q], py = qy[q], pz = qz[q];
    D rm2, RM2;
    if constexpr (CoordTraits<T>::fp) {
        const double a = (double)rmin[q], b = (double)rmax[q];
        rm2 = a * a;
        RM2 = b * b;
    } else {
        const uint64_t a = (uint64_t)rmin[q] >> t.m.dshift, b = (uint64_t)rmax[q] >> t.m.dshift;
        rm2 = a * a;
        RM2 = b * b;
    }
    int sp = 0;
    stk[sp++][lane] = 0;
    int count = 0;
    const int64_t w0 = row_ptr ? row_ptr[q] : 0;
    while (sp > 0) {
        const int node = stk[--sp][lane];
        const D d2 = node_d2(t, px, py, pz, node);
        if (d2 <= RM2) {
            if (node >= n - 1 || n == 1) {
                if (d2 >= rm2) {
                    const int leaf = (n == 1) ? 0 : node;
                    const int32_t ri = ret_index(t, leaf);
                    if (res && count < max_res) res[(size_t)q * max_res + count] = ri;
                    if (cols && w0 + count < col_cap) cols[w0 + count] = ri;
                    count++;
                }
            } else if (sp < STACK_DEPTH - 2) {
                stk[sp++][lane] = t.right[node];
                stk[sp++][lane] = t.left[node];
            }
        }
    }
    if (counts) counts[q] = count;
}

// Exclusive scan of int32 counts into int64 offsets (Q+1 entries), one block
// per 1024 counts, then block totals, then a fix-up pass.
__global__ __launch_bounds__(256) void scan_block_kernel(const int32_t *__restrict__ c, int nq,
                                                         int64_t *__restrict__ out,
                                                         int64_t *__restrict__ block_sums) {
    __shared__ int64_t wsum[4];
    const int base = blockIdx.x * 1024;
    int64_t v[4];
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = base + threadIdx.x * 4 + j;
        v[j] = (i < nq) ? (int64_t)c[i] : 0;
        s += v[j];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int64_t add = 0;
    for (int q = 0; q < w; ++q) add += wsum[q];
    int64_t run = x - s + add;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = base + threadIdx.x * 4 + j;
        if (i < nq) out[i] = run;
        run += v[j];
    }
    if (threadIdx.x == 255) block_sums[blockIdx.x] = run;
}

__global__ __launch_bounds__(64) void scan_totals_kernel(int64_t *block_sums, int nb,
                                                         int64_t *out, int nq) {
    if (threadIdx.x != 0) return;
    int64_t run = 0;
    for (int b = 0; b < nb; ++b) {
        const int64_t v = block_sums[b];
        block_sums[b] = run;
        run += v;
    }
    out[nq] = run;
}

__global__ __launch_bounds__(256) void scan_add_kernel(int64_t *out, int nq,
                                                       const int64_t *block_sums) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) out[i] += block_sums[i / 1024];
}

// ===========================================================================
// Host dispatch
// ===========================================================================
static inline unsigned blocks_for(size_t n, int per) {
    size_t g = (n + per - 1) / per;
    return (unsigned)(g ? g : 1);
}

template <typename T, bool GROUPED>
static int launch_lane_reg(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq, int k,
                           int excl, uint64_t *res, typename CoordTraits<T>::D *d,
                           uint64_t *stats) {
    const dim3 g(blocks_for(nq, LANE_BLOCK)), b(LANE_BLOCK);
#define FT_LANE_REG(MK) hipLaunchKernelGGL((lane_knn_reg_kernel<T, MK, GROUPED>), g, b, 0, s, dt, qs, nq, k, excl, res, d, stats)
    if (k <= 1) FT_LANE_REG(1);
    else if (k <= 4) FT_LANE_REG(4);
    else if (k <= 8) FT_LANE_REG(8);
    else if (k <= 16) FT_LANE_REG(16);
    else FT_LANE_REG(32);
#undef FT_LANE_REG
    FT_CHECK_LAUNCH();
    return FT_OK;
}

template <typename T, bool REF, bool GROUPED>
static int launch_lane_lds(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq, int k,
                           int excl, uint64_t *res, typename CoordTraits<T>::D *d,
                           uint64_t *stats) {
    using D = typename CoordTraits<T>::D;
    const size_t per_lane = (size_t)k * (sizeof(D) + 4) + STACK_DEPTH * 4;
    int lanes = 64;
    const size_t budget = 64 * 1024;
    while (lanes > 1 && (size_t)lanes * per_lane > budget) lanes >>= 1;
    if ((size_t)lanes * per_lane > budget)
        return fail(FT_ELDS, "k too large for local memory: reduce k");
    const size_t lds = (size_t)lanes * per_lane;
    hipLaunchKernelGGL((lane_knn_lds_kernel<T, REF, GROUPED>), dim3(blocks_for(nq, lanes)),
                       dim3(LANE_BLOCK), lds, s, dt, qs, nq, k, excl, lanes, res, d, stats);
    FT_CHECK_LAUNCH();
    return FT_OK;
}

// Device buffers for the two re-run lists, grown on demand (one per process).
struct RerunBuf {
    int32_t *list[2] = {nullptr, nullptr};
    uint32_t *count = nullptr;  // [0], [1]: list lengths
    size_t cap = 0;
    void *block = nullptr;
};
static RerunBuf g_rerun;

static int rerun_ensure(size_t n) {
    if (n <= g_rerun.cap && g_rerun.block) return FT_OK;
    if (g_rerun.block) (void)hipFree(g_rerun.block);
    g_rerun.block = nullptr;
    const size_t cap = n + n / 2 + 1024;
    void *p = nullptr;
    if (hipMalloc(&p, 2 * cap * 4 + 256) != hipSuccess)
        return fail(FT_ENOMEM, "re-run list allocation failed");
    g_rerun.block = p;
    g_rerun.list[0] = (int32_t *)p;
    g_rerun.list[1] = (int32_t *)p + cap;
    g_rerun.count = (uint32_t *)((char *)p + 2 * cap * 4);
    g_rerun.cap = cap;
    return FT_OK;
}

__global__ void zero_counts_kernel(uint32_t *p) {
    p[0] = 0;
    p[1] = 0;
}
__global__ void add_count_kernel(uint64_t *stats, const uint32_t *count) { stats[3] += *count; }

static inline int bit_length(uint64_t v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

template <typename T, int M, bool SEEDED, bool PACKED, bool LISTED>
static void launch_pk(hipStream_t s, dim3 g, const DTree<T> &dt, const QSrc<T> &qs,
                      const PacketArgs &a, const typename CoordTraits<T>::D *seed, uint64_t *res,
                      typename CoordTraits<T>::D *d) {
    hipLaunchKernelGGL((packet_knn_kernel<T, M, SEEDED, PACKED, LISTED>), g, dim3(64), 0, s, dt,
                       qs, a, seed, res, d);
}

// Runs the packet traversal and its exact re-runs:
//   pass 1  direct, seeded or not, packed keys for integer metrics;
//   pass 2  (packed only) unseeded packed run of the queries pass 1 listed;
//   pass 3  unseeded lexicographic run of what is still listed -- exact.
// The LISTED passes read list lengths on the device: no host round trip.
template <typename T, int M>
static int packet_run(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq, int k,
                      int excl, int bucket, const typename CoordTraits<T>::D *seed, uint64_t *res,
                      typename CoordTraits<T>::D *d, uint64_t *stats) {
    constexpr bool INTD = !CoordTraits<T>::fp;
    PacketArgs a{};
    a.nq = nq;
    a.k = k;
    a.excl = excl;
    a.bucket = bucket;
    a.ibits = bit_length((uint64_t)dt.n);
    a.stats = stats;
    const bool need_list = seed != nullptr || INTD;
    if (need_list) {
        FT_TRY(rerun_ensure((size_t)nq));
        hipLaunchKernelGGL(zero_counts_kernel, dim3(1), dim3(1), 0, s, g_rerun.count);
        a.rerun = g_rerun.list[0];
        a.rerun_count = g_rerun.count;
    }
    const dim3 g(blocks_for(nq, 64));
    if (seed) {
        if constexpr (INTD) launch_pk<T, M, true, true, false>(s, g, dt, qs, a, seed, res, d);
        else launch_pk<T, M, true, false, false>(s, g, dt, qs, a, seed, res, d);
    } else {
        if constexpr (INTD) launch_pk<T, M, false, true, false>(s, g, dt, qs, a, seed, res, d);
        else launch_pk<T, M, false, false, false>(s, g, dt, qs, a, seed, res, d);
    }
    FT_CHECK_LAUNCH();
    if (!need_list) return FT_OK;
    if (stats) {
        hipLaunchKernelGGL(add_count_kernel, dim3(1), dim3(1), 0, s, stats,
                           (const uint32_t *)g_rerun.count);
        FT_CHECK_LAUNCH();
    }
    const unsigned waves = std::min<unsigned>(blocks_for(nq, 64), 2048u);
    PacketArgs b = a;
    b.stats = nullptr;
    int lex_list = 0;
    if constexpr (INTD) {
        if (seed) {
            b.list = g_rerun.list[0];
            b.list_count = g_rerun.count;
            b.rerun = g_rerun.list[1];
            b.rerun_count = g_rerun.count + 1;
            launch_pk<T, M, false, true, true>(s, dim3(waves), dt, qs, b, nullptr, res, d);
            FT_CHECK_LAUNCH();
            lex_list = 1;
        }
    }
    b.list = g_rerun.list[lex_list];
    b.list_count = g_rerun.count + lex_list;
    b.rerun = nullptr;
    b.rerun_count = nullptr;
    launch_pk<T, M, false, false, true>(s, dim3(waves), dt, qs, b, nullptr, res, d);
    FT_CHECK_LAUNCH();
    return FT_OK;
}

template <typename T>
static int launch_packet(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq, int k,
                         int excl, int bucket, const typename CoordTraits<T>::D *seed,
                         uint64_t *res, typename CoordTraits<T>::D *d, uint64_t *stats) {
    if (k <= 1) return packet_run<T, 1>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
    if (k <= 4) return packet_run<T, 4>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
    if (k <= 8) return packet_run<T, 8>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
    if (k <= 16) return packet_run<T, 16>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
    return packet_run<T, 32>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
}

template <typename T>
static int knn_dispatch(hipStream_t s, const ft_tree *tree, const QSrc<T> &qs, int nq, int k,
                        int excl, uint64_t *res, void *dists_v, const ft_query_opts *opts,
                        bool external) {
    using D = typename CoordTraits<T>::D;
    D *dists = (D *)dists_v;
    const DTree<T> dt = make_dtree<T>(*tree);
    ft_query_opts o{};
    if (opts) o = *opts;
    uint64_t *stats = o.stats;
    int algo = o.algorithm;
    if (o.mode == FT_MODE_REFERENCE) {
        if (o.seed_bound) return fail(FT_EINVAL, "a seed bound needs canonical mode");
        if (external && k > 32) {
            // SortedMergeHeap tier, capacity next_pow2(max_k)
            int maxk = o.max_k > 0 ? o.max_k : 256;
            int cap = 1;
            while (cap < maxk) cap <<= 1;
            const size_t lds = (size_t)cap * (sizeof(D) + 4) + 8 + 32 * (sizeof(D) + 4) + STACK_DEPTH * 4 + 64;
            if (lds > 64 * 1024)
                return fail(FT_ELDS, "knn_query_large_k: k too large, local memory requirement exceeds device limit");
            hipLaunchKernelGGL((merge_knn_kernel<T>), dim3(nq), dim3(64), lds, s, dt, qs, nq, k,
                               cap, res, dists, stats);
            FT_CHECK_LAUNCH();
            return FT_OK;
        }
        if (algo == FT_ALGO_PLAIN)
            return launch_lane_lds<T, true, false>(s, dt, qs, nq, k, excl, res, dists, stats);
        return launch_lane_lds<T, true, true>(s, dt, qs, nq, k, excl, res, dists, stats);
    }
    // Canonical mode.
    if (algo == FT_ALGO_AUTO) algo = (k <= 32) ? FT_ALGO_PACKET : FT_ALGO_GROUPED;
    if (o.seed_bound) algo = FT_ALGO_PACKET;
    if (algo == FT_ALGO_PACKET && k <= 32) {
        const int bucket = o.leaf_bucket > 0 ? std::min(o.leaf_bucket, 64) : 32;
        return launch_packet<T>(s, dt, qs, nq, k, excl, bucket, (const D *)o.seed_bound, res,
                                dists, stats);
    }
    if (k <= 32) {
        if (algo == FT_ALGO_PLAIN)
            return launch_lane_reg<T, false>(s, dt, qs, nq, k, excl, res, dists, stats);
        return launch_lane_reg<T, true>(s, dt, qs, nq, k, excl, res, dists, stats);
    }
    if (algo == FT_ALGO_PLAIN)
        return launch_lane_lds<T, false, false>(s, dt, qs, nq, k, excl, res, dists, stats);
    return launch_lane_lds<T, false, true>(s, dt, qs, nq, k, excl, res, dists, stats);
}

// Entry points used by the C ABI layer (device pointers only).
int knn_entry(hipStream_t s, const ft_tree *tree, const void *qx, const void *qy, const void *qz,
              const int32_t *leaf_ids, int nq, int k, int excl, uint64_t *res, void *dists,
              const ft_query_opts *opts) {
    FT_DISPATCH_COORD(tree->coord_type, {
        QSrc<T> qs{(const T *)qx, (const T *)qy, (const T *)qz, leaf_ids};
        return knn_dispatch<T>(s, tree, qs, nq, k, excl, res, dists, opts, qx != nullptr);
    });
    return FT_OK;
}

int range_entry(hipStream_t s, const ft_tree *tree, const void *qx, const void *qy,
                const void *qz, const void *rmin, const void *rmax, int nq, int32_t *res,
                int32_t *counts, int max_res, const int64_t *row_ptr, int32_t *cols,
                int64_t col_cap) {
    FT_DISPATCH_COORD(tree->coord_type, {
        const DTree<T> dt = make_dtree<T>(*tree);
        hipLaunchKernelGGL(range_kernel<T>, dim3(blocks_for(nq, LANE_BLOCK)), dim3(LANE_BLOCK), 0,
                           s, dt, (const T *)qx, (const T *)qy, (const T *)qz, (const T *)rmin,
                           (const T *)rmax, nq, res, counts, max_res, row_ptr, cols, col_cap);
        FT_CHECK_LAUNCH();
    });
    return FT_OK;
}

int scan_entry(hipStream_t s, const int32_t *counts, int nq, int64_t *row_ptr) {
    const int nb = (nq + 1023) / 1024;
    int64_t *sums = nullptr;
    FT_HIP(hipMallocAsync((void **)&sums, (size_t)(nb > 0 ? nb : 1) * 8, s));
    if (nq > 0) {
        hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(256), 0, s, counts, nq, row_ptr, sums);
        hipLaunchKernelGGL(scan_totals_kernel, dim3(1), dim3(64), 0, s, sums, nb, row_ptr, nq);
        hipLaunchKernelGGL(scan_add_kernel, dim3(blocks_for(nq, 256)), dim3(256), 0, s, row_ptr,
                           nq, (const int64_t *)sums);
    } else {
        FT_HIP(hipMemsetAsync(row_ptr, 0, 8, s));
    }
    FT_CHECK_LAUNCH();
    FT_HIP(hipFreeAsync(sums, s));
    return FT_OK;
}

}  // namespace ft
