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
#include "ft_traverse.hpp"

#include <algorithm>
#include <vector>

namespace ft {

int scratch_ensure_public(Scratch &S, size_t n);


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
    __shared__ int32_t stk[LANE_STACK_CANON][LANE_BLOCK];
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
            if (sp < LANE_STACK_CANON - 2) {
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

// Per-lane DFS with an LDS heap row per lane (LocalMaxHeap, src/maxheap.hpp:340-478):
// REF = the reference's mechanics (ties as the reference picks them), else
// canonical.  LISTED: the block's lanes take entries of a device list of query
// positions (grid-stride; the list length is read on the device).
template <typename T, bool REF, bool GROUPED, bool LISTED>
__global__ __launch_bounds__(LANE_BLOCK) void lane_knn_lds_kernel(
    DTree<T> t, QSrc<T> qs, int nq, int k, int excl, int lanes, const int32_t *__restrict__ list,
    const uint32_t *__restrict__ list_count, uint64_t *__restrict__ res,
    typename CoordTraits<T>::D *__restrict__ dists, uint64_t *__restrict__ stats) {
    using D = typename CoordTraits<T>::D;
    constexpr int DEPTH = REF ? STACK_DEPTH : LANE_STACK_CANON;
    extern __shared__ __align__(16) unsigned char lds[];
    D *hd = (D *)lds;                                   // [k][lanes]
    int32_t *hi = (int32_t *)(hd + (size_t)k * lanes);  // [k][lanes]
    int32_t *stk = hi + (size_t)k * lanes;              // [DEPTH][lanes]
    (void)DEPTH;
    const int lane = threadIdx.x;
    const bool lane_ok = lane < lanes;
    const int nlist = LISTED ? (int)*list_count : nq;
    uint64_t visits = 0, drops = 0, inserts = 0;
    for (int base = blockIdx.x * lanes; base < nlist; base += (LISTED ? gridDim.x * lanes : nlist)) {
        const int slot = base + lane;
        const bool active = lane_ok && slot < nlist;
        const int q = active ? (LISTED ? list[slot] : slot) : -1;
        T px = T(0), py = T(0), pz = T(0);
        int self = -1;
        if (active) load_query(t, qs, q, px, py, pz, self);
        if (!excl) self = -1;
        LdsHeap<D, REF> h{hd, hi, lanes, lane, 0};
        lane_heap_walk<T, REF, GROUPED>(t, k, active, px, py, pz, self, h, stk, lanes, lane,
                                        visits, drops, inserts);
        if (active) {
            h.sort();
            const size_t o = (size_t)q * k;
            for (int s = 0; s < k; ++s) {
                const bool ok = s < h.count && h.ii(s) >= 0;
                res[o + s] = ok ? (uint64_t)(uint32_t)h.ii(s) : ~0ull;
                dists[o + s] = ok ? h.dd(s) : missing_dist<D>();
            }
        }
        if (!LISTED) break;
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
struct PacketArgs {
    int nq, k, excl, bucket, ibits, pad;
    const int32_t *list;        // LISTED: query positions
    const uint32_t *list_count; // LISTED: number of entries
    int32_t *rerun;             // re-run list (SEEDED or PACKED)
    uint32_t *rerun_count;
    uint64_t *stats;
};

template <typename T, int M, bool HEAP, bool SEEDED, bool PACKED, bool LISTED, int MK>
__global__ __launch_bounds__(64) void packet_knn_kernel(
    DTree<T> t, QSrc<T> qs, PacketArgs a, const typename CoordTraits<T>::D *__restrict__ seed,
    uint64_t *__restrict__ res, typename CoordTraits<T>::D *__restrict__ dists) {
    using D = typename CoordTraits<T>::D;
    using H = typename std::conditional<HEAP, TopKHeapLds<M>,
                                        typename TopKSel<D, M, PACKED>::type>::type;
    static_assert(!HEAP || PACKED, "LDS heap tier needs packed integer keys");
    static_assert(HEAP || M <= 32, "register tiers hold at most 32 entries");
    __shared__ int32_t stk[PK_STACK];
    [[maybe_unused]] uint64_t *heap_mem = nullptr;
    if constexpr (HEAP) {
        __shared__ uint64_t heap[M * 64];
        heap_mem = heap;
    }
    const int lane = threadIdx.x;
    const int k = a.k;
    const int nlist = LISTED ? (int)*a.list_count : a.nq;
    PacketCounters c;

    for (int base = blockIdx.x * 64; base < nlist; base += (LISTED ? gridDim.x * 64 : nlist)) {
        const int slot = base + lane;
        const bool active = slot < nlist;
        const int q = active ? (LISTED ? a.list[slot] : slot) : -1;
        T px = T(0), py = T(0), pz = T(0);
        int self = -1;
        if (active) load_query(t, qs, q, px, py, pz, self);
        if (!a.excl) self = -1;
        H h;
        if constexpr (HEAP) h.key = heap_mem;
        if constexpr (PACKED) h.init(k, (SEEDED && active) ? (uint64_t)seed[q] : ~0ull, a.ibits);
        else h.init(k, (SEEDED && active) ? seed[q] : dist_max<D>());
        packet_walk<T, H, MK>(t, a.bucket, active, px, py, pz, self, h, stk, c);
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
        atomicAdd((unsigned long long *)&a.stats[0], (unsigned long long)c.nodes);
        atomicAdd((unsigned long long *)&a.stats[1], (unsigned long long)c.leaves);
        if (c.overflow) atomicAdd((unsigned long long *)&a.stats[2], (unsigned long long)c.overflow);
    }
    wave_add_stats(a.stats, 4, c.inserts);
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

// Lanes per block for an LDS heap of k entries: as many as fit 64 KiB.
template <typename T, bool REF> static int lds_lanes(int k, size_t &lds) {
    using D = typename CoordTraits<T>::D;
    const size_t per_lane =
        (size_t)k * (sizeof(D) + 4) + (size_t)(REF ? STACK_DEPTH : LANE_STACK_CANON) * 4;
    int lanes = 64;
    const size_t budget = 64 * 1024;
    while (lanes > 1 && (size_t)lanes * per_lane > budget) lanes >>= 1;
    lds = (size_t)lanes * per_lane;
    return lds <= budget ? lanes : 0;
}

template <typename T, bool REF, bool GROUPED>
static int launch_lane_lds(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq, int k,
                           int excl, uint64_t *res, typename CoordTraits<T>::D *d,
                           uint64_t *stats, const int32_t *list = nullptr,
                           const uint32_t *list_count = nullptr) {
    size_t lds = 0;
    const int lanes = lds_lanes<T, REF>(k, lds);
    if (!lanes) return fail(FT_ELDS, "k too large for local memory: reduce k");
    if (list) {
        const unsigned g = std::min<unsigned>(blocks_for(nq, lanes), 4096u);
        hipLaunchKernelGGL((lane_knn_lds_kernel<T, REF, GROUPED, true>), dim3(g), dim3(LANE_BLOCK),
                           lds, s, dt, qs, nq, k, excl, lanes, list, list_count, res, d, stats);
    } else {
        hipLaunchKernelGGL((lane_knn_lds_kernel<T, REF, GROUPED, false>),
                           dim3(blocks_for(nq, lanes)), dim3(LANE_BLOCK), lds, s, dt, qs, nq, k,
                           excl, lanes, (const int32_t *)nullptr, (const uint32_t *)nullptr, res,
                           d, stats);
    }
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

template <typename T, int M, bool HEAP, bool SEEDED, bool PACKED, bool LISTED>
static void launch_pk(hipStream_t s, dim3 g, const DTree<T> &dt, const QSrc<T> &qs,
                      const PacketArgs &a, const typename CoordTraits<T>::D *seed, uint64_t *res,
                      typename CoordTraits<T>::D *d) {
    switch (metric_kind<T>(dt.m)) {
    case MK_WRAP:
        hipLaunchKernelGGL((packet_knn_kernel<T, M, HEAP, SEEDED, PACKED, LISTED, MK_WRAP>), g,
                           dim3(64), 0, s, dt, qs, a, seed, res, d);
        break;
    case MK_OPEN:
        hipLaunchKernelGGL((packet_knn_kernel<T, M, HEAP, SEEDED, PACKED, LISTED, MK_OPEN>), g,
                           dim3(64), 0, s, dt, qs, a, seed, res, d);
        break;
    default:
        hipLaunchKernelGGL((packet_knn_kernel<T, M, HEAP, SEEDED, PACKED, LISTED, MK_ANY>), g,
                           dim3(64), 0, s, dt, qs, a, seed, res, d);
        break;
    }
}

// Runs the packet traversal and its exact re-runs:
//   pass 1  direct, seeded or not, packed keys for integer metrics;
//   pass 2  (packed only) unseeded packed run of the queries pass 1 listed;
//   pass 3  unseeded lexicographic run of what is still listed -- exact.
// The LISTED passes read list lengths on the device: no host round trip.
template <typename T, int M, bool HEAP>
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
        if constexpr (INTD) launch_pk<T, M, HEAP, true, true, false>(s, g, dt, qs, a, seed, res, d);
        else launch_pk<T, M, false, true, false, false>(s, g, dt, qs, a, seed, res, d);
    } else {
        if constexpr (INTD) launch_pk<T, M, HEAP, false, true, false>(s, g, dt, qs, a, seed, res, d);
        else launch_pk<T, M, false, false, false, false>(s, g, dt, qs, a, seed, res, d);
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
            launch_pk<T, M, HEAP, false, true, true>(s, dim3(waves), dt, qs, b, nullptr, res, d);
            FT_CHECK_LAUNCH();
            lex_list = 1;
        }
    }
    b.list = g_rerun.list[lex_list];
    b.list_count = g_rerun.count + lex_list;
    b.rerun = nullptr;
    b.rerun_count = nullptr;
    if constexpr (HEAP) {
        // exact: per-lane lexicographic LDS heap over the listed queries
        return launch_lane_lds<T, false, true>(s, dt, qs, nq, k, excl, res, d, nullptr, b.list,
                                               b.list_count);
    } else {
        launch_pk<T, M, false, false, false, true>(s, dim3(waves), dt, qs, b, nullptr, res, d);
    }
    FT_CHECK_LAUNCH();
    return FT_OK;
}

template <typename T>
static int launch_packet(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq, int k,
                         int excl, int bucket, const typename CoordTraits<T>::D *seed,
                         uint64_t *res, typename CoordTraits<T>::D *d, uint64_t *stats,
                         int topk) {
    constexpr bool INTD = !CoordTraits<T>::fp;
    const bool heap = INTD && (k > 32 || topk == FT_TOPK_LDS_HEAP);
    if (!heap) {
        if (k <= 1) return packet_run<T, 1, false>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
        if (k <= 4) return packet_run<T, 4, false>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
        if (k <= 8) return packet_run<T, 8, false>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
        if (k <= 16) return packet_run<T, 16, false>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
        if (k <= 32) return packet_run<T, 32, false>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
    }
    if constexpr (INTD) {
        if (k <= 32) return packet_run<T, 32, true>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
        if (k <= 64) return packet_run<T, 64, true>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
        if (k <= 128) return packet_run<T, 128, true>(s, dt, qs, nq, k, excl, bucket, seed, res, d, stats);
    }
    return fail(FT_EINVAL, "packet traversal: k above the heap tiers");
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
    // Packet tiers: registers for k <= 32; an LDS heap per lane for integer
    // metrics up to k = 128.  Larger k (or FP metrics above 32) take the
    // per-lane LDS heap walk.
    const bool packet_ok = k <= 32 || (!CoordTraits<T>::fp && k <= 128);
    if (algo == FT_ALGO_AUTO) algo = packet_ok ? FT_ALGO_PACKET : FT_ALGO_GROUPED;
    if (o.seed_bound && packet_ok) algo = FT_ALGO_PACKET;
    if (algo == FT_ALGO_PACKET && packet_ok) {
        const int bucket = o.leaf_bucket > 0 ? std::min(o.leaf_bucket, 64) : 32;
        return launch_packet<T>(s, dt, qs, nq, k, excl, bucket, (const D *)o.seed_bound, res,
                                dists, stats, o.topk);
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
