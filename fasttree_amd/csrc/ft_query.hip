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
#include <mutex>
#include <unordered_map>
#include <vector>

namespace ft {

int scratch_ensure_public(Scratch &S, size_t n);
int radix_sort_u64(hipStream_t s, Scratch &S, const uint64_t *keys_in, size_t n,
                   uint64_t *sorted_out, int key_bits);


// ===========================================================================
// Query positions.  A "slot" is a position in the launch's query order; the
// query it holds is order[slot] (external queries sorted along the curve) or
// the slot itself.  Results are written at the query's own row.  Re-run lists
// hold slots in ascending order (compacted from per-slot flags), so a listed
// pass walks its queries in curve order too.
// ===========================================================================
struct QueryMap {
    const uint32_t *order;      // slot -> query (null: identity)
    const int32_t *list;        // LISTED: slots to run (-1: already answered)
    const uint32_t *list_count; // LISTED: number of list entries (device)
    uint8_t *redo;              // direct passes: per-slot flag, set when the slot
                                // must be re-run exactly
};

__device__ inline int slot_query(const QueryMap &qm, int slot) {
    return qm.order ? (int)qm.order[slot] : slot;
}

// ===========================================================================
// Per-lane DFS with the canonical register top-k (k <= M <= 32): the
// reference's grouped self-kNN / knn_query_small_k walk on wave64 lanes.
// Direct passes use the reference's 64-entry stack and flag a query whose
// walk would need more; LISTED passes (the exact re-runs) use 128 entries,
// and for self queries first insert the leaves within LISTED_WINDOW ranks of
// the query's own leaf, so that the walk starts with a finite bound and skips
// those leaves (the top-k set does not depend on the order of insertion).
// ===========================================================================
constexpr int LISTED_WINDOW = 32;

template <typename T, int M, bool GROUPED, bool LISTED, int MK>
__global__ __launch_bounds__(LANE_BLOCK) void lane_knn_reg_kernel(
    DTree<T> t, QSrc<T> qs, int nq, int k, int excl, QueryMap qm, uint64_t *__restrict__ res,
    typename CoordTraits<T>::D *__restrict__ dists, uint64_t *__restrict__ stats) {
    using D = typename CoordTraits<T>::D;
    constexpr int DEPTH = LISTED ? LANE_STACK_CANON : STACK_DEPTH;
    __shared__ int32_t stk[DEPTH * LANE_BLOCK];
    const int lane = threadIdx.x;
    const int nlist = LISTED ? (int)*qm.list_count : nq;
    uint64_t visits = 0, inserts = 0;
    for (int base = (LISTED ? (int)blockIdx.x : xcd_block()) * LANE_BLOCK; base < nlist;
         base += (LISTED ? gridDim.x * LANE_BLOCK : nlist)) {
        const int pos = base + lane;
        const int lslot = (LISTED && pos < nlist) ? qm.list[pos] : 0;
        const bool active = pos < nlist && (!LISTED || lslot >= 0);
        const int slot = active ? (LISTED ? lslot : pos) : 0;
        const int q = active ? slot_query(qm, slot) : 0;
        T px = T(0), py = T(0), pz = T(0);
        int self = -1;
        if (active) load_query(t, qs, q, px, py, pz, self);
        if (!excl) self = -1;
        TopK<D, M> h;
        h.init(k, dist_max<D>());
        int wlo = 0, whi = 0;
        if constexpr (LISTED) {
            if (active && self >= 0) {
                const int off = t.n - 1;
                wlo = self - LISTED_WINDOW > 0 ? self - LISTED_WINDOW : 0;
                whi = self + LISTED_WINDOW + 1 < t.n ? self + LISTED_WINDOW + 1 : t.n;
                for (int r = wlo; r < whi; ++r) {
                    if (r == self) continue;
                    const int leaf = off + r;
                    const D d2 = leaf_d2_k<T, MK>(t.m, px, py, pz, t.mnx[leaf], t.mny[leaf],
                                                  t.mnz[leaf]);
                    const int32_t ri = ret_index(t, leaf);
                    const bool take = h.beats_worst(d2, ri);
                    h.insert(take, d2, ri);
                    inserts += take ? 1 : 0;
                }
            }
        }
        const bool ovf = lane_reg_walk<T, M, GROUPED, DEPTH, MK>(t, active, px, py, pz, self, h,
                                                                 stk, LANE_BLOCK, lane, visits,
                                                                 inserts, wlo, whi);
        if (active) {
            if (!LISTED && ovf) qm.redo[slot] = 1;
            else h.store(k, res, dists, (size_t)q);
        }
        if (!LISTED) break;
    }
    wave_add_stats(stats, 0, visits);
    wave_add_stats(stats, 4, inserts);
}

// Per-lane DFS with an LDS heap row per lane (LocalMaxHeap, src/maxheap.hpp:340-478):
// REF = the reference's mechanics (ties as the reference picks them, 64-entry
// stack with its silent drop), else canonical (direct passes flag overflowing
// walks; LISTED re-runs use 128 entries).
template <typename T, bool REF, bool GROUPED, bool LISTED, int MK>
__global__ __launch_bounds__(LANE_BLOCK) void lane_knn_lds_kernel(
    DTree<T> t, QSrc<T> qs, int nq, int k, int excl, int lanes, QueryMap qm,
    uint64_t *__restrict__ res, typename CoordTraits<T>::D *__restrict__ dists,
    uint64_t *__restrict__ stats) {
    using D = typename CoordTraits<T>::D;
    constexpr int DEPTH = (LISTED && !REF) ? LANE_STACK_CANON : STACK_DEPTH;
    extern __shared__ __align__(16) unsigned char lds[];
    D *hd = (D *)lds;                                   // [k][lanes]
    int32_t *hi = (int32_t *)(hd + (size_t)k * lanes);  // [k][lanes]
    int32_t *stk = hi + (size_t)k * lanes;              // [DEPTH][lanes]
    const int lane = threadIdx.x;
    const bool lane_ok = lane < lanes;
    const int nlist = LISTED ? (int)*qm.list_count : nq;
    uint64_t visits = 0, drops = 0, inserts = 0;
    for (int base = (LISTED ? (int)blockIdx.x : xcd_block()) * lanes; base < nlist;
         base += (LISTED ? gridDim.x * lanes : nlist)) {
        const int pos = base + lane;
        const int lslot = (LISTED && lane_ok && pos < nlist) ? qm.list[pos] : 0;
        const bool active = lane_ok && pos < nlist && (!LISTED || lslot >= 0);
        const int slot = active ? (LISTED ? lslot : pos) : 0;
        const int q = active ? slot_query(qm, slot) : 0;
        T px = T(0), py = T(0), pz = T(0);
        int self = -1;
        if (active) load_query(t, qs, q, px, py, pz, self);
        if (!excl) self = -1;
        LdsHeap<D, REF> h{hd, hi, lanes, lane, 0};
        const bool ovf = lane_heap_walk<T, REF, GROUPED, DEPTH, MK>(
            t, k, active, px, py, pz, self, h, stk, lanes, lane, visits, drops, inserts);
        if (active) {
            if (!REF && !LISTED && ovf) {
                qm.redo[slot] = 1;
            } else {
                h.sort();
                const size_t o = (size_t)q * k;
                for (int s = 0; s < k; ++s) {
                    const bool ok = s < h.count && h.ii(s) >= 0;
                    res[o + s] = ok ? (uint64_t)(uint32_t)h.ii(s) : ~0ull;
                    dists[o + s] = ok ? h.dd(s) : missing_dist<D>();
                }
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
// Packet traversal: a wavefront of 64 queries (consecutive slots) shares one
// DFS (packet_walk).  Results are canonical.
// SEEDED: each lane's k-th distance starts at its seed bound; a lane that ends
// with fewer than k neighbours inside the seed flags its slot for an exact
// re-run.  PACKED: 64-bit (d2, index) keys; queries the packed form cannot
// order exactly are flagged too.  The re-run (launch_exact_listed) takes the
// compacted list of flagged slots, unseeded and lexicographic, so exact.
// ===========================================================================
struct PacketArgs {
    int nq, k, excl, bucket, ibits;
    int width;   // queries per packet (16, 32 or 64): lanes at or above it idle
    int window;  // self queries over all leaves: leaves pre-scanned on each side
    uint64_t *stats;
};

// MULTI (register tier, packed keys, k > M): ceil(k / M) passes, a separate
// instantiation so that the single-pass kernels keep their register budget.
// ===========================================================================
// Wave-per-query canonical tier (FT_ALGO_WAVE): one wavefront per query, its
// sorted list of C = 64 S entries spread over the lanes and candidates merged
// 64 leaves at a time by cross-lane bitonic networks (WaveList, wave_walk) --
// the canonical counterpart of knn_query_large_k's SortedMergeHeap
// (src/hlbvh.hpp:1232-1387, src/maxheap.hpp:226-324).  Exact in one walk for
// k up to 1024.  Four queries per 256-thread block; self and subset queries
// anchor their first leaves at their own rank, external ones at the leaf a
// greedy descent reaches.
//
// LISTED: the first step of the exact re-run of a direct pass's flagged
// slots (qm.list, *qm.list_count).  The waves take list entries in turn, and
// an entry whose query they answer is set to -1 in done_list (the list
// itself), so the per-lane re-run that follows takes only the entries whose
// walk would overflow the wave's stack (none for a tree this library built).
// ===========================================================================
template <typename T, int S, int MK, bool LISTED = false>
__global__ __launch_bounds__(256) void wave_knn_kernel(DTree<T> t, QSrc<T> qs, int nq, int k,
                                                       int excl, QueryMap qm,
                                                       uint64_t *__restrict__ res,
                                                       typename CoordTraits<T>::D *__restrict__ dists,
                                                       uint64_t *__restrict__ stats,
                                                       int32_t *__restrict__ done_list) {
    __shared__ WaveBuf wbuf[4];
    const int lane = threadIdx.x & 63;
    const int wave = uni(xcd_block() * 4 + (int)(threadIdx.x >> 6));
    WaveCounters c;
    // Answers the query of one slot; false if its walk would overflow.
    auto answer = [&](int slot) -> bool {
        const int q = uni(slot_query(qm, slot));
        T px = T(0), py = T(0), pz = T(0);
        int rank = -1;
        load_query(t, qs, q, px, py, pz, rank);
        const int self = excl ? rank : -1;
        const int anchor = rank >= 0 ? rank : descend_rank_v<T, MK>(t, px, py, pz);
        WaveList<S> L;
        L.init(k);
        if (wave_walk<T, S, MK>(t, px, py, pz, self, anchor, L, c, &wbuf[threadIdx.x >> 6]))
            return false;
        L.store(k, res, dists, (size_t)q, lane);
        return true;
    };
    if constexpr (LISTED) {
        const int cnt = (int)*qm.list_count;
        for (int e = wave; e < cnt; e += (int)gridDim.x * 4)
            if (answer(uni(qm.list[e])) && lane == 0) done_list[e] = -1;
    } else if (wave < nq) {
        if (!answer(wave) && lane == 0) qm.redo[wave] = 1;
    }
    if (stats) {  // one atomic per block and counter: a wave per query would contend
        __shared__ unsigned long long s_st[3];
        if (threadIdx.x < 3) s_st[threadIdx.x] = 0;
        __syncthreads();
        if (lane == 0) {
            atomicAdd(&s_st[0], (unsigned long long)c.nodes);
            atomicAdd(&s_st[1], (unsigned long long)c.leaves);
            atomicAdd(&s_st[2], (unsigned long long)c.inserts);
        }
        __syncthreads();
        if (threadIdx.x < 3 && s_st[threadIdx.x])
            atomicAdd((unsigned long long *)&stats[threadIdx.x == 2 ? 4 : threadIdx.x],
                      s_st[threadIdx.x]);
    }
}

// Packed-key packets: at least 4 waves per SIMD (128 VGPRs).  The k = 32
// kernel spills 72 bytes and still runs 11 % faster than at 3 waves (144
// VGPRs); the passes of 32 for k = 64 run 19 % faster than at 2 waves.  Double
// coordinates keep their natural 3 waves (forcing 4 spilled 116 bytes and was
// 32 % slower).
template <typename T, int M, bool HEAP, bool SEEDED, bool PACKED, int MK, bool MULTI = false>
__global__ __launch_bounds__(64, PACKED ? 4 : 2) void packet_knn_kernel(
    DTree<T> t, QSrc<T> qs, PacketArgs a, QueryMap qm,
    const typename CoordTraits<T>::D *__restrict__ seed, uint64_t *__restrict__ res,
    typename CoordTraits<T>::D *__restrict__ dists) {
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
    PacketCounters c;
    const int blk = xcd_block();
    const int W = a.width;
    if (blk * W >= a.nq) return;  // padding block of the XCD-ordered grid
    const int slot = blk * W + lane;
    const bool active = lane < W && slot < a.nq;
    const int q = active ? slot_query(qm, slot) : 0;
    T px = T(0), py = T(0), pz = T(0);
    int self = -1;
    if (active) load_query(t, qs, q, px, py, pz, self);
    const int rank = self;  // the query's leaf rank (self queries), else -1
    if (!a.excl) self = -1;
    // Window pre-scan.  Self queries over all leaves: the packet is leaf ranks
    // [r0, r0 + 64), so its neighbours along the curve are the leaves on
    // either side.  Subset and external queries (Morton-ordered, so a
    // packet's queries are neighbours): the run of 64 leaves around the
    // middle lane's leaf rank, or around the leaf a greedy descent reaches
    // for the middle lane's point.
    int wlo = 0, whi = 0;
    int r0 = blk * W;
    if (a.window > 0) {
        const int src = (blk * W + W / 2 < a.nq) ? W / 2 : 0;
        if (qs.qx) r0 = run_start(descend_rank<T, MK>(t, rl(px, src), rl(py, src), rl(pz, src)), t.n);
        else if (qs.leaf_ids) r0 = run_start(uni(__shfl(rank, src, 64)), t.n);
        wlo = r0 - a.window > 0 ? r0 - a.window : 0;
        whi = r0 + 64 + a.window < t.n ? r0 + 64 + a.window : t.n;
    }
    static_assert(!MULTI || !HEAP, "passes need the register tier");
    if constexpr (MULTI) {
        // k above M (register tier): ceil(k / M) passes, each taking the M
        // nearest (d2, index) above the previous pass's last one, stored in
        // place.
        bool redo = false;
        typename H::Floor lo = H::no_floor();
        for (int base = 0; base < k; base += M) {
            const int kk = k - base < M ? k - base : M;
            H h;
            // The seed is read again in every pass (an opaque copy of q keeps
            // the load in the loop): holding it across the passes would cost
            // the kernel two registers.
            int qv = q;
            asm volatile("" : "+v"(qv));
            if constexpr (PACKED) h.init(kk, (SEEDED && active) ? (uint64_t)seed[qv] : ~0ull, a.ibits);
            else h.init(kk, (SEEDED && active) ? seed[qv] : dist_max<D>());
            // one walk instance for every pass (the first with no floor)
            packet_walk<T, H, MK, true>(t, a.bucket, active, px, py, pz, self, h, stk, c, wlo,
                                        whi, r0, lo);
            if (active) {
                bool r = false;
                if constexpr (PACKED) r = h.needs_exact(kk);
                else if constexpr (SEEDED) r = h.missing(kk) > 0;
                redo = redo || r;
                if (!redo) h.store_part(kk, res, dists, (size_t)q, k, base);
            }
            lo = h.next_floor();
        }
        if (active && redo) qm.redo[slot] = 1;
    } else {
        H h;
        if constexpr (HEAP) h.key = heap_mem;
        if constexpr (PACKED) h.init(k, (SEEDED && active) ? (uint64_t)seed[q] : ~0ull, a.ibits);
        else h.init(k, (SEEDED && active) ? seed[q] : dist_max<D>());
        int flo = 0, fhi = 0;
        if constexpr (H::kSortFill) {
            // self queries over all leaves: lane = leaf rank r0 + lane; the
            // window (>= 32 leaves on each side) covers every filled rank
            if (!qs.qx && !qs.leaf_ids && a.window >= 32) {
                __shared__ FillStage<T> fst;
                c.inserts += (uint64_t)fill_from_run<T, MK>(t, h, k, active, self >= 0, rank, px,
                                                            py, pz, r0, &fst, flo, fhi);
            }
        }
        packet_walk<T, H, MK>(t, a.bucket, active, px, py, pz, self, h, stk, c, wlo, whi, r0,
                              H::no_floor(), flo, fhi);
        if (active) {
            bool redo = false;
            if constexpr (PACKED) redo = h.needs_exact(k);
            else if constexpr (SEEDED) redo = h.missing(k) > 0;
            if (redo) qm.redo[slot] = 1;
            else h.store(k, res, dists, (size_t)q);
        }
    }
    if (lane == 0 && a.stats) {
        atomicAdd((unsigned long long *)&a.stats[0], (unsigned long long)c.nodes);
        atomicAdd((unsigned long long *)&a.stats[1], (unsigned long long)c.leaves);
        if (c.overflow) atomicAdd((unsigned long long *)&a.stats[2], (unsigned long long)c.overflow);
        atomicAdd((unsigned long long *)&a.stats[5], (unsigned long long)c.rounds);
        atomicAdd((unsigned long long *)&a.stats[6], (unsigned long long)c.chains);
    }
    wave_add_stats(a.stats, 4, c.inserts);
}

// ===========================================================================
// Order-preserving compaction of per-slot flags into a slot list (ascending),
// with the count left on the device: per-tile counts, one scan block, scatter.
// ===========================================================================
constexpr int CP_TILE = 4096;  // flags per block (256 threads x 16)

__global__ __launch_bounds__(256) void flag_count_kernel(const uint8_t *__restrict__ flags, int n,
                                                         uint32_t *__restrict__ tile_counts) {
    __shared__ uint32_t s_w[4];
    const int base = blockIdx.x * CP_TILE;
    uint32_t c = 0;
    for (int i = threadIdx.x; i < CP_TILE; i += 256) {
        const int j = base + i;
        c += (j < n && flags[j]) ? 1u : 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Exclusive scan of the tile counts in place; total to *count.
__global__ __launch_bounds__(1024) void flag_scan_kernel(uint32_t *__restrict__ tile_counts,
                                                         int nt, uint32_t *__restrict__ count) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int base = 0; base < nt; base += 1024) {
        const int i = base + threadIdx.x;
        const uint32_t v = i < nt ? tile_counts[i] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[w] = x;
        __syncthreads();
        uint32_t add = s_carry;
        for (int q = 0; q < w; ++q) add += s_w[q];
        if (i < nt) tile_counts[i] = add + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = add + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = s_carry;
}

__global__ __launch_bounds__(256) void flag_scatter_kernel(const uint8_t *__restrict__ flags,
                                                           int n,
                                                           const uint32_t *__restrict__ tile_base,
                                                           int32_t *__restrict__ list) {
    __shared__ uint32_t s_w[4];
    __shared__ uint32_t s_run;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_run = tile_base[blockIdx.x];
    __syncthreads();
    const int base = blockIdx.x * CP_TILE;
    for (int r = 0; r < CP_TILE; r += 256) {
        const int j = base + r + threadIdx.x;
        const bool f = j < n && flags[j];
        const uint64_t m = __ballot(f);
        const uint32_t below = (uint32_t)__popcll(m & (lane ? (~0ull >> (64 - lane)) : 0ull));
        if (lane == 0) s_w[w] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = s_run;
        for (int q = 0; q < w; ++q) off += s_w[q];
        if (f) list[off + below] = j;
        __syncthreads();
        if (threadIdx.x == 0) s_run += s_w[0] + s_w[1] + s_w[2] + s_w[3];
        __syncthreads();
    }
}


// Range query, 64 queries per wave on one shared DFS (range_query,
// src/hlbvh.hpp:2006-2136), with the reference's results exactly:
//  * The reference's per-query DFS pushes right then left, so every query
//    meets its hits in ascending leaf rank; the shared walk is left-first too,
//    and each lane appends its hits in the order the walk meets them.
//  * The reference pushes a node's children only while its stack holds fewer
//    than 62 entries.  When it pops node v its stack holds the right siblings
//    pending along v's path, one per left turn, so that count -- the same for
//    every query that reaches v -- decides: the shared walk carries each
//    node's left turns and does not descend below a node with 62.
//  * A subtree of at most `bucket` leaves is tested leaf by leaf without
//    descending (its nodes add at most cnt - 1 left turns, so this is done
//    only when the reference could not have stopped inside it).
//  * A node is skipped only when no lane needs it (exact box distance, or the
//    conservative float32 form for integer periodic metrics); every leaf is
//    tested exactly: r_min^2 <= d2 <= r_max^2.
// Queries are taken in the order `order` (slot -> query; Morton order of the
// points inside the root box), so a wave's 64 queries are neighbours.
constexpr int RANGE_DROP_DEPTH = STACK_DEPTH - 2;  // 62

template <typename T, int MK>
__global__ __launch_bounds__(64) void range_packet_kernel(
    DTree<T> t, const T *__restrict__ qx, const T *__restrict__ qy, const T *__restrict__ qz,
    const T *__restrict__ rmin, const T *__restrict__ rmax, int nq,
    const uint32_t *__restrict__ order, int bucket, int32_t *__restrict__ res,
    int32_t *__restrict__ counts, int max_res, const int64_t *__restrict__ row_ptr,
    int32_t *__restrict__ cols, int64_t col_cap) {
    using D = typename CoordTraits<T>::D;
    __shared__ int32_t stk[PK_STACK];
    __shared__ uint8_t stk_lt[PK_STACK];
    const int lane = threadIdx.x;
    const int blk = xcd_block();
    if (blk * 64 >= nq) return;  // padding block of the XCD-ordered grid
    const int slot = blk * 64 + lane;
    const bool active = slot < nq;
    const int q = active ? (order ? (int)order[slot] : slot) : 0;
    const int n = t.n;
    const int off = n - 1;
    T px = T(0), py = T(0), pz = T(0);
    D rm2 = D(0), RM2 = D(0);
    if (active) {
        px = qx[q]; py = qy[q]; pz = qz[q];
        if constexpr (CoordTraits<T>::fp) {
            const double a = (double)rmin[q], b = (double)rmax[q];
            rm2 = a * a;
            RM2 = b * b;
        } else {
            const uint64_t a = (uint64_t)rmin[q] >> t.m.dshift, b = (uint64_t)rmax[q] >> t.m.dshift;
            rm2 = a * a;
            RM2 = b * b;
        }
    }
    constexpr bool PRE = !CoordTraits<T>::fp && MK == MK_WRAP;
    [[maybe_unused]] const int wsh = PRE ? 32 - t.m.bits : 0;
    [[maybe_unused]] uint32_t qxs = 0, qys = 0, qzs = 0;
    [[maybe_unused]] T qmx = px, qmy = py, qmz = pz;
    [[maybe_unused]] float RM2f = 0.f;
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
        RM2f = wrap_filter_bound_pre((uint64_t)RM2, wsh);
    }
    auto node_near = [&](int nd) -> bool {
        if (!active) return false;
        if constexpr (PRE) {
            const float gx = wrap_box_gap_pre<T>(qmx, qxs, t.mnx[nd], t.mxx[nd], wsh);
            const float gy = wrap_box_gap_pre<T>(qmy, qys, t.mny[nd], t.mxy[nd], wsh);
            const float gz = wrap_box_gap_pre<T>(qmz, qzs, t.mnz[nd], t.mxz[nd], wsh);
            float d = gx * gx;
            d = __builtin_fmaf(gy, gy, d);
            d = __builtin_fmaf(gz, gz, d);
            return d <= RM2f;
        } else {
            return node_d2_k<T, MK>(t, px, py, pz, nd) <= RM2;
        }
    };
    const bool writes = res || cols;
    const int64_t w0 = (active && row_ptr) ? row_ptr[q] : 0;
    int count = 0;
    // Tests leaf ranks [first, first + cnt) (cnt <= 64): lane j loads leaf j.
    auto scan = [&](int first, int cnt) {
        T bx = T(0), by = T(0), bz = T(0);
        int32_t bo = 0;
        if (lane < cnt) {
            bx = t.mnx[off + first + lane];
            by = t.mny[off + first + lane];
            bz = t.mnz[off + first + lane];
            bo = ret_index(t, off + first + lane);
        }
        uint64_t hit = 0;
        if constexpr (PRE) {
            const uint32_t ex = (uint32_t)bx << wsh, ey = (uint32_t)by << wsh,
                           ez = (uint32_t)bz << wsh;
            for (int j = 0; j < cnt; ++j) {
                const uint64_t gx = wrap_gap_pre(qxs, rl(ex, j), wsh);
                const uint64_t gy = wrap_gap_pre(qys, rl(ey, j), wsh);
                const uint64_t gz = wrap_gap_pre(qzs, rl(ez, j), wsh);
                const uint64_t d = gx * gx + gy * gy + gz * gz;
                hit |= (d >= rm2 && d <= RM2) ? (1ull << j) : 0ull;
            }
        } else {
            for (int j = 0; j < cnt; ++j) {
                const D d = leaf_d2_k<T, MK>(t.m, px, py, pz, rl(bx, j), rl(by, j), rl(bz, j));
                hit |= (d >= rm2 && d <= RM2) ? (1ull << j) : 0ull;
            }
        }
        if (!active) hit = 0;
        if (!writes) {
            count += __popcll(hit);
            return;
        }
        while (__any(hit != 0)) {  // one hit per lane per round, ascending rank
            const bool has = hit != 0;
            const int j = has ? (__ffsll((long long)hit) - 1) : 0;
            hit &= hit - 1;
            const int32_t ri = shfl_any(bo, j);
            if (has) {
                if (res && count < max_res) res[(size_t)q * max_res + count] = ri;
                if (cols && w0 + count < col_cap) cols[w0 + count] = ri;
                ++count;
            }
        }
    };
    int sp = 0;
    stk[0] = 0;
    stk_lt[0] = 0;
    sp = 1;
    while (sp > 0) {
        --sp;
        const int v = uni(stk[sp]);
        const int ltv = uni((int)stk_lt[sp]);
        if (!__any(node_near(v))) continue;
        if (n == 1) {
            scan(0, 1);
            continue;
        }
        if (v >= off) {
            scan(v - off, 1);
            continue;
        }
        const int first = uni(t.rfirst[v]);
        const int cnt = uni(t.rlast[v]) - first + 1;
        if (cnt <= bucket && ltv + cnt - 1 < RANGE_DROP_DEPTH) {
            scan(first, cnt);
            continue;
        }
        if (ltv >= RANGE_DROP_DEPTH) continue;  // the reference stops here
        // a Karras tree is less than 97 levels deep: the stack never overflows
        stk[sp] = uni(t.right[v]);
        stk_lt[sp] = (uint8_t)ltv;
        stk[sp + 1] = uni(t.left[v]);
        stk_lt[sp + 1] = (uint8_t)(ltv + 1);
        sp += 2;
    }
    if (active && counts) counts[q] = count;
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

// Block totals -> exclusive offsets in place, and the grand total into
// out[nq]: one 1024-thread block, each thread taking a contiguous run of
// ceil(nb / 1024) totals (a single lane walking 20k totals took 2 ms).
__global__ __launch_bounds__(1024) void scan_totals_kernel(int64_t *block_sums, int nb,
                                                           int64_t *out, int nq) {
    __shared__ int64_t wsum[16];
    const int per = (nb + 1023) / 1024;
    const int b0 = (int)threadIdx.x * per;
    const int b1 = b0 + per < nb ? b0 + per : nb;
    int64_t s = 0;
    for (int b = b0; b < b1; ++b) s += block_sums[b];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int64_t run = x - s;
    for (int q = 0; q < w; ++q) run += wsum[q];
    for (int b = b0; b < b1; ++b) {
        const int64_t v = block_sums[b];
        block_sums[b] = run;
        run += v;
    }
    if (threadIdx.x == 1023) out[nq] = run;
}

__global__ __launch_bounds__(256) void scan_add_kernel(int64_t *out, int nq,
                                                       const int64_t *block_sums) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) out[i] += block_sums[i / 1024];
}

// ===========================================================================
// Query order for packet launches: external points by their Morton code (21
// bits per axis) inside the tree's root box, subset leaf ranks by rank, so a
// wavefront's 64 queries are neighbours whatever order the caller used.
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256) void query_key_kernel(DTree<T> t, QSrc<T> qs, int nq,
                                                        uint64_t *__restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    if (!qs.qx) {
        keys[i] = (uint64_t)(uint32_t)qs.leaf_ids[i];
        return;
    }
    constexpr int QB = 21;
    const double side = (double)((1u << QB) - 1);
    auto quant = [&](T v, T lo, T hi) -> uint64_t {
        const double ext = (double)hi - (double)lo;
        double u = ext > 0.0 ? ((double)v - (double)lo) / ext : 0.0;
        u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
        return (uint64_t)(u * side);
    };
    const uint64_t ix = quant(qs.qx[i], t.mnx[0], t.mxx[0]);
    const uint64_t iy = quant(qs.qy[i], t.mny[0], t.mxy[0]);
    const uint64_t iz = quant(qs.qz[i], t.mnz[0], t.mxz[0]);
    uint64_t m = 0;
    for (int b = 0; b < QB; ++b) {
        m |= ((ix >> b) & 1ull) << (3 * b + 2);
        m |= ((iy >> b) & 1ull) << (3 * b + 1);
        m |= ((iz >> b) & 1ull) << (3 * b);
    }
    keys[i] = m;
}

// ===========================================================================
// Host dispatch
// ===========================================================================
static inline unsigned blocks_for(size_t n, int per) {
    size_t g = (n + per - 1) / per;
    return (unsigned)(g ? g : 1);
}

// Lanes per block for an LDS heap of k entries with a DEPTH-entry stack: as
// many as fit 64 KiB.
template <typename T> static int lds_lanes(int k, int depth, size_t &lds) {
    using D = typename CoordTraits<T>::D;
    const size_t per_lane = (size_t)k * (sizeof(D) + 4) + (size_t)depth * 4;
    int lanes = 64;
    const size_t budget = 64 * 1024;
    while (lanes > 1 && (size_t)lanes * per_lane > budget) lanes >>= 1;
    lds = (size_t)lanes * per_lane;
    return lds <= budget ? lanes : 0;
}

// Per-stream work buffers of the query paths: the re-run buffers (per-slot
// flags, the compacted slot list, per-tile counts and the list length) and
// the query-order sort buffers.  Each stream has its own set, grown on
// demand, so queries on different streams do not share memory; a stream's
// set is released by ft_stream_destroy / ft_stream_release.
//
// A graph captured on a stream holds the raw addresses of that stream's
// buffers.  While any such graph lives, a buffer that grows, or a set that is
// released, is not freed: its block joins the set's retired list, freed when
// the stream's last graph is destroyed, so a replay never touches freed
// memory.
struct RerunBuf {
    uint8_t *flags = nullptr;
    int32_t *list = nullptr;
    uint32_t *tiles = nullptr;
    uint32_t *count = nullptr;
    size_t cap = 0;
    void *block = nullptr;
};
struct QueryOrderBuf {
    Scratch S;
    uint64_t *keys = nullptr, *sorted = nullptr;
    size_t cap = 0;
};
struct StreamBufs {
    RerunBuf rr;
    QueryOrderBuf qo;
    int graphs = 0;               // live graphs captured on the stream
    bool released = false;        // out of the map, kept for its graphs
    std::vector<void *> retired;  // blocks replaced while graphs > 0
};
static std::mutex g_bufs_mu;
static std::unordered_map<hipStream_t, StreamBufs *> g_bufs;
static std::unordered_map<void *, StreamBufs *> g_graph_bufs;

static StreamBufs &stream_bufs(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_bufs_mu);
    auto it = g_bufs.find(s);
    if (it != g_bufs.end()) return *it->second;
    StreamBufs *b = new StreamBufs();
    g_bufs.emplace(s, b);
    return *b;
}

// Frees block p of stream s's set b after s's pending work, or keeps it on
// the retired list while a graph captured on s may still use it.
static int drop_block(StreamBufs &b, hipStream_t s, void *p) {
    if (!p) return FT_OK;
    {
        std::lock_guard<std::mutex> lk(g_bufs_mu);
        if (b.graphs > 0) {
            b.retired.push_back(p);
            return FT_OK;
        }
    }
    FT_HIP(hipStreamSynchronize(s));  // this stream's launches may still use it
    (void)hipFree(p);
    return FT_OK;
}

static void free_set(StreamBufs *b) {
    if (b->rr.block) (void)hipFree(b->rr.block);
    if (b->qo.keys) (void)hipFree(b->qo.keys);
    if (b->qo.S.block) (void)hipFree(b->qo.S.block);
    for (void *p : b->retired) (void)hipFree(p);
    for (auto &e : b->qo.S.ev)
        if (e) (void)hipEventDestroy(e);
    delete b;
}

// Frees the work buffers of stream s (after its pending work).  A set that a
// live graph uses leaves the map (the stream's next query makes a new one)
// and is freed with that graph.
int release_stream_bufs(hipStream_t s) {
    StreamBufs *b = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_bufs_mu);
        auto it = g_bufs.find(s);
        if (it == g_bufs.end()) return FT_OK;
        b = it->second;
        g_bufs.erase(it);
        if (b->graphs > 0) {
            b->released = true;
            return FT_OK;
        }
    }
    (void)hipStreamSynchronize(s);
    free_set(b);
    return FT_OK;
}

// A graph exec captured on stream s now holds s's buffers.
void graph_attach(hipStream_t s, void *exec) {
    StreamBufs &b = stream_bufs(s);
    std::lock_guard<std::mutex> lk(g_bufs_mu);
    ++b.graphs;
    g_graph_bufs[exec] = &b;
}

// The graph exec is destroyed: when it was its stream's last, the retired
// blocks (and a released set) are freed once the device is idle, since the
// graph may have been launched on any stream.
void graph_detach(void *exec) {
    StreamBufs *b = nullptr;
    std::vector<void *> old;
    bool drop_set = false;
    {
        std::lock_guard<std::mutex> lk(g_bufs_mu);
        auto it = g_graph_bufs.find(exec);
        if (it == g_graph_bufs.end()) return;
        b = it->second;
        g_graph_bufs.erase(it);
        if (--b->graphs > 0) return;
        old.swap(b->retired);
        drop_set = b->released;
    }
    if (old.empty() && !drop_set) return;
    (void)hipDeviceSynchronize();
    for (void *p : old) (void)hipFree(p);
    if (drop_set) free_set(b);
}

static int rerun_ensure(hipStream_t s, size_t n) {
    StreamBufs &sb = stream_bufs(s);
    RerunBuf &g_rerun = sb.rr;
    if (n <= g_rerun.cap && g_rerun.block) return FT_OK;
    FT_TRY(drop_block(sb, s, g_rerun.block));
    g_rerun.block = nullptr;
    g_rerun.cap = 0;
    const size_t cap = n + n / 2 + 1024;
    const size_t tiles = cap / CP_TILE + 2;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bytes = al(cap) + al(cap * 4) + al(tiles * 4) + 256;
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return fail(FT_ENOMEM, "re-run buffer allocation failed");
    char *c = (char *)p;
    g_rerun.block = p;
    g_rerun.flags = (uint8_t *)c;
    g_rerun.list = (int32_t *)(c + al(cap));
    g_rerun.tiles = (uint32_t *)(c + al(cap) + al(cap * 4));
    g_rerun.count = (uint32_t *)(c + al(cap) + al(cap * 4) + al(tiles * 4));
    g_rerun.cap = cap;
    return FT_OK;
}

__global__ void fill_u8_kernel(uint8_t *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
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

// Flags -> ascending slot list, count on the device.
static int compact_flags(hipStream_t s, const RerunBuf &g_rerun, int n) {
    const unsigned nt = blocks_for((size_t)n, CP_TILE);
    hipLaunchKernelGGL(flag_count_kernel, dim3(nt), dim3(256), 0, s,
                       (const uint8_t *)g_rerun.flags, n, g_rerun.tiles);
    hipLaunchKernelGGL(flag_scan_kernel, dim3(1), dim3(1024), 0, s, g_rerun.tiles, (int)nt,
                       g_rerun.count);
    hipLaunchKernelGGL(flag_scatter_kernel, dim3(nt), dim3(256), 0, s,
                       (const uint8_t *)g_rerun.flags, n, (const uint32_t *)g_rerun.tiles,
                       g_rerun.list);
    FT_CHECK_LAUNCH();
    return FT_OK;
}

// Runs the statement with M = the register tier width for k (1, 4, 8, 16, 32).
#define FT_BY_M(K, ...)                                        \
    do {                                                       \
        if ((K) <= 1) { constexpr int M = 1; __VA_ARGS__; }         \
        else if ((K) <= 4) { constexpr int M = 4; __VA_ARGS__; }    \
        else if ((K) <= 8) { constexpr int M = 8; __VA_ARGS__; }    \
        else if ((K) <= 16) { constexpr int M = 16; __VA_ARGS__; }  \
        else { constexpr int M = 32; __VA_ARGS__; }                 \
    } while (0)

// Exact re-run of the listed slots.  Up to k = 128, one wave per listed query
// first (the wave-per-query tier, LISTED); it marks the entries it answers.
// Then the per-lane walk with a 128-entry stack, unseeded, lexicographic top-k
// (registers for k <= 32, LDS heap above), over the entries left.  The flagged
// queries of a seeded pass are few and scattered (0.2 % of C2's 10M): one
// lane each left the re-run latency-bound on a few hundred waves, 1.7 ms; one
// wave each fills the chip.
template <typename T, int MK>
static int launch_exact_listed(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq,
                               int k, int excl, const uint32_t *order, const RerunBuf &g_rerun,
                               uint64_t *res, typename CoordTraits<T>::D *d) {
    const QueryMap qm{order, g_rerun.list, g_rerun.count, nullptr};
    if (k <= 128) {
        const dim3 g(xcd_grid(std::min<unsigned>(blocks_for(nq, 4), 2048u))), b(256);
        if (k <= 64)
            hipLaunchKernelGGL((wave_knn_kernel<T, 1, MK, true>), g, b, 0, s, dt, qs, nq, k, excl,
                               qm, res, d, (uint64_t *)nullptr, g_rerun.list);
        else
            hipLaunchKernelGGL((wave_knn_kernel<T, 2, MK, true>), g, b, 0, s, dt, qs, nq, k, excl,
                               qm, res, d, (uint64_t *)nullptr, g_rerun.list);
        FT_CHECK_LAUNCH();
    }
    if (k <= 32) {
        const unsigned g = std::min<unsigned>(blocks_for(nq, LANE_BLOCK), 2048u);
        FT_BY_M(k, hipLaunchKernelGGL((lane_knn_reg_kernel<T, M, true, true, MK>), dim3(g),
                                      dim3(LANE_BLOCK), 0, s, dt, qs, nq, k, excl, qm, res, d,
                                      (uint64_t *)nullptr));
    } else {
        size_t lds = 0;
        const int lanes = lds_lanes<T>(k, LANE_STACK_CANON, lds);
        if (!lanes) return fail(FT_ELDS, "k too large for local memory: reduce k");
        const unsigned g = std::min<unsigned>(blocks_for(nq, lanes), 4096u);
        hipLaunchKernelGGL((lane_knn_lds_kernel<T, false, true, true, MK>), dim3(g),
                           dim3(LANE_BLOCK), lds, s, dt, qs, nq, k, excl, lanes, qm, res, d,
                           (uint64_t *)nullptr);
    }
    FT_CHECK_LAUNCH();
    return FT_OK;
}

template <typename T, int M, bool HEAP, bool SEEDED, bool PACKED, int MK>
static void launch_pk(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, const PacketArgs &a,
                      const QueryMap &qm, const typename CoordTraits<T>::D *seed, uint64_t *res,
                      typename CoordTraits<T>::D *d) {
    const dim3 g(xcd_grid(blocks_for(a.nq, a.width)));
    if constexpr (!HEAP && M == 32) {
        if (a.k > M) {
            hipLaunchKernelGGL((packet_knn_kernel<T, M, HEAP, SEEDED, PACKED, MK, true>), g,
                               dim3(64), 0, s, dt, qs, a, qm, seed, res, d);
            return;
        }
    }
    hipLaunchKernelGGL((packet_knn_kernel<T, M, HEAP, SEEDED, PACKED, MK>), g, dim3(64), 0, s, dt,
                       qs, a, qm, seed, res, d);
}

template <typename T, int M, bool HEAP, int MK>
static void launch_packet_m(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs,
                            const PacketArgs &a, const QueryMap &qm,
                            const typename CoordTraits<T>::D *seed, uint64_t *res,
                            typename CoordTraits<T>::D *d) {
    constexpr bool INTD = !CoordTraits<T>::fp;
    if (seed) launch_pk<T, M, HEAP, true, INTD, MK>(s, dt, qs, a, qm, seed, res, d);
    else launch_pk<T, M, HEAP, false, INTD, MK>(s, dt, qs, a, qm, seed, res, d);
}

// One canonical query launch: a direct pass (packet or per-lane walk) that
// flags the slots it cannot answer exactly (packed keys that saturate, seeds
// that were too small, per-lane stacks that would overflow), compaction of
// the flags, and the exact listed re-run.  No host round trip.
template <typename T, int MK>
static int run_canonical(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq, int k,
                         int excl, int algo, int bucket, int topk, int window, int width,
                         const uint32_t *order,
                         const typename CoordTraits<T>::D *seed, uint64_t *res,
                         typename CoordTraits<T>::D *d, uint64_t *stats, const RerunBuf *rr) {
    constexpr bool INTD = !CoordTraits<T>::fp;
    if (!rr) {  // the stream's own re-run buffers, grown to nq
        FT_TRY(rerun_ensure(s, (size_t)nq));
        rr = &stream_bufs(s).rr;
    }
    const RerunBuf &g_rerun = *rr;
    hipLaunchKernelGGL(fill_u8_kernel, dim3(std::min<unsigned>(blocks_for(nq, 256), 4096u)),
                       dim3(256), 0, s, g_rerun.flags, (size_t)nq);
    const QueryMap qm{order, nullptr, nullptr, g_rerun.flags};
    if (algo == FT_ALGO_PACKET) {
        PacketArgs a{};
        a.nq = nq;
        a.k = k;
        a.excl = excl;
        a.bucket = bucket;
        a.width = width;
        a.ibits = bit_length((uint64_t)dt.n);
        // A subset packet's leaves spread over about n / nq times more ranks:
        // its window grows with that spread, up to 4x (a 20 % subset of 20M
        // clustered points at k = 64: windows of 128 / 256 / 512 / 1024 took
        // 114.7 / 112.7 / 111.9 / 113.8 ms).
        const int spread = qs.leaf_ids ? (int)std::min<int64_t>(4, std::max<int64_t>(1, dt.n / std::max(nq, 1))) : 1;
        a.window = window * spread;
        a.stats = stats;
        // k above 32 with integer coordinates: multi-pass register tier by
        // default, LDS heap tiers on request.
        const bool heap = INTD && topk == FT_TOPK_LDS_HEAP;
        if (!heap) {
            FT_BY_M(k, (launch_packet_m<T, M, false, MK>(s, dt, qs, a, qm, seed, res, d)));
        } else if constexpr (INTD) {
            if (k <= 32) launch_packet_m<T, 32, true, MK>(s, dt, qs, a, qm, seed, res, d);
            else if (k <= 64) launch_packet_m<T, 64, true, MK>(s, dt, qs, a, qm, seed, res, d);
            else if (k <= 128) launch_packet_m<T, 128, true, MK>(s, dt, qs, a, qm, seed, res, d);
            else return fail(FT_EINVAL, "packet traversal: k above the heap tiers");
        }
    } else if (algo == FT_ALGO_WAVE) {
        const dim3 g(xcd_grid(blocks_for(nq, 4))), b(256);
#define FT_WAVE_LAUNCH(S_)                                                                      \
    hipLaunchKernelGGL((wave_knn_kernel<T, S_, MK>), g, b, 0, s, dt, qs, nq, k, excl, qm, res, d, \
                       stats, (int32_t *)nullptr)
        if (k <= 64) FT_WAVE_LAUNCH(1);
        else if (k <= 128) FT_WAVE_LAUNCH(2);
        else if (k <= 256) FT_WAVE_LAUNCH(4);
        else if (k <= 512) FT_WAVE_LAUNCH(8);
        else if (k <= WAVE_MAX_K) FT_WAVE_LAUNCH(16);
        else return fail(FT_EINVAL, "wave traversal: k above 1024");
#undef FT_WAVE_LAUNCH
    } else if (k <= 32) {
        const dim3 g(xcd_grid(blocks_for(nq, LANE_BLOCK))), b(LANE_BLOCK);
        if (algo == FT_ALGO_PLAIN)
            FT_BY_M(k, hipLaunchKernelGGL((lane_knn_reg_kernel<T, M, false, false, MK>), g, b, 0, s,
                                          dt, qs, nq, k, excl, qm, res, d, stats));
        else
            FT_BY_M(k, hipLaunchKernelGGL((lane_knn_reg_kernel<T, M, true, false, MK>), g, b, 0, s,
                                          dt, qs, nq, k, excl, qm, res, d, stats));
    } else {
        size_t lds = 0;
        const int lanes = lds_lanes<T>(k, STACK_DEPTH, lds);
        if (!lanes) return fail(FT_ELDS, "k too large for local memory: reduce k");
        const dim3 g(xcd_grid(blocks_for(nq, lanes))), b(LANE_BLOCK);
        if (algo == FT_ALGO_PLAIN)
            hipLaunchKernelGGL((lane_knn_lds_kernel<T, false, false, false, MK>), g, b, lds, s, dt,
                               qs, nq, k, excl, lanes, qm, res, d, stats);
        else
            hipLaunchKernelGGL((lane_knn_lds_kernel<T, false, true, false, MK>), g, b, lds, s, dt,
                               qs, nq, k, excl, lanes, qm, res, d, stats);
    }
    FT_CHECK_LAUNCH();
    FT_TRY(compact_flags(s, g_rerun, nq));
    if (stats) {
        hipLaunchKernelGGL(add_count_kernel, dim3(1), dim3(1), 0, s, stats,
                           (const uint32_t *)g_rerun.count);
        FT_CHECK_LAUNCH();
    }
    return launch_exact_listed<T, MK>(s, dt, qs, nq, k, excl, order, g_rerun, res, d);
}

template <typename T>
static int query_order(hipStream_t s, const DTree<T> &dt, const QSrc<T> &qs, int nq,
                       const uint32_t *&order) {
    StreamBufs &sb = stream_bufs(s);
    QueryOrderBuf &g_qorder = sb.qo;
    if (g_qorder.cap < (size_t)nq) {
        FT_TRY(drop_block(sb, s, g_qorder.keys));
        g_qorder.keys = nullptr;
        g_qorder.cap = 0;
        const size_t cap = (size_t)nq + (size_t)nq / 2 + 1024;
        if (hipMalloc(&g_qorder.keys, 2 * cap * 8) != hipSuccess)
            return fail(FT_ENOMEM, "query order allocation failed");
        g_qorder.sorted = g_qorder.keys + cap;
        g_qorder.cap = cap;
    }
    if ((size_t)nq > g_qorder.S.cap_n && g_qorder.S.block) {
        // the sort scratch grows: hand its block over here, so that the
        // sort's own growth allocates a new one and frees nothing
        FT_TRY(drop_block(sb, s, g_qorder.S.block));
        g_qorder.S.block = nullptr;
        g_qorder.S.cap_n = 0;
    }
    hipLaunchKernelGGL((query_key_kernel<T>), dim3(blocks_for(nq, 256)), dim3(256), 0, s, dt, qs,
                       nq, g_qorder.keys);
    FT_CHECK_LAUNCH();
    FT_TRY(radix_sort_u64(s, g_qorder.S, g_qorder.keys, (size_t)nq, g_qorder.sorted,
                          qs.qx ? 63 : 32));
    order = g_qorder.S.perm;
    return FT_OK;
}

template <typename T>
static int knn_dispatch(hipStream_t s, const ft_tree *tree, const QSrc<T> &qs, int nq, int k,
                        int excl, uint64_t *res, void *dists_v, const ft_query_opts *opts,
                        bool external, const RerunBuf *rr = nullptr) {
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
        size_t lds = 0;
        const int lanes = lds_lanes<T>(k, STACK_DEPTH, lds);
        if (!lanes) return fail(FT_ELDS, "k too large for local memory: reduce k");
        const QueryMap qm{nullptr, nullptr, nullptr, nullptr};
        const dim3 g(xcd_grid(blocks_for(nq, lanes))), b(LANE_BLOCK);
        if (algo == FT_ALGO_PLAIN)
            hipLaunchKernelGGL((lane_knn_lds_kernel<T, true, false, false, MK_ANY>), g, b, lds, s,
                               dt, qs, nq, k, excl, lanes, qm, res, dists, stats);
        else
            hipLaunchKernelGGL((lane_knn_lds_kernel<T, true, true, false, MK_ANY>), g, b, lds, s,
                               dt, qs, nq, k, excl, lanes, qm, res, dists, stats);
        FT_CHECK_LAUNCH();
        return FT_OK;
    }
    // Canonical mode.  Packet tiers: registers for k <= 32, passes of 32 up
    // to k = 128.  Above k = 32 the wave-per-query tier, except where packets
    // of integer queries share their walks best: self queries over all leaves
    // up to k = 128, and subsets of at least a fifth of the leaves up to
    // k = 64.  Measured on 10M clustered int64 points, passes against the
    // wave tier: self k = 33 / 64 / 128 37.6 / 47.3 / 103.3 against 86.9 /
    // 106.6 / 153.8 ms; a 20 % subset 15.7 / 19.4 / 41.2 against 18.7 / 22.7 /
    // 32.3 ms; 1M external queries 14.6 / 19.1 / 40.4 against 13.0 / 15.1 /
    // 20.1 ms.  20M clustered doubles: self k = 33 / 64 176 / 249 against
    // 172 / 213 ms, a 20 % subset k = 64 98 against 45 ms, 1M external
    // queries k = 64 79 against 15 ms.  Above k = 1024, the per-lane heap
    // walk.
    const bool packet_ok = k <= 128;  // above 32: passes of 32
    if (algo == FT_ALGO_AUTO) {
        const bool int_q = !CoordTraits<T>::fp && !external;
        const bool int_self = int_q && !qs.leaf_ids;
        const bool int_dense_subset = int_q && qs.leaf_ids && (int64_t)nq * 5 >= (int64_t)dt.n;
        if (k > WAVE_MAX_K) algo = FT_ALGO_GROUPED;
        else if (k <= 32 || (int_self && k <= 128) || (int_dense_subset && k <= 64))
            algo = FT_ALGO_PACKET;
        else algo = FT_ALGO_WAVE;
    }
    if (o.seed_bound && packet_ok) algo = FT_ALGO_PACKET;
    if (algo == FT_ALGO_PACKET && !packet_ok)
        algo = k <= WAVE_MAX_K ? FT_ALGO_WAVE : FT_ALGO_GROUPED;
    const D *seed = algo == FT_ALGO_PACKET ? (const D *)o.seed_bound : nullptr;
    const int bucket = o.leaf_bucket > 0 ? std::min(o.leaf_bucket, 64) : 64;
    const uint32_t *order = nullptr;
    if (algo == FT_ALGO_WAVE && k > WAVE_MAX_K) algo = FT_ALGO_GROUPED;
    if ((algo == FT_ALGO_PACKET || algo == FT_ALGO_WAVE) && (qs.qx || qs.leaf_ids) && nq > 64)
        FT_TRY(query_order<T>(s, dt, qs, nq, order));
    // auto: 192 leaves on each side (self k = 32 at 10M clustered points: 64 /
    // 128 / 192 / 256 / 384 took 19.17 / 18.93 / 18.80 / 18.77 / 18.81 ms;
    // 192 was also as fast or faster at k = 1 to 64, for subsets and for
    // external queries)
    const int window = o.window < 0 ? 0 : (o.window == 0 ? 192 : std::min(o.window, 1 << 20));
    // queries per packet: 64 (one per lane) unless 16 or 32 is asked for
    if (o.group_width != 0 && o.group_width != 16 && o.group_width != 32 && o.group_width != 64)
        return fail(FT_EINVAL, "group_width must be 0, 16, 32 or 64");
    const int width = o.group_width ? o.group_width : 64;
    switch (metric_kind<T>(dt.m)) {
    case MK_WRAP:
        return run_canonical<T, MK_WRAP>(s, dt, qs, nq, k, excl, algo, bucket, o.topk, window,
                                         width, order, seed, res, dists, stats, rr);
    case MK_OPEN:
        return run_canonical<T, MK_OPEN>(s, dt, qs, nq, k, excl, algo, bucket, o.topk, window,
                                         width, order, seed, res, dists, stats, rr);
    default:
        return fail(FT_EINVAL, "unsupported metric: periodic integer coordinates need B <= 31");
    }
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

// ---- non-allocating self-kNN (no reference counterpart) ----------------------
// For callers that cannot allocate (code under graph capture, serving loops):
// the canonical pipeline's re-run buffers (per-slot flags, the compacted slot
// list, per-tile counts, the list length) live in one file-scope device array
// instead of per-stream allocations.  Every call on a device shares that
// device's copy, which keeps the last call's contents; the flags are cleared
// by a kernel at the start of each call and the other fields are written
// before they are read.  Calls that overlap on one device race.
constexpr size_t NA_MAX_QUERIES = (size_t)1 << 24;
constexpr size_t na_al(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t NA_OFF_LIST = na_al(NA_MAX_QUERIES);
constexpr size_t NA_OFF_TILES = NA_OFF_LIST + na_al(NA_MAX_QUERIES * 4);
constexpr size_t NA_OFF_COUNT = NA_OFF_TILES + na_al((NA_MAX_QUERIES / CP_TILE + 2) * 4);
constexpr size_t NA_BYTES = NA_OFF_COUNT + 256;  // 80 MiB + 16.5 KiB
static __device__ __align__(256) uint8_t scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_na_mu;
static void *g_na_ptr[NA_MAX_DEVICES];

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array at its own address).
static int noalloc_scratch(void **out) {
    int dev = 0;
    FT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(FT_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_na_ptr[dev] = p;
    }
    *out = g_na_ptr[dev];
    return FT_OK;
}

size_t noalloc_max_queries() { return NA_MAX_QUERIES; }

int knn_noalloc_entry(hipStream_t s, const ft_tree *tree, int k, int excl, uint64_t *res,
                      void *dists, const ft_query_opts *opts) {
    const size_t nq = tree->num_leaves;
    if (nq > NA_MAX_QUERIES)
        return fail(FT_ERANGE, "ft_self_knn_query_noalloc: more than 16777216 queries (its "
                               "fixed scratch size); use ft_self_knn_query");
    void *p = nullptr;
    FT_TRY(noalloc_scratch(&p));
    char *c = (char *)p;
    RerunBuf rr;
    rr.flags = (uint8_t *)c;
    rr.list = (int32_t *)(c + NA_OFF_LIST);
    rr.tiles = (uint32_t *)(c + NA_OFF_TILES);
    rr.count = (uint32_t *)(c + NA_OFF_COUNT);
    rr.cap = NA_MAX_QUERIES;
    FT_DISPATCH_COORD(tree->coord_type, {
        const QSrc<T> qs{nullptr, nullptr, nullptr, nullptr};
        return knn_dispatch<T>(s, tree, qs, (int)nq, k, excl, res, dists, opts, false, &rr);
    });
    return FT_OK;
}

// Range query launch: queries in the order `order` (slot -> query), packet
// walk with the metric kind fixed at compile time.
template <typename T>
static int range_walk(hipStream_t s, const DTree<T> &dt, const T *qx, const T *qy, const T *qz,
                      const T *rmin, const T *rmax, int nq, const uint32_t *order, int32_t *res,
                      int32_t *counts, int max_res, const int64_t *row_ptr, int32_t *cols,
                      int64_t col_cap) {
    if (nq <= 0) return FT_OK;
    const dim3 g(xcd_grid(blocks_for(nq, 64))), b(64);
    constexpr int BUCKET = 32;
    switch (metric_kind<T>(dt.m)) {
    case MK_WRAP:
        hipLaunchKernelGGL((range_packet_kernel<T, MK_WRAP>), g, b, 0, s, dt, qx, qy, qz, rmin,
                           rmax, nq, order, BUCKET, res, counts, max_res, row_ptr, cols, col_cap);
        break;
    case MK_OPEN:
        hipLaunchKernelGGL((range_packet_kernel<T, MK_OPEN>), g, b, 0, s, dt, qx, qy, qz, rmin,
                           rmax, nq, order, BUCKET, res, counts, max_res, row_ptr, cols, col_cap);
        break;
    default:
        hipLaunchKernelGGL((range_packet_kernel<T, MK_ANY>), g, b, 0, s, dt, qx, qy, qz, rmin,
                           rmax, nq, order, BUCKET, res, counts, max_res, row_ptr, cols, col_cap);
        break;
    }
    FT_CHECK_LAUNCH();
    return FT_OK;
}

// Queries in Morton order of the points inside the root box.
template <typename T>
static int range_launch(hipStream_t s, const ft_tree &tree, const T *qx, const T *qy, const T *qz,
                        const T *rmin, const T *rmax, int nq, int32_t *res, int32_t *counts,
                        int max_res, const int64_t *row_ptr, int32_t *cols, int64_t col_cap) {
    const DTree<T> dt = make_dtree<T>(tree);
    const QSrc<T> qs{qx, qy, qz, nullptr};
    const uint32_t *order = nullptr;
    FT_TRY(query_order<T>(s, dt, qs, nq, order));
    return range_walk<T>(s, dt, qx, qy, qz, rmin, rmax, nq, order, res, counts, max_res, row_ptr,
                         cols, col_cap);
}

int range_entry(hipStream_t s, const ft_tree *tree, const void *qx, const void *qy,
                const void *qz, const void *rmin, const void *rmax, int nq, int32_t *res,
                int32_t *counts, int max_res, const int64_t *row_ptr, int32_t *cols,
                int64_t col_cap) {
    int rc = FT_OK;
    FT_DISPATCH_COORD(tree->coord_type,
                      rc = range_launch<T>(s, *tree, (const T *)qx, (const T *)qy, (const T *)qz,
                                           (const T *)rmin, (const T *)rmax, nq, res, counts,
                                           max_res, row_ptr, cols, col_cap));
    return rc;
}

// ---------------------------------------------------------------------------
// CSR range query with the columns in one walk (no reference counterpart).
// The walk writes every query's hit count and its first CSR_POOL hits into a
// fixed-stride pool (the fixed-stride range_query output); after the scan of
// the counts, the pool rows are copied to their CSR rows, and only queries
// with more than CSR_POOL hits are walked again, writing their whole rows.
// A walk meets a query's hits in the same (ascending rank) order every time,
// so the rows equal those of a counting walk followed by a filling walk.
// ---------------------------------------------------------------------------
constexpr int CSR_POOL = 128;

// Rows of at most CSR_POOL hits from the pool into cols.  One wave per group
// of 64 rows (grid-stride over groups): the lanes load the group's counts and
// offsets in one pass, then the wave copies four rows at a time, their pool
// loads issued before their stores (one row at a time, each waiting on its
// count and offset, ran at 1.9 TB/s).
__global__ __launch_bounds__(256) void csr_copy_kernel(const int32_t *__restrict__ pool,
                                                       const int32_t *__restrict__ counts,
                                                       const int64_t *__restrict__ row_ptr, int nq,
                                                       int32_t *__restrict__ cols,
                                                       int64_t col_cap) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t g = w; g * 64 < nq; g += nw) {
        const int64_t q0 = g * 64;
        const int64_t ql = q0 + lane;
        // rows past nq count 0 hits, so nothing below touches them
        const int cl = ql < nq ? counts[ql] : 0;
        const int64_t bl = ql < nq ? row_ptr[ql] : 0;
        for (int r = 0; r < 64; r += 4) {
            int32_t v0[4], v1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = __shfl(cl, r + u, 64);
                const size_t off = (size_t)(q0 + r + u) * CSR_POOL;
                const bool fits = c <= CSR_POOL;
                v0[u] = (fits && lane < c) ? pool[off + lane] : 0;
                v1[u] = (fits && lane + 64 < c) ? pool[off + lane + 64] : 0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = __shfl(cl, r + u, 64);
                const int64_t base = __shfl(bl, r + u, 64);
                if (c > CSR_POOL) continue;
                if (lane < c && base + lane < col_cap) cols[base + lane] = v0[u];
                if (lane + 64 < c && base + lane + 64 < col_cap) cols[base + lane + 64] = v1[u];
            }
        }
    }
}

// flags[slot] = the slot's query has more hits than the pool holds.
__global__ void csr_overflow_flags_kernel(const uint32_t *__restrict__ order,
                                          const int32_t *__restrict__ counts, int nq,
                                          uint8_t *__restrict__ flags) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
        const int q = order ? (int)order[i] : i;
        flags[i] = counts[q] > CSR_POOL ? 1 : 0;
    }
}

// order2[i] = the query of the i-th listed slot.
__global__ void csr_listed_order_kernel(const uint32_t *__restrict__ order,
                                        const int32_t *__restrict__ list,
                                        const uint32_t *__restrict__ count,
                                        uint32_t *__restrict__ order2) {
    const uint32_t m = *count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        order2[i] = order ? order[list[i]] : (uint32_t)list[i];
}

int scan_entry(hipStream_t s, const int32_t *counts, int nq, int64_t *row_ptr);

template <typename T>
static int range_csr_t(hipStream_t s, const ft_tree &tree, const T *qx, const T *qy, const T *qz,
                       const T *rmin, const T *rmax, int nq, int64_t *row_ptr,
                       int32_t *(*cols_for)(int64_t total, void *ctx), void *ctx,
                       int64_t col_cap, int64_t *total_out) {
    const DTree<T> dt = make_dtree<T>(tree);
    const QSrc<T> qs{qx, qy, qz, nullptr};
    const uint32_t *order = nullptr;
    FT_TRY(query_order<T>(s, dt, qs, nq, order));
    int32_t *counts = nullptr, *pool = nullptr;
    uint32_t *order2 = nullptr;
    FT_HIP(hipMallocAsync((void **)&counts, (size_t)nq * 4, s));
    // Without room for the pool: a counting walk now and a filling walk later.
    if (hipMallocAsync((void **)&pool, (size_t)nq * CSR_POOL * 4, s) != hipSuccess) {
        (void)hipGetLastError();
        pool = nullptr;
    }
    auto release = [&]() {
        (void)hipFreeAsync(counts, s);
        if (pool) (void)hipFreeAsync(pool, s);
        if (order2) (void)hipFreeAsync(order2, s);
    };
    int rc = range_walk<T>(s, dt, qx, qy, qz, rmin, rmax, nq, order, pool, counts,
                           pool ? CSR_POOL : 0, nullptr, nullptr, 0);
    if (rc == FT_OK) rc = scan_entry(s, counts, nq, row_ptr);
    int64_t total = 0;
    if (rc == FT_OK && hipMemcpyAsync(&total, row_ptr + nq, 8, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = fail(FT_EHIP, "CSR range query: reading the total failed");
    if (rc == FT_OK && hipStreamSynchronize(s) != hipSuccess)
        rc = fail(FT_EHIP, "CSR range query: synchronize failed");
    if (rc == FT_OK) *total_out = total;
    if (rc == FT_OK && total > col_cap)
        rc = fail(FT_ERANGE, "CSR column capacity is smaller than the number of hits");
    int32_t *cols = nullptr;
    if (rc == FT_OK) {
        cols = cols_for(total, ctx);
        if (!cols) rc = fail(FT_ENOMEM, "CSR range query: column buffer");
    }
    if (rc == FT_OK && !pool) {
        rc = range_walk<T>(s, dt, qx, qy, qz, rmin, rmax, nq, order, nullptr, nullptr, 0, row_ptr,
                           cols, col_cap);
        release();
        return rc;
    }
    if (rc == FT_OK) {
        hipLaunchKernelGGL(csr_copy_kernel, dim3(std::min<unsigned>(blocks_for((nq + 63) / 64, 4), 16384u)),
                           dim3(256), 0, s, (const int32_t *)pool, (const int32_t *)counts,
                           (const int64_t *)row_ptr, nq, cols, col_cap);
        rc = rerun_ensure(s, (size_t)nq);
    }
    RerunBuf &g_rerun = stream_bufs(s).rr;
    uint32_t n_over = 0;
    if (rc == FT_OK) {
        hipLaunchKernelGGL(csr_overflow_flags_kernel, dim3(std::min<unsigned>(blocks_for(nq, 256), 4096u)),
                           dim3(256), 0, s, order, (const int32_t *)counts, nq, g_rerun.flags);
        rc = compact_flags(s, g_rerun, nq);
    }
    if (rc == FT_OK &&
        (hipMemcpyAsync(&n_over, g_rerun.count, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        rc = fail(FT_EHIP, "CSR range query: reading the overflow count failed");
    if (rc == FT_OK && n_over > 0) {
        if (hipMallocAsync((void **)&order2, (size_t)n_over * 4, s) != hipSuccess) {
            rc = fail(FT_ENOMEM, "CSR range query: overflow list allocation failed");
        } else {
            hipLaunchKernelGGL(csr_listed_order_kernel, dim3(std::min<unsigned>(blocks_for(n_over, 256), 4096u)),
                               dim3(256), 0, s, order, (const int32_t *)g_rerun.list,
                               (const uint32_t *)g_rerun.count, order2);
            if (hipGetLastError() != hipSuccess) rc = fail(FT_EHIP, "CSR range query: launch failed");
        }
        if (rc == FT_OK) {
            rc = range_walk<T>(s, dt, qx, qy, qz, rmin, rmax, (int)n_over, order2, nullptr,
                               nullptr, 0, row_ptr, cols, col_cap);
        }
    }
    release();
    return rc;
}

int range_csr_entry(hipStream_t s, const ft_tree *tree, const void *qx, const void *qy,
                    const void *qz, const void *rmin, const void *rmax, int nq, int64_t *row_ptr,
                    int32_t *(*cols_for)(int64_t total, void *ctx), void *ctx, int64_t col_cap,
                    int64_t *total) {
    int rc = FT_OK;
    FT_DISPATCH_COORD(tree->coord_type,
                      rc = range_csr_t<T>(s, *tree, (const T *)qx, (const T *)qy, (const T *)qz,
                                          (const T *)rmin, (const T *)rmax, nq, row_ptr, cols_for,
                                          ctx, col_cap, total));
    return rc;
}

int scan_entry(hipStream_t s, const int32_t *counts, int nq, int64_t *row_ptr) {
    const int nb = (nq + 1023) / 1024;
    int64_t *sums = nullptr;
    FT_HIP(hipMallocAsync((void **)&sums, (size_t)(nb > 0 ? nb : 1) * 8, s));
    if (nq > 0) {
        hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(256), 0, s, counts, nq, row_ptr, sums);
        hipLaunchKernelGGL(scan_totals_kernel, dim3(1), dim3(1024), 0, s, sums, nb, row_ptr, nq);
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
