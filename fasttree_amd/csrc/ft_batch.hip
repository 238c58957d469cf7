// ft_batch.hip -- many trees of uneven size and query requests of mixed k in
// one hipGraph launch (no reference counterpart; its closest analogue is the
// compile-time width / heap-tier dispatch of src/hlbvh.hpp:1433-1453,
// 1841-1927).
//
// One step of a batch:
//   1. builds every tree (build_bvh) on parallel graph branches, each branch
//      with its own build scratch;
//   2. sorts all query entries by (request, Morton key of the query inside the
//      tree's root box), so each wavefront's 64 queries are neighbours of one
//      request;
//   3. runs one persistent kernel per tier (register top-k of width M, LDS
//      heap of capacity M, or per-lane LDS heaps) as parallel graph nodes;
//      each drains its own work list of (request, packet) items over all
//      trees, largest measured cost first;
//   4. re-runs exactly, one lane per query, every query whose packed keys
//      could not order its results.
// The autotuner times candidate settings per (tree-size class, k tier) on the
// bound data -- tier, leaf bucket and packet width (queries per wavefront
// walk: 64, 32 or 16, the reference's sub-group width) for the queries; per size class the
// radix width of the builds' sort; the branch count of the builds -- and
// keeps the fastest.  Results are canonical (ascending (d2,
// index)), so every setting gives the same output.
#include "ft_traverse.hpp"

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace ft {

template <typename T>
int build_bvh_t(hipStream_t s, ft_tree *t, Scratch &S, const T *x, const T *y, const T *z,
                const uint32_t *id, const int8_t *ghost, size_t n, const T *bbox6,
                const uint64_t *keys);
int radix_sort_u64(hipStream_t s, Scratch &S, const uint64_t *keys_in, size_t n,
                   uint64_t *sorted_out, int key_bits);
int scratch_ensure_public(Scratch &S, size_t n);

// A request as the kernels see it (device copy).
struct BatchReq {
    int32_t tree, k;
    int32_t entry_first, num;  // entries [entry_first, entry_first + num): sorted positions too
    int32_t bucket, variant;   // tuned leaf bucket; tier kernel that runs it
    int32_t width, pad;        // queries per item (packet tiers: 16, 32 or 64)
    int64_t query_first;       // first query point
    int64_t result_first;      // first result slot
};

template <typename T> struct BatchView {
    const DTree<T> *trees;
    const BatchReq *reqs;
    int nreq;
    const uint32_t *order;  // sorted position -> entry index
    const T *qx, *qy, *qz;
    uint64_t *res;
    typename CoordTraits<T>::D *dists;
    int32_t *rerun;  // sorted positions to re-run exactly
    uint32_t *rerun_count;
    uint32_t rerun_cap;  // entries the list holds (appends beyond it are dropped)
    uint64_t *stats;
};

// Tier kernels ("variants").  PK_M: packet walk with the register top-k of
// width M; HP_M: packet walk with an LDS heap of capacity M (integer metrics);
// LANE: per-lane DFS with LDS heaps, one query per lane (any k, any metric).
// PASSES: the 32-entry register tier in passes of 32 for k above 32 (integer
// metrics).  WAVE_C: one wavefront per query with a sorted list of C entries
// across its lanes (k up to C; wave_walk).
enum Variant : int {
    V_PK1, V_PK4, V_PK8, V_PK16, V_PK32, V_HP32, V_HP64, V_HP128, V_LANE, V_PASSES,
    V_WAVE64, V_WAVE128, V_WAVE256, V_COUNT
};
static const char *const kVariantName[V_COUNT] = {
    "packet-reg1",   "packet-reg4",    "packet-reg8", "packet-reg16",       "packet-reg32",
    "packet-heap32", "packet-heap64", "packet-heap128", "lane-heap", "packet-reg32-passes",
    "wave-merge64", "wave-merge128", "wave-merge256"};
static bool is_wave(int v) { return v >= V_WAVE64 && v <= V_WAVE256; }
static int wave_cap(int v) { return 64 << (v - V_WAVE64); }

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
// Entry -> composite sort key: request index on top, then the Morton key of the
// query quantised to qb bits per axis inside the tree's root box.  Only the
// grouping matters (queries of one packet should be spatial neighbours).
template <typename T>
__global__ __launch_bounds__(256) void batch_key_kernel(BatchView<T> b, int64_t n_entries,
                                                        int rbits, int qb,
                                                        uint64_t *__restrict__ keys) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_entries) return;
    int lo = 0, hi = b.nreq - 1;  // last request with entry_first <= e
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.reqs[mid].entry_first <= e) lo = mid;
        else hi = mid - 1;
    }
    const BatchReq q = b.reqs[lo];
    const DTree<T> &t = b.trees[q.tree];
    const int64_t qi = q.query_first + (e - q.entry_first);
    const double side = (double)((1u << qb) - 1);
    auto quant = [&](T v, T lo_, T hi_) -> uint64_t {
        const double ext = (double)hi_ - (double)lo_;
        double u = ext > 0.0 ? ((double)v - (double)lo_) / ext : 0.0;
        u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
        return (uint64_t)(u * side);
    };
    const uint64_t ix = quant(b.qx[qi], t.mnx[0], t.mxx[0]);
    const uint64_t iy = quant(b.qy[qi], t.mny[0], t.mxy[0]);
    const uint64_t iz = quant(b.qz[qi], t.mnz[0], t.mxz[0]);
    uint64_t m = 0;
    for (int bit = 0; bit < qb; ++bit) {
        m |= ((ix >> bit) & 1ull) << (3 * bit + 2);
        m |= ((iy >> bit) & 1ull) << (3 * bit + 1);
        m |= ((iz >> bit) & 1ull) << (3 * bit);
    }
    keys[e] = ((uint64_t)lo << (64 - rbits)) | m;
}

__global__ void batch_reset_kernel(uint32_t *ctr, int n) {
    if ((int)threadIdx.x < n) ctr[threadIdx.x] = 0u;
}

// Work queue of a tier kernel: its item list is split into FT_XCDS contiguous
// chunks [bnd[c], bnd[c+1]) of about equal cost, one per XCD (the hardware
// hands workgroup b to XCD b % 8), so each XCD's L2 serves a contiguous run of
// requests.  A wave takes the next item of its own XCD's chunk and, once that
// is drained, of the following chunks; -1 when every chunk is drained, which
// every wave reaches.
__device__ inline int grab_item(uint32_t *ctr, const uint32_t *bnd) {
    int it = -1;
    if ((threadIdx.x & 63) == 0) {
        const int x = (int)(blockIdx.x % FT_XCDS);
        for (int t = 0; t < FT_XCDS; ++t) {
            const int cidx = (x + t) % FT_XCDS;
            const uint32_t lo = bnd[cidx], hi = bnd[cidx + 1];
            if (lo >= hi) continue;
            const uint32_t i = lo + atomicAdd(&ctr[cidx], 1u);
            if (i < hi) {
                it = (int)i;
                break;
            }
        }
    }
    return uni(__shfl(it, 0, 64));
}

// Packet tiers.  HEAP: LDS heap of capacity M, else the register top-k;
// MULTI: the register top-k in passes of M (a separate instantiation, so the
// single-pass tiers keep their register budget).  Each packet first scans
// BATCH_WINDOW leaves on each side of a run near its queries.
constexpr int BATCH_WINDOW = 128;

template <typename T, int M, bool HEAP, int MK, bool MULTI = false>
__global__ __launch_bounds__(64) void batch_packet_kernel(BatchView<T> b,
                                                          const uint2 *__restrict__ items,
                                                          const uint32_t *__restrict__ bnd,
                                                          uint32_t *__restrict__ ctr) {
    using D = typename CoordTraits<T>::D;
    constexpr bool PACKED = !CoordTraits<T>::fp;
    static_assert(!HEAP || PACKED, "LDS heap tier needs packed integer keys");
    using H = typename std::conditional<HEAP, TopKHeapLds<M>,
                                        typename TopKSel<D, M, PACKED>::type>::type;
    __shared__ int32_t stk[PK_STACK];
    [[maybe_unused]] uint64_t *heap_mem = nullptr;
    if constexpr (HEAP) {
        __shared__ uint64_t heap[M * 64];
        heap_mem = heap;
    }
    const int lane = threadIdx.x;
    PacketCounters c;
    while (true) {
        const int it = grab_item(ctr, bnd);
        if (it < 0) break;
        const uint2 w = items[it];
        const BatchReq q = b.reqs[uni((int)w.x)];
        const int j = uni((int)w.y);
        const DTree<T> t = b.trees[q.tree];
        const int k = q.k;
        const int W = q.width;
        const int e = j * W + lane;
        const bool active = lane < W && e < q.num;
        int local = 0;
        T px = T(0), py = T(0), pz = T(0);
        if (active) {
            local = (int)b.order[q.entry_first + e] - q.entry_first;
            const int64_t qi = q.query_first + local;
            px = b.qx[qi];
            py = b.qy[qi];
            pz = b.qz[qi];
        }
        // window pre-scan around the leaf a greedy descent reaches for the
        // packet's middle query (results do not depend on it)
        const int src = (j * W + W / 2 < q.num) ? W / 2 : 0;
        const int r0 = run_start(descend_rank<T, MK>(t, rl(px, src), rl(py, src), rl(pz, src)), t.n);
        const int wlo = r0 - BATCH_WINDOW > 0 ? r0 - BATCH_WINDOW : 0;
        const int whi = r0 + 64 + BATCH_WINDOW < t.n ? r0 + 64 + BATCH_WINDOW : t.n;
        bool redo = false;
        static_assert(!MULTI || !HEAP, "passes need the register tier");
        if constexpr (MULTI) {
            // k above M: passes of M, each above the previous pass's last key.
            typename H::Floor lo = H::no_floor();
            for (int base = 0; base < k; base += M) {
                const int kk = k - base < M ? k - base : M;
                H h;
                if constexpr (PACKED) h.init(kk, ~0ull, 32 - __clz(t.n));
                else h.init(kk, dist_max<D>());
                // one walk instance for every pass (the first with floor 0)
                packet_walk<T, H, MK, true>(t, q.bucket, active, px, py, pz, -1, h, stk, c, wlo,
                                            whi, r0, lo);
                if (active) {
                    if constexpr (PACKED) redo = redo || h.needs_exact(kk);
                    if (!redo)
                        h.store_part(kk, b.res + q.result_first, b.dists + q.result_first,
                                     (size_t)local, k, base);
                }
                lo = h.next_floor();
            }
        } else {
            H h;
            if constexpr (HEAP) h.key = heap_mem;
            if constexpr (PACKED) h.init(k, ~0ull, 32 - __clz(t.n));
            else h.init(k, dist_max<D>());
            packet_walk<T, H, MK>(t, q.bucket, active, px, py, pz, -1, h, stk, c, wlo, whi, r0);
            if (active) {
                if constexpr (PACKED) redo = h.needs_exact(k);
                if (!redo)
                    h.store(k, b.res + q.result_first, b.dists + q.result_first, (size_t)local);
            }
        }
        if (active && redo) {
            const uint32_t s2 = atomicAdd(b.rerun_count, 1u);
            if (s2 < b.rerun_cap) b.rerun[s2] = q.entry_first + e;
        }
    }
    if (lane == 0 && b.stats) {
        atomicAdd((unsigned long long *)&b.stats[0], (unsigned long long)c.nodes);
        atomicAdd((unsigned long long *)&b.stats[1], (unsigned long long)c.leaves);
    }
    wave_add_stats(b.stats, 4, c.inserts);
}

// Per-lane tier: `lanes` queries per item (item j covers entries
// [j*lanes, (j+1)*lanes) of its request), LDS heap rows of capacity kcap.
template <typename T, int MK>
__global__ __launch_bounds__(64) void batch_lane_kernel(BatchView<T> b,
                                                        const uint2 *__restrict__ items,
                                                        const uint32_t *__restrict__ bnd,
                                                        uint32_t *__restrict__ ctr, int lanes,
                                                        int kcap) {
    using D = typename CoordTraits<T>::D;
    extern __shared__ __align__(16) unsigned char lds[];
    D *hd = (D *)lds;
    int32_t *hi = (int32_t *)(hd + (size_t)kcap * lanes);
    int32_t *stk = hi + (size_t)kcap * lanes;
    const int lane = threadIdx.x;
    uint64_t visits = 0, drops = 0, inserts = 0;
    while (true) {
        const int it = grab_item(ctr, bnd);
        if (it < 0) break;
        const uint2 w = items[it];
        const BatchReq q = b.reqs[uni((int)w.x)];
        const int j = uni((int)w.y);
        const DTree<T> t = b.trees[q.tree];
        const int e = j * lanes + lane;
        const bool active = lane < lanes && e < q.num;
        int local = 0;
        T px = T(0), py = T(0), pz = T(0);
        if (active) {
            local = (int)b.order[q.entry_first + e] - q.entry_first;
            const int64_t qi = q.query_first + local;
            px = b.qx[qi];
            py = b.qy[qi];
            pz = b.qz[qi];
        }
        LdsHeap<D, false> h{hd, hi, lanes, lane, 0};
        lane_heap_walk<T, false, true, LANE_STACK_CANON, MK>(t, q.k, active, px, py, pz, -1, h,
                                                             stk, lanes, lane, visits, drops,
                                                             inserts);
        if (active) {
            h.sort();
            uint64_t *res = b.res + q.result_first + (size_t)local * q.k;
            D *dst = b.dists + q.result_first + (size_t)local * q.k;
            for (int s = 0; s < q.k; ++s) {
                const bool ok = s < h.count && h.ii(s) >= 0;
                res[s] = ok ? (uint64_t)(uint32_t)h.ii(s) : ~0ull;
                dst[s] = ok ? h.dd(s) : missing_dist<D>();
            }
        }
    }
    wave_add_stats(b.stats, 0, visits);
    wave_add_stats(b.stats, 4, inserts);
}

// Wave tier: one wavefront per query, `width` consecutive sorted queries of a
// request per item, each answered by wave_walk with a list of 64 S entries,
// its first leaves around the leaf a greedy descent reaches.
template <typename T, int S, int MK>
__global__ __launch_bounds__(64) void batch_wave_kernel(BatchView<T> b,
                                                        const uint2 *__restrict__ items,
                                                        const uint32_t *__restrict__ bnd,
                                                        uint32_t *__restrict__ ctr) {
    __shared__ WaveBuf wbuf;
    const int lane = threadIdx.x;
    WaveCounters c;
    while (true) {
        const int it = grab_item(ctr, bnd);
        if (it < 0) break;
        const uint2 w = items[it];
        const BatchReq q = b.reqs[uni((int)w.x)];
        const int j = uni((int)w.y);
        const DTree<T> t = b.trees[q.tree];
        const int e_end = (j + 1) * q.width < q.num ? (j + 1) * q.width : q.num;
        for (int e = j * q.width; e < e_end; ++e) {
            const int local = uni((int)b.order[q.entry_first + e] - q.entry_first);
            const int64_t qi = q.query_first + local;
            const T px = b.qx[qi], py = b.qy[qi], pz = b.qz[qi];
            const int anchor = descend_rank_v<T, MK>(t, px, py, pz);
            WaveList<S> L;
            L.init(q.k);
            if (wave_walk<T, S, MK>(t, px, py, pz, -1, anchor, L, c, &wbuf)) {
                if (lane == 0) {
                    const uint32_t s2 = atomicAdd(b.rerun_count, 1u);
                    if (s2 < b.rerun_cap) b.rerun[s2] = q.entry_first + e;
                }
            } else {
                L.store(q.k, b.res + q.result_first, b.dists + q.result_first, (size_t)local,
                        lane);
            }
        }
    }
    if (lane == 0 && b.stats) {
        atomicAdd((unsigned long long *)&b.stats[0], (unsigned long long)c.nodes);
        atomicAdd((unsigned long long *)&b.stats[1], (unsigned long long)c.leaves);
        atomicAdd((unsigned long long *)&b.stats[4], (unsigned long long)c.inserts);
    }
}

// Exact re-run of listed sorted positions: one lane per query, each lane on its
// own tree (found by binary search over the requests), lexicographic LDS heap.
template <typename T>
__global__ __launch_bounds__(64) void batch_exact_kernel(BatchView<T> b, int lanes, int kcap) {
    using D = typename CoordTraits<T>::D;
    extern __shared__ __align__(16) unsigned char lds[];
    D *hd = (D *)lds;
    int32_t *hi = (int32_t *)(hd + (size_t)kcap * lanes);
    int32_t *stk = hi + (size_t)kcap * lanes;
    const int lane = threadIdx.x;
    const int nlist = (int)min(*b.rerun_count, b.rerun_cap);
    uint64_t visits = 0, drops = 0, inserts = 0;
    for (int base = blockIdx.x * lanes; base < nlist; base += gridDim.x * lanes) {
        const int slot = base + lane;
        const bool active = lane < lanes && slot < nlist;
        int r = 0, local = 0;
        if (active) {
            const int pos = b.rerun[slot];
            int lo = 0, hi2 = b.nreq - 1;
            while (lo < hi2) {
                const int mid = (lo + hi2 + 1) >> 1;
                if (b.reqs[mid].entry_first <= pos) lo = mid;
                else hi2 = mid - 1;
            }
            r = lo;
            local = (int)b.order[pos] - b.reqs[r].entry_first;
        }
        const BatchReq q = b.reqs[r];
        const DTree<T> t = b.trees[q.tree];
        T px = T(0), py = T(0), pz = T(0);
        if (active) {
            const int64_t qi = q.query_first + local;
            px = b.qx[qi];
            py = b.qy[qi];
            pz = b.qz[qi];
        }
        LdsHeap<D, false> h{hd, hi, lanes, lane, 0};
        lane_heap_walk<T, false, false, LANE_STACK_CANON, MK_ANY>(t, active ? q.k : 0, active, px,
                                                                  py, pz, -1, h, stk, lanes, lane,
                                                                  visits, drops, inserts);
        if (active) {
            h.sort();
            uint64_t *res = b.res + q.result_first + (size_t)local * q.k;
            D *dst = b.dists + q.result_first + (size_t)local * q.k;
            for (int s = 0; s < q.k; ++s) {
                const bool ok = s < h.count && h.ii(s) >= 0;
                res[s] = ok ? (uint64_t)(uint32_t)h.ii(s) : ~0ull;
                dst[s] = ok ? h.dd(s) : missing_dist<D>();
            }
        }
    }
}

__global__ void batch_count_kernel(uint64_t *stats, const uint32_t *count) {
    if (stats) stats[3] += *count;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

static int size_class(size_t n) {
    int c = 0;
    for (size_t lim = 40000; c < 3 && n >= lim; lim *= 4) ++c;
    return c;  // < 40k, < 160k, < 640k, larger
}
static int k_tier(int k) {
    const int lim[8] = {1, 4, 8, 16, 32, 64, 128, 256};
    for (int i = 0; i < 8; ++i)
        if (k <= lim[i]) return i;
    return 7;
}
static const int kTierCap[8] = {1, 4, 8, 16, 32, 64, 128, 256};
constexpr int N_CLASS = 4, N_KTIER = 8;

struct GroupTune {
    int variant = -1, bucket = 32, width = 64;
    double ms_per_item = 0.0;  // measured cost of one item (load balancing)
    bool tuned = false;
    bool pinned = false;  // set by the caller (ft_batch_set_tier): the tuner keeps it
};
constexpr int WAVE_ITEM = 8;  // wave tiers: queries per item
constexpr int QUERY_CHAINS = 4;  // parallel chains of tier kernels in the query phase

}  // namespace ft

struct ft_batch {
    int32_t coord_type = 0;
    // trees
    std::vector<ft_tree *> trees;
    std::vector<const void *> bx, by, bz;
    std::vector<size_t> bn;
    bool build = false;
    // requests
    std::vector<ft_batch_request> reqs;
    const void *qx = nullptr, *qy = nullptr, *qz = nullptr;
    uint64_t *results = nullptr;
    void *dists = nullptr;
    int64_t n_entries = 0;
    int kmax = 0;
    uint64_t *stats = nullptr;
    // tuning
    ft::GroupTune tune[ft::N_CLASS][ft::N_KTIER];
    int radix_bits[ft::N_CLASS] = {0, 0, 0, 0};  // sort digit width per tree-size class (0: auto)
    int branches = 4;
    bool tuned = false;
    // device state
    void *dev_block = nullptr;
    size_t dev_bytes = 0;
    void *dtrees = nullptr;
    ft::BatchReq *dreqs = nullptr;
    uint2 *ditems = nullptr;
    uint32_t *dctr = nullptr;  // [V_COUNT][FT_XCDS] queue counters, then the re-run count
    uint32_t *dbnd = nullptr;  // [V_COUNT][FT_XCDS + 1] chunk bounds, then 9 for the tuner
    int32_t *drerun = nullptr;
    uint64_t *dkeys = nullptr, *dsorted = nullptr;
    int item_first[ft::V_COUNT] = {}, item_count[ft::V_COUNT] = {};
    double var_cost[ft::V_COUNT] = {};  // estimated cost of each tier's item list
    int lane_width = 64;
    int mk = 0;  // metric kind shared by every tree (MK_ANY when they differ)
    ft::Scratch sort_scratch;
    std::vector<ft::Scratch> build_scratch;
    // segmented build: every tree in one launch of each build kernel
    int seg_mode = -1;  // -1 untuned (segmented when possible), 0 per-tree builds, 1 segmented
    bool seg_ok = false;  // the trees share one key configuration (checked at upload)
    void *seg_block = nullptr;
    size_t seg_bytes = 0;
    ft::SegBuild seg{};
    ft::Scratch seg_scratch;
    std::vector<hipStream_t> side;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
    hipStream_t cap = nullptr;  // capture stream (the null stream cannot be captured)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;  // device tables need a refresh
};

namespace ft {

int fail(int code, const std::string &msg);

// LDS lanes of the per-lane tier for heap capacity kcap (canonical walk).
template <typename T> static int lane_tier_lanes(int kcap, size_t &lds) {
    using D = typename CoordTraits<T>::D;
    const size_t per_lane = (size_t)kcap * (sizeof(D) + 4) + (size_t)LANE_STACK_CANON * 4;
    int lanes = 64;
    while (lanes > 1 && (size_t)lanes * per_lane > 64 * 1024) lanes >>= 1;
    lds = (size_t)lanes * per_lane;
    return lds <= 64 * 1024 ? lanes : 0;
}

static void free_graph(ft_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static int ensure_streams(ft_batch *b, int nb) {
    while ((int)b->side.size() < nb) {
        hipStream_t st;
        FT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->side.push_back(st);
        hipEvent_t ev;
        FT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->ev_join.push_back(ev);
    }
    if (!b->ev_fork) FT_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    return FT_OK;
}

// Candidate variants for a group (class, k tier) and metric; kmax = the
// largest k of the group's requests (tier 7 holds every k above 128).
static std::vector<int> candidates(int ktier, bool fp, int kmax) {
    std::vector<int> v;
    const int cap = kTierCap[ktier];
    if (cap <= 32) v.push_back(V_PK1 + (ktier <= 4 ? ktier : 4));
    else if (cap <= 128) v.push_back(V_PASSES);  // passes of 32
    if (!fp) {
        if (cap <= 32) v.push_back(V_HP32);
        else if (cap <= 64) v.push_back(V_HP64);
        else if (cap <= 128) v.push_back(V_HP128);
    }
    v.push_back(V_LANE);
    if (cap > 32 && kmax <= 256) v.push_back(cap <= 64 ? V_WAVE64 : (cap <= 128 ? V_WAVE128 : V_WAVE256));
    return v;
}

static bool is_packet(int v) { return v != V_LANE && !is_wave(v); }
// Queries per item of a group's setting.
static int item_width(int v, int width, int lane_width) {
    return is_packet(v) ? width : (is_wave(v) ? width : lane_width);
}

// Default settings: register packets (in passes of 32 above k = 32 for integer
// metrics), per-lane heaps for larger k with FP metrics; bucket 32.
static void default_tuning(ft_batch *b) {
    const bool fp = b->coord_type == FT_COORD_F32 || b->coord_type == FT_COORD_F64;
    for (int c = 0; c < N_CLASS; ++c)
        for (int t = 0; t < N_KTIER; ++t) {
            GroupTune &g = b->tune[c][t];
            if (g.tuned || g.pinned) continue;
            g.variant = candidates(t, fp, kTierCap[t]).front();
            g.bucket = 32;
            g.width = 64;
        }
}

// Tables of the segmented build (trees, sort tiles, assembly chunks) and its
// buffers, sized here so that the step allocates nothing.  The segmented build
// needs every tree to share the curve and bits per dimension, and fewer than
// 2^31 points in all; otherwise the trees are built one by one.
template <typename T> static int upload_seg(ft_batch *b) {
    const int nt = (int)b->trees.size();
    b->seg_ok = false;
    if (!b->build || nt == 0) return FT_OK;
    const ft_config &c0 = b->trees[0]->config;
    size_t total = 0;
    for (int i = 0; i < nt; ++i) {
        const ft_config &c = b->trees[i]->config;
        if (c.sfc != c0.sfc || c.bits_per_dim != c0.bits_per_dim) return FT_OK;
        total += b->bn[i];
    }
    if (total == 0 || total + (size_t)nt * SEG_CHUNK >= (size_t)INT_MAX) return FT_OK;
    std::vector<SegTree<T>> st(nt);
    std::vector<int4> tiles;
    std::vector<int2> chunks;
    int32_t off = 0;
    for (int i = 0; i < nt; ++i) {
        const int n = (int)b->bn[i];
        SegTree<T> &q = st[i];
        q.x = (const T *)b->bx[i];
        q.y = (const T *)b->by[i];
        q.z = (const T *)b->bz[i];
        q.o = make_tree_out<T>(*b->trees[i]);
        q.sorted_keys = b->trees[i]->sorted_keys;
        q.off = off;
        q.n = n;
        q.tile0 = (int32_t)tiles.size();
        for (int a = 0; a < n; a += SEG_TILE)
            tiles.push_back(make_int4(off + a, off + std::min(a + SEG_TILE, n), i, a == 0 ? 1 : 0));
        q.ntiles = (int32_t)tiles.size() - q.tile0;
        for (int a = 0; a < n; a += SEG_CHUNK) chunks.push_back(make_int2(i, a));
        off += n;
    }
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_t = 0, s_t = al(sizeof(SegTree<T>) * nt);
    const size_t o_i = o_t + s_t, s_i = al(sizeof(int4) * tiles.size());
    const size_t o_c = o_i + s_i, s_c = al(sizeof(int2) * chunks.size());
    const size_t o_h = o_c + s_c, s_h = al(4 * 2 * (size_t)nt * SEG_HIST_WORDS);
    const size_t o_b = o_h + s_h, s_b = al((size_t)nt * 16 * sizeof(T));  // box records
    const size_t o_p = o_b + s_b, s_p = al(tiles.size() * 6 * sizeof(T));
    const size_t bytes = o_p + s_p;
    if (bytes > b->seg_bytes) {
        if (b->seg_block) FT_HIP(hipFree(b->seg_block));
        b->seg_block = nullptr;
        b->seg_bytes = 0;
        if (hipMalloc(&b->seg_block, bytes) != hipSuccess)
            return fail(FT_ENOMEM, "segmented build table allocation failed");
        b->seg_bytes = bytes;
    }
    char *base = (char *)b->seg_block;
    FT_HIP(hipMemcpy(base + o_t, st.data(), sizeof(SegTree<T>) * nt, hipMemcpyHostToDevice));
    FT_HIP(hipMemcpy(base + o_i, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice));
    FT_HIP(hipMemcpy(base + o_c, chunks.data(), sizeof(int2) * chunks.size(),
                     hipMemcpyHostToDevice));
    SegBuild &g = b->seg;
    g.trees = base + o_t;
    g.tiles = (const int4 *)(base + o_i);
    g.chunks = (const int2 *)(base + o_c);
    g.ntrees = nt;
    g.ntiles = (int)tiles.size();
    g.nchunks = (int)chunks.size();
    g.total = total;
    g.hist = (uint32_t *)(base + o_h);
    g.bbox = base + o_b;
    g.partials = base + o_p;
    FT_TRY(scratch_ensure_public(b->seg_scratch, total + (size_t)nt * SEG_CHUNK));
    b->seg_ok = true;
    return FT_OK;
}

static bool use_seg(const ft_batch *b) { return b->seg_ok && b->seg_mode != 0; }

// Device tables: trees, requests (with tuned settings), items per variant.
template <typename T> static int upload_tables(ft_batch *b) {
    using D = typename CoordTraits<T>::D;
    (void)sizeof(D);
    const int nt = (int)b->trees.size(), nr = (int)b->reqs.size();
    std::vector<DTree<T>> dts(nt);
    for (int i = 0; i < nt; ++i) dts[i] = make_dtree<T>(*b->trees[i]);
    b->mk = nt ? metric_kind<T>(dts[0].m) : MK_ANY;
    for (int i = 1; i < nt; ++i) {
        const Metric &a0 = dts[0].m, &ai = dts[i].m;
        if (metric_kind<T>(ai) != b->mk || ai.bits != a0.bits || ai.box != a0.box) b->mk = MK_ANY;
    }
    std::vector<BatchReq> rq(nr);
    std::vector<std::vector<std::pair<double, uint2>>> items(V_COUNT);
    int64_t entry = 0;
    size_t lds = 0;
    int kmax = 1;
    for (const auto &r : b->reqs) kmax = std::max(kmax, (int)r.k);
    b->kmax = kmax;
    b->lane_width = lane_tier_lanes<T>(kmax, lds);
    for (int i = 0; i < nr; ++i) {
        const ft_batch_request &r = b->reqs[i];
        const size_t n = b->build ? b->bn[r.tree] : b->trees[r.tree]->num_leaves;
        GroupTune &g = b->tune[size_class(n)][k_tier(r.k)];
        BatchReq &q = rq[i];
        q.tree = r.tree;
        q.k = r.k;
        q.entry_first = (int32_t)entry;
        q.num = (int32_t)r.num_queries;
        q.bucket = g.bucket;
        // a wave list shorter than this request's k: the per-lane tier
        q.variant = (is_wave(g.variant) && r.k > wave_cap(g.variant)) ? (int)V_LANE : g.variant;
        q.width = item_width(q.variant, g.width, b->lane_width);
        q.pad = 0;
        q.query_first = r.query_first;
        q.result_first = r.result_first;
        entry += r.num_queries;
        const int pw = q.width;
        const double cost = g.ms_per_item > 0 ? g.ms_per_item : (double)r.k * std::log2(2.0 + n);
        if (n == 0) continue;  // an empty tree: results untouched, as knn_query
        for (int j = 0; j * pw < r.num_queries; ++j)
            items[q.variant].push_back({cost, make_uint2((unsigned)i, (unsigned)j)});
    }
    b->n_entries = entry;
    std::vector<uint2> flat;
    std::vector<uint32_t> bnd((size_t)V_COUNT * (FT_XCDS + 1) + FT_XCDS + 1, 0);
    for (int v = 0; v < V_COUNT; ++v) {
        auto &L = items[v];
        std::stable_sort(L.begin(), L.end(),
                         [](const std::pair<double, uint2> &a, const std::pair<double, uint2> &c) {
                             return a.first > c.first;
                         });
        b->item_first[v] = (int)flat.size();
        b->item_count[v] = (int)L.size();
        for (auto &p : L) flat.push_back(p.second);
        // FT_XCDS contiguous chunks of about equal measured cost
        double total = 0;
        for (auto &p : L) total += p.first;
        b->var_cost[v] = total;
        uint32_t *bv = &bnd[(size_t)v * (FT_XCDS + 1)];
        double run = 0;
        int c = 1;
        for (size_t i = 0; i < L.size() && c < FT_XCDS; ++i) {
            run += L[i].first;
            while (c < FT_XCDS && run >= total * c / FT_XCDS) bv[c++] = (uint32_t)(i + 1);
        }
        for (; c <= FT_XCDS; ++c) bv[c] = (uint32_t)L.size();
    }
    // one device block: trees | reqs | items | counters | rerun | keys | sorted keys
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t ne = (size_t)std::max<int64_t>(entry, 1);
    const size_t o_t = 0, s_t = al(sizeof(DTree<T>) * std::max(nt, 1));
    const size_t o_r = o_t + s_t, s_r = al(sizeof(BatchReq) * std::max(nr, 1));
    const size_t o_i = o_r + s_r, s_i = al(sizeof(uint2) * std::max<size_t>(flat.size(), 1));
    const size_t o_c = o_i + s_i, s_c = al(4 * (V_COUNT * FT_XCDS + 1));
    const size_t o_b = o_c + s_c, s_b = al(4 * bnd.size());
    const size_t o_l = o_b + s_b, s_l = al(4 * ne);
    const size_t o_k = o_l + s_l, s_k = al(8 * ne);
    const size_t o_s = o_k + s_k, s_s = al(8 * ne);
    const size_t total = o_s + s_s;
    if (total > b->dev_bytes) {
        if (b->dev_block) FT_HIP(hipFree(b->dev_block));
        b->dev_block = nullptr;
        if (hipMalloc(&b->dev_block, total) != hipSuccess)
            return fail(FT_ENOMEM, "batch table allocation failed");
        b->dev_bytes = total;
    }
    char *base = (char *)b->dev_block;
    b->dtrees = base + o_t;
    b->dreqs = (BatchReq *)(base + o_r);
    b->ditems = (uint2 *)(base + o_i);
    b->dctr = (uint32_t *)(base + o_c);
    b->dbnd = (uint32_t *)(base + o_b);
    b->drerun = (int32_t *)(base + o_l);
    b->dkeys = (uint64_t *)(base + o_k);
    b->dsorted = (uint64_t *)(base + o_s);
    if (nt) FT_HIP(hipMemcpy(b->dtrees, dts.data(), sizeof(DTree<T>) * nt, hipMemcpyHostToDevice));
    if (nr) FT_HIP(hipMemcpy(b->dreqs, rq.data(), sizeof(BatchReq) * nr, hipMemcpyHostToDevice));
    if (!flat.empty())
        FT_HIP(hipMemcpy(b->ditems, flat.data(), sizeof(uint2) * flat.size(),
                         hipMemcpyHostToDevice));
    FT_HIP(hipMemset(b->dctr, 0, 4 * (V_COUNT * FT_XCDS + 1)));
    FT_HIP(hipMemcpy(b->dbnd, bnd.data(), 4 * bnd.size(), hipMemcpyHostToDevice));
    FT_TRY(upload_seg<T>(b));
    // The clear and copies above go through the null stream; the step runs on
    // non-blocking streams that do not wait for it.
    FT_HIP(hipDeviceSynchronize());
    FT_TRY(scratch_ensure_public(b->sort_scratch, ne));
    b->dirty = false;
    return FT_OK;
}

template <typename T> static BatchView<T> view(ft_batch *b) {
    BatchView<T> v;
    v.trees = (const DTree<T> *)b->dtrees;
    v.reqs = b->dreqs;
    v.nreq = (int)b->reqs.size();
    v.order = b->sort_scratch.perm;
    v.qx = (const T *)b->qx;
    v.qy = (const T *)b->qy;
    v.qz = (const T *)b->qz;
    v.res = b->results;
    v.dists = (typename CoordTraits<T>::D *)b->dists;
    v.rerun = b->drerun;
    v.rerun_count = b->dctr + V_COUNT * FT_XCDS;
    v.rerun_cap = (uint32_t)std::max<int64_t>(b->n_entries, 1);
    v.stats = b->stats;
    return v;
}

// Waves of a persistent kernel: as many as can be resident on the device
// (occupancy x compute units), at most one per item.
template <typename K> static unsigned persistent_waves(K kernel, size_t dyn_lds, int n_items) {
    static int n_cu = 0;
    if (!n_cu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n_cu < 1)
            n_cu = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, dyn_lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 8;
    return (unsigned)std::max(1L, std::min<long>((long)n_items, (long)per_cu * n_cu));
}

// Launches one tier kernel over `n_items` items (its work queue: chunk bounds
// `bnd`, counters `ctr`).
template <typename T, int MK>
static int launch_variant_mk(ft_batch *b, hipStream_t s, int v, const uint2 *it, int n_items,
                             const uint32_t *bnd, uint32_t *ctr) {
    const BatchView<T> bv = view<T>(b);
    constexpr bool FP = CoordTraits<T>::fp;
#define FT_BPK(M_, HEAP_, ...)                                                                  \
    do {                                                                                        \
        auto kf = batch_packet_kernel<T, M_, HEAP_, MK, ##__VA_ARGS__>;                         \
        hipLaunchKernelGGL(kf, dim3(persistent_waves(kf, 0, n_items)), dim3(64), 0, s, bv, it,  \
                           bnd, ctr);                                                           \
    } while (0)
    switch (v) {
    case V_PK1: FT_BPK(1, false); break;
    case V_PK4: FT_BPK(4, false); break;
    case V_PK8: FT_BPK(8, false); break;
    case V_PK16: FT_BPK(16, false); break;
    case V_PK32: FT_BPK(32, false); break;
    case V_PASSES: FT_BPK(32, false, true); break;
    case V_HP32:
    case V_HP64:
    case V_HP128:
        if constexpr (FP) {
            return fail(FT_EINVAL, "LDS heap packet tier needs integer coordinates");
        } else {
            if (v == V_HP32) FT_BPK(32, true);
            else if (v == V_HP64) FT_BPK(64, true);
            else FT_BPK(128, true);
        }
        break;
    case V_LANE: {
        size_t lds = 0;
        const int lanes = lane_tier_lanes<T>(b->kmax, lds);
        if (!lanes) return fail(FT_ELDS, "k too large for local memory: reduce k");
        auto kf = batch_lane_kernel<T, MK>;
        hipLaunchKernelGGL(kf, dim3(persistent_waves(kf, lds, n_items)), dim3(64), lds, s, bv, it,
                           bnd, ctr, lanes, b->kmax);
        break;
    }
    case V_WAVE64:
    case V_WAVE128:
    case V_WAVE256: {
        auto kf = v == V_WAVE64 ? batch_wave_kernel<T, 1, MK>
                                : (v == V_WAVE128 ? batch_wave_kernel<T, 2, MK> : batch_wave_kernel<T, 4, MK>);
        hipLaunchKernelGGL(kf, dim3(persistent_waves(kf, 0, n_items)), dim3(64), 0, s, bv, it, bnd,
                           ctr);
        break;
    }
    default: return fail(FT_EINVAL, "unknown batch tier");
    }
#undef FT_BPK
    FT_CHECK_LAUNCH();
    return FT_OK;
}

template <typename T>
static int launch_variant(ft_batch *b, hipStream_t s, int v, const uint2 *it, int n_items,
                          const uint32_t *bnd, uint32_t *ctr) {
    if (n_items <= 0) return FT_OK;
    if (b->mk == MK_WRAP) return launch_variant_mk<T, MK_WRAP>(b, s, v, it, n_items, bnd, ctr);
    if (b->mk == MK_OPEN) return launch_variant_mk<T, MK_OPEN>(b, s, v, it, n_items, bnd, ctr);
    return launch_variant_mk<T, MK_ANY>(b, s, v, it, n_items, bnd, ctr);
}

template <typename T> static int launch_exact(ft_batch *b, hipStream_t s) {
    size_t lds = 0;
    const int lanes = lane_tier_lanes<T>(b->kmax, lds);
    if (!lanes) return fail(FT_ELDS, "k too large for local memory: reduce k");
    hipLaunchKernelGGL((batch_exact_kernel<T>), dim3(256), dim3(64), lds, s, view<T>(b), lanes,
                       b->kmax);
    FT_CHECK_LAUNCH();
    if (b->stats) {
        hipLaunchKernelGGL(batch_count_kernel, dim3(1), dim3(1), 0, s, b->stats,
                           (const uint32_t *)(b->dctr + V_COUNT * FT_XCDS));
        FT_CHECK_LAUNCH();
    }
    return FT_OK;
}

// Tree i goes to the branch with the least work so far, largest trees first.
// Returns that order; br[i] = branch of tree i, maxn[t] = largest tree of t.
static std::vector<int> assign_branches(ft_batch *b, int nb, std::vector<int> &br,
                                        std::vector<size_t> &maxn) {
    const int nt = (int)b->trees.size();
    std::vector<int> order(nt);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->bn[a] > b->bn[c]; });
    std::vector<size_t> load(nb, 0);
    maxn.assign(nb, 0);
    br.assign(nt, 0);
    for (int i : order) {
        const int t = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        br[i] = t;
        load[t] += b->bn[i];
        maxn[t] = std::max(maxn[t], b->bn[i]);
    }
    return order;
}

// Builds every tree: segmented (one launch of each build kernel for all
// trees), or one build_bvh per tree on nb branches (nb = 1: in order on s).
template <typename T> static int enqueue_builds(ft_batch *b, hipStream_t s, int nb,
                                                bool per_tree = false) {
    const int nt = (int)b->trees.size();
    if (!b->build || nt == 0) return FT_OK;
    if (use_seg(b) && !per_tree)
        return build_segmented_t<T>(s, b->seg_scratch, b->seg, b->trees[0]->config);
    nb = std::max(1, std::min(nb, nt));
    FT_TRY(ensure_streams(b, nb));
    while ((int)b->build_scratch.size() < nb) b->build_scratch.emplace_back();
    std::vector<int> br;
    std::vector<size_t> maxn;
    const std::vector<int> order = assign_branches(b, nb, br, maxn);
    for (int t = 0; t < nb; ++t) FT_TRY(scratch_ensure_public(b->build_scratch[t], maxn[t]));
    if (nb > 1) {
        FT_HIP(hipEventRecord(b->ev_fork, s));
        for (int t = 0; t < nb; ++t) FT_HIP(hipStreamWaitEvent(b->side[t], b->ev_fork, 0));
    }
    for (int i : order) {
        ft_tree *tr = b->trees[i];
        const size_t n = b->bn[i];
        if (n == 0) continue;
        tr->num_leaves = n;
        tr->num_internal = n - 1;
        hipStream_t st = nb > 1 ? b->side[br[i]] : s;
        b->build_scratch[br[i]].radix_bits = b->radix_bits[size_class(n)];
        FT_TRY(build_bvh_t<T>(st, tr, b->build_scratch[br[i]], (const T *)b->bx[i],
                              (const T *)b->by[i], (const T *)b->bz[i], nullptr, nullptr, n,
                              nullptr, nullptr));
    }
    if (nb > 1) {
        for (int t = 0; t < nb; ++t) {
            FT_HIP(hipEventRecord(b->ev_join[t], b->side[t]));
            FT_HIP(hipStreamWaitEvent(s, b->ev_join[t], 0));
        }
    }
    return FT_OK;
}

static int bit_len(uint64_t v) {
    int r = 0;
    while (v) {
        ++r;
        v >>= 1;
    }
    return r;
}

// Query ordering (composite-key sort).
template <typename T> static int enqueue_sort(ft_batch *b, hipStream_t s) {
    if (b->n_entries == 0) return FT_OK;
    const int nreq = (int)b->reqs.size();
    const int rbits = std::max(1, bit_len((uint64_t)(nreq - 1)));
    const int qb = std::min(21, (64 - rbits) / 3);
    hipLaunchKernelGGL((batch_key_kernel<T>), dim3(cdiv(b->n_entries, 256)), dim3(256), 0, s,
                       view<T>(b), b->n_entries, rbits, qb, b->dkeys);
    FT_CHECK_LAUNCH();
    return radix_sort_u64(s, b->sort_scratch, b->dkeys, (size_t)b->n_entries, b->dsorted, 64);
}

// Tier kernels on parallel branches (or in order on s), then the exact pass.
template <typename T> static int enqueue_queries(ft_batch *b, hipStream_t s) {
    if (b->n_entries == 0) return FT_OK;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(128), 0, s, b->dctr,
                       V_COUNT * FT_XCDS + 1);
    FT_CHECK_LAUNCH();
    // Tier kernels, most expensive first, on at most QUERY_CHAINS parallel
    // chains (the hardware queues a process gets by default): each kernel
    // goes to the chain with the least estimated work so far, so the largest
    // starts at once instead of queueing behind a small one.
    std::vector<int> used;
    for (int v = 0; v < V_COUNT; ++v)
        if (b->item_count[v] > 0) used.push_back(v);
    std::stable_sort(used.begin(), used.end(),
                     [b](int x, int y) { return b->var_cost[x] > b->var_cost[y]; });
    const int nb = std::min((int)used.size(), QUERY_CHAINS);
    if (nb > 1) {
        FT_TRY(ensure_streams(b, nb));
        FT_HIP(hipEventRecord(b->ev_fork, s));
        for (int i = 0; i < nb; ++i) FT_HIP(hipStreamWaitEvent(b->side[i], b->ev_fork, 0));
    }
    std::vector<double> load(std::max(nb, 1), 0.0);
    for (int v : used) {
        const int ch = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[ch] += b->var_cost[v];
        FT_TRY(launch_variant<T>(b, nb > 1 ? b->side[ch] : s, v, b->ditems + b->item_first[v],
                                 b->item_count[v], b->dbnd + v * (FT_XCDS + 1),
                                 b->dctr + v * FT_XCDS));
    }
    if (nb > 1) {
        for (int i = 0; i < nb; ++i) {
            FT_HIP(hipEventRecord(b->ev_join[i], b->side[i]));
            FT_HIP(hipStreamWaitEvent(s, b->ev_join[i], 0));
        }
    }
    return launch_exact<T>(b, s);
}

template <typename T> static int enqueue_step(ft_batch *b, hipStream_t s) {
    if (b->dirty) FT_TRY(upload_tables<T>(b));
    FT_TRY(enqueue_builds<T>(b, s, b->branches));
    FT_TRY(enqueue_sort<T>(b, s));
    return enqueue_queries<T>(b, s);
}

static float elapsed(hipEvent_t a, hipEvent_t c) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, c);
    return ms;
}

// Autotuner.  Builds once, sorts the queries once, then for every (size class,
// k tier) group with work times each candidate tier kernel (and, for packet
// tiers, leaf bucket and then packet width) on that group's items alone,
// keeping the fastest; then, per
// size class, that class's builds with 8- and 9-bit sort digits; then the
// build phase at 1, 2, 4 and 8 branches.
// One timed setting of a group: tier kernel, leaf bucket, queries per item;
// the time of the group's items alone and their count.
struct Setting {
    int variant, bucket, width;
    float ms;
    int items;
};
static void apply_setting(GroupTune &g, const Setting &st) {
    g.variant = st.variant;
    g.bucket = st.bucket;
    g.width = st.width;
    g.ms_per_item = st.ms / std::max(st.items, 1);
    g.tuned = true;
}

template <typename T> static int tune_t(ft_batch *b, hipStream_t s) {
    const bool fp = CoordTraits<T>::fp;
    std::vector<Setting> alts[N_CLASS][N_KTIER];
    default_tuning(b);
    FT_TRY(upload_tables<T>(b));
    FT_TRY(enqueue_builds<T>(b, s, b->branches));
    FT_TRY(enqueue_sort<T>(b, s));
    FT_HIP(hipStreamSynchronize(s));
    hipEvent_t e0, e1;
    FT_HIP(hipEventCreate(&e0));
    FT_HIP(hipEventCreate(&e1));
    // Items of every group: packet tiers at 64, 32 or 16 queries per item,
    // the per-lane tier at lane_width.
    const int nr = (int)b->reqs.size();
    int rc = FT_OK;
    for (int c = 0; c < N_CLASS && rc == FT_OK; ++c) {
        for (int kt = 0; kt < N_KTIER && rc == FT_OK; ++kt) {
            std::vector<int> members;
            for (int i = 0; i < nr; ++i) {
                const auto &r = b->reqs[i];
                const size_t n = b->build ? b->bn[r.tree] : b->trees[r.tree]->num_leaves;
                if (r.num_queries > 0 && size_class(n) == c && k_tier(r.k) == kt)
                    members.push_back(i);
            }
            if (members.empty()) continue;
            int gk = 1;
            for (int i : members) gk = std::max(gk, (int)b->reqs[i].k);
            if (b->tune[c][kt].pinned) continue;
            // Times variant v at (bucket, width) on the group's items alone;
            // returns the time of the better of two runs, or a negative value
            // after an error (rc set).
            auto time_setting = [&](int v, int bk, int pw, int &n_items) -> float {
                std::vector<uint2> its;
                for (int i : members)
                    for (int j = 0; j * pw < b->reqs[i].num_queries; ++j)
                        its.push_back(make_uint2((unsigned)i, (unsigned)j));
                n_items = (int)its.size();
                // scratch item list at the end of the item table's region is not
                // guaranteed large enough: use a temporary device buffer
                uint2 *tmp = nullptr;
                if (hipMalloc(&tmp, its.size() * sizeof(uint2)) != hipSuccess) {
                    rc = fail(FT_ENOMEM, "tuner allocation failed");
                    return -1.0f;
                }
                (void)hipMemcpy(tmp, its.data(), its.size() * sizeof(uint2), hipMemcpyHostToDevice);
                // one chunk holding every item: all XCDs drain it
                std::vector<uint32_t> tb(FT_XCDS + 1, (uint32_t)its.size());
                tb[0] = 0;
                uint32_t *tbnd = b->dbnd + V_COUNT * (FT_XCDS + 1);
                (void)hipMemcpy(tbnd, tb.data(), 4 * tb.size(), hipMemcpyHostToDevice);
                // set the group's requests on the device
                for (int i : members) {
                    BatchReq q;
                    (void)hipMemcpy(&q, b->dreqs + i, sizeof q, hipMemcpyDeviceToHost);
                    q.bucket = bk;
                    q.variant = v;
                    q.width = pw;
                    (void)hipMemcpy(b->dreqs + i, &q, sizeof q, hipMemcpyHostToDevice);
                }
                // null-stream copies above, before the timed launches on s
                (void)hipDeviceSynchronize();
                float t_best = 1e30f;
                for (int rep = 0; rep < 2 && rc == FT_OK; ++rep) {
                    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(128), 0, s, b->dctr,
                                       V_COUNT * FT_XCDS + 1);
                    (void)hipEventRecord(e0, s);
                    rc = launch_variant<T>(b, s, v, tmp, (int)its.size(), tbnd, b->dctr);
                    (void)hipEventRecord(e1, s);
                    (void)hipEventSynchronize(e1);
                    t_best = std::min(t_best, elapsed(e0, e1));
                }
                (void)hipFree(tmp);
                return rc == FT_OK ? t_best : -1.0f;
            };
            std::vector<Setting> found;  // every timed setting of the group
            for (int v : candidates(kt, fp, gk)) {
                // packet tiers: leaf buckets at 64 queries per packet, then
                // 32 and 16 queries per packet with the best of those buckets
                const std::vector<int> buckets =
                    is_packet(v) ? std::vector<int>{16, 32, 64} : std::vector<int>{32};
                int v_bucket = buckets.front();
                float v_best = 1e30f;
                for (int bk : buckets) {
                    int ni = 1;
                    const int pw = is_packet(v) ? 64 : (is_wave(v) ? WAVE_ITEM : b->lane_width);
                    const float tm = time_setting(v, bk, pw, ni);
                    if (rc != FT_OK) break;
                    found.push_back({v, bk, pw, tm, ni});
                    if (tm < v_best) {
                        v_best = tm;
                        v_bucket = bk;
                    }
                }
                if (rc != FT_OK) break;
                if (!is_packet(v)) continue;
                for (int pw : {32, 16}) {
                    int ni = 1;
                    const float tm = time_setting(v, v_bucket, pw, ni);
                    if (rc != FT_OK) break;
                    found.push_back({v, v_bucket, pw, tm, ni});
                }
                if (rc != FT_OK) break;
            }
            if (rc != FT_OK || found.empty()) break;
            std::stable_sort(found.begin(), found.end(),
                             [](const Setting &x, const Setting &y) { return x.ms < y.ms; });
            // the fastest of each tier kernel, fastest first: the whole-step
            // stage below tries the runners-up
            std::vector<Setting> &alt = alts[c][kt];
            alt.clear();
            for (const Setting &st : found) {
                bool seen = false;
                for (const Setting &a : alt) seen = seen || a.variant == st.variant;
                if (!seen) alt.push_back(st);
            }
            // and the fastest tier at 64 queries per packet, when it won narrower
            if (alt.front().width != 64 && is_packet(alt.front().variant))
                for (const Setting &st : found)
                    if (st.variant == alt.front().variant && st.width == 64) {
                        alt.insert(alt.begin() + 1, st);
                        break;
                    }
            apply_setting(b->tune[c][kt], alt.front());
        }
    }
    // Whole-step refinement.  Each group was timed alone, on an otherwise idle
    // device, where a tier can win by parallelism alone (narrow packets) or
    // lose what it would gain by sharing the device (the per-lane tier leans
    // on LDS, the packet tiers on registers).  The query phase of the step,
    // every tier running at once, is timed with the isolated choices, then,
    // group by group (largest measured cost first), with each of the group's
    // other tiers; a change stays when it makes the phase faster.
    if (rc == FT_OK) {
        auto time_queries = [&]() -> float {
            float tb = 1e30f;
            if ((rc = upload_tables<T>(b)) != FT_OK) return tb;
            for (int rep = 0; rep < 2 && rc == FT_OK; ++rep) {
                (void)hipEventRecord(e0, s);
                rc = enqueue_queries<T>(b, s);
                (void)hipEventRecord(e1, s);
                (void)hipEventSynchronize(e1);
                tb = std::min(tb, elapsed(e0, e1));
            }
            return tb;
        };
        std::vector<std::pair<double, int>> order;
        for (int c = 0; c < N_CLASS; ++c)
            for (int kt = 0; kt < N_KTIER; ++kt)
                if (!alts[c][kt].empty()) order.push_back({-(double)alts[c][kt].front().ms, c * N_KTIER + kt});
        std::sort(order.begin(), order.end());
        float t_cur = time_queries();
        for (const auto &o : order) {
            if (rc != FT_OK) break;
            const int c = o.second / N_KTIER, kt = o.second % N_KTIER;
            GroupTune keep = b->tune[c][kt];
            for (size_t a = 1; a < alts[c][kt].size() && a < 4 && rc == FT_OK; ++a) {
                const GroupTune prev = b->tune[c][kt];
                apply_setting(b->tune[c][kt], alts[c][kt][a]);
                const float t = time_queries();
                if (t < 0.99f * t_cur) {
                    t_cur = t;
                    keep = b->tune[c][kt];
                } else {
                    b->tune[c][kt] = prev;
                }
            }
            b->tune[c][kt] = keep;
        }
    }
    // Radix width per size class: that class's builds, in order on s, with
    // 8- and 9-bit digits (the sorted order is the same).
    if (rc == FT_OK && b->build) {
        const int nt = (int)b->trees.size();
        if (b->build_scratch.empty()) b->build_scratch.emplace_back();
        Scratch &S0 = b->build_scratch[0];
        for (int c = 0; c < N_CLASS && rc == FT_OK; ++c) {
            size_t maxn = 0;
            for (int i = 0; i < nt; ++i)
                if (b->bn[i] > 0 && size_class(b->bn[i]) == c) maxn = std::max(maxn, b->bn[i]);
            if (maxn == 0) continue;
            rc = scratch_ensure_public(S0, maxn);
            float best = 1e30f;
            int best_bits = 8;
            for (int bits : {8, 9}) {
                S0.radix_bits = bits;
                float t_best = 1e30f;
                for (int rep = 0; rep < 2 && rc == FT_OK; ++rep) {
                    (void)hipEventRecord(e0, s);
                    for (int i = 0; i < nt && rc == FT_OK; ++i) {
                        const size_t n = b->bn[i];
                        if (n == 0 || size_class(n) != c) continue;
                        rc = build_bvh_t<T>(s, b->trees[i], S0, (const T *)b->bx[i],
                                            (const T *)b->by[i], (const T *)b->bz[i], nullptr,
                                            nullptr, n, nullptr, nullptr);
                    }
                    (void)hipEventRecord(e1, s);
                    (void)hipEventSynchronize(e1);
                    t_best = std::min(t_best, elapsed(e0, e1));
                }
                if (t_best < best) {
                    best = t_best;
                    best_bits = bits;
                }
            }
            b->radix_bits[c] = best_bits;
        }
        S0.radix_bits = 0;
    }
    // Build phase: one build per tree on 1, 2, 4 or 8 branches, against the
    // segmented build of every tree at once.
    if (rc == FT_OK && b->build) {
        double best = 1e30;
        int best_nb = 1;
        for (int nb : {1, 2, 4, 8}) {
            if (nb > 1 && nb > (int)b->trees.size()) break;
            float t_best = 1e30f;
            for (int rep = 0; rep < 2 && rc == FT_OK; ++rep) {
                (void)hipEventRecord(e0, s);
                rc = enqueue_builds<T>(b, s, nb, true);
                (void)hipEventRecord(e1, s);
                (void)hipEventSynchronize(e1);
                t_best = std::min(t_best, elapsed(e0, e1));
            }
            if (t_best < best) {
                best = t_best;
                best_nb = nb;
            }
        }
        b->branches = best_nb;
        b->seg_mode = 0;
        if (b->seg_ok && rc == FT_OK) {
            float t_seg = 1e30f;
            for (int rep = 0; rep < 2 && rc == FT_OK; ++rep) {
                (void)hipEventRecord(e0, s);
                rc = build_segmented_t<T>(s, b->seg_scratch, b->seg, b->trees[0]->config);
                (void)hipEventRecord(e1, s);
                (void)hipEventSynchronize(e1);
                t_seg = std::min(t_seg, elapsed(e0, e1));
            }
            if (t_seg < best) b->seg_mode = 1;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    FT_TRY(rc);
    b->tuned = true;
    b->dirty = true;
    free_graph(b);
    return upload_tables<T>(b);
}

// Sizes every table, stream and scratch buffer the step uses, so that the
// step itself (and a graph capture of it) allocates nothing.
template <typename T> static int prepare_t(ft_batch *b) {
    if (b->dirty) FT_TRY(upload_tables<T>(b));
    FT_TRY(ensure_streams(b, std::max(b->branches, (int)V_COUNT)));
    if (b->build && !b->trees.empty()) {
        const int nb = std::max(1, std::min(b->branches, (int)b->trees.size()));
        std::vector<int> br;
        std::vector<size_t> maxn;
        assign_branches(b, nb, br, maxn);
        while ((int)b->build_scratch.size() < nb) b->build_scratch.emplace_back();
        for (int t = 0; t < nb; ++t) FT_TRY(scratch_ensure_public(b->build_scratch[t], maxn[t]));
    }
    return FT_OK;
}

// Records the step on the batch's own capture stream; the graph is then
// launched on any stream (also the null stream, which cannot be captured).
template <typename T> static int capture_t(ft_batch *b, hipStream_t caller) {
    free_graph(b);
    if (!b->tuned) default_tuning(b);
    FT_TRY(prepare_t<T>(b));
    if (!b->cap) FT_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    FT_HIP(hipStreamSynchronize(caller));  // table uploads done before recording
    const hipStream_t s = b->cap;
    FT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    const int rc = enqueue_step<T>(b, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != FT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess)
        return fail(FT_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    b->graph = g;
    FT_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return FT_OK;
}

}  // namespace ft

using namespace ft;

#define FT_BATCH_DISPATCH(b, CALL)                                                   \
    switch ((b)->coord_type) {                                                       \
    case FT_COORD_F32: { using T = float; return CALL; }                             \
    case FT_COORD_F64: { using T = double; return CALL; }                            \
    case FT_COORD_U32: { using T = uint32_t; return CALL; }                          \
    case FT_COORD_U64: { using T = uint64_t; return CALL; }                          \
    default: return fail(FT_EINVAL, "unknown coordinate type");                      \
    }

extern "C" {

int ft_batch_create(ft_batch **out, int32_t coord_type) {
    if (!out) return fail(FT_EINVAL, "null batch");
    if (coord_type < FT_COORD_F32 || coord_type > FT_COORD_U64)
        return fail(FT_EINVAL, "unknown coordinate type");
    ft_batch *b = new ft_batch;
    b->coord_type = coord_type;
    *out = b;
    return FT_OK;
}

int ft_batch_destroy(ft_batch *b) {
    if (!b) return FT_OK;
    (void)hipDeviceSynchronize();
    free_graph(b);
    for (hipStream_t st : b->side) (void)hipStreamDestroy(st);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (hipEvent_t ev : b->ev_join) (void)hipEventDestroy(ev);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->dev_block) (void)hipFree(b->dev_block);
    if (b->sort_scratch.block) (void)hipFree(b->sort_scratch.block);
    if (b->seg_scratch.block) (void)hipFree(b->seg_scratch.block);
    if (b->seg_block) (void)hipFree(b->seg_block);
    for (auto &S : b->build_scratch)
        if (S.block) (void)hipFree(S.block);
    delete b;
    return FT_OK;
}

int ft_batch_set_trees(ft_batch *b, int num_trees, ft_tree *const *trees, const void *const *x,
                       const void *const *y, const void *const *z, const size_t *n) {
    if (!b || num_trees < 0 || (num_trees > 0 && !trees)) return fail(FT_EINVAL, "bad tree list");
    const bool build = x != nullptr;
    if (build && (!y || !z || !n)) return fail(FT_EINVAL, "build inputs need x, y, z and n");
    b->trees.assign(trees, trees + num_trees);
    b->build = build;
    b->bx.clear(); b->by.clear(); b->bz.clear(); b->bn.clear();
    for (int i = 0; i < num_trees; ++i) {
        ft_tree *t = b->trees[i];
        if (!t || !t->block) return fail(FT_EINVAL, "tree not allocated");
        if (t->coord_type != b->coord_type)
            return fail(FT_EINVAL, "tree coordinate type differs from the batch's");
        if (build) {
            if (n[i] > t->capacity)
                return fail(FT_EINVAL, "tree capacity is smaller than the particle count");
            if (n[i] > (size_t)(INT_MAX / 2))
                return fail(FT_ERANGE, "Particle count n exceeds maximum supported tree size (INT_MAX / 2).");
            b->bx.push_back(x[i]); b->by.push_back(y[i]); b->bz.push_back(z[i]);
            b->bn.push_back(n[i]);
            t->num_leaves = n[i];  // the batch builds it with this many particles
            t->num_internal = n[i] ? n[i] - 1 : 0;
        }
    }
    b->dirty = true;
    free_graph(b);
    return FT_OK;
}

int ft_batch_set_queries(ft_batch *b, int num_requests, const ft_batch_request *reqs,
                         const void *qx, const void *qy, const void *qz, uint64_t *results,
                         void *dists) {
    if (!b || num_requests < 0 || (num_requests > 0 && !reqs))
        return fail(FT_EINVAL, "bad request list");
    int64_t total = 0;
    for (int i = 0; i < num_requests; ++i) {
        const ft_batch_request &r = reqs[i];
        if (r.tree < 0 || r.tree >= (int)b->trees.size())
            return fail(FT_EINVAL, "request names a tree outside the batch");
        if (r.k < 1 || r.k > 256) return fail(FT_EINVAL, "batch requests need 1 <= k <= 256");
        if (r.num_queries < 0 || r.query_first < 0 || r.result_first < 0)
            return fail(FT_EINVAL, "negative query range");
        total += r.num_queries;
    }
    if (total > (int64_t)INT_MAX / 2) return fail(FT_ERANGE, "too many batch queries");
    if (total > 0 && (!qx || !qy || !qz || !results || !dists))
        return fail(FT_EINVAL, "null query or result pointer");
    b->reqs.assign(reqs, reqs + num_requests);
    b->qx = qx; b->qy = qy; b->qz = qz;
    b->results = results;
    b->dists = dists;
    b->dirty = true;
    free_graph(b);
    return FT_OK;
}

int ft_batch_set_stats(ft_batch *b, uint64_t *stats) {
    if (!b) return fail(FT_EINVAL, "null batch");
    b->stats = stats;
    b->dirty = true;
    free_graph(b);
    return FT_OK;
}

int ft_batch_set_tier(ft_batch *b, int32_t n_class, int32_t k_tier, const char *tier,
                      int32_t bucket, int32_t queries_per_item) {
    if (!b || !tier) return fail(FT_EINVAL, "null batch or tier name");
    if (n_class < 0 || n_class >= N_CLASS || k_tier < 0 || k_tier >= N_KTIER)
        return fail(FT_EINVAL, "size class must be 0..3 and k tier 0..7");
    int v = -1;
    for (int i = 0; i < V_COUNT; ++i)
        if (std::strcmp(tier, kVariantName[i]) == 0) v = i;
    if (v < 0) return fail(FT_EINVAL, std::string("unknown batch tier: ") + tier);
    const bool fp = b->coord_type == FT_COORD_F32 || b->coord_type == FT_COORD_F64;
    const std::vector<int> ok = candidates(k_tier, fp, kTierCap[k_tier]);
    if (std::find(ok.begin(), ok.end(), v) == ok.end())
        return fail(FT_EINVAL, std::string("tier ") + tier + " does not serve this k tier");
    if (bucket < 0 || bucket > 64) return fail(FT_EINVAL, "leaf bucket must be 0..64");
    int w = queries_per_item;
    if (is_packet(v)) {
        if (w == 0) w = 64;
        if (w != 16 && w != 32 && w != 64)
            return fail(FT_EINVAL, "packet tiers take 16, 32 or 64 queries per item");
    } else if (is_wave(v)) {
        if (w == 0) w = WAVE_ITEM;
        if (w < 1 || w > 64) return fail(FT_EINVAL, "wave tiers take 1..64 queries per item");
    }
    GroupTune &g = b->tune[n_class][k_tier];
    g.variant = v;
    g.bucket = bucket ? bucket : 32;
    g.width = w;
    g.ms_per_item = 0.0;
    g.tuned = true;
    g.pinned = true;
    b->dirty = true;
    free_graph(b);
    return FT_OK;
}

int ft_batch_tune(ft_batch *b, ft_stream s) {
    if (!b) return fail(FT_EINVAL, "null batch");
    FT_BATCH_DISPATCH(b, tune_t<T>(b, as_stream(s)));
}

int ft_batch_run(ft_batch *b, ft_stream s) {
    if (!b) return fail(FT_EINVAL, "null batch");
    if (!b->tuned) default_tuning(b);
    FT_BATCH_DISPATCH(b, enqueue_step<T>(b, as_stream(s)));
}

int ft_batch_capture(ft_batch *b, ft_stream s) {
    if (!b) return fail(FT_EINVAL, "null batch");
    FT_BATCH_DISPATCH(b, capture_t<T>(b, as_stream(s)));
}

int ft_batch_launch(ft_batch *b, ft_stream s) {
    if (!b) return fail(FT_EINVAL, "null batch");
    if (!b->exec) FT_TRY(ft_batch_capture(b, s));
    FT_HIP(hipGraphLaunch(b->exec, as_stream(s)));
    return FT_OK;
}

// Writes the chosen settings as JSON text: per group (size class, k tier) the
// tier kernel, leaf bucket and measured cost per item; the build branch count.
int ft_batch_tuning(ft_batch *b, char *buf, size_t len) {
    if (!b || !buf || len == 0) return fail(FT_EINVAL, "null batch or buffer");
    if (!b->tuned) default_tuning(b);
    std::string out = "{\"tuned\": ";
    out += b->tuned ? "true" : "false";
    out += ", \"build\": \"";
    out += use_seg(b) ? "segmented" : "per-tree";
    out += "\", \"build_branches\": " + std::to_string(b->branches);
    static const char *const cls[N_CLASS] = {"<40k", "<160k", "<640k", ">=640k"};
    out += ", \"radix_bits\": {";
    for (int c = 0; c < N_CLASS; ++c)
        out += std::string(c ? ", " : "") + "\"" + cls[c] + "\": " + std::to_string(b->radix_bits[c]);
    out += "}, \"groups\": [";
    bool first = true;
    for (int c = 0; c < N_CLASS; ++c)
        for (int t = 0; t < N_KTIER; ++t) {
            const GroupTune &g = b->tune[c][t];
            if (!g.tuned) continue;
            char line[256];
            std::snprintf(line, sizeof line,
                          "%s{\"n_class\": \"%s\", \"k_max\": %d, \"tier\": \"%s\", "
                          "\"bucket\": %d, \"queries_per_item\": %d, \"ms_per_item\": %.6f}",
                          first ? "" : ", ", cls[c], kTierCap[t], kVariantName[g.variant],
                          g.bucket, item_width(g.variant, g.width, b->lane_width),
                          g.ms_per_item);
            out += line;
            first = false;
        }
    out += "]}";
    if (out.size() + 1 > len) return fail(FT_ERANGE, "tuning text does not fit the buffer");
    std::memcpy(buf, out.c_str(), out.size() + 1);
    return FT_OK;
}

}  // extern "C"
