/*
 * fasttree_c.h -- C ABI of the MI355X (gfx950) HLBVH build and neighbour-query path.
 *
 * Every entry point is plain C: pointers, sizes and a stream handle.  Coordinates
 * are passed as `const void *` and interpreted with the coordinate type recorded
 * in the tree (FT_COORD_*), which replaces the reference's compile-time
 * `coord_t` (src/hlbvh.hpp:202-209).  Distances are `double` for FP coordinates
 * and `uint64_t` for integer coordinates, as the reference's `dist_t`
 * (src/sfc.peano_hilbert.hpp:185-193, src/sfc.morton.hpp:18-23).
 *
 * The reference's compile-time switches (SFC_TYPE_*, BITS_PER_DIMENSION,
 * PERIODIC_BC, RETURN_ORIG_INDICES; CMakeLists.txt:63-125) become the runtime
 * ft_config stored in the tree when it is allocated.
 *
 * Pointers may be device pointers (hipMalloc / ft_malloc) or plain host
 * pointers; host pointers are staged through device memory and copied back,
 * as the reference's ensure_device_readable / copy_back_and_free do
 * (src/hlbvh.hpp:125-200).  Calls with device pointers are stream-ordered and
 * asynchronous; calls that stage host memory synchronise the stream.
 *
 * Return value: FT_OK or an FT_E* code; ft_last_error() gives the message.  The
 * C++ wrapper (fasttree/hlbvh.hpp) turns the codes back into the reference's
 * exceptions.
 */
#ifndef FASTTREE_C_H
#define FASTTREE_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define FT_OK 0
#define FT_EINVAL 1   /* reference: std::invalid_argument (k > _MAX_K_), bad argument   */
#define FT_ERANGE 2   /* reference: std::runtime_error (n > INT_MAX/2), size limits       */
#define FT_ELDS 3     /* reference: std::runtime_error (local memory requirement)         */
#define FT_ENOMEM 4   /* device allocation failed                                          */
#define FT_EHIP 5     /* HIP runtime error                                                  */

/* ---- coordinate types (reference coord_t) ----------------------------- */
#define FT_COORD_F32 0
#define FT_COORD_F64 1
#define FT_COORD_U32 2 /* POSITIONS_IN_32BIT + FASTTREE_INTEGER_COORDS */
#define FT_COORD_U64 3 /* POSITIONS_IN_64BIT + FASTTREE_INTEGER_COORDS */

/* ---- space-filling curves --------------------------------------------- */
#define FT_SFC_MORTON 0
#define FT_SFC_PEANO_HILBERT 1

/* ---- query semantics --------------------------------------------------- */
#define FT_MODE_CANONICAL 0 /* results ordered by (d2, returned index); exact top-k set        */
#define FT_MODE_REFERENCE 1 /* the reference's heap mechanics: same tie picks, 64-entry stack  */

#define FT_ALGO_AUTO 0
#define FT_ALGO_PLAIN 1   /* one lane per query, independent DFS (reference knn_query_small_k) */
#define FT_ALGO_GROUPED 2 /* SFC-grouped lockstep DFS with a wave vote (reference self-kNN)    */
#define FT_ALGO_PACKET 3  /* one shared traversal per wavefront of SFC-adjacent queries        */
#define FT_ALGO_WAVE 4    /* one wavefront per query: leaves tested 64 at a time, a sorted      */
                          /* list spread over the lanes, batches merged by cross-lane bitonic   */
                          /* networks (canonical SortedMergeHeap); k up to 1024                 */

#define FT_TOPK_AUTO 0      /* = FT_TOPK_REGISTERS                                        */
#define FT_TOPK_REGISTERS 1 /* sorted (d2, index) entries in registers, up to 32; a larger */
                            /* k (integer metrics) runs in passes of 32, each above the   */
                            /* last key of the one before                                  */
#define FT_TOPK_LDS_HEAP 2  /* a max-heap of packed keys per lane in LDS (integer metrics)*/

/* Stream handle: a hipStream_t passed as an opaque pointer.  NULL = null stream. */
typedef void *ft_stream;

/* Replaces the reference's compile-time configuration. */
typedef struct ft_config {
    int32_t sfc;                 /* FT_SFC_MORTON | FT_SFC_PEANO_HILBERT                   */
    int32_t bits_per_dim;        /* B in [1, 21]: 3B-bit keys                              */
    int32_t periodic;            /* PERIODIC_BC: integer coords wrap at 2^B                */
    int32_t return_orig_indices; /* RETURN_ORIG_INDICES                                    */
    double box_length;           /* FP coords with periodic=1 only: period L (no reference */
                                 /* counterpart; the reference ignores PERIODIC_BC for FP) */
} ft_config;

/*
 * TreeSoA (src/hlbvh.hpp:609-673): the first fourteen fields keep the
 * reference's names and order.  Internal nodes are [0, n-2] (root 0), leaves
 * [n-1, 2n-2].  The remaining fields are extensions with no reference
 * counterpart.
 */
typedef struct ft_tree {
    void *min_x, *max_x; /* coord_t[2n-1] */
    void *min_y, *max_y;
    void *min_z, *max_z;
    int32_t *left_child;  /* [n-1] */
    int32_t *right_child; /* [n-1] */
    int32_t *parent;      /* [2n-1], root -1 */
    uint32_t *id;         /* [2n-1], set on leaves */
    int8_t *is_ghost;     /* [2n-1], set on leaves */
    int32_t *orig_idx;    /* [2n-1], set on leaves */
    size_t num_leaves;
    size_t num_internal;
    /* ---- extensions ---- */
    int32_t coord_type;    /* FT_COORD_* */
    int32_t reserved0;
    ft_config config;
    uint64_t *sorted_keys; /* [n] SFC keys in leaf order                          */
    int32_t *range_first;  /* [n-1] first leaf rank under each internal node      */
    int32_t *range_last;   /* [n-1] last leaf rank under each internal node       */
    size_t capacity;       /* leaves the allocation can hold (TreePool semantics) */
    void *block;           /* single backing allocation owned by the library      */
} ft_tree;

/* Persistent build scratch (BvhScratch, src/hlbvh.hpp:720-745). Opaque. */
typedef struct ft_scratch ft_scratch;

/* Per-call query options.  NULL = all defaults. */
typedef struct ft_query_opts {
    int32_t max_k;       /* the reference's _MAX_K_ template argument; 0 = that entry   */
                         /* point's reference default (32 self, 256 subset/knn)          */
    int32_t mode;        /* FT_MODE_*                                                   */
    int32_t algorithm;   /* FT_ALGO_*                                                   */
    int32_t group_width; /* FT_ALGO_PACKET: queries that share one walk (16, 32 or 64; */
                         /* 0 = 64), the reference's sub-group width W, one wavefront   */
                         /* per group; other algorithms walk per lane.  Same results.   */
    int32_t leaf_bucket; /* FT_ALGO_PACKET: subtrees with at most this many leaves are    */
                         /* scanned as one contiguous run of leaf slots (1..64, 0 = auto) */
    int32_t topk;        /* FT_TOPK_*: top-k container of the packet traversal            */
    const void *seed_bound; /* optional per-query squared-distance bound (dist_t[num_q]):*/
                            /* the search is first limited to it; queries that find     */
                            /* fewer than k neighbours inside it are re-run unbounded.  */
                            /* Canonical mode only.  Results are unchanged by it.       */
    uint64_t *stats;     /* optional device uint64_t[8] counters, accumulated:           */
                         /* [0] nodes visited [1] leaves tested [2] stack overflows     */
                         /* [3] queries re-run exactly [4] heap insertions              */
                         /* packet walks, per wave: [5] insertion rounds [6] rounds     */
                         /* that ran the insertion chain                                */
    int32_t window;      /* FT_ALGO_PACKET self queries: leaves scanned on each side of  */
                         /* a packet's own run (subsets: the run around its middle      */
                         /* lane's leaf, widened by up to 4x for sparse subsets) before */
                         /* the walk (0 = auto: 192, negative = none)                   */
    int32_t reserved2;
} ft_query_opts;

/* ---- library / device utilities --------------------------------------- */
const char *ft_last_error(void);
const char *ft_version(void);
int ft_device_count(int *count);
int ft_malloc(void **ptr, size_t bytes);
int ft_free(void *ptr);
/* h2d / d2h return once the copy is done.  d2d and memset are ordered on s; with */
/* s = NULL they also complete before returning.                                 */
int ft_memcpy_h2d(void *dst, const void *src, size_t bytes, ft_stream s);
int ft_memcpy_d2h(void *dst, const void *src, size_t bytes, ft_stream s);
int ft_memcpy_d2d(void *dst, const void *src, size_t bytes, ft_stream s);
int ft_memset(void *dst, int value, size_t bytes, ft_stream s);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned): the       */
/* known-size reference against which hardware byte counters are scaled.         */
int ft_copy_kernel(void *dst, const void *src, size_t bytes, ft_stream s);
/* Queries keep small work buffers per stream (re-run lists, query ordering), so calls  */
/* on different streams are independent; ft_stream_destroy waits for the stream's work  */
/* and frees them.  ft_stream_release does only the latter: call it before destroying a */
/* stream that was not made by ft_stream_create (no counterpart in the reference).      */
int ft_stream_create(ft_stream *s);
int ft_stream_destroy(ft_stream s);
int ft_stream_release(ft_stream s);
/* hipGraph capture of this library's calls on a created stream, e.g. a build_bvh +     */
/* self_knn_query step replayed on moved particles (no reference counterpart).  Run the  */
/* step once eagerly first: calls that would allocate, synchronise or stage host         */
/* pointers during a capture fail it.  ft_graph_end instantiates the captured graph.     */
/* A graph uses the work buffers of the stream it was captured on.  They stay allocated  */
/* until the graph is destroyed, also when a later, larger query on that stream grows    */
/* them or ft_stream_release / ft_stream_destroy drops them, so a replay never reads     */
/* freed memory.  Do not replay a graph while its stream runs other queries (they share  */
/* those buffers).  Trees and build scratch are the caller's: a graph that captured a    */
/* build or query must not be replayed after ft_tree_ensure or a larger build grew them. */
int ft_graph_begin(ft_stream s);
int ft_graph_end(ft_stream s, void **exec);
int ft_graph_launch(void *exec, ft_stream s);
int ft_graph_destroy(void *exec);
int ft_stream_synchronize(ft_stream s);
int ft_event_create(void **ev);
int ft_event_destroy(void *ev);
int ft_event_record(void *ev, ft_stream s);
int ft_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---- tree and scratch lifetime ---------------------------------------- */
/* TreeSoA(q, n) (src/hlbvh.hpp:632-650).  Allocates device memory for n leaves. */
int ft_tree_alloc(ft_tree *tree, size_t n, int32_t coord_type, const ft_config *cfg, ft_stream s);
/* TreePool::ensure (src/hlbvh.hpp:682-715): reuse the allocation when n fits,  */
/* else reallocate with capacity n + n/2 + 1024.                                */
int ft_tree_ensure(ft_tree *tree, size_t n, int32_t coord_type, const ft_config *cfg, ft_stream s);
/* TreeSoA::free(q) (src/hlbvh.hpp:657-672). */
int ft_tree_free(ft_tree *tree, ft_stream s);
/* Serialised trees, for checkpoints and transfers (no reference counterpart; the     */
/* middle section is the reference's MPI transfer sequence, test/test_mpi.cpp:76-100): */
/* a 64-byte header (magic "FTTREE01", coord_type, config), u64 [num_leaves,            */
/* num_internal], min_x..max_z, left_child, right_child, parent, then id, is_ghost,     */
/* orig_idx, sorted_keys, range_first, range_last.  Buffers may be host or device       */
/* memory; both calls return when the copies are done.  Deserialising (re)allocates     */
/* the tree as ft_tree_ensure does.                                                     */
int ft_tree_serialized_size(const ft_tree *tree, size_t *bytes);
int ft_tree_serialize(ft_stream s, const ft_tree *tree, void *buf, size_t bytes);
int ft_tree_deserialize(ft_stream s, ft_tree *tree, const void *buf, size_t bytes);
/* BvhScratch (src/hlbvh.hpp:720-745). */
int ft_scratch_create(ft_scratch **scratch);
int ft_scratch_destroy(ft_scratch *scratch);
/* Stage timing of the next builds: on = record HIP events between build stages. */
int ft_scratch_set_profiling(ft_scratch *scratch, int on);
/* ms of the last profiled build per stage: bbox, encode, sort, assemble (leaf gather,  */
/* topology and in-chunk boxes), assemble_top (nodes spanning chunks).                  */
int ft_scratch_stage_ms(ft_scratch *scratch, double *ms, int max_stages, int *n_stages);
/* Digit width of the next builds' radix sort: 8 or 9 bits per pass (63-bit keys: 8  */
/* or 7 passes), or 0 (default) for the width with fewer passes, 8 on a tie.  The     */
/* sorted order is the same; only the speed differs.  No reference counterpart (its   */
/* sort is oneDPL's); the batched tuner sets it per tree-size class.                  */
int ft_scratch_set_radix_bits(ft_scratch *scratch, int bits);

/* ---- build ------------------------------------------------------------- */
/* compute_bbox (src/hlbvh.hpp:462-578).  out6 (host) = min_x,max_x,min_y,max_y,min_z,max_z. */
int ft_compute_bbox(ft_stream s, int32_t coord_type, const void *x, const void *y, const void *z,
                    size_t n, void *out6);
/* sfc_encode (src/hlbvh.hpp:357-423).  bbox6 (host, coord_t[6]) is used by FP coords only. */
int ft_sfc_encode(ft_stream s, int32_t coord_type, const ft_config *cfg, const void *x,
                  const void *y, const void *z, size_t n, uint64_t *keys, const void *bbox6);
/* build_tree (src/hlbvh.hpp:773-982): Karras topology + refit over sorted keys.  The   */
/* sorted coordinate/attribute arrays are optional as in the reference (NULL = leaves  */
/* already hold them).                                                                 */
int ft_build_tree(ft_stream s, ft_tree *tree, const uint64_t *sorted_keys, const void *sx,
                  const void *sy, const void *sz, const uint32_t *sid, const int8_t *sghost,
                  const int32_t *sorig);
/* fill_particle_coords (test/benchmark/BenchmarkUtils.hpp:164-205): snapshot positions   */
/* (in_type FT_COORD_F32, as the HDF5 PartType0/Coordinates read gives them, or          */
/* FT_COORD_F64) to tree coordinates x, y, z of coord_type, n of each.  Integer           */
/* coord_type: one cube over all three axes, box_min = the smallest coordinate,           */
/* box_size = (largest - box_min) * 1.0001 (1.0 when all are equal), then each value is   */
/* float_to_int_rep(v, box_min, 1 / box_size) (src/sfc.peano_hilbert.hpp:495-503) with    */
/* bits_per_dim bits ([1, 32] for uint32, [1, 52] for uint64).  FP coord_type: a plain    */
/* cast, box (0, 1).  box_out (nullable, host or device): double[2] = {box_min,           */
/* box_size}.  Host arrays are staged.  Positions must be finite; the sign of a zero      */
/* box_min may differ from the reference's, which keeps the first of equal minima.        */
int ft_fill_particle_coords(ft_stream s, int32_t in_type, const void *px, const void *py,
                            const void *pz, size_t n, int32_t coord_type, int bits_per_dim,
                            void *x, void *y, void *z, double *box_out);
/* build_bvh (src/hlbvh.hpp:2165-2302).  id / is_ghost may be NULL (iota / zeros).  bbox6 */
/* (host coord_t[6]) and keys (n unsorted SFC keys) are optional precomputed inputs.     */
/* scratch may be NULL (a temporary one is used, as the reference's 3-arg overload).     */
int ft_build_bvh(ft_stream s, ft_tree *tree, ft_scratch *scratch, const void *x, const void *y,
                 const void *z, const uint32_t *id, const int8_t *is_ghost, size_t n,
                 const void *bbox6, const uint64_t *keys);

/* ---- queries ------------------------------------------------------------ */
/* self_knn_query (src/hlbvh.hpp:1929-1954).  results: uint64_t[n*k] (size_t in the     */
/* reference), dists: dist_t[n*k].  Missing neighbours: index UINT64_MAX, distance       */
/* -DBL_MAX (FP) or 0 (integer), as type_identity_max (src/hlbvh.hpp:59-69).             */
int ft_self_knn_query(ft_stream s, const ft_tree *tree, int k, uint64_t *results, void *dists,
                      int exclude_self, const ft_query_opts *opts);
/* Non-allocating self_knn_query: no counterpart in the reference.  For callers that     */
/* cannot allocate device memory (code under hipGraph capture, tight serving loops,       */
/* harnesses that forbid hipMalloc).  Same arguments and results as ft_self_knn_query;    */
/* the re-run buffers live in a fixed device array of 80 MiB per GPU, held until the      */
/* process ends, so at most 16777216 queries (num_leaves; FT_ERANGE above).  results,     */
/* dists, seed_bound and stats must be device (or pinned host) memory (FT_EINVAL          */
/* otherwise).  Every call on one GPU shares that array: calls that overlap on one GPU    */
/* (on different streams or threads) race.                                                */
int ft_self_knn_query_noalloc(ft_stream s, const ft_tree *tree, int k, uint64_t *results,
                              void *dists, int exclude_self, const ft_query_opts *opts);
/* self_knn_query_subset (src/hlbvh.hpp:1956-1983): leaf ranks, results in query order. */
int ft_self_knn_query_subset(ft_stream s, const ft_tree *tree, const int32_t *leaf_ids,
                             int num_ids, int k, uint64_t *results, void *dists,
                             int exclude_self, const ft_query_opts *opts);
/* knn_query (src/hlbvh.hpp:1405-1472): external query points. */
int ft_knn_query(ft_stream s, const ft_tree *tree, const void *qx, const void *qy,
                 const void *qz, int k, int num_queries, uint64_t *results, void *dists,
                 const ft_query_opts *opts);
/* range_query (src/hlbvh.hpp:2006-2136): fixed stride output, counts = total hits. */
int ft_range_query(ft_stream s, const ft_tree *tree, const void *qx, const void *qy,
                   const void *qz, const void *r_min, const void *r_max, int num_queries,
                   int32_t *results, int32_t *result_counts, int max_results_per_query);
/* Range query with CSR output (no reference counterpart): row_offsets int64_t[Q+1],   */
/* cols int32_t[capacity].  Pass cols = NULL to get only counts/offsets (total in      */
/* row_offsets[Q]); hits beyond col_capacity return FT_ERANGE after filling offsets.   */
/* With cols, one walk fills both (it uses 512 bytes of scratch per query; rows of   */
/* more than 128 hits are walked a second time).                                      */
int ft_range_query_csr(ft_stream s, const ft_tree *tree, const void *qx, const void *qy,
                       const void *qz, const void *r_min, const void *r_max, int num_queries,
                       int64_t *row_offsets, int32_t *cols, int64_t col_capacity);

/* ---- learned search radius (no reference counterpart) --------------------- */
/* A 12-64-64-64-1 ReLU MLP (bf16 weights, fp32 accumulation on MFMA) predicts     */
/* log d2_k of a self query from tree-density features; its prediction plus a      */
/* margin seeds ft_query_opts.seed_bound.  Results are unchanged by the seed.      */
typedef struct ft_mlp ft_mlp;
int ft_mlp_num_params(void);
int ft_mlp_create(ft_mlp **model);
int ft_mlp_destroy(ft_mlp *model);
/* params: ft_mlp_num_params() floats; norm: 34 floats (feature mean[16],          */
/* 1/std[16], target mean, target std) -- host memory; NULL keeps the current.     */
int ft_mlp_set(ft_mlp *model, const float *params, const float *norm, double margin);
int ft_mlp_get(ft_mlp *model, float *params, float *norm, double *margin);
/* Feature rows float[nq][16] (device) for self queries (leaf_ids NULL = all).     */
int ft_knn_features(ft_stream s, const ft_tree *tree, const int32_t *leaf_ids, int nq, int k,
                    float *features);
/* Predicted log d2_k, float[n] (device pointers).                                 */
int ft_mlp_forward(ft_stream s, ft_mlp *model, const float *features, int n, float *pred);
/* Gradient float[num_params] and loss of the mean squared error of the             */
/* standardised target over n samples (device pointers).                           */
int ft_mlp_grad(ft_stream s, ft_mlp *model, const float *features, const float *target, int n,
                float *grad, float *loss);
/* One Adam step over n samples (device pointers).                                 */
int ft_mlp_train_step(ft_stream s, ft_mlp *model, const float *features, const float *target,
                      int n, float lr);
int ft_mlp_last_loss(ft_stream s, ft_mlp *model, float *loss);
/* Seed bounds dist_t[nq] (device) for self queries: features + forward + margin.  */
/* scratch_features: nq * 16 floats of device memory.                              */
int ft_mlp_seed(ft_stream s, ft_mlp *model, const ft_tree *tree, const int32_t *leaf_ids,
                int nq, int k, float *scratch_features, void *seed);

/* ---- batched trees and mixed-k queries in one graph (no reference counterpart) ---- */
/* Many trees of uneven size, and query requests of mixed k over them, run as one    */
/* step: every tree is built (build_bvh), the queries are ordered along each tree's  */
/* curve, and persistent kernels per heap tier drain a work list over all trees.     */
/* The step can be captured in a hipGraph and replayed with one launch.  Its closest */
/* reference analogue is the compile-time group-width / heap-tier dispatch          */
/* (src/hlbvh.hpp:1433-1453, 1841-1927).  Results equal knn_query's canonical        */
/* results on each tree.  Pointers given to a batch must stay valid while it is used.*/
typedef struct ft_batch ft_batch;
typedef struct ft_batch_request {
    int32_t tree;         /* index of the tree the queries search                    */
    int32_t k;            /* neighbours per query, 1..256                            */
    int64_t query_first;  /* first query point (index into qx/qy/qz)                 */
    int64_t num_queries;
    int64_t result_first; /* first result slot (element index into results/dists)    */
} ft_batch_request;
int ft_batch_create(ft_batch **batch, int32_t coord_type);
int ft_batch_destroy(ft_batch *batch);
/* trees: allocated trees of the batch's coordinate type.  x/y/z/n: per-tree build    */
/* inputs (device pointers); x = NULL takes the trees as built (queries only).        */
int ft_batch_set_trees(ft_batch *batch, int num_trees, ft_tree *const *trees,
                       const void *const *x, const void *const *y, const void *const *z,
                       const size_t *n);
/* Query points (device, coord_t) and outputs (device): results uint64_t, dists      */
/* dist_t, row-major per query at each request's result_first.                        */
int ft_batch_set_queries(ft_batch *batch, int num_requests, const ft_batch_request *reqs,
                         const void *qx, const void *qy, const void *qz, uint64_t *results,
                         void *dists);
/* Optional device uint64_t[8] counters as in ft_query_opts.stats.                    */
int ft_batch_set_stats(ft_batch *batch, uint64_t *stats);
/* Autotuner: runs the step's parts on the bound data and keeps, per (tree-size      */
/* class, k tier), the fastest tier kernel and leaf bucket, and the fastest build    */
/* branch count.  Synchronises the stream.                                            */
int ft_batch_tune(ft_batch *batch, ft_stream s);
/* Pins the setting of one (tree-size class, k tier) group, which the autotuner then  */
/* keeps: size classes 0..3 = trees below 40k, 160k, 640k leaves and larger; k tiers  */
/* 0..7 = k up to 1, 4, 8, 16, 32, 64, 128 and above.  tier: a tier name as           */
/* ft_batch_tuning prints it ("packet-reg32", "packet-reg32-passes", "lane-heap",     */
/* "wave-merge128", ...) that serves the k tier; bucket: packet leaf bucket (1..64,   */
/* 0 = 32); queries_per_item: packet tiers 16, 32 or 64, wave tiers 1..64 (0 = the    */
/* default).  No reference counterpart (the reference fixes these at compile time).   */
int ft_batch_set_tier(ft_batch *batch, int32_t n_class, int32_t k_tier, const char *tier,
                      int32_t bucket, int32_t queries_per_item);
/* One step, launched kernel by kernel (no graph).                                    */
int ft_batch_run(ft_batch *batch, ft_stream s);
/* Records the step into a hipGraph (allocates nothing while recording).             */
int ft_batch_capture(ft_batch *batch, ft_stream s);
/* Replays the recorded step: one graph launch (captures first if needed).           */
int ft_batch_launch(ft_batch *batch, ft_stream s);
/* The settings in use, as JSON text.                                                  */
int ft_batch_tuning(ft_batch *batch, char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* FASTTREE_C_H */
