// ft_capi.hip -- the C ABI (include/fasttree_c.h): argument checks with the
// reference's error behaviour, host-pointer staging, tree / scratch lifetime.
#include "ft_internal.hpp"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <limits>
#include <mutex>
#include <vector>

namespace ft {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// Defined in ft_build.hip / ft_query.hip.
template <typename T>
int build_bvh_t(hipStream_t s, ft_tree *t, Scratch &S, const T *x, const T *y, const T *z,
                const uint32_t *id, const int8_t *ghost, size_t n, const T *bbox6,
                const uint64_t *keys);
template <typename T>
int compute_bbox_t(hipStream_t s, Scratch &S, const T *x, const T *y, const T *z, size_t n,
                   T *out6);
template <typename T>
int sfc_encode_t(hipStream_t s, Scratch &S, const ft_config &cfg, const T *x, const T *y,
                 const T *z, size_t n, uint64_t *keys, const T *bbox6);
template <typename T>
int build_tree_t(hipStream_t s, ft_tree *t, Scratch &S, const uint64_t *sorted_keys, const T *sx,
                 const T *sy, const T *sz, const uint32_t *sid, const int8_t *sghost,
                 const int32_t *sorig);
int collect_profile(Scratch &S);
int knn_entry(hipStream_t s, const ft_tree *tree, const void *qx, const void *qy, const void *qz,
              const int32_t *leaf_ids, int nq, int k, int excl, uint64_t *res, void *dists,
              const ft_query_opts *opts);
int range_entry(hipStream_t s, const ft_tree *tree, const void *qx, const void *qy,
                const void *qz, const void *rmin, const void *rmax, int nq, int32_t *res,
                int32_t *counts, int max_res, const int64_t *row_ptr, int32_t *cols,
                int64_t col_cap);
int scan_entry(hipStream_t s, const int32_t *counts, int nq, int64_t *row_ptr);
int release_stream_bufs(hipStream_t s);
void graph_attach(hipStream_t s, void *exec);
int knn_noalloc_entry(hipStream_t s, const ft_tree *tree, int k, int excl, uint64_t *res,
                      void *dists, const ft_query_opts *opts);
size_t noalloc_max_queries();
size_t snapshot_work_doubles();
int snapshot_entry(hipStream_t s, int in_type, const void *px, const void *py, const void *pz,
                   size_t n, int coord_type, int bits, void *x, void *y, void *z, double *work);
void graph_detach(void *exec);
int range_csr_entry(hipStream_t s, const ft_tree *tree, const void *qx, const void *qy,
                    const void *qz, const void *rmin, const void *rmax, int nq, int64_t *row_ptr,
                    int32_t *(*cols_for)(int64_t total, void *ctx), void *ctx, int64_t col_cap,
                    int64_t *total);

// ---------------------------------------------------------------------------
// Staging
// ---------------------------------------------------------------------------
bool is_device_accessible(const void *p) {
    if (!p) return true;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // plain host memory: clear the lookup error
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
           a.type == hipMemoryTypeHost || a.type == hipMemoryTypeUnified;
}

// Releases what st still owns and clears it.
static void stage_reset(Staged &st) {
    if (st.owned) (void)hipFreeAsync(st.dev, st.stream);
    st.dev = nullptr;
    st.host_src = nullptr;
    st.host_dst = nullptr;
    st.bytes = 0;
    st.owned = false;
    st.stream = nullptr;
}

int stage_in(const void *p, size_t bytes, hipStream_t s, Staged &st) {
    stage_reset(st);
    if (!p || bytes == 0 || is_device_accessible(p)) {
        st.dev = const_cast<void *>(p);
        return FT_OK;
    }
    void *d = nullptr;
    if (hipMallocAsync(&d, bytes, s) != hipSuccess) return fail(FT_ENOMEM, "staging allocation failed");
    if (hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, s) != hipSuccess) {
        (void)hipFreeAsync(d, s);
        return fail(FT_EHIP, "staging copy to the device failed");
    }
    st.dev = d;
    st.host_src = p;
    st.bytes = bytes;
    st.owned = true;
    st.stream = s;
    return FT_OK;
}

int stage_out(void *p, size_t bytes, hipStream_t s, Staged &st) {
    stage_reset(st);
    if (!p || bytes == 0 || is_device_accessible(p)) {
        st.dev = p;
        return FT_OK;
    }
    void *d = nullptr;
    if (hipMallocAsync(&d, bytes, s) != hipSuccess) return fail(FT_ENOMEM, "staging allocation failed");
    st.dev = d;
    st.host_dst = p;
    st.bytes = bytes;
    st.owned = true;
    st.stream = s;
    return FT_OK;
}

int unstage(Staged &st, hipStream_t s, bool copy_back) {
    if (!st.owned) return FT_OK;
    if (copy_back && st.host_dst) {
        FT_HIP(hipMemcpyAsync(st.host_dst, st.dev, st.bytes, hipMemcpyDeviceToHost, s));
        FT_HIP(hipStreamSynchronize(s));
    }
    FT_HIP(hipFreeAsync(st.dev, s));
    st.owned = false;
    return FT_OK;
}

// Process-wide scratch for the helper entry points (compute_bbox, sfc_encode,
// build_tree).  Calls that use it are serialised by a mutex.
static Scratch g_util;
static std::mutex g_util_mu;

static int valid_config(const ft_config *c, int ct) {
    if (!c) return fail(FT_EINVAL, "null configuration");
    if (c->sfc != FT_SFC_MORTON && c->sfc != FT_SFC_PEANO_HILBERT)
        return fail(FT_EINVAL, "unknown space-filling curve");
    if (c->bits_per_dim < 1 || c->bits_per_dim > 21)
        return fail(FT_EINVAL, "bits_per_dim must be in [1, 21] (3B-bit keys in 64 bits)");
    if (ct == FT_COORD_F32 && c->sfc == FT_SFC_MORTON && c->bits_per_dim > 23)
        return fail(FT_EINVAL, "BITS_PER_DIMENSION exceeds the mantissa precision of float");
    if (c->periodic && (ct == FT_COORD_F32 || ct == FT_COORD_F64) && !(c->box_length > 0.0))
        return fail(FT_EINVAL, "periodic FP coordinates need box_length > 0");
    return FT_OK;
}

}  // namespace ft

using namespace ft;

extern "C" {

const char *ft_last_error(void) { return g_err.c_str(); }
const char *ft_version(void) { return "fasttree-mi355x 0.1.0 (gfx950)"; }

int ft_device_count(int *count) {
    FT_HIP(hipGetDeviceCount(count));
    return FT_OK;
}
int ft_malloc(void **ptr, size_t bytes) {
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) return fail(FT_ENOMEM, "hipMalloc failed");
    return FT_OK;
}
int ft_free(void *ptr) {
    FT_HIP(hipFree(ptr));
    return FT_OK;
}
int ft_memcpy_h2d(void *dst, const void *src, size_t bytes, ft_stream s) {
    FT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(s)));
    FT_HIP(hipStreamSynchronize(as_stream(s)));
    return FT_OK;
}
int ft_memcpy_d2h(void *dst, const void *src, size_t bytes, ft_stream s) {
    FT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(s)));
    FT_HIP(hipStreamSynchronize(as_stream(s)));
    return FT_OK;
}
// Without a stream these run on the null stream, which non-blocking streams
// do not wait for: they then complete before returning.
int ft_memcpy_d2d(void *dst, const void *src, size_t bytes, ft_stream s) {
    FT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(s)));
    if (!s) FT_HIP(hipStreamSynchronize(nullptr));
    return FT_OK;
}
int ft_memset(void *dst, int value, size_t bytes, ft_stream s) {
    FT_HIP(hipMemsetAsync(dst, value, bytes, as_stream(s)));
    if (!s) FT_HIP(hipStreamSynchronize(nullptr));
    return FT_OK;
}

// A plain streaming copy kernel (16 bytes per lane per step): the known-size
// reference for hardware byte counters (bytes read = bytes written = n).
__global__ __launch_bounds__(256) void copy16_kernel(const uint4 *__restrict__ src,
                                                     uint4 *__restrict__ dst, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int ft_copy_kernel(void *dst, const void *src, size_t bytes, ft_stream s) {
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return fail(FT_EINVAL, "ft_copy_kernel: 16-byte aligned sizes and pointers only");
    const size_t n16 = bytes / 16;
    if (!n16) return FT_OK;
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(s),
                       (const uint4 *)src, (uint4 *)dst, n16);
    FT_CHECK_LAUNCH();
    return FT_OK;
}
int ft_stream_create(ft_stream *s) {
    hipStream_t h;
    FT_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = (ft_stream)h;
    return FT_OK;
}
int ft_stream_release(ft_stream s) { return release_stream_bufs(as_stream(s)); }
int ft_stream_destroy(ft_stream s) {
    (void)release_stream_bufs(as_stream(s));
    FT_HIP(hipStreamDestroy(as_stream(s)));
    return FT_OK;
}
int ft_stream_synchronize(ft_stream s) {
    FT_HIP(hipStreamSynchronize(as_stream(s)));
    return FT_OK;
}
// hipGraph capture of the calls issued on a stream (relaxed mode: other
// threads may keep allocating).  Calls that would allocate, copy to the host
// or synchronise (growing a tree or scratch, staging host pointers, the CSR
// range query) end the capture with an error; run the step once eagerly
// first so that every buffer has its size.
int ft_graph_begin(ft_stream s) {
    if (!s) return fail(FT_EINVAL, "graph capture needs a created stream, not the null stream");
    FT_HIP(hipStreamBeginCapture(as_stream(s), hipStreamCaptureModeRelaxed));
    return FT_OK;
}
int ft_graph_end(ft_stream s, void **exec) {
    if (!s || !exec) return fail(FT_EINVAL, "null argument");
    hipGraph_t g = nullptr;
    FT_HIP(hipStreamEndCapture(as_stream(s), &g));
    hipGraphExec_t e = nullptr;
    const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (rc != hipSuccess) return fail(FT_EHIP, "hipGraphInstantiate failed");
    graph_attach(as_stream(s), (void *)e);  // the stream's work buffers outlive the graph
    *exec = (void *)e;
    return FT_OK;
}
int ft_graph_launch(void *exec, ft_stream s) {
    if (!exec) return fail(FT_EINVAL, "null graph");
    FT_HIP(hipGraphLaunch((hipGraphExec_t)exec, as_stream(s)));
    return FT_OK;
}
int ft_graph_destroy(void *exec) {
    if (!exec) return FT_OK;
    const hipError_t rc = hipGraphExecDestroy((hipGraphExec_t)exec);
    graph_detach(exec);
    FT_HIP(rc);
    return FT_OK;
}
int ft_event_create(void **ev) {
    hipEvent_t e;
    FT_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return FT_OK;
}
int ft_event_destroy(void *ev) {
    FT_HIP(hipEventDestroy((hipEvent_t)ev));
    return FT_OK;
}
int ft_event_record(void *ev, ft_stream s) {
    FT_HIP(hipEventRecord((hipEvent_t)ev, as_stream(s)));
    return FT_OK;
}
int ft_event_elapsed_ms(void *start, void *stop, float *ms) {
    FT_HIP(hipEventSynchronize((hipEvent_t)stop));
    FT_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return FT_OK;
}

// ---- tree --------------------------------------------------------------------
static int tree_carve(ft_tree *t, size_t cap, int32_t ct, ft_stream s) {
    const size_t cs = coord_size(ct);
    const size_t nodes = 2 * cap - 1, inner = cap > 1 ? cap - 1 : 1;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    size_t o_c[6];
    for (int a = 0; a < 6; ++a) {
        o_c[a] = off;
        off += al(nodes * cs);
    }
    const size_t o_l = off; off += al(inner * 4);
    const size_t o_r = off; off += al(inner * 4);
    const size_t o_p = off; off += al(nodes * 4);
    const size_t o_id = off; off += al(nodes * 4);
    const size_t o_g = off; off += al(nodes);
    const size_t o_o = off; off += al(nodes * 4);
    const size_t o_k = off; off += al(cap * 8);
    const size_t o_rf = off; off += al(inner * 4);
    const size_t o_rl = off; off += al(inner * 4);
    void *blk = nullptr;
    if (hipMalloc(&blk, off) != hipSuccess) return fail(FT_ENOMEM, "tree allocation failed");
    char *b = (char *)blk;
    t->min_x = b + o_c[0]; t->max_x = b + o_c[1];
    t->min_y = b + o_c[2]; t->max_y = b + o_c[3];
    t->min_z = b + o_c[4]; t->max_z = b + o_c[5];
    t->left_child = (int32_t *)(b + o_l);
    t->right_child = (int32_t *)(b + o_r);
    t->parent = (int32_t *)(b + o_p);
    t->id = (uint32_t *)(b + o_id);
    t->is_ghost = (int8_t *)(b + o_g);
    t->orig_idx = (int32_t *)(b + o_o);
    t->sorted_keys = (uint64_t *)(b + o_k);
    t->range_first = (int32_t *)(b + o_rf);
    t->range_last = (int32_t *)(b + o_rl);
    t->block = blk;
    t->capacity = cap;
    (void)s;
    return FT_OK;
}

int ft_tree_alloc(ft_tree *tree, size_t n, int32_t coord_type, const ft_config *cfg, ft_stream s) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    std::memset(tree, 0, sizeof(*tree));
    FT_TRY(valid_config(cfg, coord_type));
    if (coord_type < FT_COORD_F32 || coord_type > FT_COORD_U64)
        return fail(FT_EINVAL, "unknown coordinate type");
    tree->coord_type = coord_type;
    tree->config = *cfg;
    tree->num_leaves = n;
    tree->num_internal = n > 0 ? n - 1 : 0;
    if (n == 0) return FT_OK;
    if (n > (size_t)(INT_MAX / 2))
        return fail(FT_ERANGE, "Particle count n exceeds maximum supported tree size (INT_MAX / 2).");
    return tree_carve(tree, n, coord_type, s);
}

int ft_tree_ensure(ft_tree *tree, size_t n, int32_t coord_type, const ft_config *cfg, ft_stream s) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    FT_TRY(valid_config(cfg, coord_type));
    if (n > (size_t)(INT_MAX / 2))
        return fail(FT_ERANGE, "Particle count n exceeds maximum supported tree size (INT_MAX / 2).");
    if (!tree->block || n > tree->capacity || tree->coord_type != coord_type) {
        if (tree->block) FT_HIP(hipFree(tree->block));
        tree->block = nullptr;
        const size_t cap = n + n / 2 + 1024;
        FT_TRY(tree_carve(tree, cap, coord_type, s));
    }
    tree->coord_type = coord_type;
    tree->config = *cfg;
    tree->num_leaves = n;
    tree->num_internal = n > 0 ? n - 1 : 0;
    return FT_OK;
}

int ft_tree_free(ft_tree *tree, ft_stream s) {
    (void)s;
    if (!tree) return FT_OK;
    if (tree->block) FT_HIP(hipFree(tree->block));
    const int32_t ct = tree->coord_type;
    const ft_config cfg = tree->config;
    std::memset(tree, 0, sizeof(*tree));
    tree->coord_type = ct;
    tree->config = cfg;
    return FT_OK;
}

// ---- serialised trees (no reference counterpart) --------------------------------
// Layout: a 64-byte header (magic "FTTREE01", coord_type, reserved, the
// ft_config), then the reference's transfer sequence (test/test_mpi.cpp:76-100):
// u64 [num_leaves, num_internal], min_x, max_x, min_y, max_y, min_z, max_z
// (2n-1 coordinates each), left_child, right_child (n-1 int32), parent (2n-1
// int32); then id (u32), is_ghost (i8), orig_idx (int32) over 2n-1 nodes,
// sorted_keys (n u64), range_first, range_last (n-1 int32).  Native byte order.
namespace {
constexpr char kTreeMagic[8] = {'F', 'T', 'T', 'R', 'E', 'E', '0', '1'};
struct TreeHeader {
    char magic[8];
    int32_t coord_type;
    int32_t reserved;
    ft_config config;
    uint8_t pad[64 - 16 - sizeof(ft_config)];
};
static_assert(sizeof(TreeHeader) == 64, "tree header is 64 bytes");

struct TreePart {
    const void *ptr;
    size_t bytes;
};
// The arrays of a tree of n leaves in serialised order (pointers into `t`).
static std::vector<TreePart> tree_parts(const ft_tree &t, size_t n) {
    const size_t cs = coord_size(t.coord_type);
    const size_t m = n ? 2 * n - 1 : 0, in = n ? n - 1 : 0;
    return {{t.min_x, m * cs},      {t.max_x, m * cs},          {t.min_y, m * cs},
            {t.max_y, m * cs},      {t.min_z, m * cs},          {t.max_z, m * cs},
            {t.left_child, in * 4}, {t.right_child, in * 4},    {t.parent, m * 4},
            {t.id, m * 4},          {t.is_ghost, m},            {t.orig_idx, m * 4},
            {t.sorted_keys, n * 8}, {t.range_first, in * 4},    {t.range_last, in * 4}};
}
static size_t tree_bytes(int32_t coord_type, size_t n) {
    ft_tree t{};
    t.coord_type = coord_type;
    size_t b = sizeof(TreeHeader) + 16;
    for (const TreePart &p : tree_parts(t, n)) b += p.bytes;
    return b;
}
}  // namespace

// Checks, on the host, every index a traversal follows in a serialised tree
// of n leaves (buffer `in`, host or device memory): children in [0, 2n-1),
// every node but the root the child of exactly one internal node whose label
// its parent entry names (so the nodes reachable from the root form a tree and
// every walk ends), and leaf ranges 0 <= first <= last < n.
static int check_topology(const char *in, int32_t coord_type, size_t n, hipStream_t hs) {
    if (n < 2) return FT_OK;
    ft_tree t{};
    t.coord_type = coord_type;
    const std::vector<TreePart> parts = tree_parts(t, n);
    std::vector<size_t> off(parts.size());
    size_t at = sizeof(TreeHeader) + 16;
    for (size_t i = 0; i < parts.size(); ++i) {
        off[i] = at;
        at += parts[i].bytes;
    }
    const size_t in_n = n - 1, m = 2 * n - 1;
    std::vector<int32_t> left(in_n), right(in_n), parent(m), rf(in_n), rl(in_n);
    auto get = [&](std::vector<int32_t> &v, int part) -> int {
        FT_HIP(hipMemcpyAsync(v.data(), in + off[part], v.size() * 4, hipMemcpyDefault, hs));
        return FT_OK;
    };
    FT_TRY(get(left, 6));
    FT_TRY(get(right, 7));
    FT_TRY(get(parent, 8));
    FT_TRY(get(rf, 13));
    FT_TRY(get(rl, 14));
    FT_HIP(hipStreamSynchronize(hs));
    std::vector<uint8_t> seen(m, 0);
    for (size_t i = 0; i < in_n; ++i) {
        for (const int32_t c : {left[i], right[i]}) {
            if (c <= 0 || (size_t)c >= m || seen[c] || parent[c] != (int32_t)i)
                return fail(FT_EINVAL, "serialised tree: corrupt child or parent index");
            seen[c] = 1;
        }
        if (rf[i] < 0 || rf[i] > rl[i] || (size_t)rl[i] >= n)
            return fail(FT_EINVAL, "serialised tree: corrupt leaf range");
    }
    if (parent[0] != -1) return fail(FT_EINVAL, "serialised tree: corrupt root");
    return FT_OK;
}

int ft_tree_serialized_size(const ft_tree *tree, size_t *bytes) {
    if (!tree || !bytes) return fail(FT_EINVAL, "null argument");
    *bytes = tree_bytes(tree->coord_type, tree->num_leaves);
    return FT_OK;
}

int ft_tree_serialize(ft_stream s, const ft_tree *tree, void *buf, size_t bytes) {
    if (!tree || !buf) return fail(FT_EINVAL, "null argument");
    const size_t n = tree->num_leaves;
    if (bytes < tree_bytes(tree->coord_type, n))
        return fail(FT_ERANGE, "serialisation buffer is smaller than ft_tree_serialized_size");
    const hipStream_t hs = as_stream(s);
    TreeHeader h{};
    std::memcpy(h.magic, kTreeMagic, 8);
    h.coord_type = tree->coord_type;
    h.config = tree->config;
    const uint64_t dims[2] = {(uint64_t)n, (uint64_t)tree->num_internal};
    char *o = (char *)buf;
    FT_HIP(hipMemcpyAsync(o, &h, sizeof h, hipMemcpyDefault, hs));
    FT_HIP(hipMemcpyAsync(o + sizeof h, dims, 16, hipMemcpyDefault, hs));
    size_t at = sizeof h + 16;
    for (const TreePart &p : tree_parts(*tree, n)) {
        if (p.bytes) FT_HIP(hipMemcpyAsync(o + at, p.ptr, p.bytes, hipMemcpyDefault, hs));
        at += p.bytes;
    }
    FT_HIP(hipStreamSynchronize(hs));  // the header and dims live on this stack frame
    return FT_OK;
}

int ft_tree_deserialize(ft_stream s, ft_tree *tree, const void *buf, size_t bytes) {
    if (!tree || !buf) return fail(FT_EINVAL, "null argument");
    const hipStream_t hs = as_stream(s);
    if (bytes < sizeof(TreeHeader) + 16) return fail(FT_ERANGE, "serialised tree is truncated");
    TreeHeader h{};
    uint64_t dims[2] = {0, 0};
    const char *in = (const char *)buf;
    FT_HIP(hipMemcpyAsync(&h, in, sizeof h, hipMemcpyDefault, hs));
    FT_HIP(hipMemcpyAsync(dims, in + sizeof h, 16, hipMemcpyDefault, hs));
    FT_HIP(hipStreamSynchronize(hs));
    if (std::memcmp(h.magic, kTreeMagic, 8) != 0) return fail(FT_EINVAL, "not a serialised tree");
    const size_t n = (size_t)dims[0];
    if (dims[1] != (n ? n - 1 : 0)) return fail(FT_EINVAL, "serialised tree: inconsistent sizes");
    if (h.coord_type < FT_COORD_F32 || h.coord_type > FT_COORD_U64)
        return fail(FT_EINVAL, "serialised tree: unknown coordinate type");
    if (bytes < tree_bytes(h.coord_type, n)) return fail(FT_ERANGE, "serialised tree is truncated");
    if (n > (size_t)(INT_MAX / 2)) return fail(FT_EINVAL, "serialised tree: too many leaves");
    FT_TRY(valid_config(&h.config, h.coord_type));
    FT_TRY(check_topology(in, h.coord_type, n, hs));
    FT_TRY(ft_tree_ensure(tree, n, h.coord_type, &h.config, s));
    size_t at = sizeof h + 16;
    for (const TreePart &p : tree_parts(*tree, n)) {
        if (p.bytes) FT_HIP(hipMemcpyAsync((void *)p.ptr, in + at, p.bytes, hipMemcpyDefault, hs));
        at += p.bytes;
    }
    FT_HIP(hipStreamSynchronize(hs));
    return FT_OK;
}

// ---- scratch -------------------------------------------------------------------
int ft_scratch_create(ft_scratch **scratch) {
    if (!scratch) return fail(FT_EINVAL, "null scratch pointer");
    *scratch = new ft_scratch();
    return FT_OK;
}
int ft_scratch_destroy(ft_scratch *scratch) {
    if (!scratch) return FT_OK;
    Scratch &S = scratch->s;
    if (S.block) (void)hipFree(S.block);
    for (auto &e : S.ev)
        if (e) (void)hipEventDestroy(e);
    delete scratch;
    return FT_OK;
}
int ft_scratch_set_profiling(ft_scratch *scratch, int on) {
    if (!scratch) return fail(FT_EINVAL, "null scratch");
    scratch->s.profiling = on;
    return FT_OK;
}
int ft_scratch_set_radix_bits(ft_scratch *scratch, int bits) {
    if (!scratch) return fail(FT_EINVAL, "null scratch");
    if (bits != 0 && bits != 8 && bits != 9) return fail(FT_EINVAL, "radix bits must be 0, 8 or 9");
    scratch->s.radix_bits = bits;
    return FT_OK;
}
int ft_scratch_stage_ms(ft_scratch *scratch, double *ms, int max_stages, int *n_stages) {
    if (!scratch) return fail(FT_EINVAL, "null scratch");
    FT_TRY(collect_profile(scratch->s));
    const int ns = scratch->s.n_stages;
    for (int i = 0; i < ns && i < max_stages; ++i) ms[i] = scratch->s.stage_ms[i];
    if (n_stages) *n_stages = ns;
    return FT_OK;
}

// ---- build ---------------------------------------------------------------------
int ft_compute_bbox(ft_stream s, int32_t coord_type, const void *x, const void *y, const void *z,
                    size_t n, void *out6) {
    if (!out6) return fail(FT_EINVAL, "null output");
    if (n == 0) {
        // type_identity_min / type_identity_max (src/hlbvh.hpp:47-69, :466-474)
        FT_DISPATCH_COORD(coord_type, {
            T *o = (T *)out6;
            T lo, hi;
            if constexpr (CoordTraits<T>::fp) {
                lo = std::numeric_limits<T>::max();
                hi = -std::numeric_limits<T>::max();
            } else {
                lo = std::numeric_limits<T>::max();
                hi = 0;
            }
            o[0] = lo; o[1] = hi; o[2] = lo; o[3] = hi; o[4] = lo; o[5] = hi;
        });
        return FT_OK;
    }
    const hipStream_t hs = as_stream(s);
    const size_t bytes = n * coord_size(coord_type);
    Staged sx, sy, sz;
    FT_TRY(stage_in(x, bytes, hs, sx));
    FT_TRY(stage_in(y, bytes, hs, sy));
    FT_TRY(stage_in(z, bytes, hs, sz));
    int rc;
    {
        std::lock_guard<std::mutex> lk(g_util_mu);
        FT_DISPATCH_COORD(coord_type, rc = compute_bbox_t<T>(hs, g_util, (const T *)sx.dev,
                                                             (const T *)sy.dev, (const T *)sz.dev,
                                                             n, (T *)out6));
    }
    unstage(sx, hs, false);
    unstage(sy, hs, false);
    unstage(sz, hs, false);
    return rc;
}

int ft_sfc_encode(ft_stream s, int32_t coord_type, const ft_config *cfg, const void *x,
                  const void *y, const void *z, size_t n, uint64_t *keys, const void *bbox6) {
    FT_TRY(valid_config(cfg, coord_type));
    if (n == 0) return FT_OK;
    const hipStream_t hs = as_stream(s);
    const size_t bytes = n * coord_size(coord_type);
    Staged sx, sy, sz, sk;
    FT_TRY(stage_in(x, bytes, hs, sx));
    FT_TRY(stage_in(y, bytes, hs, sy));
    FT_TRY(stage_in(z, bytes, hs, sz));
    FT_TRY(stage_out(keys, n * 8, hs, sk));
    int rc;
    {
        std::lock_guard<std::mutex> lk(g_util_mu);
        FT_DISPATCH_COORD(coord_type, rc = sfc_encode_t<T>(hs, g_util, *cfg, (const T *)sx.dev,
                                                           (const T *)sy.dev, (const T *)sz.dev,
                                                           n, (uint64_t *)sk.dev,
                                                           (const T *)bbox6));
    }
    if (rc == FT_OK) rc = unstage(sk, hs, true);
    unstage(sx, hs, false);
    unstage(sy, hs, false);
    unstage(sz, hs, false);
    return rc;
}

int ft_build_tree(ft_stream s, ft_tree *tree, const uint64_t *sorted_keys, const void *sx,
                  const void *sy, const void *sz, const uint32_t *sid, const int8_t *sghost,
                  const int32_t *sorig) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    const size_t n = tree->num_leaves;
    if (n == 0) return FT_OK;
    if (!tree->block) return fail(FT_EINVAL, "tree is not allocated");
    const hipStream_t hs = as_stream(s);
    const size_t bytes = n * coord_size(tree->coord_type);
    Staged k, x, y, z, i, g, o;
    FT_TRY(stage_in(sorted_keys, n * 8, hs, k));
    FT_TRY(stage_in(sx, bytes, hs, x));
    FT_TRY(stage_in(sy, bytes, hs, y));
    FT_TRY(stage_in(sz, bytes, hs, z));
    FT_TRY(stage_in(sid, n * 4, hs, i));
    FT_TRY(stage_in(sghost, n, hs, g));
    FT_TRY(stage_in(sorig, n * 4, hs, o));
    int rc;
    {
        std::lock_guard<std::mutex> lk(g_util_mu);
        FT_DISPATCH_COORD(tree->coord_type,
                          rc = build_tree_t<T>(hs, tree, g_util, (const uint64_t *)k.dev,
                                               (const T *)x.dev, (const T *)y.dev,
                                               (const T *)z.dev, (const uint32_t *)i.dev,
                                               (const int8_t *)g.dev, (const int32_t *)o.dev));
    }
    for (Staged *st : {&k, &x, &y, &z, &i, &g, &o}) unstage(*st, hs, false);
    return rc;
}

int ft_fill_particle_coords(ft_stream s, int32_t in_type, const void *px, const void *py,
                            const void *pz, size_t n, int32_t coord_type, int bits_per_dim,
                            void *x, void *y, void *z, double *box_out) {
    if (in_type != FT_COORD_F32 && in_type != FT_COORD_F64)
        return fail(FT_EINVAL, "snapshot positions must be float or double");
    if (coord_type < FT_COORD_F32 || coord_type > FT_COORD_U64)
        return fail(FT_EINVAL, "unknown coordinate type");
    const bool integer = coord_type == FT_COORD_U32 || coord_type == FT_COORD_U64;
    const int max_bits = coord_type == FT_COORD_U32 ? 32 : 52;
    if (integer && (bits_per_dim < 1 || bits_per_dim > max_bits))
        return fail(FT_EINVAL, "bits_per_dim must be in [1, 32] for uint32 and [1, 52] for "
                               "uint64 coordinates");
    const hipStream_t hs = as_stream(s);
    const double unit_box[2] = {0.0, 1.0};
    if ((n == 0 || !integer) && box_out) {  // the reference's box for these cases
        if (!is_device_accessible(box_out)) {
            box_out[0] = unit_box[0];
            box_out[1] = unit_box[1];
        } else {
            FT_HIP(hipMemcpyAsync(box_out, unit_box, 16, hipMemcpyDefault, hs));
            FT_HIP(hipStreamSynchronize(hs));  // unit_box lives on this stack frame
        }
    }
    if (n == 0) return FT_OK;
    if (!px || !py || !pz || !x || !y || !z) return fail(FT_EINVAL, "null coordinate array");
    const size_t ib = n * (in_type == FT_COORD_F32 ? 4 : 8), ob = n * coord_size(coord_type);
    Staged sx, sy, sz, ox, oy, oz;
    FT_TRY(stage_in(px, ib, hs, sx));
    FT_TRY(stage_in(py, ib, hs, sy));
    FT_TRY(stage_in(pz, ib, hs, sz));
    FT_TRY(stage_out(x, ob, hs, ox));
    FT_TRY(stage_out(y, ob, hs, oy));
    FT_TRY(stage_out(z, ob, hs, oz));
    double *work = nullptr;
    if (hipMallocAsync((void **)&work, snapshot_work_doubles() * 8, hs) != hipSuccess)
        return fail(FT_ENOMEM, "fill_particle_coords: work buffer allocation failed");
    int rc = snapshot_entry(hs, in_type, sx.dev, sy.dev, sz.dev, n, coord_type, bits_per_dim,
                            ox.dev, oy.dev, oz.dev, work);
    if (rc == FT_OK && integer && box_out) {
        if (hipMemcpyAsync(box_out, work, 16, hipMemcpyDefault, hs) != hipSuccess ||
            (!is_device_accessible(box_out) && hipStreamSynchronize(hs) != hipSuccess))
            rc = fail(FT_EHIP, "fill_particle_coords: copying the box failed");
    }
    (void)hipFreeAsync(work, hs);
    if (rc == FT_OK) rc = unstage(ox, hs, true);
    if (rc == FT_OK) rc = unstage(oy, hs, true);
    if (rc == FT_OK) rc = unstage(oz, hs, true);
    return rc;
}

int ft_build_bvh(ft_stream s, ft_tree *tree, ft_scratch *scratch, const void *x, const void *y,
                 const void *z, const uint32_t *id, const int8_t *is_ghost, size_t n,
                 const void *bbox6, const uint64_t *keys) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    if (n == 0) return FT_OK;
    if (n > (size_t)(INT_MAX / 2))
        return fail(FT_ERANGE, "Particle count n exceeds maximum supported tree size (INT_MAX / 2).");
    if (!tree->block || tree->capacity < n)
        return fail(FT_EINVAL, "tree capacity is smaller than the particle count");
    FT_TRY(valid_config(&tree->config, tree->coord_type));
    tree->num_leaves = n;
    tree->num_internal = n - 1;
    const hipStream_t hs = as_stream(s);
    const size_t bytes = n * coord_size(tree->coord_type);
    Staged sx, sy, sz, si, sg, sk;
    FT_TRY(stage_in(x, bytes, hs, sx));
    FT_TRY(stage_in(y, bytes, hs, sy));
    FT_TRY(stage_in(z, bytes, hs, sz));
    FT_TRY(stage_in(id, n * 4, hs, si));
    FT_TRY(stage_in(is_ghost, n, hs, sg));
    FT_TRY(stage_in(keys, n * 8, hs, sk));
    ft_scratch *tmp = nullptr;
    if (!scratch) {
        FT_TRY(ft_scratch_create(&tmp));
        scratch = tmp;
    }
    int rc;
    FT_DISPATCH_COORD(tree->coord_type,
                      rc = build_bvh_t<T>(hs, tree, scratch->s, (const T *)sx.dev,
                                          (const T *)sy.dev, (const T *)sz.dev,
                                          (const uint32_t *)si.dev, (const int8_t *)sg.dev, n,
                                          (const T *)bbox6, (const uint64_t *)sk.dev));
    for (Staged *st : {&sx, &sy, &sz, &si, &sg, &sk}) unstage(*st, hs, false);
    if (tmp) {
        hipStreamSynchronize(hs);
        ft_scratch_destroy(tmp);
    }
    return rc;
}

// ---- queries -------------------------------------------------------------------
// The reference's message for k > _MAX_K_ names the entry point
// (src/hlbvh.hpp:1942, :1970, :1445).
static int check_k(int k, int max_k, const char *entry) {
    if (k < 0) return fail(FT_EINVAL, "k must be non-negative");
    if (k > max_k) {
        char msg[160];
        std::snprintf(msg, sizeof msg,
                      "k exceeds _MAX_K_ template parameter. Instantiate %s with larger _MAX_K_.",
                      entry);
        return fail(FT_EINVAL, msg);
    }
    return FT_OK;
}

static int knn_common(ft_stream s, const ft_tree *tree, const void *qx, const void *qy,
                      const void *qz, const int32_t *leaf_ids, int nq, int k, uint64_t *results,
                      void *dists, int excl, const ft_query_opts *opts) {
    const hipStream_t hs = as_stream(s);
    const size_t cs = coord_size(tree->coord_type);
    const size_t outn = (size_t)nq * (size_t)k;
    Staged sx, sy, sz, sl, sr, sd, sb;
    FT_TRY(stage_in(qx, (size_t)nq * cs, hs, sx));
    FT_TRY(stage_in(qy, (size_t)nq * cs, hs, sy));
    FT_TRY(stage_in(qz, (size_t)nq * cs, hs, sz));
    FT_TRY(stage_in(leaf_ids, (size_t)nq * 4, hs, sl));
    FT_TRY(stage_out(results, outn * 8, hs, sr));
    FT_TRY(stage_out(dists, outn * 8, hs, sd));
    ft_query_opts o{};
    if (opts) o = *opts;
    if (o.seed_bound) {
        FT_TRY(stage_in(o.seed_bound, (size_t)nq * 8, hs, sb));
        o.seed_bound = sb.dev;
    }
    int rc = knn_entry(hs, tree, sx.dev, sy.dev, sz.dev, (const int32_t *)sl.dev, nq, k, excl,
                       (uint64_t *)sr.dev, sd.dev, &o);
    if (rc == FT_OK) rc = unstage(sr, hs, true);
    if (rc == FT_OK) rc = unstage(sd, hs, true);
    for (Staged *st : {&sx, &sy, &sz, &sl, &sb, &sr, &sd}) unstage(*st, hs, false);
    return rc;
}

int ft_self_knn_query(ft_stream s, const ft_tree *tree, int k, uint64_t *results, void *dists,
                      int exclude_self, const ft_query_opts *opts) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    const size_t n = tree->num_leaves;
    if (n == 0) return FT_OK;
    FT_TRY(check_k(k, (opts && opts->max_k > 0) ? opts->max_k : 32, "self_knn_query"));
    if (k == 0) return FT_OK;
    return knn_common(s, tree, nullptr, nullptr, nullptr, nullptr, (int)n, k, results, dists,
                      exclude_self ? 1 : 0, opts);
}

// Same results as ft_self_knn_query, with no allocation on any path: device
// (or pinned host) buffers only, since staging host memory would allocate.
int ft_self_knn_query_noalloc(ft_stream s, const ft_tree *tree, int k, uint64_t *results,
                              void *dists, int exclude_self, const ft_query_opts *opts) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    const size_t n = tree->num_leaves;
    if (n == 0) return FT_OK;
    FT_TRY(check_k(k, (opts && opts->max_k > 0) ? opts->max_k : 32, "self_knn_query"));
    if (k == 0) return FT_OK;
    if (n > noalloc_max_queries())
        return fail(FT_ERANGE, "ft_self_knn_query_noalloc: more than 16777216 queries (its "
                               "fixed scratch size); use ft_self_knn_query");
    if (!results || !dists) return fail(FT_EINVAL, "null results or dists");
    const void *dev_only[4] = {results, dists, opts ? (const void *)opts->seed_bound : nullptr,
                               opts ? (const void *)opts->stats : nullptr};
    for (const void *p : dev_only)
        if (!is_device_accessible(p))
            return fail(FT_EINVAL, "ft_self_knn_query_noalloc: results, dists, seed_bound and "
                                   "stats must be device memory (host buffers need staging)");
    return knn_noalloc_entry(as_stream(s), tree, k, exclude_self ? 1 : 0, results, dists, opts);
}

int ft_self_knn_query_subset(ft_stream s, const ft_tree *tree, const int32_t *leaf_ids,
                             int num_ids, int k, uint64_t *results, void *dists, int exclude_self,
                             const ft_query_opts *opts) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    if (num_ids <= 0) return FT_OK;
    FT_TRY(check_k(k, (opts && opts->max_k > 0) ? opts->max_k : 256, "self_knn_query_subset"));
    if (k == 0 || tree->num_leaves == 0) return FT_OK;
    if (!leaf_ids) return fail(FT_EINVAL, "null leaf_ids");
    return knn_common(s, tree, nullptr, nullptr, nullptr, leaf_ids, num_ids, k, results, dists,
                      exclude_self ? 1 : 0, opts);
}

int ft_knn_query(ft_stream s, const ft_tree *tree, const void *qx, const void *qy,
                 const void *qz, int k, int num_queries, uint64_t *results, void *dists,
                 const ft_query_opts *opts) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    if (tree->num_leaves == 0 || num_queries <= 0) return FT_OK;
    FT_TRY(check_k(k, (opts && opts->max_k > 0) ? opts->max_k : 256, "knn_query"));
    if (k == 0) return FT_OK;
    if (!qx || !qy || !qz) return fail(FT_EINVAL, "null query coordinates");
    return knn_common(s, tree, qx, qy, qz, nullptr, num_queries, k, results, dists, 0, opts);
}

int ft_range_query(ft_stream s, const ft_tree *tree, const void *qx, const void *qy,
                   const void *qz, const void *r_min, const void *r_max, int num_queries,
                   int32_t *results, int32_t *result_counts, int max_results_per_query) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    if (tree->num_leaves == 0 || num_queries <= 0) return FT_OK;
    if (max_results_per_query < 0) return fail(FT_EINVAL, "negative max_results_per_query");
    const hipStream_t hs = as_stream(s);
    const size_t qb = (size_t)num_queries * coord_size(tree->coord_type);
    Staged sx, sy, sz, sa, sb, sr, sc;
    FT_TRY(stage_in(qx, qb, hs, sx));
    FT_TRY(stage_in(qy, qb, hs, sy));
    FT_TRY(stage_in(qz, qb, hs, sz));
    FT_TRY(stage_in(r_min, qb, hs, sa));
    FT_TRY(stage_in(r_max, qb, hs, sb));
    FT_TRY(stage_out(results, (size_t)num_queries * max_results_per_query * 4, hs, sr));
    FT_TRY(stage_out(result_counts, (size_t)num_queries * 4, hs, sc));
    int rc = range_entry(hs, tree, sx.dev, sy.dev, sz.dev, sa.dev, sb.dev, num_queries,
                         (int32_t *)sr.dev, (int32_t *)sc.dev, max_results_per_query, nullptr,
                         nullptr, 0);
    if (rc == FT_OK) rc = unstage(sr, hs, true);
    if (rc == FT_OK) rc = unstage(sc, hs, true);
    for (Staged *st : {&sx, &sy, &sz, &sa, &sb, &sr, &sc}) unstage(*st, hs, false);
    return rc;
}

int ft_range_query_csr(ft_stream s, const ft_tree *tree, const void *qx, const void *qy,
                       const void *qz, const void *r_min, const void *r_max, int num_queries,
                       int64_t *row_offsets, int32_t *cols, int64_t col_capacity) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    if (num_queries <= 0) return FT_OK;
    if (!row_offsets) return fail(FT_EINVAL, "null row_offsets");
    const hipStream_t hs = as_stream(s);
    if (tree->num_leaves == 0) {
        std::vector<int64_t> z((size_t)num_queries + 1, 0);
        Staged so;
        FT_TRY(stage_out(row_offsets, z.size() * 8, hs, so));
        FT_HIP(hipMemcpyAsync(so.dev, z.data(), z.size() * 8, hipMemcpyHostToDevice, hs));
        FT_HIP(hipStreamSynchronize(hs));
        return unstage(so, hs, true);
    }
    const size_t qb = (size_t)num_queries * coord_size(tree->coord_type);
    Staged sx, sy, sz, sa, sb, so, sc;
    FT_TRY(stage_in(qx, qb, hs, sx));
    FT_TRY(stage_in(qy, qb, hs, sy));
    FT_TRY(stage_in(qz, qb, hs, sz));
    FT_TRY(stage_in(r_min, qb, hs, sa));
    FT_TRY(stage_in(r_max, qb, hs, sb));
    FT_TRY(stage_out(row_offsets, ((size_t)num_queries + 1) * 8, hs, so));
    int rc = FT_OK;
    if (cols) {
        // counts, offsets and columns from one walk; the column buffer is
        // staged once the total is known
        struct ColsCtx {
            int32_t *user;
            Staged *sc;
            hipStream_t hs;
        } cctx{cols, &sc, hs};
        auto cols_for = [](int64_t total, void *p) -> int32_t * {
            ColsCtx *c = (ColsCtx *)p;
            if (stage_out(c->user, (size_t)(total > 0 ? total : 1) * 4, c->hs, *c->sc) != FT_OK)
                return nullptr;
            return (int32_t *)c->sc->dev;
        };
        int64_t total = 0;
        rc = range_csr_entry(hs, tree, sx.dev, sy.dev, sz.dev, sa.dev, sb.dev, num_queries,
                             (int64_t *)so.dev, cols_for, &cctx, col_capacity, &total);
        if (rc == FT_OK) rc = unstage(sc, hs, true);
    } else {
        int32_t *counts = nullptr;
        FT_HIP(hipMallocAsync((void **)&counts, (size_t)num_queries * 4, hs));
        rc = range_entry(hs, tree, sx.dev, sy.dev, sz.dev, sa.dev, sb.dev, num_queries, nullptr,
                         counts, 0, nullptr, nullptr, 0);
        if (rc == FT_OK) rc = scan_entry(hs, counts, num_queries, (int64_t *)so.dev);
        hipFreeAsync(counts, hs);
    }
    int rc2 = unstage(so, hs, true);
    if (rc == FT_OK) rc = rc2;
    for (Staged *st : {&sx, &sy, &sz, &sa, &sb, &sc}) unstage(*st, hs, false);
    return rc;
}

}  // extern "C"
