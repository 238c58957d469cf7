// fasttree/hlbvh.hpp -- C++ drop-in for FastTree's HLBVH build / query API on
// MI355X, over the C ABI in fasttree_c.h (link libfasttree_amd.so).
//
// The names, argument order and exceptions follow src/hlbvh.hpp of the
// reference; a hipStream_t replaces the in-order sycl::queue.  The same
// compile-time switches select the types and the curve:
//   FASTTREE_INTEGER_COORDS + POSITIONS_IN_32BIT / POSITIONS_IN_64BIT,
//   COORDS_TYPE_DOUBLE (else float), SFC_TYPE_MORTON / SFC_TYPE_PEANO_HILBERT,
//   BITS_PER_DIMENSION (default 21), PERIODIC_BC, RETURN_ORIG_INDICES,
//   FASTTREE_BOX_LENGTH (FP periodic period; no reference counterpart).
// Tree arrays live in device memory; host pointers passed to build_bvh,
// knn_query and range_query are staged as the reference stages them.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../fasttree_c.h"

namespace fasttree {

#if defined(FASTTREE_INTEGER_COORDS)
#if defined(POSITIONS_IN_64BIT)
using coord_t = std::uint64_t;
#else
using coord_t = std::uint32_t;
#endif
using dist_t = std::uint64_t;
#elif defined(COORDS_TYPE_DOUBLE)
using coord_t = double;
using dist_t = double;
#else
using coord_t = float;
using dist_t = double;
#endif

#ifndef BITS_PER_DIMENSION
#define FASTTREE_BITS 21
#else
#define FASTTREE_BITS BITS_PER_DIMENSION
#endif
#ifndef FASTTREE_BOX_LENGTH
#define FASTTREE_BOX_LENGTH 0.0
#endif

namespace detail {
template <typename T> constexpr int coord_code();
template <> constexpr int coord_code<float>() { return FT_COORD_F32; }
template <> constexpr int coord_code<double>() { return FT_COORD_F64; }
template <> constexpr int coord_code<std::uint32_t>() { return FT_COORD_U32; }
template <> constexpr int coord_code<std::uint64_t>() { return FT_COORD_U64; }

inline ft_config config() {
    ft_config c{};
#if defined(SFC_TYPE_MORTON)
    c.sfc = FT_SFC_MORTON;
#else
    c.sfc = FT_SFC_PEANO_HILBERT;
#endif
    c.bits_per_dim = FASTTREE_BITS;
#if defined(PERIODIC_BC)
    c.periodic = 1;
#endif
#if defined(RETURN_ORIG_INDICES)
    c.return_orig_indices = 1;
#endif
    c.box_length = FASTTREE_BOX_LENGTH;
    return c;
}

// Status codes back to the reference's exception types.
inline void check(int rc) {
    if (rc == FT_OK) return;
    const std::string msg = ft_last_error();
    if (rc == FT_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
}  // namespace detail

// particles<T> (src/hlbvh.hpp:26-34).
template <typename T> struct particles {
    std::vector<T> pos_x, pos_y, pos_z;
    std::vector<std::uint32_t> id;
    std::vector<std::int8_t> is_ghost;
};

// BoundingBox (src/hlbvh.hpp:78-110); empty box = (max, lowest) identities.
template <typename FloatT> struct BoundingBox {
    FloatT min_x, max_x, min_y, max_y, min_z, max_z;
    BoundingBox(FloatT a, FloatT b, FloatT c, FloatT d, FloatT e, FloatT f)
        : min_x(a), max_x(b), min_y(c), max_y(d), min_z(e), max_z(f) {}
    BoundingBox()
        : min_x(std::numeric_limits<FloatT>::max()),
          max_x(std::numeric_limits<FloatT>::is_integer ? FloatT(0) : -std::numeric_limits<FloatT>::max()),
          min_y(min_x), max_y(max_x), min_z(min_x), max_z(max_x) {}
};

// TreeSoA (src/hlbvh.hpp:609-673): same field names; device memory.
struct TreeSoA {
    coord_t *min_x = nullptr, *max_x = nullptr;
    coord_t *min_y = nullptr, *max_y = nullptr;
    coord_t *min_z = nullptr, *max_z = nullptr;
    int *left_child = nullptr, *right_child = nullptr, *parent = nullptr;
    std::uint32_t *id = nullptr;
    std::int8_t *is_ghost = nullptr;
    int *orig_idx = nullptr;
    std::size_t num_leaves = 0, num_internal = 0;
    ft_tree c{};

    TreeSoA(hipStream_t q, std::size_t n) {
        const ft_config cfg = detail::config();
        detail::check(ft_tree_alloc(&c, n, detail::coord_code<coord_t>(), &cfg, (ft_stream)q));
        sync();
    }
    void free(hipStream_t q) {
        detail::check(ft_tree_free(&c, (ft_stream)q));
        sync();
    }
    // Refreshes the typed views after the library (re)sized the tree.
    void sync() {
        min_x = (coord_t *)c.min_x; max_x = (coord_t *)c.max_x;
        min_y = (coord_t *)c.min_y; max_y = (coord_t *)c.max_y;
        min_z = (coord_t *)c.min_z; max_z = (coord_t *)c.max_z;
        left_child = c.left_child; right_child = c.right_child; parent = c.parent;
        id = c.id; is_ghost = c.is_ghost; orig_idx = c.orig_idx;
        num_leaves = c.num_leaves; num_internal = c.num_internal;
    }
};

// TreePool (src/hlbvh.hpp:682-715).
struct TreePool {
    TreeSoA *tree = nullptr;
    std::size_t capacity = 0;
    TreeSoA &ensure(hipStream_t q, std::size_t n_total) {
        if (!tree) tree = new TreeSoA(q, 0);
        const ft_config cfg = detail::config();
        detail::check(ft_tree_ensure(&tree->c, n_total, detail::coord_code<coord_t>(), &cfg,
                                     (ft_stream)q));
        tree->sync();
        capacity = tree->c.capacity;
        return *tree;
    }
    void free_all(hipStream_t q) {
        if (tree) {
            tree->free(q);
            delete tree;
            tree = nullptr;
        }
        capacity = 0;
    }
};

// BvhScratch (src/hlbvh.hpp:720-745).
struct BvhScratch {
    ft_scratch *s = nullptr;
    BvhScratch() { detail::check(ft_scratch_create(&s)); }
    // Radix sort digit width of the builds: 8 or 9 bits, 0 = fewer passes.
    void set_radix_bits(int bits) { detail::check(ft_scratch_set_radix_bits(s, bits)); }
    void free_all(hipStream_t) {
        ft_scratch_destroy(s);
        s = nullptr;
    }
};

template <typename FloatT>
inline BoundingBox<FloatT> compute_bbox(hipStream_t q, const FloatT *x, const FloatT *y,
                                        const FloatT *z, std::size_t n) {
    FloatT o[6];
    if (n == 0) return BoundingBox<FloatT>();
    detail::check(ft_compute_bbox((ft_stream)q, detail::coord_code<FloatT>(), x, y, z, n, o));
    return BoundingBox<FloatT>(o[0], o[1], o[2], o[3], o[4], o[5]);
}

// Keys are the 64-bit sort keys (the reference's to_sort_key of sfc_key).
template <typename FloatT>
inline void sfc_encode(hipStream_t q, const FloatT *x, const FloatT *y, const FloatT *z,
                       std::size_t n, std::uint64_t *keys, const BoundingBox<FloatT> &bbox) {
    const ft_config cfg = detail::config();
    const FloatT b[6] = {bbox.min_x, bbox.max_x, bbox.min_y, bbox.max_y, bbox.min_z, bbox.max_z};
    detail::check(ft_sfc_encode((ft_stream)q, detail::coord_code<FloatT>(), &cfg, x, y, z, n, keys, b));
}

inline void build_tree(hipStream_t q, TreeSoA &tree, const std::uint64_t *sorted_keys,
                       const coord_t *sx = nullptr, const coord_t *sy = nullptr,
                       const coord_t *sz = nullptr, const std::uint32_t *sid = nullptr,
                       const std::int8_t *sghost = nullptr, const int *sorig = nullptr) {
    detail::check(ft_build_tree((ft_stream)q, &tree.c, sorted_keys, sx, sy, sz, sid, sghost, sorig));
}

inline void build_bvh(hipStream_t q, const particles<coord_t> &p, TreeSoA &tree,
                      BvhScratch &scratch, BoundingBox<coord_t> *bbox = nullptr,
                      std::uint64_t *key = nullptr) {
    const std::size_t n = p.pos_x.size();
    if (n == 0) return;
    coord_t b[6];
    if (bbox) {
        b[0] = bbox->min_x; b[1] = bbox->max_x; b[2] = bbox->min_y;
        b[3] = bbox->max_y; b[4] = bbox->min_z; b[5] = bbox->max_z;
    }
    detail::check(ft_build_bvh((ft_stream)q, &tree.c, scratch.s, p.pos_x.data(), p.pos_y.data(),
                               p.pos_z.data(), p.id.empty() ? nullptr : p.id.data(),
                               p.is_ghost.empty() ? nullptr : p.is_ghost.data(), n,
                               bbox ? b : nullptr, key));
    tree.sync();
}

inline void build_bvh(hipStream_t q, const particles<coord_t> &p, TreeSoA &tree,
                      BoundingBox<coord_t> *bbox = nullptr, std::uint64_t *key = nullptr) {
    BvhScratch scratch;
    build_bvh(q, p, tree, scratch, bbox, key);
    (void)hipStreamSynchronize(q);
    scratch.free_all(q);
}

inline ft_query_opts default_opts(int max_k) {
    ft_query_opts o{};
    o.max_k = max_k;
    return o;
}

template <int _MAX_K_ = 32>
inline void self_knn_query(hipStream_t q, const TreeSoA &tree, int k, std::size_t *results,
                           dist_t *result_dists, bool exclude_self = true) {
    const ft_query_opts o = default_opts(_MAX_K_);
    detail::check(ft_self_knn_query((ft_stream)q, &tree.c, k, (std::uint64_t *)results,
                                    result_dists, exclude_self ? 1 : 0, &o));
}

template <int _MAX_K_ = 256>
inline void self_knn_query_subset(hipStream_t q, const TreeSoA &tree, const int *leaf_ids,
                                  int num_ids, int k, std::size_t *results, dist_t *result_dists,
                                  bool exclude_self = true) {
    const ft_query_opts o = default_opts(_MAX_K_);
    detail::check(ft_self_knn_query_subset((ft_stream)q, &tree.c, leaf_ids, num_ids, k,
                                           (std::uint64_t *)results, result_dists,
                                           exclude_self ? 1 : 0, &o));
}

template <int _MAX_K_ = 256>
inline void knn_query(hipStream_t q, const TreeSoA &tree, const coord_t *qx, const coord_t *qy,
                      const coord_t *qz, int k, int num_queries, std::size_t *results,
                      dist_t *result_dists) {
    const ft_query_opts o = default_opts(_MAX_K_);
    detail::check(ft_knn_query((ft_stream)q, &tree.c, qx, qy, qz, k, num_queries,
                               (std::uint64_t *)results, result_dists, &o));
}

inline void range_query(hipStream_t q, const TreeSoA &tree, const coord_t *qx, const coord_t *qy,
                        const coord_t *qz, const coord_t *r_min, const coord_t *r_max,
                        int num_queries, int *results, int *result_counts,
                        int max_results_per_query) {
    detail::check(ft_range_query((ft_stream)q, &tree.c, qx, qy, qz, r_min, r_max, num_queries,
                                 results, result_counts, max_results_per_query));
}

// The snapshot conversion of the reference's benchmark utilities
// (test/benchmark/BenchmarkUtils.hpp:82-87, 164-205), run on the GPU (null
// stream): ParticleData holds what the HDF5 read of PartType0 gives; the HDF5
// read itself is not part of this library.
namespace bench_utils {
struct ParticleData {
    std::vector<float> pos_x, pos_y, pos_z;
    std::vector<float> vel_x, vel_y, vel_z;
    std::vector<float> masses;
    std::size_t count = 0;
};

template <typename T>
inline void fill_particle_coords(const ParticleData &data, particles<T> &p, double &out_box_min,
                                 double &out_box_size) {
    const std::size_t n = data.count;
    p.pos_x.resize(n);
    p.pos_y.resize(n);
    p.pos_z.resize(n);
    double box[2] = {0.0, 1.0};
    detail::check(ft_fill_particle_coords(nullptr, FT_COORD_F32, data.pos_x.data(),
                                          data.pos_y.data(), data.pos_z.data(), n,
                                          detail::coord_code<T>(), FASTTREE_BITS, p.pos_x.data(),
                                          p.pos_y.data(), p.pos_z.data(), box));
    out_box_min = box[0];
    out_box_size = box[1];
}
}  // namespace bench_utils

static_assert(sizeof(std::size_t) == sizeof(std::uint64_t), "results are 64-bit indices");

}  // namespace fasttree
