// A reference-shaped C++ caller of the drop-in header (include/fasttree/hlbvh.hpp),
// after the dataset runner of test/test_self_knn.cpp:567-682: build_bvh on a
// TreeSoA, self_knn_query and self_knn_query_subset (the reference's subset:
// every n/17-th leaf plus 0, n/2 and n-1) for k = 16 and 32, exclude_self on
// and off, results in device memory copied back; then knn_query over the
// subset's points.  The configuration comes from the -D flags (the reference's
// macros).  It is built from source by fasttree_amd/build.py and run by
// tests/test_gpu_header.py, which checks every output against the oracle.
//
// usage: selfknn_runner <points.bin> <out_dir>
//   points.bin: uint64 n, then x[n], y[n], z[n] as coord_t
//   writes <out_dir>/<name>_idx.bin (uint64) and <name>_d2.bin (dist_t) per run
#include <fasttree/hlbvh.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace fasttree;

static void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

template <typename V> static void write_vec(const std::string &path, const std::vector<V> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(V), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

// Runs one query into device buffers of nq * k and writes both outputs.
template <typename F>
static void run(hipStream_t q, const std::string &out, const std::string &name, std::size_t nq,
                int k, F &&query) {
    std::size_t *d_res = nullptr;
    dist_t *d_dists = nullptr;
    hip_check(hipMalloc((void **)&d_res, nq * k * sizeof(std::size_t)), "hipMalloc results");
    hip_check(hipMalloc((void **)&d_dists, nq * k * sizeof(dist_t)), "hipMalloc dists");
    query(d_res, d_dists);
    hip_check(hipStreamSynchronize(q), "query");
    std::vector<std::size_t> res(nq * k);
    std::vector<dist_t> dists(nq * k);
    hip_check(hipMemcpy(res.data(), d_res, res.size() * sizeof(std::size_t), hipMemcpyDeviceToHost),
              "copy results");
    hip_check(hipMemcpy(dists.data(), d_dists, dists.size() * sizeof(dist_t), hipMemcpyDeviceToHost),
              "copy dists");
    write_vec(out + "/" + name + "_idx.bin", res);
    write_vec(out + "/" + name + "_d2.bin", dists);
    hip_check(hipFree(d_res), "hipFree");
    hip_check(hipFree(d_dists), "hipFree");
}

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <points.bin> <out_dir>\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    std::uint64_t n = 0;
    if (!f || std::fread(&n, 8, 1, f) != 1) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    particles<coord_t> p;
    for (std::vector<coord_t> *a : {&p.pos_x, &p.pos_y, &p.pos_z}) {
        a->resize(n);
        if (std::fread(a->data(), sizeof(coord_t), n, f) != n) {
            std::fprintf(stderr, "short read\n");
            return 2;
        }
    }
    std::fclose(f);
    const std::string out = argv[2];

    hipStream_t q = nullptr;
    hip_check(hipStreamCreate(&q), "hipStreamCreate");
    TreeSoA tree(q, n);
    build_bvh(q, p, tree);
    hip_check(hipStreamSynchronize(q), "build_bvh");

    std::vector<int> subset;
    for (std::size_t i = 0; i < n; i += std::max<std::size_t>(1, n / 17)) subset.push_back((int)i);
    subset.push_back(0);
    subset.push_back((int)(n - 1));
    subset.push_back((int)(n / 2));
    std::sort(subset.begin(), subset.end());
    subset.erase(std::unique(subset.begin(), subset.end()), subset.end());
    write_vec(out + "/subset_ids.bin", subset);
    int *d_ids = nullptr;
    hip_check(hipMalloc((void **)&d_ids, subset.size() * sizeof(int)), "hipMalloc ids");
    hip_check(hipMemcpy(d_ids, subset.data(), subset.size() * sizeof(int), hipMemcpyHostToDevice),
              "copy ids");

    // the subset's points as external queries (host arrays: knn_query stages them)
    std::vector<coord_t> qx, qy, qz;
    for (int r : subset) {
        qx.push_back(p.pos_x[r]);
        qy.push_back(p.pos_y[r]);
        qz.push_back(p.pos_z[r]);
    }

    int failures = 0;
    for (int k : {16, 32}) {
        for (bool excl : {true, false}) {
            const std::string tag = "k" + std::to_string(k) + (excl ? "_excl" : "_incl");
            run(q, out, "self_" + tag, n, k, [&](std::size_t *r, dist_t *d) {
                self_knn_query(q, tree, k, r, d, excl);
            });
            run(q, out, "subset_" + tag, subset.size(), k, [&](std::size_t *r, dist_t *d) {
                self_knn_query_subset(q, tree, d_ids, (int)subset.size(), k, r, d, excl);
            });
        }
        run(q, out, "ext_k" + std::to_string(k), subset.size(), k, [&](std::size_t *r, dist_t *d) {
            knn_query(q, tree, qx.data(), qy.data(), qz.data(), k, (int)subset.size(), r, d);
        });
    }
    // k above the default template cap throws the reference's exception
    try {
        run(q, out, "over_cap", 1, 33, [&](std::size_t *r, dist_t *d) {
            self_knn_query(q, tree, 33, r, d, true);
        });
        std::printf("no exception for k > _MAX_K_\n");
        ++failures;
    } catch (const std::invalid_argument &e) {
        std::printf("invalid_argument: %s\n", e.what());
    }
    hip_check(hipFree(d_ids), "hipFree");
    tree.free(q);
    hip_check(hipStreamDestroy(q), "hipStreamDestroy");
    std::printf("n=%llu subset=%zu failures=%d\n", (unsigned long long)n, subset.size(), failures);
    return failures ? 1 : 0;
}
