// Compile-and-link check of the C++ drop-in header (include/fasttree/hlbvh.hpp)
// in the configuration named by the -D flags on the command line.  Without a
// GPU it only exercises host logic: error mapping and empty-input paths.
#include <fasttree/hlbvh.hpp>

#include <cstdio>
#include <stdexcept>

using namespace fasttree;

int main() {
    int failures = 0;
    // k above the template cap is rejected before any device work, with the
    // reference's exception type (an empty tree returns first, as there).
    try {
        ft_tree t{};
        t.num_leaves = 5;
        ft_query_opts o = default_opts(32);
        detail::check(ft_self_knn_query(nullptr, &t, 33, nullptr, nullptr, 1, &o));
        std::printf("no exception for k > _MAX_K_\n");
        ++failures;
    } catch (const std::invalid_argument &e) {
        std::printf("invalid_argument: %s\n", e.what());
    } catch (const std::exception &e) {
        std::printf("wrong exception type: %s\n", e.what());
        ++failures;
    }
    // Empty input returns silently, as in the reference.
    particles<coord_t> p;
    BoundingBox<coord_t> b = compute_bbox<coord_t>(nullptr, nullptr, nullptr, nullptr, 0);
    (void)b;
    // The templates instantiate with the reference's defaults.
    void (*sk)(hipStream_t, const TreeSoA &, int, std::size_t *, dist_t *, bool) =
        &self_knn_query<32>;
    void (*ss)(hipStream_t, const TreeSoA &, const int *, int, int, std::size_t *, dist_t *,
               bool) = &self_knn_query_subset<256>;
    void (*kq)(hipStream_t, const TreeSoA &, const coord_t *, const coord_t *, const coord_t *,
               int, int, std::size_t *, dist_t *) = &knn_query<256>;
    (void)sk; (void)ss; (void)kq; (void)p;
    // An empty snapshot converts to nothing with the reference's box (0, 1).
    bench_utils::ParticleData empty;
    particles<coord_t> pe;
    double box_min = -1.0, box_size = -1.0;
    bench_utils::fill_particle_coords(empty, pe, box_min, box_size);
    if (box_min != 0.0 || box_size != 1.0 || !pe.pos_x.empty()) {
        std::printf("empty snapshot: box (%g, %g)\n", box_min, box_size);
        ++failures;
    }
    std::printf("failures: %d\n", failures);
    return failures ? 1 : 0;
}
