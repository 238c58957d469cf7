// Host check of the float32 pre-filters of the integer periodic metric.
//  * wrap_filter_near, and wrap_sq_pre on pre-shifted coordinates against
//    wrap_filter_bound_pre: for every tested (query, leaf, bound) with exact
//    squared distance d <= w the leaf passes (no false rejection), including
//    bounds equal to d, gaps at the wrap (+-2^(B-1)) and B up to 31 where gaps
//    exceed float's 24-bit mantissa; passes with d > w are counted (rare).
//  * wrap_gap_pre on pre-shifted coordinates equals wrap_abs32 exactly.
//  * wrap_box_gap_pre: the float box distance never exceeds the scaled bound
//    when box_d2's exact distance is within it, for 64-bit coordinates with
//    and without bits above B (the inside test reads the full values).
#include "../../fasttree_amd/csrc/ft_traverse.hpp"
#include <cstdio>
#include <random>
using namespace ft;
int main() {
  std::mt19937_64 g(11);
  long bad = 0, extra = 0, tests = 0, gap_bad = 0, box_bad = 0, box_tests = 0;
  for (int B : {5, 10, 21, 24, 25, 31}) {
    const int sh = 32 - B;
    Metric m{B, 1, 0, 0, 0.0};
    const uint32_t mask = (uint32_t)((1ull << B) - 1);
    for (int it = 0; it < 300000; ++it) {
      uint32_t q[3], p[3];
      for (int a = 0; a < 3; ++a) {
        q[a] = (uint32_t)g() & mask;
        const int mode = (int)(g() % 4);
        if (mode == 0) p[a] = (uint32_t)g() & mask;                          // anywhere
        else if (mode == 1) p[a] = (q[a] + (uint32_t)(g() % 64) - 32) & mask; // close
        else if (mode == 2) p[a] = (q[a] + (1u << (B - 1))) & mask;          // at the wrap
        else p[a] = (q[a] + (uint32_t)(g() % (1u << (B > 12 ? B - 8 : 4)))) & mask;
      }
      const uint64_t d = leaf_d2<uint32_t>(m, q[0], q[1], q[2], p[0], p[1], p[2]);
      for (int a = 0; a < 3; ++a)
        if (wrap_gap_pre(q[a] << sh, p[a] << sh, sh) != wrap_abs32(q[a], p[a], sh)) ++gap_bad;
      const float dpre = wrap_sq_pre(q[0] << sh, q[1] << sh, q[2] << sh, p[0] << sh, p[1] << sh,
                                     p[2] << sh);
      for (int v = 0; v < 4; ++v) {
        uint64_t w = d;                   // the bound equal to d ...
        if (v == 1) w = d ? d - 1 : 0;    // ... just below ...
        if (v == 2) w = d + 1;            // ... just above ...
        if (v == 3) w = d + (d >> 10);    // ... well above
        const bool pass = wrap_filter_near(q[0], q[1], q[2], p[0], p[1], p[2], sh,
                                           wrap_filter_bound(w));
        const bool pass_pre = dpre <= wrap_filter_bound_pre(w, sh);
        tests += 2;
        if (d <= w && !pass) { if (bad < 5) printf("rejected B=%d d=%llu w=%llu\n", B, (unsigned long long)d, (unsigned long long)w); ++bad; }
        if (d <= w && !pass_pre) { if (bad < 5) printf("pre rejected B=%d d=%llu w=%llu\n", B, (unsigned long long)d, (unsigned long long)w); ++bad; }
        if (d > w && pass) ++extra;
        if (d > w && pass_pre) ++extra;
      }
      // boxes: corners around the leaf, 64-bit coordinates, sometimes with high bits set
      if (it % 4 == 0) {
        uint64_t lo[3], hi[3], qq[3];
        for (int a = 0; a < 3; ++a) {
          const uint64_t ext = g() % (1ull << (B > 14 ? B - 6 : B - 1));
          lo[a] = (p[a] >= ext / 2) ? p[a] - ext / 2 : 0;
          hi[a] = lo[a] + ext;
          if (hi[a] > mask) hi[a] = mask;
          const uint64_t high = (g() % 8 == 0) ? ((g() % 4 + 1) << 40) : 0;  // unmasked input
          lo[a] += high; hi[a] += high;
          qq[a] = q[a] + ((g() % 8 == 0) ? ((g() % 4 + 1) << 40) : 0);
        }
        const uint64_t db = box_d2<uint64_t>(m, qq[0], qq[1], qq[2], lo[0], hi[0], lo[1], hi[1],
                                             lo[2], hi[2]);
        float f = 0.f;
        for (int a = 0; a < 3; ++a) {
          const float ga = wrap_box_gap_pre<uint64_t>(qq[a] & mask, (uint32_t)qq[a] << sh, lo[a],
                                                      hi[a], sh);
          f = a == 0 ? ga * ga : __builtin_fmaf(ga, ga, f);
        }
        for (uint64_t w : {db, db + 1, db + (db >> 8)}) {
          ++box_tests;
          if (!(f <= wrap_filter_bound_pre(w, sh))) {
            if (box_bad < 5) printf("box rejected B=%d d=%llu w=%llu\n", B, (unsigned long long)db, (unsigned long long)w);
            ++box_bad;
          }
        }
      }
    }
  }
  printf("tests %ld, false rejections: %ld, passes beyond the bound: %ld, gap mismatches: %ld, "
         "box tests %ld, box false rejections: %ld\n", tests, bad, extra, gap_bad, box_tests, box_bad);
  return (bad != 0 || gap_bad != 0 || box_bad != 0);
}
