// Host check of the float32 pre-filter of the integer periodic metric: for
// every tested (query, leaf, bound) with exact squared distance d <= w the
// filter passes the leaf (no false rejection), including bounds equal to d,
// gaps at the wrap (+-2^(B-1)) and B up to 31 where gaps exceed float's 24-bit
// mantissa.  Also counts how often it passes a leaf with d > w (must be rare).
#include "../../fasttree_amd/csrc/ft_traverse.hpp"
#include <cstdio>
#include <random>
using namespace ft;
int main() {
  std::mt19937_64 g(11);
  long bad = 0, extra = 0, tests = 0;
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
      for (int v = 0; v < 4; ++v) {
        uint64_t w = d;                   // the bound equal to d ...
        if (v == 1) w = d ? d - 1 : 0;    // ... just below ...
        if (v == 2) w = d + 1;            // ... just above ...
        if (v == 3) w = d + (d >> 10);    // ... well above
        const bool pass = wrap_filter_near(q[0], q[1], q[2], p[0], p[1], p[2], sh,
                                           wrap_filter_bound(w));
        ++tests;
        if (d <= w && !pass) { if (bad < 5) printf("rejected B=%d d=%llu w=%llu\n", B, (unsigned long long)d, (unsigned long long)w); ++bad; }
        if (d > w && pass) ++extra;
      }
    }
  }
  printf("tests %ld, false rejections: %ld, passes beyond the bound: %ld\n", tests, bad, extra);
  return bad != 0;
}
