// Host check: the short point-distance formula (leaf_d2) returns bit-identical
// values to the box formula with a degenerate box (point_d2 = node_distance_sq,
// src/hlbvh.hpp:998-1055) for every coordinate type and metric, including
// coordinates with bits above B and points on the periodic wrap.
#include "../../fasttree_amd/csrc/ft_internal.hpp"
#include <random>
#include <cstdio>
using namespace ft;
template <typename T> int run(std::mt19937_64& g, int periodic, int bits, bool wide) {
  Metric m{bits, periodic, bits > 31 ? bits - 31 : 0, 0, 100.0};
  int bad = 0;
  for (int it = 0; it < 400000; ++it) {
    T a[3], b[3];
    for (int i = 0; i < 3; ++i) {
      if constexpr (CoordTraits<T>::fp) { a[i] = (T)(std::uniform_real_distribution<double>(0, 100)(g)); b[i] = (T)(std::uniform_real_distribution<double>(0, 100)(g)); if (it % 7 == 0) b[i] = a[i]; }
      else { uint64_t mask = wide ? ~0ull : ((1ull << bits) - 1); a[i] = (T)(g() & mask); b[i] = (T)(g() & mask); if (it % 7 == 0) b[i] = a[i]; if (it % 11 == 0) b[i] = (T)(a[i] + ((uint64_t)1 << bits)); }
    }
    auto r1 = point_d2<T>(m, a[0], a[1], a[2], b[0], b[1], b[2]);
    auto r2 = leaf_d2<T>(m, a[0], a[1], a[2], b[0], b[1], b[2]);
    if (memcmp(&r1, &r2, sizeof(r1)) != 0) { if (bad < 3) printf("mismatch\n"); bad++; }
  }
  return bad;
}
int main() {
  std::mt19937_64 g(7); int bad = 0;
  for (int p = 0; p < 2; ++p) for (int b : {5, 10, 21}) {
    bad += run<uint64_t>(g, p, b, false); bad += run<uint64_t>(g, p, b, true);
    bad += run<uint32_t>(g, p, b, false); bad += run<uint32_t>(g, p, b, true);
    bad += run<double>(g, p, b, false); bad += run<float>(g, p, b, false);
  }
  printf("total mismatches: %d\n", bad); return bad != 0;
}
