// Host check of the float32 pre-filter for double coordinates (fpf_*): for
// random query/leaf pairs in a periodic box and in open space (coordinates of
// magnitude up to M, gaps from 1e-7 to the box scale, pairs at the wrap), the
// filter passes every leaf whose exact double squared distance d is at most
// the bound w (w = d, the next double up, and slightly above), and with a
// floor lo <= d never puts it below the floor; passes beyond the bound are
// counted (the filter's slack).
#include "../../fasttree_amd/csrc/ft_traverse.hpp"
#include <cmath>
#include <cstdio>
#include <random>
using namespace ft;
int main() {
  std::mt19937_64 g(17);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long tests = 0, bad_hi = 0, bad_lo = 0, slack = 0;
  for (int wrap = 0; wrap < 2; ++wrap) {
    for (double L : {1.0, 100.0, 12345.678}) {
      Metric m{21, wrap, 0, 0, L};
      const double M = wrap ? L : 3.0 * L;  // open space: coordinates in [-M, M]
      const double E = fpf_err(M, L, wrap != 0);
      for (int it = 0; it < 400000; ++it) {
        double p[3], q[3];
        for (int a = 0; a < 3; ++a) {
          p[a] = wrap ? U(g) * L : (2.0 * U(g) - 1.0) * M;
          const int mode = (int)(g() % 4);
          double gap = 0.0;
          if (mode == 0) gap = (U(g) - 0.5) * L;                       // anywhere
          else if (mode == 1) gap = (U(g) - 0.5) * L * 1e-4;           // close
          else if (mode == 2) gap = (U(g) - 0.5) * L * 1e-7;           // very close
          else gap = wrap ? (0.5 + (U(g) - 0.5) * 1e-6) * L : gap;   // at the wrap
          q[a] = p[a] + gap;
          if (wrap) q[a] = std::fmod(q[a] + L, L);
          else if (q[a] > M) q[a] = M;
          else if (q[a] < -M) q[a] = -M;
        }
        const double d = leaf_d2<double>(m, p[0], p[1], p[2], q[0], q[1], q[2]);
        const float fd = fpf_sq(fpf_gap((float)p[0], (float)q[0], (float)L, wrap != 0),
                                fpf_gap((float)p[1], (float)q[1], (float)L, wrap != 0),
                                fpf_gap((float)p[2], (float)q[2], (float)L, wrap != 0));
        for (double w : {d, std::nextafter(d, 1e300), d * (1.0 + 1e-9)}) {
          ++tests;
          if (!(fd <= fpf_upper(w, E))) {
            if (bad_hi < 5) printf("rejected: wrap %d L %g d %.17g w %.17g fd %.9g\n", wrap, L, d, w, fd);
            ++bad_hi;
          }
        }
        for (double lo : {d, std::nextafter(d, 0.0), d * (1.0 - 1e-9)}) {
          if (!(fd >= fpf_lower(lo, E))) {
            if (bad_lo < 5) printf("below floor: wrap %d L %g d %.17g lo %.17g fd %.9g\n", wrap, L, d, lo, fd);
            ++bad_lo;
          }
        }
        if (fd <= fpf_upper(d * 0.999, E) && d > 0) ++slack;
      }
    }
  }
  printf("fp filter tests %ld, false rejections: %ld, false floor rejections: %ld, "
         "passes below 0.999 d: %ld\n", tests, bad_hi, bad_lo, slack);
  return (bad_hi != 0 || bad_lo != 0);
}
