// Brute-force check of pa::circle_cols (host build of the device helper):
// for random and boundary-exact pupils, the interval it returns equals the set
// of columns where pa::in_circle holds, row by row.  Prints "ok <rows>" or the
// first mismatch.
#include <cstdio>
#include <cstdint>
#include <cmath>

#include "../pyatmosphere_amd/csrc/pa_internal.h"

static uint64_t state = 0x9E3779B97F4A7C15ull;
static double uni() {  // splitmix64 -> [0, 1)
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (z >> 11) * 0x1.0p-53;
}

static int check(int n, double delta, double a, double sx, double sy, long& rows) {
  const double a2 = a * a;
  for (int r = 0; r < n; ++r) {
    const double dy = pa::grid_offset(r, n, delta, -sy);
    int lo, hi;
    pa::circle_cols(n, delta, dy, sx, a2, lo, hi);
    for (int c = 0; c < n; ++c) {
      const bool in = pa::in_circle(pa::grid_offset(c, n, delta, sx), dy, a2);
      if (in != (c >= lo && c <= hi)) {
        std::printf("mismatch n=%d delta=%.17g a=%.17g sx=%.17g sy=%.17g row=%d col=%d lo=%d hi=%d\n", n, delta,
                    a, sx, sy, r, c, lo, hi);
        return 1;
      }
    }
    ++rows;
  }
  return 0;
}

int main() {
  long rows = 0;
  const int ns[] = {64, 256, 512, 2048};
  for (int t = 0; t < 400; ++t) {
    const int n = ns[t % 4];
    const double delta = 1e-4 * std::pow(10.0, 2.0 * uni());
    const double span = n * delta;
    double a, sx, sy;
    switch (t % 5) {
      case 0:   // random
        a = span * uni() * 0.6; sx = span * (uni() - 0.5); sy = span * (uni() - 0.5); break;
      case 1:   // radius a whole number of cells, centred on a grid point: edge through grid points
        a = delta * (int)(1 + uni() * n / 3); sx = delta * (int)((uni() - 0.5) * n / 2);
        sy = delta * (int)((uni() - 0.5) * n / 2); break;
      case 2:   // centre half-way between grid points
        a = delta * (0.5 + (int)(uni() * n / 3)); sx = delta * (0.5 + (int)((uni() - 0.5) * n / 2));
        sy = delta * (0.5 + (int)((uni() - 0.5) * n / 2)); break;
      case 3:   // tiny pupils (at most a few cells), and the delta/2 pupil
        a = (t % 2) ? delta / 2 : delta * 3 * uni(); sx = delta * (uni() - 0.5) * 4; sy = delta * (uni() - 0.5) * 4; break;
      default:  // partly or wholly off the grid
        a = span * uni(); sx = span * (uni() * 2 - 1); sy = span * (uni() * 2 - 1); break;
    }
    if (check(n, delta, a, sx, sy, rows)) return 1;
  }
  std::printf("ok %ld\n", rows);
  return 0;
}
