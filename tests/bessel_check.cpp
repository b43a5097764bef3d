// Prints pa::bessel_j<32>(z) (host build of the device routine) for the z
// values given on stdin, one line per z: z J_0 .. J_31.
#include <cstdio>
#include <cmath>

#include "../pyatmosphere_amd/csrc/pa_internal.h"

int main() {
  double z;
  while (std::scanf("%lf", &z) == 1) {
    double j[32];
    pa::bessel_j<32>(z, j);
    std::printf("%.17g", z);
    for (int m = 0; m < 32; ++m) std::printf(" %.17g", j[m]);
    std::printf("\n");
  }
  return 0;
}
