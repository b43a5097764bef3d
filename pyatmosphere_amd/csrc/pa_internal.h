// Internal declarations shared by the HIP translation units of libpa_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pa_hip.h"

namespace pa {

constexpr int kMaxPupils = PA_MAX_PUPILS;     // pupils evaluated per channel in one pass
constexpr int kMaxSums = 6 + kMaxPupils;      // partial sums per tile
constexpr int kChParams = 4;                  // per channel: delta, shift_x, shift_y, spare
constexpr int kMoments = PA_NUM_MOMENTS;      // finalized values per element (see pa_hip.h)
constexpr int kMaxDevices = 64;               // devices whose static scratch address is kept

// Pupil membership (pupils.py:8-10): (x - sx)^2 + (y + sy)^2 <= a^2 with
// dx = x - sx, dy = y + sy, every operation rounded on its own as in the
// float64 numpy expression (no fused multiply-add).
__host__ __device__ inline bool in_circle(double dx, double dy, double a2) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy <= a2;
}

// Grid coordinate minus a shift, (c - n/2) delta - s, with the product and the
// difference each rounded (numpy's x - shift on the float64 grid).  The row
// offset y + sy of the pupil mask is grid_offset(r, n, delta, -sy).
__host__ __device__ inline double grid_offset(int c, int n, double delta, double s) {
#pragma clang fp contract(off)
  return (double)(c - n / 2) * delta - s;
}

// The columns c in [lo, hi] of a grid row that lie inside a pupil, where
// x_c = (c - n/2) delta and dy = grid_offset(r, n, delta, -sy) is the row's offset.  in_circle is
// unimodal in c (x_c, the subtraction, the square and the sum are each
// monotone), so the set is one interval: its ends are estimated with a square
// root and then moved until they satisfy the exact predicate.  lo > hi: empty.
__host__ __device__ inline void circle_cols(int n, double delta, double dy, double sx, double a2, int& lo, int& hi) {
#pragma clang fp contract(off)
  lo = 0;
  hi = -1;
  // the sign of the rounded a^2 - (dy^2 rounded) is exact: negative means the
  // predicate fails at dx = 0 and so everywhere
  const double rem = a2 - dy * dy;
  if (!(rem >= 0.0)) return;
  auto inside = [&](int c) { return in_circle(grid_offset(c, n, delta, sx), dy, a2); };
  // the predicate is smallest at the column nearest x = sx (or a neighbour)
  const int c0 = (int)fmin(fmax(rint(sx / delta) + n / 2, 0.0), n - 1.0);
  int cm = -1;
  for (int d = -1; d <= 1 && cm < 0; ++d)
    if (c0 + d >= 0 && c0 + d < n && inside(c0 + d)) cm = c0 + d;
  if (cm < 0) return;
  const double h = sqrt(rem);
  int l = (int)fmin(fmax(ceil((sx - h) / delta) + n / 2, 0.0), (double)cm);
  int r = (int)fmax(fmin(floor((sx + h) / delta) + n / 2, n - 1.0), (double)cm);
  while (l > 0 && inside(l - 1)) --l;
  while (!inside(l)) ++l;          // stops at cm at the latest
  while (r < n - 1 && inside(r + 1)) ++r;
  while (!inside(r)) --r;
  lo = l;
  hi = r;
}

// Bessel functions J_m(z), m < M <= 32, |z| <= 14, into out[0 .. M):
// Miller's backward recurrence j_{m-1} = (2m / z) j_m - j_{m+1} from
// j_49 = 0, j_48 = 1e-300, normalised by J_0 + 2 sum_{k>=1} J_{2k} = 1, with
// J_m(-z) = (-1)^m J_m(z).  J_48(z) is below 4e-21 of the largest value kept
// for |z| <= 14, and the recurrence grows by at most 48! (2/|z|)^48 down to
// j_0, which stays in range for |z| > 1e-11; below that J_0 = 1 - z^2/4,
// J_1 = z/2, J_2 = z^2/8 hold to 1e-34 and higher orders vanish.
template <int M>
__host__ __device__ inline void bessel_j(double z, double* __restrict__ out) {
  static_assert(M <= 32, "recurrence start chosen for M <= 32, |z| <= 14");
  const double az = fabs(z);
  if (az <= 1e-11) {
    out[0] = 1.0 - 0.25 * z * z;
    if (M > 1) out[1] = 0.5 * z;
    if (M > 2) out[2] = 0.125 * z * z;
    for (int m = 3; m < M; ++m) out[m] = 0.0;
    return;
  }
  const double h = 2.0 / az;
  double v[M];                              // unnormalised j_m, m < M (registers: the loop is unrolled)
  double jp = 0.0, jc = 1e-300, se = 0.0;   // j_{m+1}, j_m, sum of j_{2k}, k >= 1
#pragma unroll
  for (int m = 48; m >= 1; --m) {
    const double jn = m * h * jc - jp;      // j_{m-1}
    jp = jc;
    jc = jn;
    const int i = m - 1;
    if (i >= 2 && !(i & 1)) se += jc;
    if (i < M) v[i] = jc;
  }
  const double inv = 1.0 / (jc + 2.0 * se);
#pragma unroll
  for (int m = 0; m < M; ++m) out[m] = v[m] * ((z < 0.0 && (m & 1)) ? -inv : inv);
}

int set_error(int code, const char* msg);
int check_launch(const char* what);

// step passes (pa_step.hip)
int row_first(int n, float2* work, const float2* in, const float2* src, const int* chan,
              const float2* tw, int batch, hipStream_t s);
int col_step(int n, float2* work, const float2* htab, const int* chan, const int* nsteps, int step,
             int max_steps, const float2* tw, int cols, int batch, hipStream_t s);
int col_groups(int n, int* out);
int row_mid(int n, float2* work, const float* ang, const int* nsteps, int step, const float2* tw, int batch,
            hipStream_t s);
int source_fresnel(int n, float2* sep, const float2* src, const float2* htab, const int* chan, int max_steps,
                   const float2* tw, int batch, hipStream_t s);
int row_mid_source(int n, float2* work, const float2* sep, const float* ang, const float2* tw, int batch,
                   hipStream_t s);
int expmi_eval(int count, const float* x, float2* out, hipStream_t s);
int row_last(int n, const float2* work, float2* out, double* part, const double* chparam,
             const double* pup, int npup, const double* pup_moments, const int* chan, const float2* tw,
             int batch, hipStream_t s);
int tile_rows(int n);
int fft_screen(int n, const float2* coeff, float2* work, const double2* sub, int nsub, const double2* mean,
               const float2* tw, float* phi, float* phi_im, float* ang, int batch, hipStream_t s);

}  // namespace pa
