// Sparse-spectral phase-screen synthesis (phase_screens.py:93-136; grids.py:98-119).
//
// phi(r,c) = Re sum_k a_k exp(i alpha_k(c)) exp(i beta_k(r)),
//   alpha_k(c) = 2 pi fx_k x_c,  beta_k(r) = 2 pi fy_k y_r.
// With p_k(c) = e^{i alpha} - 1 and q_k(r) = e^{i beta} - 1 the exact identity
//   e^{i alpha} e^{i beta} = 1 + p + q + p q
// splits the screen into
//   phi0      = Re sum a_k                       (scalar, float64)
//   g(c)      = Re sum a_k p_k(c)                (1-D, float64)
//   h(r)      = Re sum a_k q_k(r)                (1-D, float64)
//   phi2(r,c) = Re sum (a_k q_k(r)) p_k(c)       (2-D, GEMM on MFMA)
// The large piston/tilt content of the screen lives in the float64 parts; the
// 2-D remainder is small, so a split-precision GEMM reaches fp32 accuracy.
// The step pass multiplies by e^{-i(phi0 + h)} e^{-i g} e^{-i phi2}.
// p and q come from a float64 sincos of the first angle of a run of rows (or
// columns) and a float64 product recurrence along the run; their float64
// error (~1e-15) is far below the resolution of the split GEMM operands.
//
// The GEMM is real with K = 2P: A[r][2k] = Re(a q), A[r][2k+1] = -Im(a q),
// B[c][2k] = Re p, B[c][2k+1] = Im p, phi2 = A B^T.
#include "pa_fft.h"
#include "pa_internal.h"

namespace pa {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef float pa_f4v __attribute__((ext_vector_type(4)));
typedef float pa_f2v __attribute__((ext_vector_type(2)));
typedef _Float16 pa_h2v __attribute__((ext_vector_type(2)));

constexpr double kTwoPi = 6.283185307179586476925;

static inline int kpad(int points) { return ((2 * points + 31) / 32) * 32; }

// polynomial block of the screen GEMM (k_screen_lowrank_part, k_screen_lowrank_sum)
constexpr int kLR = PA_LOWRANK_TERMS;
constexpr double kLowRankZ = 14.0;

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11: 10 rounds, multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    if (i) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
           (uint32_t)p0};
  }
  return c;
}

// Word w of the stream of (realization rid, screen): counter (rid_lo, rid_hi, screen, w/4).
__device__ __forceinline__ uint32_t stream_word(uint64_t rid, int screen, uint32_t w, uint64_t seed) {
  const U4 o = philox10(U4{(uint32_t)rid, (uint32_t)(rid >> 32), (uint32_t)screen, w >> 2},
                        (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint32_t s = w & 3;
  return s == 0 ? o.x : s == 1 ? o.y : s == 2 ? o.z : o.w;
}

__device__ __forceinline__ double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386962890625e-10; }

// Stream words: 0 = radial jitter, 1 + k = angle k, 1 + P + 2k, 2 + P + 2k =
// Box-Muller pair of amplitude k.  Point k of element (channel ch, realization
// id); U = the jitter, shared by every point.
__device__ __forceinline__ void draw_point(int P, const float* __restrict__ base,
                                           const float* __restrict__ sqrt_psd, int ch, uint64_t id,
                                           uint64_t seed, int screen, int k, float U, float& rho, float& theta,
                                           float2& value) {
#pragma clang fp contract(off)
  const float f = base[(size_t)ch * P + k];
  const float fp = k ? base[(size_t)ch * P + k - 1] : 0.0f;
  const double f2 = (double)f * (double)f, fp2 = (double)fp * (double)fp;
  rho = (float)sqrt(fp2 + (double)U * (f2 - fp2));
  theta = (float)(kTwoPi * u01(stream_word(id, screen, 1 + k, seed)));
  const double u1 = u01(stream_word(id, screen, 1 + P + 2 * k, seed));
  const double u2 = u01(stream_word(id, screen, 2 + P + 2 * k, seed));
  const double R = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincos(kTwoPi * u2, &sn, &cs);
  const float sq = sqrt_psd[(size_t)ch * P + k];
  value = make_float2((float)(R * cs) * sq, (float)(R * sn) * sq);
}

__global__ void k_spectral_draw(int P, const float* __restrict__ base, const float* __restrict__ sqrt_psd,
                                const int* __restrict__ chan, const uint64_t* __restrict__ rid,
                                uint64_t seed, int screen, float* __restrict__ rho,
                                float* __restrict__ theta, float2* __restrict__ value) {
  const int b = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  const int ch = chan ? chan[b] : 0;
  const uint64_t id = rid[b];
  const float U = (float)u01(stream_word(id, screen, 0, seed));
  float r, t;
  float2 v;
  draw_point(P, base, sqrt_psd, ch, id, seed, screen, k, U, r, t, v);
  rho[(size_t)b * P + k] = r;
  theta[(size_t)b * P + k] = t;
  value[(size_t)b * P + k] = v;
}

// ------------------------------------------------------------------ prep
// Per point, in float64: (fx, fy) = rho (cos theta, sin theta) and the
// amplitude a' = a e^{2 pi i (fx x0 + fy y0)} that moves the screen shift
// (x0, y0) (phase_screens.py:112-113) into the spectrum, so that every later
// kernel works on the unshifted grid x = (c - n/2) delta and the split
// phi = phi0 + g + h + phi2 is taken about the grid centre whatever the
// shift: fxy[b][k] = (fx, fy, Re a', Im a').  phi0 = Re sum a'; operand scales.
// Returns |a'|.
__device__ __forceinline__ double prep_point(double rr, double th, float2 a, double x0, double y0, double2& f,
                                             double2& amp) {
#pragma clang fp contract(off)
  double sn, cs;
  sincos(th, &sn, &cs);
  const double fx = rr * cs, fy = rr * sn;
  double ar = a.x, ai = a.y;
  if (x0 != 0.0 || y0 != 0.0) {
    double ss, sc;
    sincos(kTwoPi * (fx * x0 + fy * y0), &ss, &sc);
    const double t = ar * sc - ai * ss;
    ai = ar * ss + ai * sc;
    ar = t;
  }
  f = make_double2(fx, fy);
  amp = make_double2(ar, ai);
  return sqrt(ar * ar + ai * ai);
}

// Block reduction of the 256 per-thread (sum, max) pairs in s_sum / s_max
// (tree, fixed order) and the element's sparam entries.
__device__ __forceinline__ void prep_finish(double* __restrict__ s_sum, double* __restrict__ s_max, double* sp,
                                            int mode) {
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double sa = 1.0, sb = 1.0;
    if (mode != PA_GEMM_F32) {
      // A entries: |a q| <= 2 max|a| scaled to <= 2^14 (lowered by
      // k_screen_lowrank_sum if the polynomial block needs it); B entries:
      // |p| <= 2 and |T_m| <= 1, scaled to <= 2^14
      const double m = 2.0 * s_max[0];
      sa = m > 0 ? ldexp(1.0, 14 - (int)ceil(log2(m))) : 1.0;
      sb = ldexp(1.0, 13);
    }
    sp[0] = s_sum[0];
    sp[1] = sa;
    sp[2] = sb;
    sp[3] = 1.0 / (sa * sb);
    sp[4] = 0.0;   // no polynomial block until pa_screen_tables finds one
    sp[5] = 0.0;
    sp[6] = sa;    // the direct points' A scale
  }
}

__global__ void __launch_bounds__(256)
k_screen_prep(int P, const float* __restrict__ rho, const float* __restrict__ theta,
              const float2* __restrict__ value, double2* __restrict__ fxy,
              double* __restrict__ sparam, int mode, const double* __restrict__ chparam,
              const int* __restrict__ chan, const int* __restrict__ nsteps, int step) {
  const int b = blockIdx.x;
  if (nsteps && step >= nsteps[b]) return;
  const int ch = chan ? chan[b] : 0;
  const double x0 = chparam[ch * kChParams + 1], y0 = chparam[ch * kChParams + 2];
  __shared__ double s_sum[256], s_max[256];
  double sum = 0, amax = 0;
  for (int k = threadIdx.x; k < P; k += 256) {
    double2 f, a;
    const double am = prep_point(rho[(size_t)b * P + k], theta[(size_t)b * P + k], value[(size_t)b * P + k], x0,
                                 y0, f, a);
    fxy[((size_t)b * P + k) * 2] = f;
    fxy[((size_t)b * P + k) * 2 + 1] = a;
    sum += a.x;
    amax = fmax(amax, am);
  }
  s_sum[threadIdx.x] = sum;
  s_max[threadIdx.x] = amax;
  __syncthreads();
  prep_finish(s_sum, s_max, sparam + (size_t)b * PA_SCREEN_PARAMS, mode);
}

// The spectral draw and the prep in one launch (one block of 1024 threads per
// element): the same per-point arithmetic and the same summation order as
// k_spectral_draw followed by k_screen_prep (thread t < 256 adds points t,
// t + 256, t + 512, ... in turn), so fxy and sparam are identical; the
// draw's rho, theta, value are not stored.
constexpr int kDrawPrepThreads = 1024;
__global__ void __launch_bounds__(kDrawPrepThreads)
k_screen_draw_prep(int P, const float* __restrict__ base, const float* __restrict__ sqrt_psd,
                   const uint64_t* __restrict__ rid, uint64_t seed, int screen, double2* __restrict__ fxy,
                   double* __restrict__ sparam, int mode, const double* __restrict__ chparam,
                   const int* __restrict__ chan, const int* __restrict__ nsteps, int step) {
  const int b = blockIdx.x;
  if (nsteps && step >= nsteps[b]) return;
  const int ch = chan ? chan[b] : 0;
  const uint64_t id = rid[b];
  const double x0 = chparam[ch * kChParams + 1], y0 = chparam[ch * kChParams + 2];
  const float U = (float)u01(stream_word(id, screen, 0, seed));
  __shared__ double s_re[kDrawPrepThreads], s_am[kDrawPrepThreads];
  double sum = 0, amax = 0;
  for (int k0 = 0; k0 < P; k0 += kDrawPrepThreads) {
    const int k = k0 + threadIdx.x;
    double re = 0.0, am = 0.0;
    if (k < P) {
      float r, t;
      float2 v;
      draw_point(P, base, sqrt_psd, ch, id, seed, screen, k, U, r, t, v);
      double2 f, a;
      am = prep_point(r, t, v, x0, y0, f, a);
      fxy[((size_t)b * P + k) * 2] = f;
      fxy[((size_t)b * P + k) * 2 + 1] = a;
      re = a.x;
    }
    s_re[threadIdx.x] = re;
    s_am[threadIdx.x] = am;
    __syncthreads();
    if (threadIdx.x < 256) {
#pragma unroll
      for (int i = 0; i < kDrawPrepThreads / 256; ++i) {
        if (k0 + threadIdx.x + 256 * i < P) {
          sum += s_re[threadIdx.x + 256 * i];
          amax = fmax(amax, s_am[threadIdx.x + 256 * i]);
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 256) {
    s_re[threadIdx.x] = sum;
    s_am[threadIdx.x] = amax;
  }
  __syncthreads();
  prep_finish(s_re, s_am, sparam + (size_t)b * PA_SCREEN_PARAMS, mode);
}

// ------------------------------------------------------------------ polynomial block
// Points whose plane waves turn by at most kLowRankZ rad between the grid
// centre and its edge (X = n delta / 2) are not tabulated column by column.
// With x_c = X t_c (t_c = (c - n/2)/(n/2); the screen shift is in the
// amplitudes a', k_screen_prep) and likewise y_r = X s_r, the Jacobi-Anger
// expansion e^{i z t} = J_0(z) + 2 sum_{m>=1} i^m J_m(z) T_m(t) (T_m:
// Chebyshev polynomials) gives
//   e^{i alpha x_c} - 1 = sum_m P_m T_m(t_c),  P_0 = J_0(alpha X) - 1,
//                         P_m = 2 i^m J_m(alpha X)  (m >= 1),
// and the same for beta.  For points smooth along both axes, their share of
// phi2 = Re sum_k (a' q)(r) p(c) is
//   sum_{n,m} T_n(s_r) T_m(t_c) Re G[n][m],  G[n][m] = sum_k a'_k Q_kn P_km;
// for points smooth along the columns only (|alpha| X <= Z), only p expands:
//   sum_m T_m(t_c) Re V_m(r),  V_m(r) = sum_k (a' q)_k(r) P_km.
// Both share kLR = 32 K columns: B = T_m(t_c), A = sum_n T_n(s_r) Re G[n][m] +
// Re V_m(r) (Clenshaw; V from the tables' own plane waves q), instead of two
// columns per point.  |P_m|, |Q_n| <= 2, and |z| <= 14 leaves a truncation
// below 2 sum_{m>=32} |J_m(14)| ~ 2.4e-9 of the term, two orders below the
// 2^-22 of the split-fp16 operands (|z| <= 10: 1e-13, for ~6% more K).
// k_screen_classify groups the points of each element in fxy (a stable
// permutation): slots [0, n2) smooth along both axes, [n2, d0) along the
// columns only (at most kColLowMax, and as many as the B operand's spare room
// holds), then the rest, tabulated directly.  d0 is a multiple of 16 (the
// GEMM's K chunks; points past it stay direct), 0 below 32 (no block, fxy left
// as drawn); kbeg = 2 d0 - 32 is where the GEMM starts.  Then two passes over
// chunks of kLRChunk slots below d0.  k_screen_lowrank_part, grid (element,
// chunk): Bessel values of its points, the float64 partial sums of G (slots
// < n2), of the h coefficients Re sum_k a'_k Q_kn (slots < n2) and of the g
// coefficients Re sum_k a'_k P_km (slots < d0) (kLRSums values, the sparam
// order from offset 8), and the P_km of the column-smooth slots for the
// tables; k_screen_lowrank_sum adds the chunks' partials in chunk order.
constexpr int kLRChunk = 64;                                          // points per chunk
constexpr int kLRSums = kLR * kLR + 2 * kLR;                          // G, h and g sums
constexpr int kLRCoefH = 8 + kLR * kLR, kLRCoefG = kLRCoefH + kLR;   // sparam offsets
constexpr int kLRThreads = 512;                                       // waves 0-3: G, 4-7: h and g
static_assert(PA_LOWRANK_TERMS == 32, "the block is 32 K columns (kbeg = 2 d0 - 32)");
static_assert(2 * kLRChunk <= kLRThreads && kLRThreads == 512 && kLRChunk == 64, "polynomial block split");
constexpr int kLRMaxPoints = 1024;   // cap on d0
constexpr int kLRMaxChunks = kLRMaxPoints / kLRChunk;
static inline int lowrank_chunks(int P) { return (std::min(P, kLRMaxPoints) + kLRChunk - 1) / kLRChunk; }
constexpr int kColLowMax = 64;       // column-smooth points per element (LDS rows of the A tables)
constexpr int kClassThreads = 1024;   // points that may move: the first 1024 (the low radii)
// sparam slots of the grouping: [4] kbeg, [5] d0, [7] n2
constexpr int kSpKbeg = 4, kSpD0 = 5, kSpN2 = 7;

typedef double d4v __attribute__((ext_vector_type(4)));

// Exclusive count of the block's threads below this one with flag set, and
// the total into `total`.
__device__ __forceinline__ int block_rank(bool flag, int& total, int* s_w) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(flag);
  const int below = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) s_w[wave] = __popcll(m);
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kClassThreads / 64; ++w) {
    const int c = s_w[w];
    pre += w < wave ? c : 0;
    tot += c;
  }
  __syncthreads();   // s_w is reused by the next call
  total = tot;
  return pre + below;
}

__global__ void __launch_bounds__(kClassThreads)
k_screen_classify(int n, int P, double2* __restrict__ fxy, const double* __restrict__ chparam,
                  const int* __restrict__ chan, double* __restrict__ sparam, int nc_max,
                  const int* __restrict__ nsteps, int step) {
  const int b = blockIdx.x;
  if (nsteps && step >= nsteps[b]) return;
  const int ch = chan ? chan[b] : 0;
  const double X = 0.5 * n * chparam[ch * kChParams + 0];
  double* sp = sparam + (size_t)b * PA_SCREEN_PARAMS;
  const int G = min(P, kClassThreads), k = threadIdx.x;
  __shared__ int s_w[kClassThreads / 64];
  double2 f = make_double2(0.0, 0.0), a = make_double2(0.0, 0.0);
  int cls = 3;
  if (k < G) {
    f = fxy[((size_t)b * P + k) * 2];
    a = fxy[((size_t)b * P + k) * 2 + 1];
    const bool col = fabs(kTwoPi * f.x) * X <= kLowRankZ, row = fabs(kTwoPi * f.y) * X <= kLowRankZ;
    cls = col && row ? 0 : col ? 1 : 2;
  }
  int n2, n1;
  const int e0 = block_rank(cls == 0, n2, s_w);
  const int e1 = block_rank(cls == 1, n1, s_w);
  const int nc = min(n1, nc_max);
  int d0 = min((n2 + nc) / 16 * 16, kLRMaxPoints);
  if (d0 < 32) d0 = 0;
  if (threadIdx.x == 0) {
    sp[kSpKbeg] = d0 ? 2 * d0 - 32 : 0;
    sp[kSpD0] = d0;
    sp[kSpN2] = min(n2, d0);
  }
  if (!d0 || k >= G) return;   // no block: the points stay as drawn
  // every point was loaded (into registers) before the barriers above
  const int slot = cls == 0 ? e0 : (cls == 1 && e1 < nc) ? n2 + e1 : n2 + nc + (k - e0 - min(e1, nc));
  fxy[((size_t)b * P + slot) * 2] = f;
  fxy[((size_t)b * P + slot) * 2 + 1] = a;
}

__global__ void __launch_bounds__(kLRThreads)
k_screen_lowrank_part(int n, int P, const double2* __restrict__ fxy, const double* __restrict__ chparam,
                      const int* __restrict__ chan, const double* __restrict__ sparam, double* __restrict__ part,
                      double2* __restrict__ pcoef, int nc_max, const int* __restrict__ nsteps, int step) {
  const int b = blockIdx.x, k0 = blockIdx.y * kLRChunk;
  if (nsteps && step >= nsteps[b]) return;
  const int ch = chan ? chan[b] : 0;
  const double delta = chparam[ch * kChParams + 0];
  const double X = 0.5 * n * delta;
  const double* sp = sparam + (size_t)b * PA_SCREEN_PARAMS;
  const int d0 = (int)sp[kSpD0], n2 = (int)sp[kSpN2];
  if (k0 >= d0) return;   // the whole block: no chunk points in the polynomial block
  __shared__ double2 s_p[kLRChunk][kLR];   // P_km
  __shared__ double2 s_q[kLRChunk][kLR];   // a'_k Q_kn (zero for column-smooth slots)
  __shared__ double2 s_a[kLRChunk];        // a'_k
  const int cnt = min(kLRChunk, d0 - k0);   // a multiple of 16
  // (point j, side): Bessel values; side 1 (alpha, columns) gives P_km =
  // 2 i^m J_m (m >= 1) or J_0 - 1, side 0 (beta, rows) a'_k Q_kn; points
  // j >= cnt of a partial chunk, and side 0 of the column-smooth slots, are zero
  if (threadIdx.x < 2 * kLRChunk) {
    const int j = threadIdx.x % kLRChunk, side = threadIdx.x / kLRChunk, slot = k0 + j;
    double J[kLR];
    double2 a = make_double2(0.0, 0.0);
    if (j < cnt && (side == 1 || slot < n2)) {
      const double2 f = fxy[((size_t)b * P + slot) * 2];
      a = fxy[((size_t)b * P + slot) * 2 + 1];
      bessel_j<kLR>(kTwoPi * (side == 0 ? f.y : f.x) * X, J);
    } else {
#pragma unroll
      for (int m = 0; m < kLR; ++m) J[m] = m == 0 ? 1.0 : 0.0;
    }
    // a column-smooth slot's P_km is kept for the tables
    double2* pc = (side == 1 && j < cnt && slot >= n2) ? pcoef + ((size_t)b * nc_max + (slot - n2)) * kLR : nullptr;
#pragma unroll
    for (int m = 0; m < kLR; ++m) {
      const double v = m == 0 ? J[0] - 1.0 : 2.0 * J[m];
      // times i^m
      const double cr = (m & 3) == 0 ? v : (m & 3) == 2 ? -v : 0.0;
      const double ci = (m & 3) == 1 ? v : (m & 3) == 3 ? -v : 0.0;
      if (side == 1) {
        s_p[j][m] = make_double2(cr, ci);
        if (pc) pc[m] = make_double2(cr, ci);
      } else {
        s_q[j][m] = make_double2(a.x * cr - a.y * ci, a.x * ci + a.y * cr);
      }
    }
    if (side == 1) s_a[j] = a;
  }
  __syncthreads();
  double* out = part + ((size_t)b * gridDim.y + blockIdx.y) * kLRSums;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave < 4) {
    // G[n][m] = sum_k A[n][k] B[k][m] over K = 2 cnt on v_mfma_f64_16x16x4f64,
    // A[n][2j] = Re(a'Q)_jn, A[n][2j+1] = -Im(a'Q)_jn, B[2j][m] = Re P_jm,
    // B[2j+1][m] = Im P_jm; wave w: the 16 x 16 tile (rows 16 (w / 2),
    // columns 16 (w % 2)).  Lane l supplies A[l % 16][l / 16] and
    // B[l / 16][l % 16] of each K step of 4 and holds D[4 i + l / 16][l % 16].
    const int n0 = 16 * (wave >> 1), m0 = 16 * (wave & 1);
    const int r = lane & 15, kk = lane >> 4, jo = kk >> 1;
    const bool im = kk & 1;
    d4v acc = {0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < cnt; j += 2) {
      const double2 q = s_q[j + jo][n0 + r], p = s_p[j + jo][m0 + r];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(im ? -q.y : q.x, im ? p.y : p.x, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(n0 + 4 * i + kk) * kLR + m0 + r] = acc[i];
  } else {
    // h (t < 32) and g sums: 4 lanes per sum, 16 points each, then the 4
    // lanes' values added by xor shuffles
    const int t = threadIdx.x - 256, o = t >> 2, sub = t & 3;
    double v = 0.0;
    for (int j = sub; j < cnt; j += 4) {
      if (o < kLR) {
        v += s_q[j][o].x;                                        // h: Re sum a' Q_kn
      } else {
        const double2 aa = s_a[j], pp = s_p[j][o - kLR];
        v += aa.x * pp.x - aa.y * pp.y;                          // g: Re sum a' P_km
      }
    }
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    if (sub == 0) out[kLR * kLR + o] = v;
  }
}

// One thread per sum: the chunks' partials in chunk order.  Then the A scale:
// the direct points' scale sp[6], lowered (split GEMM modes) if the block's
// column bounds sum_n |Re G[n][m]| + 2 sum_t |a'_t| |P_tm| (on its A entries,
// |T_n| <= 1, |q| <= 2) would pass 2^14, into sp[1] with 1/(A scale B scale)
// in sp[3] for the GEMM.
__global__ void __launch_bounds__(1024)
k_screen_lowrank_sum(int P, const double2* __restrict__ fxy, double* __restrict__ sparam,
                     const double* __restrict__ part, const double2* __restrict__ pcoef, int nc_max, int chunks,
                     int mode, const int* __restrict__ nsteps, int step) {
  const int b = blockIdx.x;
  if (nsteps && step >= nsteps[b]) return;
  double* sp = sparam + (size_t)b * PA_SCREEN_PARAMS;
  const int d0 = (int)sp[kSpD0], n2 = (int)sp[kSpN2];
  if (!d0) return;   // prep's scales stand
  const int used = (d0 + kLRChunk - 1) / kLRChunk;   // <= chunks <= kLRMaxChunks
  __shared__ double s_g[kLR * kLR];
  for (int e = threadIdx.x; e < kLRSums; e += 1024) {
    const double* pb = part + (size_t)b * chunks * kLRSums + e;
    double v[kLRMaxChunks];
#pragma unroll
    for (int c = 0; c < kLRMaxChunks; ++c) v[c] = c < used ? pb[(size_t)c * kLRSums] : 0.0;
    double sum = 0.0;
#pragma unroll
    for (int c = 0; c < kLRMaxChunks; ++c) sum += v[c];
    sp[8 + e] = sum;
    if (e < kLR * kLR) s_g[e] = sum;
  }
  if (mode == PA_GEMM_F32) return;
  __syncthreads();
  // |Re V_m(r)| <= sum_t |a'_t| |q_t(r)| |P_tm| with |q| <= 2: the terms
  // 2 |a'_t| |P_tm| of column m = tid % 32 over t = tid / 32 (+ 32), then
  // added over t in order
  __shared__ double s_v[32][kLR];
  {
    const int m = threadIdx.x % kLR, t = threadIdx.x / kLR;
    double v = 0.0;
    for (int tt = t; tt < d0 - n2; tt += 1024 / kLR) {
      const double2 a = fxy[((size_t)b * P + n2 + tt) * 2 + 1];
      const double2 pm = pcoef[((size_t)b * nc_max + tt) * kLR + m];
      v += 2.0 * sqrt(a.x * a.x + a.y * a.y) * sqrt(pm.x * pm.x + pm.y * pm.y);
    }
    s_v[t][m] = v;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    double colsum = 0.0;
    if (threadIdx.x < kLR) {
      for (int q = 0; q < kLR; ++q) colsum += fabs(s_g[q * kLR + threadIdx.x]);
      for (int t = 0; t < 1024 / kLR; ++t) colsum += s_v[t][threadIdx.x];
    }
    for (int off = 32; off > 0; off >>= 1) colsum = fmax(colsum, __shfl_xor(colsum, off, 64));
    if (threadIdx.x == 0) {
      double scale = sp[6];
      if (colsum > 0.0) scale = fmin(scale, ldexp(1.0, 14 - (int)ceil(log2(colsum))));
      sp[1] = scale;
      sp[3] = 1.0 / (scale * sp[2]);
    }
  }
}

// ------------------------------------------------------------------ operand stores
__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Write the pair (v0, v1) at [row][2k] of a K-contiguous float32 operand.
__device__ __forceinline__ void put_pair_f32(void* op, int Kp, int row, int k, double v0, double v1) {
  *reinterpret_cast<float2*>(reinterpret_cast<float*>(op) + (size_t)row * Kp + 2 * k) =
      make_float2((float)v0, (float)v1);
}

// Split (v0, v1) into hi and lo halves (fp16 or bf16): returns the packed hi
// pair, the lo pair in `lo`.
template <int MODE>
__device__ __forceinline__ uint32_t split_pack(double v0, double v1, uint32_t& lo) {
  float f0 = (float)v0, f1 = (float)v1;
  // opaque to the optimizer: otherwise the two roundings (double -> float ->
  // half) are merged into one double -> half rounding, which gfx950 has no
  // instruction for (a ~25-instruction software sequence per value)
  asm volatile("" : "+v"(f0), "+v"(f1));
  if constexpr (MODE == PA_GEMM_F16X3) {
    // pair conversions (v_cvt_pk_f16_f32, round to nearest even like the
    // scalar ones): hi = f16(f), lo = f16(f - hi)
    const pa_f2v f = {f0, f1};
    const pa_h2v h = __builtin_convertvector(f, pa_h2v);
    const pa_h2v l = __builtin_convertvector(f - __builtin_convertvector(h, pa_f2v), pa_h2v);
    lo = __builtin_bit_cast(uint32_t, l);
    return __builtin_bit_cast(uint32_t, h);
  } else {
    const uint16_t h0 = bf16_rne(f0), h1 = bf16_rne(f1);
    lo = (uint32_t)bf16_rne(f0 - bf16_to_f(h0)) | ((uint32_t)bf16_rne(f1 - bf16_to_f(h1)) << 16);
    return (uint32_t)h0 | ((uint32_t)h1 << 16);
  }
}

// ------------------------------------------------------------------ tables
// One block per RT = TC * RG consecutive rows (AXIS = 0: A operand, h, eh) or
// columns (AXIS = 1: B operand, g, eg) of element b; threads stride over k.
// Along the block's rows the plane wave advances by a fixed angle, so each
// thread forms e^{i t} for its RT points from one float64 sincos of the first
// angle and one of the step (the product recurrence loses ~1e-16 per step);
// the operand is staged and written TC rows at a time.  The
// recurrence runs on w = z e^{it} with z = scale a (AXIS 0) or z = scale
// (AXIS 1), so an entry scale a (e^{it} - 1) or scale (e^{it} - 1) is w - z;
// the 1-D sums are accumulated times `scale` (a power of two, divided out
// exactly at the end).  RT (4, 8, 16 or 32) trades sincos work and per-block
// set-up against the number of blocks.
template <int MODE, int AXIS, int TC, int RG>
__global__ void __launch_bounds__(256, 4)
k_screen_tables(int n, int P, int Kp, const double2* __restrict__ fxy,
                double* __restrict__ sparam, const double* __restrict__ chparam,
                const int* __restrict__ chan, void* __restrict__ op, double* __restrict__ hg,
                float2* __restrict__ esc, const double2* __restrict__ pcoef, int nc_max,
                const int* __restrict__ nsteps, int step) {
  constexpr int RT = TC * RG;
  static_assert(RT == 4 || RT == 8 || RT == 16 || RT == 32, "rows per block");
  const int b = blockIdx.y;
  if (nsteps && step >= nsteps[b]) return;
  const int i0 = blockIdx.x * RT;  // first row (AXIS 0) or column (AXIS 1)
  const int ch = chan ? chan[b] : 0;
  const double delta = chparam[ch * kChParams + 0];
  const double coord0 = (double)(i0 - n / 2) * delta;   // the shift is in the amplitudes (k_screen_prep)
  double* sp = sparam + (size_t)b * PA_SCREEN_PARAMS;
  // polynomial block (k_screen_classify): slots below klow = d0 are carried
  // by it, slots [n2, klow) are column-smooth
  const int kbeg = (int)sp[kSpKbeg], klow = (int)sp[kSpD0], n2 = (int)sp[kSpN2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // operand scales (k_screen_prep, the A scale lowered by k_screen_lowrank_sum
  // for the polynomial block); AXIS 0 stages G in LDS for the block's columns
  const double scale = sp[AXIS == 0 ? 1 : 2];
  __shared__ double s_g[AXIS == 0 ? kLR * kLR : 1];
  if constexpr (AXIS == 0) {
    if (klow)
      for (int e = threadIdx.x; e < kLR * kLR; e += 256) s_g[e] = sp[8 + e];
  }
  constexpr bool SPLIT = MODE != PA_GEMM_F32;
  const size_t plane_bytes = (size_t)n * Kp * (SPLIT ? 2 : 4);
  char* opb = reinterpret_cast<char*>(op) + (size_t)b * plane_bytes * (SPLIT ? 2 : 1);
  // Split modes stage the block's TC rows x 256 points (packed hi and lo
  // pairs) in LDS, then write them out as 16-byte pieces of the K-blocked
  // operand: the TC rows of one 16-element chunk are one contiguous run of
  // 32 TC bytes.  Rows are padded so that 8 consecutive pieces read from 8
  // distinct 16-byte bank groups.
  constexpr int SROW = 256 + 8;
  // (before the direct points the same LDS stages the column-smooth slots'
  // plane waves)
  __shared__ __attribute__((aligned(16))) uint32_t s_stage[2 * TC * SROW];
  uint32_t* s_hi = s_stage;
  uint32_t* s_lo = s_stage + TC * SROW;
  // AXIS 0, first: the column-smooth slots, CS at a time.  Their P_km go to
  // LDS (s_pc) while their plane waves (a' q)(r) for the block's rows are
  // formed by the same recurrence as the direct points, thread (slot t, row
  // group g) from two float64 sincos at the group's first row, and staged
  // unscaled (s_aq, in the operand staging LDS); then each thread adds their
  // share of V_m(r) to its entries of the polynomial columns (vacc, entry e =
  // tid + 256 i, column m = tid % 32) and thread j < RT their share of h(r)
  // (hsh), in slot order.
  constexpr int NV = (RT * kLR + 255) / 256;
  constexpr int CS = 16, RGS = RT >= 16 ? 16 : RT, RPG = RT / RGS;   // slots, row groups, rows per group
  static_assert(RT * CS * 16 <= 2 * TC * SROW * 4, "column-smooth staging fits the operand staging");
  __shared__ double2 s_pc[AXIS == 0 ? CS * kLR : 1];
  double vacc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) vacc[i] = 0.0;
  double hsh = 0.0;
  if constexpr (AXIS == 0) {
    double2* s_aq = reinterpret_cast<double2*>(s_stage);
    const double2* pc = pcoef + (size_t)b * nc_max * kLR;
    const int ncs = klow - n2, t = threadIdx.x % CS, g = threadIdx.x / CS;
    // the next sub-chunk's P_km entries (e = tid, tid + 256) and this thread's
    // slot are fetched one sub-chunk ahead, so their latency overlaps the work
    static_assert(CS * kLR == 2 * 256, "two P_km entries per thread and sub-chunk");
    double2 pn[2], fn = make_double2(0.0, 0.0), an = make_double2(0.0, 0.0);
    auto fetch = [&](int t0) {
      const int cnt = min(CS, ncs - t0);
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = threadIdx.x + 256 * u;
        pn[u] = e < cnt * kLR ? pc[(size_t)t0 * kLR + e] : make_double2(0.0, 0.0);
      }
      if (t < cnt) {
        fn = fxy[((size_t)b * P + n2 + t0 + t) * 2];
        an = fxy[((size_t)b * P + n2 + t0 + t) * 2 + 1];
      }
    };
    if (ncs > 0) fetch(0);
    for (int t0 = 0; t0 < ncs; t0 += CS) {
      const int cnt = min(CS, ncs - t0);
      const double2 f = fn, a = an;
      __syncthreads();   // the staging LDS is free
      s_pc[threadIdx.x] = pn[0];
      s_pc[threadIdx.x + 256] = pn[1];
      if (t0 + CS < ncs) fetch(t0 + CS);
      if (t < cnt && g < RGS) {
        double s0, c0, sd, cd;
        sincos(kTwoPi * f.y * (coord0 + (double)(g * RPG) * delta), &s0, &c0);
        sincos(kTwoPi * f.y * delta, &sd, &cd);
        double wr = a.x * c0 - a.y * s0, wi = a.x * s0 + a.y * c0;
#pragma unroll
        for (int j = 0; j < RPG; ++j) {
          s_aq[(g * RPG + j) * CS + t] = make_double2(wr - a.x, wi - a.y);
          const double nr = wr * cd - wi * sd;
          wi = wr * sd + wi * cd;
          wr = nr;
        }
      }
      __syncthreads();
      const int m = threadIdx.x % kLR;
#pragma unroll 4
      for (int tt = 0; tt < cnt; ++tt) {
        const double2 pm = s_pc[tt * kLR + m];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          const int e = threadIdx.x + 256 * i;
          if (e < RT * kLR) {
            const double2 q = s_aq[(e / kLR) * CS + tt];
            vacc[i] += q.x * pm.x - q.y * pm.y;
          }
        }
      }
      if ((int)threadIdx.x < RT)
#pragma nounroll
        for (int tt = 0; tt < cnt; ++tt) hsh += s_aq[threadIdx.x * CS + tt].x;
    }
    __syncthreads();   // before the direct points reuse the staging LDS
  }
  double acc[RT];   // this thread's share of the 1-D sums, times scale
#pragma unroll
  for (int j = 0; j < RT; ++j) acc[j] = 0;
  // points below klow are carried by the polynomial block and the Chebyshev
  // sums added below
  for (int k0 = (klow / 256) * 256; k0 < Kp / 2; k0 += 256) {
    const int k = k0 + threadIdx.x;
    const int W = min(256, Kp / 2 - k0);  // points in this group (a multiple of 16)
    const bool live = k < Kp / 2 && k >= klow;
    // padding points (k >= P) get zero amplitude and frequency: exact zeros
    double s0 = 0.0, c0 = 1.0, sd = 0.0, cd = 1.0;
    double2 a = make_double2(0.0, 0.0);
    if (live && k < P) {
      const double2 f = fxy[((size_t)b * P + k) * 2];
      const double fk = AXIS == 0 ? f.y : f.x;
      sincos(kTwoPi * fk * coord0, &s0, &c0);
      sincos(kTwoPi * fk * delta, &sd, &cd);
      a = fxy[((size_t)b * P + k) * 2 + 1];
    }
    const double zr = AXIS == 0 ? a.x * scale : scale, zi = AXIS == 0 ? a.y * scale : 0.0;
    double wr = zr * c0 - zi * s0, wi = zr * s0 + zi * c0;
#pragma unroll
    for (int g = 0; g < RG; ++g) {
      if (live) {
#pragma unroll
        for (int j = 0; j < TC; ++j) {
          const double v0 = wr - zr, pi = wi - zi;
          double v1;
          if constexpr (AXIS == 0) {
            acc[g * TC + j] += v0;
            v1 = -pi;
          } else {
            acc[g * TC + j] += a.x * v0 - a.y * pi;
            v1 = pi;
          }
          if constexpr (SPLIT) {
            uint32_t lo;
            s_hi[j * SROW + threadIdx.x] = split_pack<MODE>(v0, v1, lo);
            s_lo[j * SROW + threadIdx.x] = lo;
          } else {
            put_pair_f32(opb, Kp, i0 + g * TC + j, k, v0, v1);
          }
          const double nr = wr * cd - wi * sd;
          wi = wr * sd + wi * cd;
          wr = nr;
        }
      }
      if constexpr (SPLIT) {
        __syncthreads();
        // piece p: half h (4 points) of row j of chunk c (8 points) of this
        // group; chunks of 8 points below klow (a multiple of 16) are not written
        const int c0 = max(0, (klow - k0) / 8), r0 = i0 + g * TC;
        for (int pc = threadIdx.x + 2 * TC * c0; pc < TC * (W / 8) * 2; pc += 256) {
          const int h = pc & 1, j = (pc >> 1) % TC, c = (pc >> 1) / TC;
          const int so = j * SROW + c * 8 + h * 4;
          const size_t go = (((size_t)(k0 / 8 + c) * n + r0 + j) * 32) + h * 16;
          *reinterpret_cast<u4v*>(opb + go) = *reinterpret_cast<const u4v*>(&s_hi[so]);
          *reinterpret_cast<u4v*>(opb + plane_bytes + go) = *reinterpret_cast<const u4v*>(&s_lo[so]);
        }
        __syncthreads();
      }
    }
  }
  // the polynomial block's kLR = 32 columns [kbeg, kbeg + 32) of these rows
  // (AXIS 0: sum_n T_n(s) Re G[n][m] by Clenshaw + Re V_m(r)) or columns (AXIS
  // 1: T_m(t)); thread e: row e / 32, column e % 32, the even lane of each
  // pair stores both values (RT * 32 is a multiple of 64, so the loop is
  // uniform over each wave)
  if (klow) {
    if constexpr (AXIS == 0) __syncthreads();   // s_g
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = threadIdx.x + 256 * i;
      if (e >= RT * kLR) break;
      const int j = e / kLR, m = e % kLR;
      const double u = (double)(i0 + j - n / 2) / (double)(n / 2);
      double val;
      if constexpr (AXIS == 0) {
        double b1 = 0.0, b2 = 0.0;
#pragma unroll
        for (int q = kLR - 1; q >= 1; --q) {
          const double b0 = s_g[q * kLR + m] + 2.0 * u * b1 - b2;
          b2 = b1;
          b1 = b0;
        }
        val = s_g[m] + u * b1 - b2 + vacc[i];
      } else {
        double t0 = 1.0, t1 = u;
        for (int q = 1; q < m; ++q) {
          const double t2 = 2.0 * u * t1 - t0;
          t0 = t1;
          t1 = t2;
        }
        val = m == 0 ? 1.0 : t1;
      }
      const double v0 = val * scale;
      const double v1 = __shfl_down(v0, 1, 64);
      if (!(m & 1)) {
        const int pr = m / 2;
        if constexpr (SPLIT) {
          uint32_t lo;
          const uint32_t hi = split_pack<MODE>(v0, v1, lo);
          const size_t go = ((size_t)(kbeg / 16 + pr / 8) * n + i0 + j) * 32 + (2 * pr % 16) * 2;
          *reinterpret_cast<uint32_t*>(opb + go) = hi;
          *reinterpret_cast<uint32_t*>(opb + plane_bytes + go) = lo;
        } else {
          put_pair_f32(opb, Kp, i0 + j, kbeg / 2 + pr, v0, v1);
        }
      }
    }
  }
  // reduce the RT sums over the block.  Within a wave, halving exchanges:
  // at offset 32 a lane keeps the half of its sums selected by bit 5 of its
  // index and adds the partner's copy of that half, and so on down to one sum
  // per lane; lanes then hold the sum of row jr(lane) over 64 / RT lanes,
  // finished by xor shuffles over the remaining offsets.
  constexpr int LG = RT == 4 ? 2 : RT == 8 ? 3 : RT == 16 ? 4 : 5;
#pragma unroll
  for (int s = 0; s < LG; ++s) {
    const int half = RT >> (s + 1), off = 32 >> s;
    const bool up = lane & off;
#pragma unroll
    for (int t = 0; t < half; ++t) {
      const double keep = up ? acc[t + half] : acc[t];
      const double send = up ? acc[t] : acc[t + half];
      acc[t] = keep + __shfl_xor(send, off, 64);
    }
  }
  double tot = acc[0];
#pragma unroll
  for (int off = 32 >> LG; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
  // row of this lane's sum: bit 5 chose the upper half first, so it is the
  // most significant bit of the row index
  int jr = 0;
#pragma unroll
  for (int s = 0; s < LG; ++s) jr |= ((lane >> (5 - s)) & 1) << (LG - 1 - s);
  __shared__ double red[4][RT];
  if ((lane & ((32 >> (LG - 1)) - 1)) == 0) red[wave][jr] = tot;
  __syncthreads();
  if (threadIdx.x < RT) {
    const int j = threadIdx.x, i = i0 + j;
    double v = (red[0][j] + red[1][j] + red[2][j] + red[3][j]) * (1.0 / scale);
    if constexpr (AXIS == 0) v += hsh;   // the column-smooth slots' share
    // the polynomial points' share of h(r) (AXIS 0) or g(c): sum_n T_n(u) c[n]
    // by Clenshaw
    if (klow) {
      const double* cc = sp + (AXIS == 0 ? kLRCoefH : kLRCoefG);
      const double u = (double)(i - n / 2) / (double)(n / 2);
      double b1 = 0.0, b2 = 0.0;
#pragma unroll
      for (int q = kLR - 1; q >= 1; --q) {
        const double b0 = cc[q] + 2.0 * u * b1 - b2;
        b2 = b1;
        b1 = b0;
      }
      v += cc[0] + u * b1 - b2;
    }
    hg[((size_t)b * 2 + AXIS) * n + i] = v;
    const double ph = AXIS == 0 ? v + sp[0] : v;  // fold the piston into the row factor
    double sn, cs;
    sincos(ph, &sn, &cs);
    esc[((size_t)b * 2 + AXIS) * n + i] = make_float2((float)cs, (float)-sn);
  }
}

// ------------------------------------------------------------------ GEMM
// phi2 tile = A_tile B_tile^T over K = Kp (a multiple of 32), batched over the
// elements of the batch in one 1-D grid.  Block tiles are visited in an
// XCD-aware order: consecutive block ids land on different XCDs (round robin),
// so the remap gives each XCD a contiguous range of tiles; tiles are walked in
// 4 x 8 super-tiles so one XCD's blocks share A and B panels in its L2.
template <int NTM, int NTN>
__device__ __forceinline__ void gemm_tile(int L, int total, int& b, int& ty, int& tx) {
  constexpr int GY = NTM >= 4 ? 4 : NTM, GX = NTN >= 8 ? 8 : NTN;
  const int logical = (total % 8 == 0) ? (L % 8) * (total / 8) + L / 8 : L;
  b = logical / (NTM * NTN);
  const int t = logical % (NTM * NTN);
  const int st = t / (GY * GX), i = t % (GY * GX);
  ty = (st / (NTN / GX)) * GY + i / GX;
  tx = (st % (NTN / GX)) * GX + i % GX;
}

// Epilogue shared by the GEMM kernels: phi2 = acc / (sA sB); writes, each
// nullable and in the work layout, the real phi2, the screen angle
// phi2 + (ah[r] + ag[c]) (ah, ag: the angles of the row and column factors,
// in [-pi, pi]; e^{-i angle} is the screen factor the row pass applies) and
// the complex screen factor eh[r] eg[c] e^{-i phi2}.  The wave's tile starts
// at global (m0 + r0, n0 + c0); seh / seg hold the block's row / column
// factors (eh[m0 ..], eg[n0 ..]) and sah / sag their angles, in LDS.  A 32x32 MFMA block leaves element
// (8 q + 4 (lane / 32) + u, lane % 32) in acc[4 q + u]: rows row..row+3 are
// consecutive inside one row tile (T >= 4) or two (T == 2), so each lane
// stores runs of 4 (or 2) values.
// Streaming (non-temporal) vector stores of the epilogue: the screen is read
// once, by the next row pass, and its lines should not displace the operand
// panels that the XCD's next tiles read from L2.
__device__ __forceinline__ void st_nt(float* p, float a, float b, float c, float d) {
  const pa_f4v v = {a, b, c, d};
  __builtin_nontemporal_store(v, reinterpret_cast<pa_f4v*>(p));
}
__device__ __forceinline__ void st_nt(float* p, float a, float b) {
  const pa_f2v v = {a, b};
  __builtin_nontemporal_store(v, reinterpret_cast<pa_f2v*>(p));
}

template <int N, int MI, int NI>
__device__ __forceinline__ void gemm_epilogue(const f16v (&acc)[MI][NI], int m0, int n0, int r0, int c0, int lane,
                                              float inv, float* __restrict__ out, float* __restrict__ ao,
                                              float2* __restrict__ fo, const float2* seh, const float2* seg,
                                              const float* sah, const float* sag) {
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      const int lc = c0 + j * 32 + (lane & 31), col = n0 + lc;
      const float2 gc = seg[lc];
      const float agc = sag[lc];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int lr = r0 + i * 32 + 8 * q + 4 * (lane >> 5), row = m0 + lr;
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = acc[i][j][4 * q + u] * inv;
        if (out) {
          if constexpr (Geo<N>::T >= 4) {
            st_nt(out + tiled<N>(row, col), v[0], v[1], v[2], v[3]);
          } else {
            st_nt(out + tiled<N>(row, col), v[0], v[1]);
            st_nt(out + tiled<N>(row + 2, col), v[2], v[3]);
          }
        }
        if (ao) {
          float a[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) a[u] = v[u] + (sah[lr + u] + agc);
          if constexpr (Geo<N>::T >= 4) {
            st_nt(ao + tiled<N>(row, col), a[0], a[1], a[2], a[3]);
          } else {
            st_nt(ao + tiled<N>(row, col), a[0], a[1]);
            st_nt(ao + tiled<N>(row + 2, col), a[2], a[3]);
          }
        }
        if (fo) {
          float2 f[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) f[u] = cmul(cmul(seh[lr + u], gc), expmi(v[u]));
          if constexpr (Geo<N>::T >= 4) {
            float* d = reinterpret_cast<float*>(fo + tiled<N>(row, col));
            st_nt(d, f[0].x, f[0].y, f[1].x, f[1].y);
            st_nt(d + 4, f[2].x, f[2].y, f[3].x, f[3].y);
          } else {
            st_nt(reinterpret_cast<float*>(fo + tiled<N>(row, col)), f[0].x, f[0].y, f[1].x, f[1].y);
            st_nt(reinterpret_cast<float*>(fo + tiled<N>(row + 2, col)), f[2].x, f[2].y, f[3].x, f[3].y);
          }
        }
      }
    }
}

// ---- split-precision (f16x3 / bf16x3) GEMM
// Operands are K-blocked: plane (hi or lo) of element b holds, for each
// 16-wide K chunk q, the N rows' 32-byte pieces back to back:
//   byte ((q N + row) 16 + k % 16) 2 of the plane, q = k / 16,
// so one chunk of a block's rows is a contiguous run of HBM.
// Geometry: BM x BN block tile, WGM x WGN waves, each wave owning a
// (BM/WGM) x (BN/WGN) tile of 32x32 MFMA blocks.  Grids of 512 and up use
// 256 x 256 block tiles (8 waves of 128 x 64); smaller grids 128 x 128 (or
// N x N) with 4 waves.
// CFG 1: 256 x 256 block tiles (8 waves of 128 x 64) for grids of 512 and up;
// CFG 2: 128 x 128 (or N x N) tiles with 4 waves of 64 x 64, 4x the blocks.
template <int N, int CFG>
struct SplitGeo {
  static constexpr bool BIG = N >= 512 && CFG == 1;
  static constexpr int BM = BIG ? 256 : (N < 128 ? N : 128);
  static constexpr int BN = BM;
  static constexpr int WGM = 2, WGN = BIG ? 4 : 2;
  static constexpr int NT = 64 * WGM * WGN;
  static constexpr int WM = BM / WGM, WN = BN / WGN, MI = WM / 32, NI = WN / 32;
  static constexpr int NTM = N / BM, NTN = N / BN;
  // one LDS stage: one K chunk of the block's rows, A hi | A lo | B hi | B lo,
  // 32 bytes per row, moved by 1-KB (32-row) segments
  static constexpr int STAGE = 64 * (BM + BN);
  static constexpr int SEGS = 2 * (BM + BN) / 32;
  static constexpr int SEG_PER_WAVE = SEGS / (NT / 64);
  static constexpr int DEPTH = 4;                 // LDS stages: 2 chunks in flight, 1 in registers
  static_assert(MI >= 1 && NI >= 1, "wave tile");
  static_assert(SEGS % (NT / 64) == 0, "segments per wave");
  static_assert(DEPTH * STAGE <= 160 * 1024 && DEPTH >= 3, "LDS stages");
  static constexpr int FPT = (BM + BN + NT - 1) / NT;   // row/column factors staged per thread
};

// Within a stage, row r's two 16-byte halves sit at positions h ^ sw16(r).
// Any 8 aligned rows read at one h touch 8 distinct 16-byte groups of a
// 128-byte bank sweep, any 16 aligned rows 16 distinct groups of 256 bytes.
__device__ __forceinline__ int sw16(int r) { return ((r >> 2) ^ (r >> 3)) & 1; }

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// s_waitcnt immediate that waits until at most v vector-memory operations of
// this wave are outstanding (expcnt and lgkmcnt left at their maxima).
constexpr unsigned vmcnt_imm(unsigned v) { return (v & 15u) | ((v >> 4) << 14) | (7u << 4) | (15u << 8); }

// 16 bytes per lane from global memory straight into LDS: lane l writes LDS
// bytes lds + 16 l .. lds + 16 l + 15 (lds is wave-uniform, passed in M0).
// Written as an asm statement so that the compiler does not add its own
// conservative wait for all LDS DMA before every LDS read: completion is
// tracked explicitly with vmcnt waits and barriers by the caller, and the
// memory clobber keeps LDS accesses from moving across the issue.
// M0 is a reserved register (the compiler sets it before each of its own
// uses), so naming it in the clobber list only draws a diagnostic.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void dma16(const void* src, uint32_t lds) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(src), "s"(lds)
               : "memory", "m0");
}
#pragma clang diagnostic pop

__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) char*)p;
}

// Issue this wave's segments of K chunk q into the LDS stage at `stage`:
// segment s covers 32 rows of one plane (A hi, A lo, B hi, B lo in order).
template <class G, int NROWS>
__device__ __forceinline__ void split_issue(uint32_t stage, const char* ga, const char* gb, size_t plane,
                                            int lofs, int wave, int q) {
#pragma unroll
  for (int t = 0; t < G::SEG_PER_WAVE; ++t) {
    const int sgi = wave * G::SEG_PER_WAVE + t;   // uniform
    const bool isa = sgi < 2 * G::BM / 32;
    const int s2 = isa ? sgi : sgi - 2 * G::BM / 32;
    const int rows = isa ? G::BM / 32 : G::BN / 32;
    const int pl = s2 / rows, rb = s2 % rows;
    const char* src = (isa ? ga : gb) + pl * plane + ((size_t)q * NROWS + rb * 32) * 32 + lofs;
    dma16(src, stage + sgi * 1024);
  }
}

// One 16-wide K chunk of the wave tile: its fragments read out of LDS
// (split_frags), then 3 MFMAs per 32x32 block, the small cross terms first and
// the hi x hi product last (split_mma).
template <class G>
struct SplitFrags {
  u4v ahi[G::MI], alo[G::MI], bhi[G::NI], blo[G::NI];
};
template <class G>
__device__ __forceinline__ void split_frags(SplitFrags<G>& f, const char* cA, const char* cB, int fo) {
#pragma unroll
  for (int i = 0; i < G::MI; ++i) {
    f.ahi[i] = *reinterpret_cast<const u4v*>(cA + i * 32 * 32 + fo);
    f.alo[i] = *reinterpret_cast<const u4v*>(cA + (G::BM + i * 32) * 32 + fo);
  }
#pragma unroll
  for (int j = 0; j < G::NI; ++j) {
    f.bhi[j] = *reinterpret_cast<const u4v*>(cB + j * 32 * 32 + fo);
    f.blo[j] = *reinterpret_cast<const u4v*>(cB + (G::BN + j * 32) * 32 + fo);
  }
}
template <class G, int MODE>
__device__ __forceinline__ void split_mma(f16v (&acc)[G::MI][G::NI], const SplitFrags<G>& f) {
#pragma unroll
  for (int i = 0; i < G::MI; ++i)
#pragma unroll
    for (int j = 0; j < G::NI; ++j) {
      if constexpr (MODE == PA_GEMM_F16X3) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8, f.alo[i]),
                                                           __builtin_bit_cast(h8, f.bhi[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8, f.ahi[i]),
                                                           __builtin_bit_cast(h8, f.blo[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8, f.ahi[i]),
                                                           __builtin_bit_cast(h8, f.bhi[j]), acc[i][j], 0, 0, 0);
      } else {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8, f.alo[i]),
                                                            __builtin_bit_cast(b8, f.bhi[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8, f.ahi[i]),
                                                            __builtin_bit_cast(b8, f.blo[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8, f.ahi[i]),
                                                            __builtin_bit_cast(b8, f.bhi[j]), acc[i][j], 0, 0, 0);
      }
    }
}

template <int N, int MODE, int CFG>
__global__ void __launch_bounds__((SplitGeo<N, CFG>::NT))
k_screen_gemm_split(const void* __restrict__ opa, const void* __restrict__ opb,
                    const double* __restrict__ sparam, const float2* __restrict__ esc, float* __restrict__ phi2,
                    float* __restrict__ ang, float2* __restrict__ fac, int Kp, const int* __restrict__ nsteps,
                    int step, int batch) {
  using G = SplitGeo<N, CFG>;
  constexpr int BM = G::BM, MI = G::MI, NI = G::NI, DEPTH = G::DEPTH;

  int b, ty, tx;
  gemm_tile<G::NTM, G::NTN>(blockIdx.x, G::NTM * G::NTN * batch, b, ty, tx);
  if (nsteps && step >= nsteps[b]) return;
  __shared__ __attribute__((aligned(1024))) char lds[DEPTH * G::STAGE];
  const uint32_t lbase = __builtin_amdgcn_readfirstlane(lds_addr(lds));

  const int m0 = ty * BM, n0 = tx * G::BN;
  const size_t plane = (size_t)N * Kp * 2;          // bytes of one operand plane
  const char* ga = reinterpret_cast<const char*>(opa) + (size_t)b * 2 * plane + (size_t)m0 * 32;
  const char* gb = reinterpret_cast<const char*>(opb) + (size_t)b * 2 * plane + (size_t)n0 * 32;

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave / G::WGN, wn = wave % G::WGN;
  // DMA source of lane l within a segment: row l / 2, the half stored at position l % 2
  const int lofs = (lane >> 1) * 32 + (((lane & 1) ^ sw16(lane >> 1)) << 4);
  // fragment reads: lane reads row lane % 32 of a 32-row block, K half lane / 32
  const int fr = lane & 31;
  const int fo = fr * 32 + (((lane >> 5) ^ sw16(fr)) << 4);
  const int oa = wm * G::WM * 32, ob = 2 * BM * 32 + wn * G::WN * 32;

  // the block's row and column factors, fetched now and parked in LDS for
  // the epilogue once the stages are free
  const float2* eh = esc + (size_t)b * 2 * N;
  float2 fct[G::FPT];
#pragma unroll
  for (int u = 0; u < G::FPT; ++u) {
    const int i = tid + u * G::NT;
    fct[u] = i < BM ? eh[m0 + i] : (i < BM + G::BN ? eh[N + n0 + i - BM] : make_float2(0.f, 0.f));
  }

  f16v acc[MI][NI];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

  // Stage q % DEPTH holds chunk q.  The fragments of chunk kt + 1 are read
  // out of LDS while chunk kt is multiplied from registers, and chunks kt + 2
  // .. kt + DEPTH - 1 are in flight: at step kt every wave waits for its own
  // loads of chunk kt + 1, the barrier publishes them and frees the stage of
  // chunk kt - 1 (its fragments were read at step kt - 2 and multiplied at
  // step kt - 1), which receives chunk kt + DEPTH - 1.  Loads past the last
  // chunk re-read it (uniform count of outstanding loads, no branch) into
  // stages whose fragments are not used.
  // K columns [kbeg, Kp): chunks kt0 .. kt0 + nc - 1 (kbeg: polynomial block)
  const int kt0 = min(max((int)sparam[(size_t)b * PA_SCREEN_PARAMS + 4] / 16, 0), Kp / 16 - 1);
  const int nc = Kp / 16 - kt0;
#pragma unroll
  for (int q = 0; q < DEPTH - 1; ++q)
    split_issue<G, N>(lbase + q * G::STAGE, ga, gb, plane, lofs, wave, kt0 + min(q, nc - 1));
  // The two fragment sets alternate (the loop is unrolled by two), and
  // scheduling barriers keep each step's LDS reads ahead of its MFMAs.
  auto kstep = [&](int kt, const SplitFrags<G>& cur, SplitFrags<G>& nxt) {
    __builtin_amdgcn_s_waitcnt(vmcnt_imm((DEPTH - 3) * G::SEG_PER_WAVE));
    __syncthreads();
    split_issue<G, N>(lbase + ((kt + DEPTH - 1) % DEPTH) * G::STAGE, ga, gb, plane, lofs, wave,
                      kt0 + min(kt + DEPTH - 1, nc - 1));
    if (kt + 1 < nc) {
      const char* st = lds + ((kt + 1) % DEPTH) * G::STAGE;
      split_frags<G>(nxt, st + oa, st + ob, fo);
    }
    __builtin_amdgcn_sched_barrier(0);
    split_mma<G, MODE>(acc, cur);
    __builtin_amdgcn_sched_barrier(0);
  };
  SplitFrags<G> fa, fb;
  __builtin_amdgcn_s_waitcnt(vmcnt_imm((DEPTH - 2) * G::SEG_PER_WAVE));
  __syncthreads();
  split_frags<G>(fa, lds + oa, lds + ob, fo);
  int kt = 0;
  for (; kt + 1 < nc; kt += 2) {
    kstep(kt, fa, fb);
    kstep(kt + 1, fb, fa);
  }
  if (kt < nc) kstep(kt, fa, fb);
  // no DMA may still be writing this block's LDS (and every wave is done
  // reading the stages) before the factors go in
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
  __syncthreads();
  float2* sf = reinterpret_cast<float2*>(lds);
  float* sa = reinterpret_cast<float*>(sf + BM + G::BN);
#pragma unroll
  for (int u = 0; u < G::FPT; ++u)
    if (tid + u * G::NT < BM + G::BN) {
      sf[tid + u * G::NT] = fct[u];
      sa[tid + u * G::NT] = -atan2f(fct[u].y, fct[u].x);
    }
  __syncthreads();
  gemm_epilogue<N, MI, NI>(acc, m0, n0, wm * G::WM, wn * G::WN, lane, (float)sparam[(size_t)b * PA_SCREEN_PARAMS + 3],
                           phi2 ? phi2 + (size_t)b * N * N : nullptr, ang ? ang + (size_t)b * N * N : nullptr,
                           fac ? fac + (size_t)b * N * N : nullptr, sf, sf + BM, sa, sa + BM);
}

// ---- float32 GEMM (v_mfma_f32_32x32x2f32): the full-precision reference mode
template <int N>
struct F32Geo {
  static constexpr int BM = N < 128 ? N : 128, BN = BM;
  static constexpr int NTM = N / BM, NTN = N / BN;
};

template <int N>
__global__ void __launch_bounds__(256)
k_screen_gemm_f32(const void* __restrict__ opa, const void* __restrict__ opb,
                  const double* __restrict__ sparam, const float2* __restrict__ esc, float* __restrict__ phi2,
                  float* __restrict__ ang, float2* __restrict__ fac, int Kp, const int* __restrict__ nsteps,
                  int step, int batch) {
  constexpr int BM = F32Geo<N>::BM, BN = F32Geo<N>::BN;
  constexpr int WM = BM / 2, WN = BN / 2, MI = WM / 32, NI = WN / 32;
  constexpr int BK = 32;
  constexpr int LDSROW = BK * 4 + 16;          // padded LDS row (bytes)
  constexpr int CH_ROW = BK * 4 / 16;          // 16-byte chunks per row
  constexpr int CPA = BM * CH_ROW / 256, CPB = BN * CH_ROW / 256;
  static_assert(CPA >= 1 && CPB >= 1 && (BM * CH_ROW) % 256 == 0, "chunking");
  constexpr int STA = BM * LDSROW, STB = BN * LDSROW;

  int b, ty, tx;
  gemm_tile<F32Geo<N>::NTM, F32Geo<N>::NTN>(blockIdx.x, F32Geo<N>::NTM * F32Geo<N>::NTN * batch, b, ty, tx);
  if (nsteps && step >= nsteps[b]) return;
  __shared__ __attribute__((aligned(16))) char sA[2 * STA];
  __shared__ __attribute__((aligned(16))) char sB[2 * STB];

  const int m0 = ty * BM, n0 = tx * BN;
  const size_t rowb = (size_t)Kp * 4;
  const char* ga = reinterpret_cast<const char*>(opa) + (size_t)b * N * rowb + m0 * rowb;
  const char* gb = reinterpret_cast<const char*>(opb) + (size_t)b * N * rowb + n0 * rowb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = tid / CH_ROW, kc = tid % CH_ROW;
  constexpr int RPI = 256 / CH_ROW;

  u4v ra[CPA], rb[CPB];
  auto gload = [&](size_t koff) {
#pragma unroll
    for (int i = 0; i < CPA; ++i)
      ra[i] = *reinterpret_cast<const u4v*>(ga + (lr + RPI * i) * rowb + kc * 16 + koff);
#pragma unroll
    for (int i = 0; i < CPB; ++i)
      rb[i] = *reinterpret_cast<const u4v*>(gb + (lr + RPI * i) * rowb + kc * 16 + koff);
  };
  auto sstore = [&](char* a, char* bb) {
#pragma unroll
    for (int i = 0; i < CPA; ++i) *reinterpret_cast<u4v*>(a + (lr + RPI * i) * LDSROW + kc * 16) = ra[i];
#pragma unroll
    for (int i = 0; i < CPB; ++i) *reinterpret_cast<u4v*>(bb + (lr + RPI * i) * LDSROW + kc * 16) = rb[i];
  };

  f16v acc[MI][NI];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

  // K columns [kbeg, Kp) (kbeg: polynomial block, a multiple of 32)
  const int kt0 = min(max((int)sparam[(size_t)b * PA_SCREEN_PARAMS + 4] / BK, 0), Kp / BK - 1);
  gload((size_t)kt0 * BK * 4);
  sstore(sA, sB);
  __syncthreads();
  const int nk = Kp / BK;
  for (int kt = kt0; kt < nk; ++kt) {
    const int cur = (kt - kt0) & 1;
    gload((size_t)min(kt + 1, nk - 1) * BK * 4);
    __builtin_amdgcn_sched_barrier(0);   // keep the prefetch ahead of the MFMA work
    const char* cA = sA + cur * STA;
    const char* cB = sB + cur * STB;
#pragma unroll
    for (int ks = 0; ks < BK / 2; ++ks) {
      float a[MI], bb[NI];
#pragma unroll
      for (int i = 0; i < MI; ++i)
        a[i] = *reinterpret_cast<const float*>(cA + (wm * WM + i * 32 + (lane & 31)) * LDSROW +
                                               (2 * ks + (lane >> 5)) * 4);
#pragma unroll
      for (int j = 0; j < NI; ++j)
        bb[j] = *reinterpret_cast<const float*>(cB + (wn * WN + j * 32 + (lane & 31)) * LDSROW +
                                                (2 * ks + (lane >> 5)) * 4);
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], bb[j], acc[i][j], 0, 0, 0);
    }
    sstore(sA + (1 - cur) * STA, sB + (1 - cur) * STB);
    __syncthreads();
  }

  // the stages are free after the last barrier: park the factors there
  const float2* eh = esc + (size_t)b * 2 * N;
  float2* sf = reinterpret_cast<float2*>(sA);
  float* sa = reinterpret_cast<float*>(sf + BM + BN);
  if (tid < BM + BN) {
    const float2 e = tid < BM ? eh[m0 + tid] : eh[N + n0 + tid - BM];
    sf[tid] = e;
    sa[tid] = -atan2f(e.y, e.x);
  }
  __syncthreads();
  gemm_epilogue<N, MI, NI>(acc, m0, n0, wm * WM, wn * WN, lane, (float)sparam[(size_t)b * PA_SCREEN_PARAMS + 3],
                           phi2 ? phi2 + (size_t)b * N * N : nullptr, ang ? ang + (size_t)b * N * N : nullptr,
                           fac ? fac + (size_t)b * N * N : nullptr, sf, sf + BM, sa, sa + BM);
}

// ------------------------------------------------------------------ compose
template <int N>
__global__ void __launch_bounds__(256)
k_screen_compose(const double* __restrict__ sparam, const double* __restrict__ hg,
                 const float* __restrict__ phi2, float* __restrict__ phi) {
  const int b = blockIdx.y;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int r = (int)(idx / N), c = (int)(idx % N);
  const double v = sparam[(size_t)b * PA_SCREEN_PARAMS] + hg[(size_t)b * 2 * N + r] +
                   hg[((size_t)b * 2 + 1) * N + c] + (double)phi2[(size_t)b * N * N + tiled<N>(r, c)];
  phi[(size_t)b * N * N + idx] = (float)v;
}

// ------------------------------------------------------------------ launchers
int operand_bytes(int n, int points, int mode, long long* out) {
  if (mode < 0 || mode > 2) return set_error(PA_ERR_ARG, "unknown GEMM mode");
  const long long planes = mode == PA_GEMM_F32 ? 1 : 2;
  const long long es = mode == PA_GEMM_F32 ? 4 : 2;
  *out = planes * es * (long long)n * kpad(points);
  return PA_OK;
}

int spectral_draw(int P, const float* base, const float* sqrt_psd, const int* chan, const uint64_t* rid,
                  uint64_t seed, int screen, float* rho, float* theta, float2* value, int batch,
                  hipStream_t s) {
  dim3 g((P + 255) / 256, batch);
  hipLaunchKernelGGL(k_spectral_draw, g, dim3(256), 0, s, P, base, sqrt_psd, chan, rid, seed, screen,
                     rho, theta, value);
  return check_launch("spectral_draw");
}

int screen_draw_prep(int P, const float* base, const float* sqrt_psd, const int* chan, const uint64_t* rid,
                     uint64_t seed, int screen, double* fxy, double* sparam, int mode, const double* chparam,
                     const int* nsteps, int step, int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_screen_draw_prep, dim3(batch), dim3(kDrawPrepThreads), 0, s, P, base, sqrt_psd, rid, seed,
                     screen, reinterpret_cast<double2*>(fxy), sparam, mode, chparam, chan, nsteps, step);
  return check_launch("screen_draw_prep");
}

int screen_prep(int P, const float* rho, const float* theta, const float2* value, double* fxy,
                double* sparam, int mode, const double* chparam, const int* chan, const int* nsteps, int step,
                int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_screen_prep, dim3(batch), dim3(256), 0, s, P, rho, theta, value,
                     reinterpret_cast<double2*>(fxy), sparam, mode, chparam, chan, nsteps, step);
  return check_launch("screen_prep");
}

template <int MODE, int TC, int RG>
static void tables_launch(int n, int P, double* fxy, double* sparam,
                          const double* chparam, const int* chan, void* opa, void* opb, double* hg, float2* esc,
                          int nc_max, const int* nsteps, int step, int batch, hipStream_t s) {
  const int Kp = kpad(P);
  // the polynomial block's partial sums, then the column-smooth points' P_km
  // (nc_max per element), go to the B operand buffer, which
  // k_screen_tables<.., 1, ..> overwrites after k_screen_lowrank_sum and
  // k_screen_tables<.., 0, ..> have read them
  const int chunks = lowrank_chunks(P);
  double* part = reinterpret_cast<double*>(opb);
  double2* pcoef = reinterpret_cast<double2*>(part + (size_t)batch * chunks * kLRSums);
  double2* f2 = reinterpret_cast<double2*>(fxy);
  hipLaunchKernelGGL(k_screen_classify, dim3(batch), dim3(kClassThreads), 0, s, n, P, f2, chparam, chan, sparam,
                     nc_max, nsteps, step);
  hipLaunchKernelGGL(k_screen_lowrank_part, dim3(batch, chunks), dim3(kLRThreads), 0, s, n, P, f2, chparam, chan,
                     sparam, part, pcoef, nc_max, nsteps, step);
  hipLaunchKernelGGL(k_screen_lowrank_sum, dim3(batch), dim3(1024), 0, s, P, f2, sparam, part, pcoef, nc_max,
                     chunks, MODE, nsteps, step);
  dim3 g(n / (TC * RG), batch);
  hipLaunchKernelGGL((k_screen_tables<MODE, 0, TC, RG>), g, dim3(256), 0, s, n, P, Kp, f2, sparam, chparam, chan,
                     opa, hg, esc, pcoef, nc_max, nsteps, step);
  hipLaunchKernelGGL((k_screen_tables<MODE, 1, TC, RG>), g, dim3(256), 0, s, n, P, Kp, f2, sparam, chparam, chan,
                     opb, hg, esc, pcoef, nc_max, nsteps, step);
}

// rows per table block: 32 when that still gives >= 4 blocks per CU, else 8
// or 4 (measured at 2048^2: 32 rows 0.14 ms per 32 screens, 8 rows 0.16)
static int tables_rows_default(int n, int batch) {
  const long long blocks32 = (long long)(n / 32) * batch, blocks8 = (long long)(n / 8) * batch;
  return blocks32 >= 1024 ? 32 : (blocks8 >= 1024 ? 8 : 4);
}

template <int MODE>
static int tables_mode(int n, int P, double* fxy, double* sparam,
                       const double* chparam, const int* chan, void* opa, void* opb, double* hg,
                       float2* esc, int rows, int nc_max, const int* nsteps, int step, int batch, hipStream_t s) {
  if (rows == 0) rows = tables_rows_default(n, batch);
#define TL_(TC, RG) tables_launch<MODE, TC, RG>(n, P, fxy, sparam, chparam, chan, opa, opb, hg, esc, nc_max, \
                                                nsteps, step, batch, s)
  switch (rows) {
    case 4: TL_(4, 1); break;
    case 8: TL_(8, 1); break;
    case 16: TL_(8, 2); break;
    case 32: TL_(8, 4); break;
    default: return set_error(PA_ERR_ARG, "screen_tables: rows per block must be 0, 4, 8, 16 or 32");
  }
#undef TL_
  return check_launch("screen_tables");
}

int screen_tables(int n, int P, double* fxy, double* sparam,
                  const double* chparam, const int* chan, void* opa, void* opb, double* hg,
                  float2* esc, int mode, int rows, const int* nsteps, int step, int batch, hipStream_t s) {
  long long opbytes = 0;
  if (const int rc = operand_bytes(n, P, mode, &opbytes)) return rc;
  // (below 32 points there is no polynomial block and nothing is written)
  const long long partial = (long long)lowrank_chunks(P) * kLRSums * (long long)sizeof(double);
  if (P >= 32 && partial > opbytes)
    return set_error(PA_ERR_SIZE, "screen_tables: B operand too small for the polynomial block's partial sums");
  // column-smooth points per element: as many P_km rows as the B operand's
  // room after the partial sums holds, at most kColLowMax
  const int nc_max = (int)std::max(0LL, std::min<long long>(kColLowMax, (opbytes - partial) / (kLR * 16)));
  switch (mode) {
    case PA_GEMM_F16X3:
      return tables_mode<PA_GEMM_F16X3>(n, P, fxy, sparam, chparam, chan, opa, opb, hg, esc, rows, nc_max,
                                        nsteps, step, batch, s);
    case PA_GEMM_BF16X3:
      return tables_mode<PA_GEMM_BF16X3>(n, P, fxy, sparam, chparam, chan, opa, opb, hg, esc, rows, nc_max,
                                         nsteps, step, batch, s);
    case PA_GEMM_F32:
      return tables_mode<PA_GEMM_F32>(n, P, fxy, sparam, chparam, chan, opa, opb, hg, esc, rows, nc_max,
                                      nsteps, step, batch, s);
    default:
      return set_error(PA_ERR_ARG, "unknown GEMM mode");
  }
}

template <int N, int MODE, int CFG>
static void gemm_split_launch(int Kp, const void* opa, const void* opb, const double* sparam, const float2* esc,
                              float* phi2, float* ang, float2* fac, const int* nsteps, int step, int batch, hipStream_t s) {
  using G = SplitGeo<N, CFG>;
  hipLaunchKernelGGL((k_screen_gemm_split<N, MODE, CFG>), dim3(G::NTM * G::NTN * batch), dim3(G::NT), 0, s, opa,
                     opb, sparam, esc, phi2, ang, fac, Kp, nsteps, step, batch);
}

// GEMM tile: the 256 x 256 tiles when they give at least two blocks per CU
static int gemm_tile_default(int n, int batch) {
  return (n >= 512 && (long long)(n / 256) * (n / 256) * batch >= 512) ? 1 : 2;
}

template <int N>
static int gemm_n(int P, const void* opa, const void* opb, const double* sparam, const float2* esc, float* phi2,
                  float* ang, float2* fac, int mode, int tile, const int* nsteps, int step, int batch, hipStream_t s) {
  const int Kp = kpad(P);
  if (tile == 0) tile = gemm_tile_default(N, batch);
  if (tile != 1 && tile != 2) return set_error(PA_ERR_ARG, "screen_gemm: tile must be 0, 1 or 2");
  if (N < 512) tile = 2;   // one geometry below 512
  switch (mode) {
    case PA_GEMM_F16X3:
      if (tile == 1) gemm_split_launch<N, PA_GEMM_F16X3, 1>(Kp, opa, opb, sparam, esc, phi2, ang, fac, nsteps, step, batch, s);
      else gemm_split_launch<N, PA_GEMM_F16X3, 2>(Kp, opa, opb, sparam, esc, phi2, ang, fac, nsteps, step, batch, s);
      break;
    case PA_GEMM_BF16X3:
      if (tile == 1) gemm_split_launch<N, PA_GEMM_BF16X3, 1>(Kp, opa, opb, sparam, esc, phi2, ang, fac, nsteps, step, batch, s);
      else gemm_split_launch<N, PA_GEMM_BF16X3, 2>(Kp, opa, opb, sparam, esc, phi2, ang, fac, nsteps, step, batch, s);
      break;
    case PA_GEMM_F32: {
      const dim3 gf(F32Geo<N>::NTM * F32Geo<N>::NTN * batch);
      hipLaunchKernelGGL((k_screen_gemm_f32<N>), gf, dim3(256), 0, s, opa, opb, sparam, esc, phi2, ang, fac, Kp, nsteps,
                         step, batch);
      break;
    }
    default:
      return set_error(PA_ERR_ARG, "unknown GEMM mode");
  }
  return check_launch("screen_gemm");
}

template <int N>
static int compose_n(const double* sparam, const double* hg, const float* phi2, float* phi, int batch,
                     hipStream_t s) {
  dim3 g((N * N) / 256, batch);
  hipLaunchKernelGGL(k_screen_compose<N>, g, dim3(256), 0, s, sparam, hg, phi2, phi);
  return check_launch("screen_compose");
}

#define PA_DISPATCH_N(n, CALL)      \
  switch (n) {                      \
    case 64: return CALL(64);       \
    case 128: return CALL(128);     \
    case 256: return CALL(256);     \
    case 512: return CALL(512);     \
    case 1024: return CALL(1024);   \
    case 2048: return CALL(2048);   \
    case 4096: return CALL(4096);   \
    default: return set_error(PA_ERR_SIZE, "grid size must be a power of two in [64, 4096]"); \
  }

int screen_gemm(int n, int P, const void* opa, const void* opb, const double* sparam, const float2* esc,
                float* phi2, float* ang, float2* fac, int mode, int tile, const int* nsteps, int step, int batch,
                hipStream_t s) {
#define C_(NN) gemm_n<NN>(P, opa, opb, sparam, esc, phi2, ang, fac, mode, tile, nsteps, step, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}

int screen_compose(int n, const double* sparam, const double* hg, const float* phi2, float* phi,
                   int batch, hipStream_t s) {
#define C_(NN) compose_n<NN>(sparam, hg, phi2, phi, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}

}  // namespace pa
