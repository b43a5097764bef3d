// Fused split-step passes over batches of N x N complex64 fields.
//
// A realization with S screens runs as
//   row_first  : source (or input field) -> forward row FFT           -> tiled work
//   col_step(i): forward column FFT -> * H_i -> inverse column FFT    (i = 0..S)
//   row_mid(i) : inverse row FFT -> * e^{-i phi_i} -> forward row FFT (i = 0..S-1)
//   row_last   : inverse row FFT -> row-major field and/or moment partial sums
// which is ifft2(H fft2(.)) per Fresnel step (theory/vacuum.py:5-7 with the
// shifts of utils.py:42-50 cancelled) interleaved with the screen multiply of
// pathes.py:72-73.  H_i = s_i h_i[r] h_i[c] is separable, s_i = e^{ikd}/N^2.
// The screen factor eh[r] eg[c] e^{-i phi2[r,c]} is formed by the screen GEMM's
// epilogue (pa_screen.hip) and stored as complex64 in the work layout.
#include "pa_fft.h"
#include "pa_internal.h"

// Minimum waves per SIMD requested for the fused row pass (4: two blocks per CU).
#ifndef PA_ROWMID_WAVES
#define PA_ROWMID_WAVES 4
#endif

namespace pa {

// Thread -> (line, lane) maps. Row passes: line = row inside the tile, the
// line index runs fastest so that a wave touches one contiguous run of the
// tile. Column passes: (row-in-tile, column, row-tile) order for the same.
template <int N>
__device__ __forceinline__ void row_map(int tid, int& t, int& l) {
  constexpr int T = Geo<N>::T;
  t = tid % T;
  l = tid / T;
}
template <int N>
__device__ __forceinline__ void col_map(int tid, int& cc, int& l) {
  constexpr int T = Geo<N>::T, C = Geo<N>::C;
  const int lo = tid % T;
  cc = (tid / T) % C;
  const int hi = tid / (T * C);
  l = lo + T * hi;
}

// ------------------------------------------------------------------ row_first
// work[b] <- FFT_rows(u0) where u0 = in[b] (row-major) or src_y[r] * src_x[c].
template <int N>
__global__ void __launch_bounds__(Geo<N>::ROW_THREADS)
k_row_first(float2* __restrict__ work, const float2* __restrict__ in,
            const float2* __restrict__ src, const int* __restrict__ chan,
            const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T;
  __shared__ float2 lds[T * Geo<N>::LINE_ROW];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  load_twiddles<N>(s_tw, tw, threadIdx.x, Geo<N>::ROW_THREADS);
  int t, l;
  row_map<N>(threadIdx.x, t, l);
  const int b = blockIdx.y;
  const int r = blockIdx.x * T + t;
  float2 v[E];
  if (in) {
    const float2* row = in + (size_t)b * N * N + (size_t)r * N;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = row[l + L * e];
  } else {
    const int ch = chan ? chan[b] : 0;
    const float2* sy = src + (size_t)ch * 2 * N;
    const float2* sx = sy + N;
    const float2 gy = sy[r];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cmul(gy, sx[l + L * e]);
  }
  fft_line<N, -1>(v, lds + t * Geo<N>::LINE_ROW, l, s_tw);
  float2* out = work + (size_t)b * N * N;
#pragma unroll
  for (int e = 0; e < E; ++e) out[tiled<N>(r, l + L * e)] = v[e];
}

// ------------------------------------------------------------------ col_step
// Column FFT, transfer function, inverse column FFT, in place on work[b].
// htab layout: [channel][step][2][N] (row factor, then column factor holding s).
template <int N>
__global__ void __launch_bounds__(Geo<N>::COL_THREADS)
k_col_step(float2* __restrict__ work, const float2* __restrict__ htab,
           const int* __restrict__ chan, const int* __restrict__ nsteps, int step, int max_steps,
           const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, C = Geo<N>::C;
  __shared__ float2 lds[C * Geo<N>::LINE_COL];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  const int b = blockIdx.y;
  if (nsteps && step >= nsteps[b] + 1) return;  // this element has no step `step`
  load_twiddles<N>(s_tw, tw, threadIdx.x, Geo<N>::COL_THREADS);
  int cc, l;
  col_map<N>(threadIdx.x, cc, l);
  const int c = blockIdx.x * C + cc;
  float2* f = work + (size_t)b * N * N;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = f[tiled<N>(l + L * e, c)];
  // the transfer factors are fetched with the field, ahead of the transform
  const int ch = chan ? chan[b] : 0;
  const float2* hr = htab + ((size_t)ch * max_steps + step) * 2 * N;
  const float2 hc = hr[N + c];
  float2 h[E];
#pragma unroll
  for (int e = 0; e < E; ++e) h[e] = hr[l + L * e];
  float2* line = lds + cc * Geo<N>::LINE_COL;
  fft_line<N, -1>(v, line, l, s_tw);
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = cmul(v[e], cmul(h[e], hc));
  fft_line<N, +1>(v, line, l, s_tw);
  asm volatile("" : "+v"(l));  // recompute the store addresses instead of keeping them live
#pragma unroll
  for (int e = 0; e < E; ++e) f[tiled<N>(l + L * e, c)] = v[e];
}

// ------------------------------------------------------------------ row_mid
// Inverse row FFT, screen multiply, forward row FFT, in place on work[b].
// The screen factor e^{-i ang(r,c)} is formed here from the float32 screen
// angle ang[b] (work layout; pa_screen_gemm: phi2 plus the row and column
// angles reduced to [-pi, pi]).
template <int N>
__global__ void __launch_bounds__(Geo<N>::ROW_THREADS, PA_ROWMID_WAVES)
k_row_mid(float2* __restrict__ work, const float* __restrict__ ang, const int* __restrict__ nsteps, int step,
          const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T;
  __shared__ float2 lds[T * Geo<N>::LINE_ROW];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  const int b = blockIdx.y;
  if (nsteps && step >= nsteps[b]) return;  // no screen `step` for this element
  load_twiddles<N>(s_tw, tw, threadIdx.x, Geo<N>::ROW_THREADS);
  int t, l;
  row_map<N>(threadIdx.x, t, l);
  const int r = blockIdx.x * T + t;
  float2* f = work + (size_t)b * N * N;
  const float* an = ang + (size_t)b * N * N;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = f[tiled<N>(r, l + L * e)];
  float2* line = lds + t * Geo<N>::LINE_ROW;
  fft_line<N, +1>(v, line, l, s_tw);
  asm volatile("" : "+v"(l));
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = cmul(v[e], expmi_fast(an[tiled<N>(r, l + L * e)]));
  fft_line<N, -1>(v, line, l, s_tw);
  asm volatile("" : "+v"(l));  // recompute the store addresses instead of keeping them live
#pragma unroll
  for (int e = 0; e < E; ++e) f[tiled<N>(r, l + L * e)] = v[e];
}

// ------------------------------------------------------------------ row_last
// Inverse row FFT; writes the row-major field (out != null) and/or per-tile
// partial sums of the intensity moments (part != null):
//   part[b][tile][k], k = 0: sum I, 1: sum I x, 2: sum I x^2, 3: sum I y,
//   4: sum I x y, 5: sum I y^2, 6 + p: sum of I inside pupil p
// with x = (c - N/2) delta, y = (r - N/2) delta in float64.
// pup layout: [channel][npup][3] = (radius, shift_x, shift_y) float64.
template <int N>
__global__ void __launch_bounds__(Geo<N>::ROW_THREADS)
k_row_last(const float2* __restrict__ work, float2* __restrict__ out, double* __restrict__ part,
           const double* __restrict__ chparam, const double* __restrict__ pup, int npup,
           const int* __restrict__ chan, const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T;
  __shared__ float2 lds[T * Geo<N>::LINE_ROW];
  __shared__ double red[Geo<N>::ROW_THREADS / 64][kMaxSums];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  load_twiddles<N>(s_tw, tw, threadIdx.x, Geo<N>::ROW_THREADS);
  int t, l;
  row_map<N>(threadIdx.x, t, l);
  const int b = blockIdx.y;
  const int r = blockIdx.x * T + t;
  __shared__ int2 s_pc[T * kMaxPupils];
  static_assert(Geo<N>::ROW_THREADS >= T * kMaxPupils, "one lane per (row, pupil)");
  // the channel's spacing is fetched before the transform, so that its
  // (dependent) load does not stall the reduction after it
  const int ch = chan ? chan[b] : 0;
  const double delta = chparam[ch * kChParams + 0];
  const float2* f = work + (size_t)b * N * N;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = f[tiled<N>(r, l + L * e)];
  // While the field loads are in flight, T * npup lanes find the column
  // interval of each pupil in each of the block's rows (visible to every
  // thread after the transform's barriers).
  if (part && threadIdx.x < T * npup) {
    const int tt = threadIdx.x / npup, p = threadIdx.x - tt * npup;
    const double* pp = pup + ((size_t)ch * kMaxPupils + p) * 3;
    const double a = pp[0];
    int lo, hi;
    circle_cols(N, delta, grid_offset((int)blockIdx.x * T + tt, N, delta, -pp[2]), pp[1], a * a, lo, hi);
    s_pc[tt * kMaxPupils + p] = make_int2(lo, hi);
  }
  fft_line<N, +1>(v, lds + t * Geo<N>::LINE_ROW, l, s_tw);
  if (out) {
    float2* o = out + (size_t)b * N * N + (size_t)r * N;
#pragma unroll
    for (int e = 0; e < E; ++e) o[l + L * e] = v[e];
  }
  if (!part) return;
  const double y = (double)(r - N / 2) * delta;
  // each pupil covers one interval of this row's columns
  int plo[kMaxPupils], phi[kMaxPupils];
#pragma unroll
  for (int p = 0; p < kMaxPupils; ++p) {
    const int2 q = p < npup ? s_pc[t * kMaxPupils + p] : make_int2(0, -1);
    plo[p] = q.x;
    phi[p] = q.y;
  }
  double s0 = 0, s1 = 0, s2 = 0;
  double sp[kMaxPupils];
#pragma unroll
  for (int p = 0; p < kMaxPupils; ++p) sp[p] = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int c = l + L * e;
    const double x = (double)(c - N / 2) * delta;
    const double I = (double)v[e].x * v[e].x + (double)v[e].y * v[e].y;
    s0 += I;
    s1 += I * x;
    s2 += I * x * x;
#pragma unroll
    for (int p = 0; p < kMaxPupils; ++p)
      if (p < npup) sp[p] += (c >= plo[p] && c <= phi[p]) ? I : 0.0;
  }
  double acc[kMaxSums];
  acc[0] = s0;
  acc[1] = s1;
  acc[2] = s2;
  acc[3] = s0 * y;
  acc[4] = s1 * y;
  acc[5] = s0 * y * y;
#pragma unroll
  for (int p = 0; p < kMaxPupils; ++p) acc[6 + p] = sp[p];
  const int ns = 6 + npup;
  constexpr int NT = Geo<N>::ROW_THREADS, NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* dst = part + ((size_t)b * (N / T) + blockIdx.x) * kMaxSums;
  if constexpr (NT * kMaxSums <= T * Geo<N>::LINE_ROW) {
    // Through LDS (the transform's line buffer, free once every wave is past
    // it): one wave per sum adds the block's per-thread values in a fixed
    // order, then across its lanes.
    __syncthreads();
    double* sr = reinterpret_cast<double*>(lds);
#pragma unroll
    for (int k = 0; k < kMaxSums; ++k)
      if (k < ns) sr[k * NT + threadIdx.x] = acc[k];
    __syncthreads();
    for (int k = wave; k < ns; k += NW) {
      double a = 0;
#pragma unroll
      for (int j = 0; j < NW; ++j) a += sr[k * NT + j * 64 + lane];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
      if (lane == 0) dst[k] = a;
    }
  } else {
    // line buffer too small: across lanes first, then the waves in a fixed order
#pragma unroll
    for (int k = 0; k < kMaxSums; ++k) {
      if (k < ns) {
        double a = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane == 0) red[wave][k] = a;
      }
    }
    __syncthreads();
    if (threadIdx.x < ns) {
      double a = 0;
      for (int w = 0; w < NW; ++w) a += red[w][threadIdx.x];
      dst[threadIdx.x] = a;
    }
  }
}

// ------------------------------------------------------------------ FFT phase screens
// The centred inverse transform of FFTPhaseScreen (phase_screens.py:52-53 with
// utils.py:47-50): phi(r, c) = sum_{ky, kx} cn[ky + N/2][kx + N/2]
// e^{2 pi i (kx (c - N/2) + ky (r - N/2)) / N}.  With N even the two
// ifftshifts are the sign (-1)^(index) on input and output ((-1)^(N/2) = 1 for
// N >= 4), so it is an unnormalised inverse DFT between two sign flips:
// k_fft_rows<N, +1, true> (row-major input, sign, inverse row FFT -> work) then
// k_col_fft<N, +1> (inverse column FFT in place), then k_fft_screen_finish.
template <int N, int DIR>
__global__ void __launch_bounds__(Geo<N>::ROW_THREADS)
k_fft_rows_signed(float2* __restrict__ work, const float2* __restrict__ in, const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T;
  __shared__ float2 lds[T * Geo<N>::LINE_ROW];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  load_twiddles<N>(s_tw, tw, threadIdx.x, Geo<N>::ROW_THREADS);
  int t, l;
  row_map<N>(threadIdx.x, t, l);
  const int b = blockIdx.y;
  const int r = blockIdx.x * T + t;
  const float2* row = in + (size_t)b * N * N + (size_t)r * N;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int c = l + L * e;
    const float2 x = row[c];
    v[e] = ((r + c) & 1) ? make_float2(-x.x, -x.y) : x;
  }
  fft_line<N, DIR>(v, lds + t * Geo<N>::LINE_ROW, l, s_tw);
  float2* out = work + (size_t)b * N * N;
#pragma unroll
  for (int e = 0; e < E; ++e) out[tiled<N>(r, l + L * e)] = v[e];
}

template <int N, int DIR>
__global__ void __launch_bounds__(Geo<N>::COL_THREADS)
k_col_fft(float2* __restrict__ work, const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, C = Geo<N>::C;
  __shared__ float2 lds[C * Geo<N>::LINE_COL];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  load_twiddles<N>(s_tw, tw, threadIdx.x, Geo<N>::COL_THREADS);
  int cc, l;
  col_map<N>(threadIdx.x, cc, l);
  const int c = blockIdx.x * C + cc;
  float2* f = work + (size_t)blockIdx.y * N * N;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = f[tiled<N>(l + L * e, c)];
  fft_line<N, DIR>(v, lds + cc * Geo<N>::LINE_COL, l, s_tw);
#pragma unroll
  for (int e = 0; e < E; ++e) f[tiled<N>(l + L * e, c)] = v[e];
}

// phi = sign * work + subharmonics - mean (phase_screens.py:55-67): the 3 x 3
// plane waves of each subharmonic level enter as nsub separable terms
// sum_j ey[j](r) g[j](c) (float64, built on the host: ey[j](r) =
// e^{2 pi i f_j y_r}, g[j](c) = sum_i cn_ij e^{2 pi i f_i x_c}); mean: the
// screen's mean over the grid (float64 complex, per element).  Outputs
// (nullable): the real part row-major, the imaginary part row-major, the real
// part as a screen angle in the work layout (what pa_row_mid applies).
template <int N>
__global__ void __launch_bounds__(256)
k_fft_screen_finish(const float2* __restrict__ work, const double2* __restrict__ sub, int nsub,
                    const double2* __restrict__ mean, float* __restrict__ phi, float* __restrict__ phi_im,
                    float* __restrict__ ang) {
  const int b = blockIdx.y;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;   // row-major element
  const int r = (int)(idx / N), c = (int)(idx % N);
  const size_t tix = tiled<N>(r, c);
  const float2 w = work[(size_t)b * N * N + tix];
  const double sg = ((r + c) & 1) ? -1.0 : 1.0;
  double re = sg * w.x - mean[b].x, im = sg * w.y - mean[b].y;
  const double2* sb = sub + (size_t)b * 2 * nsub * N;
  for (int j = 0; j < nsub; ++j) {
    const double2 ey = sb[(size_t)j * N + r], g = sb[(size_t)(nsub + j) * N + c];
    re += ey.x * g.x - ey.y * g.y;
    im += ey.x * g.y + ey.y * g.x;
  }
  const size_t o = (size_t)b * N * N;
  if (phi) phi[o + idx] = (float)re;
  if (phi_im) phi_im[o + idx] = (float)im;
  if (ang) ang[o + tix] = (float)re;
}

// ------------------------------------------------------------------ launchers
template <int N>
static int launch_row_first(float2* work, const float2* in, const float2* src, const int* chan,
                            const float2* tw, int batch, hipStream_t s) {
  dim3 g(N / Geo<N>::T, batch);
  hipLaunchKernelGGL(k_row_first<N>, g, dim3(Geo<N>::ROW_THREADS), 0, s, work, in, src, chan, tw);
  return check_launch("row_first");
}
template <int N>
static int launch_col_step(float2* work, const float2* htab, const int* chan, const int* nsteps,
                           int step, int max_steps, const float2* tw, int batch, hipStream_t s) {
  dim3 g(N / Geo<N>::C, batch);
  hipLaunchKernelGGL(k_col_step<N>, g, dim3(Geo<N>::COL_THREADS), 0, s, work, htab, chan, nsteps,
                     step, max_steps, tw);
  return check_launch("col_step");
}
template <int N>
static int launch_row_mid(float2* work, const float* ang, const int* nsteps, int step, const float2* tw,
                          int batch, hipStream_t s) {
  dim3 g(N / Geo<N>::T, batch);
  hipLaunchKernelGGL(k_row_mid<N>, g, dim3(Geo<N>::ROW_THREADS), 0, s, work, ang, nsteps, step, tw);
  return check_launch("row_mid");
}
template <int N>
static int launch_row_last(const float2* work, float2* out, double* part, const double* chparam,
                           const double* pup, int npup, const int* chan, const float2* tw,
                           int batch, hipStream_t s) {
  dim3 g(N / Geo<N>::T, batch);
  hipLaunchKernelGGL(k_row_last<N>, g, dim3(Geo<N>::ROW_THREADS), 0, s, work, out, part, chparam,
                     pup, npup, chan, tw);
  return check_launch("row_last");
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

int row_first(int n, float2* work, const float2* in, const float2* src, const int* chan,
              const float2* tw, int batch, hipStream_t s) {
#define C_(NN) launch_row_first<NN>(work, in, src, chan, tw, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}
int col_step(int n, float2* work, const float2* htab, const int* chan, const int* nsteps, int step,
             int max_steps, const float2* tw, int batch, hipStream_t s) {
#define C_(NN) launch_col_step<NN>(work, htab, chan, nsteps, step, max_steps, tw, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}
int row_mid(int n, float2* work, const float* ang, const int* nsteps, int step, const float2* tw, int batch,
            hipStream_t s) {
#define C_(NN) launch_row_mid<NN>(work, ang, nsteps, step, tw, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}
int row_last(int n, const float2* work, float2* out, double* part, const double* chparam,
             const double* pup, int npup, const int* chan, const float2* tw, int batch,
             hipStream_t s) {
#define C_(NN) launch_row_last<NN>(work, out, part, chparam, pup, npup, chan, tw, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}

template <int N>
static int launch_fft_screen(const float2* coeff, float2* work, const double2* sub, int nsub, const double2* mean,
                             const float2* tw, float* phi, float* phi_im, float* ang, int batch, hipStream_t s) {
  hipLaunchKernelGGL((k_fft_rows_signed<N, +1>), dim3(N / Geo<N>::T, batch), dim3(Geo<N>::ROW_THREADS), 0, s, work,
                     coeff, tw);
  hipLaunchKernelGGL((k_col_fft<N, +1>), dim3(N / Geo<N>::C, batch), dim3(Geo<N>::COL_THREADS), 0, s, work, tw);
  hipLaunchKernelGGL(k_fft_screen_finish<N>, dim3(N * N / 256, batch), dim3(256), 0, s, work, sub, nsub, mean, phi,
                     phi_im, ang);
  return check_launch("fft_screen");
}

int fft_screen(int n, const float2* coeff, float2* work, const double2* sub, int nsub, const double2* mean,
               const float2* tw, float* phi, float* phi_im, float* ang, int batch, hipStream_t s) {
#define C_(NN) launch_fft_screen<NN>(coeff, work, sub, nsub, mean, tw, phi, phi_im, ang, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}

int tile_rows(int n) {
#define C_(NN) Geo<NN>::T
  PA_DISPATCH_N(n, C_)
#undef C_
}

}  // namespace pa
