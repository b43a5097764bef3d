// Fused split-step passes over batches of N x N complex64 fields.
//
// A realization with S screens runs as
//   row_first  : source (or input field) -> forward row FFT           -> tiled work
//   col_step(i): forward column FFT -> * H_i -> inverse column FFT    (i = 0..S)
//   row_mid(i) : inverse row FFT -> * e^{-i phi_i} -> forward row FFT (i = 0..S-1)
//   row_last   : inverse row FFT -> row-major field and/or moment partial sums
// or, from the Gaussian source with S >= 1, source_fresnel -> row_mid_source in
// place of row_first -> col_step(0) -> row_mid(0) (the separable first step).
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
template <int N, int C = Geo<N>::C>
__device__ __forceinline__ void col_map(int tid, int& cc, int& l) {
  constexpr int T = Geo<N>::T;
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
    const rsrc_t ri = make_rsrc(in + (size_t)b * N * N, N * N * 8);
    const int vo = (r * N + l) * 8;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = buf_ld2<kBufNT>(ri, vo, e * L * 8);
  } else {
    const int ch = chan ? chan[b] : 0;
    const float2* sy = src + (size_t)ch * 2 * N;
    const float2* sx = sy + N;
    const float2 gy = sy[r];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cmul(gy, sx[l + L * e]);
  }
  fft_line<N, -1>(v, lds + t * Geo<N>::LINE_ROW, l, s_tw);
  const rsrc_t rw = make_rsrc(work + (size_t)b * N * N, N * N * 8);
  const int wo = ((blockIdx.x * N + l) * T + t) * 8;   // tiled(r, l), r = blockIdx.x T + t
#pragma unroll
  for (int e = 0; e < E; ++e) buf_st2<kBufNT>(v[e], rw, wo, e * L * T * 8);
}

// ------------------------------------------------------------------ separable first step
// The Gaussian source is the outer product src_y[r] src_x[c] and the first
// Fresnel step's transfer function H_0 = h_0[fy] h_0[N + fx] (htab, step 0) is
// separable too, so up to the first screen the field is an outer product:
//   ifft2(H_0 fft2(src_y src_x^T)) = uy ux^T,
//   uy = IFFT(h_0[0, N) FFT(src_y)),  ux = IFFT(h_0[N, 2N) FFT(src_x))
// (unnormalised transforms; the scalar and 1/N^2 sit in the factors).  The
// schedule row_first -> col_step(0) -> row_mid(0) then becomes source_fresnel
// (two 1-D lines per element) -> row_mid_source, which forms the field of the
// first screen plane from the two vectors: the first row pass, the first
// column pass, and the first screen pass's field read and inverse transform
// drop out.  One block per (axis, element): sep[b][0] = uy, sep[b][1] = ux.
template <int N>
__global__ void __launch_bounds__(Geo<N>::L)
k_source_fresnel(float2* __restrict__ sep, const float2* __restrict__ src, const float2* __restrict__ htab,
                 const int* __restrict__ chan, int max_steps, const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L;
  __shared__ float2 lds[Geo<N>::PADN];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  const int axis = blockIdx.x, b = blockIdx.y, l = threadIdx.x;
  load_twiddles<N>(s_tw, tw, l, L);
  const int ch = chan ? chan[b] : 0;
  const float2* sv = src + ((size_t)ch * 2 + axis) * N;                        // src_y, then src_x
  const float2* hv = htab + (size_t)ch * max_steps * 2 * N + (size_t)axis * N;  // h_0[fy], then h_0[fx]
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = sv[l + L * e];
  fft_line<N, -1>(v, lds, l, s_tw);
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = cmul(v[e], hv[l + L * e]);
  fft_line<N, +1>(v, lds, l, s_tw);
  float2* o = sep + ((size_t)b * 2 + axis) * N;
#pragma unroll
  for (int e = 0; e < E; ++e) o[l + L * e] = v[e];
}

// The first screen pass on the separable field: u = uy[r] ux[c], times the
// screen factor exp(-i ang[b](r, c)), forward row FFT -> tiled work.  Every
// element must have a first screen.  Persistent like k_row_mid: the next
// item's angles (HBM) and vectors (L2) are fetched before the current
// transform.
template <int N>
__global__ void __launch_bounds__(Geo<N>::ROW_THREADS, PA_ROWMID_WAVES)
k_row_mid_source(float2* __restrict__ work, const float2* __restrict__ sep, const float* __restrict__ ang,
                 const float2* __restrict__ tw, int items) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T, TILES = N / Geo<N>::T;
  __shared__ float2 lds[T * Geo<N>::LINE_ROW];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  int t, l;
  row_map<N>(threadIdx.x, t, l);
  float2* line = lds + t * Geo<N>::LINE_ROW;
  float2 nx[E], ny = make_float2(0.f, 0.f);
  float na[E];
  auto fetch = [&](int it) {
    const int b = it / TILES, tile = it - b * TILES;
    const float2* uy = sep + (size_t)b * 2 * N;
    ny = uy[tile * T + t];
#pragma unroll
    for (int e = 0; e < E; ++e) nx[e] = uy[N + l + L * e];
    const rsrc_t ra = make_rsrc(ang + (size_t)b * N * N, N * N * 4);
    const int idx = (tile * N + l) * T + t;
#pragma unroll
    for (int e = 0; e < E; ++e) na[e] = buf_ld1<kBufNT>(ra, idx * 4, e * L * T * 4);
  };
  int item = blockIdx.x;
  if (item < items) fetch(item);
  load_twiddles<N>(s_tw, tw, threadIdx.x, Geo<N>::ROW_THREADS);
  for (; item < items; item += gridDim.x) {
    const int b = item / TILES, tile = item - b * TILES;
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cmul(cmul(ny, nx[e]), expmi_fast(na[e]));
    if (item + (int)gridDim.x < items) fetch(item + gridDim.x);
    fft_line<N, -1>(v, line, l, s_tw);
    const rsrc_t rf = make_rsrc(work + (size_t)b * N * N, N * N * 8);
    const int idx = (tile * N + l) * T + t;
#pragma unroll
    for (int e = 0; e < E; ++e) buf_st2<kBufNT>(v[e], rf, idx * 8, e * L * T * 8);
    __syncthreads();   // the line buffer is rewritten by the next item's first pass
  }
}

// ------------------------------------------------------------------ col_step
// Column FFT, transfer function, inverse column FFT, in place on work[b].
// htab layout: [channel][step][2][N] (row factor, then column factor holding s).
// One block per item (element b, group of C columns), items element-major, so
// the blocks the device holds at once cover one element's field; the point
// (l + L e, c) of the tiled field is at the lane offset of e = 0 plus e L N
// points (L is a multiple of T), one buffer load per point with the stride in
// the instruction's SGPR offset.  (Measured: a persistent form that prefetches
// the next item, and non-temporal hints, are both slower for this pass.)
template <int N, int C>
__global__ void __launch_bounds__(C * Geo<N>::L)
k_col_step(float2* __restrict__ work, const float2* __restrict__ htab,
           const int* __restrict__ chan, const int* __restrict__ nsteps, int step, int max_steps,
           const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T, G = N / C;
  static_assert((C * L) % 64 == 0 && C * L <= 1024 && N % C == 0, "column group");
  __shared__ float2 lds[C * Geo<N>::LINE_COL];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  const int item = blockIdx.x;
  const int b = item / G;
  if (nsteps && step >= nsteps[b] + 1) return;  // this element has no step `step`
  int cc, l;
  col_map<N, C>(threadIdx.x, cc, l);
  const int c = (item - b * G) * C + cc;
  const rsrc_t rf = make_rsrc(work + (size_t)b * N * N, N * N * 8);
  const int fo = (((l / T) * N + c) * T + l % T) * 8;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = buf_ld2<0>(rf, fo, e * L * N * 8);
  // the transfer factors (L2-resident) are fetched with the field, ahead of the transform
  const int ch = chan ? chan[b] : 0;
  const float2* hr = htab + ((size_t)ch * max_steps + step) * 2 * N;
  const rsrc_t rh = make_rsrc(hr, 2 * N * 8);
  const float2 hc = buf_ld2<0>(rh, c * 8, N * 8);
  float2 h[E];
#pragma unroll
  for (int e = 0; e < E; ++e) h[e] = buf_ld2<0>(rh, l * 8, e * L * 8);
  load_twiddles<N>(s_tw, tw, threadIdx.x, C * L);
  float2* line = lds + cc * Geo<N>::LINE_COL;
  fft_line<N, -1>(v, line, l, s_tw);
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = cmul(v[e], cmul(h[e], hc));
  fft_line<N, +1>(v, line, l, s_tw);
#pragma unroll
  for (int e = 0; e < E; ++e) buf_st2<0>(v[e], rf, fo, e * L * N * 8);
}

// ------------------------------------------------------------------ row_mid
// Inverse row FFT, screen multiply, forward row FFT, in place on work[b].
// The screen factor e^{-i ang(r,c)} is formed here from the float32 screen
// angle ang[b] (work layout; pa_screen_gemm: phi2 plus the row and column
// angles reduced to [-pi, pi]).  Persistent: a block walks items (element b,
// row tile) with stride gridDim.x (the blocks the device holds at once) and
// issues the next item's field loads before the current transforms, so that
// they stay in flight through them; loads and stores are non-temporal (the
// field streams through once per pass).  The point (tile T + t, l + L e) of
// the tiled field is at (tile N + l) T + t + e L T.
template <int N>
__device__ __forceinline__ void row_load(float2 (&v)[Geo<N>::E], const float2* __restrict__ work, int item, int t,
                                         int l) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T, TILES = N / Geo<N>::T;
  const int b = item / TILES, tile = item - b * TILES;
  const rsrc_t rf = make_rsrc(work + (size_t)b * N * N, N * N * 8);
  const int idx = (tile * N + l) * T + t;
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = buf_ld2<kBufNT>(rf, idx * 8, e * L * T * 8);
}

template <int N>
__global__ void __launch_bounds__(Geo<N>::ROW_THREADS, PA_ROWMID_WAVES)
k_row_mid(float2* __restrict__ work, const float* __restrict__ ang, const int* __restrict__ nsteps, int step,
          const float2* __restrict__ tw, int items) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T, TILES = N / Geo<N>::T;
  __shared__ float2 lds[T * Geo<N>::LINE_ROW];
  __shared__ float2 s_tw[TwTable<N>::SIZE];
  int t, l;
  row_map<N>(threadIdx.x, t, l);
  float2* line = lds + t * Geo<N>::LINE_ROW;
  // elements with fewer screens have no screen `step`
  auto active = [&](int it) { return !nsteps || step < nsteps[it / TILES]; };
  int item = blockIdx.x;
  while (item < items && !active(item)) item += gridDim.x;
  float2 nx[E];
  if (item < items) row_load<N>(nx, work, item, t, l);
  load_twiddles<N>(s_tw, tw, threadIdx.x, Geo<N>::ROW_THREADS);
  while (item < items) {
    const int b = item / TILES, tile = item - b * TILES;
    const int idx = (tile * N + l) * T + t;
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nx[e];
    // this item's angles, used after the first transform
    const rsrc_t ra = make_rsrc(ang + (size_t)b * N * N, N * N * 4);
    float a[E];
#pragma unroll
    for (int e = 0; e < E; ++e) a[e] = buf_ld1<kBufNT>(ra, idx * 4, e * L * T * 4);
    int next = item + gridDim.x;
    while (next < items && !active(next)) next += gridDim.x;
    if (next < items) row_load<N>(nx, work, next, t, l);
    fft_line<N, +1>(v, line, l, s_tw);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cmul(v[e], expmi_fast(a[e]));
    fft_line<N, -1>(v, line, l, s_tw);
    const rsrc_t rf = make_rsrc(work + (size_t)b * N * N, N * N * 8);
#pragma unroll
    for (int e = 0; e < E; ++e) buf_st2<kBufNT>(v[e], rf, idx * 8, e * L * T * 8);
    item = next;
    __syncthreads();   // the line buffer is rewritten by the next item's first pass
  }
}

// ------------------------------------------------------------------ row_last
// Inverse row FFT; writes the row-major field (out != null) and/or per-tile
// partial sums of the intensity moments (part != null):
//   part[b][tile][k], k = 0: sum I, 1: sum I x, 2: sum I x^2, 3: sum I y,
//   4: sum I x y, 5: sum I y^2, 6 + p: sum of I inside pupil p
// with x = (c - N/2) delta, y = (r - N/2) delta in float64.
// pup layout: [channel][npup][3] = (radius, shift_x, shift_y) float64.
// Persistent over items (element b, row tile) with the next item's field
// loads issued before the current transform.
template <int N>
__global__ void __launch_bounds__(Geo<N>::ROW_THREADS, PA_ROWMID_WAVES)
k_row_last(const float2* __restrict__ work, float2* __restrict__ out, double* __restrict__ part,
           const double* __restrict__ chparam, const double* __restrict__ pup, int npup,
           const double* __restrict__ pup_moments, const int* __restrict__ chan, const float2* __restrict__ tw,
           int items) {
  constexpr int E = Geo<N>::E, L = Geo<N>::L, T = Geo<N>::T, TILES = N / Geo<N>::T;
  __shared__ float2 lds[T * Geo<N>::LINE_ROW];
  __shared__ double s_red[(Geo<N>::ROW_THREADS / 64) * Geo<N>::T * (3 + kMaxPupils)];
  __shared__ float2 s_tw[TwTable<N, true>::SIZE];
  __shared__ int2 s_pc[T * kMaxPupils];
  static_assert(Geo<N>::ROW_THREADS >= T * kMaxPupils, "one lane per (row, pupil)");
  constexpr int NT = Geo<N>::ROW_THREADS;
  int t, l;
  row_map<N>(threadIdx.x, t, l);
  // The last T * npup threads find the column interval of each pupil in each
  // row of item `it` (one (row, pupil) each).  They do it for the next item
  // while the block reduces the current one (those threads belong to the last
  // wave, idle in that phase when 6 + npup < NT / 64), so the search is off the
  // critical path; the intervals are read after the next transform's barriers.
  auto find_intervals = [&](int it) {
    const int j = (int)threadIdx.x - (NT - T * npup);
    if (!part || j < 0 || it >= items) return;
    const int bb = it / TILES, tl = it - bb * TILES;
    const int ch = chan ? chan[bb] : 0;
    const double dl = chparam[ch * kChParams + 0];
    const int tt = j / npup, p = j - tt * npup;
    const double* pp = pup + ((size_t)ch * kMaxPupils + p) * 3;
    const double a = pp[0];
    // centroid-tracked pupils take the element's own (mean_x, mean_y)
    const double sx = pup_moments ? pup_moments[(size_t)bb * kMoments + 1] : pp[1];
    const double sy = pup_moments ? pup_moments[(size_t)bb * kMoments + 2] : pp[2];
    int lo, hi;
    circle_cols(N, dl, grid_offset(tl * T + tt, N, dl, -sy), sx, a * a, lo, hi);
    s_pc[tt * kMaxPupils + p] = make_int2(lo, hi);
  };
  int item = blockIdx.x;
  float2 nx[E];
  if (item < items) row_load<N>(nx, work, item, t, l);
  load_twiddles<N, true>(s_tw, tw, threadIdx.x, NT);
  find_intervals(item);
  for (; item < items; item += gridDim.x) {
    const int b = item / TILES, tile = item - b * TILES;
    const int r = tile * T + t;
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nx[e];
    if (item + (int)gridDim.x < items) row_load<N>(nx, work, item + gridDim.x, t, l);
    fft_line<N, +1, true>(v, lds + t * Geo<N>::LINE_ROW, l, s_tw);
    if (out) {
      const rsrc_t ro = make_rsrc(out + (size_t)b * N * N, N * N * 8);
#pragma unroll
      for (int e = 0; e < E; ++e) buf_st2<kBufNT>(v[e], ro, (r * N + l) * 8, e * L * 8);
    }
    if (part) {
      // Per-thread sums of I, x I, x^2 I and each pupil's I over the thread's
      // columns; added over the wave's lanes of the same row (xor shuffles
      // over the lane bits above the row's, T | 64), then over the waves and,
      // with the row's y weights, over the tile's T rows by one thread per
      // partial sum, in a fixed order: one barrier per item.
      constexpr int NW = NT / 64, NR = 3 + kMaxPupils;
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      // chparam may be null when no partial sums are asked for (pa_row_last's contract)
      const double delta = chparam[(chan ? chan[b] : 0) * kChParams + 0];
      // The thread's columns are c = l + L e, x_e = xl + (e - E/2) Ld with
      // xl = l delta, Ld = L delta (L E = N), so the x-moments follow from the
      // sums of I, (e - E/2) I and (e - E/2)^2 I, whose weights are constants.
      // The columns inside a pupil's interval [lo, hi] are one run of e.
      auto e_range = [&](int p, int& elo, int& ehi) {
        const int2 q = s_pc[t * kMaxPupils + p];
        elo = q.x <= l ? 0 : (q.x - l + L - 1) / L;
        ehi = q.y < l ? -1 : (q.y - l) / L;
      };
      auto emit = [&](int k, double a) {
#pragma unroll
        for (int off = 32; off >= T; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane < T) s_red[(wave * T + lane) * NR + k] = a;
      };
      double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        constexpr int H = E / 2;
        const double I = (double)v[e].x * v[e].x + (double)v[e].y * v[e].y;
        s0 += I;
        s1 += (double)(e - H) * I;
        s2 += (double)((e - H) * (e - H)) * I;
      }
      const double xl = (double)l * delta, Ld = (double)L * delta;
      emit(0, s0);
      emit(1, xl * s0 + Ld * s1);
      emit(2, xl * xl * s0 + 2.0 * xl * Ld * s1 + Ld * Ld * s2);
      // Most row tiles miss a pupil altogether: its sum is formed only in
      // waves where some lane's run of e is non-empty.
      for (int p = 0; p < npup; ++p) {
        int elo, ehi;
        e_range(p, elo, ehi);
        double sp = 0;
        if (__ballot(elo <= ehi)) {
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const double I = (double)v[e].x * v[e].x + (double)v[e].y * v[e].y;
            sp += (e >= elo && e <= ehi) ? I : 0.0;
          }
        }
        emit(3 + p, sp);
      }
      __syncthreads();
      find_intervals(item + gridDim.x);
      // partial sum k of the tile: 0 I, 1 I x, 2 I x^2, 3 I y, 4 I x y,
      // 5 I y^2, 6 + p pupil p
      const int k = threadIdx.x;
      if (k < 6 + npup) {
        const int src = k < 3 ? k : k < 6 ? (k == 4 ? 1 : 0) : k - 3;
        const int pw = k < 3 || k >= 6 ? 0 : k == 5 ? 2 : 1;   // power of y
        double a = 0;
        for (int tt = 0; tt < T; ++tt) {
          double rs = 0;
          for (int w = 0; w < NW; ++w) rs += s_red[(w * T + tt) * NR + src];
          const double y = (double)(tile * T + tt - N / 2) * delta;
          a += pw == 0 ? rs : pw == 1 ? rs * y : rs * y * y;
        }
        part[((size_t)b * TILES + tile) * kMaxSums + k] = a;
      }
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
// Grid of a persistent pass: the blocks the device holds at once (occupancy
// times compute units, looked up once per kernel and device), at most `items`.
template <typename K>
static int resident_grid(K kernel, int threads, int items) {
  static int cached[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
  if (!cached[dev]) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), threads, 0) !=
            hipSuccess || per_cu <= 0)
      per_cu = 1;
    (void)hipGetLastError();
    cached[dev] = cus * per_cu;
  }
  return items < cached[dev] ? items : cached[dev];
}

// Columns per column-pass block: the layout's default C, or C / 2 or 2 C where
// the block stays a whole number of waves of at most 1024 threads (a launch
// shape: the results do not depend on it).
constexpr bool col_group_ok(int n, int l, int cc) {
  return cc >= 1 && n % cc == 0 && (cc * l) % 64 == 0 && cc * l <= 1024;
}
template <int N>
static int col_group_choices(int* out) {
  constexpr int C = Geo<N>::C, L = Geo<N>::L;
  int k = 0;
  if (col_group_ok(N, L, C)) out[k++] = C;
  if (col_group_ok(N, L, C / 2) && C / 2 >= 1) out[k++] = C / 2;
  if (col_group_ok(N, L, 2 * C)) out[k++] = 2 * C;
  return k;
}

template <int N, int CC>
static int launch_col_group(float2* work, const float2* htab, const int* chan, const int* nsteps, int step,
                            int max_steps, const float2* tw, int batch, hipStream_t s) {
  if constexpr (col_group_ok(N, Geo<N>::L, CC)) {
    hipLaunchKernelGGL((k_col_step<N, CC>), dim3((N / CC) * batch), dim3(CC * Geo<N>::L), 0, s, work, htab, chan,
                       nsteps, step, max_steps, tw);
    return check_launch("col_step");
  } else {
    return set_error(PA_ERR_ARG, "col_step: unsupported column group for this grid size");
  }
}

template <int N>
static int launch_col_step(float2* work, const float2* htab, const int* chan, const int* nsteps,
                           int step, int max_steps, const float2* tw, int cols, int batch, hipStream_t s) {
  constexpr int C = Geo<N>::C;
  if (cols == 0 || cols == C) return launch_col_group<N, C>(work, htab, chan, nsteps, step, max_steps, tw, batch, s);
  if (C / 2 >= 1 && cols == C / 2)
    return launch_col_group<N, (C / 2 >= 1 ? C / 2 : 1)>(work, htab, chan, nsteps, step, max_steps, tw, batch, s);
  if (cols == 2 * C) return launch_col_group<N, 2 * C>(work, htab, chan, nsteps, step, max_steps, tw, batch, s);
  return set_error(PA_ERR_ARG, "col_step: cols must be 0 or one of pa_col_groups(n)");
}
template <int N>
static int launch_row_mid(float2* work, const float* ang, const int* nsteps, int step, const float2* tw,
                          int batch, hipStream_t s) {
  const int items = (N / Geo<N>::T) * batch;
  const int g = resident_grid(k_row_mid<N>, Geo<N>::ROW_THREADS, items);
  hipLaunchKernelGGL(k_row_mid<N>, dim3(g), dim3(Geo<N>::ROW_THREADS), 0, s, work, ang, nsteps, step, tw, items);
  return check_launch("row_mid");
}
template <int N>
static int launch_source_fresnel(float2* sep, const float2* src, const float2* htab, const int* chan, int max_steps,
                                 const float2* tw, int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_source_fresnel<N>, dim3(2, batch), dim3(Geo<N>::L), 0, s, sep, src, htab, chan, max_steps,
                     tw);
  return check_launch("source_fresnel");
}
template <int N>
static int launch_row_mid_source(float2* work, const float2* sep, const float* ang, const float2* tw, int batch,
                                 hipStream_t s) {
  const int items = (N / Geo<N>::T) * batch;
  const int g = resident_grid(k_row_mid_source<N>, Geo<N>::ROW_THREADS, items);
  hipLaunchKernelGGL(k_row_mid_source<N>, dim3(g), dim3(Geo<N>::ROW_THREADS), 0, s, work, sep, ang, tw, items);
  return check_launch("row_mid_source");
}
template <int N>
static int launch_row_last(const float2* work, float2* out, double* part, const double* chparam,
                           const double* pup, int npup, const double* pup_moments, const int* chan,
                           const float2* tw, int batch, hipStream_t s) {
  const int items = (N / Geo<N>::T) * batch;
  const int g = resident_grid(k_row_last<N>, Geo<N>::ROW_THREADS, items);
  hipLaunchKernelGGL(k_row_last<N>, dim3(g), dim3(Geo<N>::ROW_THREADS), 0, s, work, out, part, chparam,
                     pup, npup, pup_moments, chan, tw, items);
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
             int max_steps, const float2* tw, int cols, int batch, hipStream_t s) {
#define C_(NN) launch_col_step<NN>(work, htab, chan, nsteps, step, max_steps, tw, cols, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}

int col_groups(int n, int* out) {
#define C_(NN) col_group_choices<NN>(out)
  PA_DISPATCH_N(n, C_)
#undef C_
}
int row_mid(int n, float2* work, const float* ang, const int* nsteps, int step, const float2* tw, int batch,
            hipStream_t s) {
#define C_(NN) launch_row_mid<NN>(work, ang, nsteps, step, tw, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}
int source_fresnel(int n, float2* sep, const float2* src, const float2* htab, const int* chan, int max_steps,
                   const float2* tw, int batch, hipStream_t s) {
#define C_(NN) launch_source_fresnel<NN>(sep, src, htab, chan, max_steps, tw, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}
int row_mid_source(int n, float2* work, const float2* sep, const float* ang, const float2* tw, int batch,
                   hipStream_t s) {
#define C_(NN) launch_row_mid_source<NN>(work, sep, ang, tw, batch, s)
  PA_DISPATCH_N(n, C_)
#undef C_
}
int row_last(int n, const float2* work, float2* out, double* part, const double* chparam,
             const double* pup, int npup, const double* pup_moments, const int* chan, const float2* tw,
             int batch, hipStream_t s) {
#define C_(NN) launch_row_last<NN>(work, out, part, chparam, pup, npup, pup_moments, chan, tw, batch, s)
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
