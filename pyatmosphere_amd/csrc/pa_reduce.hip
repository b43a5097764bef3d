// Intensity-moment reductions, pupils and small utility kernels.
// Moments follow measures.py:1-38 (sums of |u|^2 times a coordinate weight,
// times delta^2) and simulations/beam.py:26-33 (mean_x2_r); pupils follow
// pupils.py:8-13 with the mask (x - sx)^2 + (y + sy)^2 <= a^2.
#include <mutex>

#include "pa_fft.h"
#include "pa_internal.h"

namespace pa {

// Per-tile partial sums over row-major fields (same format as k_row_last).
__global__ void __launch_bounds__(256)
k_field_partials(int n, int T, const float2* __restrict__ u, double* __restrict__ part,
                 const double* __restrict__ chparam, const double* __restrict__ pup, int npup,
                 const double* __restrict__ pup_moments, const int* __restrict__ chan) {
  const int b = blockIdx.y, tile = blockIdx.x;
  const int ch = chan ? chan[b] : 0;
  const double delta = chparam[ch * kChParams + 0];
  double rad2[kMaxPupils], psx[kMaxPupils], psy[kMaxPupils];
#pragma unroll
  for (int p = 0; p < kMaxPupils; ++p) {
    const double* pp = pup + ((size_t)ch * kMaxPupils + p) * 3;
    rad2[p] = p < npup ? pp[0] * pp[0] : -1.0;
    psx[p] = p < npup ? pp[1] : 0.0;
    psy[p] = p < npup ? pp[2] : 0.0;
    if (pup_moments && p < npup) {
      psx[p] = pup_moments[(size_t)b * kMoments + 1];
      psy[p] = pup_moments[(size_t)b * kMoments + 2];
    }
  }
  double acc[kMaxSums];
#pragma unroll
  for (int k = 0; k < kMaxSums; ++k) acc[k] = 0;
  const float2* f = u + (size_t)b * n * n;
  const int count = T * n;
  for (int i = threadIdx.x; i < count; i += 256) {
    const int r = tile * T + i / n, c = i % n;
    const float2 v = f[(size_t)r * n + c];
    const double x = (double)(c - n / 2) * delta, y = (double)(r - n / 2) * delta;
    const double I = (double)v.x * v.x + (double)v.y * v.y;
    acc[0] += I;
    acc[1] += I * x;
    acc[2] += I * x * x;
    acc[3] += I * y;
    acc[4] += I * x * y;
    acc[5] += I * y * y;
#pragma unroll
    for (int p = 0; p < kMaxPupils; ++p) {
      if (in_circle(grid_offset(c, n, delta, psx[p]), grid_offset(r, n, delta, -psy[p]), rad2[p])) acc[6 + p] += I;
    }
  }
  __shared__ double red[4][kMaxSums];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kMaxSums; ++k) {
    double a = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) red[wave][k] = a;
  }
  __syncthreads();
  if (threadIdx.x < kMaxSums)
    part[((size_t)b * (n / T) + tile) * kMaxSums + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Sum tiles (fixed-order tree over 256 threads) and form the measures of
// measures.py / beam.py.
__global__ void __launch_bounds__(256)
k_moments_finalize(int ntiles, const double* __restrict__ part, const double* __restrict__ chparam,
                   int npup, const int* __restrict__ chan, double* __restrict__ out) {
  const int b = blockIdx.x;
  __shared__ double red[kMaxSums][256];
  __shared__ double S[kMaxSums];
  const int ns = 6 + npup;
  for (int k = 0; k < ns; ++k) {
    double a = 0;
    for (int t = threadIdx.x; t < ntiles; t += 256) a += part[((size_t)b * ntiles + t) * kMaxSums + k];
    red[k][threadIdx.x] = a;
  }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < ns; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < kMaxSums) S[threadIdx.x] = threadIdx.x < ns ? red[threadIdx.x][0] : 0.0;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int ch = chan ? chan[b] : 0;
  const double delta = chparam[ch * kChParams + 0];
  const double d2 = delta * delta;
  double* o = out + (size_t)b * kMoments;
  const double eta = S[0] * d2, mx = S[1] * d2, my = -S[3] * d2;
  const double mx2 = S[2] * d2, mxy = -S[4] * d2, my2 = S[5] * d2;
  o[0] = eta;
  o[1] = mx;
  o[2] = my;
  o[3] = mx2;
  o[4] = mxy;
  o[5] = my2;
  const double r0 = sqrt(mx * mx + my * my);
  const double cx = mx / r0, sx = my / r0;  // NaN at r0 == 0, as in beam.py:27-30
  o[6] = cx * cx * mx2 + 2.0 * cx * sx * mxy + sx * sx * my2;
  for (int p = 0; p < kMaxPupils; ++p) o[7 + p] = p < npup ? S[6 + p] * d2 : 0.0;
}

__global__ void k_apply_pupil(int n, float2* __restrict__ u, double delta, double rad2, double sx,
                              double sy) {
  const int b = blockIdx.y;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  if (!in_circle(grid_offset(c, n, delta, sx), grid_offset(r, n, delta, -sy), rad2))
    u[(size_t)b * n * n + idx] = make_float2(0.f, 0.f);
}

__global__ void k_expmi(int count, const float* __restrict__ x, float2* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < count) out[i] = expmi_fast(x[i]);
}

int expmi_eval(int count, const float* x, float2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_expmi, dim3((count + 255) / 256), dim3(256), 0, s, count, x, out);
  return check_launch("expmi");
}

__global__ void k_twiddles(int n, float2* __restrict__ tw) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  double s, c;
  sincospi(2.0 * k / n, &s, &c);
  tw[k] = make_float2((float)c, (float)-s);
}

__global__ void k_outer(int n, float2* __restrict__ out, const float2* __restrict__ gy,
                        const float2* __restrict__ gx) {
  const int b = blockIdx.y;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  out[(size_t)b * n * n + idx] = cmul(gy[(size_t)b * n + r], gx[(size_t)b * n + c]);
}

__global__ void k_intensity(size_t count, const float2* __restrict__ u, float* __restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= count) return;
  const float2 v = u[idx];
  out[idx] = v.x * v.x + v.y * v.y;
}

int outer(int n, float2* out, const float2* gy, const float2* gx, int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_outer, dim3((n * n + 255) / 256, batch), dim3(256), 0, s, n, out, gy, gx);
  return check_launch("outer");
}

int intensity(long long count, const float2* u, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_intensity, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, (size_t)count, u, out);
  return check_launch("intensity");
}

int field_partials(int n, const float2* u, double* part, const double* chparam, const double* pup,
                   int npup, const double* pup_moments, const int* chan, int batch, hipStream_t s) {
  const int T = tile_rows(n);
  if (T <= 0) return T;
  hipLaunchKernelGGL(k_field_partials, dim3(n / T, batch), dim3(256), 0, s, n, T, u, part, chparam,
                     pup, npup, pup_moments, chan);
  return check_launch("field_partials");
}

int moments_finalize(int n, const double* part, const double* chparam, int npup, const int* chan,
                     double* out, int batch, hipStream_t s) {
  const int T = tile_rows(n);
  if (T <= 0) return T;
  hipLaunchKernelGGL(k_moments_finalize, dim3(batch), dim3(256), 0, s, n / T, part, chparam, npup,
                     chan, out);
  return check_launch("moments_finalize");
}

// ------------------------------------------------------------------ no-allocation moments
// The per-tile partials of pa_field_moments_static live in this file-scope
// device array instead of a caller's buffer.  Each device that runs the entry
// point holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device (the
// array's name is a host-side stand-in, not a device address).  Every call on
// a device shares that device's array, so calls that overlap on one device
// (different streams, or threads) race; calls in one stream do not.  Nothing
// needs clearing: k_field_partials writes every partial k_moments_finalize
// reads.
constexpr long long kStaticScratchBytes = 8ll << 20;
static __device__ __align__(256) double scratch[kStaticScratchBytes / sizeof(double)];

long long static_scratch_bytes() { return kStaticScratchBytes; }

long long moments_scratch_bytes(int n, int batch) {
  const int T = tile_rows(n);
  if (T <= 0) return T;
  return (long long)batch * (n / T) * kMaxSums * (long long)sizeof(double);
}

static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int static_scratch(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return set_error(PA_ERR_HIP, "field_moments_static: no current HIP device");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_error(PA_ERR_HIP, "field_moments_static: hipGetSymbolAddress of the scratch array failed");
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PA_OK;
}

int field_moments_static(int n, const float2* u, const double* chparam, const double* pup, int npup,
                         const double* pup_moments, const int* chan, double* out, int batch, hipStream_t s) {
  double* part = nullptr;
  int rc = static_scratch(&part);
  if (rc) return rc;
  rc = field_partials(n, u, part, chparam, pup, npup, pup_moments, chan, batch, s);
  if (rc) return rc;
  return moments_finalize(n, part, chparam, npup, chan, out, batch, s);
}

// ------------------------------------------------------------------ ensemble accumulation
// Per-channel Monte Carlo ensembles on the device (simulations/beam.py:35-71;
// simulations/pdt.py:28-48): each realization's finalized moments add to its
// channel's float64 sums of the BeamResult quantities and their squares, and
// each pupil's transmittance to a 64-bin integer histogram over [0, 1].  One
// thread per channel walks the batch in order, so the sums are deterministic
// (no atomics); a channel stops taking beam samples at beam_cap and PDT
// samples at pdt_cap (caps < 0: no limit).  Quantities per sample, in the
// order of PA_ENS_*: eta, mean_x, mean_y, mean_x2, mean_xy, mean_y2,
// mean_x2_r, bw2_x = mean_x^2, bw2_y = mean_y^2, lt2 = 4 mean_x2,
// st2 = lt2 - 4 bw2_x (beam.py:35-47, evaluated with the same roundings).
// Binning follows numpy.histogram(eta, 64, range=(0, 1)): values outside
// [0, 1] (and NaN) are dropped, eta = 1 falls into the last bin, and the bin
// of eta is floor(64 eta) (edges k/64 are exact in float64).
__global__ void __launch_bounds__(256)
k_ensemble_accumulate(int batch, const double* __restrict__ mom, const int* __restrict__ slot_chan, int nchan,
                      int npup, int beam_cap, int pdt_cap, double* __restrict__ acc,
                      long long* __restrict__ counts, unsigned long long* __restrict__ hist) {
  const int ch = blockIdx.x * blockDim.x + threadIdx.x;
  if (ch >= nchan) return;
  double* a = acc + (size_t)ch * 2 * PA_ENS_QUANTITIES;
  long long nb = counts[2 * ch], np_ = counts[2 * ch + 1];
  unsigned long long* h = hist + (size_t)ch * kMaxPupils * PA_PDT_BINS;
  for (int b = 0; b < batch; ++b) {
    if ((slot_chan ? slot_chan[b] : 0) != ch) continue;
    const double* m = mom + (size_t)b * kMoments;
    if (beam_cap < 0 || nb < beam_cap) {
#pragma clang fp contract(off)
      double q[PA_ENS_QUANTITIES];
      for (int k = 0; k < 7; ++k) q[k] = m[k];
      q[7] = m[1] * m[1];
      q[8] = m[2] * m[2];
      q[9] = 4.0 * m[3];
      q[10] = q[9] - 4.0 * q[7];
      for (int k = 0; k < PA_ENS_QUANTITIES; ++k) {
        a[k] += q[k];
        a[PA_ENS_QUANTITIES + k] += q[k] * q[k];
      }
      ++nb;
    }
    if (pdt_cap < 0 || np_ < pdt_cap) {
      for (int p = 0; p < npup; ++p) {
        const double e = m[7 + p];
        if (!(e >= 0.0 && e <= 1.0)) continue;
        int bin = (int)(e * PA_PDT_BINS);
        if (bin == PA_PDT_BINS) bin = PA_PDT_BINS - 1;
        ++h[p * PA_PDT_BINS + bin];
      }
      ++np_;
    }
  }
  counts[2 * ch] = nb;
  counts[2 * ch + 1] = np_;
}

int ensemble_accumulate(int batch, const double* mom, const int* slot_chan, int nchan, int npup, int beam_cap,
                        int pdt_cap, double* acc, long long* counts, unsigned long long* hist, hipStream_t s) {
  hipLaunchKernelGGL(k_ensemble_accumulate, dim3((nchan + 255) / 256), dim3(256), 0, s, batch, mom, slot_chan,
                     nchan, npup, beam_cap, pdt_cap, acc, counts, hist);
  return check_launch("ensemble_accumulate");
}

int apply_pupil(int n, float2* u, double delta, double radius, double sx, double sy, int batch,
                hipStream_t s) {
  hipLaunchKernelGGL(k_apply_pupil, dim3((n * n + 255) / 256, batch), dim3(256), 0, s, n, u, delta,
                     radius * radius, sx, sy);
  return check_launch("apply_pupil");
}

// ------------------------------------------------------------------ structure function
// Row structure function of a screen (theory/phase_screens/sf.py:4-8 averaged
// over rows as in simulations/structure_function.py:11-12):
//   D[r - 1] = (1 / rows) sum_rows (1 / (n - r)) sum_{c < n - r} (phi[c] - phi[c + r])^2.
// k_sf_partials: one block per kSFRows rows of element b, the rows in LDS;
// thread t takes lags r = 1 + t, 1 + t + 256, ...; per (row, lag) the squares
// are summed in float32 over four interleaved accumulators (the reference's own
// arithmetic is float32), then in float64 over the block's rows; the block's
// n - 1 sums go to part[b][group][r - 1].  k_sf_finalize adds the groups in
// order and divides by (n - r) n.
constexpr int kSFRows = 4;

__global__ void __launch_bounds__(256)
k_sf_partials(int n, const float* __restrict__ phi, double* __restrict__ part) {
  extern __shared__ float s_rows[];   // kSFRows x n
  const int b = blockIdx.y, g = blockIdx.x;
  const float* src = phi + ((size_t)b * n + (size_t)g * kSFRows) * n;
  for (int i = threadIdx.x; i < kSFRows * n; i += 256) s_rows[i] = src[i];
  __syncthreads();
  double* dst = part + ((size_t)b * (n / kSFRows) + g) * (n - 1);
  for (int r = 1 + threadIdx.x; r < n; r += 256) {
    double acc = 0.0;
    for (int q = 0; q < kSFRows; ++q) {
      const float* row = s_rows + q * n;
      float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
      const int m = n - r;
      int c = 0;
      for (; c + 3 < m; c += 4) {
        const float d0 = row[c] - row[c + r], d1 = row[c + 1] - row[c + 1 + r];
        const float d2 = row[c + 2] - row[c + 2 + r], d3 = row[c + 3] - row[c + 3 + r];
        a0 = fmaf(d0, d0, a0);
        a1 = fmaf(d1, d1, a1);
        a2 = fmaf(d2, d2, a2);
        a3 = fmaf(d3, d3, a3);
      }
      for (; c < m; ++c) {
        const float d = row[c] - row[c + r];
        a0 = fmaf(d, d, a0);
      }
      acc += (double)((a0 + a1) + (a2 + a3));
    }
    dst[r - 1] = acc;
  }
}

__global__ void __launch_bounds__(256)
k_sf_finalize(int n, const double* __restrict__ part, double* __restrict__ out) {
  const int b = blockIdx.y, r = 1 + blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const int groups = n / kSFRows;
  const double* p = part + (size_t)b * groups * (n - 1) + (r - 1);
  double v = 0.0;
  for (int g = 0; g < groups; ++g) v += p[(size_t)g * (n - 1)];
  out[(size_t)b * (n - 1) + r - 1] = v / ((double)(n - r) * n);
}

long long sf_scratch_bytes(int n, int batch) {
  return (long long)batch * (n / kSFRows) * (n - 1) * (long long)sizeof(double);
}

int structure_function(int n, const float* phi, double* part, double* out, int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_sf_partials, dim3(n / kSFRows, batch), dim3(256), kSFRows * n * sizeof(float), s, n, phi,
                     part);
  hipLaunchKernelGGL(k_sf_finalize, dim3((n - 1 + 255) / 256, batch), dim3(256), 0, s, n, part, out);
  return check_launch("structure_function");
}

int fill_twiddles(int n, float2* tw, hipStream_t s) {
  hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, s, n, tw);
  return check_launch("twiddles");
}

}  // namespace pa
