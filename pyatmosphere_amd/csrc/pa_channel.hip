// Channel-level entry points: one opaque plan per channel that owns the device
// tables (source factors, per-step transfer factors, twiddles, pupils,
// spectral base and sqrt psd) and a workspace for batches of up to max_batch
// realizations, and runs a whole batch (PhaseScreensPath.generator followed by
// the measures, pathes.py:61-75 and measures.py:1-38) with one call.  The
// plan is built on the host in float64 exactly as the package's Python layer
// builds its tables (source: theory/sources.py:16-18; transfer function:
// theory/vacuum.py:5-7 on the unshifted frequency axis with the inverse-FFT
// norm and the segment gain folded into the column factor; pathes.py:27-40
// for zero-length steps), and is the one place in the library that allocates
// device memory (at create; run allocates nothing).
#include <complex>
#include <vector>

#include "pa_internal.h"

namespace pa {

int operand_bytes(int n, int points, int mode, long long* out);
int spectral_draw(int P, const float* base, const float* sqrt_psd, const int* chan, const uint64_t* rid,
                  uint64_t seed, int screen, float* rho, float* theta, float2* value, int batch, hipStream_t s);
int screen_prep(int P, const float* rho, const float* theta, const float2* value, double* fxy, double* sparam,
                int mode, const double* chparam, const int* chan, const int* nsteps, int step, int batch,
                hipStream_t s);
int screen_draw_prep(int P, const float* base, const float* sqrt_psd, const int* chan, const uint64_t* rid,
                     uint64_t seed, int screen, double* fxy, double* sparam, int mode, const double* chparam,
                     const int* nsteps, int step, int batch, hipStream_t s);
int screen_tables(int n, int P, double* fxy, double* sparam, const double* chparam, const int* chan,
                  void* opa, void* opb, double* hg, float2* esc, int mode, int rows, const int* nsteps, int step,
                  int batch, hipStream_t s);
int screen_gemm(int n, int P, const void* opa, const void* opb, const double* sparam, const float2* esc,
                float* phi2, float* ang, float2* fac, int mode, int tile, const int* nsteps, int step, int batch,
                hipStream_t s);
int moments_finalize(int n, const double* part, const double* chparam, int npup, const int* chan, double* out,
                     int batch, hipStream_t s);
int fill_twiddles(int n, float2* tw, hipStream_t s);

}  // namespace pa

struct pa_channel {
  int n = 0, points = 0, screens = 0, max_batch = 0, mode = 0, npup = 0;
  std::vector<void*> allocs;
  // tables
  float2 *src = nullptr, *htab = nullptr, *tw = nullptr;
  double *chparam = nullptr, *pup = nullptr;
  float *base = nullptr, *sqrt_psd = nullptr;
  // workspace
  float2 *work = nullptr, *esc = nullptr, *sep = nullptr;
  float* ang = nullptr;
  void *opa = nullptr, *opb = nullptr;
  double *fxy = nullptr, *sparam = nullptr, *hg = nullptr, *part = nullptr;
  int* chan = nullptr;
};

namespace {

using namespace pa;
using cd = std::complex<double>;

template <typename T>
int alloc(pa_channel* c, T** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 1) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(PA_ERR_HIP, "channel_create: device allocation failed");
  }
  c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return PA_OK;
}

template <typename T>
int upload(T* dst, const std::vector<T>& src, hipStream_t s) {
  if (hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(PA_ERR_HIP, "channel_create: table upload failed");
  }
  return PA_OK;
}

void release(pa_channel* c) {
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
}

float2 to_c64(cd z) { return make_float2((float)z.real(), (float)z.imag()); }

}  // namespace

#define STREAM(s) reinterpret_cast<hipStream_t>(s)
#define CHECK(expr)        \
  do {                     \
    const int rc_ = (expr); \
    if (rc_) return rc_;   \
  } while (0)

extern "C" {

int pa_channel_create(pa_channel** out, int n, double delta, double wvl, double w0, double F0, double length,
                      int nscreens, const double* positions, const double* gains, int points, const float* base,
                      const float* psd, int npup, const double* pupils, double shift_x, double shift_y,
                      int max_batch, int mode, void* stream) {
  if (!out) return set_error(PA_ERR_ARG, "channel_create: null output");
  *out = nullptr;
  if (!(n >= 64 && n <= 4096 && (n & (n - 1)) == 0))
    return set_error(PA_ERR_SIZE, "grid size must be a power of two in [64, 4096]");
  if (nscreens < 0 || (nscreens > 0 && !positions)) return set_error(PA_ERR_ARG, "channel_create: bad screens");
  if (points <= 0 || !base || !psd) return set_error(PA_ERR_ARG, "channel_create: spectral base and psd needed");
  if (npup < 0 || npup > PA_MAX_PUPILS || (npup > 0 && !pupils))
    return set_error(PA_ERR_ARG, "channel_create: pupils");
  if (max_batch <= 0 || mode < 0 || mode > 2 || !(delta > 0) || !(wvl > 0) || !(w0 > 0))
    return set_error(PA_ERR_ARG, "channel_create: batch, mode, spacing, wavelength and waist must be valid");
  hipStream_t s = STREAM(stream);
  pa_channel* c = new pa_channel;
  c->n = n;
  c->points = points;
  c->screens = nscreens;
  c->max_batch = max_batch;
  c->mode = mode;
  c->npup = npup;
  const int steps = nscreens + 1;
  const double k = 2 * M_PI / wvl;
  // grid axis (grids.py:40-53): (i - n/2) delta
  std::vector<double> ax(n);
  for (int i = 0; i < n; ++i) ax[i] = (double)(i - n / 2) * delta;
  // separable Gaussian source: sqrt(2/pi)/w0 exp(-(1/w0^2 + i k/(2 F0)) r^2)
  std::vector<float2> src(2 * n);
  const double curv = std::isinf(F0) ? 0.0 : k / 2 / F0;
  for (int i = 0; i < n; ++i) {
    const cd g = std::exp(-cd(1 / (w0 * w0), curv) * (ax[i] * ax[i]));
    src[i] = to_c64(std::sqrt(2 / M_PI) / w0 * g);
    src[n + i] = to_c64(g);
  }
  // Fresnel steps d_0 .. d_S (pathes.py:67-75) and their transfer factors:
  // row h(f_r), column s h(f_c) with h = exp(-i pi lambda d f^2) on
  // fftfreq(n, delta) and s = exp(i k d) / n^2 times the step's gain
  std::vector<double> d(steps);
  for (int i = 0; i < nscreens; ++i) d[i] = i > 0 ? positions[i] - positions[i - 1] : positions[0];
  d[nscreens] = length - (nscreens > 0 ? positions[nscreens - 1] : 0.0);
  std::vector<double> f(n);
  const double df = 1.0 / (n * delta);   // numpy.fft.fftfreq: integer index times 1 / (n d)
  for (int i = 0; i < n; ++i) f[i] = (double)(i < (n + 1) / 2 ? i : i - n) * df;
  std::vector<float2> htab((size_t)steps * 2 * n);
  for (int st = 0; st < steps; ++st) {
    const double gain = gains ? gains[st] : 1.0;
    const cd sc = d[st] > 0 ? std::exp(cd(0, k * d[st])) / (double)n / (double)n * gain : cd(gain / n / n, 0);
    for (int i = 0; i < n; ++i) {
      // the argument in numpy's order of operations, ((-pi d) lambda) f^2
      const cd h = d[st] > 0 ? std::exp(cd(0, ((-M_PI * d[st]) * wvl) * (f[i] * f[i]))) : cd(1, 0);
      htab[(size_t)st * 2 * n + i] = to_c64(h);
      htab[(size_t)st * 2 * n + n + i] = to_c64(sc * h);
    }
  }
  std::vector<double> chp = {delta, shift_x, shift_y, 0.0};
  std::vector<double> pup((size_t)PA_MAX_PUPILS * 3, 0.0);
  for (int p = 0; p < npup * 3; ++p) pup[p] = pupils[p];
  std::vector<float> bs(base, base + points), sq(points);
  for (int i = 0; i < points; ++i) sq[i] = std::sqrt(psd[i]);
  long long opb = 0;
  int rc = operand_bytes(n, points, mode, &opb);
  const int T = tile_rows(n);
  const size_t B = (size_t)max_batch, nn = (size_t)n * n;
  if (!rc) rc = alloc(c, &c->src, 2 * (size_t)n);
  if (!rc) rc = alloc(c, &c->htab, htab.size());
  if (!rc) rc = alloc(c, &c->tw, (size_t)n);
  if (!rc) rc = alloc(c, &c->chparam, (size_t)PA_CH_PARAMS);
  if (!rc) rc = alloc(c, &c->pup, pup.size());
  if (!rc) rc = alloc(c, &c->base, (size_t)points);
  if (!rc) rc = alloc(c, &c->sqrt_psd, (size_t)points);
  if (!rc) rc = alloc(c, &c->work, B * nn);
  if (!rc) rc = alloc(c, &c->ang, B * nn);
  if (!rc) rc = alloc(c, reinterpret_cast<unsigned char**>(&c->opa), B * (size_t)opb);
  if (!rc) rc = alloc(c, reinterpret_cast<unsigned char**>(&c->opb), B * (size_t)opb);
  if (!rc) rc = alloc(c, &c->fxy, B * points * 4);
  if (!rc) rc = alloc(c, &c->sparam, B * PA_SCREEN_PARAMS);
  if (!rc) rc = alloc(c, &c->hg, B * 2 * n);
  if (!rc) rc = alloc(c, &c->esc, B * 2 * n);
  if (!rc) rc = alloc(c, &c->sep, B * 2 * n);
  if (!rc) rc = alloc(c, &c->part, B * (n / T) * PA_PART_SUMS);
  if (!rc) rc = alloc(c, &c->chan, B);
  if (!rc) rc = upload(c->src, src, s);
  if (!rc) rc = upload(c->htab, htab, s);
  if (!rc) rc = upload(c->chparam, chp, s);
  if (!rc) rc = upload(c->pup, pup, s);
  if (!rc) rc = upload(c->base, bs, s);
  if (!rc) rc = upload(c->sqrt_psd, sq, s);
  if (!rc) rc = upload(c->chan, std::vector<int>(B, 0), s);
  if (!rc) rc = fill_twiddles(n, c->tw, s);
  if (rc) {
    release(c);
    delete c;
    return rc;
  }
  *out = c;
  return PA_OK;
}

int pa_channel_run(pa_channel* c, int batch, const float* rho, const float* theta, const void* value,
                   const uint64_t* rid, uint64_t seed, const void* field_in, void* field_out, double* moments,
                   void* stream) {
  if (!c) return set_error(PA_ERR_ARG, "channel_run: null plan");
  if (batch <= 0 || batch > c->max_batch) return set_error(PA_ERR_ARG, "channel_run: batch outside 1..max_batch");
  const bool host_draws = rho || theta || value;
  if (host_draws && !(rho && theta && value)) return set_error(PA_ERR_ARG, "channel_run: rho, theta, value together");
  if (!host_draws && c->screens > 0 && !rid) return set_error(PA_ERR_ARG, "channel_run: device draws need rid");
  hipStream_t s = STREAM(stream);
  const int n = c->n, P = c->points, steps = c->screens + 1;
  // from the source with a screen to come, the first Fresnel step as two 1-D
  // propagations (the source and H_0 are separable; pa_source_fresnel)
  const bool sep = !field_in && c->screens > 0;
  if (sep) {
    CHECK(source_fresnel(n, c->sep, c->src, c->htab, c->chan, steps, c->tw, batch, s));
  } else {
    CHECK(row_first(n, c->work, reinterpret_cast<const float2*>(field_in), field_in ? nullptr : c->src, c->chan,
                    c->tw, batch, s));
  }
  for (int i = 0; i < steps; ++i) {
    if (!(sep && i == 0)) CHECK(col_step(n, c->work, c->htab, c->chan, nullptr, i, steps, c->tw, 0, batch, s));
    if (i == c->screens) break;
    if (host_draws) {
      const size_t o = (size_t)i * batch * P;
      CHECK(screen_prep(P, rho + o, theta + o, reinterpret_cast<const float2*>(value) + o, c->fxy, c->sparam,
                        c->mode, c->chparam, c->chan, nullptr, i, batch, s));
    } else {
      CHECK(screen_draw_prep(P, c->base, c->sqrt_psd, c->chan, rid, seed, i, c->fxy, c->sparam, c->mode,
                             c->chparam, nullptr, i, batch, s));
    }
    CHECK(screen_tables(n, P, c->fxy, c->sparam, c->chparam, c->chan, c->opa, c->opb, c->hg, c->esc, c->mode, 0,
                        nullptr, i, batch, s));
    CHECK(screen_gemm(n, P, c->opa, c->opb, c->sparam, c->esc, nullptr, c->ang, nullptr, c->mode, 0, nullptr, i,
                      batch, s));
    if (sep && i == 0) {
      CHECK(row_mid_source(n, c->work, c->sep, c->ang, c->tw, batch, s));
    } else {
      CHECK(row_mid(n, c->work, c->ang, nullptr, i, c->tw, batch, s));
    }
  }
  CHECK(row_last(n, c->work, reinterpret_cast<float2*>(field_out), moments ? c->part : nullptr, c->chparam, c->pup,
                 c->npup, nullptr, c->chan, c->tw, batch, s));
  if (moments) CHECK(moments_finalize(n, c->part, c->chparam, c->npup, c->chan, moments, batch, s));
  return PA_OK;
}

int pa_channel_destroy(pa_channel* c) {
  if (!c) return PA_OK;
  release(c);
  delete c;
  return PA_OK;
}

}  // extern "C"
