// extern "C" boundary of libpa_hip.so (declared in include/pa_hip.h).
// Validates arguments on the host, forwards to the kernel launchers, and
// turns every failure into a negative status plus a thread-local message.
#include <stdio.h>
#include <string.h>

#include "pa_internal.h"

namespace pa {

static thread_local char g_err[512] = "";

int set_error(int code, const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    return PA_ERR_HIP;
  }
  return PA_OK;
}

// launchers defined in the other translation units
int operand_bytes(int n, int points, int mode, long long* out);
long long sf_scratch_bytes(int n, int batch);
int structure_function(int n, const float* phi, double* part, double* out, int batch, hipStream_t s);
int spectral_draw(int P, const float* base, const float* sqrt_psd, const int* chan, const uint64_t* rid,
                  uint64_t seed, int screen, float* rho, float* theta, float2* value, int batch,
                  hipStream_t s);
int screen_prep(int P, const float* rho, const float* theta, const float2* value, double* fxy,
                double* sparam, int mode, const double* chparam, const int* chan, const int* nsteps, int step,
                int batch, hipStream_t s);
int screen_draw_prep(int P, const float* base, const float* sqrt_psd, const int* chan, const uint64_t* rid,
                     uint64_t seed, int screen, double* fxy, double* sparam, int mode, const double* chparam,
                     const int* nsteps, int step, int batch, hipStream_t s);
int screen_tables(int n, int P, double* fxy, double* sparam,
                  const double* chparam, const int* chan, void* opa, void* opb, double* hg,
                  float2* esc, int mode, int rows, const int* nsteps, int step, int batch, hipStream_t s);
int screen_gemm(int n, int P, const void* opa, const void* opb, const double* sparam, const float2* esc,
                float* phi2, float* ang, float2* fac, int mode, int tile, const int* nsteps, int step, int batch,
                hipStream_t s);
int screen_compose(int n, const double* sparam, const double* hg, const float* phi2, float* phi,
                   int batch, hipStream_t s);
int field_partials(int n, const float2* u, double* part, const double* chparam, const double* pup,
                   int npup, const double* pup_moments, const int* chan, int batch, hipStream_t s);
int moments_finalize(int n, const double* part, const double* chparam, int npup, const int* chan,
                     double* out, int batch, hipStream_t s);
int apply_pupil(int n, float2* u, double delta, double radius, double sx, double sy, int batch,
                hipStream_t s);
int ensemble_accumulate(int batch, const double* mom, const int* slot_chan, int nchan, int npup, int beam_cap,
                        int pdt_cap, double* acc, long long* counts, unsigned long long* hist, hipStream_t s);
int hbm_stream(int kind, const void* src, void* dst, long long bytes, int unroll, int nt, hipStream_t s);
long long static_scratch_bytes();
long long moments_scratch_bytes(int n, int batch);
int field_moments_static(int n, const float2* u, const double* chparam, const double* pup, int npup,
                         const double* pup_moments, const int* chan, double* out, int batch, hipStream_t s);
int fill_twiddles(int n, float2* tw, hipStream_t s);
int outer(int n, float2* out, const float2* gy, const float2* gx, int batch, hipStream_t s);
int intensity(long long count, const float2* u, float* out, hipStream_t s);

int mlp_gemm_f32(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb, int tb, float* C,
                 int ldc, float alpha, float beta, hipStream_t s);
int mlp_bias_act(int M, int N, float* Z, const float* bias, float* Y, int act, hipStream_t s);
int mlp_act_backward(int M, int N, const float* Z, const float* dY, float* dZ, hipStream_t s);
int mlp_bias_grad(int M, int N, const float* dZ, float* db, hipStream_t s);
int mlp_head_loss(int M, int NB, const float* out, const float* target, const float* tvar, float wv, float wc,
                  float* loss, float* dout, hipStream_t s);
int mlp_adam(int n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
             int t, hipStream_t s);
int mlp_to_bf16(long long n, const float* x, unsigned short* y, hipStream_t s);
int mlp_layer_bf16(int M, int N, int K, const float* X, const unsigned short* W, const float* bias, float* Y,
                   int act, hipStream_t s);

static bool pow2_in_range(int n) { return n >= 64 && n <= 4096 && (n & (n - 1)) == 0; }

}  // namespace pa

using namespace pa;

#define REQUIRE(cond, msg)                           \
  do {                                               \
    if (!(cond)) return set_error(PA_ERR_ARG, msg);  \
  } while (0)
#define REQUIRE_N(n) \
  do { if (!pow2_in_range(n)) return set_error(PA_ERR_SIZE, "grid size must be a power of two in [64, 4096]"); } while (0)
#define STREAM(s) reinterpret_cast<hipStream_t>(s)

extern "C" {

const char* pa_last_error(void) { return g_err; }
int pa_version(void) { return 1; }

int pa_tile_rows(int n) {
  REQUIRE_N(n);
  return tile_rows(n);
}

long long pa_operand_bytes(int n, int points, int mode) {
  if (!pow2_in_range(n) || points <= 0) return PA_ERR_SIZE;
  long long v = 0;
  const int rc = operand_bytes(n, points, mode, &v);
  return rc ? rc : v;
}

long long pa_structure_function_scratch(int n, int batch) {
  if (!pow2_in_range(n) || batch <= 0) return PA_ERR_SIZE;
  return sf_scratch_bytes(n, batch);
}

int pa_structure_function(int n, const float* phi, double* scratch, double* out, int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(batch > 0, "structure_function: batch must be positive");
  REQUIRE(phi && scratch && out, "structure_function: null argument");
  return structure_function(n, phi, scratch, out, batch, STREAM(stream));
}

int pa_fill_twiddles(int n, void* tw, void* stream) {
  REQUIRE_N(n);
  REQUIRE(tw, "tw is null");
  return fill_twiddles(n, reinterpret_cast<float2*>(tw), STREAM(stream));
}

int pa_row_first(int n, void* work, const void* in, const void* src, const int* chan, const void* tw,
                 int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(work && tw && batch > 0, "row_first: null work/tw or batch <= 0");
  REQUIRE(in || src, "row_first: need an input field or source factors");
  return row_first(n, reinterpret_cast<float2*>(work), reinterpret_cast<const float2*>(in),
                   reinterpret_cast<const float2*>(src), chan, reinterpret_cast<const float2*>(tw),
                   batch, STREAM(stream));
}

int pa_col_step(int n, void* work, const void* htab, const int* chan, const int* nsteps, int step,
                int max_steps, const void* tw, int cols, int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(work && htab && tw && batch > 0, "col_step: null argument or batch <= 0");
  REQUIRE(step >= 0 && step < max_steps, "col_step: step out of range");
  return col_step(n, reinterpret_cast<float2*>(work), reinterpret_cast<const float2*>(htab), chan,
                  nsteps, step, max_steps, reinterpret_cast<const float2*>(tw), cols, batch, STREAM(stream));
}

int pa_col_groups(int n, int* out) {
  REQUIRE_N(n);
  REQUIRE(out, "col_groups: null output");
  return col_groups(n, out);
}

int pa_row_mid(int n, void* work, const float* ang, const int* nsteps, int step, const void* tw, int batch,
               void* stream) {
  REQUIRE_N(n);
  REQUIRE(work && ang && tw && batch > 0, "row_mid: null argument or batch <= 0");
  return row_mid(n, reinterpret_cast<float2*>(work), ang, nsteps, step, reinterpret_cast<const float2*>(tw), batch,
                 STREAM(stream));
}

int pa_source_fresnel(int n, void* sep, const void* src, const void* htab, const int* chan, int max_steps,
                      const void* tw, int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(sep && src && htab && tw && batch > 0, "source_fresnel: null argument or batch <= 0");
  REQUIRE(max_steps >= 1, "source_fresnel: max_steps < 1");
  return source_fresnel(n, reinterpret_cast<float2*>(sep), reinterpret_cast<const float2*>(src),
                        reinterpret_cast<const float2*>(htab), chan, max_steps, reinterpret_cast<const float2*>(tw),
                        batch, STREAM(stream));
}

int pa_row_mid_source(int n, void* work, const void* sep, const float* ang, const void* tw, int batch,
                      void* stream) {
  REQUIRE_N(n);
  REQUIRE(work && sep && ang && tw && batch > 0, "row_mid_source: null argument or batch <= 0");
  return row_mid_source(n, reinterpret_cast<float2*>(work), reinterpret_cast<const float2*>(sep), ang,
                        reinterpret_cast<const float2*>(tw), batch, STREAM(stream));
}

int pa_expmi(int count, const float* x, void* out, void* stream) {
  REQUIRE(count >= 0, "expmi: count < 0");
  REQUIRE(count == 0 || (x && out), "expmi: null argument");
  if (count == 0) return PA_OK;
  return expmi_eval(count, x, reinterpret_cast<float2*>(out), STREAM(stream));
}

int pa_row_last(int n, const void* work, void* out, double* part, const double* chparam,
                const double* pup, int npup, const double* pup_moments, const int* chan, const void* tw, int batch,
                void* stream) {
  REQUIRE_N(n);
  REQUIRE(work && tw && batch > 0, "row_last: null argument or batch <= 0");
  REQUIRE(out || part, "row_last: nothing to produce");
  REQUIRE(!part || chparam, "row_last: partial sums need chparam");
  REQUIRE(npup >= 0 && npup <= PA_MAX_PUPILS, "row_last: too many pupils");
  REQUIRE(npup == 0 || pup, "row_last: pupils table is null");
  return row_last(n, reinterpret_cast<const float2*>(work), reinterpret_cast<float2*>(out), part,
                  chparam, pup, npup, pup_moments, chan, reinterpret_cast<const float2*>(tw), batch, STREAM(stream));
}

int pa_field_partials(int n, const void* u, double* part, const double* chparam, const double* pup,
                      int npup, const double* pup_moments, const int* chan, int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(u && part && chparam && batch > 0, "field_partials: null argument or batch <= 0");
  REQUIRE(npup >= 0 && npup <= PA_MAX_PUPILS, "field_partials: too many pupils");
  REQUIRE(npup == 0 || pup, "field_partials: pupils table is null");
  return field_partials(n, reinterpret_cast<const float2*>(u), part, chparam, pup, npup, pup_moments,
                        chan, batch, STREAM(stream));
}

int pa_moments_finalize(int n, const double* part, const double* chparam, int npup, const int* chan,
                        double* out, int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(part && chparam && out && batch > 0, "moments_finalize: null argument or batch <= 0");
  REQUIRE(npup >= 0 && npup <= PA_MAX_PUPILS, "moments_finalize: too many pupils");
  return moments_finalize(n, part, chparam, npup, chan, out, batch, STREAM(stream));
}

int pa_ensemble_accumulate(int batch, const double* moments, const int* slot_channel, int channels, int npup,
                           int beam_cap, int pdt_cap, double* acc, long long* counts, uint64_t* hist,
                           void* stream) {
  REQUIRE(batch > 0 && channels > 0, "ensemble_accumulate: batch and channels must be positive");
  REQUIRE(moments && acc && counts && hist, "ensemble_accumulate: null argument");
  REQUIRE(npup >= 0 && npup <= PA_MAX_PUPILS, "ensemble_accumulate: too many pupils");
  return ensemble_accumulate(batch, moments, slot_channel, channels, npup, beam_cap, pdt_cap, acc, counts,
                             reinterpret_cast<unsigned long long*>(hist), STREAM(stream));
}

int pa_field_moments_static(int n, const void* u, const double* chparam, const double* pup, int npup,
                            const double* pup_moments, const int* chan, double* out, int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(batch > 0, "field_moments_static: batch must be positive");
  if (moments_scratch_bytes(n, batch) > static_scratch_bytes())
    return set_error(PA_ERR_SIZE, "field_moments_static: batch * (n / tile rows) * PA_PART_SUMS * 8 bytes of "
                                  "partials exceed the library's static scratch (pa_field_moments_static_max)");
  REQUIRE(u && chparam && out, "field_moments_static: null argument");
  REQUIRE(npup >= 0 && npup <= PA_MAX_PUPILS, "field_moments_static: too many pupils");
  REQUIRE(npup == 0 || pup, "field_moments_static: pupils table is null");
  return field_moments_static(n, reinterpret_cast<const float2*>(u), chparam, pup, npup, pup_moments, chan, out,
                              batch, STREAM(stream));
}

int pa_hbm_stream(int kind, const void* src, void* dst, long long bytes, int unroll, int nontemporal,
                  void* stream) {
  REQUIRE(kind >= 0 && kind <= 2, "hbm_stream: kind must be 0 (copy), 1 (read) or 2 (write)");
  REQUIRE(bytes > 0 && bytes % 16 == 0, "hbm_stream: bytes must be a positive multiple of 16");
  REQUIRE((kind == 2 || src) && dst, "hbm_stream: null argument");
  return hbm_stream(kind, src, dst, bytes, unroll, nontemporal, STREAM(stream));
}

int pa_fft_screen(int n, const void* coeff, void* work, const void* sub, int nsub, const void* mean, const void* tw,
                  float* phi, float* phi_im, float* ang, int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(batch > 0 && nsub >= 0, "fft_screen: batch must be positive and nsub non-negative");
  REQUIRE(coeff && work && tw && mean && (nsub == 0 || sub), "fft_screen: null argument");
  REQUIRE(phi || phi_im || ang, "fft_screen: no output");
  return fft_screen(n, reinterpret_cast<const float2*>(coeff), reinterpret_cast<float2*>(work),
                    reinterpret_cast<const double2*>(sub), nsub, reinterpret_cast<const double2*>(mean),
                    reinterpret_cast<const float2*>(tw), phi, phi_im, ang, batch, STREAM(stream));
}

int pa_field_moments_static_max(int n) {
  REQUIRE_N(n);
  return (int)(static_scratch_bytes() / moments_scratch_bytes(n, 1));
}

int pa_spectral_draw(int points, const float* base, const float* sqrt_psd, const int* chan,
                     const uint64_t* rid, uint64_t seed, int screen, float* rho, float* theta,
                     void* value, int batch, void* stream) {
  REQUIRE(points > 0 && batch > 0, "spectral_draw: points and batch must be positive");
  REQUIRE(base && sqrt_psd && rid && rho && theta && value, "spectral_draw: null argument");
  return spectral_draw(points, base, sqrt_psd, chan, rid, seed, screen, rho, theta,
                       reinterpret_cast<float2*>(value), batch, STREAM(stream));
}

int pa_screen_prep(int points, const float* rho, const float* theta, const void* value, double* fxy,
                   double* sparam, int mode, const double* chparam, const int* chan, const int* nsteps, int step,
                   int batch, void* stream) {
  REQUIRE(points > 0 && batch > 0, "screen_prep: points and batch must be positive");
  REQUIRE(rho && theta && value && fxy && sparam && chparam, "screen_prep: null argument");
  REQUIRE(mode >= 0 && mode <= 2, "screen_prep: unknown GEMM mode");
  return screen_prep(points, rho, theta, reinterpret_cast<const float2*>(value), fxy, sparam, mode, chparam,
                     chan, nsteps, step, batch, STREAM(stream));
}

int pa_screen_draw_prep(int points, const float* base, const float* sqrt_psd, const int* chan, const uint64_t* rid,
                        uint64_t seed, int screen, double* fxy, double* sparam, int mode, const double* chparam,
                        const int* nsteps, int step, int batch, void* stream) {
  REQUIRE(points > 0 && batch > 0, "screen_draw_prep: points and batch must be positive");
  REQUIRE(base && sqrt_psd && rid && fxy && sparam && chparam, "screen_draw_prep: null argument");
  REQUIRE(mode >= 0 && mode <= 2, "screen_draw_prep: unknown GEMM mode");
  return screen_draw_prep(points, base, sqrt_psd, chan, rid, seed, screen, fxy, sparam, mode, chparam, nsteps, step,
                          batch, STREAM(stream));
}

int pa_screen_tables(int n, int points, double* fxy, double* sparam, const double* chparam, const int* chan,
                     void* opa, void* opb, double* hg, void* esc, int mode, int rows, const int* nsteps, int step,
                     int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(points > 0 && batch > 0, "screen_tables: points and batch must be positive");
  REQUIRE(fxy && sparam && chparam && opa && opb && hg && esc, "screen_tables: null argument");
  return screen_tables(n, points, fxy, sparam, chparam, chan, opa, opb, hg, reinterpret_cast<float2*>(esc), mode,
                       rows, nsteps, step, batch, STREAM(stream));
}

int pa_screen_gemm(int n, int points, const void* opa, const void* opb, const double* sparam, const void* esc,
                   float* phi2, float* ang, void* fac, int mode, int tile, const int* nsteps, int step, int batch,
                   void* stream) {
  REQUIRE_N(n);
  REQUIRE(points > 0 && batch > 0, "screen_gemm: points and batch must be positive");
  REQUIRE(opa && opb && sparam && esc && (phi2 || ang || fac), "screen_gemm: null argument");
  REQUIRE(mode >= 0 && mode <= 2, "screen_gemm: unknown GEMM mode");
  return screen_gemm(n, points, opa, opb, sparam, reinterpret_cast<const float2*>(esc), phi2, ang,
                     reinterpret_cast<float2*>(fac), mode, tile, nsteps, step, batch, STREAM(stream));
}

int pa_screen_compose(int n, const double* sparam, const double* hg, const float* phi2, float* phi,
                      int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(sparam && hg && phi2 && phi && batch > 0, "screen_compose: null argument");
  return screen_compose(n, sparam, hg, phi2, phi, batch, STREAM(stream));
}

int pa_apply_pupil(int n, void* u, double delta, double radius, double shift_x, double shift_y,
                   int batch, void* stream) {
  REQUIRE_N(n);
  REQUIRE(u && batch > 0, "apply_pupil: null field or batch <= 0");
  return apply_pupil(n, reinterpret_cast<float2*>(u), delta, radius, shift_x, shift_y, batch,
                     STREAM(stream));
}

int pa_outer(int n, void* out, const void* gy, const void* gx, int batch, void* stream) {
  REQUIRE(n > 0 && n <= 65536 && out && gy && gx && batch > 0, "outer: bad argument");
  return outer(n, reinterpret_cast<float2*>(out), reinterpret_cast<const float2*>(gy),
               reinterpret_cast<const float2*>(gx), batch, STREAM(stream));
}

int pa_intensity(long long count, const void* u, float* out, void* stream) {
  REQUIRE(count > 0 && u && out, "intensity: bad argument");
  return intensity(count, reinterpret_cast<const float2*>(u), out, STREAM(stream));
}

int pa_mlp_gemm_f32(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb, int tb,
                    float* C, int ldc, float alpha, float beta, void* stream) {
  REQUIRE(M > 0 && N > 0 && K > 0 && A && B && C, "mlp_gemm_f32: bad argument");
  REQUIRE(lda >= (ta ? M : K) && ldb >= (tb ? K : N) && ldc >= N, "mlp_gemm_f32: leading dimension too small");
  return mlp_gemm_f32(M, N, K, A, lda, ta, B, ldb, tb, C, ldc, alpha, beta, STREAM(stream));
}
int pa_mlp_bias_act(int M, int N, float* Z, const float* bias, float* Y, int act, void* stream) {
  REQUIRE(M > 0 && N > 0 && Z && bias && Y, "mlp_bias_act: bad argument");
  return mlp_bias_act(M, N, Z, bias, Y, act, STREAM(stream));
}
int pa_mlp_act_backward(int M, int N, const float* Z, const float* dY, float* dZ, void* stream) {
  REQUIRE(M > 0 && N > 0 && Z && dY && dZ, "mlp_act_backward: bad argument");
  return mlp_act_backward(M, N, Z, dY, dZ, STREAM(stream));
}
int pa_mlp_bias_grad(int M, int N, const float* dZ, float* db, void* stream) {
  REQUIRE(M > 0 && N > 0 && dZ && db, "mlp_bias_grad: bad argument");
  return mlp_bias_grad(M, N, dZ, db, STREAM(stream));
}
int pa_mlp_head_loss(int M, int NB, const float* out, const float* target, const float* tvar, float wv, float wc,
                     float* loss, float* dout, void* stream) {
  REQUIRE(M > 0 && NB > 0 && NB <= 64 && out && target && tvar && (loss || dout), "mlp_head_loss: bad argument");
  return mlp_head_loss(M, NB, out, target, tvar, wv, wc, loss, dout, STREAM(stream));
}
int pa_mlp_adam(int n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                int t, void* stream) {
  REQUIRE(n > 0 && w && g && m && v && t >= 1, "mlp_adam: bad argument");
  return mlp_adam(n, w, g, m, v, lr, b1, b2, eps, t, STREAM(stream));
}
int pa_mlp_to_bf16(long long n, const float* x, void* y, void* stream) {
  REQUIRE(n > 0 && x && y, "mlp_to_bf16: bad argument");
  return mlp_to_bf16(n, x, reinterpret_cast<unsigned short*>(y), STREAM(stream));
}
int pa_mlp_layer_bf16(int M, int N, int K, const float* X, const void* W, const float* bias, float* Y, int act,
                      void* stream) {
  REQUIRE(M > 0 && N > 0 && K > 0 && X && W && bias && Y, "mlp_layer_bf16: bad argument");
  REQUIRE(K % 16 == 0 && N % 32 == 0, "mlp_layer_bf16: K must be a multiple of 16 and N of 32");
  return mlp_layer_bf16(M, N, K, X, reinterpret_cast<const unsigned short*>(W), bias, Y, act, STREAM(stream));
}

}  // extern "C"
