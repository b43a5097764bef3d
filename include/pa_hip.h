/* C ABI of libpa_hip.so: the MI355X (gfx950) split-step phase-screen path.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer owned by the caller; the
 *     library allocates nothing.  Complex arrays are interleaved float32
 *     pairs (complex64).  `stream` is a hipStream_t passed as void* (NULL =
 *     the default stream).  Nothing synchronises: every call only enqueues.
 *   - Return 0 on success, a negative PA_ERR_* code on error; the message
 *     of the last error on the calling thread is pa_last_error().
 *   - Batches: `batch` independent realizations are processed per call;
 *     per-element arrays are indexed [b][...].  `chan` (nullable = all 0)
 *     maps an element to a channel row of the per-channel tables; `nsteps`
 *     (nullable = all active) holds the screen count S_b of each element,
 *     so elements with fewer screens skip the steps they do not have.
 *   - Field layout between passes ("work"): rows are grouped in tiles of
 *     T = pa_tile_rows(n) rows, element (r, c) at ((r/T)*n + c)*T + r%T.
 *     Fields handed in or out by the caller are row-major [b][r][c].
 *   - Grid coordinates: x_c = (c - n/2) delta, y_r = (r - n/2) delta.
 *
 * Each entry point cites the pyAtmosphere interface it replaces
 * (paths relative to the pyAtmosphere source tree).
 */
#ifndef PA_HIP_H
#define PA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PA_MAX_PUPILS 8
/* finalized values per element: eta, mean_x, mean_y, mean_x2, mean_xy,
 * mean_y2, mean_x2_r, then eta inside each pupil */
#define PA_NUM_MOMENTS (7 + PA_MAX_PUPILS)
/* per-tile partial sums written by pa_row_last / pa_field_partials */
#define PA_PART_SUMS (6 + PA_MAX_PUPILS)
/* per-sample quantities of pa_ensemble_accumulate and its histogram bins */
#define PA_ENS_QUANTITIES 11
#define PA_PDT_BINS 64
/* per-channel parameter row: delta, shift_x, shift_y, spare (float64) */
#define PA_CH_PARAMS 4
/* Low-frequency points folded into the screen GEMM as a polynomial block:
 * PA_LOWRANK_TERMS Chebyshev terms per axis (see pa_screen_tables). */
#define PA_LOWRANK_TERMS 32
/* per-element screen parameter row (float64): [0] phi0, [1] A scale,
 * [2] B scale, [3] 1/(A scale * B scale), [6] the direct points' A scale
 * (pa_screen_prep, which also clears [4] and [5]); [4] kbeg: first K column
 * the GEMM reads, [5] d0: number of leading points of fxy (after
 * pa_screen_tables has grouped them) folded into the polynomial block, [7]
 * n2: how many of those are smooth along both axes, [8 + PA_LOWRANK_TERMS n +
 * m] Re G[n][m], [8 + PA_LOWRANK_TERMS^2 + n] and [8 + PA_LOWRANK_TERMS^2 +
 * PA_LOWRANK_TERMS + m] the Chebyshev coefficients of the block points' shares
 * of h(r) (first n2) and g(c) (all d0), and [1], [3] lowered when the block's
 * entries need a smaller A scale (pa_screen_tables) */
#define PA_SCREEN_PARAMS (8 + PA_LOWRANK_TERMS * PA_LOWRANK_TERMS + 2 * PA_LOWRANK_TERMS)

enum {
  PA_OK = 0,
  PA_ERR_ARG = -1,
  PA_ERR_SIZE = -2,
  PA_ERR_HIP = -3,
};

/* Screen-synthesis GEMM precision modes. */
enum {
  PA_GEMM_F16X3 = 0,  /* fp16 hi/lo split, 3 MFMA products, fp32 accumulate */
  PA_GEMM_BF16X3 = 1, /* bf16 hi/lo split, 3 MFMA products, fp32 accumulate */
  PA_GEMM_F32 = 2,    /* fp32 MFMA                                          */
};

const char* pa_last_error(void);
int pa_version(void);
/* rows per tile of the work layout for grid size n (64..4096, power of 2) */
int pa_tile_rows(int n);
/* bytes of one screen-GEMM operand buffer (A or B) per element */
long long pa_operand_bytes(int n, int points, int mode);

/* Twiddle table tw[k] = exp(-2 pi i k / n), k < n, computed in float64. */
int pa_fill_twiddles(int n, void* tw, void* stream);

/* Source + forward row FFT (sources.py:21-23; theory/sources.py:16-18;
 * utils.py:42-44).  u0 = in[b] (row-major) when `in` is non-null, else the
 * separable Gaussian src[chan][0][r] * src[chan][1][c]. */
int pa_row_first(int n, void* work, const void* in, const void* src, const int* chan,
                 const void* tw, int batch, void* stream);

/* Column FFT, transfer function H = htab[chan][step][0][r] * htab[chan][step][1][c],
 * inverse column FFT (theory/vacuum.py:5-7; pathes.py:27-40).  Skipped for
 * elements with step > nsteps[b].  `cols`: columns per block, a launch shape
 * (0 = the library's choice, else one of pa_col_groups(n)); results do not
 * depend on it. */
int pa_col_step(int n, void* work, const void* htab, const int* chan, const int* nsteps,
                int step, int max_steps, const void* tw, int cols, int batch, void* stream);
/* The column-group sizes pa_col_step takes for grid size n (at most 3, the
 * default first) into out; returns their number. */
int pa_col_groups(int n, int* out);

/* Inverse row FFT, u *= exp(-i phi) (pathes.py:72-73) with the screen factor
 * formed in the pass as exp(-i ang[b](r,c)) from the float32 screen angle of
 * pa_screen_gemm (work layout; exp as pa_expmi), forward row FFT.  Skipped
 * for step >= nsteps[b]. */
int pa_row_mid(int n, void* work, const float* ang, const int* nsteps, int step, const void* tw, int batch,
               void* stream);

/* The first Fresnel step of the Gaussian source as two 1-D propagations
 * (sources.py:21-23 and theory/vacuum.py:5-7 with the transfer factors of
 * step 0; pathes.py:61-75's first step): the source src[chan] = (src_y, src_x)
 * and H_0 = htab[chan][0][0][fy] * htab[chan][0][1][fx] are separable, so the
 * field at the first screen plane is sep[b][0][r] * sep[b][1][c] with
 * sep[b][0] = IFFT(htab[..][0] FFT(src_y)), sep[b][1] = IFFT(htab[..][1] FFT(src_x))
 * (unnormalised transforms, complex64, sep: batch x 2 x n).  Replaces
 * pa_row_first + pa_col_step(step 0) when every element has a screen. */
int pa_source_fresnel(int n, void* sep, const void* src, const void* htab, const int* chan, int max_steps,
                      const void* tw, int batch, void* stream);

/* pa_row_mid of the first screen on the separable field of pa_source_fresnel:
 * u = sep[b][0][r] * sep[b][1][c], u *= exp(-i ang[b](r,c)) (pathes.py:72-73),
 * forward row FFT -> work (tiled).  Every element must have a first screen. */
int pa_row_mid_source(int n, void* work, const void* sep, const float* ang, const void* tw, int batch,
                      void* stream);

/* out[i] = exp(-i x[i]) (complex64) as pa_row_mid forms the screen factor:
 * reduction by 2 pi in float32, then the hardware sine and cosine; absolute
 * error below 5e-7 for |x| <= 512.  No reference counterpart (a check of the
 * device arithmetic). */
int pa_expmi(int count, const float* x, void* out, void* stream);

/* Inverse row FFT (utils.py:47-50); row-major field to `out` (nullable) and
 * per-tile intensity sums to `part` (nullable) [b][n/T][PA_PART_SUMS]
 * (measures.py:1-38; pupils.py:8-13).  pup: [chan][PA_MAX_PUPILS][3] =
 * (radius, shift_x, shift_y); when `pup_moments` is non-null each element's
 * pupils are centred on its own centroid (mean_x, mean_y) read from
 * pup_moments[b] (TrackedPDTResult, simulations/pdt.py:63-66), so a second
 * call on the same work buffer gives the tracked transmittances without
 * storing the field.  `work` is not modified. */
int pa_row_last(int n, const void* work, void* out, double* part, const double* chparam,
                const double* pup, int npup, const double* pup_moments, const int* chan, const void* tw, int batch,
                void* stream);

/* Per-tile intensity sums of row-major fields u[b] (measures.py:1-38 on a
 * given `output`; simulations/pdt.py:21-26).  pup as above; when
 * `pup_moments` is non-null each element's pupils are shifted by its own
 * centroid (mean_x, mean_y) read from pup_moments[b] (simulations/pdt.py:63-66). */
int pa_field_partials(int n, const void* u, double* part, const double* chparam,
                      const double* pup, int npup, const double* pup_moments,
                      const int* chan, int batch, void* stream);

/* Sum the per-tile partials (fixed order) into out[b][PA_NUM_MOMENTS]:
 * eta, mean_x, mean_y, mean_x2, mean_xy, mean_y2, mean_x2_r, eta_pupil[p]
 * (measures.py:7-38; simulations/beam.py:26-33). */
int pa_moments_finalize(int n, const double* part, const double* chparam, int npup,
                        const int* chan, double* out, int batch, void* stream);

/* Per-channel Monte Carlo ensembles on the device (simulations/beam.py:35-71;
 * simulations/pdt.py:28-48).  For every element b of a batch of finalized
 * moments[b][PA_NUM_MOMENTS] whose channel is c = slot_channel[b] (nullable =
 * all 0, c < channels): while counts[c][0] < beam_cap, add the float64
 * quantities q = (eta, mean_x, mean_y, mean_x2, mean_xy, mean_y2, mean_x2_r,
 * mean_x^2, mean_y^2, 4 mean_x2, 4 mean_x2 - 4 mean_x^2) to acc[c][0][.] and
 * q^2 to acc[c][1][.] (PA_ENS_QUANTITIES each) and count the sample; while
 * counts[c][1] < pdt_cap, add each pupil's eta to its PA_PDT_BINS-bin integer
 * histogram hist[c][p][.] over [0, 1] (numpy.histogram binning: out-of-range
 * values dropped, eta = 1 in the last bin) and count the sample.  Caps < 0:
 * unlimited.  Elements are taken in batch order (deterministic sums); the
 * caller zeroes acc, counts and hist once before the first batch. */
int pa_ensemble_accumulate(int batch, const double* moments, const int* slot_channel, int channels, int npup,
                           int beam_cap, int pdt_cap, double* acc, long long* counts, uint64_t* hist,
                           void* stream);

/* pa_field_partials followed by pa_moments_finalize without the caller's
 * partials buffer: no counterpart in pyAtmosphere.  For callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops): the
 * per-tile partials go to a fixed 8 MiB device array inside the library (one
 * per GPU, held until the process ends), so the call allocates nothing; u,
 * chparam, pup, pup_moments, chan and out[b][PA_NUM_MOMENTS] are as for
 * pa_field_partials / pa_moments_finalize, and the results are identical.
 * batch must not exceed pa_field_moments_static_max(n) (fields per call: 146
 * at n = 2048, 36 at 4096), else PA_ERR_SIZE.  Calls on one GPU share its
 * array: calls that overlap on one GPU (other streams or threads) race. */
int pa_field_moments_static(int n, const void* u, const double* chparam, const double* pup, int npup,
                            const double* pup_moments, const int* chan, double* out, int batch, void* stream);
int pa_field_moments_static_max(int n);

/* Device Philox4x32-10 spectral draw for screen `screen` of realizations
 * rid[b] (grids.py:98-107; phase_screens.py:93-106): jitter, theta and the
 * complex Gaussian amplitudes times sqrt(psd).  base, sqrt_psd: [chan][P]. */
int pa_spectral_draw(int points, const float* base, const float* sqrt_psd, const int* chan,
                     const uint64_t* rid, uint64_t seed, int screen, float* rho, float* theta,
                     void* value, int batch, void* stream);

/* Screen synthesis (phase_screens.py:108-136, 25-28): plane-wave tables with
 * float64 argument reduction, the split 1-D parts in float64 and the 2-D
 * remainder as a GEMM on MFMA.
 *   prep  : fxy[b][P][4] (float64: fx, fy and the amplitude a' = a
 *           exp(2 pi i (fx x0 + fy y0)) that carries the channel's screen shift
 *           (x0, y0) = chparam[1..2], so that the split is taken about the grid
 *           centre), sparam[b][PA_SCREEN_PARAMS]
 *   tables: operands A, B (pa_operand_bytes each per element; B also holds
 *           the polynomial block's partial sums until it is written),
 *           hg[b][2][n] float64 (h(r), g(c)), esc[b][2][n] complex64.
 *           Points whose phase changes by at most 10 rad between the grid
 *           centre and edge on both axes (n2 of them) enter the 2-D
 *           remainder as sum_{n,m < PA_LOWRANK_TERMS} T_n(s_r) T_m(t_c) Re G[n][m]
 *           (T: Chebyshev polynomials; s, t: row and column offsets from the
 *           centre over n/2; Jacobi-Anger expansion, truncated below 1e-13
 *           relative), and points that do so along the columns only (up to
 *           64 more) as sum_m T_m(t_c) Re V_m(r), V_m(r) = sum_k (a' q)_k(r)
 *           P_km, in the same 32 K columns [kbeg, kbeg + 32), kbeg = 2 d0 -
 *           32, instead of 2 columns each, with their shares of h and g as
 *           Chebyshev sums or plane-wave sums.  To that end the call groups
 *           the points of fxy in place (a stable permutation: the d0 block
 *           points first, d0 a multiple of 16, then the rest); below 32 block
 *           points kbeg = d0 = 0, every point is direct and fxy is unchanged.
 *   gemm  : phi2 = A B^T over K columns [kbeg, Kp), each output nullable and
 *           in the work layout: the real float32 phi2[b]; the float32 screen
 *           angle ang[b](r,c) = phi2 + (ah[r] + ag[c]), ah and ag the angles
 *           of esc[b][0][r] = exp(-i (phi0 + h(r))) and esc[b][1][c] =
 *           exp(-i g(c)) in [-pi, pi], so exp(-i ang) = exp(-i phi) (what
 *           pa_row_mid applies); the complex screen factor fac[b](r,c) =
 *           esc[b][0][r] esc[b][1][c] exp(-i phi2)
 * Launch shapes (tuning only; results do not depend on them): `rows` = rows
 * per table block (4, 8, 16, 32; 0 = the library's choice for n and batch);
 * `tile` = GEMM block tile (1: 256 x 256 with 8 waves, n >= 512; 2: 128 x 128
 * with 4 waves; 0 = the library's choice). */
int pa_screen_prep(int points, const float* rho, const float* theta, const void* value,
                   double* fxy, double* sparam, int mode, const double* chparam, const int* chan,
                   const int* nsteps, int step, int batch, void* stream);
/* pa_spectral_draw followed by pa_screen_prep in one launch (grids.py:98-107;
 * phase_screens.py:93-113): the same fxy and sparam bit for bit, without
 * storing rho, theta and value; elements with step >= nsteps[b] are skipped. */
int pa_screen_draw_prep(int points, const float* base, const float* sqrt_psd, const int* chan,
                        const uint64_t* rid, uint64_t seed, int screen, double* fxy, double* sparam, int mode,
                        const double* chparam, const int* nsteps, int step, int batch, void* stream);
int pa_screen_tables(int n, int points, double* fxy, double* sparam, const double* chparam,
                     const int* chan, void* opa, void* opb, double* hg, void* esc, int mode, int rows,
                     const int* nsteps, int step, int batch, void* stream);
int pa_screen_gemm(int n, int points, const void* opa, const void* opb, const double* sparam,
                   const void* esc, float* phi2, float* ang, void* fac, int mode, int tile,
                   const int* nsteps, int step, int batch, void* stream);
/* Full screen phi = phi0 + h(r) + g(c) + phi2 as a row-major float32 array
 * (the value PhaseScreen.generate returns, phase_screens.py:25-28). */
int pa_screen_compose(int n, const double* sparam, const double* hg, const float* phi2,
                      float* phi, int batch, void* stream);

/* FFT phase screen with subharmonics (FFTPhaseScreen.generate_phase_screen,
 * phase_screens.py:42-67): from the caller's coefficients coeff[b] (n x n
 * complex64, row-major on the centred f-grid, as get_cn_coefficients draws
 * them) the centred inverse transform ifft2(cn, 1) (utils.py:47-50) in
 * `work` (n x n complex64 per element, scratch), then the subharmonic plane
 * waves as nsub separable float64 terms sum_j ey[j](r) g[j](c) with sub[b] =
 * (ey[0..nsub), g[0..nsub)) of n complex128 each, minus mean[b] (complex128,
 * the screen's mean over the grid).  Outputs (nullable, not all null): the
 * real screen phi[b] and its imaginary part phi_im[b] (row-major float32), and
 * the real screen as a float32 screen angle in the work layout for
 * pa_row_mid. */
int pa_fft_screen(int n, const void* coeff, void* work, const void* sub, int nsub, const void* mean, const void* tw,
                  float* phi, float* phi_im, float* ang, int batch, void* stream);

/* Row structure function of float32 screens phi[b][n][n] (row-major, as
 * pa_screen_compose writes them): out[b][r - 1] = the mean over rows of
 * sum_{c < n - r} (phi[c] - phi[c + r])^2 / (n - r), r = 1 .. n - 1
 * (theory/phase_screens/sf.py:4-8 averaged over rows as in
 * simulations/structure_function.py:11-12); float32 squares summed per
 * (row, lag), float64 over rows.  scratch: pa_structure_function_scratch(n,
 * batch) bytes of device memory. */
long long pa_structure_function_scratch(int n, int batch);
int pa_structure_function(int n, const float* phi, double* scratch, double* out, int batch, void* stream);

/* u[b] *= mask(pupil) for row-major fields (pupils.py:8-13). */
int pa_apply_pupil(int n, void* u, double delta, double radius, double shift_x, double shift_y,
                   int batch, void* stream);

/* out[b][r][c] = gy[b][r] * gx[b][c]: the separable Gaussian source field
 * (theory/sources.py:16-18 on RectGrid.get_rho2, grids.py:74-76). */
int pa_outer(int n, void* out, const void* gy, const void* gx, int batch, void* stream);

/* out[i] = |u[i]|^2 (measures.py:1-4). */
int pa_intensity(long long count, const void* u, float* out, void* stream);

/* HBM streaming over `bytes` (a multiple of 16) in 16-byte vectors, `unroll`
 * (1, 2, 4, 8, 16) vectors in flight per thread, optionally non-temporal:
 * kind 0 copies src to dst, 1 reads src (dst: a 4-byte sink), 2 fills dst.
 * No pyAtmosphere counterpart: the measured bandwidth ceiling that the field
 * passes are compared with. */
int pa_hbm_stream(int kind, const void* src, void* dst, long long bytes, int unroll, int nontemporal,
                  void* stream);

/* ---- channel plans: one call per batch (PhaseScreensPath.generator followed
 * by the measures, pathes.py:61-75 and measures.py:1-38, for SSPhaseScreen
 * screens of one channel).  The plan computes the channel's tables on the
 * host in float64 (Gaussian source, theory/sources.py:16-18; transfer
 * function per Fresnel step, theory/vacuum.py:5-7, identity for zero-length
 * steps, pathes.py:30-40; the amplitude gain of each step, nullable = 1) and
 * allocates them and a workspace for max_batch realizations: the only entry
 * points that allocate device memory (create allocates, destroy frees, run
 * allocates nothing).  Host arrays: positions[nscreens] (pathes.py:80-89),
 * gains[nscreens + 1], base and psd[points] (float32, the log-polar radii and
 * annulus variances, grids.py:94-96 and phase_screens.py:82-91),
 * pupils[npup][3] = (radius, shift_x, shift_y).  A plan is used from one
 * stream at a time. */
typedef struct pa_channel pa_channel;
int pa_channel_create(pa_channel** out, int n, double delta, double wvl, double w0, double F0, double length,
                      int nscreens, const double* positions, const double* gains, int points, const float* base,
                      const float* psd, int npup, const double* pupils, double shift_x, double shift_y,
                      int max_batch, int mode, void* stream);
/* Propagate `batch` realizations.  Spectra: device arrays rho, theta (float32)
 * and value (complex64), each [nscreens][batch][points] as drawn on the host
 * in the reference order (grids.py:98-107, phase_screens.py:93-106), or all
 * null for device Philox draws keyed by (seed, rid[b], screen) (rid: device
 * [batch]).  field_in (nullable = the Gaussian source) and field_out
 * (nullable) are row-major device fields [batch][n][n]; moments (nullable)
 * receives [batch][PA_NUM_MOMENTS] as pa_moments_finalize. */
int pa_channel_run(pa_channel* plan, int batch, const float* rho, const float* theta, const void* value,
                   const uint64_t* rid, uint64_t seed, const void* field_in, void* field_out, double* moments,
                   void* stream);
int pa_channel_destroy(pa_channel* plan);

/* ---- transmittance-PDT surrogate (no pyAtmosphere counterpart: the learned
 * model that maps channel parameters to the PDT that simulations/pdt.py
 * samples).  Row-major float32 unless noted. */
/* C = alpha op(A) op(B) + beta C; op(A) M x K (A is K x M when ta), op(B) K x N
 * (B is N x K when tb). */
int pa_mlp_gemm_f32(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb, int tb,
                    float* C, int ldc, float alpha, float beta, void* stream);
/* Z += bias (per column); Y = SiLU(Z) when act, else Y = Z. */
int pa_mlp_bias_act(int M, int N, float* Z, const float* bias, float* Y, int act, void* stream);
/* dZ = dY * SiLU'(Z). */
int pa_mlp_act_backward(int M, int N, const float* Z, const float* dY, float* dZ, void* stream);
/* db = column sums of dZ. */
int pa_mlp_bias_grad(int M, int N, const float* dZ, float* db, void* stream);
/* Per row of out[M][NB+1] (NB <= 64): softmax cross-entropy over NB bins +
 * wc * squared distance between predicted and target CDFs + wv * (log-variance
 * error)^2; per-row loss and/or gradient (divided by M). */
int pa_mlp_head_loss(int M, int NB, const float* out, const float* target, const float* tvar, float wv, float wc,
                     float* loss, float* dout, void* stream);
/* Adam step t (1-based) with bias correction. */
int pa_mlp_adam(int n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                int t, void* stream);
/* float32 -> bf16 (round to nearest even); y holds n uint16. */
int pa_mlp_to_bf16(long long n, const float* x, void* y, void* stream);
/* Y = act(X W^T + bias) on bf16 MFMA with float32 accumulation; W bf16 [N][K];
 * K multiple of 16, N multiple of 32. */
int pa_mlp_layer_bf16(int M, int N, int K, const float* X, const void* W, const float* bias, float* Y, int act,
                      void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PA_HIP_H */
