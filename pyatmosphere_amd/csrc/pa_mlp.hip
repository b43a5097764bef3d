// Kernels of the transmittance-PDT surrogate: a small MLP trained and evaluated
// on the device.  Training runs in float32 (LDS-tiled GEMMs on the vector
// units, which are exact enough for gradient checks against a float64 host
// MLP); inference GEMMs run on bf16 MFMA with float32 accumulation and fused
// bias + activation.  Activation: SiLU, z * sigmoid(z).  Output head: NB
// logits (softmax PDT over NB transmittance bins) and one log-variance.
#include "pa_internal.h"

namespace pa {

typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ fp32 GEMM
// C[M][N] = alpha * op(A) op(B) + beta * C, row-major, op = transpose when t* != 0:
// op(A) is M x K (A is M x K, or K x M when ta), op(B) is K x N (B is K x N, or N x K when tb).
constexpr int GT = 64, GK = 16;

__global__ void __launch_bounds__(256)
k_gemm_f32(int M, int N, int K, const float* __restrict__ A, int lda, int ta, const float* __restrict__ B,
           int ldb, int tb, float* __restrict__ C, int ldc, float alpha, float beta) {
  __shared__ float sa[GK][GT + 4];
  __shared__ float sb[GK][GT + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int i = threadIdx.x; i < GK * GT; i += 256) {
      const int kk = i / GT, mm = i % GT;
      const int m = m0 + mm, k = k0 + kk;
      float va = 0.f;
      if (m < M && k < K) va = ta ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
      sa[kk][mm] = va;
      const int n = n0 + mm;
      float vb = 0.f;
      if (n < N && k < K) vb = tb ? B[(size_t)n * ldb + k] : B[(size_t)k * ldb + n];
      sb[kk][mm] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sa[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sb[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + ty * 4 + i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tx * 4 + j;
      if (n >= N) continue;
      float* c = C + (size_t)m * ldc + n;
      *c = alpha * acc[i][j] + (beta != 0.f ? beta * *c : 0.f);
    }
  }
}

// ------------------------------------------------------------------ elementwise
__device__ __forceinline__ float sigm(float z) { return 1.f / (1.f + expf(-z)); }

// Z += bias (row broadcast); Y = act(Z) when act != 0 else Y = Z.
__global__ void k_bias_act(int M, int N, float* __restrict__ Z, const float* __restrict__ bias,
                           float* __restrict__ Y, int act) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)M * N) return;
  const float z = Z[i] + bias[i % N];
  Z[i] = z;
  Y[i] = act ? z * sigm(z) : z;
}

// dZ = dY * act'(Z)
__global__ void k_act_backward(int M, int N, const float* __restrict__ Z, const float* __restrict__ dY,
                               float* __restrict__ dZ) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)M * N) return;
  const float z = Z[i], s = sigm(z);
  dZ[i] = dY[i] * s * (1.f + z * (1.f - s));
}

// db[n] = sum_m dZ[m][n] (fixed order per column)
__global__ void k_bias_grad(int M, int N, const float* __restrict__ dZ, float* __restrict__ db) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dZ[(size_t)m * N + n];
  db[n] = s;
}

// Head: row m of out holds NB logits and one log-variance.  Loss per row:
//   -sum_j t_j log p_j + wc sum_j (P_j - T_j)^2 + wv (lv - tv)^2
// with p = softmax(logits), P and T the cumulative sums of p and t (a ranked
// probability score, the squared distance between the CDFs the KS metric
// compares).  Gradient (divided by M): with g_i = 2 wc sum_{j >= i} (P_j - T_j),
// dL/dl_i = p_i - t_i + p_i (g_i - sum_k p_k g_k); dL/dlv = 2 wv (lv - tv).
// One wave per row; NB <= 64 (one bin per lane).
__global__ void __launch_bounds__(64)
k_head_loss(int M, int NB, const float* __restrict__ out, const float* __restrict__ target,
            const float* __restrict__ tvar, float wv, float wc, float* __restrict__ loss, float* __restrict__ dout) {
  const int m = blockIdx.x;
  const int j = threadIdx.x;
  const bool on = j < NB;
  const float* l = out + (size_t)m * (NB + 1);
  const float lj = on ? l[j] : -3.4e38f;
  float mx = lj;
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  const float ej = on ? expf(lj - mx) : 0.f;
  float se = ej;
  for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o, 64);
  const float lse = mx + logf(se);
  const float pj = ej / se;
  const float tj = on ? target[(size_t)m * NB + j] : 0.f;
  // inclusive prefix sums of p and t over the lanes
  float P = pj, T = tj;
  for (int o = 1; o < 64; o <<= 1) {
    const float a = __shfl_up(P, o, 64), b = __shfl_up(T, o, 64);
    if (j >= o) {
      P += a;
      T += b;
    }
  }
  const float dcdf = on ? P - T : 0.f;
  // suffix sums of (P - T)
  float G = dcdf;
  for (int o = 1; o < 64; o <<= 1) {
    const float a = __shfl_down(G, o, 64);
    if (j + o < 64) G += a;
  }
  const float gj = 2.f * wc * G;
  float pg = pj * gj;
  for (int o = 32; o > 0; o >>= 1) pg += __shfl_xor(pg, o, 64);
  float part = (on ? -tj * (lj - lse) : 0.f) + wc * dcdf * dcdf;
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
  if (on && dout) dout[(size_t)m * (NB + 1) + j] = (pj - tj + pj * (gj - pg)) / M;
  if (j == 0) {
    const float d = l[NB] - tvar[m];
    if (loss) loss[m] = part + wv * d * d;
    if (dout) dout[(size_t)m * (NB + 1) + NB] = 2.f * wv * d / M;
  }
}

// Adam with bias correction; w, m, v float32.
__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ x, unsigned short* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// ------------------------------------------------------------------ bf16 MFMA inference layer
// Y[M][N] = act(X[M][K] W[N][K]^T + bias), X float32 (rounded to bf16 on load),
// W bf16, fp32 accumulation.  One wave per 32x32 output tile, 4 waves per block
// covering 32 rows x 128 columns; K multiple of 16, N multiple of 32 (padded by caller).
__global__ void __launch_bounds__(256)
k_layer_bf16(int M, int N, int K, const float* __restrict__ X, const unsigned short* __restrict__ W,
             const float* __restrict__ bias, float* __restrict__ Y, int act) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 128 + wave * 32;
  if (n0 >= N) return;
  typedef float f16v __attribute__((ext_vector_type(16)));
  f16v acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  const int row = m0 + (lane & 31);
  const int col = n0 + (lane & 31);
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int kb = k0 + 8 * (lane >> 5);
    us8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      a[j] = row < M ? f2bf(X[(size_t)row * K + kb + j]) : (unsigned short)0;
      b[j] = W[(size_t)col * K + kb + j];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), acc,
                                                  0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int r = m0 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
    if (r >= M) continue;
    const float z = acc[v] + bias[col];
    Y[(size_t)r * N + col] = act ? z * sigm(z) : z;
  }
}

// ------------------------------------------------------------------ launchers
int mlp_gemm_f32(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb, int tb, float* C,
                 int ldc, float alpha, float beta, hipStream_t s) {
  dim3 g((N + GT - 1) / GT, (M + GT - 1) / GT);
  hipLaunchKernelGGL(k_gemm_f32, g, dim3(256), 0, s, M, N, K, A, lda, ta, B, ldb, tb, C, ldc, alpha, beta);
  return check_launch("mlp_gemm_f32");
}
int mlp_bias_act(int M, int N, float* Z, const float* bias, float* Y, int act, hipStream_t s) {
  hipLaunchKernelGGL(k_bias_act, dim3((unsigned)(((size_t)M * N + 255) / 256)), dim3(256), 0, s, M, N, Z, bias,
                     Y, act);
  return check_launch("mlp_bias_act");
}
int mlp_act_backward(int M, int N, const float* Z, const float* dY, float* dZ, hipStream_t s) {
  hipLaunchKernelGGL(k_act_backward, dim3((unsigned)(((size_t)M * N + 255) / 256)), dim3(256), 0, s, M, N, Z,
                     dY, dZ);
  return check_launch("mlp_act_backward");
}
int mlp_bias_grad(int M, int N, const float* dZ, float* db, hipStream_t s) {
  hipLaunchKernelGGL(k_bias_grad, dim3((N + 255) / 256), dim3(256), 0, s, M, N, dZ, db);
  return check_launch("mlp_bias_grad");
}
int mlp_head_loss(int M, int NB, const float* out, const float* target, const float* tvar, float wv, float wc,
                  float* loss, float* dout, hipStream_t s) {
  hipLaunchKernelGGL(k_head_loss, dim3(M), dim3(64), 0, s, M, NB, out, target, tvar, wv, wc, loss, dout);
  return check_launch("mlp_head_loss");
}
int mlp_adam(int n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
             int t, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, eps, c1, c2);
  return check_launch("mlp_adam");
}
int mlp_to_bf16(long long n, const float* x, unsigned short* y, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (size_t)n, x, y);
  return check_launch("mlp_to_bf16");
}
int mlp_layer_bf16(int M, int N, int K, const float* X, const unsigned short* W, const float* bias, float* Y,
                   int act, hipStream_t s) {
  dim3 g((N + 127) / 128, (M + 31) / 32);
  hipLaunchKernelGGL(k_layer_bf16, g, dim3(256), 0, s, M, N, K, X, W, bias, Y, act);
  return check_launch("mlp_layer_bf16");
}

}  // namespace pa
