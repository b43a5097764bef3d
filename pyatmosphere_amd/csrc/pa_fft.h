// In-register Stockham FFT machinery for CDNA4 (gfx950).
//
// One "line" (a row or a column of the N x N field) of N complex64 points is
// spread over L = N/E threads, each thread holding E points in registers in
// the cyclic distribution v[e] = x[l + L*e].  Every radix pass of the
// Stockham autosort algorithm reads its butterfly inputs in exactly that
// cyclic distribution, so the first pass needs no exchange after a coalesced
// load, the last pass leaves the result in the same distribution for a
// coalesced store, and an inverse transform followed by a pointwise multiply
// and a forward transform needs no exchange at all between the two.
// Intermediate passes exchange through a padded LDS line buffer.
//
// Radix plans (E = elements per thread, 16 unless noted):
//   64: 8x8 (E=8)        128: 16x8       256: 16x16
//   512: 8x8x8 (E=8)     1024: 16x8x8    2048: 16x16x8    4096: 16x16x16
// Blocks stay at <= 512 threads so a kernel may use 128 VGPRs at 4 waves/SIMD.
#pragma once
#include <hip/hip_runtime.h>

namespace pa {

__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// a * (DIR * i): DIR = -1 is the forward kernel e^{-2 pi i nk/N}.
template <int DIR>
__host__ __device__ __forceinline__ float2 rot(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int DIR>
__host__ __device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__host__ __device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
  float2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
  float2 t2 = cadd(v1, v3), t3 = rot<DIR>(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

// 8-point DFT as two 4-point DFTs and the W8 combine.
template <int DIR>
__host__ __device__ __forceinline__ void dft8(float2 (&v)[8]) {
  constexpr float c = 0.70710678118654752440f;
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // W8^1 = (c, DIR c), W8^2 = DIR i, W8^3 = (-c, DIR c)
  float2 w1 = make_float2(c * (o1.x - DIR * o1.y), c * (o1.y + DIR * o1.x));
  float2 w2 = rot<DIR>(o2);
  float2 w3 = make_float2(c * (-o3.x - DIR * o3.y), c * (-o3.y + DIR * o3.x));
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1);
  v[5] = csub(e1, w1);
  v[2] = cadd(e2, w2);
  v[6] = csub(e2, w2);
  v[3] = cadd(e3, w3);
  v[7] = csub(e3, w3);
}

// W16^e for e in [0, 16): cos(2 pi e/16), DIR * sin(2 pi e/16) with DIR folded by caller.
__host__ __device__ __forceinline__ constexpr float cos16(int e) {
  return e == 0 ? 1.0f : e == 1 ? 0.92387953251128675613f : e == 2 ? 0.70710678118654752440f
       : e == 3 ? 0.38268343236508977173f : e == 4 ? 0.0f : e == 5 ? -0.38268343236508977173f
       : e == 6 ? -0.70710678118654752440f : e == 7 ? -0.92387953251128675613f : e == 8 ? -1.0f
       : e == 9 ? -0.92387953251128675613f : 0.0f;
}
__host__ __device__ __forceinline__ constexpr float sin16(int e) {
  return e == 0 ? 0.0f : e == 1 ? 0.38268343236508977173f : e == 2 ? 0.70710678118654752440f
       : e == 3 ? 0.92387953251128675613f : e == 4 ? 1.0f : e == 5 ? 0.92387953251128675613f
       : e == 6 ? 0.70710678118654752440f : e == 7 ? 0.38268343236508977173f : e == 8 ? 0.0f
       : e == 9 ? -0.38268343236508977173f : 0.0f;
}

template <int DIR, int EXP>
__host__ __device__ __forceinline__ float2 tw16(float2 a) {
  if constexpr (EXP == 0) {
    return a;
  } else if constexpr (EXP == 4) {
    return rot<DIR>(a);
  } else {
    constexpr float wc = cos16(EXP);
    constexpr float ws = DIR * sin16(EXP);
    return make_float2(a.x * wc - a.y * ws, a.x * ws + a.y * wc);
  }
}

// 16-point DFT as 4x4: n = 4 n1 + n2, k = k1 + 4 k2.
template <int DIR>
__host__ __device__ __forceinline__ void dft16(float2 (&v)[16]) {
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<DIR>(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
  // v[n2 + 4 k1] now holds Y[n2][k1]; twiddle by W16^{n2 k1}
  v[5] = tw16<DIR, 1>(v[5]);
  v[9] = tw16<DIR, 2>(v[9]);
  v[13] = tw16<DIR, 3>(v[13]);
  v[6] = tw16<DIR, 2>(v[6]);
  v[10] = tw16<DIR, 4>(v[10]);
  v[14] = tw16<DIR, 6>(v[14]);
  v[7] = tw16<DIR, 3>(v[7]);
  v[11] = tw16<DIR, 6>(v[11]);
  v[15] = tw16<DIR, 9>(v[15]);
  float2 o[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 a = v[4 * k1 + 0], b = v[4 * k1 + 1], c = v[4 * k1 + 2], d = v[4 * k1 + 3];
    dft4<DIR>(a, b, c, d);
    o[k1] = a;
    o[k1 + 4] = b;
    o[k1 + 8] = c;
    o[k1 + 12] = d;
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = o[i];
}

template <int R, int DIR>
__host__ __device__ __forceinline__ void dft(float2 (&u)[R]) {
  if constexpr (R == 2) {
    dft2<DIR>(u[0], u[1]);
  } else if constexpr (R == 4) {
    dft4<DIR>(u[0], u[1], u[2], u[3]);
  } else if constexpr (R == 8) {
    dft8<DIR>(u);
  } else {
    static_assert(R == 16, "radix");
    dft16<DIR>(u);
  }
}

// ------------------------------------------------------------------ plans
template <int N>
struct Plan;
// E: points per thread; R0..R2: radices (0 = unused); T: rows per tile of the
// field layout and per row-pass block; C: columns per column-pass block.
template <> struct Plan<64>   { static constexpr int E = 8,  NP = 2, R0 = 8,  R1 = 8,  R2 = 0,  T = 8,  C = 16; };
template <> struct Plan<128>  { static constexpr int E = 16, NP = 2, R0 = 16, R1 = 8,  R2 = 0,  T = 8,  C = 16; };
template <> struct Plan<256>  { static constexpr int E = 16, NP = 2, R0 = 16, R1 = 16, R2 = 0,  T = 16, C = 16; };
template <> struct Plan<512>  { static constexpr int E = 8,  NP = 3, R0 = 8,  R1 = 8,  R2 = 8,  T = 8,  C = 8; };
template <> struct Plan<1024> { static constexpr int E = 16, NP = 3, R0 = 16, R1 = 8,  R2 = 8,  T = 8,  C = 8; };
template <> struct Plan<2048> { static constexpr int E = 16, NP = 3, R0 = 16, R1 = 16, R2 = 8,  T = 4,  C = 4; };
template <> struct Plan<4096> { static constexpr int E = 16, NP = 3, R0 = 16, R1 = 16, R2 = 16, T = 2,  C = 2; };

template <int N>
struct Geo {
  static constexpr int E = Plan<N>::E;
  static constexpr int L = N / E;                 // threads per line
  static constexpr int T = Plan<N>::T;
  static constexpr int C = Plan<N>::C;
  // Padded LDS lines: one pad element per 16, then a line stride chosen so that
  // the lines whose threads share a 16-lane LDS phase land on different banks.
  // A 16-lane phase of a row pass holds T lines x 16/T lanes (line stride
  // 16/T mod 16 complex elements); of a column pass 16/T lines x T lanes
  // (stride T mod 16).
  static constexpr int PADN = N + N / 16;
  static constexpr int LINE_BASE = ((PADN + 31) / 32) * 32;
  static constexpr int LINE_ROW = LINE_BASE + (16 / T) % 16;
  static constexpr int LINE_COL = LINE_BASE + T % 16;
  static constexpr int ROW_THREADS = T * L;
  static constexpr int COL_THREADS = C * L;
  // the thread maps of the row and column passes need T | L and whole waves
  static_assert(L % T == 0, "tile height must divide the threads per line");
  static_assert(N % T == 0 && N % C == 0, "tiles must divide the grid");
  static_assert(ROW_THREADS % 64 == 0 && ROW_THREADS <= 1024, "row block size");
  static_assert(COL_THREADS % 64 == 0 && COL_THREADS <= 1024, "column block size");
  static_assert(T == 2 || T % 4 == 0, "the GEMM epilogue stores 4 (or 2) rows of a tile at once");
};

__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }

// sin and cos of a float32 angle x: Cody-Waite reduction to r in [-pi/4, pi/4]
// with pi/2 split in three float32 parts applied by FMA, quadrant q, then
// Taylor polynomials through r^9 (sin) and r^10 (cos), whose truncation error at
// |r| = pi/4 is below 2e-9.  Absolute error < 1e-7 for |x| <= 4096.
// Returns (cos x, -sin x) = e^{-ix}.
__device__ __forceinline__ float2 expmi(float x) {
  const float q = rintf(x * 0.63661977236758134308f);          // 2/pi
  float r = fmaf(-q, 1.57079637050628662109375f, x);           // pi/2 = p1 + p2 + p3
  r = fmaf(-q, -4.37113900018624283e-8f, r);
  r = fmaf(-q, -1.7763568394002504647e-15f, r);
  const float r2 = r * r;
  float s = fmaf(r2, 2.7557319e-6f, -1.9841270e-4f);           // 1/9!, -1/7!
  s = fmaf(r2, s, 8.3333333e-3f);                              // 1/5!
  s = fmaf(r2, s, -1.6666667e-1f);                             // -1/3!
  s = fmaf(r2 * r, s, r);
  float c = fmaf(r2, -2.7557319e-7f, 2.4801587e-5f);           // -1/10!, 1/8!
  c = fmaf(r2, c, -1.3888889e-3f);                             // -1/6!
  c = fmaf(r2, c, 4.1666668e-2f);                              // 1/4!
  c = fmaf(r2, c, -0.5f);
  c = fmaf(r2, c, 1.0f);
  const int qi = (int)q & 3;
  const float sn = qi == 0 ? s : qi == 1 ? c : qi == 2 ? -s : -c;
  const float cs = qi == 0 ? c : qi == 1 ? -s : qi == 2 ? -c : s;
  return make_float2(cs, -sn);
}

// e^{-ix} as the row pass forms the screen factor: Cody-Waite reduction by
// 2 pi (split in two float32 parts applied by FMA) to r in [-pi, pi], then the
// hardware sine and cosine of r / (2 pi) revolutions.  Measured on gfx950
// against float64: max |error| 4.0e-7, rms 1.1e-7 for |x| <= 512 (the 2-D
// screen remainder is at most ~50 rad).
__device__ __forceinline__ float2 expmi_fast(float x) {
  const float q = rintf(x * 0.15915494309189535f);          // 1 / (2 pi)
  float r = fmaf(-q, 6.28318548202514648f, x);               // 2 pi = hi + lo
  r = fmaf(-q, -1.7484555314695172e-7f, r);
  const float t = r * 0.15915494309189535f;
  return make_float2(__builtin_amdgcn_cosf(t), -__builtin_amdgcn_sinf(t));
}

// Twiddles W^k = e^{-2 pi i k/N} are read from a two-level table in LDS:
// tw[k & 63] * tw[TW_LO + (k >> 6)], i.e. W^(k mod 64) and W^(64 (k div 64)),
// each the float32 rounding of a float64 value (one extra rounding per twiddle).
// The second pass (NS = R0, radix R1) has only R0 (R1 - 1) distinct twiddles,
// W^{jm r N/(R0 R1)} for jm < R0, 0 < r < R1.  With PASS2 they follow the
// two-level part as a table of their own, formed once per block by the same
// lookups and products the other passes use per butterfly group
// (pass_twiddle), and the second pass reads each twiddle instead of forming
// it.  The persistent last row pass uses it (its blocks form the table once
// for many lines); in the other passes it did not pay, so they keep the
// per-group twiddles.
constexpr int TW_LO = 64;
template <int N, bool PASS2 = false>
struct TwTable {
  static constexpr int LO = 64 + (N + 63) / 64;
  static constexpr int NS2 = Plan<N>::R0, R2 = Plan<N>::R1;
  static constexpr int SIZE = LO + (PASS2 ? NS2 * (R2 - 1) : 0);
};

// W^{k r} (DIR < 0) from the full table tw: W^k = W^{64 (k div 64)} W^{k mod 64}
// and W^{4k} likewise, then W^{k r} = W^{4k (r div 4)} W^{k (r mod 4)} with the
// powers formed by complex products.  4k < N for the radices used (R >= 4).
template <int R>
__device__ __forceinline__ float2 pass_twiddle(const float2* __restrict__ tw, int k, int r) {
  float2 wb[4], wa[4];
  wb[1] = cmul(tw[64 * (k >> 6)], tw[k & 63]);
  wb[2] = cmul(wb[1], wb[1]);
  wb[3] = cmul(wb[2], wb[1]);
  const int k4 = 4 * k;
  wa[1] = cmul(tw[64 * (k4 >> 6)], tw[k4 & 63]);
  wa[2] = cmul(wa[1], wa[1]);
  wa[3] = cmul(wa[2], wa[1]);
  // r is a run-time value: select rather than index (an indexed private array
  // would live in scratch memory)
  const int rb = r & 3, ra = r >> 2;
  const float2 b = rb == 1 ? wb[1] : rb == 2 ? wb[2] : wb[3];
  const float2 a = ra == 1 ? wa[1] : ra == 2 ? wa[2] : wa[3];
  return ra == 0 ? b : rb == 0 ? a : cmul(a, b);
}

// Fill the block's LDS twiddle table from the full float64-rounded table.
template <int N, bool PASS2 = false>
__device__ __forceinline__ void load_twiddles(float2* s_tw, const float2* __restrict__ tw, int tid, int nthreads) {
  using TT = TwTable<N, PASS2>;
  for (int i = tid; i < TT::SIZE; i += nthreads) {
    if (i < TT::LO) {
      const int k = i < 64 ? i : 64 * (i - 64);
      s_tw[i] = k < N ? tw[k] : make_float2(1.f, 0.f);
    } else {
      const int q = i - TT::LO, jm = q / (TT::R2 - 1), r = q - jm * (TT::R2 - 1) + 1;
      s_tw[i] = pass_twiddle<TT::R2>(tw, jm * (N / (TT::NS2 * TT::R2)), r);
    }
  }
}

// One Stockham pass with radix R over a line held cyclically in v.
// NS: product of the radices already applied. LAST: leave output in registers.
// tw: the block's LDS twiddle table (load_twiddles).  PASS2: the second pass
// reads its twiddles from the block's table (TwTable<N, true>,
// load_twiddles<N, true>) instead of forming them per butterfly group.
template <int N, int R, int NS, int DIR, bool LAST, bool PASS2 = false>
__device__ __forceinline__ void stockham_pass(float2 (&v)[Geo<N>::E], float2* lds, int l,
                                              const float2* __restrict__ tw) {
  constexpr int E = Geo<N>::E;
  constexpr int L = Geo<N>::L;
  constexpr int M = E / R;
  // an opaque copy of the lane index: keeps the compiler from carrying twiddles
  // or addresses of this pass over to a later transform (register pressure)
  asm volatile("" : "+v"(l));
#pragma unroll
  for (int m = 0; m < M; ++m) {
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[m + r * M];
    const int j = l + L * m;
    const int jm = j & (NS - 1);
    if constexpr (PASS2 && NS > 1 && NS == TwTable<N>::NS2) {
      // the second pass: twiddles from the block's table (conjugated for the
      // inverse transform)
      const float2* t2 = tw + TwTable<N>::LO + jm * (R - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = t2[r - 1];
        if constexpr (DIR > 0) w.y = -w.y;
        u[r] = cmul(u[r], w);
      }
    } else if constexpr (NS > 1) {
      // twiddles W^{k r}, r < R, from two table lookups (W^k and W^{4k}) and
      // products: W^{k r} = W^{4k (r / 4)} W^{k (r % 4)}, each at most three
      // roundings from its exact value
      const int k = jm * (N / (NS * R));
      float2 wb[4], wa[R > 4 ? R / 4 : 1];
      wb[1] = cmul(tw[TW_LO + (k >> 6)], tw[k & 63]);
      if constexpr (DIR > 0) wb[1].y = -wb[1].y;
      wb[2] = cmul(wb[1], wb[1]);
      wb[3] = cmul(wb[2], wb[1]);
      if constexpr (R > 4) {
        const int k4 = 4 * k;
        wa[1] = cmul(tw[TW_LO + (k4 >> 6)], tw[k4 & 63]);
        if constexpr (DIR > 0) wa[1].y = -wa[1].y;
        if constexpr (R > 8) {
          wa[2] = cmul(wa[1], wa[1]);
          wa[3] = cmul(wa[2], wa[1]);
        }
      }
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w;
        if (r < 4) w = wb[r];
        else if ((r & 3) == 0) w = wa[r >> 2];
        else w = cmul(wa[r >> 2], wb[r & 3]);
        u[r] = cmul(u[r], w);
      }
    }
    dft<R, DIR>(u);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[m + r * M] = u[r];
    } else {
      // lpad(base + r NS) = lpad(base) + lpad(r NS): base mod 16 plus
      // r NS mod 16 stays below 16 (base mod 16 = jm < NS when NS divides 16
      // and 16 divides NS R; base is a multiple of 8 when NS = 1, R = 8;
      // r NS mod 16 = 0 when 16 divides NS), so the pad offsets are
      // constants and the R stores share one address register
      static_assert(NS % 16 == 0 || (16 % NS == 0 && ((NS * R) % 16 == 0 || (NS == 1 && R == 8))),
                    "padded store offsets");
      const unsigned ju = (unsigned)j;
      const int pb = lpad((int)((ju / NS) * NS * R + (ju & (NS - 1))));
#pragma unroll
      for (int r = 0; r < R; ++r) lds[pb + lpad(r * NS)] = u[r];
    }
  }
  if constexpr (!LAST) {
    // likewise lpad(l + L e) = lpad(l) + lpad(L e) for l < L (16 divides L,
    // or L divides 16): one address register, constant offsets
    static_assert(L % 16 == 0 || 16 % L == 0, "padded load offsets");
    const int pl = lpad(l);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = lds[pl + lpad(L * e)];
    __syncthreads();
  }
}

// Full unnormalised transform of one line (all lines of a block in lockstep).
// PASS2: tw is a TwTable<N, true> filled by load_twiddles<N, true>.
template <int N, int DIR, bool PASS2 = false>
__device__ __forceinline__ void fft_line(float2 (&v)[Geo<N>::E], float2* lds, int l,
                                         const float2* __restrict__ tw) {
  using P = Plan<N>;
  if constexpr (P::NP == 2) {
    stockham_pass<N, P::R0, 1, DIR, false>(v, lds, l, tw);
    stockham_pass<N, P::R1, P::R0, DIR, true, PASS2>(v, lds, l, tw);
  } else {
    stockham_pass<N, P::R0, 1, DIR, false>(v, lds, l, tw);
    stockham_pass<N, P::R1, P::R0, DIR, false, PASS2>(v, lds, l, tw);
    stockham_pass<N, P::R2, P::R0 * P::R1, DIR, true>(v, lds, l, tw);
  }
}

// Buffer access to one element's N x N field: a buffer resource over the
// field (SGPRs) plus one lane offset (one VGPR); the E points of a thread sit
// at constant strides that go into the instruction's SGPR offset, so a line's
// loads and stores need no per-point address registers.  AUX 2 marks the
// access non-temporal (the field is streamed once per pass).
typedef unsigned int pa_v2u __attribute__((ext_vector_type(2)));
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kBufNT = 2;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
template <int AUX>
__device__ __forceinline__ float2 buf_ld2(rsrc_t r, int voff, int soff) {
  const pa_v2u x = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX);
  return make_float2(__uint_as_float(x.x), __uint_as_float(x.y));
}
template <int AUX>
__device__ __forceinline__ float buf_ld1(rsrc_t r, int voff, int soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
}
template <int AUX>
__device__ __forceinline__ void buf_st2(float2 v, rsrc_t r, int voff, int soff) {
  pa_v2u x;
  x.x = __float_as_uint(v.x);
  x.y = __float_as_uint(v.y);
  __builtin_amdgcn_raw_buffer_store_b64(x, r, voff, soff, AUX);
}

// Field layout in HBM between passes ("row-tiled"): T consecutive rows form a
// tile stored column-interleaved, element (r, c) at ((r/T)*N + c)*T + r%T.
// A row pass block owns one tile (one contiguous T*N run); a column pass block
// owns C columns, i.e. one contiguous T*C run per tile.
template <int N>
__device__ __forceinline__ size_t tiled(int r, int c) {
  constexpr int T = Geo<N>::T;
  return ((size_t)(r / T) * N + c) * T + (r % T);
}

}  // namespace pa
