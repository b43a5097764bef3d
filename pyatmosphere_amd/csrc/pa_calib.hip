// HBM streaming kernels that measure the practical bandwidth ceiling of the
// device, against which the field passes are judged: a copy (read + write),
// a read-only sum and a write-only fill, each over 16-byte vectors with U
// independent vectors per thread in flight.  Each block streams one
// contiguous chunk of blockDim * U vectors (wave-contiguous 1 KiB runs per
// vector slot), the grid covers the buffer once.  `nt` selects non-temporal
// loads and stores (the buffers are far larger than the 256 MB MALL).
#include "pa_internal.h"

namespace pa {

// a native 16-byte vector (the non-temporal builtins take clang vector types)
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_hbm_copy(const v4u* __restrict__ src, v4u* __restrict__ dst, size_t n) {
  const size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x;
  v4u v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t i = base + (size_t)u * 256;
    if (i < n) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t i = base + (size_t)u * 256;
    if (i < n) {
      if (NT) __builtin_nontemporal_store(v[u], dst + i);
      else dst[i] = v[u];
    }
  }
}

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_hbm_read(const v4u* __restrict__ src, size_t n, unsigned* __restrict__ sink) {
  const size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x;
  unsigned acc = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t i = base + (size_t)u * 256;
    if (i < n) {
      const v4u v = NT ? __builtin_nontemporal_load(src + i) : src[i];
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
  }
  // a store the compiler cannot drop, taken by (almost) no thread
  if (acc == 0x9e3779b9u) sink[0] = acc;
}

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_hbm_write(v4u* __restrict__ dst, size_t n, unsigned value) {
  const size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x;
  const v4u v = {value, value, value, value};
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t i = base + (size_t)u * 256;
    if (i < n) {
      if (NT) __builtin_nontemporal_store(v, dst + i);
      else dst[i] = v;
    }
  }
}

template <int U, bool NT>
static int launch_stream(int kind, const void* src, void* dst, size_t vecs, hipStream_t s) {
  const dim3 grid((unsigned)((vecs + 256 * U - 1) / (256 * U)));
  if (kind == 0)
    hipLaunchKernelGGL((k_hbm_copy<U, NT>), grid, dim3(256), 0, s, static_cast<const v4u*>(src),
                       static_cast<v4u*>(dst), vecs);
  else if (kind == 1)
    hipLaunchKernelGGL((k_hbm_read<U, NT>), grid, dim3(256), 0, s, static_cast<const v4u*>(src), vecs,
                       static_cast<unsigned*>(dst));
  else
    hipLaunchKernelGGL((k_hbm_write<U, NT>), grid, dim3(256), 0, s, static_cast<v4u*>(dst), vecs, 0x3f800000u);
  return check_launch("hbm_stream");
}

int hbm_stream(int kind, const void* src, void* dst, long long bytes, int unroll, int nt, hipStream_t s) {
  const size_t vecs = (size_t)bytes / 16;
#define PA_STREAM_CASE(UU)                                                  \
  case UU:                                                                  \
    return nt ? launch_stream<UU, true>(kind, src, dst, vecs, s)            \
              : launch_stream<UU, false>(kind, src, dst, vecs, s);
  switch (unroll) {
    PA_STREAM_CASE(1)
    PA_STREAM_CASE(2)
    PA_STREAM_CASE(4)
    PA_STREAM_CASE(8)
    PA_STREAM_CASE(16)
    default:
      return set_error(PA_ERR_ARG, "hbm_stream: unroll must be 1, 2, 4, 8 or 16");
  }
#undef PA_STREAM_CASE
}

}  // namespace pa
