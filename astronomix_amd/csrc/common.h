// Device helpers and host error plumbing shared by the finite-volume and
// finite-difference translation units of libamx.so.
#pragma once

#include <hip/hip_runtime.h>

#include "amx.h"

namespace amx {

constexpr int NSLOT = 64;  // partial-max slots of the block reductions

__device__ __forceinline__ int wrapi(int i, int n) {
  while (i < 0) i += n;
  while (i >= n) i -= n;
  return i;
}

// Hardware reciprocal / square root (v_rcp_f32, v_sqrt_f32: <= 1 ulp) for
// data-dependent divides and roots in the stencil kernels.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ int swizzle_block(int bid, int nblk) {
  // consecutive logical tiles on the same XCD (8 XCDs, round-robin dispatch)
  if ((nblk & 7) == 0) return (bid & 7) * (nblk >> 3) + (bid >> 3);
  return bid;
}

__device__ __forceinline__ unsigned canon_bits(float m) {
  return m != m ? 0x7fc00000u : __float_as_uint(m);
}

__device__ __forceinline__ unsigned wave_max_u(unsigned v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned w = (unsigned)__shfl_xor((int)v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// Max over the block of one value per thread; the result is valid in thread 0.
// Every thread of the block must call it (it contains a barrier).
template <int NW>
__device__ __forceinline__ unsigned block_max_u(unsigned v, unsigned* sred) {
  v = wave_max_u(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sred[w] = v;
  __syncthreads();
  unsigned m = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NW; ++i) m = sred[i] > m ? sred[i] : m;
  }
  return m;
}

__device__ __forceinline__ unsigned slots_max(const unsigned* p) {
  unsigned m = 0;
  for (int i = 0; i < NSLOT; ++i) m = p[i] > m ? p[i] : m;
  return m;
}

}  // namespace amx

// Last error of the calling thread (defined in fv_mhd.hip); returns s.
amx_status amx_fail(amx_status s, const char* fmt, const char* what = "");

#define AMX_HIPCHK(expr)                                                                         \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) return amx_fail(AMX_ERR_HIP, "HIP error: %s", hipGetErrorString(_e));  \
  } while (0)
