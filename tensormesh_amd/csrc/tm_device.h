// Device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tm_common.h"

namespace tmx {

constexpr int kWave = 64;

// Consecutive workgroups are dispatched round-robin over the 8 XCDs of an
// MI355X.  Remap so that each XCD receives one contiguous range of logical
// block ids: neighbouring plan blocks share halo nodes, so they should share
// an L2.  Bijective for any n.
__device__ __forceinline__ int xcd_remap(int bid, int n) {
  constexpr int nx = 8;
  const int q = n / nx, r = n % nx;
  const int x = bid % nx, k = bid / nx;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Block-wide sum; every thread gets the result.  red must hold blockDim/64 doubles.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) s += red[i];
  return s;
}

// Loads that bypass the per-CU cache (data written by other workgroups of the
// same kernel).
__device__ __forceinline__ double load_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace tmx

#define TM_HIP(call)                                                                \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      tmx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                      \
      return TM_ERR_HIP;                                                            \
    }                                                                               \
  } while (0)
