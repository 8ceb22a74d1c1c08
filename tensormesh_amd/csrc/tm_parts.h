// Deterministic grid-wide reductions without atomics: every workgroup of a
// producer kernel writes its partial sum to part[slot * stride + blockIdx.x];
// a later kernel (kernel boundaries order the writes) sums the partials of a
// slot in a fixed order with its first wave.
#pragma once
#include "tm_device.h"

namespace tmx {

template <int NS>
__device__ __forceinline__ void sum_parts(const double* part, int stride, const int (&which)[NS], int n,
                                          double (&out)[NS]) {
  __shared__ double sh[NS];
  if (threadIdx.x < kWave) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      double s = 0.0;
      for (int k = threadIdx.x; k < n; k += kWave) s += part[which[i] * stride + k];
      s = wave_sum(s);
      if (threadIdx.x == 0) sh[i] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NS; ++i) out[i] = sh[i];
}

template <int NT, int NS>
__device__ __forceinline__ void write_parts(double* part, int stride, const int (&which)[NS], double (&loc)[NS]) {
  __shared__ double red[NT / kWave];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double b = block_sum<NT>(loc[i], red);
    if (threadIdx.x == 0) part[which[i] * stride + blockIdx.x] = b;
  }
}

}  // namespace tmx
