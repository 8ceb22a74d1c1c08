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
      // eight loads in flight per lane; the adds keep the order k ascending
      const double* p = part + which[i] * stride;
      double s = 0.0;
      int k = threadIdx.x;
      for (; k + 7 * kWave < n; k += 8 * kWave) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[k + u * kWave];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
      }
      for (; k < n; k += kWave) s += p[k];
      s = wave_sum(s);
      if (threadIdx.x == 0) sh[i] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NS; ++i) out[i] = sh[i];
}

// All NS block sums in one pass: wave sums, one workgroup barrier, then
// lane i < NS adds the wave sums of value i in wave order (the same order as
// one block_sum per value, so the partials are unchanged bit for bit).
template <int NT, int NS>
__device__ __forceinline__ void write_parts(double* part, int stride, const int (&which)[NS], double (&loc)[NS]) {
  __shared__ double red[NT / kWave][NS];
  double v[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) v[i] = wave_sum(loc[i]);
  if (threadIdx.x % kWave == 0)
#pragma unroll
    for (int i = 0; i < NS; ++i) red[threadIdx.x / kWave][i] = v[i];
  __syncthreads();
  if (threadIdx.x < NS) {
    double b = 0.0;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) b += red[w][threadIdx.x];
    int slot = which[0];
#pragma unroll
    for (int i = 1; i < NS; ++i)
      if ((int)threadIdx.x == i) slot = which[i];
    part[slot * stride + blockIdx.x] = b;
  }
  __syncthreads();  // red may be reused by a following call
}

}  // namespace tmx
