// CSR row kernels shared by the Krylov solvers: S lanes per (block) row,
// 4-deep unrolled accumulation, XCD-contiguous row slabs.
#pragma once
#include "tm_device.h"

namespace tmx {

// out[i] += sum over entries of one block row.  Each lane takes entries
// sl, sl + S, sl + 2S, ... in batches of U: the U column loads of a batch are
// issued together (predicated past the row end: those lanes load nothing and
// add zero), then the U value and x loads, so a lane keeps U independent
// chains in flight instead of walking its entries one by one.
template <typename VT, int S, int BS, int U = (S <= 4 ? 4 : 2)>
__device__ __forceinline__ void row_accum(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                          const VT* __restrict__ v, const double* __restrict__ x, int64_t row,
                                          int sl, double (&out)[BS]) {
  const int b = rp[row], e = rp[row + 1];
  for (int k0 = b + sl; k0 < e; k0 += U * S) {
    int64_t c[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = k0 + u * S < e;
      c[u] = ok[u] ? ci[k0 + u * S] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int i = 0; i < BS; ++i)
#pragma unroll
        for (int j = 0; j < BS; ++j) {
          const double a = ok[u] ? (double)v[(int64_t)(k0 + u * S) * BS * BS + i * BS + j] : 0.0;
          const double xv = ok[u] ? x[c[u] * BS + j] : 0.0;
          out[i] += a * xv;
        }
  }
}

// Rows of logical workgroup L: the contiguous range [L*chunk, (L+1)*chunk),
// with L = XCD-remapped block id, so each XCD sweeps one contiguous slab of
// rows and the x entries of neighbouring rows stay in its L2.
struct RowChunk {
  int64_t r0, r1;
};
__device__ __forceinline__ RowChunk row_chunk(int64_t nrows) {
  const int L = tmx::xcd_remap(blockIdx.x, gridDim.x);
  const int64_t chunk = (nrows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)L * chunk;
  return RowChunk{r0, r0 + chunk < nrows ? r0 + chunk : nrows};
}

template <int S, int BS>
__device__ __forceinline__ void sub_reduce(double (&o)[BS]) {
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int d = S / 2; d > 0; d >>= 1) o[i] += __shfl_xor(o[i], d, kWave);
}

}  // namespace tmx
