// CSR row kernels shared by the Krylov solvers: S lanes per (block) row,
// 4-deep unrolled accumulation, XCD-contiguous row slabs.
#pragma once
#include <type_traits>

#include "tm_device.h"

namespace tmx {

typedef int32_t i32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef int16_t i16x4u __attribute__((ext_vector_type(4), aligned(2)));

// Column index of CSR entry k of row `row`: ci is either the int32 column
// array or an int16 array of column - row (a compact copy for matrices whose
// entries all lie within 32767 of the diagonal, 2 bytes per entry less).
__device__ __forceinline__ int64_t col_of(const int32_t* __restrict__ ci, int64_t, int k) { return ci[k]; }
__device__ __forceinline__ int64_t col_of(const int16_t* __restrict__ ci, int64_t row, int k) {
  return row + ci[k];
}
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));

// out[i] += sum over entries of one block row.  lim is an entry index at or
// past the row's end and at most nnz (the caller's rp[last row + 1]): no load
// reaches lim.
//
// FP32 scalar rows (S <= 4): lane sl takes the 4 consecutive entries from
// b + 4 sl (then + 4 S, ...), read as one 16-byte column load and one 16-byte
// value load (dword-aligned types; gfx950 serves unaligned multi-dword global
// loads), so one load instruction of a wave covers whole consecutive cache
// lines.  Entries past the row end are masked (the load may run into the next
// row, never to lim); a group of 4 that would reach lim falls back to
// single loads.  A lane adds its entries in order.
//
// Other rows: each lane takes entries sl, sl + S, sl + 2S, ... in batches of U:
// the U column loads of a batch are issued together (predicated past the row
// end: those lanes load nothing and add zero), then the U value and x loads,
// so a lane keeps U independent chains in flight.  FP32 3x3 blocks read their
// 36 bytes as two 16-byte loads and one 4-byte load, and the three x entries
// as one 16-byte and one 8-byte load (same products, same order).
template <typename VT, int S, int BS, int U = (S <= 4 ? 4 : 2), typename CI = int32_t>
__device__ __forceinline__ void row_accum(const int32_t* __restrict__ rp, const CI* __restrict__ ci,
                                          const VT* __restrict__ v, const double* __restrict__ x, int64_t row,
                                          int sl, int lim, double (&out)[BS]) {
  const int b = rp[row], e = rp[row + 1];
  if constexpr (std::is_same<VT, float>::value && BS == 1 && S <= 4) {
    for (int k0 = b + 4 * sl; k0 < e; k0 += 4 * S) {
      int c[4];
      float a[4];
      if (k0 + 4 <= lim) {
        if constexpr (std::is_same<CI, int16_t>::value) {
          const i16x4u cv = *(const i16x4u*)(ci + k0);  // one 8-byte load
          c[0] = (int)row + cv.x, c[1] = (int)row + cv.y, c[2] = (int)row + cv.z, c[3] = (int)row + cv.w;
        } else {
          const i32x4u cv = *(const i32x4u*)(ci + k0);
          c[0] = cv.x, c[1] = cv.y, c[2] = cv.z, c[3] = cv.w;
        }
        const f32x4u av = *(const f32x4u*)(v + k0);
        a[0] = av.x, a[1] = av.y, a[2] = av.z, a[3] = av.w;
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          c[u] = k0 + u < e ? (int)col_of(ci, row, k0 + u) : 0;
          a[u] = k0 + u < e ? v[k0 + u] : 0.0f;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const bool ok = k0 + u < e;
        const double xv = x[ok ? c[u] : 0];
        out[0] += ok ? (double)a[u] * xv : 0.0;
      }
    }
    return;
  }
  if constexpr (std::is_same<VT, float>::value && BS == 3) {
    for (int k0 = b + sl; k0 < e; k0 += U * S) {
      int64_t c[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ok[u] = k0 + u * S < e;
        c[u] = ok[u] ? col_of(ci, row, k0 + u * S) : 0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        const float* vb = v + (int64_t)(k0 + u * S) * 9;
        const f32x4u a0 = *(const f32x4u*)vb;
        const f32x4u a1 = *(const f32x4u*)(vb + 4);
        const float a8 = vb[8];
        const double* xb = x + c[u] * 3;
        const f64x2u x01 = *(const f64x2u*)xb;
        const double x2 = xb[2];
        out[0] += (double)a0.x * x01.x;
        out[0] += (double)a0.y * x01.y;
        out[0] += (double)a0.z * x2;
        out[1] += (double)a0.w * x01.x;
        out[1] += (double)a1.x * x01.y;
        out[1] += (double)a1.y * x2;
        out[2] += (double)a1.z * x01.x;
        out[2] += (double)a1.w * x01.y;
        out[2] += (double)a8 * x2;
      }
    }
    return;
  }
  for (int k0 = b + sl; k0 < e; k0 += U * S) {
    int64_t c[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = k0 + u * S < e;
      c[u] = ok[u] ? col_of(ci, row, k0 + u * S) : 0;
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
