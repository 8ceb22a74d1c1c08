// Learned preconditioner for gfx950: a U-shaped edge message-passing network
// over nested node aggregates, and the flexible CG that applies it.
//
// Features are node-major [n][64] BF16 rows (128 B per node).  One
// message-passing layer at a level with normalised operator A^ (unit
// diagonal, Galerkin-coarsened on coarse levels) is
//     a = A^ h                      (edge messages weighted by matrix entries)
//     u = [h | a] W + b             (W: 128 x 64, BF16 MFMA, FP32 accumulate)
//     h' = h + act(u)               (act: identity on the first n_lin channels,
//                                    ReLU on the rest)
// and is one fused kernel: a 64-node tile gathers its messages into LDS, four
// waves multiply the [64 x 128] tile by W with v_mfma_f32_16x16x32_bf16 and
// write h' (and, for training, a and u).  The backward pass is one persistent
// MFMA kernel (input gradients and per-workgroup weight-gradient partials)
// plus a message-transpose kernel; every weight gradient is reduced in a
// fixed order, so training steps are deterministic.
//
// MFMA operand layout (16x16x32 BF16, wave64): lane l holds A[l%16][8(l/16)..+7],
// B[8(l/16)..+7][l%16]; the FP32 result holds D[4(l/16)+v][l%16], v = 0..3.
#include <cmath>
#include <cstdint>

#include "tm_device.h"
#include "tm_parts.h"
#include "tm_spmv.h"

namespace {

using tmx::kWave;
constexpr int F = 64;          // feature width
constexpr int XP = 2 * F + 8;  // LDS pitch (BF16 elements) of a [h | a] row
constexpr int GP = F + 8;      // LDS pitch of a 64-wide BF16 row
constexpr int TILE = 64;       // node rows per tile (4 waves x 16 rows)
constexpr int64_t kSmallLevel = 65536;  // inference layers below this many rows use 16-row tiles
constexpr int NT = 256;
constexpr int NTF = 512;  // flexible-CG vector kernels
constexpr int kMaxParts = 1024;
constexpr int NWP = 2 * F * F + F;  // weight-gradient partial of one layer: dW (128 x 64) + db (64)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t u) { return __builtin_bit_cast(float, (uint32_t)u << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
}
__device__ __forceinline__ bf16x8 lds8(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *(const uint4*)p); }
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float act(float u, int j, int n_lin) { return j < n_lin ? u : fmaxf(u, 0.0f); }
__device__ __forceinline__ float dact(float u, int j, int n_lin) { return (j < n_lin || u > 0.0f) ? 1.0f : 0.0f; }
__device__ __forceinline__ bool stopped(const int* stop) { return stop != nullptr && *stop != 0; }

// ---------------------------------------------------------------------------
// Forward message-passing layer
//
// A workgroup owns one 64-row tile.  The CSR entries of its rows are staged in
// LDS (up to kEnt entries; larger tiles read them from global memory), each
// wave gathers the messages of 16 rows with 4 lanes per row and 16 features
// per lane (32 B of the neighbour's row per lane), then the tile [h | a] is
// multiplied by W on MFMA.
// ---------------------------------------------------------------------------
constexpr int kEnt = 1024;

__device__ __forceinline__ void load16(const uint16_t* __restrict__ h, int64_t j, int f0, float (&o)[16]) {
  const uint4 a = *(const uint4*)&h[j * F + f0];
  const uint4 b = *(const uint4*)&h[j * F + f0 + 8];
  const uint32_t w8[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    o[2 * t] = bf2f((uint16_t)(w8[t] & 0xffffu));
    o[2 * t + 1] = bf2f((uint16_t)(w8[t] >> 16));
  }
}

// 16 features of a row as BF16 into LDS (32 B)
__device__ __forceinline__ void store16(uint16_t* p, const float (&v)[16]) {
  uint32_t pk[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) pk[u] = pack2(v[2 * u], v[2 * u + 1]);
  *(uint4*)p = uint4{pk[0], pk[1], pk[2], pk[3]};
  *(uint4*)(p + 8) = uint4{pk[4], pk[5], pk[6], pk[7]};
}

__device__ __forceinline__ void load_wt(uint16_t* W, const uint16_t* __restrict__ wt) {
  for (int c = threadIdx.x; c < F * 16; c += NT) {
    const int j = c / 16, k8 = (c % 16) * 8;
    *(uint4*)&W[j * XP + k8] = *(const uint4*)&wt[j * 2 * F + k8];
  }
}

// acc[jt] = X[16 rows of this wave][0:128] x W^T[jt*16 .. +16][0:128]^T with the
// B operands read straight from the global W^T ([64][128], 16 KB: every
// workgroup reads the same rows, so they come from the L1/L2 caches and the
// LDS stays free for more resident workgroups)
__device__ __forceinline__ void tile_mfma_gw(const uint16_t* X, const uint16_t* __restrict__ wt, int wv, int l,
                                             f32x4 (&acc)[4]) {
#pragma unroll
  for (int jt = 0; jt < 4; ++jt) acc[jt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int ar = wv * 16 + (l % 16), ak = 8 * (l / 16);
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const bf16x8 a = lds8(&X[ar * XP + kk * 32 + ak]);
#pragma unroll
    for (int jt = 0; jt < 4; ++jt)
      acc[jt] = mfma(a, __builtin_bit_cast(bf16x8, *(const uint4*)&wt[(jt * 16 + (l % 16)) * 2 * F + kk * 32 + ak]),
                     acc[jt]);
  }
}

// acc[jt] = X[16 rows of this wave][0:128] x W^T[jt*16 .. +16][0:128]^T
__device__ __forceinline__ void tile_mfma(const uint16_t* X, const uint16_t* W, int wv, int l, f32x4 (&acc)[4]) {
#pragma unroll
  for (int jt = 0; jt < 4; ++jt) acc[jt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int ar = wv * 16 + (l % 16), ak = 8 * (l / 16);
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const bf16x8 a = lds8(&X[ar * XP + kk * 32 + ak]);
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) acc[jt] = mfma(a, lds8(&W[(jt * 16 + (l % 16)) * XP + kk * 32 + ak]), acc[jt]);
  }
}

template <bool SAVE>
__global__ __launch_bounds__(NT) void k_mp(int64_t n, const int32_t* __restrict__ rp,
                                           const int32_t* __restrict__ ci, const float* __restrict__ av,
                                           const uint16_t* __restrict__ h, const uint16_t* __restrict__ wt,
                                           const float* __restrict__ bias, int n_lin, uint16_t* __restrict__ out,
                                           uint16_t* __restrict__ s_agg, uint16_t* __restrict__ s_u,
                                           const int* stop) {
  if (stopped(stop)) return;
  __shared__ __align__(16) uint16_t X[TILE * XP];
  __shared__ int rps[TILE + 1];
  __shared__ int ec[kEnt];
  __shared__ float ev[kEnt];
  const int tid = threadIdx.x, wv = tid / kWave, l = tid % kWave;
  const int64_t r0 = (int64_t)tmx::xcd_remap(blockIdx.x, gridDim.x) * TILE;
  for (int c = tid; c <= TILE; c += NT) rps[c] = rp[r0 + c < n ? r0 + c : n];
  __syncthreads();
  const int e0 = rps[0], ecount = rps[TILE] - e0;
  const bool staged = ecount <= kEnt;
  if (staged)
    for (int c = tid; c < ecount; c += NT) {
      ec[c] = ci[e0 + c];
      ev[c] = av[e0 + c];
    }
  __syncthreads();
  const int ri = wv * 16 + l / 4, f0 = (l % 4) * 16;
  {
    const int64_t row = r0 + ri;
    float hv[16], a[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      a[u] = 0.0f;
      hv[u] = 0.0f;
    }
    if (row < n) {
      load16(h, row, f0, hv);
      const int kb = rps[ri] - e0, ke = rps[ri + 1] - e0;
      // four neighbour rows in flight per lane (the gathers are latency-bound)
      for (int k = kb; k < ke; k += 4) {
        int c[4];
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const bool in = k + q < ke;
          c[q] = in ? (staged ? ec[k + q] : ci[e0 + k + q]) : (int)row;
          v[q] = in ? (staged ? ev[k + q] : av[e0 + k + q]) : 0.0f;
        }
        float hq[4][16];
#pragma unroll
        for (int q = 0; q < 4; ++q) load16(h, c[q], f0, hq[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int u = 0; u < 16; ++u) a[u] += v[q] * hq[q][u];
      }
    }
    store16(&X[ri * XP + f0], hv);
    store16(&X[ri * XP + F + f0], a);
    if (SAVE && row < n) {
      *(uint4*)&s_agg[row * F + f0] = *(const uint4*)&X[ri * XP + F + f0];
      *(uint4*)&s_agg[row * F + f0 + 8] = *(const uint4*)&X[ri * XP + F + f0 + 8];
    }
  }
  __syncthreads();
  f32x4 acc[4];
  tile_mfma_gw(X, wt, wv, l, acc);
#pragma unroll
  for (int jt = 0; jt < 4; ++jt) {
    const int j = jt * 16 + (l % 16);
    const float bj = bias[j];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = wv * 16 + 4 * (l / 16) + v;
      const float u = acc[jt][v] + bj;
      const float hn = bf2f(X[i * XP + j]) + act(u, j, n_lin);
      X[i * XP + j] = f2bf(hn);
      if (SAVE && r0 + i < n) s_u[(r0 + i) * F + j] = f2bf(u);
    }
  }
  __syncthreads();
  for (int c = l; c < 16 * 8; c += kWave) {
    const int i = wv * 16 + c / 8, k8 = (c % 8) * 8;
    if (r0 + i < n) *(uint4*)&out[(r0 + i) * F + k8] = *(const uint4*)&X[i * XP + k8];
  }
}

// Inference layer for small levels (few rows, so k_mp's 64-row tiles leave
// most CUs idle and each lane walks ~7 batches of dependent gathers): 16-row
// tiles, 16 lanes per row with 4 features each, 8 neighbour rows in flight.
// Each feature's message is summed over the row's entries in CSR order and
// rounded like k_mp, and the MFMA sees the same [h | a] tile, so the result
// is bit-identical to k_mp<false>; wave w forms output columns 16w..16w+15.
constexpr int TS = 16;  // rows per tile
__global__ __launch_bounds__(NT) void k_mp_small(int64_t n, const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ci, const float* __restrict__ av,
                                                 const uint16_t* __restrict__ h, const uint16_t* __restrict__ wt,
                                                 const float* __restrict__ bias, int n_lin,
                                                 uint16_t* __restrict__ out, const int* stop) {
  if (stopped(stop)) return;
  __shared__ __align__(16) uint16_t X[TS * XP];
  const int tid = threadIdx.x, wv = tid / kWave, l = tid % kWave;
  const int64_t r0 = (int64_t)blockIdx.x * TS;
  const int ri = tid / 16, f0 = (tid % 16) * 4;
  const int64_t row = r0 + ri;
  float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  uint2 hv = uint2{0u, 0u};
  if (row < n) {
    hv = *(const uint2*)&h[row * F + f0];
    const int kb = rp[row], ke = rp[row + 1];
    for (int k = kb; k < ke; k += 8) {
      int c[8];
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const bool in = k + q < ke;
        c[q] = in ? ci[k + q] : (int)row;
        v[q] = in ? av[k + q] : 0.0f;
      }
      uint2 hq[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) hq[q] = *(const uint2*)&h[(int64_t)c[q] * F + f0];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        a[0] += v[q] * bf2f((uint16_t)(hq[q].x & 0xffffu));
        a[1] += v[q] * bf2f((uint16_t)(hq[q].x >> 16));
        a[2] += v[q] * bf2f((uint16_t)(hq[q].y & 0xffffu));
        a[3] += v[q] * bf2f((uint16_t)(hq[q].y >> 16));
      }
    }
  }
  *(uint2*)&X[ri * XP + f0] = hv;
  *(uint2*)&X[ri * XP + F + f0] = uint2{pack2(a[0], a[1]), pack2(a[2], a[3])};
  __syncthreads();
  // wave wv: columns jt = wv of the 16-row block
  f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  {
    const int ar = l % 16, ak = 8 * (l / 16), jt = wv;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const bf16x8 xa = lds8(&X[ar * XP + kk * 32 + ak]);
      acc = mfma(xa, __builtin_bit_cast(bf16x8, *(const uint4*)&wt[(jt * 16 + (l % 16)) * 2 * F + kk * 32 + ak]),
                 acc);
    }
  }
  __syncthreads();  // every wave has read X before the epilogue overwrites it
  {
    const int j = wv * 16 + (l % 16);
    const float bj = bias[j];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 4 * (l / 16) + v;
      const float u = acc[v] + bj;
      X[i * XP + j] = f2bf(bf2f(X[i * XP + j]) + act(u, j, n_lin));
    }
  }
  __syncthreads();
  if (row < n) *(uint2*)&out[row * F + f0] = *(const uint2*)&X[ri * XP + f0];
}

// ---------------------------------------------------------------------------
// Finest level at inference, fused: unpool + message passing + decode.
//
// With the linear encoder e_j = x_j0 We0 + x_j1 We1 + be, the unpooled level-0
// features are h_j = e_j + w_j hc[agg_j] and their messages
//     a_i = sum_j A^_ij h_j
//         = (A^ x0)_i We0 + (A^ x1)_i We1 + (A^ 1)_i be + sum_J B_iJ hc[J]
// with B = A^ P^ (B_iJ = sum over neighbours j of i in aggregate J of
// A^_ij w_j, formed once per matrix).  So a row needs two scalar products
// over its CSR row, its few B entries (coarse feature rows, shared by the
// rows of one aggregate) and its own coarse row - no 64-wide level-0 tensor.
// One 64-row tile per workgroup; the gathers of a row are issued together.
// ---------------------------------------------------------------------------
struct FinalArgs {
  const int32_t* rp;
  const int32_t* ci;
  const float* av;
  const float2* xs;   // (d r / s, lk) per node, written by k_encode_pool
  const float* a1;    // row sums of A^
  const int32_t* agg;
  const float* w;
  const uint16_t* hc;
  const int32_t* brp;
  const int32_t* bcol;
  const float* bv;
  const float* we;
  const float* be;
  const uint16_t* wt;
  const float* bias;
  int n_lin;
  const float* wd;
  const float* bd;
  const float* d;
  const double* scale;
  double* z;
};

__global__ __launch_bounds__(NT) void k_final(int64_t n, FinalArgs A, const int* stop) {
  if (stopped(stop)) return;
  __shared__ __align__(16) uint16_t X[TILE * XP];
  __shared__ __align__(16) uint16_t W[F * XP];
  const int tid = threadIdx.x, wv = tid / kWave, l = tid % kWave;
  const int64_t r0 = (int64_t)tmx::xcd_remap(blockIdx.x, gridDim.x) * TILE;
  load_wt(W, A.wt);
  const int ri = wv * 16 + l / 4, qd = l % 4, f0 = qd * 16;
  const int64_t row = r0 + ri;
  float hv[16], a[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    hv[u] = 0.0f;
    a[u] = 0.0f;
  }
  // scalar products (A^ x0)_i, (A^ x1)_i: each of the 4 lanes of the row takes
  // every 4th entry, column indices first so the xs gathers overlap
  float ax0 = 0.0f, ax1 = 0.0f;
  if (row < n) {
    const int kb = A.rp[row], ke = A.rp[row + 1];
    int c[4];
    float v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = kb + qd + 4 * u;
      c[u] = k < ke ? A.ci[k] : -1;
      v[u] = k < ke ? A.av[k] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (c[u] >= 0) {
        const float2 xj = A.xs[c[u]];
        ax0 += v[u] * xj.x;
        ax1 += v[u] * xj.y;
      }
    for (int k = kb + qd + 16; k < ke; k += 4) {
      const float2 xj = A.xs[A.ci[k]];
      ax0 += A.av[k] * xj.x;
      ax1 += A.av[k] * xj.y;
    }
  }
  ax0 += __shfl_xor(ax0, 1, kWave);
  ax0 += __shfl_xor(ax0, 2, kWave);
  ax1 += __shfl_xor(ax1, 1, kWave);
  ax1 += __shfl_xor(ax1, 2, kWave);
  if (row < n) {
    const float2 xi = A.xs[row];
    const float wi = A.w[row], ai = A.a1[row];
    const int bb = A.brp[row], be_ = A.brp[row + 1];
    float hc[16];
    load16(A.hc, A.agg[row], f0, hc);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const float e = xi.x * A.we[f0 + u] + xi.y * A.we[F + f0 + u] + A.be[f0 + u];
      hv[u] = e + wi * hc[u];
      a[u] = ax0 * A.we[f0 + u] + ax1 * A.we[F + f0 + u] + ai * A.be[f0 + u];
    }
    int k = bb;
    for (; k + 1 < be_; k += 2) {
      const int j0 = A.bcol[k], j1 = A.bcol[k + 1];
      const float b0 = A.bv[k], b1 = A.bv[k + 1];
      float g0[16], g1[16];
      load16(A.hc, j0, f0, g0);
      load16(A.hc, j1, f0, g1);
#pragma unroll
      for (int u = 0; u < 16; ++u) a[u] += b0 * g0[u] + b1 * g1[u];
    }
    if (k < be_) {
      const float b0 = A.bv[k];
      float g0[16];
      load16(A.hc, A.bcol[k], f0, g0);
#pragma unroll
      for (int u = 0; u < 16; ++u) a[u] += b0 * g0[u];
    }
  }
  store16(&X[ri * XP + f0], hv);
  store16(&X[ri * XP + F + f0], a);
  __syncthreads();
  f32x4 acc[4];
  tile_mfma(X, W, wv, l, acc);
  float yp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int jt = 0; jt < 4; ++jt) {
    const int j = jt * 16 + (l % 16);
    const float bj = A.bias[j], wdj = A.wd[j];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = wv * 16 + 4 * (l / 16) + v;
      const float hn = bf2f(f2bf(bf2f(X[i * XP + j]) + act(acc[jt][v] + bj, j, A.n_lin)));
      yp[v] += hn * wdj;
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) yp[v] += __shfl_xor(yp[v], m, kWave);
  if ((l % 16) == 0) {
    const double s = *A.scale;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t rr = r0 + wv * 16 + 4 * (l / 16) + v;
      if (rr < n) A.z[rr] = s * (double)A.d[rr] * (double)(yp[v] + A.bd[0]);
    }
  }
}

// ---------------------------------------------------------------------------
// Backward message-passing layer
//   g_u = g_out * act'(u)
//   g_in = g_out + g_u W_h^T   (this kernel)  + A^ (g_u W_a^T)  (k_msg_add)
//   dW += [h | a]^T g_u, db += sum g_u    (per-workgroup partials)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_mp_bwd(int64_t n, const float* __restrict__ g_out,
                                               const uint16_t* __restrict__ h, const uint16_t* __restrict__ agg,
                                               const uint16_t* __restrict__ u, const uint16_t* __restrict__ w,
                                               int n_lin, float* __restrict__ g_in, float* __restrict__ g_msg,
                                               float* __restrict__ part) {
  __shared__ __align__(16) uint16_t Wl[2 * F * GP];  // W: [128 inputs][64 outputs]
  __shared__ __align__(16) uint16_t Xs[TILE * XP];
  __shared__ __align__(16) uint16_t Gb[TILE * GP];
  __shared__ float Gf[TILE * F];
  const int tid = threadIdx.x, wv = tid / kWave, l = tid % kWave;
  for (int c = tid; c < 2 * F * 8; c += NT) {
    const int k = c / 8, j8 = (c % 8) * 8;
    *(uint4*)&Wl[k * GP + j8] = *(const uint4*)&w[k * F + j8];
  }
  f32x4 dw[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) dw[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float db = 0.0f;
  const int64_t ntiles = (n + TILE - 1) / TILE;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t r0 = t * TILE;
    __syncthreads();
    for (int c = tid; c < TILE * 8; c += NT) {
      const int i = c / 8, k8 = (c % 8) * 8;
      const int64_t row = r0 + i;
      uint4 vh = {0u, 0u, 0u, 0u}, va = {0u, 0u, 0u, 0u};
      if (row < n) {
        vh = *(const uint4*)&h[row * F + k8];
        va = *(const uint4*)&agg[row * F + k8];
      }
      *(uint4*)&Xs[i * XP + k8] = vh;
      *(uint4*)&Xs[i * XP + F + k8] = va;
    }
    for (int c = tid; c < TILE * F; c += NT) {
      const int i = c / F, j = c % F;
      const int64_t row = r0 + i;
      float g = 0.0f;
      if (row < n) g = g_out[row * F + j] * dact(bf2f(u[row * F + j]), j, n_lin);
      Gf[i * F + j] = g;
      Gb[i * GP + j] = f2bf(g);
    }
    __syncthreads();
    // g_u (16 rows of this wave) x W^T -> [16 x 128]
    {
      f32x4 acc[8];
#pragma unroll
      for (int ct = 0; ct < 8; ++ct) acc[ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      const int ar = wv * 16 + (l % 16), ak = 8 * (l / 16);
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        const bf16x8 a = lds8(&Gb[ar * GP + kk * 32 + ak]);
#pragma unroll
        for (int ct = 0; ct < 8; ++ct) acc[ct] = mfma(a, lds8(&Wl[(ct * 16 + (l % 16)) * GP + kk * 32 + ak]), acc[ct]);
      }
#pragma unroll
      for (int ct = 0; ct < 8; ++ct) {
        const int c = ct * 16 + (l % 16);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t row = r0 + wv * 16 + 4 * (l / 16) + v;
          if (row < n) {
            if (c < F)
              g_in[row * F + c] = g_out[row * F + c] + acc[ct][v];
            else
              g_msg[row * F + (c - F)] = acc[ct][v];
          }
        }
      }
    }
    // dW[k][j] += sum_i X[i][k] g_u[i][j]: this wave owns k in [32 wv, 32 wv + 32)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const int i0 = kk * 32 + 8 * (l / 16);
      bf16x8 bop[4];
#pragma unroll
      for (int jt = 0; jt < 4; ++jt)
#pragma unroll
        for (int v = 0; v < 8; ++v) bop[jt][v] = __builtin_bit_cast(__bf16, Gb[(i0 + v) * GP + jt * 16 + (l % 16)]);
#pragma unroll
      for (int kt = 0; kt < 2; ++kt) {
        bf16x8 aop;
        const int k = (2 * wv + kt) * 16 + (l % 16);
#pragma unroll
        for (int v = 0; v < 8; ++v) aop[v] = __builtin_bit_cast(__bf16, Xs[(i0 + v) * XP + k]);
#pragma unroll
        for (int jt = 0; jt < 4; ++jt) dw[kt][jt] = mfma(aop, bop[jt], dw[kt][jt]);
      }
    }
    if (tid < F)
      for (int i = 0; i < TILE; ++i) db += Gf[i * F + tid];
  }
  float* pp = part + (int64_t)blockIdx.x * NWP;
#pragma unroll
  for (int kt = 0; kt < 2; ++kt)
#pragma unroll
    for (int jt = 0; jt < 4; ++jt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int k = (2 * wv + kt) * 16 + 4 * (l / 16) + v, j = jt * 16 + (l % 16);
        pp[k * F + j] = dw[kt][jt][v];
      }
  if (tid < F) pp[2 * F * F + tid] = db;
}

// g_in[i] += sum_k A^_ik g_msg[col_k]   (A^ symmetric: the transpose of the
// message aggregation is the aggregation itself).  One wave per row.
__global__ __launch_bounds__(NT) void k_msg_add(int64_t n, const int32_t* __restrict__ rp,
                                                const int32_t* __restrict__ ci, const float* __restrict__ av,
                                                const float* __restrict__ g_msg, float* __restrict__ g_in) {
  const int l = threadIdx.x % kWave;
  for (int64_t row = blockIdx.x * (int64_t)(NT / kWave) + threadIdx.x / kWave; row < n;
       row += (int64_t)gridDim.x * (NT / kWave)) {
    float s = 0.0f;
    const int b = rp[row], e = rp[row + 1];
    int k = b;
    for (; k + 3 < e; k += 4) {
      int c[4];
      float v[4], g[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        c[q] = ci[k + q];
        v[q] = av[k + q];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) g[q] = g_msg[(int64_t)c[q] * F + l];
#pragma unroll
      for (int q = 0; q < 4; ++q) s += v[q] * g[q];
    }
    for (; k < e; ++k) s += av[k] * g_msg[(int64_t)ci[k] * F + l];
    g_in[row * F + l] += s;
  }
}

// out[c] (+)= sum_b part[b][c] in block order.
__global__ __launch_bounds__(NT) void k_reduce_parts(int nparts, int width, const float* __restrict__ part,
                                                     float* __restrict__ out, int accumulate) {
  for (int c = blockIdx.x * NT + threadIdx.x; c < width; c += gridDim.x * NT) {
    float s = 0.0f;
    for (int b = 0; b < nparts; ++b) s += part[(int64_t)b * width + c];
    out[c] = accumulate ? out[c] + s : s;
  }
}

// ---------------------------------------------------------------------------
// Aggregation levels: pool (restriction) and unpool (prolongation + skip)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ float ld(const T* p);
template <>
__device__ __forceinline__ float ld<uint16_t>(const uint16_t* p) { return bf2f(*p); }
template <>
__device__ __forceinline__ float ld<float>(const float* p) { return *p; }
template <typename T>
__device__ __forceinline__ void st(T* p, float v);
template <>
__device__ __forceinline__ void st<uint16_t>(uint16_t* p, float v) { *p = f2bf(v); }
template <>
__device__ __forceinline__ void st<float>(float* p, float v) { *p = v; }

template <typename T>
struct Vec8;  // 8 features: one 16 B (BF16) or 32 B (FP32) chunk
template <>
struct Vec8<uint16_t> {
  __device__ __forceinline__ static void load(const uint16_t* p, float (&o)[8]) {
    const uint4 a = *(const uint4*)p;
    const uint32_t w4[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      o[2 * t] = bf2f((uint16_t)(w4[t] & 0xffffu));
      o[2 * t + 1] = bf2f((uint16_t)(w4[t] >> 16));
    }
  }
  __device__ __forceinline__ static void store(uint16_t* p, const float (&v)[8]) {
    *(uint4*)p = uint4{pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7])};
  }
};
template <>
struct Vec8<float> {
  __device__ __forceinline__ static void load(const float* p, float (&o)[8]) {
    const float4 a = *(const float4*)p, b = *(const float4*)(p + 4);
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
  }
  __device__ __forceinline__ static void store(float* p, const float (&v)[8]) {
    *(float4*)p = float4{v[0], v[1], v[2], v[3]};
    *(float4*)(p + 4) = float4{v[4], v[5], v[6], v[7]};
  }
};

// out[I][f] = sum over children i of I: w_i h[i][f].  8 lanes per coarse node.
template <typename T>
__global__ __launch_bounds__(NT) void k_pool(int64_t nc, const int32_t* __restrict__ cp,
                                             const int32_t* __restrict__ child, const float* __restrict__ w,
                                             const T* __restrict__ h, T* __restrict__ out, const int* stop) {
  if (stopped(stop)) return;
  const int f0 = (threadIdx.x % 8) * 8;
  for (int64_t I = blockIdx.x * (int64_t)(NT / 8) + threadIdx.x / 8; I < nc; I += (int64_t)gridDim.x * (NT / 8)) {
    float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = cp[I]; k < cp[I + 1]; ++k) {
      const int i = child[k];
      const float wi = w[i];
      float v[8];
      Vec8<T>::load(&h[(int64_t)i * F + f0], v);
#pragma unroll
      for (int t = 0; t < 8; ++t) s[t] += wi * v[t];
    }
    Vec8<T>::store(&out[I * F + f0], s);
  }
}

// out[i][f] = skip[i][f] + w_i hc[agg_i][f]   (skip may be null; out may alias skip)
template <typename T>
__global__ __launch_bounds__(NT) void k_unpool(int64_t nf, const int32_t* __restrict__ agg,
                                               const float* __restrict__ w, const T* __restrict__ hc,
                                               const T* skip, T* out, const int* stop) {
  if (stopped(stop)) return;
  const int64_t total = nf * 8;
  for (int64_t t = blockIdx.x * (int64_t)NT + threadIdx.x; t < total; t += (int64_t)gridDim.x * NT) {
    const int64_t i = t / 8;
    const int f0 = (int)(t % 8) * 8;
    const float wi = w[i];
    float v[8], sk[8];
    Vec8<T>::load(&hc[(int64_t)agg[i] * F + f0], v);
    if (skip) {
      Vec8<T>::load(&skip[i * F + f0], sk);
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = sk[u] + wi * v[u];
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] *= wi;
    }
    Vec8<T>::store(&out[i * F + f0], v);
  }
}

// ---------------------------------------------------------------------------
// Encoder / decoder
//   x_i = (d_i r_i / s, lk_i), s = sqrt(sum (d r)^2 / n)  (d = diag^-1/2)
//   h0[i][f] = act(x_i0 We[0][f] + x_i1 We[1][f] + be[f])
//   y_i = h[i] . wd + bd,   z_i = s d_i y_i
// ---------------------------------------------------------------------------
__device__ __forceinline__ double scale_from_parts(const double* ss_part, int nparts, int64_t n) {
  const int which[1] = {0};
  double t[1];
  tmx::sum_parts<1>(ss_part, kMaxParts, which, nparts, t);
  const double s = n > 0 ? sqrt(t[0] / (double)n) : 0.0;
  return s > 0.0 ? s : 1.0;
}

// s = sqrt(sum of the partials / n) into *s_out (single workgroup)
__global__ __launch_bounds__(NT) void k_scale(int64_t n, const double* __restrict__ ss_part, int nparts,
                                              double* __restrict__ s_out) {
  const double s = scale_from_parts(ss_part, nparts, n);
  if (threadIdx.x == 0) *s_out = s;
}

__global__ __launch_bounds__(NT) void k_sumsq(int64_t n, const double* __restrict__ r, const float* __restrict__ d,
                                              double* __restrict__ ss_part) {
  double loc[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const double v = (double)d[i] * r[i];
    loc[0] += v * v;
  }
  const int which[1] = {0};
  tmx::write_parts<NT, 1>(ss_part, kMaxParts, which, loc);
}

__global__ __launch_bounds__(NT) void k_encode(int64_t n, const double* __restrict__ r, const float* __restrict__ d,
                                               const float* __restrict__ lk, const double* __restrict__ ss_part,
                                               int nparts, const float* __restrict__ we, const float* __restrict__ be,
                                               uint16_t* __restrict__ h0, double* __restrict__ s_out,
                                               const int* stop) {
  if (stopped(stop)) return;
  const double s = scale_from_parts(ss_part, nparts, n);
  if (blockIdx.x == 0 && threadIdx.x == 0) *s_out = s;
  const float inv = (float)(1.0 / s);
  const int q = threadIdx.x % 8, f0 = q * 8;
  for (int64_t i = blockIdx.x * (int64_t)(NT / 8) + threadIdx.x / 8; i < n; i += (int64_t)gridDim.x * (NT / 8)) {
    const float x0 = (float)((double)d[i] * r[i]) * inv, x1 = lk[i];
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = x0 * we[f0 + t] + x1 * we[F + f0 + t] + be[f0 + t];
    Vec8<uint16_t>::store(&h0[i * F + f0], v);
  }
}

// Fused encoder + first pooling (inference).  The encoder is linear, so
//   out[I] = (sum w_i x_i0) We0 + (sum w_i x_i1) We1 + (sum w_i) be
// over the children i of I; also writes x_i = (d_i r_i / s, lk_i) for the
// finest-level kernel.  8 lanes per coarse node.
__global__ __launch_bounds__(NT) void k_encode_pool(int64_t n, int64_t nc, const int32_t* __restrict__ cp,
                                                    const int32_t* __restrict__ child, const float* __restrict__ w,
                                                    const double* __restrict__ r, const float* __restrict__ d,
                                                    const float* __restrict__ lk, const double* __restrict__ ss_part,
                                                    int nparts, const float* __restrict__ we,
                                                    const float* __restrict__ be, uint16_t* __restrict__ out,
                                                    float2* __restrict__ xs, double* __restrict__ s_out,
                                                    const int* stop) {
  if (stopped(stop)) return;
  // nparts == 0: the caller already stored s (the flexible-CG check kernel does)
  const double s = nparts > 0 ? scale_from_parts(ss_part, nparts, n) : *s_out;
  if (nparts > 0 && blockIdx.x == 0 && threadIdx.x == 0) *s_out = s;
  const float inv = (float)(1.0 / s);
  const int q = threadIdx.x % 8, f0 = q * 8;
  for (int64_t I0 = blockIdx.x * (int64_t)(NT / 8); I0 < nc; I0 += (int64_t)gridDim.x * (NT / 8)) {
    const int64_t I = I0 + threadIdx.x / 8;
    float s0 = 0.0f, s1 = 0.0f, sw = 0.0f;
    if (I < nc)
      for (int k = cp[I] + q; k < cp[I + 1]; k += 8) {
        const int i = child[k];
        float x0, x1;
        if (r) {
          x0 = (float)((double)d[i] * r[i]) * inv;
          x1 = lk ? lk[i] : 0.0f;
          xs[i] = float2{x0, x1};
        } else {  // inputs already in xs (k_sym0_down)
          const float2 v = xs[i];
          x0 = v.x;
          x1 = v.y;
        }
        const float wi = w[i];
        s0 += wi * x0;
        s1 += wi * x1;
        sw += wi;
      }
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
      s0 += __shfl_xor(s0, m, kWave);
      s1 += __shfl_xor(s1, m, kWave);
      sw += __shfl_xor(sw, m, kWave);
    }
    if (I < nc) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = s0 * we[f0 + t] + s1 * we[F + f0 + t] + sw * be[f0 + t];
      Vec8<uint16_t>::store(&out[I * F + f0], v);
    }
  }
}

// dWe, dbe partials: [3][64] per workgroup (linear encoder)
__global__ __launch_bounds__(NT) void k_encode_bwd(int64_t n, const double* __restrict__ r,
                                                   const float* __restrict__ d, const float* __restrict__ lk,
                                                   const double* __restrict__ s_in, const float* __restrict__ g_h0,
                                                   float* __restrict__ part) {
  __shared__ float red[NT / kWave][3 * F];
  const int f = threadIdx.x % F, wv = threadIdx.x / F;
  const float inv = (float)(1.0 / *s_in);
  float a0 = 0.0f, a1 = 0.0f, ab = 0.0f;
  for (int64_t i = blockIdx.x * (int64_t)(NT / F) + wv; i < n; i += (int64_t)gridDim.x * (NT / F)) {
    const float x0 = (float)((double)d[i] * r[i]) * inv, x1 = lk[i];
    const float g = g_h0[i * F + f];
    a0 += x0 * g;
    a1 += x1 * g;
    ab += g;
  }
  red[wv][f] = a0;
  red[wv][F + f] = a1;
  red[wv][2 * F + f] = ab;
  __syncthreads();
  if (threadIdx.x < 3 * F) {
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) sum += red[k][threadIdx.x];
    part[(int64_t)blockIdx.x * 3 * F + threadIdx.x] = sum;
  }
}

// y (normalised output, optional), z = s d y (FP64).  8 lanes per row.
__global__ __launch_bounds__(NT) void k_decode(int64_t n, const uint16_t* __restrict__ h,
                                               const float* __restrict__ wd, const float* __restrict__ bd,
                                               const float* __restrict__ d, const double* __restrict__ s_in,
                                               float* __restrict__ y, double* __restrict__ z, const int* stop) {
  if (stopped(stop)) return;
  const double s = *s_in;
  const int qd = threadIdx.x % 8, f0 = qd * 8;
  float wl[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) wl[t] = wd[f0 + t];
  const float b = bd[0];
  for (int64_t i0 = blockIdx.x * (int64_t)(NT / 8); i0 < n; i0 += (int64_t)gridDim.x * (NT / 8)) {
    const int64_t i = i0 + threadIdx.x / 8;
    float acc = 0.0f;
    if (i < n) {
      float v[8];
      Vec8<uint16_t>::load(&h[i * F + f0], v);
#pragma unroll
      for (int t = 0; t < 8; ++t) acc += v[t] * wl[t];
    }
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    if (qd == 0 && i < n) {
      const float yi = acc + b;
      if (y) y[i] = yi;
      z[i] = s * (double)d[i] * (double)yi;
    }
  }
}

// Finest level at inference when no layer runs there: unpool with the
// encoder output as skip, then decode:
//   y_i = (e_i + w_i hc[agg_i]) . wd + bd,  e_i = x_i0 We0 + x_i1 We1 + be,
//   z_i = s d_i y_i.
// The decoder is linear, so it is applied to the coarse rows first
// (u[J] = hc[J] . wd, one dot per coarse node) and the finest level only
// gathers one scalar per node.
__global__ __launch_bounds__(NT) void k_row_dot(int64_t nc, const uint16_t* __restrict__ hc,
                                                const float* __restrict__ wd, float* __restrict__ u,
                                                const int* stop) {
  if (stopped(stop)) return;
  const int q = threadIdx.x % 8, f0 = q * 8;
  float wl[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) wl[t] = wd[f0 + t];
  for (int64_t J0 = blockIdx.x * (int64_t)(NT / 8); J0 < nc; J0 += (int64_t)gridDim.x * (NT / 8)) {
    const int64_t J = J0 + threadIdx.x / 8;
    float acc = 0.0f;
    if (J < nc) {
      float v[8];
      Vec8<uint16_t>::load(&hc[J * F + f0], v);
#pragma unroll
      for (int t = 0; t < 8; ++t) acc += v[t] * wl[t];
    }
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    if (q == 0 && J < nc) u[J] = acc;
  }
}

__global__ __launch_bounds__(NT) void k_unpool_decode(int64_t n, const float2* __restrict__ xs,
                                                      const int32_t* __restrict__ agg, const float* __restrict__ w,
                                                      const float* __restrict__ u, const float* __restrict__ we,
                                                      const float* __restrict__ be, const float* __restrict__ wd,
                                                      const float* __restrict__ bd, const float* __restrict__ d,
                                                      const double* __restrict__ s_in, double* __restrict__ z,
                                                      const int* stop) {
  if (stopped(stop)) return;
  __shared__ float cs[3];
  if (threadIdx.x < kWave) {
    const int f = threadIdx.x;
    float c0 = we[f] * wd[f], c1 = we[F + f] * wd[f], cb = be[f] * wd[f];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      c0 += __shfl_xor(c0, o, kWave);
      c1 += __shfl_xor(c1, o, kWave);
      cb += __shfl_xor(cb, o, kWave);
    }
    if (f == 0) {
      cs[0] = c0;
      cs[1] = c1;
      cs[2] = cb + bd[0];
    }
  }
  __syncthreads();
  const double s = *s_in;
  const float c0 = cs[0], c1 = cs[1], cb = cs[2];
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const float2 xi = xs[i];
    z[i] = s * (double)d[i] * (double)(xi.x * c0 + xi.y * c1 + cb + w[i] * u[agg[i]]);
  }
}

// ---------------------------------------------------------------------------
// Symmetric network with a layer pair on level 0, at inference.
// The encoder output x (x) we is rank one, so the level-0 layer
// L0(h) = h + h W_h + (A^ h) W_n maps it to  x (x) a + (A^x) (x) b  with
// a = we + we W_h, b = we W_n, and the decoder of the adjoint layer reads
//   y_i = u_i . a + (A^ (u . b))_i,   u = x (x) a + (A^x) (x) b + w hc[agg]
// (hc: level-1 features of the up path).  So the whole finest level is two
// scalar CSR products and a few per-node scalars:
//   k_sym0_down : xs_i = (x_i, (A^x)_i),  x = d r / s, from u = d^2 r and A's
//                 own values: (A^x)_i = d_i (A u)_i / s (A^ = D A D, D = diag d)
//   (k_encode_pool with xs as input pools x (x) a + (A^x) (x) b)
//   k_row_dot2  : ua[J] = hc[J] . a,  ub[J] = hc[J] . b
//   k_sym0_up   : t_j = d_j (x_j ab + (A^x)_j bb + w_j ub[agg_j])  (pre-scaled by d);
//                 y_i = x_i aa + (A^x)_i ab + w_i ua[agg_i] + d_i (A t)_i,  z = s d y
// Using A's values (which the flexible-CG SpMV streams anyway) instead of a
// second, normalised copy keeps one value array per iteration in the caches.
// where aa = a.a, ab = a.b, bb = b.b.
// ---------------------------------------------------------------------------
// (A^ v)_i for float v on 4-lane groups with the unrolled row loop
template <int S>
__device__ __forceinline__ float row_dot_f(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                           const float* __restrict__ av, const float* __restrict__ v, int64_t row,
                                           int sl) {
  const int b = rp[row], e = rp[row + 1];
  float acc = 0.0f;
  constexpr int U = S <= 4 ? 4 : 2;  // entries per lane per batch, loads issued together
  for (int k0 = b + sl; k0 < e; k0 += U * S) {
    int c[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = k0 + u * S < e;
      c[u] = ok[u] ? ci[k0 + u * S] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc += ok[u] ? av[k0 + u * S] * v[c[u]] : 0.0f;
  }
#pragma unroll
  for (int o = S / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
  return acc;
}

// u = d^2 r (FP32) when the flexible-CG update has not written it
__global__ __launch_bounds__(NT) void k_sym0_dr(int64_t n, const double* __restrict__ r, const float* __restrict__ d,
                                                float* __restrict__ u, const int* stop) {
  if (stopped(stop)) return;
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
    u[i] = (float)((double)d[i] * ((double)d[i] * r[i]));
}

// xs_i = (x_i, (A^x)_i), x = d r / s, from u = d^2 r (FP32) and A's values av
template <int S>
__global__ __launch_bounds__(NT) void k_sym0_down(int64_t n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci, const float* __restrict__ av,
                                                  const float* __restrict__ d, const float* __restrict__ u,
                                                  const double* __restrict__ s_in, float2* __restrict__ xs,
                                                  const int* stop) {
  if (stopped(stop)) return;
  const float inv = (float)(1.0 / *s_in);
  const int sl = threadIdx.x % S;
  const tmx::RowChunk rc = tmx::row_chunk(n);
  for (int64_t i = rc.r0 + threadIdx.x / S; i < rc.r1; i += NT / S) {
    const float ui = sl == 0 ? u[i] : 0.0f, di = sl == 0 ? d[i] : 1.0f;  // issued before the gathers
    const float acc = row_dot_f<S>(rp, ci, av, u, i, sl);
    if (sl == 0) xs[i] = float2{ui / di * inv, di * acc * inv};
  }
}

__global__ __launch_bounds__(NT) void k_row_dot2(int64_t nc, const uint16_t* __restrict__ hc,
                                                 const float* __restrict__ va, const float* __restrict__ vb,
                                                 float* __restrict__ ua, float* __restrict__ ub, const int* stop) {
  if (stopped(stop)) return;
  const int q = threadIdx.x % 8, f0 = q * 8;
  float la[8], lb[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    la[t] = va[f0 + t];
    lb[t] = vb[f0 + t];
  }
  for (int64_t J0 = blockIdx.x * (int64_t)(NT / 8); J0 < nc; J0 += (int64_t)gridDim.x * (NT / 8)) {
    const int64_t J = J0 + threadIdx.x / 8;
    float sa = 0.0f, sb = 0.0f;
    if (J < nc) {
      float v[8];
      Vec8<uint16_t>::load(&hc[J * F + f0], v);
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        sa += v[t] * la[t];
        sb += v[t] * lb[t];
      }
    }
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) {
      sa += __shfl_xor(sa, o, kWave);
      sb += __shfl_xor(sb, o, kWave);
    }
    if (q == 0 && J < nc) {
      ua[J] = sa;
      ub[J] = sb;
    }
  }
}

__global__ __launch_bounds__(NT) void k_sym0_t(int64_t n, const float2* __restrict__ xs,
                                               const int32_t* __restrict__ agg, const float* __restrict__ w,
                                               const float* __restrict__ ub, const float* __restrict__ c3,
                                               const float* __restrict__ d, float* __restrict__ t, const int* stop) {
  if (stopped(stop)) return;
  const float ab = c3[1], bb = c3[2];
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const float2 x = xs[i];
    t[i] = d[i] * (x.x * ab + x.y * bb + w[i] * ub[agg[i]]);
  }
}

template <int S>
__global__ __launch_bounds__(NT) void k_sym0_up(int64_t n, const int32_t* __restrict__ rp,
                                                const int32_t* __restrict__ ci, const float* __restrict__ av,
                                                const float2* __restrict__ xs, const int32_t* __restrict__ agg,
                                                const float* __restrict__ w, const float* __restrict__ ua,
                                                const float* __restrict__ t, const float* __restrict__ c3,
                                                const float* __restrict__ d, const double* __restrict__ s_in,
                                                double* __restrict__ z, const int* stop) {
  if (stopped(stop)) return;
  const float aa = c3[0], ab = c3[1];
  const double s = *s_in;
  const int sl = threadIdx.x % S;
  const tmx::RowChunk rc = tmx::row_chunk(n);
  for (int64_t i = rc.r0 + threadIdx.x / S; i < rc.r1; i += NT / S) {
    // the row's own terms do not depend on the product: issue them first
    float2 x = float2{0.0f, 0.0f};
    float wu = 0.0f, di = 0.0f;
    if (sl == 0) {
      x = xs[i];
      wu = w[i] * ua[agg[i]];
      di = d[i];
    }
    const float acc = row_dot_f<S>(rp, ci, av, t, i, sl);
    if (sl == 0) {
      const float y = x.x * aa + x.y * ab + wu + di * acc;
      z[i] = s * (double)di * (double)y;
    }
  }
}

// g_h[i][f] = g_y_i wd[f]; partials dwd[f] = sum h[i][f] g_y_i, dbd = sum g_y_i
__global__ __launch_bounds__(NT) void k_decode_bwd(int64_t n, const uint16_t* __restrict__ h,
                                                   const float* __restrict__ wd, const float* __restrict__ g_y,
                                                   float* __restrict__ g_h, float* __restrict__ part) {
  __shared__ float red[NT / kWave][F + 1];
  const int f = threadIdx.x % F, wv = threadIdx.x / F;
  const float wf = wd[f];
  float a = 0.0f, ab = 0.0f;
  for (int64_t i = blockIdx.x * (int64_t)(NT / F) + wv; i < n; i += (int64_t)gridDim.x * (NT / F)) {
    const float g = g_y[i];
    g_h[i * F + f] = g * wf;
    a += bf2f(h[i * F + f]) * g;
    ab += g;
  }
  red[wv][f] = a;
  if (f == 0) red[wv][F] = ab;
  __syncthreads();
  if (threadIdx.x <= F) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) s += red[k][threadIdx.x];
    part[(int64_t)blockIdx.x * (F + 1) + threadIdx.x] = s;
  }
}

// Training loss (energy functional of the normalised system A^ y = r~):
//   L = (y.A^y / 2 - y.r~) / n,  g_y = (A^ y - r~) / n,  r~ = d r / s
// Partials: [0] y.A^y, [1] y.r~, [2] e~.r~ with e~ = e / (d s) (e optional).
__global__ __launch_bounds__(NT) void k_energy_loss(int64_t n, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci, const float* __restrict__ av,
                                                    const float* __restrict__ y, const double* __restrict__ r,
                                                    const float* __restrict__ d, const double* __restrict__ s_in,
                                                    const double* __restrict__ e, float* __restrict__ g_y,
                                                    double* __restrict__ part) {
  const double s = *s_in;
  const double invn = n > 0 ? 1.0 / (double)n : 0.0;
  double loc[3] = {0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    double ay = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) ay += (double)av[k] * (double)y[ci[k]];
    const double rt = (double)d[i] * r[i] / s;
    g_y[i] = (float)((ay - rt) * invn);
    loc[0] += (double)y[i] * ay;
    loc[1] += (double)y[i] * rt;
    if (e) loc[2] += e[i] / ((double)d[i] * s) * rt;
  }
  const int which[3] = {0, 1, 2};
  tmx::write_parts<NT, 3>(part, kMaxParts, which, loc);
}

// Scale-invariant loss: the A^-norm error left after the optimal step along y,
//   L = 1 - (y.r~)^2 / ((y.A^y)(e~.r~)),
//   dL/dy = c ((1 - t) r~ - t n g),  t = y.r~ / y.A^y,  c = -2 y.r~ / (y.A^y e~.r~),
// applied in place to g = (A^y - r~) / n written by k_energy_loss, whose
// partial sums [y.A^y | y.r~ | e~.r~] give the scalars.
__global__ __launch_bounds__(NT) void k_cosine_grad(int64_t n, const double* __restrict__ r,
                                                    const float* __restrict__ d, const double* __restrict__ s_in,
                                                    const double* __restrict__ part, int nparts,
                                                    float* __restrict__ g_y) {
  const int which[3] = {0, 1, 2};
  double t3[3];
  tmx::sum_parts<3>(part, kMaxParts, which, nparts, t3);
  const double yay = t3[0], yr = t3[1], er = t3[2];
  const bool ok = yay > 0.0 && er > 0.0;
  const double t = ok ? yr / yay : 0.0;
  const double c = ok ? -2.0 * yr / (yay * er) : 0.0;
  const double s = *s_in;
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const double rt = (double)d[i] * r[i] / s;
    g_y[i] = (float)(c * ((1.0 - t) * rt - t * (double)n * (double)g_y[i]));
  }
}

// ---------------------------------------------------------------------------
// Level set-up for one matrix
// ---------------------------------------------------------------------------
// d_i = A_ii^-1/2 (1 where the diagonal is missing or not positive)
// Row kernels of the level set-up: G lanes per row, so a wave's loads and
// stores cover G consecutive entries of each of its rows (coalesced); the
// outer loop is uniform over the workgroup, so the shuffles see whole groups.
constexpr int kRowG = 16;

__global__ __launch_bounds__(NT) void k_level_diag(int64_t n, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci, const float* __restrict__ v,
                                                   float* __restrict__ d) {
  const int sub = threadIdx.x % kRowG;
  for (int64_t i0 = blockIdx.x * (int64_t)(NT / kRowG); i0 < n; i0 += (int64_t)gridDim.x * (NT / kRowG)) {
    const int64_t i = i0 + threadIdx.x / kRowG;
    double a = 0.0;
    if (i < n)
      for (int k = rp[i] + sub; k < rp[i + 1]; k += kRowG)
        if (ci[k] == i) a += (double)v[k];
#pragma unroll
    for (int o = kRowG / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, kRowG);
    if (sub == 0 && i < n) d[i] = a > 0.0 ? (float)(1.0 / sqrt(a)) : 1.0f;
  }
}

__global__ __launch_bounds__(NT) void k_level_normalize(int64_t n, const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci, const float* __restrict__ v,
                                                        const float* __restrict__ d, float* __restrict__ vn) {
  const int sub = threadIdx.x % kRowG;
  for (int64_t i0 = blockIdx.x * (int64_t)(NT / kRowG); i0 < n; i0 += (int64_t)gridDim.x * (NT / kRowG)) {
    const int64_t i = i0 + threadIdx.x / kRowG;
    if (i >= n) continue;
    const float di = d[i];
    for (int k = rp[i] + sub; k < rp[i + 1]; k += kRowG) vn[k] = v[k] * di * d[ci[k]];
  }
}

// Galerkin coarse values: vc[s] = sum over fine entries mapped to coarse entry s
__global__ __launch_bounds__(NT) void k_level_coarsen(int64_t nnz_c, const int32_t* __restrict__ seg,
                                                      const int32_t* __restrict__ perm,
                                                      const float* __restrict__ vf, float* __restrict__ vc) {
  for (int64_t s = blockIdx.x * (int64_t)NT + threadIdx.x; s < nnz_c; s += (int64_t)gridDim.x * NT) {
    double a = 0.0;
    for (int k = seg[s]; k < seg[s + 1]; ++k) a += (double)vf[perm[k]];
    vc[s] = (float)a;
  }
}

// pooling weights w_i = d_c[agg_i] / d_i  (= sqrt(A_ii / Ac_II))
__global__ __launch_bounds__(NT) void k_level_poolw(int64_t nf, const int32_t* __restrict__ agg,
                                                    const float* __restrict__ df, const float* __restrict__ dc,
                                                    float* __restrict__ w) {
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < nf; i += (int64_t)gridDim.x * NT)
    w[i] = dc[agg[i]] / df[i];
}

// B = A^ P^ on the finest level: bv[s] = sum over the fine entries k mapped to
// B entry s of A^_k w_col(k)   (fixed order through perm / seg)
__global__ __launch_bounds__(NT) void k_level_prolong(int64_t nnz_b, const int32_t* __restrict__ seg,
                                                      const int32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ ci, const float* __restrict__ vn,
                                                      const float* __restrict__ w, float* __restrict__ bv) {
  for (int64_t s = blockIdx.x * (int64_t)NT + threadIdx.x; s < nnz_b; s += (int64_t)gridDim.x * NT) {
    double a = 0.0;
    for (int k = seg[s]; k < seg[s + 1]; ++k) {
      const int e = perm[k];
      a += (double)vn[e] * (double)w[ci[e]];
    }
    bv[s] = (float)a;
  }
}

__global__ __launch_bounds__(NT) void k_row_sums(int64_t n, const int32_t* __restrict__ rp,
                                                 const float* __restrict__ v, float* __restrict__ out) {
  const int sub = threadIdx.x % kRowG;
  for (int64_t i0 = blockIdx.x * (int64_t)(NT / kRowG); i0 < n; i0 += (int64_t)gridDim.x * (NT / kRowG)) {
    const int64_t i = i0 + threadIdx.x / kRowG;
    double a = 0.0;
    if (i < n)
      for (int k = rp[i] + sub; k < rp[i + 1]; k += kRowG) a += (double)v[k];
#pragma unroll
    for (int o = kRowG / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, kRowG);
    if (sub == 0 && i < n) out[i] = (float)a;
  }
}

// Adam (bias-corrected), FP32 master weights; entries with mask[i] == 0 stay fixed
__global__ __launch_bounds__(NT) void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                                             float* __restrict__ m, float* __restrict__ v,
                                             const float* __restrict__ mask, float lr, float b1, float b2, float eps,
                                             float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    if (mask && mask[i] == 0.0f) continue;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// ---------------------------------------------------------------------------
// Flexible CG truncated to the last m directions (FCG(m), Notay 2000):
//   z = M(r); p = z - sum_j beta_j p_j, beta_j = (z, q_j) / (p_j, q_j)
//   q = A p; alpha = (p, r) / (p, q); x += alpha p; r -= alpha q
// over a ring of m stored (p_j, q_j = A p_j), evaluated with one reduction
// phase per iteration: w = A z and the dots (z,q_j), (p_j,r), (z,r), (z,w)
// come out of one SpMV kernel, a single workgroup turns their partial sums
// into the betas and alpha (the stored directions being mutually A-orthogonal)
//   q = w - sum beta_j q_j,  (p,r) = (z,r) - sum beta_j (p_j,r),
//   (p,q) = (z,w) - 2 sum beta_j (z,q_j) + sum beta_j^2 (p_j,q_j)
// and one update kernel applies everything.  Unfilled slots hold zeros with
// (p_j,q_j) = 1, so their beta_j is 0.
// ---------------------------------------------------------------------------
constexpr int kFcgMaxM = 8;
struct FcgScalars {
  double bb, tol2, relres, s, pq_new, alpha;
  double pq[kFcgMaxM];
  double beta[kFcgMaxM];
};
struct FcgFlags {
  int status;  // 0 running, 1 converged, 2 breakdown, 3 max iterations
  int iter;
  long long maxiter;
};
// partial slots: (z,q_j) [m] | (p_j,r) [m] | zr | zw | rr | bb | ss
constexpr int S_ZQ = 0, S_PR = kFcgMaxM, S_ZR = 2 * kFcgMaxM, S_ZW = S_ZR + 1, S_RR = S_ZR + 2, S_BB = S_ZR + 3,
              S_SS = S_ZR + 4;
constexpr int kFcgSlots = S_SS + 1;

struct FcgWork {
  double *r, *z, *w, *P, *Q, *part;  // P, Q: [m][n]
  float* dr;                         // d r in FP32 (input of a finest-level CSR product)
  FcgScalars* s;
  FcgFlags* f;
  int m;
};

FcgWork fcg_carve(void* work, int64_t n, int m) {
  FcgWork w;
  double* b = (double*)work;
  w.r = b;
  w.z = b + n;
  w.w = b + 2 * n;
  w.P = b + 3 * n;
  w.Q = w.P + (int64_t)m * n;
  w.part = w.Q + (int64_t)m * n;
  w.s = (FcgScalars*)(w.part + kFcgSlots * kMaxParts);
  w.f = (FcgFlags*)(w.s + 1);
  w.dr = (float*)(w.f + 1);
  w.m = m;
  return w;
}

// START: w = A x (the residual is formed by k_fcg_init).  Otherwise w = A z
// with partials (z,q_j), (p_j,r), (z,r), (z,w).
template <int S, int M, bool START>
__global__ __launch_bounds__(NTF) void k_fcg_spmv(int64_t n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci, const float* __restrict__ v,
                                                  const double* __restrict__ x, FcgWork w) {
  if (!START && w.f->status != 0) return;
  constexpr int R = 2;
  constexpr int64_t stride = NTF / S;
  const int sl = threadIdx.x % S;
  const tmx::RowChunk rc = tmx::row_chunk(n);
  const int lim = rp[rc.r1];
  const double* src = START ? x : w.z;
  double loc[2 * M + 2];
#pragma unroll
  for (int k = 0; k < 2 * M + 2; ++k) loc[k] = 0.0;
  for (int64_t row = rc.r0 + threadIdx.x / S; row < rc.r1; row += stride * R) {
    // the row's own vector entries do not depend on the product: load them
    // before the gathers so their latency overlaps
    double zi[R], ri[R], qi[R][M], pi[R][M];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = row + r * stride;
      zi[r] = ri[r] = 0.0;
#pragma unroll
      for (int j = 0; j < M; ++j) qi[r][j] = pi[r][j] = 0.0;
      if (!START && sl == 0 && i < rc.r1) {
        zi[r] = w.z[i];
        ri[r] = w.r[i];
#pragma unroll
        for (int j = 0; j < M; ++j) {
          qi[r][j] = w.Q[(int64_t)j * n + i];
          if (M > 1) pi[r][j] = w.P[(int64_t)j * n + i];
        }
      }
    }
    double o[R][1];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      o[r][0] = 0.0;
      if (row + r * stride < rc.r1) tmx::row_accum<float, S, 1>(rp, ci, v, src, row + r * stride, sl, lim, o[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      tmx::sub_reduce<S, 1>(o[r]);
      const int64_t i = row + r * stride;
      if (sl == 0 && i < rc.r1) {
        w.w[i] = o[r][0];
        if (!START) {
#pragma unroll
          for (int j = 0; j < M; ++j) {
            loc[j] += zi[r] * qi[r][j];
            // with one direction (p_old, r) = 0: alpha made the new residual orthogonal to it
            if (M > 1) loc[M + j] += pi[r][j] * ri[r];
          }
          loc[2 * M] += zi[r] * ri[r];
          loc[2 * M + 1] += zi[r] * o[r][0];
        }
      }
    }
  }
  if (!START) {
    int which[2 * M + 2];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      which[j] = S_ZQ + j;
      which[M + j] = S_PR + j;
    }
    which[2 * M] = S_ZR;
    which[2 * M + 1] = S_ZW;
    tmx::write_parts<NTF, 2 * M + 2>(w.part, kMaxParts, which, loc);
  }
}

// r = b - w (w = A x0), P = Q = 0; partials (r,r), (b,b), (d r)^2
template <int M>
__global__ __launch_bounds__(NTF) void k_fcg_init(int64_t n, const double* __restrict__ b, const float* __restrict__ d,
                                                  FcgWork w) {
  double loc[3] = {0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)NTF + threadIdx.x; i < n; i += (int64_t)gridDim.x * NTF) {
    const double ri = b[i] - w.w[i];
    w.r[i] = ri;
#pragma unroll
    for (int j = 0; j < M; ++j) {
      w.P[(int64_t)j * n + i] = 0.0;
      w.Q[(int64_t)j * n + i] = 0.0;
    }
    loc[0] += ri * ri;
    loc[1] += b[i] * b[i];
    const double dr = (double)d[i] * ri;
    w.dr[i] = (float)((double)d[i] * dr);  // u = d^2 r for the fused finest level
    loc[2] += dr * dr;
  }
  const int which[3] = {S_RR, S_BB, S_SS};
  tmx::write_parts<NTF, 3>(w.part, kMaxParts, which, loc);
}

__global__ __launch_bounds__(NT) void k_fcg_ctl0(FcgWork w, int64_t n, int nparts, double rtol2, double atol2,
                                                 long long maxiter) {
  const int which[3] = {S_RR, S_BB, S_SS};
  double t[3];
  tmx::sum_parts<3>(w.part, kMaxParts, which, nparts, t);
  if (threadIdx.x == 0) {
    const double sc = n > 0 ? sqrt(t[2] / (double)n) : 0.0;
    w.s->s = sc > 0.0 ? sc : 1.0;
    w.s->bb = t[1];
    w.s->tol2 = fmax(rtol2 * t[1], atol2);
    for (int j = 0; j < kFcgMaxM; ++j) w.s->pq[j] = 1.0;
    w.s->pq_new = 1.0;
    w.s->relres = t[1] > 0.0 ? sqrt(t[0] / t[1]) : 0.0;
    w.f->iter = 0;
    w.f->maxiter = maxiter;
    w.f->status = (t[0] <= w.s->tol2) ? 1 : 0;
  }
}

// p, q into ring slot iter % m; x += alpha p, r -= alpha q; partials (r,r), (d r)^2.
// Every workgroup first forms the betas and alpha from the SpMV partials (one
// wave per slot, lanes strided over the partials in a fixed order, so every
// workgroup holds the same bits); workgroup 0 records them for k_fcg_check.
template <int M>
__global__ __launch_bounds__(NTF) void k_fcg_update(int64_t n, double* __restrict__ x, const float* __restrict__ d,
                                                    FcgWork w, int nparts) {
  if (w.f->status != 0) return;
  __shared__ double sh[2 * M + 2];
  __shared__ double sc[M + 2];
  {
    const int wv = threadIdx.x / kWave, l = threadIdx.x % kWave;
    for (int k = wv; k < 2 * M + 2; k += NTF / kWave) {
      const int slot = k < M ? S_ZQ + k : (k < 2 * M ? S_PR + (k - M) : (k == 2 * M ? S_ZR : S_ZW));
      double a = 0.0;
      for (int b = l; b < nparts; b += kWave) a += w.part[slot * kMaxParts + b];
      a = tmx::wave_sum(a);
      if (l == 0) sh[k] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double pr = sh[2 * M], pq = sh[2 * M + 1];
      for (int j = 0; j < M; ++j) {
        const double bj = sh[j] / w.s->pq[j];
        sc[j] = bj;
        pr -= bj * sh[M + j];
        pq += bj * (bj * w.s->pq[j] - 2.0 * sh[j]);
      }
      const bool okq = pq > 0.0 && __builtin_isfinite(pq) && __builtin_isfinite(pr);
      sc[M] = okq ? pr / pq : 0.0;
      sc[M + 1] = okq ? pq : -1.0;
      if (blockIdx.x == 0) {
        for (int j = 0; j < M; ++j) w.s->beta[j] = sc[j];
        w.s->alpha = sc[M];
        w.s->pq_new = sc[M + 1];
      }
    }
    __syncthreads();
  }
  const double alpha = sc[M];
  const bool ok = sc[M + 1] > 0.0;
  double beta[M];
#pragma unroll
  for (int j = 0; j < M; ++j) beta[j] = sc[j];
  const int slot = w.f->iter % M;
  double loc[2] = {0.0, 0.0};
  if (ok) {
    for (int64_t i = blockIdx.x * (int64_t)NTF + threadIdx.x; i < n; i += (int64_t)gridDim.x * NTF) {
      double pi = w.z[i], qi = w.w[i];
#pragma unroll
      for (int j = 0; j < M; ++j) {
        pi -= beta[j] * w.P[(int64_t)j * n + i];
        qi -= beta[j] * w.Q[(int64_t)j * n + i];
      }
      w.P[(int64_t)slot * n + i] = pi;
      w.Q[(int64_t)slot * n + i] = qi;
      x[i] += alpha * pi;
      const double ri = w.r[i] - alpha * qi;
      w.r[i] = ri;
      loc[0] += ri * ri;
      const double dr = (double)d[i] * ri;
      w.dr[i] = (float)((double)d[i] * dr);  // u = d^2 r for the fused finest level
      loc[1] += dr * dr;
    }
  }
  const int which[2] = {S_RR, S_SS};
  tmx::write_parts<NTF, 2>(w.part, kMaxParts, which, loc);
}

// Single workgroup: (p,q) of the new ring slot, iteration count, convergence,
// and the input scale s = rms(d r) of the next preconditioner application.
__global__ __launch_bounds__(NT) void k_fcg_check(FcgWork w, int64_t n, int nparts) {
  if (w.f->status != 0) return;
  __shared__ double red[NT / kWave][2];
  double a = 0.0, c = 0.0;
  for (int k = threadIdx.x; k < nparts; k += NT) {
    a += w.part[S_RR * kMaxParts + k];
    c += w.part[S_SS * kMaxParts + k];
  }
  a = tmx::wave_sum(a);
  c = tmx::wave_sum(c);
  if (threadIdx.x % kWave == 0) {
    red[threadIdx.x / kWave][0] = a;
    red[threadIdx.x / kWave][1] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double rr = 0.0, ss = 0.0;
    for (int k = 0; k < NT / kWave; ++k) {
      rr += red[k][0];
      ss += red[k][1];
    }
    const double sc = n > 0 ? sqrt(ss / (double)n) : 0.0;
    w.s->s = sc > 0.0 ? sc : 1.0;  // input scale of the next preconditioner application
    const double pq = w.s->pq_new;
    const int it = w.f->iter;
    w.f->iter = it + 1;
    if (!(pq > 0.0) || !__builtin_isfinite(pq) || !__builtin_isfinite(rr)) {
      w.f->status = 2;
      return;
    }
    w.s->pq[it % w.m] = pq;
    w.s->relres = w.s->bb > 0.0 ? sqrt(rr / w.s->bb) : 0.0;
    if (rr <= w.s->tol2)
      w.f->status = 1;
    else if (it + 1 >= w.f->maxiter)
      w.f->status = 3;
  }
}

int grid_of(int64_t work, int per_block, int cap) {
  int64_t g = (work + per_block - 1) / per_block;
  return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

// one grid for every flexible-CG vector kernel: their partial slots hold G entries
int fcg_grid(int64_t n) { return grid_of(n, NTF, kMaxParts); }

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" int tm_gnn_mp_forward(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                                 const uint16_t* h, const uint16_t* w_t, const float* bias, int n_linear,
                                 uint16_t* out, uint16_t* save_msg, uint16_t* save_pre, const int* stop,
                                 void* stream) {
  TM_REQUIRE(n >= 0 && row_ptr && col_idx && values && h && w_t && bias && out, TM_ERR_ARG, "null argument");
  TM_REQUIRE((save_msg == nullptr) == (save_pre == nullptr), TM_ERR_ARG, "save_msg and save_pre go together");
  TM_REQUIRE(out != h, TM_ERR_ARG, "tm_gnn_mp_forward: out must not alias h");
  if (n == 0) return TM_OK;
  const int64_t tiles = (n + TILE - 1) / TILE;
  TM_REQUIRE(tiles <= INT32_MAX, TM_ERR_LIMIT, "too many rows");
  hipStream_t st = (hipStream_t)stream;
  if (save_msg)
    hipLaunchKernelGGL(k_mp<true>, dim3((unsigned)tiles), dim3(NT), 0, st, n, row_ptr, col_idx, values, h, w_t,
                       bias, n_linear, out, save_msg, save_pre, stop);
  else if (n < kSmallLevel)
    hipLaunchKernelGGL(k_mp_small, dim3((unsigned)((n + TS - 1) / TS)), dim3(NT), 0, st, n, row_ptr, col_idx, values,
                       h, w_t, bias, n_linear, out, stop);
  else
    hipLaunchKernelGGL(k_mp<false>, dim3((unsigned)tiles), dim3(NT), 0, st, n, row_ptr, col_idx, values, h, w_t,
                       bias, n_linear, out, save_msg, save_pre, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_encode_pool(int64_t n, int64_t n_coarse, const int32_t* child_ptr, const int32_t* child,
                                  const float* w, const double* r, const float* d, const float* lk,
                                  const double* ss_part, int nparts, const float* we, const float* be, uint16_t* out,
                                  float* xs, double* scale, const int* stop, void* stream) {
  TM_REQUIRE(n >= 0 && n_coarse >= 0 && child_ptr && child && w && d && we && be && out && xs && scale,
             TM_ERR_ARG, "null argument");
  TM_REQUIRE(r || nparts == 0, TM_ERR_ARG, "encode_pool: without r the scale must already be stored (nparts 0)");
  TM_REQUIRE(nparts == 0 || ss_part, TM_ERR_ARG, "encode_pool: partials missing");
  TM_REQUIRE(nparts >= 0 && nparts <= kMaxParts, TM_ERR_ARG, "encode_pool: bad partial count %d", nparts);
  const int g = grid_of(n_coarse, NT / 8, 8192);
  hipLaunchKernelGGL(k_encode_pool, dim3(g), dim3(NT), 0, (hipStream_t)stream, n, n_coarse, child_ptr, child, w, r,
                     d, lk, ss_part, nparts, we, be, out, (float2*)xs, scale, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_final(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                            const float* xs, const float* row_sums, const int32_t* agg, const float* w,
                            const uint16_t* hc, const int32_t* b_row_ptr, const int32_t* b_col, const float* b_values,
                            const float* we, const float* be, const uint16_t* w_t, const float* bias, int n_linear,
                            const float* wd, const float* bd, const float* d, const double* scale, double* z,
                            const int* stop, void* stream) {
  TM_REQUIRE(n >= 0 && row_ptr && col_idx && values && xs && row_sums && agg && w && hc && b_row_ptr && b_col &&
                 b_values && we && be && w_t && bias && wd && bd && d && scale && z,
             TM_ERR_ARG, "null argument");
  if (n == 0) return TM_OK;
  const int64_t tiles = (n + TILE - 1) / TILE;
  TM_REQUIRE(tiles <= INT32_MAX, TM_ERR_LIMIT, "too many rows");
  FinalArgs A{row_ptr, col_idx, values, (const float2*)xs, row_sums, agg, w, hc, b_row_ptr, b_col, b_values, we, be,
              w_t, bias, n_linear, wd, bd, d, scale, z};
  hipLaunchKernelGGL(k_final, dim3((unsigned)tiles), dim3(NT), 0, (hipStream_t)stream, n, A, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_unpool_decode(int64_t n, int64_t n_coarse, const float* xs, const int32_t* agg,
                                    const float* w, const uint16_t* hc, const float* we, const float* be,
                                    const float* wd, const float* bd, const float* d, const double* scale,
                                    float* coarse_dot, double* z, const int* stop, void* stream) {
  TM_REQUIRE(n >= 0 && n_coarse >= 0 && xs && agg && w && hc && we && be && wd && bd && d && scale && coarse_dot &&
                 z,
             TM_ERR_ARG, "null argument");
  if (n == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (n_coarse > 0)
    hipLaunchKernelGGL(k_row_dot, dim3(grid_of(n_coarse, NT / 8, 8192)), dim3(NT), 0, st, n_coarse, hc, wd,
                       coarse_dot, stop);
  hipLaunchKernelGGL(k_unpool_decode, dim3(grid_of(n, NT, 8192)), dim3(NT), 0, st, n, (const float2*)xs, agg, w,
                     coarse_dot, we, be, wd, bd, d, scale, z, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_sym0_down(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                                const double* r, const float* d, float* u, int u_ready, const double* scale,
                                float* xs, const int* stop, void* stream) {
  TM_REQUIRE(n >= 0 && row_ptr && col_idx && values && d && u && scale && xs, TM_ERR_ARG, "null argument");
  TM_REQUIRE(u_ready || r, TM_ERR_ARG, "sym0_down: r needed when u is not ready");
  if (n == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (!u_ready)
    hipLaunchKernelGGL(k_sym0_dr, dim3(grid_of(n, NT, 8192)), dim3(NT), 0, st, n, r, d, u, stop);
  hipLaunchKernelGGL(k_sym0_down<4>, dim3(grid_of(n, NT / 4 * 4, 4096)), dim3(NT), 0, st, n, row_ptr, col_idx,
                     values, d, u, scale, (float2*)xs, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_sym0_up(int64_t n, int64_t n_coarse, const int32_t* row_ptr, const int32_t* col_idx,
                              const float* values, const float* xs, const int32_t* agg, const float* w,
                              const uint16_t* hc, const float* va, const float* vb, const float* c3,
                              const float* d, const double* scale, float* ua, float* ub, float* t, double* z,
                              const int* stop, void* stream) {
  TM_REQUIRE(n >= 0 && n_coarse >= 0 && row_ptr && col_idx && values && xs && agg && w && hc && va && vb && c3 &&
                 d && scale && ua && ub && t && z,
             TM_ERR_ARG, "null argument");
  if (n == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_row_dot2, dim3(grid_of(n_coarse, NT / 8, 8192)), dim3(NT), 0, st, n_coarse, hc, va, vb, ua,
                     ub, stop);
  hipLaunchKernelGGL(k_sym0_t, dim3(grid_of(n, NT, 8192)), dim3(NT), 0, st, n, (const float2*)xs, agg, w, ub, c3, d,
                     t, stop);
  hipLaunchKernelGGL(k_sym0_up<4>, dim3(grid_of(n, NT / 4 * 4, 4096)), dim3(NT), 0, st, n, row_ptr, col_idx,
                     values, (const float2*)xs, agg, w, ua, t, c3, d, scale, z, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_level_prolong(int64_t nnz_b, const int32_t* seg, const int32_t* perm, const int32_t* col_idx,
                                    const float* values_normalized, const float* w, float* b_values, void* stream) {
  TM_REQUIRE(nnz_b >= 0 && seg && perm && col_idx && values_normalized && w && b_values, TM_ERR_ARG, "null argument");
  if (nnz_b == 0) return TM_OK;
  hipLaunchKernelGGL(k_level_prolong, dim3(grid_of(nnz_b, NT, 16384)), dim3(NT), 0, (hipStream_t)stream, nnz_b, seg,
                     perm, col_idx, values_normalized, w, b_values);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_row_sums(int64_t n, const int32_t* row_ptr, const float* values, float* out, void* stream) {
  TM_REQUIRE(n >= 0 && row_ptr && values && out, TM_ERR_ARG, "null argument");
  if (n == 0) return TM_OK;
  hipLaunchKernelGGL(k_row_sums, dim3(grid_of(n, NT / kRowG, 8192)), dim3(NT), 0, (hipStream_t)stream, n, row_ptr, values,
                     out);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int64_t tm_gnn_part_floats(void) { return (int64_t)kMaxParts * NWP; }

extern "C" int tm_gnn_mp_backward(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                                  const float* g_out, const uint16_t* h, const uint16_t* msg, const uint16_t* pre,
                                  const uint16_t* w, int n_linear, float* g_in, float* g_msg, float* part,
                                  float* grad_w, void* stream) {
  TM_REQUIRE(n >= 0 && row_ptr && col_idx && values && g_out && h && msg && pre && w && g_in && g_msg && part &&
                 grad_w,
             TM_ERR_ARG, "null argument");
  TM_REQUIRE(g_in != g_out, TM_ERR_ARG, "tm_gnn_mp_backward: g_in must not alias g_out");
  hipStream_t st = (hipStream_t)stream;
  const int G = grid_of((n + TILE - 1) / TILE, 1, 512);
  if (n > 0) {
    hipLaunchKernelGGL(k_mp_bwd, dim3(G), dim3(NT), 0, st, n, g_out, h, msg, pre, w, n_linear, g_in, g_msg, part);
    hipLaunchKernelGGL(k_msg_add, dim3(grid_of(n, NT / kWave, 8192)), dim3(NT), 0, st, n, row_ptr, col_idx, values,
                       g_msg, g_in);
    hipLaunchKernelGGL(k_reduce_parts, dim3(grid_of(NWP, NT, 64)), dim3(NT), 0, st, G, NWP, part, grad_w, 1);
  }
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_pool(int64_t n_coarse, const int32_t* child_ptr, const int32_t* child, const float* w,
                           const void* h, void* out, int dtype, const int* stop, void* stream) {
  TM_REQUIRE(n_coarse >= 0 && child_ptr && child && w && h && out, TM_ERR_ARG, "null argument");
  TM_REQUIRE(dtype == TM_BF16 || dtype == TM_F32, TM_ERR_ARG, "pool: dtype must be bf16 or f32");
  if (n_coarse == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  const int g = grid_of(n_coarse, NT / 8, 8192);
  if (dtype == TM_BF16)
    hipLaunchKernelGGL(k_pool<uint16_t>, dim3(g), dim3(NT), 0, st, n_coarse, child_ptr, child, w,
                       (const uint16_t*)h, (uint16_t*)out, stop);
  else
    hipLaunchKernelGGL(k_pool<float>, dim3(g), dim3(NT), 0, st, n_coarse, child_ptr, child, w, (const float*)h,
                       (float*)out, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_unpool(int64_t n_fine, const int32_t* agg, const float* w, const void* hc, const void* skip,
                             void* out, int dtype, const int* stop, void* stream) {
  TM_REQUIRE(n_fine >= 0 && agg && w && hc && out, TM_ERR_ARG, "null argument");
  TM_REQUIRE(dtype == TM_BF16 || dtype == TM_F32, TM_ERR_ARG, "unpool: dtype must be bf16 or f32");
  if (n_fine == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  const int g = grid_of(n_fine * 8, NT, 16384);
  if (dtype == TM_BF16)
    hipLaunchKernelGGL(k_unpool<uint16_t>, dim3(g), dim3(NT), 0, st, n_fine, agg, w, (const uint16_t*)hc,
                       (const uint16_t*)skip, (uint16_t*)out, stop);
  else
    hipLaunchKernelGGL(k_unpool<float>, dim3(g), dim3(NT), 0, st, n_fine, agg, w, (const float*)hc,
                       (const float*)skip, (float*)out, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_sumsq(int64_t n, const double* r, const float* d, double* ss_part, int* nparts,
                            void* stream) {
  TM_REQUIRE(n >= 0 && r && d && ss_part && nparts, TM_ERR_ARG, "null argument");
  const int g = grid_of(n, NT * 4, kMaxParts);
  *nparts = g;
  hipLaunchKernelGGL(k_sumsq, dim3(g), dim3(NT), 0, (hipStream_t)stream, n, r, d, ss_part);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_scale(int64_t n, const double* ss_part, int nparts, double* scale, void* stream) {
  TM_REQUIRE(n >= 0 && ss_part && scale && nparts >= 1 && nparts <= kMaxParts, TM_ERR_ARG, "bad scale arguments");
  hipLaunchKernelGGL(k_scale, dim3(1), dim3(NT), 0, (hipStream_t)stream, n, ss_part, nparts, scale);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_encode(int64_t n, const double* r, const float* d, const float* lk, const double* ss_part,
                             int nparts, const float* we, const float* be, uint16_t* h0, double* scale,
                             const int* stop, void* stream) {
  TM_REQUIRE(n >= 0 && r && d && lk && ss_part && we && be && h0 && scale, TM_ERR_ARG, "null argument");
  TM_REQUIRE(nparts >= 1 && nparts <= kMaxParts, TM_ERR_ARG, "encode: bad partial count %d", nparts);
  const int g = grid_of(n, NT / 8, 4096);
  hipLaunchKernelGGL(k_encode, dim3(g), dim3(NT), 0, (hipStream_t)stream, n, r, d, lk, ss_part, nparts, we, be,
                     h0, scale, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_encode_backward(int64_t n, const double* r, const float* d, const float* lk,
                                      const double* scale, const float* g_h0, float* part, float* grad,
                                      void* stream) {
  TM_REQUIRE(n >= 0 && r && d && lk && scale && g_h0 && part && grad, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int G = grid_of(n, NT / F * 16, kMaxParts);
  hipLaunchKernelGGL(k_encode_bwd, dim3(G), dim3(NT), 0, st, n, r, d, lk, scale, g_h0, part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(NT), 0, st, G, 3 * F, part, grad, 1);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_decode(int64_t n, const uint16_t* h, const float* wd, const float* bd, const float* d,
                             const double* scale, float* y, double* z, const int* stop, void* stream) {
  TM_REQUIRE(n >= 0 && h && wd && bd && d && scale && z, TM_ERR_ARG, "null argument");
  const int g = grid_of(n, NT / 8, 8192);
  hipLaunchKernelGGL(k_decode, dim3(g), dim3(NT), 0, (hipStream_t)stream, n, h, wd, bd, d, scale, y, z, stop);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_decode_backward(int64_t n, const uint16_t* h, const float* wd, const float* g_y, float* g_h,
                                      float* part, float* grad, void* stream) {
  TM_REQUIRE(n >= 0 && h && wd && g_y && g_h && part && grad, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int G = grid_of(n, NT / F * 16, kMaxParts);
  hipLaunchKernelGGL(k_decode_bwd, dim3(G), dim3(NT), 0, st, n, h, wd, g_y, g_h, part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(NT), 0, st, G, F + 1, part, grad, 1);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_energy_loss(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                                  const float* y, const double* r, const float* d, const double* scale,
                                  const double* e, float* g_y, double* part, int* nparts, void* stream) {
  TM_REQUIRE(n >= 0 && row_ptr && col_idx && values && y && r && d && scale && g_y && part && nparts, TM_ERR_ARG,
             "null argument");
  const int g = grid_of(n, NT * 4, kMaxParts);
  *nparts = g;
  hipLaunchKernelGGL(k_energy_loss, dim3(g), dim3(NT), 0, (hipStream_t)stream, n, row_ptr, col_idx, values, y, r, d,
                     scale, e, g_y, part);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_cosine_grad(int64_t n, const double* r, const float* d, const double* scale,
                                  const double* part, int nparts, float* g_y, void* stream) {
  TM_REQUIRE(n >= 0 && r && d && scale && part && g_y, TM_ERR_ARG, "null argument");
  TM_REQUIRE(nparts >= 1 && nparts <= kMaxParts, TM_ERR_ARG, "cosine_grad: bad partial count %d", nparts);
  hipLaunchKernelGGL(k_cosine_grad, dim3(grid_of(n, NT * 4, 1024)), dim3(NT), 0, (hipStream_t)stream, n, r, d,
                     scale, part, nparts, g_y);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_level_setup(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                                  float* d, float* values_normalized, void* stream) {
  TM_REQUIRE(n >= 0 && row_ptr && col_idx && values && d, TM_ERR_ARG, "null argument");
  if (n == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  const int g = grid_of(n, NT / kRowG, 8192);
  hipLaunchKernelGGL(k_level_diag, dim3(g), dim3(NT), 0, st, n, row_ptr, col_idx, values, d);
  if (values_normalized)
    hipLaunchKernelGGL(k_level_normalize, dim3(g), dim3(NT), 0, st, n, row_ptr, col_idx, values, d,
                       values_normalized);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_level_coarsen(int64_t nnz_coarse, const int32_t* seg, const int32_t* perm,
                                    const float* fine_values, float* coarse_values, void* stream) {
  TM_REQUIRE(nnz_coarse >= 0 && seg && perm && fine_values && coarse_values, TM_ERR_ARG, "null argument");
  if (nnz_coarse == 0) return TM_OK;
  hipLaunchKernelGGL(k_level_coarsen, dim3(grid_of(nnz_coarse, NT, 16384)), dim3(NT), 0, (hipStream_t)stream,
                     nnz_coarse, seg, perm, fine_values, coarse_values);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_pool_weights(int64_t n_fine, const int32_t* agg, const float* d_fine, const float* d_coarse,
                                   float* w, void* stream) {
  TM_REQUIRE(n_fine >= 0 && agg && d_fine && d_coarse && w, TM_ERR_ARG, "null argument");
  if (n_fine == 0) return TM_OK;
  hipLaunchKernelGGL(k_level_poolw, dim3(grid_of(n_fine, NT, 8192)), dim3(NT), 0, (hipStream_t)stream, n_fine, agg,
                     d_fine, d_coarse, w);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_gnn_adam(int64_t n, float* param, const float* grad, float* m, float* v, const float* mask,
                           float lr, float beta1, float beta2, float eps, int64_t step, void* stream) {
  TM_REQUIRE(n >= 0 && param && grad && m && v && step >= 1, TM_ERR_ARG, "bad adam arguments");
  if (n == 0) return TM_OK;
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid_of(n, NT, 1024)), dim3(NT), 0, (hipStream_t)stream, n, param, grad, m, v, mask,
                     lr, beta1, beta2, eps, c1, c2);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

// ---- flexible CG ----------------------------------------------------------
extern "C" int64_t tm_fcg_work_size(int64_t n, int m) {
  if (!(m == 1 || m == 2 || m == 4 || m == 8)) return -1;
  return (int64_t)sizeof(double) * ((3 + 2 * (int64_t)m) * n + kFcgSlots * kMaxParts) +
         (int64_t)sizeof(FcgScalars) + (int64_t)sizeof(FcgFlags) + (int64_t)sizeof(float) * n + 64;
}

extern "C" int tm_fcg_layout(int64_t n, int m, void* work, tm_fcg_view* view) {
  TM_REQUIRE(n >= 0 && work && view, TM_ERR_ARG, "null argument");
  TM_REQUIRE(m == 1 || m == 2 || m == 4 || m == 8, TM_ERR_ARG, "fcg: m must be 1, 2, 4 or 8 (got %d)", m);
  FcgWork w = fcg_carve(work, n, m);
  view->r = w.r;
  view->z = w.z;
  view->ss_part = w.part + S_SS * kMaxParts;
  view->scale = &w.s->s;
  view->relres = &w.s->relres;
  view->status = &w.f->status;
  view->iter = &w.f->iter;
  view->nparts = fcg_grid(n);
  view->reserved = 0;
  view->dr = w.dr;
  return TM_OK;
}

template <int M>
int fcg_start_m(int64_t n, const int32_t* rp, const int32_t* ci, const float* v, const double* b, const double* x,
                const float* d, FcgWork w, double rtol, double atol, int64_t maxiter, hipStream_t st) {
  const int G = fcg_grid(n);
  TM_HIP(hipMemsetAsync(w.f, 0, sizeof(FcgFlags), st));
  hipLaunchKernelGGL((k_fcg_spmv<4, M, true>), dim3(G), dim3(NTF), 0, st, n, rp, ci, v, x, w);
  hipLaunchKernelGGL(k_fcg_init<M>, dim3(G), dim3(NTF), 0, st, n, b, d, w);
  hipLaunchKernelGGL(k_fcg_ctl0, dim3(1), dim3(NT), 0, st, w, n, G, rtol * rtol, atol * atol, (long long)maxiter);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_fcg_start(int64_t n, int m, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                            const double* b, const double* x, const float* d, void* work, double rtol, double atol,
                            int64_t maxiter, void* stream) {
  TM_REQUIRE(n > 0 && row_ptr && col_idx && values && b && x && d && work, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  FcgWork w = fcg_carve(work, n, m);
  switch (m) {
    case 1: return fcg_start_m<1>(n, row_ptr, col_idx, values, b, x, d, w, rtol, atol, maxiter, st);
    case 2: return fcg_start_m<2>(n, row_ptr, col_idx, values, b, x, d, w, rtol, atol, maxiter, st);
    case 4: return fcg_start_m<4>(n, row_ptr, col_idx, values, b, x, d, w, rtol, atol, maxiter, st);
    case 8: return fcg_start_m<8>(n, row_ptr, col_idx, values, b, x, d, w, rtol, atol, maxiter, st);
  }
  tmx::set_error("fcg: m must be 1, 2, 4 or 8 (got %d)", m);
  return TM_ERR_ARG;
}

template <int M>
int fcg_iterate_m(int64_t n, const int32_t* rp, const int32_t* ci, const float* v, double* x, const float* d,
                  FcgWork w, hipStream_t st) {
  const int G = fcg_grid(n);
  hipLaunchKernelGGL((k_fcg_spmv<4, M, false>), dim3(G), dim3(NTF), 0, st, n, rp, ci, v, x, w);
  hipLaunchKernelGGL(k_fcg_update<M>, dim3(G), dim3(NTF), 0, st, n, x, d, w, G);
  hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(NT), 0, st, w, n, G);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

// One FCG iteration after the preconditioner wrote z = M(r) into the view.
extern "C" int tm_fcg_iterate(int64_t n, int m, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                              double* x, const float* d, void* work, void* stream) {
  TM_REQUIRE(n > 0 && row_ptr && col_idx && values && x && d && work, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  FcgWork w = fcg_carve(work, n, m);
  switch (m) {
    case 1: return fcg_iterate_m<1>(n, row_ptr, col_idx, values, x, d, w, st);
    case 2: return fcg_iterate_m<2>(n, row_ptr, col_idx, values, x, d, w, st);
    case 4: return fcg_iterate_m<4>(n, row_ptr, col_idx, values, x, d, w, st);
    case 8: return fcg_iterate_m<8>(n, row_ptr, col_idx, values, x, d, w, st);
  }
  tmx::set_error("fcg: m must be 1, 2, 4 or 8 (got %d)", m);
  return TM_ERR_ARG;
}
