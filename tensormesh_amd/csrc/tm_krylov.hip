// CSR SpMV and Jacobi-preconditioned CG for gfx950.
//
// Matrix values are FP32 (or FP64), vectors and every reduction FP64.  One PCG
// iteration is three stream-ordered kernels:
//   K1  q = A p and (p, q)               -> alpha = rz / pq
//   K2  x += alpha p, r -= alpha q, z = D^-1 r, (r, z) and (r, r)
//                                        -> beta, convergence test
//   K3  p = D^-1 r + beta p
// Each reduction is finished inside its own kernel by the last workgroup to
// arrive (fixed-order sum of the per-workgroup partials, so the result does
// not depend on arrival order).  Scalars live in device memory; kernels return
// immediately once the device-side status flag is set, so the host only reads
// the flag every few iterations.
#include <vector>

#include "tm_device.h"

namespace {

using tmx::kWave;
constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

// Device-side PCG state (doubles), followed by integer flags.
struct PcgScalars {
  double rz, pq, rr, alpha, beta, bb, tol2, relres;
};
struct PcgFlags {
  unsigned int counter[2];
  int status;  // 0 running, 1 converged, 2 breakdown (p.Ap <= 0), 3 max iterations
  int iter;
  long long maxiter;
};

struct PcgWork {
  double* r;
  double* p;
  double* q;
  double* dinv;
  double* part;  // [3][kMaxBlocks]
  PcgScalars* s;
  PcgFlags* f;
};

PcgWork carve_work(void* work, int64_t n) {
  PcgWork w;
  double* d = (double*)work;
  w.r = d;
  w.p = d + n;
  w.q = d + 2 * n;
  w.dinv = d + 3 * n;
  w.part = d + 4 * n;
  w.s = (PcgScalars*)(w.part + 3 * kMaxBlocks);
  w.f = (PcgFlags*)(w.s + 1);
  return w;
}

// out[i] = sum over the block row of B[i][j] x[col*BS + j], reduced over the S
// lanes that share the row (BS = 1: plain CSR; BS = 3: 3x3 blocks stored
// row-major per CSR entry).
template <typename VT, int S, int BS>
__device__ __forceinline__ void row_dot(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                        const VT* __restrict__ v, const double* __restrict__ x, int64_t row,
                                        int sl, double (&out)[BS]) {
  const int b = rp[row], e = rp[row + 1];
#pragma unroll
  for (int i = 0; i < BS; ++i) out[i] = 0.0;
  for (int k = b + sl; k < e; k += S) {
    const int64_t c = ci[k];
    double xc[BS];
#pragma unroll
    for (int j = 0; j < BS; ++j) xc[j] = x[c * BS + j];
#pragma unroll
    for (int i = 0; i < BS; ++i)
#pragma unroll
      for (int j = 0; j < BS; ++j) out[i] += (double)v[(int64_t)k * BS * BS + i * BS + j] * xc[j];
  }
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int o = S / 2; o > 0; o >>= 1) out[i] += __shfl_xor(out[i], o, kWave);
}

template <typename VT, int S, int BS>
__global__ __launch_bounds__(kThreads) void k_spmv(int64_t nrows, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci, const VT* __restrict__ v,
                                                  const double* __restrict__ x, double* __restrict__ y) {
  const int sl = threadIdx.x % S;
  const int64_t per = kThreads / S;
  for (int64_t row = blockIdx.x * per + threadIdx.x / S; row < nrows; row += (int64_t)gridDim.x * per) {
    double o[BS];
    row_dot<VT, S, BS>(rp, ci, v, x, row, sl, o);
    if (sl == 0)
#pragma unroll
      for (int i = 0; i < BS; ++i) y[row * BS + i] = o[i];
  }
}

// Last workgroup to arrive finishes the reduction of NS partial sums.
template <int NS>
__device__ __forceinline__ bool finish_reduce(double (&local)[NS], double* part, unsigned int* counter,
                                              double (&tot)[NS]) {
  __shared__ double red[kThreads / kWave];
  __shared__ bool last;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double b = tmx::block_sum<kThreads>(local[i], red);
    if (threadIdx.x == 0) part[i * kMaxBlocks + blockIdx.x] = b;
  }
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    double s = 0.0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += kThreads) s += tmx::load_agent(part + i * kMaxBlocks + k);
    tot[i] = tmx::block_sum<kThreads>(s, red);
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

template <typename VT, int BS>
__global__ __launch_bounds__(kThreads) void k_diag_inv(int64_t nrows, const int32_t* __restrict__ rp,
                                                      const int32_t* __restrict__ ci, const VT* __restrict__ v,
                                                      double* __restrict__ dinv) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * kThreads) {
    double d[BS];
#pragma unroll
    for (int c = 0; c < BS; ++c) d[c] = 1.0;  // missing diagonal -> 1 (torch_solve.py:24-45)
    bool found = false;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) {
#pragma unroll
        for (int c = 0; c < BS; ++c) {
          const double vk = (double)v[(int64_t)k * BS * BS + c * BS + c];
          d[c] = found ? d[c] + vk : vk;
        }
        found = true;
      }
#pragma unroll
    for (int c = 0; c < BS; ++c) dinv[i * BS + c] = 1.0 / d[c];
  }
}

// r = b - q (q = A x0), p = D^-1 r, rz = (r, p), rr = (r, r), bb = (b, b)
__global__ __launch_bounds__(kThreads) void k_pcg_init(int64_t n, const double* __restrict__ b, PcgWork w,
                                                      double rtol2, double atol2) {
  double loc[3] = {0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double bi = b[i];
    const double ri = bi - w.q[i];
    const double zi = w.dinv[i] * ri;
    w.r[i] = ri;
    w.p[i] = zi;
    loc[0] += ri * zi;
    loc[1] += ri * ri;
    loc[2] += bi * bi;
  }
  double tot[3];
  if (finish_reduce<3>(loc, w.part, &w.f->counter[0], tot) && threadIdx.x == 0) {
    PcgScalars* s = w.s;
    s->rz = tot[0];
    s->rr = tot[1];
    s->bb = tot[2];
    s->tol2 = fmax(rtol2 * tot[2], atol2);
    s->relres = tot[2] > 0.0 ? sqrt(tot[1] / tot[2]) : 0.0;
    w.f->iter = 0;
    w.f->status = (tot[1] <= s->tol2) ? 1 : 0;
  }
}

template <typename VT, int S, int BS>
__global__ __launch_bounds__(kThreads) void k_pcg_spmv(int64_t nrows, const int32_t* __restrict__ rp,
                                                      const int32_t* __restrict__ ci, const VT* __restrict__ v,
                                                      PcgWork w) {
  if (w.f->status != 0) return;
  const int sl = threadIdx.x % S;
  const int64_t per = kThreads / S;
  double loc[1] = {0.0};
  for (int64_t row = blockIdx.x * per + threadIdx.x / S; row < nrows; row += (int64_t)gridDim.x * per) {
    double o[BS];
    row_dot<VT, S, BS>(rp, ci, v, w.p, row, sl, o);
    if (sl == 0) {
#pragma unroll
      for (int i = 0; i < BS; ++i) {
        w.q[row * BS + i] = o[i];
        loc[0] += o[i] * w.p[row * BS + i];
      }
    }
  }
  double tot[1];
  if (finish_reduce<1>(loc, w.part, &w.f->counter[0], tot) && threadIdx.x == 0) {
    w.s->pq = tot[0];
    if (!(tot[0] > 0.0)) {
      w.f->status = 2;
    } else {
      w.s->alpha = w.s->rz / tot[0];
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_pcg_update(int64_t n, double* __restrict__ x, PcgWork w) {
  if (w.f->status != 0) return;
  const double alpha = w.s->alpha;
  double loc[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    x[i] += alpha * w.p[i];
    const double ri = w.r[i] - alpha * w.q[i];
    w.r[i] = ri;
    loc[0] += ri * w.dinv[i] * ri;
    loc[1] += ri * ri;
  }
  double tot[2];
  if (finish_reduce<2>(loc, w.part, &w.f->counter[1], tot) && threadIdx.x == 0) {
    PcgScalars* s = w.s;
    s->beta = tot[0] / s->rz;
    s->rz = tot[0];
    s->rr = tot[1];
    s->relres = s->bb > 0.0 ? sqrt(tot[1] / s->bb) : 0.0;
    const int it = ++w.f->iter;
    if (tot[1] <= s->tol2)
      w.f->status = 1;
    else if (it >= w.f->maxiter)
      w.f->status = 3;
  }
}

__global__ __launch_bounds__(kThreads) void k_pcg_dir(int64_t n, PcgWork w) {
  if (w.f->status != 0) return;
  const double beta = w.s->beta;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    w.p[i] = w.dinv[i] * w.r[i] + beta * w.p[i];
}

__global__ void k_pcg_setmax(PcgWork w, long long maxiter) {
  w.f->counter[0] = 0;
  w.f->counter[1] = 0;
  w.f->maxiter = maxiter;
  w.f->status = 0;
  w.f->iter = 0;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_gather(int64_t n, const int32_t* __restrict__ idx,
                                                    const T* __restrict__ src, T* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    dst[i] = src[idx[i]];
}

int grid_for(int64_t n, int per_block) {
  int64_t g = (n + per_block - 1) / per_block;
  return (int)(g < 1 ? 1 : (g > kMaxBlocks ? kMaxBlocks : g));
}

template <typename VT, int S, int BS>
int pcg_run(int64_t nrows, const int32_t* rp, const int32_t* ci, const VT* v, const double* b, double* x, PcgWork w,
            double rtol, double atol, int64_t maxiter, int check_every, int64_t* iters, double* relres,
            hipStream_t st) {
  const int64_t n = nrows * BS;
  const int gv = grid_for(n, kThreads * 4);
  const int gs = grid_for(nrows, kThreads / S);
  hipLaunchKernelGGL(k_pcg_setmax, dim3(1), dim3(1), 0, st, w, (long long)maxiter);
  hipLaunchKernelGGL((k_diag_inv<VT, BS>), dim3(grid_for(nrows, kThreads)), dim3(kThreads), 0, st, nrows, rp, ci, v,
                     w.dinv);
  hipLaunchKernelGGL((k_spmv<VT, S, BS>), dim3(gs), dim3(kThreads), 0, st, nrows, rp, ci, v, x, w.q);
  hipLaunchKernelGGL(k_pcg_init, dim3(gv), dim3(kThreads), 0, st, n, b, w, rtol * rtol, atol * atol);
  TM_HIP(hipGetLastError());
  PcgFlags hf{};
  PcgScalars hs{};
  int chunk = check_every > 0 ? check_every : 16;
  for (;;) {
    TM_HIP(hipMemcpyAsync(&hf, w.f, sizeof(hf), hipMemcpyDeviceToHost, st));
    TM_HIP(hipStreamSynchronize(st));
    if (hf.status != 0) break;
    for (int k = 0; k < chunk; ++k) {
      hipLaunchKernelGGL((k_pcg_spmv<VT, S, BS>), dim3(gs), dim3(kThreads), 0, st, nrows, rp, ci, v, w);
      hipLaunchKernelGGL(k_pcg_update, dim3(gv), dim3(kThreads), 0, st, n, x, w);
      hipLaunchKernelGGL(k_pcg_dir, dim3(gv), dim3(kThreads), 0, st, n, w);
    }
    TM_HIP(hipGetLastError());
    if (check_every <= 0 && chunk < 256) chunk *= 2;
  }
  TM_HIP(hipMemcpyAsync(&hs, w.s, sizeof(hs), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  if (iters) *iters = hf.iter;
  if (relres) *relres = hs.relres;
  if (hf.status == 2) {
    tmx::set_error("PCG breakdown: p.Ap <= 0 at iteration %d (matrix not SPD?)", hf.iter);
    return TM_ERR_NOTCONV;
  }
  if (hf.status == 3) {
    tmx::set_error("PCG did not converge in %lld iterations (relres %.3e)", (long long)maxiter, hs.relres);
    return TM_ERR_NOTCONV;
  }
  return TM_OK;
}

// Lanes per row: the smallest power of two covering the mean row length.
int lanes_for(int64_t n, int64_t nnz) {
  const int64_t avg = n > 0 ? (nnz + n - 1) / n : 1;
  if (avg <= 4) return 4;
  if (avg <= 8) return 8;
  if (avg <= 16) return 16;
  return 32;
}

template <typename VT, int BS>
int spmv_dispatch(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const VT* v, const double* x,
                  double* y, hipStream_t st) {
  const int S = lanes_for(n, nnz);
  const int g = grid_for(n, kThreads / S);
  switch (S) {
    case 4: hipLaunchKernelGGL((k_spmv<VT, 4, BS>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
    case 8: hipLaunchKernelGGL((k_spmv<VT, 8, BS>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
    case 16: hipLaunchKernelGGL((k_spmv<VT, 16, BS>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
    default: hipLaunchKernelGGL((k_spmv<VT, 32, BS>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
  }
  TM_HIP(hipGetLastError());
  return TM_OK;
}

template <typename VT, int BS>
int pcg_dispatch(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const VT* v, const double* b,
                 double* x, PcgWork w, double rtol, double atol, int64_t maxiter, int check_every,
                 int64_t* iters, double* relres, hipStream_t st) {
  switch (lanes_for(n, nnz)) {
    case 4: return pcg_run<VT, 4, BS>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    case 8: return pcg_run<VT, 8, BS>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    case 16: return pcg_run<VT, 16, BS>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    default: return pcg_run<VT, 32, BS>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
  }
}

}  // namespace

extern "C" int tm_spmv(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
                       int value_dtype, const double* x, double* y, void* stream) {
  TM_REQUIRE(n >= 0 && row_ptr && col_idx && values && x && y, TM_ERR_ARG, "null argument");
  if (n == 0) return TM_OK;
  int32_t nnz_h = 0;
  hipStream_t st = (hipStream_t)stream;
  TM_HIP(hipMemcpyAsync(&nnz_h, row_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  if (value_dtype == TM_F32) return spmv_dispatch<float, 1>(n, nnz_h, row_ptr, col_idx, (const float*)values, x, y, st);
  return spmv_dispatch<double, 1>(n, nnz_h, row_ptr, col_idx, (const double*)values, x, y, st);
}

extern "C" int tm_spmv_bsr3(int64_t n_block_rows, const int32_t* row_ptr, const int32_t* col_idx,
                            const void* values, int value_dtype, const double* x, double* y, void* stream) {
  TM_REQUIRE(n_block_rows >= 0 && row_ptr && col_idx && values && x && y, TM_ERR_ARG, "null argument");
  if (n_block_rows == 0) return TM_OK;
  int32_t nnz_h = 0;
  hipStream_t st = (hipStream_t)stream;
  TM_HIP(hipMemcpyAsync(&nnz_h, row_ptr + n_block_rows, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  if (value_dtype == TM_F32)
    return spmv_dispatch<float, 3>(n_block_rows, nnz_h, row_ptr, col_idx, (const float*)values, x, y, st);
  return spmv_dispatch<double, 3>(n_block_rows, nnz_h, row_ptr, col_idx, (const double*)values, x, y, st);
}

extern "C" int tm_gather(int64_t n, const int32_t* idx, const void* src, void* dst, int elem_bytes,
                         void* stream) {
  TM_REQUIRE(n >= 0 && (elem_bytes == 4 || elem_bytes == 8), TM_ERR_ARG, "bad gather arguments");
  if (n == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  const int g = grid_for(n, kThreads * 4);
  if (elem_bytes == 4)
    hipLaunchKernelGGL((k_gather<float>), dim3(g), dim3(kThreads), 0, st, n, idx, (const float*)src, (float*)dst);
  else
    hipLaunchKernelGGL((k_gather<double>), dim3(g), dim3(kThreads), 0, st, n, idx, (const double*)src,
                       (double*)dst);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int64_t tm_pcg_work_size(int64_t n) {
  return (int64_t)sizeof(double) * (4 * n + 3 * kMaxBlocks) + (int64_t)sizeof(PcgScalars) +
         (int64_t)sizeof(PcgFlags) + 64;
}

extern "C" int tm_pcg_jacobi(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
                             int value_dtype, const double* b, double* x, void* work, double rtol, double atol,
                             int64_t maxiter, int check_every, int64_t* iters, double* relres, void* stream) {
  TM_REQUIRE(n > 0 && row_ptr && col_idx && values && b && x && work, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  int32_t nnz_h = 0;
  TM_HIP(hipMemcpyAsync(&nnz_h, row_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  PcgWork w = carve_work(work, n);
  if (value_dtype == TM_F32)
    return pcg_dispatch<float, 1>(n, nnz_h, row_ptr, col_idx, (const float*)values, b, x, w, rtol, atol, maxiter,
                                  check_every, iters, relres, st);
  return pcg_dispatch<double, 1>(n, nnz_h, row_ptr, col_idx, (const double*)values, b, x, w, rtol, atol, maxiter,
                                 check_every, iters, relres, st);
}

extern "C" int tm_pcg_jacobi_bsr3(int64_t n_block_rows, const int32_t* row_ptr, const int32_t* col_idx,
                                  const void* values, int value_dtype, const double* b, double* x, void* work,
                                  double rtol, double atol, int64_t maxiter, int check_every, int64_t* iters,
                                  double* relres, void* stream) {
  TM_REQUIRE(n_block_rows > 0 && row_ptr && col_idx && values && b && x && work, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  int32_t nnz_h = 0;
  TM_HIP(hipMemcpyAsync(&nnz_h, row_ptr + n_block_rows, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  PcgWork w = carve_work(work, 3 * n_block_rows);
  if (value_dtype == TM_F32)
    return pcg_dispatch<float, 3>(n_block_rows, nnz_h, row_ptr, col_idx, (const float*)values, b, x, w, rtol, atol,
                                  maxiter, check_every, iters, relres, st);
  return pcg_dispatch<double, 3>(n_block_rows, nnz_h, row_ptr, col_idx, (const double*)values, b, x, w, rtol, atol,
                                 maxiter, check_every, iters, relres, st);
}
