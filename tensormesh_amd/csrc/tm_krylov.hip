// CSR SpMV and Jacobi-preconditioned CG for gfx950.
//
// Matrix values are FP32 (or FP64), vectors and every reduction FP64.  One PCG
// iteration is three stream-ordered kernels:
//   K1  q = A p and (p, q)               -> alpha = rz / pq
//   K2  r -= alpha q, z = D^-1 r, (r, z) and (r, r)
//                                        -> beta, convergence test
//   K3  x += alpha p, p = D^-1 r + beta p
// (x is updated in K3, which reads p anyway: one vector pass fewer.)
// Each kernel writes one partial sum per workgroup; the next kernel, which
// needs the scalar, sums the partials in a fixed order (so the result does not
// depend on scheduling and no atomics or grid-wide fences are needed).
// Scalars live in device memory; kernels return immediately once the
// device-side status flag is set, so the host only reads the flag (and the
// residual) every few iterations, as many as the convergence rate predicts.
#include <cstdlib>
#include <mutex>
#include <vector>

#include "tm_device.h"
#include "tm_parts.h"
#include "tm_spmv.h"

namespace {

using tmx::kWave;
using tmx::col_of;
using tmx::row_accum;
using tmx::row_chunk;
using tmx::RowChunk;
using tmx::sub_reduce;
constexpr int kThreads = 512;
constexpr int kMaxBlocks = 1024;

// Device-side PCG state.  Scalars are produced by one designated workgroup of
// a kernel and only consumed by later kernels (kernel boundaries order them).
struct PcgScalars {
  double rz, rz_prev, bb, tol2, relres, alpha;
};
// K3 runs iff K2 of the same iteration ran to the end: K1 clears `step` (its
// workgroups never read it), K2 sets it (never read in K2), K3 only reads it.
// K3 cannot use `status` for this, since its own workgroup 0 may set it while
// other workgroups of the launch are still to start.
struct PcgFlags {
  int status;  // 0 running, 1 converged, 2 breakdown (p.Ap <= 0), 3 max iterations,
               // 4 true residual missed after the restarts (BiCGSTAB)
  int iter;
  long long maxiter;
  int step;
};

struct PcgWork {
  double* r;
  double* p;
  double* q;
  double* dinv;
  double* part;  // [4][kMaxBlocks]: (p,q) | (r,z) | (r,r) | (b,b)
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
  w.s = (PcgScalars*)(w.part + 4 * kMaxBlocks);
  w.f = (PcgFlags*)(w.s + 1);
  return w;
}

constexpr int P_PQ = 0, P_RZ = 1, P_RR = 2, P_BB = 3;

// y = A x; each S-lane group handles R rows per trip (rows row, row + stride, ...).
template <typename VT, int S, int BS, int R, typename CI = int32_t>
__global__ __launch_bounds__(kThreads) void k_spmv(int64_t nrows, const int32_t* __restrict__ rp,
                                                  const CI* __restrict__ ci, const VT* __restrict__ v,
                                                  const double* __restrict__ x, double* __restrict__ y) {
  const int sl = threadIdx.x % S;
  constexpr int64_t stride = kThreads / S;
  const RowChunk rc = row_chunk(nrows);
  const int lim = rp[rc.r1];
  for (int64_t row = rc.r0 + threadIdx.x / S; row < rc.r1; row += stride * R) {
    double o[R][BS];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int i = 0; i < BS; ++i) o[r][i] = 0.0;
      if (row + r * stride < rc.r1)
        row_accum<VT, S, BS, (S <= 4 ? 4 : 2), CI>(rp, ci, v, x, row + r * stride, sl, lim, o[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      sub_reduce<S, BS>(o[r]);
      if (sl == 0 && row + r * stride < rc.r1)
#pragma unroll
        for (int i = 0; i < BS; ++i) y[(row + r * stride) * BS + i] = o[r][i];
    }
  }
}

// Sum of the first n entries of each of NS partial arrays, computed by wave 0 in
// a fixed order (lane-strided, then a shuffle tree) and broadcast through LDS:
// every workgroup of a kernel obtains bit-identical values.
template <int NS>
__device__ __forceinline__ void sum_parts(const double* part, const int (&which)[NS], int n, double (&out)[NS]) {
  tmx::sum_parts<NS>(part, kMaxBlocks, which, n, out);
}

template <int NS>
__device__ __forceinline__ void write_parts(double* part, const int (&which)[NS], double (&loc)[NS]) {
  tmx::write_parts<kThreads, NS>(part, kMaxBlocks, which, loc);
}

template <typename VT, int BS, typename CI = int32_t>
__global__ __launch_bounds__(kThreads) void k_diag_inv(int64_t nrows, const int32_t* __restrict__ rp,
                                                      const CI* __restrict__ ci, const VT* __restrict__ v,
                                                      double* __restrict__ dinv) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * kThreads) {
    double d[BS];
#pragma unroll
    for (int c = 0; c < BS; ++c) d[c] = 1.0;  // missing diagonal -> 1 (torch_solve.py:24-45)
    bool found = false;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (col_of(ci, i, k) == i) {
#pragma unroll
        for (int c = 0; c < BS; ++c) {
          const double vk = (double)v[(int64_t)k * BS * BS + c * BS + c];
          d[c] = found ? d[c] + vk : vk;
        }
        found = true;
      }
    // a stored diagonal below 1e-12 in magnitude counts as 1, as a missing one
    // does (the Jacobi set-up of bicgstab_py, torch_solve.py:177-182)
#pragma unroll
    for (int c = 0; c < BS; ++c) dinv[i * BS + c] = 1.0 / (fabs(d[c]) < 1e-12 ? 1.0 : d[c]);
  }
}

// r = b - q (q = A x0), p = D^-1 r; partials (r,z), (r,r), (b,b)
__global__ __launch_bounds__(kThreads) void k_pcg_init(int64_t n, const double* __restrict__ b, PcgWork w) {
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
  const int which[3] = {P_RZ, P_RR, P_BB};
  write_parts<3>(w.part, which, loc);
}

__global__ __launch_bounds__(kThreads) void k_pcg_ctl0(PcgWork w, int nparts, double rtol2, double atol2,
                                                      long long maxiter) {
  const int which[3] = {P_RZ, P_RR, P_BB};
  double t[3];
  sum_parts<3>(w.part, which, nparts, t);
  if (threadIdx.x == 0) {
    w.s->rz = t[0];
    w.s->rz_prev = t[0];
    w.s->bb = t[2];
    w.s->tol2 = fmax(rtol2 * t[2], atol2);
    w.s->relres = t[2] > 0.0 ? sqrt(t[1] / t[2]) : 0.0;
    w.f->iter = 0;
    w.f->maxiter = maxiter;
    w.f->step = 0;
    w.f->status = (t[1] <= w.s->tol2) ? 1 : 0;
  }
}

// K1: q = A p, partial (p, q)
template <typename VT, int S, int BS, typename CI = int32_t>
__global__ __launch_bounds__(kThreads) void k_pcg_spmv(int64_t nrows, const int32_t* __restrict__ rp,
                                                      const CI* __restrict__ ci, const VT* __restrict__ v,
                                                      PcgWork w) {
  if (blockIdx.x == 0 && threadIdx.x == 0) w.f->step = 0;
  if (w.f->status != 0) return;
  constexpr int R = 1;
  const int sl = threadIdx.x % S;
  constexpr int64_t stride = kThreads / S;
  const RowChunk rc = row_chunk(nrows);
  const int lim = rp[rc.r1];
  double loc[1] = {0.0};
  for (int64_t row = rc.r0 + threadIdx.x / S; row < rc.r1; row += stride * R) {
    double o[R][BS];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int i = 0; i < BS; ++i) o[r][i] = 0.0;
      if (row + r * stride < rc.r1)
        row_accum<VT, S, BS, (S <= 4 ? 4 : 2), CI>(rp, ci, v, w.p, row + r * stride, sl, lim, o[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      sub_reduce<S, BS>(o[r]);
      const int64_t rr = row + r * stride;
      if (sl == 0 && rr < rc.r1)
#pragma unroll
        for (int i = 0; i < BS; ++i) {
          w.q[rr * BS + i] = o[r][i];
          loc[0] += o[r][i] * w.p[rr * BS + i];
        }
    }
  }
  const int which[1] = {P_PQ};
  write_parts<1>(w.part, which, loc);
}

// K2: alpha = rz / pq, r -= alpha q, partials (r, z), (r, r)
__global__ __launch_bounds__(kThreads) void k_pcg_update(int64_t n, PcgWork w, int nparts) {
  if (w.f->status != 0) return;
  const int which_in[1] = {P_PQ};
  double pq[1];
  sum_parts<1>(w.part, which_in, nparts, pq);
  const double rz = w.s->rz;
  const bool ok = pq[0] > 0.0;
  const double alpha = ok ? rz / pq[0] : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.s->rz_prev = rz;
    w.s->alpha = alpha;
    if (!ok)
      w.f->status = 2;
    else
      w.f->step = 1;
  }
  double loc[2] = {0.0, 0.0};
  if (ok) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
      const double ri = w.r[i] - alpha * w.q[i];
      w.r[i] = ri;
      loc[0] += ri * w.dinv[i] * ri;
      loc[1] += ri * ri;
    }
  }
  const int which[2] = {P_RZ, P_RR};
  write_parts<2>(w.part, which, loc);
}

// K3: x += alpha p; beta = rz_new / rz_old, convergence test, p = D^-1 r + beta p
__global__ __launch_bounds__(kThreads) void k_pcg_dir(int64_t n, double* __restrict__ x, PcgWork w, int nparts) {
  if (w.f->step == 0) return;
  const int which[2] = {P_RZ, P_RR};
  double t[2];
  sum_parts<2>(w.part, which, nparts, t);
  const double rz_new = t[0], rr = t[1];
  const double tol2 = w.s->tol2;
  const bool conv = rr <= tol2;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int it = w.f->iter + 1;
    w.f->iter = it;
    w.s->rz = rz_new;
    w.s->relres = w.s->bb > 0.0 ? sqrt(rr / w.s->bb) : 0.0;
    if (conv)
      w.f->status = 1;
    else if (it >= w.f->maxiter)
      w.f->status = 3;
  }
  const double alpha = w.s->alpha;
  if (conv) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
      x[i] += alpha * w.p[i];
    return;
  }
  const double beta = rz_new / w.s->rz_prev;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double pi = w.p[i];
    x[i] += alpha * pi;
    w.p[i] = w.dinv[i] * w.r[i] + beta * pi;
  }
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

// Iterations to enqueue before the host next reads the device status.  The
// kernels of an iteration that starts after convergence return at once, but
// each still costs a dispatch, so the chunk follows the convergence rate seen
// over the last chunk: it ends at the predicted convergence iteration (at
// least 1, at most 256; 16 for the first chunk, doubling while the residual
// does not fall).  check_every > 0 fixes the chunk.
struct Poller {
  int check_every;
  int chunk;
  int64_t maxiter;
  int it_prev = -1;
  double rel_prev = 0.0;
  Poller(int ce, int64_t mi) : check_every(ce), chunk(cap(ce > 0 ? ce : 16, 0, mi)), maxiter(mi) {}
  // launches past maxiter only set the status: never queue more than are left
  static int cap(int c, int it, int64_t mi) {
    if (mi <= 0) return c;
    const int64_t left = mi - it;
    return left < 1 ? 1 : (left < c ? (int)left : c);
  }
  int next(int it, double rel, double target) {
    int c = check_every > 0 ? check_every : (chunk < 256 ? 2 * chunk : 256);
    if (check_every <= 0 && it_prev >= 0 && it > it_prev && rel > 0.0 && rel_prev > 0.0 && rel < rel_prev &&
        target > 0.0) {
      // Krylov convergence speeds up, so the rate since the last read
      // overestimates the iterations left: queue half of them (at least 16),
      // all of them once 16 or fewer remain.  On recorded Jacobi-PCG residual
      // histories this leaves a few queued iterations past convergence per
      // solve instead of tens, for a few more status reads.
      const double rate = log(rel / rel_prev) / (double)(it - it_prev);  // < 0
      const double rem = log(target / rel) / rate;                       // iterations left
      if (rem <= 16.0)
        c = rem < 1.0 ? 1 : (int)ceil(rem);
      else
        c = rem > 512.0 ? 256 : (int)fmax(16.0, ceil(0.5 * rem));
    }
    it_prev = it;
    rel_prev = rel;
    if (check_every <= 0) chunk = c;
    return cap(c, it, maxiter);
  }
};

template <typename VT, int S, int BS, typename CI = int32_t>
int pcg_run(int64_t nrows, const int32_t* rp, const CI* ci, const VT* v, const double* b, double* x, PcgWork w,
            double rtol, double atol, int64_t maxiter, int check_every, int64_t* iters, double* relres,
            hipStream_t st) {
  const int64_t n = nrows * BS;
  // one grid size for every PCG kernel: the partial arrays have G entries
  const int G = grid_for(n, kThreads * 2);
  hipLaunchKernelGGL((k_diag_inv<VT, BS, CI>), dim3(grid_for(nrows, kThreads)), dim3(kThreads), 0, st, nrows, rp, ci,
                     v, w.dinv);
  hipLaunchKernelGGL((k_spmv<VT, S, BS, 2, CI>), dim3(G), dim3(kThreads), 0, st, nrows, rp, ci, v, x, w.q);
  hipLaunchKernelGGL(k_pcg_init, dim3(G), dim3(kThreads), 0, st, n, b, w);
  hipLaunchKernelGGL(k_pcg_ctl0, dim3(1), dim3(kThreads), 0, st, w, G, rtol * rtol, atol * atol, (long long)maxiter);
  TM_HIP(hipGetLastError());
  // scalars and flags are adjacent: one read gives the status and the residual
  struct {
    PcgScalars s;
    PcgFlags f;
  } h{};
  static_assert(sizeof(h) == sizeof(PcgScalars) + sizeof(PcgFlags), "PCG state layout");
  PcgFlags& hf = h.f;
  PcgScalars& hs = h.s;
  Poller poll(check_every, maxiter);
  int chunk = poll.chunk;
  for (;;) {
    TM_HIP(hipMemcpyAsync(&h, w.s, sizeof(h), hipMemcpyDeviceToHost, st));
    TM_HIP(hipStreamSynchronize(st));
    if (hf.status != 0) break;
    if (hf.iter > 0) chunk = poll.next(hf.iter, hs.relres, hs.bb > 0.0 ? sqrt(hs.tol2 / hs.bb) : 0.0);
    for (int k = 0; k < chunk; ++k) {
      hipLaunchKernelGGL((k_pcg_spmv<VT, S, BS, CI>), dim3(G), dim3(kThreads), 0, st, nrows, rp, ci, v, w);
      hipLaunchKernelGGL(k_pcg_update, dim3(G), dim3(kThreads), 0, st, n, w, G);
      hipLaunchKernelGGL(k_pcg_dir, dim3(G), dim3(kThreads), 0, st, n, x, w, G);
    }
    TM_HIP(hipGetLastError());
  }
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

// Lanes per row: TM_SPMV_LANES when set (2, 4, 8, 16, 32), otherwise 4 lanes
// for rows of up to 16 entries (each lane issues its up to 4 loads at once;
// measured 5 % faster per PCG iteration than 8 lanes on the 15-entry P1 rows),
// then the smallest power of two covering half the mean row length.
int lanes_for(int64_t n, int64_t nnz) {
  static const int forced = [] {
    const char* e = getenv("TM_SPMV_LANES");
    return e ? atoi(e) : 0;
  }();
  if (forced == 2 || forced == 4 || forced == 8 || forced == 16 || forced == 32) return forced;
  const int64_t avg = n > 0 ? (nnz + n - 1) / n : 1;
  if (avg <= 16) return 4;
  if (avg <= 32) return 8;
  if (avg <= 64) return 16;
  return 32;
}

template <typename VT, int BS>
int spmv_dispatch(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const VT* v, const double* x,
                  double* y, hipStream_t st) {
  const int S = lanes_for(n, nnz);
  const int g = grid_for(n, kThreads / S * 2);
  switch (S) {
    case 2: hipLaunchKernelGGL((k_spmv<VT, 2, BS, 2>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
    case 4: hipLaunchKernelGGL((k_spmv<VT, 4, BS, 2>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
    case 8: hipLaunchKernelGGL((k_spmv<VT, 8, BS, 2>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
    case 16: hipLaunchKernelGGL((k_spmv<VT, 16, BS, 2>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
    default: hipLaunchKernelGGL((k_spmv<VT, 32, BS, 2>), dim3(g), dim3(kThreads), 0, st, n, rp, ci, v, x, y); break;
  }
  TM_HIP(hipGetLastError());
  return TM_OK;
}

template <typename VT, int BS, typename CI = int32_t>
int pcg_dispatch(int64_t n, int64_t nnz, const int32_t* rp, const CI* ci, const VT* v, const double* b,
                 double* x, PcgWork w, double rtol, double atol, int64_t maxiter, int check_every,
                 int64_t* iters, double* relres, hipStream_t st) {
  switch (lanes_for(n, nnz)) {
    case 2: return pcg_run<VT, 2, BS, CI>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    case 4: return pcg_run<VT, 4, BS, CI>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    case 8: return pcg_run<VT, 8, BS, CI>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    case 16:
      return pcg_run<VT, 16, BS, CI>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    default:
      return pcg_run<VT, 32, BS, CI>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
  }
}

// Column deltas col - row as int16; fits[0] is cleared when an entry lies
// more than 32767 from the diagonal (the caller sets it to 1 first).
__global__ __launch_bounds__(kThreads) void k_delta16(int64_t nrows, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ ci, int16_t* __restrict__ cd,
                                                     int* __restrict__ fits) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * kThreads) {
    bool ok = true;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int64_t d = (int64_t)ci[k] - i;
      ok = ok && d >= -32767 && d <= 32767;
      cd[k] = (int16_t)(ok ? d : 0);
    }
    if (!ok) atomicAnd(fits, 0);
  }
}


// ---------------------------------------------------------------------------
// Jacobi-preconditioned BiCGSTAB (right preconditioning, M = diag(A)) for
// non-symmetric systems: the algorithm of sparse/solve/torch_solve.py:130-248
// with device-side scalars.  One iteration is five kernels:
//   B1  v = A phat, (r0h, v)
//   B2  alpha = rho / (r0h, v); s = r - alpha v, shat = M^-1 s
//   B3  t = A shat, (t, t), (t, s)
//   B4  omega = (t, s) / (t, t); x += alpha phat + omega shat; r = s - omega t;
//       (r, r), (r0h, r)
//   B5  convergence test; beta = (rho_new / rho)(alpha / omega);
//       p = r + beta (p - omega v), phat = M^-1 p
// Every scalar is summed from the producer's partials by each consumer in a
// fixed order; the values a later kernel needs (rho, alpha) are stored by
// workgroup 0 of the kernel that formed them, and never overwritten in the
// kernel that reads them.
struct BcScalars {
  double rho, alpha, bb, tol2, relres;
};
struct BcWork {
  double *r, *r0h, *p, *v, *s, *t, *phat, *shat, *minv, *part;
  BcScalars* sc;
  PcgFlags* f;
};
constexpr int B_R0V = 0, B_TT = 1, B_TS = 2, B_RR = 3, B_RHO = 4, B_BB = 5;

BcWork carve_bc(void* work, int64_t n) {
  BcWork w;
  double* d = (double*)work;
  w.r = d;
  w.r0h = d + n;
  w.p = d + 2 * n;
  w.v = d + 3 * n;
  w.s = d + 4 * n;
  w.t = d + 5 * n;
  w.phat = d + 6 * n;
  w.shat = d + 7 * n;
  w.minv = d + 8 * n;
  w.part = d + 9 * n;
  w.sc = (BcScalars*)(w.part + 6 * kMaxBlocks);
  w.f = (PcgFlags*)(w.sc + 1);
  return w;
}

// y = A x with one partial (a, y) per workgroup (a = r0h for B1, y = t -> (t,t), (t,s) for B3)
template <typename VT, int S, bool TWO>
__global__ __launch_bounds__(kThreads) void k_bc_spmv(int64_t nrows, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ ci, const VT* __restrict__ v,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     const double* __restrict__ a, BcWork w, int slot0) {
  if (w.f->status != 0) return;
  const int sl = threadIdx.x % S;
  constexpr int64_t stride = kThreads / S;
  const RowChunk rc = row_chunk(nrows);
  const int lim = rp[rc.r1];
  double loc[2] = {0.0, 0.0};
  for (int64_t row = rc.r0 + threadIdx.x / S; row < rc.r1; row += stride) {
    double o[1] = {0.0};
    row_accum<VT, S, 1>(rp, ci, v, x, row, sl, lim, o);
    sub_reduce<S, 1>(o);
    if (sl == 0) {
      y[row] = o[0];
      if (TWO) {
        loc[0] += o[0] * o[0];
        loc[1] += o[0] * a[row];
      } else {
        loc[0] += o[0] * a[row];
      }
    }
  }
  if (TWO) {
    const int which[2] = {slot0, slot0 + 1};
    write_parts<2>(w.part, which, loc);
  } else {
    double l1[1] = {loc[0]};
    const int which[1] = {slot0};
    write_parts<1>(w.part, which, l1);
  }
}

// r = b - v (v = A x0), r0h = r, p = r, phat = M^-1 r; partials (r,r), (r0h,r), (b,b)
__global__ __launch_bounds__(kThreads) void k_bc_init(int64_t n, const double* __restrict__ b, BcWork w) {
  double loc[3] = {0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double bi = b[i];
    const double ri = bi - w.v[i];
    w.r[i] = ri;
    w.r0h[i] = ri;
    w.p[i] = ri;
    w.phat[i] = w.minv[i] * ri;
    loc[0] += ri * ri;
    loc[1] += ri * ri;
    loc[2] += bi * bi;
  }
  const int which[3] = {B_RR, B_RHO, B_BB};
  write_parts<3>(w.part, which, loc);
}

__global__ __launch_bounds__(kThreads) void k_bc_ctl0(BcWork w, int nparts, double rtol2, double atol2,
                                                     long long maxiter) {
  const int which[3] = {B_RR, B_RHO, B_BB};
  double t[3];
  sum_parts<3>(w.part, which, nparts, t);
  if (threadIdx.x == 0) {
    w.sc->bb = t[2];
    w.sc->tol2 = fmax(rtol2 * t[2], atol2);
    w.sc->relres = t[2] > 0.0 ? sqrt(t[0] / t[2]) : 0.0;
    w.f->iter = 0;
    w.f->maxiter = maxiter;
    w.f->step = 0;
    w.f->status = (t[0] <= w.sc->tol2) ? 1 : 0;
  }
}

// After the recursion reports convergence: with v = A x and r = b - v
// recomputed by k_bc_init, test the true residual (keeps the iteration count);
// a miss restarts the recursion from this r (residual replacement) while
// `restart` is set, and otherwise ends the solve with status 4.
__global__ __launch_bounds__(kThreads) void k_bc_recheck(BcWork w, int nparts, int restart) {
  const int which[1] = {B_RR};
  double t[1];
  sum_parts<1>(w.part, which, nparts, t);
  if (threadIdx.x == 0) {
    w.sc->relres = w.sc->bb > 0.0 ? sqrt(t[0] / w.sc->bb) : 0.0;
    w.f->step = 0;
    w.f->status = (t[0] <= w.sc->tol2) ? 1 : (!restart ? 4 : (w.f->iter >= w.f->maxiter ? 3 : 0));
  }
}

// B2: alpha = rho / (r0h, v) with rho = (r0h, r) of the current r; s, shat
__global__ __launch_bounds__(kThreads) void k_bc_s(int64_t n, BcWork w, int nparts) {
  if (w.f->status != 0) return;
  const int which[2] = {B_R0V, B_RHO};
  double t[2];
  sum_parts<2>(w.part, which, nparts, t);
  const double denom = t[0], rho = t[1];
  const bool ok = denom != 0.0 && __builtin_isfinite(denom) && __builtin_isfinite(rho);
  const double alpha = ok ? rho / denom : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.sc->rho = rho;
    w.sc->alpha = alpha;
    if (!ok) w.f->status = 2;
  }
  if (!ok) return;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double si = w.r[i] - alpha * w.v[i];
    w.s[i] = si;
    w.shat[i] = w.minv[i] * si;
  }
}

// B4: omega, x and r updates; partials (r,r), (r0h,r).  `step` tells B5 that
// this kernel ran (B5 cannot rely on the status word, see PcgFlags).
__global__ __launch_bounds__(kThreads) void k_bc_x(int64_t n, double* __restrict__ x, BcWork w, int nparts) {
  if (blockIdx.x == 0 && threadIdx.x == 0) w.f->step = 0;
  if (w.f->status != 0) return;
  const int which_in[2] = {B_TT, B_TS};
  double t[2];
  sum_parts<2>(w.part, which_in, nparts, t);
  const double omega = t[0] != 0.0 ? t[1] / t[0] : 0.0;
  const double alpha = w.sc->alpha;
  double loc[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double si = w.s[i];
    x[i] += alpha * w.phat[i] + omega * w.shat[i];
    const double ri = si - omega * w.t[i];
    w.r[i] = ri;
    loc[0] += ri * ri;
    loc[1] += w.r0h[i] * ri;
  }
  const int which[2] = {B_RR, B_RHO};
  write_parts<2>(w.part, which, loc);
  if (blockIdx.x == 0 && threadIdx.x == 0) w.f->step = 1;
}

// B5: convergence test and the new direction
__global__ __launch_bounds__(kThreads) void k_bc_dir(int64_t n, BcWork w, int nparts) {
  if (w.f->step == 0) return;
  const int which[4] = {B_RR, B_RHO, B_TT, B_TS};
  double t[4];
  sum_parts<4>(w.part, which, nparts, t);
  const double rr = t[0], rho_new = t[1];
  const double omega = t[2] != 0.0 ? t[3] / t[2] : 0.0;
  const bool conv = rr <= w.sc->tol2;
  const bool stall = !conv && (omega == 0.0 || !__builtin_isfinite(rho_new));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int it = w.f->iter + 1;
    w.f->iter = it;
    w.sc->relres = w.sc->bb > 0.0 ? sqrt(rr / w.sc->bb) : 0.0;
    if (conv)
      w.f->status = 1;
    else if (stall)
      w.f->status = 2;
    else if (it >= w.f->maxiter)
      w.f->status = 3;
  }
  if (conv || stall) return;
  const double beta = (rho_new / w.sc->rho) * (w.sc->alpha / omega);
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double pi = w.r[i] + beta * (w.p[i] - omega * w.v[i]);
    w.p[i] = pi;
    w.phat[i] = w.minv[i] * pi;
  }
}

template <typename VT, int S>
int bc_run(int64_t n, const int32_t* rp, const int32_t* ci, const VT* v, const double* b, double* x, BcWork w,
           double rtol, double atol, int64_t maxiter, int check_every, int64_t* iters, double* relres,
           hipStream_t st) {
  const int G = grid_for(n, kThreads * 2);
  hipLaunchKernelGGL((k_diag_inv<VT, 1>), dim3(grid_for(n, kThreads)), dim3(kThreads), 0, st, n, rp, ci, v,
                     w.minv);
  hipLaunchKernelGGL((k_spmv<VT, S, 1, 2>), dim3(G), dim3(kThreads), 0, st, n, rp, ci, v, x, w.v);
  hipLaunchKernelGGL(k_bc_init, dim3(G), dim3(kThreads), 0, st, n, b, w);
  hipLaunchKernelGGL(k_bc_ctl0, dim3(1), dim3(kThreads), 0, st, w, G, rtol * rtol, atol * atol, (long long)maxiter);
  TM_HIP(hipGetLastError());
  struct {
    BcScalars s;
    PcgFlags f;
  } h{};
  static_assert(sizeof(h) == sizeof(BcScalars) + sizeof(PcgFlags), "BiCGSTAB state layout");
  PcgFlags& hf = h.f;
  BcScalars& hs = h.s;
  Poller poll(check_every, maxiter);
  int chunk = poll.chunk;
  int restarts = 0;
  for (;;) {
    TM_HIP(hipMemcpyAsync(&h, w.sc, sizeof(h), hipMemcpyDeviceToHost, st));
    TM_HIP(hipStreamSynchronize(st));
    if (hf.status == 1 && hf.iter > 0) {
      // the recursive residual met the tolerance: check the true one (x is
      // final here); after three restarts a miss ends the solve
      const int restart = restarts < 3 ? 1 : 0;
      ++restarts;
      hipLaunchKernelGGL((k_spmv<VT, S, 1, 2>), dim3(G), dim3(kThreads), 0, st, n, rp, ci, v, x, w.v);
      hipLaunchKernelGGL(k_bc_init, dim3(G), dim3(kThreads), 0, st, n, b, w);
      hipLaunchKernelGGL(k_bc_recheck, dim3(1), dim3(kThreads), 0, st, w, G, restart);
      TM_HIP(hipGetLastError());
      TM_HIP(hipMemcpyAsync(&h, w.sc, sizeof(h), hipMemcpyDeviceToHost, st));
      TM_HIP(hipStreamSynchronize(st));
      poll = Poller(check_every, maxiter);  // the recursion restarts from the true residual
      chunk = Poller::cap(poll.chunk, hf.iter, maxiter);
    } else if (hf.iter > 0) {
      chunk = poll.next(hf.iter, hs.relres, hs.bb > 0.0 ? sqrt(hs.tol2 / hs.bb) : 0.0);
    }
    if (hf.status != 0) break;
    for (int k = 0; k < chunk; ++k) {
      hipLaunchKernelGGL((k_bc_spmv<VT, S, false>), dim3(G), dim3(kThreads), 0, st, n, rp, ci, v, w.phat, w.v,
                         w.r0h, w, B_R0V);
      hipLaunchKernelGGL(k_bc_s, dim3(G), dim3(kThreads), 0, st, n, w, G);
      hipLaunchKernelGGL((k_bc_spmv<VT, S, true>), dim3(G), dim3(kThreads), 0, st, n, rp, ci, v, w.shat, w.t,
                         w.s, w, B_TT);
      hipLaunchKernelGGL(k_bc_x, dim3(G), dim3(kThreads), 0, st, n, x, w, G);
      hipLaunchKernelGGL(k_bc_dir, dim3(G), dim3(kThreads), 0, st, n, w, G);
    }
    TM_HIP(hipGetLastError());
  }
  if (iters) *iters = hf.iter;
  if (relres) *relres = hs.relres;
  if (hf.status == 2) {
    tmx::set_error("BiCGSTAB breakdown at iteration %d (relres %.3e)", hf.iter, hs.relres);
    return TM_ERR_NOTCONV;
  }
  if (hf.status == 3) {
    tmx::set_error("BiCGSTAB did not converge in %lld iterations (relres %.3e)", (long long)maxiter, hs.relres);
    return TM_ERR_NOTCONV;
  }
  if (hf.status == 4) {
    tmx::set_error("BiCGSTAB: the true residual misses the tolerance after %d restarts (relres %.3e)",
                   restarts - 1, hs.relres);
    return TM_ERR_NOTCONV;
  }
  return TM_OK;
}

template <typename VT>
int bc_dispatch(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const VT* v, const double* b, double* x,
                BcWork w, double rtol, double atol, int64_t maxiter, int check_every, int64_t* iters, double* relres,
                hipStream_t st) {
  switch (lanes_for(n, nnz)) {
    case 2: return bc_run<VT, 2>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    case 4: return bc_run<VT, 4>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    case 8: return bc_run<VT, 8>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    case 16: return bc_run<VT, 16>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
    default: return bc_run<VT, 32>(n, rp, ci, v, b, x, w, rtol, atol, maxiter, check_every, iters, relres, st);
  }
}

}  // namespace

extern "C" int tm_spmv(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                       const void* values, int value_dtype, const double* x, double* y, void* stream) {
  TM_REQUIRE(n >= 0 && nnz >= 0 && row_ptr && col_idx && values && x && y, TM_ERR_ARG, "null argument");
  if (n == 0) return TM_OK;
  const int64_t nnz_h = nnz;
  hipStream_t st = (hipStream_t)stream;
  if (value_dtype == TM_F32) return spmv_dispatch<float, 1>(n, nnz_h, row_ptr, col_idx, (const float*)values, x, y, st);
  return spmv_dispatch<double, 1>(n, nnz_h, row_ptr, col_idx, (const double*)values, x, y, st);
}

extern "C" int tm_spmv_bsr3(int64_t n_block_rows, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                            const void* values, int value_dtype, const double* x, double* y, void* stream) {
  TM_REQUIRE(n_block_rows >= 0 && nnz >= 0 && row_ptr && col_idx && values && x && y, TM_ERR_ARG,
             "null argument");
  if (n_block_rows == 0) return TM_OK;
  const int64_t nnz_h = nnz;
  hipStream_t st = (hipStream_t)stream;
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
  return (int64_t)sizeof(double) * (4 * n + 4 * kMaxBlocks) + (int64_t)sizeof(PcgScalars) +
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

extern "C" int tm_csr_delta16(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, int16_t* col_delta,
                              int* fits, void* stream) {
  TM_REQUIRE(n > 0 && row_ptr && col_idx && col_delta && fits, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int one = 1;
  TM_HIP(hipMemcpyAsync(fits, &one, sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_delta16, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, st, n, row_ptr, col_idx, col_delta,
                     fits);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_pcg_jacobi_d16(int64_t n, const int32_t* row_ptr, const int16_t* col_delta, const void* values,
                                 int value_dtype, const double* b, double* x, void* work, double rtol, double atol,
                                 int64_t maxiter, int check_every, int64_t* iters, double* relres, void* stream) {
  TM_REQUIRE(n > 0 && row_ptr && col_delta && values && b && x && work, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  int32_t nnz_h = 0;
  TM_HIP(hipMemcpyAsync(&nnz_h, row_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  PcgWork w = carve_work(work, n);
  if (value_dtype == TM_F32)
    return pcg_dispatch<float, 1, int16_t>(n, nnz_h, row_ptr, col_delta, (const float*)values, b, x, w, rtol, atol,
                                           maxiter, check_every, iters, relres, st);
  return pcg_dispatch<double, 1, int16_t>(n, nnz_h, row_ptr, col_delta, (const double*)values, b, x, w, rtol, atol,
                                          maxiter, check_every, iters, relres, st);
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

extern "C" int64_t tm_bicgstab_work_size(int64_t n) {
  return (int64_t)sizeof(double) * (9 * n + 6 * kMaxBlocks) + (int64_t)sizeof(BcScalars) +
         (int64_t)sizeof(PcgFlags) + 64;
}

extern "C" int tm_bicgstab_jacobi(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                                  const void* values, int value_dtype, const double* b, double* x, void* work,
                                  double rtol, double atol, int64_t maxiter, int check_every, int64_t* iters,
                                  double* relres, void* stream) {
  TM_REQUIRE(n > 0 && nnz >= 0 && row_ptr && col_idx && values && b && x && work, TM_ERR_ARG, "null argument");
  TM_REQUIRE(maxiter > 0, TM_ERR_ARG, "maxiter must be positive");
  hipStream_t st = (hipStream_t)stream;
  BcWork w = carve_bc(work, n);
  if (value_dtype == TM_F32)
    return bc_dispatch<float>(n, nnz, row_ptr, col_idx, (const float*)values, b, x, w, rtol, atol, maxiter,
                              check_every, iters, relres, st);
  return bc_dispatch<double>(n, nnz, row_ptr, col_idx, (const double*)values, b, x, w, rtol, atol, maxiter,
                             check_every, iters, relres, st);
}

// ---------------------------------------------------------------------------
// Allocation-free Jacobi-PCG: the workspace of tm_pcg_jacobi (r, p, q, D^-1,
// the partial sums, scalars and flags) lives in a static device array of
// fixed size instead of caller memory.  Each GPU holds its own copy of the
// array at its own address; host code only ever passes the address that
// hipGetSymbolAddress returned for the current device (the symbol's host-side
// name is not a device address).  Every kernel writes each workspace word
// before reading it in the same solve (D^-1, q = A x0, r and p in the set-up
// kernels, partials by their producers, scalars and flags by k_pcg_ctl0), so
// what an earlier call left there is never read.
namespace {

constexpr int64_t kNoAllocMaxRows = TM_PCG_NOALLOC_MAX_ROWS;
constexpr size_t kNoAllocBytes = sizeof(double) * (4 * kNoAllocMaxRows + 4 * kMaxBlocks) + sizeof(PcgScalars) +
                                 sizeof(PcgFlags) + 64;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(hipStream_t st, void** out) {
  int dev = -1;
  TM_HIP(hipGetDevice(&dev));
  TM_REQUIRE(dev >= 0 && dev < kMaxDevices, TM_ERR_LIMIT, "device ordinal %d beyond %d", dev, kMaxDevices);
  if (st) {
    int sdev = -1;
    TM_HIP(hipStreamGetDevice(st, &sdev));
    TM_REQUIRE(sdev == dev, TM_ERR_ARG, "the stream belongs to device %d, the current device is %d", sdev, dev);
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TM_REQUIRE(p != nullptr, TM_ERR_HIP, "no device address for the PCG scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TM_OK;
}

}  // namespace

extern "C" int64_t tm_pcg_noalloc_max_rows(void) { return kNoAllocMaxRows; }

extern "C" int tm_pcg_jacobi_noalloc(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
                                     int value_dtype, const double* b, double* x, double rtol, double atol,
                                     int64_t maxiter, int check_every, int64_t* iters, double* relres,
                                     void* stream) {
  TM_REQUIRE(n <= kNoAllocMaxRows, TM_ERR_LIMIT,
             "tm_pcg_jacobi_noalloc: %lld rows exceed the static workspace of %lld rows", (long long)n,
             (long long)kNoAllocMaxRows);
  TM_REQUIRE(n > 0 && row_ptr && col_idx && values && b && x, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  void* work = nullptr;
  const int rc = scratch_address(st, &work);
  if (rc != TM_OK) return rc;
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
