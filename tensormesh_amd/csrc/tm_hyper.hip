// Compressible neo-Hookean P1 tetrahedra on gfx950: energy, residual, tangent
// (3x3 blocks on the scalar CSR pattern) and the matrix-free tangent action
// used by the adjoint.  Same owner-computes block plan as the scalar kernels,
// with 72-byte (3x3 FP64) accumulators per CSR entry.
//
// With g_a = grad phi_a, F = I + sum_b u_b g_b^T, J = det F, w_a = F^-T g_a:
//   P g_a           = mu F g_a + (lam ln J - mu) w_a
//   K_ab[i][k] / V  = mu (g_a . g_b) d_ik + lam w_a,i w_b,k - (lam ln J - mu) w_a,k w_b,i
//   psi             = mu/2 (|F|^2 - 3) - mu ln J + lam/2 (ln J)^2
//
// Only the off-diagonal blocks are accumulated per element: every element's
// gradients sum to zero (and so do its w_a), so sum_b K_ab = 0 and each owned
// row's diagonal block is minus the sum of its off-diagonal blocks, formed
// once per row after the element loop (27 instead of 36 LDS atomics per
// owned vertex).
//
// Isotropic linear elasticity is the same tangent at u = 0 (F = I, J = 1,
// ln J = 0, w_a = g_a):
//   K_ab[i][k] / V  = mu (g_a . g_b) d_ik + lam g_a,i g_b,k + mu g_a,k g_b,i
// i.e. B_a^T C B_b with the Voigt strain-displacement matrices and Hooke's
// C; the LIN variants take that path without reading a displacement.
#include "tm_device.h"

namespace {

using tmx::kWave;
constexpr int kThreads = 256;

struct PlanArgs {
  const int32_t* blk;
  const int32_t* halo;
  const uint32_t* meta;
  const uint4* recs;
  const int32_t* rec_elem;
  const int4* runs;
  int nblk;
};

__host__ __device__ constexpr size_t a16(size_t x) { return (x + 15) & ~size_t(15); }

struct LdsV {
  double* acc;   // [nacc * 9]
  double* xs;    // [3][nh]
  double* us;    // [3][nh]
  double* accr;  // [nown * 3]
  uint32_t* ms;  // [nown]
};

__device__ __forceinline__ LdsV carve_v(char* s, int nacc, int nh, int nown) {
  LdsV L;
  size_t o = 0;
  L.acc = (double*)(s + o);
  o = a16(o + 72 * (size_t)nacc);
  L.xs = (double*)(s + o);
  o = a16(o + 24 * (size_t)nh);
  L.us = (double*)(s + o);
  o = a16(o + 24 * (size_t)nh);
  L.accr = (double*)(s + o);
  o = a16(o + 24 * (size_t)nown);
  L.ms = (uint32_t*)(s + o);
  return L;
}

__device__ __forceinline__ int rec_local(uint64_t lo, int a) { return (int)((lo >> (10 * a)) & 1023u); }
__device__ __forceinline__ int rec_rank4(uint64_t lo, uint64_t hi, int a, int j) {
  const int bit = 40 + 6 * (a * 3 + j);
  return bit < 64 ? (int)((lo >> bit) & 63u) : (int)((hi >> (bit - 64)) & 63u);
}

// Per-element kinematics of a P1 tetrahedron.
struct Kin {
  double g[4][3];   // grad phi_a
  double F[3][3];
  double w[4][3];   // F^-T g_a
  double vol, J, lnJ;
};

template <bool LIN = false>
__device__ __forceinline__ void kinematics(const double (&x)[4][3], const double (&u)[4][3], Kin& k) {
  double e1[3], e2[3], e3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    e1[d] = x[1][d] - x[0][d];
    e2[d] = x[2][d] - x[0][d];
    e3[d] = x[3][d] - x[0][d];
  }
  double c[4][3];
  c[1][0] = e2[1] * e3[2] - e2[2] * e3[1];
  c[1][1] = e2[2] * e3[0] - e2[0] * e3[2];
  c[1][2] = e2[0] * e3[1] - e2[1] * e3[0];
  c[2][0] = e3[1] * e1[2] - e3[2] * e1[1];
  c[2][1] = e3[2] * e1[0] - e3[0] * e1[2];
  c[2][2] = e3[0] * e1[1] - e3[1] * e1[0];
  c[3][0] = e1[1] * e2[2] - e1[2] * e2[1];
  c[3][1] = e1[2] * e2[0] - e1[0] * e2[2];
  c[3][2] = e1[0] * e2[1] - e1[1] * e2[0];
  const double det = e1[0] * c[1][0] + e1[1] * c[1][1] + e1[2] * c[1][2];
  const double idet = 1.0 / det;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[0][d] = -(c[1][d] + c[2][d] + c[3][d]);
#pragma unroll
    for (int a = 0; a < 4; ++a) k.g[a][d] = c[a][d] * idet;
  }
  k.vol = fabs(det) / 6.0;
  if (LIN) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) k.F[i][j] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) k.w[a][i] = k.g[a][i];
    k.J = 1.0;
    k.lnJ = 0.0;
    return;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double h = (i == j) ? 1.0 : 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) h += u[a][i] * k.g[a][j];
      k.F[i][j] = h;
    }
  const double (&F)[3][3] = k.F;
  // cofactor matrix C = J F^-T
  double C[3][3];
  C[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
  C[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
  C[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
  C[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
  C[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
  C[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
  C[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
  C[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
  C[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
  k.J = F[0][0] * C[0][0] + F[0][1] * C[0][1] + F[0][2] * C[0][2];
  k.lnJ = log(k.J);
  const double iJ = 1.0 / k.J;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) k.w[a][i] = (C[i][0] * k.g[a][0] + C[i][1] * k.g[a][1] + C[i][2] * k.g[a][2]) * iJ;
}

template <typename ET>
__device__ __forceinline__ double elem_E(const ET* E, double E0, int e) {
  return E ? (double)E[e] : E0;
}

template <typename ET>
__device__ __forceinline__ void stage_v(const LdsV& L, const int32_t* halo, const uint32_t* meta, int nacc, int nh,
                                        int nown, const double* X, const double* U, bool zero_acc) {
  if (zero_acc) {  // 16-byte stores (9 nacc doubles; the region is 16-byte aligned)
    for (int i = threadIdx.x; i < 9 * nacc / 2; i += kThreads) ((double2*)L.acc)[i] = make_double2(0.0, 0.0);
    if (((9 * nacc) & 1) && threadIdx.x == 0) L.acc[9 * nacc - 1] = 0.0;
  }
  for (int h = threadIdx.x; h < nh; h += kThreads) {
    const int64_t g = halo[h];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      L.xs[d * nh + h] = X[g * 3 + d];
      if (U) L.us[d * nh + h] = U[g * 3 + d];
    }
  }
  for (int a = threadIdx.x; a < nown; a += kThreads) {
    L.ms[a] = meta[a];
    L.accr[3 * a] = 0.0;
    L.accr[3 * a + 1] = 0.0;
    L.accr[3 * a + 2] = 0.0;
  }
}

// Tangent values (optional), residual (optional) and energy (optional).  LIN:
// linear-elastic stiffness (u = 0; U is not read, only WK is meaningful).
template <typename ET, typename OT, bool WK, bool WR, bool LIN = false>
__global__ __launch_bounds__(kThreads) void k_nh_assemble(PlanArgs p, const double* __restrict__ X,
                                                         const double* __restrict__ U, const ET* __restrict__ E,
                                                         double E0, double nu, OT* __restrict__ kvals,
                                                         double* __restrict__ R, double* __restrict__ energy) {
  extern __shared__ __align__(16) char smem[];
  __shared__ double red[kThreads / kWave];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, runoff = bi1.y, nrun = bi1.z, nacc = bi1.w;
  // without the tangent no CSR accumulators are carved, so residual and
  // energy calls also run on plans built without them
  const LdsV L = carve_v(smem, WK ? nacc : 0, nh, nown);
  // a lane's first record and its element's modulus do not depend on the
  // staged nodes: their loads are issued first
  const int t0 = threadIdx.x;
  const uint4 rc0 = t0 < nrec ? p.recs[roff + t0] : make_uint4(0, 0, 0, 0);
  const double E0v = (E && t0 < nrec) ? (double)E[p.rec_elem[roff + t0]] : E0;
  stage_v<ET>(L, p.halo + hoff, p.meta + hoff, nacc, nh, nown, X, LIN ? nullptr : U, WK);
  __syncthreads();
  const double mu_f = 1.0 / (2.0 * (1.0 + nu)), lam_f = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double en = 0.0;
  for (int r = threadIdx.x; r < nrec; r += kThreads) {
    const uint4 rc = r == t0 ? rc0 : p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    const uint64_t hi = (uint64_t)rc.z | ((uint64_t)rc.w << 32);
    int l[4];
    double x[4][3], u[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      l[a] = rec_local(lo, a);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        x[a][d] = L.xs[d * nh + l[a]];
        u[a][d] = LIN ? 0.0 : L.us[d * nh + l[a]];
      }
    }
    Kin k;
    kinematics<LIN>(x, u, k);
    const double Ee = r == t0 ? E0v : elem_E(E, E0, p.rec_elem[roff + r]);
    const double mu = mu_f * Ee, lam = lam_f * Ee;
    const double s = lam * k.lnJ - mu;
    if (energy) {
      int owned = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a) owned += l[a] < nown;
      double f2 = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) f2 += k.F[i][j] * k.F[i][j];
      const double psi = 0.5 * mu * (f2 - 3.0) - mu * k.lnJ + 0.5 * lam * k.lnJ * k.lnJ;
      en += 0.25 * owned * k.vol * psi;
    }
    if (WR) {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (l[a] >= nown) continue;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double pg = mu * (k.F[i][0] * k.g[a][0] + k.F[i][1] * k.g[a][1] + k.F[i][2] * k.g[a][2]) + s * k.w[a][i];
          atomicAdd(L.accr + 3 * l[a] + i, k.vol * pg);
        }
      }
    }
    if (WK) {
      // Off-diagonal blocks only: the diagonal block is minus their row sum
      // (sum_b g_b = 0, so sum_b K_ab = 0), formed after the loop.  K_ba is
      // the transpose of K_ab, so each unordered pair is evaluated once and
      // added to row a as K_ab and to row b as K_ab^T (each where owned).
      const double muv = mu * k.vol, lamv = lam * k.vol, sv = s * k.vol;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int bb = a + 1; bb < 4; ++bb) {
          const bool oa = l[a] < nown, ob = l[bb] < nown;
          if (!(oa || ob)) continue;
          const double gg = muv * (k.g[a][0] * k.g[bb][0] + k.g[a][1] * k.g[bb][1] + k.g[a][2] * k.g[bb][2]);
          double v[3][3];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
              v[i][kk] = (i == kk ? gg : 0.0) + lamv * k.w[a][i] * k.w[bb][kk] - sv * k.w[a][kk] * k.w[bb][i];
          if (oa) {  // rank of bb in row a: bb is the (bb-1)-th vertex other than a
            double* blk = L.acc + 9 * ((int)(L.ms[l[a]] & 0xffffu) + rec_rank4(lo, hi, a, bb - 1));
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int kk = 0; kk < 3; ++kk) atomicAdd(blk + 3 * i + kk, v[i][kk]);
          }
          if (ob) {  // rank of a in row bb: a < bb is the a-th vertex other than bb
            double* blk = L.acc + 9 * ((int)(L.ms[l[bb]] & 0xffffu) + rec_rank4(lo, hi, bb, a));
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int kk = 0; kk < 3; ++kk) atomicAdd(blk + 3 * i + kk, v[kk][i]);
          }
        }
    }
  }
  if (energy) {
    const double t = tmx::block_sum<kThreads>(en, red);
    if (threadIdx.x == 0) atomicAdd(energy, t);
  }
  __syncthreads();
  if (WK) {
    for (int t = threadIdx.x; t < 9 * nown; t += kThreads) {
      const int a = t / 9, c = t % 9;
      const uint32_t m = L.ms[a];
      const int base = (int)(m & 0xffffu), dr = (int)((m >> 16) & 0xffu), len = (int)(m >> 24);
      if (len == 0) continue;
      double sum = 0.0;
      for (int k = 0; k < len; ++k) sum += (k == dr) ? 0.0 : L.acc[9 * (base + k) + c];
      L.acc[9 * (base + dr) + c] = -sum;
    }
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (int q = w; q < nrun; q += kThreads / kWave) {
      const int4 ru = p.runs[runoff + q];
      const int64_t v0 = (int64_t)ru.x * 9;
      const int a0 = ru.y * 9, len = ru.z * 9;
      for (int t = lane; t < len; t += kWave) kvals[v0 + t] = (OT)L.acc[a0 + t];
    }
  }
  if (WR) {
    for (int a = threadIdx.x; a < nown; a += kThreads) {
      const int64_t g = p.halo[hoff + a];
      R[3 * g] = L.accr[3 * a];
      R[3 * g + 1] = L.accr[3 * a + 1];
      R[3 * g + 2] = L.accr[3 * a + 2];
    }
  }
}

// Y = K(u) lam (matrix-free tangent action) and optionally dE_e += sum_{a owned} lam_a . R_a / E_e.
template <typename ET>
__global__ __launch_bounds__(kThreads) void k_nh_apply(PlanArgs p, const double* __restrict__ X,
                                                      const double* __restrict__ U, const ET* __restrict__ E,
                                                      double E0, double nu, const double* __restrict__ lamv,
                                                      double* __restrict__ Y, double* __restrict__ dE) {
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x;
  // lam is staged in the acc region as [3][nh]
  const LdsV L = carve_v(smem, 0, nh, nown);
  double* lv = (double*)(L.ms + nown);
  lv = (double*)(((uintptr_t)lv + 15) & ~(uintptr_t)15);
  stage_v<ET>(L, p.halo + hoff, p.meta + hoff, 0, nh, nown, X, U, false);
  for (int h = threadIdx.x; h < nh; h += kThreads) {
    const int64_t g = p.halo[hoff + h];
#pragma unroll
    for (int d = 0; d < 3; ++d) lv[d * nh + h] = lamv[g * 3 + d];
  }
  __syncthreads();
  const double mu_f = 1.0 / (2.0 * (1.0 + nu)), lam_f = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  for (int r = threadIdx.x; r < nrec; r += kThreads) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    int l[4];
    double x[4][3], u[4][3], lm[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      l[a] = rec_local(lo, a);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        x[a][d] = L.xs[d * nh + l[a]];
        u[a][d] = L.us[d * nh + l[a]];
        lm[a][d] = lv[d * nh + l[a]];
      }
    }
    Kin k;
    kinematics(x, u, k);
    const int e = p.rec_elem[roff + r];
    const double Ee = elem_E(E, E0, e);
    const double mu = mu_f * Ee, lam = lam_f * Ee;
    const double s = lam * k.lnJ - mu;
    // per element: sum_b K_ab lam_b = V [mu (g_a . G) + lam w_a (w . lam) - s (W^T... )]
    double wl = 0.0;   // sum_b w_b . lam_b
    double Gl[3] = {0.0, 0.0, 0.0};  // sum_b lam_b (g_a . g_b) -> needs a, so keep M = sum_b lam_b g_b^T
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // sum_b lam_b g_b^T
    double Wt[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // sum_b w_b lam_b^T
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        wl += k.w[bb][i] * lm[bb][i];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          M[i][j] += lm[bb][i] * k.g[bb][j];
          Wt[i][j] += k.w[bb][i] * lm[bb][j];
        }
      }
    }
    (void)Gl;
    double de = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (l[a] >= nown) continue;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        // (K lam)_a,i = V [ mu sum_b (g_a.g_b) lam_b,i + lam w_a,i (sum_b w_b.lam_b) - s sum_b w_a,k lam_b,k w_b,i ]
        const double t1 = mu * (M[i][0] * k.g[a][0] + M[i][1] * k.g[a][1] + M[i][2] * k.g[a][2]);
        const double t2 = lam * k.w[a][i] * wl;
        const double t3 = s * (Wt[i][0] * k.w[a][0] + Wt[i][1] * k.w[a][1] + Wt[i][2] * k.w[a][2]);
        atomicAdd(L.accr + 3 * l[a] + i, k.vol * (t1 + t2 - t3));
        if (dE) {
          const double pg = mu * (k.F[i][0] * k.g[a][0] + k.F[i][1] * k.g[a][1] + k.F[i][2] * k.g[a][2]) + s * k.w[a][i];
          de += lm[a][i] * k.vol * pg;
        }
      }
    }
    if (dE && de != 0.0) atomicAdd(dE + e, de / Ee);
  }
  __syncthreads();
  for (int a = threadIdx.x; a < nown; a += kThreads) {
    const int64_t g = p.halo[hoff + a];
    Y[3 * g] = L.accr[3 * a];
    Y[3 * g + 1] = L.accr[3 * a + 1];
    Y[3 * g + 2] = L.accr[3 * a + 2];
  }
}

PlanArgs plan_args(const tm_plan_dev* d) {
  return PlanArgs{d->blk, d->halo, d->meta, (const uint4*)d->recs, d->rec_elem, (const int4*)d->runs, d->n_blocks};
}

template <typename K, typename... A>
int launch(K kernel, const tm_plan_dev* plan, int lds, hipStream_t st, A... args) {
  if (plan->n_blocks == 0) return TM_OK;
  if (lds > 65536) TM_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  hipLaunchKernelGGL(kernel, dim3(plan->n_blocks), dim3(kThreads), lds, st, args...);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

template <typename ET, typename OT>
int nh_dispatch(const tm_plan_dev* plan, const double* X, const double* U, const ET* E, double E0, double nu,
                OT* kv, double* R, double* en, hipStream_t st) {
  PlanArgs p = plan_args(plan);
  if (kv && R) return launch(k_nh_assemble<ET, OT, true, true>, plan, plan->lds_bytes, st, p, X, U, E, E0, nu, kv, R, en);
  if (kv) return launch(k_nh_assemble<ET, OT, true, false>, plan, plan->lds_bytes, st, p, X, U, E, E0, nu, kv, R, en);
  if (R) return launch(k_nh_assemble<ET, OT, false, true>, plan, plan->lds_bytes, st, p, X, U, E, E0, nu, kv, R, en);
  return launch(k_nh_assemble<ET, OT, false, false>, plan, plan->lds_bytes, st, p, X, U, E, E0, nu, kv, R, en);
}

}  // namespace

extern "C" int tm_asm_nh(const tm_plan_dev* plan, const double* coords, const double* u, const void* E_elem,
                         int E_dtype, double E0, double nu, void* kvals, int kv_dtype, double* R, double* energy,
                         void* stream) {
  TM_REQUIRE(plan && coords && u, TM_ERR_ARG, "null argument");
  TM_REQUIRE(nu > -1.0 && nu < 0.5, TM_ERR_ARG, "Poisson ratio %g outside (-1, 0.5)", nu);
  TM_REQUIRE(!kvals || plan->entry_bytes == 0 || plan->entry_bytes >= 72, TM_ERR_ARG,
             "the tangent needs a plan with 72-byte accumulators (this one has %d)", plan->entry_bytes);
  hipStream_t st = (hipStream_t)stream;
  if (E_elem == nullptr || E_dtype == TM_F64) {
    const double* E = (const double*)E_elem;
    if (kv_dtype == TM_F32) return nh_dispatch(plan, coords, u, E, E0, nu, (float*)kvals, R, energy, st);
    return nh_dispatch(plan, coords, u, E, E0, nu, (double*)kvals, R, energy, st);
  }
  const float* E = (const float*)E_elem;
  if (kv_dtype == TM_F32) return nh_dispatch(plan, coords, u, E, E0, nu, (float*)kvals, R, energy, st);
  return nh_dispatch(plan, coords, u, E, E0, nu, (double*)kvals, R, energy, st);
}

extern "C" int tm_asm_linear_elastic(const tm_plan_dev* plan, const double* coords, const void* E_elem, int E_dtype,
                                     double E0, double nu, void* kvals, int kv_dtype, void* stream) {
  TM_REQUIRE(plan && coords && kvals, TM_ERR_ARG, "null argument");
  TM_REQUIRE(nu > -1.0 && nu < 0.5, TM_ERR_ARG, "Poisson ratio %g outside (-1, 0.5)", nu);
  TM_REQUIRE(kv_dtype == TM_F32 || kv_dtype == TM_F64, TM_ERR_ARG, "stiffness values must be f32 or f64");
  TM_REQUIRE(plan->entry_bytes == 0 || plan->entry_bytes >= 72, TM_ERR_ARG,
             "the stiffness needs a plan with 72-byte accumulators (this one has %d)", plan->entry_bytes);
  hipStream_t st = (hipStream_t)stream;
  const PlanArgs p = plan_args(plan);
  if (E_elem == nullptr || E_dtype == TM_F64) {
    const double* E = (const double*)E_elem;
    if (kv_dtype == TM_F32)
      return launch(k_nh_assemble<double, float, true, false, true>, plan, plan->lds_bytes, st, p, coords,
                    (const double*)nullptr, E, E0, nu, (float*)kvals, (double*)nullptr, (double*)nullptr);
    return launch(k_nh_assemble<double, double, true, false, true>, plan, plan->lds_bytes, st, p, coords,
                  (const double*)nullptr, E, E0, nu, (double*)kvals, (double*)nullptr, (double*)nullptr);
  }
  TM_REQUIRE(E_dtype == TM_F32, TM_ERR_ARG, "E must be f32 or f64");
  const float* E = (const float*)E_elem;
  if (kv_dtype == TM_F32)
    return launch(k_nh_assemble<float, float, true, false, true>, plan, plan->lds_bytes, st, p, coords,
                  (const double*)nullptr, E, E0, nu, (float*)kvals, (double*)nullptr, (double*)nullptr);
  return launch(k_nh_assemble<float, double, true, false, true>, plan, plan->lds_bytes, st, p, coords,
                (const double*)nullptr, E, E0, nu, (double*)kvals, (double*)nullptr, (double*)nullptr);
}

extern "C" int tm_nh_apply(const tm_plan_dev* plan, const double* coords, const double* u, const void* E_elem,
                           int E_dtype, double E0, double nu, const double* lam, double* Y, double* dE,
                           void* stream) {
  TM_REQUIRE(plan && coords && u && lam && Y, TM_ERR_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  PlanArgs p = plan_args(plan);
  if (E_elem == nullptr || E_dtype == TM_F64)
    return launch(k_nh_apply<double>, plan, plan->lds_bytes, st, p, coords, u, (const double*)E_elem, E0, nu, lam, Y, dE);
  return launch(k_nh_apply<float>, plan, plan->lds_bytes, st, p, coords, u, (const float*)E_elem, E0, nu, lam, Y, dE);
}
