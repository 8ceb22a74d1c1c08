// P1 owner-computes assembly kernels for gfx950.
//
// One workgroup owns one plan block: a compact set of CSR rows (mesh nodes)
// plus every element touching them.  The workgroup
//   1. gathers the coordinates / nodal coefficient of the block's halo nodes
//      into LDS and zeroes an FP64 accumulator holding the owned rows' CSR
//      entries,
//   2. walks the block's element records (one element per lane, 16-byte
//      records carrying 10-bit halo-local ids and the 6-bit in-row rank of
//      every (owned row, column) pair), computes the element geometry in FP64
//      (cofactor vectors c_b = |J| grad phi_b and det J) and adds the owned
//      rows of the element matrix into LDS with ds_add_f64,
//   3. writes each owned row once, in runs of consecutive rows whose CSR
//      ranges are contiguous (coalesced stores, optionally narrowed to FP32).
// No global atomics, no sort, every CSR entry written exactly once.
#include "tm_device.h"

namespace {

using tmx::kWave;
constexpr int kThreads = 256;

struct PlanArgs {
  const int32_t* blk;
  const int32_t* halo;
  const uint32_t* meta;
  const uint4* recs;
  const int4* runs;
  int nblk;
};

struct QuadArg {
  double wsum;
  double W[4];
  double M[16];
};

__host__ __device__ constexpr size_t a16(size_t x) { return (x + 15) & ~size_t(15); }

// LDS regions of one block (sizes in elements), identical in every kernel.
struct Lds {
  double* acc;   // [nacc] owned-row CSR accumulators
  double* xs;    // [DIM][nh] halo coordinates (SoA)
  double* cs;    // [nh] halo coefficient / adjoint accumulator
  double* accb;  // [nown] owned-node vector accumulator
  uint32_t* ms;  // [nown] owned-row meta (acc offset | diag rank << 16)
};

__device__ __forceinline__ Lds carve(char* s, int nacc, int nh, int nown, int dim) {
  Lds L;
  size_t o = 0;
  L.acc = (double*)(s + o);
  o = a16(o + 8 * (size_t)nacc);
  L.xs = (double*)(s + o);
  o = a16(o + 8 * (size_t)dim * nh);
  L.cs = (double*)(s + o);
  o = a16(o + 8 * (size_t)nh);
  L.accb = (double*)(s + o);
  o = a16(o + 8 * (size_t)nown);
  L.ms = (uint32_t*)(s + o);
  return L;
}

template <typename T>
__device__ __forceinline__ double ld(const T* p, int64_t i) {
  return (double)p[i];
}

__device__ __forceinline__ int rec_local(uint64_t lo, int a) { return (int)((lo >> (10 * a)) & 1023u); }

template <int NB>
__device__ __forceinline__ int rec_rank(uint64_t lo, uint64_t hi, int a, int j) {
  const int bit = 40 + 6 * (a * (NB - 1) + j);
  return bit < 64 ? (int)((lo >> bit) & 63u) : (int)((hi >> (bit - 64)) & 63u);
}

// Cofactor vectors c[b] (grad phi_b = c[b] / det) and det J of a P1 simplex.
template <int DIM>
__device__ __forceinline__ void p1_geometry(const double (&x)[DIM + 1][DIM], double (&c)[DIM + 1][DIM],
                                            double& det);

template <>
__device__ __forceinline__ void p1_geometry<3>(const double (&x)[4][3], double (&c)[4][3], double& det) {
  double e1[3], e2[3], e3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    e1[d] = x[1][d] - x[0][d];
    e2[d] = x[2][d] - x[0][d];
    e3[d] = x[3][d] - x[0][d];
  }
  c[1][0] = e2[1] * e3[2] - e2[2] * e3[1];
  c[1][1] = e2[2] * e3[0] - e2[0] * e3[2];
  c[1][2] = e2[0] * e3[1] - e2[1] * e3[0];
  c[2][0] = e3[1] * e1[2] - e3[2] * e1[1];
  c[2][1] = e3[2] * e1[0] - e3[0] * e1[2];
  c[2][2] = e3[0] * e1[1] - e3[1] * e1[0];
  c[3][0] = e1[1] * e2[2] - e1[2] * e2[1];
  c[3][1] = e1[2] * e2[0] - e1[0] * e2[2];
  c[3][2] = e1[0] * e2[1] - e1[1] * e2[0];
  det = e1[0] * c[1][0] + e1[1] * c[1][1] + e1[2] * c[1][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) c[0][d] = -(c[1][d] + c[2][d] + c[3][d]);
}

template <>
__device__ __forceinline__ void p1_geometry<2>(const double (&x)[3][2], double (&c)[3][2], double& det) {
  const double e1x = x[1][0] - x[0][0], e1y = x[1][1] - x[0][1];
  const double e2x = x[2][0] - x[0][0], e2y = x[2][1] - x[0][1];
  c[1][0] = e2y;
  c[1][1] = -e2x;
  c[2][0] = -e1y;
  c[2][1] = e1x;
  det = e1x * e2y - e1y * e2x;
  c[0][0] = -(c[1][0] + c[2][0]);
  c[0][1] = -(c[1][1] + c[2][1]);
}

template <int DIM>
__device__ __forceinline__ double dotc(const double* a, const double* b) {
  double s = a[0] * b[0];
#pragma unroll
  for (int d = 1; d < DIM; ++d) s += a[d] * b[d];
  return s;
}

// Stages the block's halo coordinates (and optionally a nodal coefficient)
// into LDS and zeroes the accumulators.
template <int DIM, typename CT, bool COEF>
__device__ __forceinline__ void stage(const Lds& L, const int32_t* __restrict__ halo, const uint32_t* __restrict__ meta,
                                      int nacc, int nh, int nown, const double* __restrict__ X,
                                      const CT* __restrict__ coef) {
  for (int i = threadIdx.x; i < nacc; i += kThreads) L.acc[i] = 0.0;
  for (int h = threadIdx.x; h < nh; h += kThreads) {
    const int64_t g = halo[h];
#pragma unroll
    for (int d = 0; d < DIM; ++d) L.xs[d * nh + h] = X[g * DIM + d];
    L.cs[h] = COEF ? ld(coef, g) : 0.0;
  }
  for (int a = threadIdx.x; a < nown; a += kThreads) {
    L.ms[a] = meta[a];
    L.accb[a] = 0.0;
  }
}

// FORM: 0 laplace, 1 kappa-laplace (nodal coef), 2 mass; all multiplied by scale.
template <int NB, int FORM, typename CT, typename OT>
__global__ __launch_bounds__(kThreads) void k_asm_matrix(PlanArgs p, const double* __restrict__ X,
                                                        const CT* __restrict__ coef, double scale,
                                                        QuadArg q, OT* __restrict__ values) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, runoff = bi1.y, nrun = bi1.z, nacc = bi1.w;
  const Lds L = carve(smem, nacc, nh, nown, DIM);
  stage<DIM, CT, FORM == 1>(L, p.halo + hoff, p.meta + hoff, nacc, nh, nown, X, coef);
  __syncthreads();

  for (int r = threadIdx.x; r < nrec; r += kThreads) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    const uint64_t hi = (uint64_t)rc.z | ((uint64_t)rc.w << 32);
    int l[NB];
    double x[NB][DIM];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
#pragma unroll
      for (int d = 0; d < DIM; ++d) x[a][d] = L.xs[d * nh + l[a]];
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    const double adet = fabs(det);
    double s;
    if (FORM == 0) {
      s = scale * q.wsum / adet;
    } else if (FORM == 1) {
      double k = 0.0;
#pragma unroll
      for (int a = 0; a < NB; ++a) k += q.W[a] * L.cs[l[a]];
      s = scale * k / adet;
    } else {
      s = scale * adet;
    }
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      if (l[a] >= nown) continue;
      const uint32_t m = L.ms[l[a]];
      double* row = L.acc + (m & 0xffffu);
      double kaa = FORM == 2 ? s * q.M[a * NB + a] : s * dotc<DIM>(c[a], c[a]);
      atomicAdd(row + (m >> 16), kaa);
      int j = 0;
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        if (bb == a) continue;
        const double kab = FORM == 2 ? s * q.M[a * NB + bb] : s * dotc<DIM>(c[a], c[bb]);
        atomicAdd(row + rec_rank<NB>(lo, hi, a, j), kab);
        ++j;
      }
    }
  }
  __syncthreads();

  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int k = w; k < nrun; k += kThreads / kWave) {
    const int4 ru = p.runs[runoff + k];
    for (int t = lane; t < ru.z; t += kWave) values[(int64_t)ru.x + t] = (OT)L.acc[ru.y + t];
  }
}

// FORM: 0 constant load c, 1 nodal f.
template <int NB, int FORM, typename CT>
__global__ __launch_bounds__(kThreads) void k_asm_load(PlanArgs p, const double* __restrict__ X,
                                                      const CT* __restrict__ f, double cval, QuadArg q,
                                                      double* __restrict__ out) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w, nrec = bi1.x;
  const Lds L = carve(smem, 0, nh, nown, DIM);
  stage<DIM, CT, FORM == 1>(L, p.halo + hoff, p.meta + hoff, 0, nh, nown, X, f);
  __syncthreads();
  for (int r = threadIdx.x; r < nrec; r += kThreads) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    int l[NB];
    double x[NB][DIM];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
#pragma unroll
      for (int d = 0; d < DIM; ++d) x[a][d] = L.xs[d * nh + l[a]];
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    const double adet = fabs(det);
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      if (l[a] >= nown) continue;
      double v;
      if (FORM == 0) {
        v = cval * q.W[a];
      } else {
        v = 0.0;
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) v += q.M[a * NB + cc] * L.cs[l[cc]];
      }
      atomicAdd(L.accb + l[a], adet * v);
    }
  }
  __syncthreads();
  for (int a = threadIdx.x; a < nown; a += kThreads) out[p.halo[hoff + a]] = L.accb[a];
}

// dkappa_c += sum_e W_c / |det| * sum_{a owned, b} g[slot(a, b)] c_a . c_b
template <int NB, typename GT>
__global__ __launch_bounds__(kThreads) void k_kappa_adjoint(PlanArgs p, const double* __restrict__ X,
                                                           const GT* __restrict__ g, QuadArg q,
                                                           double* __restrict__ dkappa) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, runoff = bi1.y, nrun = bi1.z, nacc = bi1.w;
  const Lds L = carve(smem, nacc, nh, nown, DIM);
  stage<DIM, double, false>(L, p.halo + hoff, p.meta + hoff, 0, nh, nown, X, nullptr);
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int k = w; k < nrun; k += kThreads / kWave) {
    const int4 ru = p.runs[runoff + k];
    for (int t = lane; t < ru.z; t += kWave) L.acc[ru.y + t] = (double)g[(int64_t)ru.x + t];
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nrec; r += kThreads) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    const uint64_t hi = (uint64_t)rc.z | ((uint64_t)rc.w << 32);
    int l[NB];
    double x[NB][DIM];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
#pragma unroll
      for (int d = 0; d < DIM; ++d) x[a][d] = L.xs[d * nh + l[a]];
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      if (l[a] >= nown) continue;
      const uint32_t m = L.ms[l[a]];
      const double* row = L.acc + (m & 0xffffu);
      t += row[m >> 16] * dotc<DIM>(c[a], c[a]);
      int j = 0;
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        if (bb == a) continue;
        t += row[rec_rank<NB>(lo, hi, a, j)] * dotc<DIM>(c[a], c[bb]);
        ++j;
      }
    }
    t /= fabs(det);
#pragma unroll
    for (int a = 0; a < NB; ++a) atomicAdd(L.cs + l[a], q.W[a] * t);
  }
  __syncthreads();
  for (int h = threadIdx.x; h < nh; h += kThreads) {
    const double v = L.cs[h];
    if (v != 0.0) atomicAdd(dkappa + p.halo[hoff + h], v);
  }
}

PlanArgs plan_args(const tm_plan_dev* d) {
  return PlanArgs{d->blk, d->halo, d->meta, (const uint4*)d->recs, (const int4*)d->runs, d->n_blocks};
}

QuadArg quad_arg(const tm_quad* q) {
  QuadArg a;
  a.wsum = q->wsum;
  for (int i = 0; i < 4; ++i) a.W[i] = q->W[i];
  for (int i = 0; i < 16; ++i) a.M[i] = q->M[i];
  return a;
}

template <typename K, typename... A>
int launch_plan_kernel(K kernel, const tm_plan_dev* plan, hipStream_t st, A... args) {
  if (plan->n_blocks == 0) return TM_OK;
  if (plan->lds_bytes > 65536) {
    TM_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               plan->lds_bytes));
  }
  hipLaunchKernelGGL(kernel, dim3(plan->n_blocks), dim3(kThreads), plan->lds_bytes, st, args...);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

template <int NB, typename OT>
int dispatch_matrix(int form, const tm_plan_dev* plan, const double* X, const void* coef, int cdt,
                    double scale, QuadArg q, OT* values, hipStream_t st) {
  PlanArgs p = plan_args(plan);
  if (form == TM_FORM_LAPLACE)
    return launch_plan_kernel(k_asm_matrix<NB, 0, float, OT>, plan, st, p, X, (const float*)nullptr,
                              scale, q, values);
  if (form == TM_FORM_MASS)
    return launch_plan_kernel(k_asm_matrix<NB, 2, float, OT>, plan, st, p, X, (const float*)nullptr,
                              scale, q, values);
  if (form == TM_FORM_KAPPA_LAPLACE) {
    TM_REQUIRE(coef != nullptr, TM_ERR_ARG, "kappa-Laplace needs a nodal coefficient");
    if (cdt == TM_F32)
      return launch_plan_kernel(k_asm_matrix<NB, 1, float, OT>, plan, st, p, X, (const float*)coef,
                                scale, q, values);
    return launch_plan_kernel(k_asm_matrix<NB, 1, double, OT>, plan, st, p, X, (const double*)coef,
                              scale, q, values);
  }
  TM_REQUIRE(false, TM_ERR_ARG, "unknown form %d", form);
}

}  // namespace

extern "C" int tm_asm_p1_matrix(int form, int nb, int dim, const tm_plan_dev* plan,
                                const double* coords, const void* coef, int coef_dtype,
                                double scale, const tm_quad* quad, void* values, int value_dtype,
                                void* stream) {
  TM_REQUIRE(plan && quad && coords && values, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG,
             "only P1 triangles (nb=3, dim=2) and tetrahedra (nb=4, dim=3) are supported");
  hipStream_t st = (hipStream_t)stream;
  QuadArg q = quad_arg(quad);
  if (value_dtype == TM_F32) {
    if (nb == 3) return dispatch_matrix<3>(form, plan, coords, coef, coef_dtype, scale, q, (float*)values, st);
    return dispatch_matrix<4>(form, plan, coords, coef, coef_dtype, scale, q, (float*)values, st);
  }
  if (nb == 3) return dispatch_matrix<3>(form, plan, coords, coef, coef_dtype, scale, q, (double*)values, st);
  return dispatch_matrix<4>(form, plan, coords, coef, coef_dtype, scale, q, (double*)values, st);
}

extern "C" int tm_asm_p1_load(int form, int nb, int dim, const tm_plan_dev* plan, const double* coords,
                              const void* f, int f_dtype, double c, const tm_quad* quad, double* b,
                              void* stream) {
  TM_REQUIRE(plan && quad && coords && b, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG, "only P1 simplices");
  hipStream_t st = (hipStream_t)stream;
  QuadArg q = quad_arg(quad);
  PlanArgs p = plan_args(plan);
  if (form == TM_LOAD_CONST) {
    if (nb == 3) return launch_plan_kernel(k_asm_load<3, 0, float>, plan, st, p, coords, (const float*)nullptr, c, q, b);
    return launch_plan_kernel(k_asm_load<4, 0, float>, plan, st, p, coords, (const float*)nullptr, c, q, b);
  }
  TM_REQUIRE(form == TM_LOAD_NODAL && f != nullptr, TM_ERR_ARG, "bad load form");
  if (f_dtype == TM_F32) {
    if (nb == 3) return launch_plan_kernel(k_asm_load<3, 1, float>, plan, st, p, coords, (const float*)f, c, q, b);
    return launch_plan_kernel(k_asm_load<4, 1, float>, plan, st, p, coords, (const float*)f, c, q, b);
  }
  if (nb == 3) return launch_plan_kernel(k_asm_load<3, 1, double>, plan, st, p, coords, (const double*)f, c, q, b);
  return launch_plan_kernel(k_asm_load<4, 1, double>, plan, st, p, coords, (const double*)f, c, q, b);
}

extern "C" int tm_asm_p1_kappa_adjoint(int nb, int dim, const tm_plan_dev* plan, const double* coords,
                                       const void* g, int g_dtype, const tm_quad* quad, double* dkappa,
                                       void* stream) {
  TM_REQUIRE(plan && quad && coords && g && dkappa, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG, "only P1 simplices");
  hipStream_t st = (hipStream_t)stream;
  QuadArg q = quad_arg(quad);
  PlanArgs p = plan_args(plan);
  if (g_dtype == TM_F32) {
    if (nb == 3) return launch_plan_kernel(k_kappa_adjoint<3, float>, plan, st, p, coords, (const float*)g, q, dkappa);
    return launch_plan_kernel(k_kappa_adjoint<4, float>, plan, st, p, coords, (const float*)g, q, dkappa);
  }
  if (nb == 3) return launch_plan_kernel(k_kappa_adjoint<3, double>, plan, st, p, coords, (const double*)g, q, dkappa);
  return launch_plan_kernel(k_kappa_adjoint<4, double>, plan, st, p, coords, (const double*)g, q, dkappa);
}
