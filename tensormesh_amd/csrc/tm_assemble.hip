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

// LDS regions of one block, identical in every kernel:
//   acc   [nacc]   owned-row CSR accumulators (FP64)
//   nodes [nh]     halo nodes as double4: 3D {x, y, z, coef}, 2D {x, y, coef, -}
//   accb  [nown]   owned-node vector accumulator (FP64)
//   ms    [nown]   owned-row meta: acc offset | diag rank << 16
struct Lds {
  double* acc;
  double4* nodes;
  double* accb;
  uint32_t* ms;
};

__device__ __forceinline__ Lds carve(char* s, int nacc, int nh, int nown) {
  Lds L;
  size_t o = 0;
  L.acc = (double*)(s + o);
  o = a16(o + 8 * (size_t)nacc);
  L.nodes = (double4*)(s + o);
  o = a16(o + 32 * (size_t)nh);
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

// In-row rank of vertex b in the row of owned vertex a (a != b).
template <int NB>
__device__ __forceinline__ int rec_rank(uint64_t lo, uint64_t hi, int a, int b) {
  const int j = b - (b > a ? 1 : 0);
  const int bit = 40 + 6 * (a * (NB - 1) + j);
  return bit < 64 ? (int)((lo >> bit) & 63u) : (int)((hi >> (bit - 64)) & 63u);
}

template <int DIM>
__device__ __forceinline__ void node_xyz(const double4& n, double (&x)[DIM]);
template <>
__device__ __forceinline__ void node_xyz<3>(const double4& n, double (&x)[3]) {
  x[0] = n.x;
  x[1] = n.y;
  x[2] = n.z;
}
template <>
__device__ __forceinline__ void node_xyz<2>(const double4& n, double (&x)[2]) {
  x[0] = n.x;
  x[1] = n.y;
}
template <int DIM>
__device__ __forceinline__ double node_coef(const double4& n) {
  return DIM == 3 ? n.w : n.z;
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

// 1/x to full FP64 precision: hardware reciprocal estimate refined by two
// Newton steps (each doubles the correct bits).
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// Stages the block's halo nodes (and optionally a nodal coefficient) into LDS
// and zeroes the accumulators.
template <int NT, int DIM, typename CT, bool COEF>
__device__ __forceinline__ void stage(const Lds& L, const int32_t* __restrict__ halo, const uint32_t* __restrict__ meta,
                                      int nacc, int nh, int nown, const double* __restrict__ X,
                                      const CT* __restrict__ coef) {
  for (int i = threadIdx.x; i < nacc; i += NT) L.acc[i] = 0.0;
  for (int h = threadIdx.x; h < nh; h += NT) {
    const int64_t g = halo[h];
    const double cf = COEF ? ld(coef, g) : 0.0;
    double4 n;
    if (DIM == 3)
      n = make_double4(X[g * 3], X[g * 3 + 1], X[g * 3 + 2], cf);
    else
      n = make_double4(X[g * 2], X[g * 2 + 1], cf, 0.0);
    L.nodes[h] = n;
  }
  for (int a = threadIdx.x; a < nown; a += NT) {
    L.ms[a] = meta[a];
    L.accb[a] = 0.0;
  }
}

// Sets the diagonal of every owned row to minus the sum of its off-diagonal
// entries (exact for forms whose element rows sum to zero: sum_b c_b = 0).
template <int NT>
__device__ __forceinline__ void rowsum_diagonal(const Lds& L, int nown, int nacc) {
  for (int a = threadIdx.x; a < nown; a += NT) {
    const uint32_t m = L.ms[a];
    const int b0 = (int)(m & 0xffffu);
    const int b1 = a + 1 < nown ? (int)(L.ms[a + 1] & 0xffffu) : nacc;
    if (b1 <= b0) continue;
    const int d = b0 + (int)(m >> 16);
    double s = 0.0;
    for (int k = b0; k < b1; ++k) s += (k == d) ? 0.0 : L.acc[k];
    L.acc[d] = -s;
  }
}

// FORM: 0 laplace, 1 kappa-laplace (nodal coef), 2 mass; all multiplied by scale.
template <int NB, int FORM, typename CT, typename OT, int NT>
__global__ __launch_bounds__(NT) void k_asm_matrix(PlanArgs p, const double* __restrict__ X,
                                                  const CT* __restrict__ coef, double scale, QuadArg q,
                                                  OT* __restrict__ values) {
  constexpr int DIM = NB - 1;
  constexpr bool ROWSUM = FORM != 2;  // diagonal from the zero row sum
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, runoff = bi1.y, nrun = bi1.z, nacc = bi1.w;
  const Lds L = carve(smem, nacc, nh, nown);
  stage<NT, DIM, CT, FORM == 1>(L, p.halo + hoff, p.meta + hoff, nacc, nh, nown, X, coef);
  __syncthreads();

  uint4 nxt = threadIdx.x < nrec ? p.recs[roff + threadIdx.x] : make_uint4(0, 0, 0, 0);
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = nxt;
    if (r + NT < nrec) nxt = p.recs[roff + r + NT];  // prefetch the next record
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    const uint64_t hi = (uint64_t)rc.z | ((uint64_t)rc.w << 32);
    int l[NB];
    double x[NB][DIM], cf[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      const double4 n = L.nodes[l[a]];
      node_xyz<DIM>(n, x[a]);
      cf[a] = node_coef<DIM>(n);
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    const double adet = fabs(det);
    double s;
    if (FORM == 0) {
      s = scale * q.wsum * recip(adet);
    } else if (FORM == 1) {
      double k = 0.0;
#pragma unroll
      for (int a = 0; a < NB; ++a) k += q.W[a] * cf[a];
      s = scale * k * recip(adet);
    } else {
      s = scale * adet;
    }
    bool own[NB];
    uint32_t m[NB];
    bool all = true;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      own[a] = l[a] < nown;
      all = all && own[a];
      m[a] = own[a] ? L.ms[l[a]] : 0u;
    }
    if (__all(all)) {
      // every vertex owned in every lane: no branches (records are sorted so
      // that the block's interior elements form whole waves)
      if (!ROWSUM) {
#pragma unroll
        for (int a = 0; a < NB; ++a) atomicAdd(L.acc + (m[a] & 0xffffu) + (m[a] >> 16), s * q.M[a * NB + a]);
      }
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int bb = a + 1; bb < NB; ++bb) {
          const double kab = FORM == 2 ? s * q.M[a * NB + bb] : s * dotc<DIM>(c[a], c[bb]);
          atomicAdd(L.acc + (m[a] & 0xffffu) + rec_rank<NB>(lo, hi, a, bb), kab);
          atomicAdd(L.acc + (m[bb] & 0xffffu) + rec_rank<NB>(lo, hi, bb, a), kab);
        }
      continue;
    }
    if (!ROWSUM) {
#pragma unroll
      for (int a = 0; a < NB; ++a)
        if (own[a]) atomicAdd(L.acc + (m[a] & 0xffffu) + (m[a] >> 16), s * q.M[a * NB + a]);
    }
    // each symmetric pair once: K_ab = K_ba
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int bb = a + 1; bb < NB; ++bb) {
        if (!(own[a] || own[bb])) continue;
        const double kab = FORM == 2 ? s * q.M[a * NB + bb] : s * dotc<DIM>(c[a], c[bb]);
        if (own[a]) atomicAdd(L.acc + (m[a] & 0xffffu) + rec_rank<NB>(lo, hi, a, bb), kab);
        if (own[bb]) atomicAdd(L.acc + (m[bb] & 0xffffu) + rec_rank<NB>(lo, hi, bb, a), kab);
      }
  }
  __syncthreads();
  if (ROWSUM) {
    rowsum_diagonal<NT>(L, nown, nacc);
    __syncthreads();
  }

  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int k = w; k < nrun; k += NT / kWave) {
    const int4 ru = p.runs[runoff + k];
    for (int t = lane; t < ru.z; t += kWave) values[(int64_t)ru.x + t] = (OT)L.acc[ru.y + t];
  }
}

// FORM: 0 constant load c, 1 nodal f.
template <int NB, int FORM, typename CT, int NT>
__global__ __launch_bounds__(NT) void k_asm_load(PlanArgs p, const double* __restrict__ X,
                                                const CT* __restrict__ f, double cval, QuadArg q,
                                                double* __restrict__ out) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w, nrec = bi1.x;
  const Lds L = carve(smem, 0, nh, nown);
  stage<NT, DIM, CT, FORM == 1>(L, p.halo + hoff, p.meta + hoff, 0, nh, nown, X, f);
  __syncthreads();
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    int l[NB];
    double x[NB][DIM], cf[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      const double4 n = L.nodes[l[a]];
      node_xyz<DIM>(n, x[a]);
      cf[a] = node_coef<DIM>(n);
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
        for (int cc = 0; cc < NB; ++cc) v += q.M[a * NB + cc] * cf[cc];
      }
      atomicAdd(L.accb + l[a], adet * v);
    }
  }
  __syncthreads();
  for (int a = threadIdx.x; a < nown; a += NT) out[p.halo[hoff + a]] = L.accb[a];
}

// dkappa_c += sum_e W_c / |det| * sum_{a owned, b} g[slot(a, b)] c_a . c_b
// (per-node accumulator: the 4th double of each LDS node, unused by this kernel)
template <int NB, typename GT, int NT>
__global__ __launch_bounds__(NT) void k_kappa_adjoint(PlanArgs p, const double* __restrict__ X,
                                                     const GT* __restrict__ g, QuadArg q,
                                                     double* __restrict__ dkappa) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, runoff = bi1.y, nrun = bi1.z, nacc = bi1.w;
  const Lds L = carve(smem, nacc, nh, nown);
  stage<NT, DIM, double, false>(L, p.halo + hoff, p.meta + hoff, 0, nh, nown, X, nullptr);
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int k = w; k < nrun; k += NT / kWave) {
    const int4 ru = p.runs[runoff + k];
    for (int t = lane; t < ru.z; t += kWave) L.acc[ru.y + t] = (double)g[(int64_t)ru.x + t];
  }
  __syncthreads();
  double* acc_node = (double*)L.nodes;  // element 4*h + 3 of node h
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    const uint64_t hi = (uint64_t)rc.z | ((uint64_t)rc.w << 32);
    int l[NB];
    double x[NB][DIM];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      node_xyz<DIM>(L.nodes[l[a]], x[a]);
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
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        if (bb == a) continue;
        t += row[rec_rank<NB>(lo, hi, a, bb)] * dotc<DIM>(c[a], c[bb]);
      }
    }
    t *= recip(fabs(det));
#pragma unroll
    for (int a = 0; a < NB; ++a) atomicAdd(acc_node + 4 * l[a] + 3, q.W[a] * t);
  }
  __syncthreads();
  for (int h = threadIdx.x; h < nh; h += NT) {
    const double v = acc_node[4 * h + 3];
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
int launch_plan_kernel(K kernel, int nt, const tm_plan_dev* plan, hipStream_t st, A... args) {
  if (plan->n_blocks == 0) return TM_OK;
  if (plan->lds_bytes > 65536) {
    TM_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               plan->lds_bytes));
  }
  hipLaunchKernelGGL(kernel, dim3(plan->n_blocks), dim3(nt), plan->lds_bytes, st, args...);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

// Workgroup size of the plan kernels: plan->threads (256, 512 or 1024; 0 = 256).
#define TM_BY_THREADS(plan, KERNEL_NT, ...)                                        \
  do {                                                                             \
    const int nt_ = (plan)->threads == 1024 ? 1024 : ((plan)->threads == 512 ? 512 : 256); \
    if (nt_ == 1024) { constexpr int NT = 1024; return launch_plan_kernel(KERNEL_NT, NT, plan, __VA_ARGS__); } \
    if (nt_ == 512) { constexpr int NT = 512; return launch_plan_kernel(KERNEL_NT, NT, plan, __VA_ARGS__); }   \
    { constexpr int NT = 256; return launch_plan_kernel(KERNEL_NT, NT, plan, __VA_ARGS__); }                   \
  } while (0)

template <int NB, typename OT>
int dispatch_matrix(int form, const tm_plan_dev* plan, const double* X, const void* coef, int cdt,
                    double scale, QuadArg q, OT* values, hipStream_t st) {
  PlanArgs p = plan_args(plan);
  if (form == TM_FORM_LAPLACE)
    TM_BY_THREADS(plan, (k_asm_matrix<NB, 0, float, OT, NT>), st, p, X, (const float*)nullptr, scale, q, values);
  if (form == TM_FORM_MASS)
    TM_BY_THREADS(plan, (k_asm_matrix<NB, 2, float, OT, NT>), st, p, X, (const float*)nullptr, scale, q, values);
  if (form == TM_FORM_KAPPA_LAPLACE) {
    TM_REQUIRE(coef != nullptr, TM_ERR_ARG, "kappa-Laplace needs a nodal coefficient");
    if (cdt == TM_F32)
      TM_BY_THREADS(plan, (k_asm_matrix<NB, 1, float, OT, NT>), st, p, X, (const float*)coef, scale, q, values);
    TM_BY_THREADS(plan, (k_asm_matrix<NB, 1, double, OT, NT>), st, p, X, (const double*)coef, scale, q, values);
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
    if (nb == 3) return launch_plan_kernel(k_asm_load<3, 0, float, 256>, 256, plan, st, p, coords, (const float*)nullptr, c, q, b);
    return launch_plan_kernel(k_asm_load<4, 0, float, 256>, 256, plan, st, p, coords, (const float*)nullptr, c, q, b);
  }
  TM_REQUIRE(form == TM_LOAD_NODAL && f != nullptr, TM_ERR_ARG, "bad load form");
  if (f_dtype == TM_F32) {
    if (nb == 3) return launch_plan_kernel(k_asm_load<3, 1, float, 256>, 256, plan, st, p, coords, (const float*)f, c, q, b);
    return launch_plan_kernel(k_asm_load<4, 1, float, 256>, 256, plan, st, p, coords, (const float*)f, c, q, b);
  }
  if (nb == 3) return launch_plan_kernel(k_asm_load<3, 1, double, 256>, 256, plan, st, p, coords, (const double*)f, c, q, b);
  return launch_plan_kernel(k_asm_load<4, 1, double, 256>, 256, plan, st, p, coords, (const double*)f, c, q, b);
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
    if (nb == 3) return launch_plan_kernel(k_kappa_adjoint<3, float, 256>, 256, plan, st, p, coords, (const float*)g, q, dkappa);
    return launch_plan_kernel(k_kappa_adjoint<4, float, 256>, 256, plan, st, p, coords, (const float*)g, q, dkappa);
  }
  if (nb == 3) return launch_plan_kernel(k_kappa_adjoint<3, double, 256>, 256, plan, st, p, coords, (const double*)g, q, dkappa);
  return launch_plan_kernel(k_kappa_adjoint<4, double, 256>, 256, plan, st, p, coords, (const double*)g, q, dkappa);
}
