// Boundary-facet load vectors for P1 simplices (FacetAssembler of the
// reference, assemble/facet_assembler.py:123-277).
//
// One thread per selected facet (an element plus the local index of the
// vertex opposite the facet).  The integrand is linear in the cell basis
// function: F0(q) v + F1(q) . grad v, F0 / F1 tabulated by the caller at the
// facet quadrature points (H components each).  The thread computes the facet
// measure |F| in FP64 (edge length in 2D; twice the face area in 3D, the
// Jacobian of the map from the reference triangle), the cell gradients when
// F1 is given, the cell basis values at the quadrature points (the facet's
// barycentric coordinates; zero for the opposite vertex) and adds
//   b[node_a * H + h] += sum_q w_q |F| (F0[q][h] phi_a(q) + F1[q][h] . grad phi_a)
// with FP64 atomics.
#include "tm_device.h"

namespace {

constexpr int kFacetThreads = 256;
constexpr int kMaxQ = 16;
constexpr int kMaxH = 3;

struct FacetRule {
  double bary[kMaxQ][3];  // facet barycentric coordinates of each point
  double w[kMaxQ];        // weights (sum 1 on the unit interval, 1/2 on the reference triangle)
  int nq;
};

// PASS 0: plain FP64 atomics into b.  Deterministic mode: PASS 1 takes the
// largest |contribution| of every entry of b into rmax (bit patterns; non-
// negative doubles order like them), PASS 2 adds each contribution rounded to
// a multiple of 2^k with 2^(k+44) above that maximum: sums of up to 512 such
// terms are exact, so b does not depend on the order of the atomics.
template <int NB, int PASS>
__global__ __launch_bounds__(kFacetThreads) void k_facet_load(int64_t nf, const double* __restrict__ X,
                                                             const int64_t* __restrict__ cells,
                                                             const int32_t* __restrict__ opp, FacetRule rule,
                                                             int H, const double* __restrict__ F0,
                                                             const double* __restrict__ F1,
                                                             double* __restrict__ b, double* __restrict__ rmax) {
  constexpr int DIM = NB - 1;
  const int64_t f = blockIdx.x * (int64_t)kFacetThreads + threadIdx.x;
  if (f >= nf) return;
  int64_t node[NB];
  double x[NB][DIM];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    node[a] = cells[f * NB + a];
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[a][d] = X[node[a] * DIM + d];
  }
  const int o = opp[f];
  int fv[DIM];  // facet vertices: the cell vertices other than o, ascending
  {
    int k = 0;
#pragma unroll
    for (int a = 0; a < NB; ++a)
      if (a != o && k < DIM) fv[k++] = a;
  }
  double meas;
  if (DIM == 2) {
    const double ex = x[fv[1]][0] - x[fv[0]][0], ey = x[fv[1]][1] - x[fv[0]][1];
    meas = sqrt(ex * ex + ey * ey);
  } else {
    double e1[3], e2[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      e1[d] = x[fv[1]][d] - x[fv[0]][d];
      e2[d] = x[fv[DIM - 1]][d] - x[fv[0]][d];
    }
    const double cx = e1[1] * e2[2] - e1[2] * e2[1];
    const double cy = e1[2] * e2[0] - e1[0] * e2[2];
    const double cz = e1[0] * e2[1] - e1[1] * e2[0];
    meas = sqrt(cx * cx + cy * cy + cz * cz);
  }
  // cell gradients grad phi_a = c_a / det (only when F1 is given)
  double g[NB][DIM];
  if (F1) {
    double J[DIM][DIM];  // J[i][j] = x_{j+1,i} - x_{0,i}
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) J[i][j] = x[j + 1][i] - x[0][i];
    double inv[DIM][DIM], det;
    if (DIM == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      inv[0][0] = J[1][1] / det;
      inv[0][1] = -J[0][1] / det;
      inv[1][0] = -J[1][0] / det;
      inv[1][1] = J[0][0] / det;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      inv[0][0] = c00 / det;
      inv[1][0] = c01 / det;
      inv[2][0] = c02 / det;
      inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
      inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
      inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    }
    // reference gradients: phi_0 = 1 - sum xi, phi_k = xi_{k-1}; grad phi = J^-T grad_xi phi
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          const double gr = a == 0 ? -1.0 : (a - 1 == i ? 1.0 : 0.0);
          s += inv[i][j] * gr;
        }
        g[a][j] = s;
      }
  }
  double acc[NB][kMaxH];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int h = 0; h < kMaxH; ++h) acc[a][h] = 0.0;
  for (int q = 0; q < rule.nq; ++q) {
    const double fxw = rule.w[q] * meas;
    double phi[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) phi[a] = 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) phi[fv[k]] = rule.bary[q][k];
#pragma unroll
    for (int h = 0; h < kMaxH; ++h) {
      if (h >= H) break;
      const double f0 = F0[(f * rule.nq + q) * H + h];
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        double v = f0 * phi[a];
        if (F1) {
          const double* f1 = F1 + ((f * rule.nq + q) * H + h) * DIM;
#pragma unroll
          for (int j = 0; j < DIM; ++j) v += f1[j] * g[a][j];
        }
        acc[a][h] += fxw * v;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int h = 0; h < kMaxH; ++h) {
      if (h >= H) continue;
      const int64_t i = node[a] * H + h;
      if (PASS == 0) {
        atomicAdd(b + i, acc[a][h]);
      } else if (PASS == 1) {
        atomicMax((unsigned long long*)(rmax + i), (unsigned long long)__double_as_longlong(fabs(acc[a][h])));
      } else {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(rmax[i]);
        const int k = bits ? (int)((bits >> 52) & 0x7ffull) - 1023 + 1 - 44 : 0;
        atomicAdd(b + i, ldexp(rint(ldexp(acc[a][h], -k)), k));
      }
    }
}

}  // namespace

template <int NB>
void launch_facets(int pass, int grid, hipStream_t st, int64_t nf, const double* X, const int64_t* cells,
                   const int32_t* opp, const FacetRule& rule, int H, const double* F0, const double* F1, double* b,
                   double* rmax) {
  if (pass == 0)
    hipLaunchKernelGGL((k_facet_load<NB, 0>), dim3(grid), dim3(kFacetThreads), 0, st, nf, X, cells, opp, rule, H, F0,
                       F1, b, rmax);
  else if (pass == 1)
    hipLaunchKernelGGL((k_facet_load<NB, 1>), dim3(grid), dim3(kFacetThreads), 0, st, nf, X, cells, opp, rule, H, F0,
                       F1, b, rmax);
  else
    hipLaunchKernelGGL((k_facet_load<NB, 2>), dim3(grid), dim3(kFacetThreads), 0, st, nf, X, cells, opp, rule, H, F0,
                       F1, b, rmax);
}

extern "C" int tm_facet_load(int nb, int dim, int64_t n_facets, const double* coords, const int64_t* cells,
                             const int32_t* opposite, int n_q, const double* q_bary, const double* q_w, int H,
                             const double* F0, const double* F1, double* b, int flags, double* scratch,
                             void* stream) {
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG,
             "only P1 triangles (nb=3, dim=2) and tetrahedra (nb=4, dim=3) are supported");
  TM_REQUIRE(n_facets >= 0 && n_q >= 1 && n_q <= kMaxQ && H >= 1 && H <= kMaxH, TM_ERR_ARG,
             "bad facet load arguments (n_q %d, H %d)", n_q, H);
  if (n_facets == 0) return TM_OK;
  TM_REQUIRE(coords && cells && opposite && q_bary && q_w && F0 && b, TM_ERR_ARG, "null argument");
  FacetRule rule{};
  rule.nq = n_q;
  for (int q = 0; q < n_q; ++q) {
    rule.w[q] = q_w[q];
    for (int k = 0; k < dim; ++k) rule.bary[q][k] = q_bary[q * dim + k];
  }
  TM_REQUIRE((flags & ~TM_ASM_DETERMINISTIC) == 0, TM_ERR_ARG, "unknown flags");
  const bool det = (flags & TM_ASM_DETERMINISTIC) != 0;
  TM_REQUIRE(!det || scratch, TM_ERR_ARG, "the deterministic mode needs a zeroed scratch array of N*H doubles");
  hipStream_t st = (hipStream_t)stream;
  const int grid = (int)((n_facets + kFacetThreads - 1) / kFacetThreads);
  for (int pass = det ? 1 : 0; pass <= (det ? 2 : 0); ++pass) {
    if (nb == 3)
      launch_facets<3>(pass, grid, st, n_facets, coords, cells, opposite, rule, H, F0, F1, b, scratch);
    else
      launch_facets<4>(pass, grid, st, n_facets, coords, cells, opposite, rule, H, F0, F1, b, scratch);
    TM_HIP(hipGetLastError());
  }
  return TM_OK;
}
