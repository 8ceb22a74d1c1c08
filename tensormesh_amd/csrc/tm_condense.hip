// Dirichlet static condensation on the device (Condenser._compute_layout,
// __call__, condense_rhs and recover of tensormesh/operator/condense.py:
// 182-220, 238-309, 388-458).
//
// Layout of a canonical CSR matrix (rows ascending, columns ascending in a
// row) and a Dirichlet mask:
//   * DOF map: an inner DOF's local id is its rank among the inner DOFs in
//     increasing global order, an outer DOF's its rank among the outer DOFs;
//   * inner-inner (ii) and inner-outer (io) edges in the original edge order,
//     as condensed CSR views (local columns) plus the edge ids into the full
//     value array.
// The ranks come from one exclusive prefix sum over the rows of three
// counters (inner flag, ii entries, io entries): per-workgroup sums, one
// workgroup scanning those sums, then per-workgroup scans with the offsets.
// Integer arithmetic only, so the result is exact and deterministic.
#include "tm_device.h"

namespace {

constexpr int NT = 256;     // rows per workgroup in the count/offset/fill kernels
constexpr int NS = 1024;    // lanes of the single-workgroup scan of the partial sums

struct CondWork {
  int32_t* cnt;    // [3][n]  inner flag, ii entries, io entries of each row
  int32_t* off;    // [3][n]  exclusive prefix sums of cnt
  int32_t* bsum;   // [3][nb] per-workgroup sums, scanned in place
  int32_t* tot;    // [4]     n_inner, nnz_ii, nnz_io, overflow flag
};

__host__ __device__ inline int64_t n_blocks(int64_t n) { return (n + NT - 1) / NT; }
__host__ __device__ inline size_t a256(size_t x) { return (x + 255) & ~size_t(255); }

__host__ __device__ inline CondWork cond_carve(void* work, int64_t n) {
  char* p = (char*)work;
  const int64_t nb = n_blocks(n);
  CondWork w;
  size_t o = 0;
  w.cnt = (int32_t*)(p + o);
  o = a256(o + 12 * (size_t)n);
  w.off = (int32_t*)(p + o);
  o = a256(o + 12 * (size_t)n);
  w.bsum = (int32_t*)(p + o);
  o = a256(o + 12 * (size_t)nb);
  w.tot = (int32_t*)(p + o);
  return w;
}

// per row: inner flag and the number of inner / outer columns of an inner row
__global__ __launch_bounds__(NT) void k_cond_count(int64_t n, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci, const uint8_t* __restrict__ dir,
                                                   CondWork w) {
  __shared__ int32_t red[3][NT / tmx::kWave];
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  int32_t c[3] = {0, 0, 0};
  if (i < n && !dir[i]) {
    c[0] = 1;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
      if (dir[ci[k]])
        ++c[2];
      else
        ++c[1];
    }
  }
  if (i < n)
#pragma unroll
    for (int j = 0; j < 3; ++j) w.cnt[(int64_t)j * n + i] = c[j];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    int32_t v = c[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, tmx::kWave);
    if (threadIdx.x % tmx::kWave == 0) red[j][threadIdx.x / tmx::kWave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < NT / tmx::kWave; ++k) s += red[threadIdx.x][k];
    w.bsum[(int64_t)threadIdx.x * n_blocks(n) + blockIdx.x] = s;
  }
}

// One workgroup: exclusive scan of the per-workgroup sums (three columns) in
// place; totals to w.tot and, when given, to sizes = {n_inner, n_outer,
// nnz_ii, nnz_io}.
__global__ __launch_bounds__(NS) void k_cond_scan_blocks(int64_t n, CondWork w, int64_t* sizes) {
  __shared__ int64_t part[NS];
  const int64_t nb = n_blocks(n);
  const int64_t per = (nb + NS - 1) / NS;
  const int64_t b0 = threadIdx.x * per, b1 = b0 + per < nb ? b0 + per : nb;
  int64_t tot[3];
  for (int j = 0; j < 3; ++j) {
    int32_t* col = w.bsum + (int64_t)j * nb;
    int64_t s = 0;
    for (int64_t b = b0; b < b1; ++b) s += col[b];
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of the NS thread sums
    for (int o = 1; o < NS; o <<= 1) {
      const int64_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    int64_t run = part[threadIdx.x] - s;  // exclusive prefix of this thread's range
    for (int64_t b = b0; b < b1; ++b) {
      const int32_t v = col[b];
      col[b] = (int32_t)run;
      run += v;
    }
    tot[j] = part[NS - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const bool over = tot[1] >= (int64_t(1) << 31) || tot[2] >= (int64_t(1) << 31);
    w.tot[0] = (int32_t)tot[0];
    w.tot[1] = (int32_t)tot[1];
    w.tot[2] = (int32_t)tot[2];
    w.tot[3] = over ? 1 : 0;
    if (sizes) {
      sizes[0] = tot[0];
      sizes[1] = n - tot[0];
      sizes[2] = tot[1];
      sizes[3] = tot[2];
    }
  }
}

// Per-workgroup exclusive scans plus the workgroup offsets; the DOF map.
__global__ __launch_bounds__(NT) void k_cond_offsets(int64_t n, const uint8_t* __restrict__ dir, CondWork w,
                                                     int32_t* __restrict__ local) {
  __shared__ int32_t s[3][NT];
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  const int64_t nb = n_blocks(n);
#pragma unroll
  for (int j = 0; j < 3; ++j) s[j][threadIdx.x] = i < n ? w.cnt[(int64_t)j * n + i] : 0;
  __syncthreads();
  for (int o = 1; o < NT; o <<= 1) {
    int32_t v[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = threadIdx.x >= o ? s[j][threadIdx.x - o] : 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 3; ++j) s[j][threadIdx.x] += v[j];
    __syncthreads();
  }
  if (i < n) {
    int32_t off[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      off[j] = w.bsum[(int64_t)j * nb + blockIdx.x] + s[j][threadIdx.x] - w.cnt[(int64_t)j * n + i];
      w.off[(int64_t)j * n + i] = off[j];
    }
    local[i] = dir[i] ? (int32_t)(i - off[0]) : off[0];
  }
}

// Condensed CSR views, edge ids, DOF lists, optional COO (int64) and edge
// masks.  Thread per row; an inner row writes its ii and io ranges in edge
// order, which keeps each row's local columns ascending.
__global__ __launch_bounds__(NT) void k_cond_fill(int64_t n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci, const uint8_t* __restrict__ dir,
                                                  CondWork w, const int32_t* __restrict__ local,
                                                  int32_t* __restrict__ rp_ii, int32_t* __restrict__ ci_ii,
                                                  int32_t* __restrict__ e_ii, int32_t* __restrict__ rp_io,
                                                  int32_t* __restrict__ ci_io, int32_t* __restrict__ e_io,
                                                  int32_t* __restrict__ inner_dof, int32_t* __restrict__ outer_dof,
                                                  int64_t* __restrict__ coo_ii, int64_t* __restrict__ coo_io,
                                                  uint8_t* __restrict__ m_ii, uint8_t* __restrict__ m_io) {
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  const int32_t n_inner = w.tot[0], nnz_ii = w.tot[1], nnz_io = w.tot[2];
  if (i == 0) {
    rp_ii[n_inner] = nnz_ii;
    rp_io[n_inner] = nnz_io;
  }
  if (i >= n) return;
  const int32_t b = rp[i], e = rp[i + 1];
  if (dir[i]) {
    outer_dof[local[i]] = (int32_t)i;
    if (m_ii)
      for (int32_t k = b; k < e; ++k) {
        m_ii[k] = 0;
        m_io[k] = 0;
      }
    return;
  }
  const int32_t li = local[i];
  int32_t o = w.off[(int64_t)1 * n + i], p = w.off[(int64_t)2 * n + i];
  rp_ii[li] = o;
  rp_io[li] = p;
  inner_dof[li] = (int32_t)i;
  for (int32_t k = b; k < e; ++k) {
    const int32_t c = ci[k];
    const bool outer = dir[c] != 0;
    if (!outer) {
      ci_ii[o] = local[c];
      e_ii[o] = k;
      if (coo_ii) {
        coo_ii[o] = li;
        coo_ii[(int64_t)nnz_ii + o] = local[c];
      }
      ++o;
    } else {
      ci_io[p] = local[c];
      e_io[p] = k;
      if (coo_io) {
        coo_io[p] = li;
        coo_io[(int64_t)nnz_io + p] = local[c];
      }
      ++p;
    }
    if (m_ii) {
      m_ii[k] = outer ? 0 : 1;
      m_io[k] = outer ? 1 : 0;
    }
  }
}

// out[li] = rhs[inner_dof[li]] - sum_k v_io[k] u_outer[ci_io[k]]
template <typename VT>
__global__ __launch_bounds__(NT) void k_cond_rhs(int64_t n_inner, const int32_t* __restrict__ inner_dof,
                                                 const double* __restrict__ rhs, const int32_t* __restrict__ rp_io,
                                                 const int32_t* __restrict__ ci_io, const VT* __restrict__ v_io,
                                                 const double* __restrict__ u_outer, double* __restrict__ out) {
  for (int64_t li = (int64_t)blockIdx.x * NT + threadIdx.x; li < n_inner; li += (int64_t)gridDim.x * NT) {
    double s = 0.0;
    if (u_outer)
      for (int32_t k = rp_io[li]; k < rp_io[li + 1]; ++k) s += (double)v_io[k] * u_outer[ci_io[k]];
    out[li] = rhs[inner_dof[li]] - s;
  }
}

// out[i] = u_outer[local[i]] on Dirichlet DOFs (0 without u_outer), u_inner[local[i]] elsewhere
__global__ __launch_bounds__(NT) void k_cond_recover(int64_t n, const uint8_t* __restrict__ dir,
                                                     const int32_t* __restrict__ local,
                                                     const double* __restrict__ u_inner,
                                                     const double* __restrict__ u_outer, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const int32_t l = local[i];
    out[i] = dir[i] ? (u_outer ? u_outer[l] : 0.0) : u_inner[l];
  }
}

int grid_rows(int64_t n) {
  const int64_t g = (n + NT - 1) / NT;
  return (int)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

}  // namespace

extern "C" int64_t tm_condense_work_size(int64_t n) {
  if (n < 0) return -1;
  return (int64_t)(a256(12 * (size_t)n) * 2 + a256(12 * (size_t)n_blocks(n)) + 256);
}

extern "C" int tm_condense_count(int64_t n, const int32_t* row_ptr, const int32_t* col_idx,
                                 const uint8_t* dirichlet, void* work, int32_t* local, int64_t* sizes,
                                 void* stream) {
  TM_REQUIRE(n > 0 && row_ptr && col_idx && dirichlet && work && local, TM_ERR_ARG, "null argument");
  TM_REQUIRE(n < (int64_t(1) << 31), TM_ERR_LIMIT, "more than 2^31 - 1 DOFs");
  hipStream_t st = (hipStream_t)stream;
  const CondWork w = cond_carve(work, n);
  const int64_t nb = n_blocks(n);
  TM_REQUIRE(nb < (int64_t(1) << 31), TM_ERR_LIMIT, "too many rows");
  hipLaunchKernelGGL(k_cond_count, dim3((unsigned)nb), dim3(NT), 0, st, n, row_ptr, col_idx, dirichlet, w);
  hipLaunchKernelGGL(k_cond_scan_blocks, dim3(1), dim3(NS), 0, st, n, w, sizes);
  hipLaunchKernelGGL(k_cond_offsets, dim3((unsigned)nb), dim3(NT), 0, st, n, dirichlet, w, local);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_condense_fill(int64_t n, const int32_t* row_ptr, const int32_t* col_idx,
                                const uint8_t* dirichlet, const void* work, const int32_t* local, int32_t* rp_ii,
                                int32_t* ci_ii, int32_t* e_ii, int32_t* rp_io, int32_t* ci_io, int32_t* e_io,
                                int32_t* inner_dof, int32_t* outer_dof, int64_t* coo_ii, int64_t* coo_io,
                                uint8_t* is_inner_edge, uint8_t* is_ou2in_edge, void* stream) {
  TM_REQUIRE(n > 0 && row_ptr && col_idx && dirichlet && work && local && rp_ii && rp_io && inner_dof &&
                 outer_dof,
             TM_ERR_ARG, "null argument");
  TM_REQUIRE((is_inner_edge == nullptr) == (is_ou2in_edge == nullptr), TM_ERR_ARG,
             "give both edge masks or neither");
  hipStream_t st = (hipStream_t)stream;
  const CondWork w = cond_carve(const_cast<void*>(work), n);
  hipLaunchKernelGGL(k_cond_fill, dim3((unsigned)n_blocks(n)), dim3(NT), 0, st, n, row_ptr, col_idx, dirichlet, w,
                     local, rp_ii, ci_ii, e_ii, rp_io, ci_io, e_io, inner_dof, outer_dof, coo_ii, coo_io,
                     is_inner_edge, is_ou2in_edge);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_condense_rhs(int64_t n_inner, const int32_t* inner_dof, const double* rhs, const int32_t* rp_io,
                               const int32_t* ci_io, const void* v_io, int v_dtype, const double* u_outer,
                               double* out, void* stream) {
  TM_REQUIRE(n_inner >= 0 && inner_dof && rhs && rp_io && out, TM_ERR_ARG, "null argument");
  TM_REQUIRE(u_outer == nullptr || (ci_io && v_io), TM_ERR_ARG, "Dirichlet values need the inner-outer block");
  TM_REQUIRE(v_dtype == TM_F32 || v_dtype == TM_F64, TM_ERR_ARG, "values must be f32 or f64");
  if (n_inner == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (v_dtype == TM_F32)
    hipLaunchKernelGGL((k_cond_rhs<float>), dim3(grid_rows(n_inner)), dim3(NT), 0, st, n_inner, inner_dof, rhs,
                       rp_io, ci_io, (const float*)v_io, u_outer, out);
  else
    hipLaunchKernelGGL((k_cond_rhs<double>), dim3(grid_rows(n_inner)), dim3(NT), 0, st, n_inner, inner_dof, rhs,
                       rp_io, ci_io, (const double*)v_io, u_outer, out);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_condense_recover(int64_t n, const uint8_t* dirichlet, const int32_t* local,
                                   const double* u_inner, const double* u_outer, double* out, void* stream) {
  TM_REQUIRE(n >= 0 && dirichlet && local && u_inner && out, TM_ERR_ARG, "null argument");
  if (n == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_cond_recover, dim3(grid_rows(n)), dim3(NT), 0, st, n, dirichlet, local, u_inner, u_outer,
                     out);
  TM_HIP(hipGetLastError());
  return TM_OK;
}
