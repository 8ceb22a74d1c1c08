// Canonical CSR pattern of a P1 mesh built on the device
// (ElementAssembler.from_mesh pattern + P_E, tensormesh/assemble/
// element_assembler.py:838-884: the unique (row, col) pairs of every element's
// vertex pairs, rows ascending, columns ascending; P_E[e][i][j] = position of
// (conn[e][i], conn[e][j]) in that order).
//
//   1. incidence: per-node element counts (integer atomics), exclusive scan,
//      fill through per-node cursors - the order inside a node's list is
//      arbitrary, which the next step makes irrelevant;
//   2. one wave per row gathers the vertices of the row's elements into LDS
//      (the candidates), marks each value's first occurrence and ranks the
//      marked values: rank = number of distinct candidates below it.  A count
//      pass gives the row lengths, a scan the row pointers, a fill pass writes
//      each distinct column at row_ptr + rank - sorted and duplicate-free;
//   3. P_E by binary search of each element pair in its row.
// Everything is integer work, so the pattern is exact and does not depend on
// the order of the atomics.
#include "tm_device.h"

namespace {

constexpr int NT = 256;          // threads per workgroup
constexpr int WPB = NT / 64;     // rows (waves) per workgroup in the row kernels
constexpr int kMaxCand = 1024;   // candidates per row (incident elements x nb)
constexpr int NS = 1024;         // lanes of the single-workgroup scan

struct PatWork {
  int32_t* inc_ptr;  // [N + 1]
  int32_t* cursor;   // [N]
  int32_t* inc;      // [E * nb]
  int32_t* len;      // [N]
  int32_t* bsum;     // [nblk]
  int32_t* flag;     // [4]: [0] over-long row, [1] node id out of range,
                     //      [2] element pair missing from the pattern, [3] scan total
};

__host__ __device__ inline size_t a256(size_t x) { return (x + 255) & ~size_t(255); }
__host__ __device__ inline int64_t scan_blocks(int64_t n) { return (n + NT - 1) / NT; }

__host__ __device__ inline PatWork pat_carve(void* work, int64_t E, int nb, int64_t N) {
  char* p = (char*)work;
  PatWork w;
  size_t o = 0;
  w.inc_ptr = (int32_t*)(p + o);
  o = a256(o + 4 * (size_t)(N + 1));
  w.cursor = (int32_t*)(p + o);
  o = a256(o + 4 * (size_t)N);
  w.inc = (int32_t*)(p + o);
  o = a256(o + 4 * (size_t)E * nb);
  w.len = (int32_t*)(p + o);
  o = a256(o + 4 * (size_t)N);
  w.bsum = (int32_t*)(p + o);
  o = a256(o + 4 * (size_t)scan_blocks(N + 1));
  w.flag = (int32_t*)(p + o);
  return w;
}

__global__ __launch_bounds__(NT) void k_pat_zero(int64_t n, int32_t* __restrict__ a, int32_t* __restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) a[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < 4) flag[threadIdx.x] = 0;
}

__global__ __launch_bounds__(NT) void k_pat_count(int64_t E, int nb, int64_t N, const int64_t* __restrict__ conn,
                                                  int32_t* __restrict__ cnt, int32_t* __restrict__ flag) {
  for (int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x; t < E * nb; t += (int64_t)gridDim.x * NT) {
    const int64_t v = conn[t];
    if (v < 0 || v >= N) {
      atomicOr(flag + 1, 1);
      continue;
    }
    atomicAdd(cnt + v, 1);
  }
}

__global__ __launch_bounds__(NT) void k_pat_fill_inc(int64_t E, int nb, int64_t N, const int64_t* __restrict__ conn,
                                                     const int32_t* __restrict__ inc_ptr,
                                                     int32_t* __restrict__ cursor, int32_t* __restrict__ inc) {
  for (int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x; t < E * nb; t += (int64_t)gridDim.x * NT) {
    const int64_t v = conn[t];
    if (v < 0 || v >= N) continue;
    const int32_t pos = atomicAdd(cursor + v, 1);
    inc[inc_ptr[v] + pos] = (int32_t)(t / nb);
  }
}

// Exclusive scan of a[0, n) into out[0, n] (out[n] = total): per-workgroup
// sums, one workgroup scanning them, per-workgroup scans with the offsets.
__global__ __launch_bounds__(NT) void k_scan_sums(int64_t n, const int32_t* __restrict__ a,
                                                  int32_t* __restrict__ bsum) {
  __shared__ int32_t red[NT / 64];
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  int32_t v = i < n ? a[i] : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) s += red[k];
    bsum[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(NS) void k_scan_block_sums(int64_t nblk, int32_t* __restrict__ bsum,
                                                        int32_t* __restrict__ total) {
  __shared__ int64_t part[NS];
  const int64_t per = (nblk + NS - 1) / NS;
  const int64_t b0 = threadIdx.x * per, b1 = b0 + per < nblk ? b0 + per : nblk;
  int64_t s = 0;
  for (int64_t b = b0; b < b1; ++b) s += bsum[b];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < NS; o <<= 1) {
    const int64_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = part[threadIdx.x] - s;
  for (int64_t b = b0; b < b1; ++b) {
    const int32_t v = bsum[b];
    bsum[b] = (int32_t)run;
    run += v;
  }
  if (threadIdx.x == NS - 1 && total) *total = part[NS - 1] < (int64_t(1) << 31) ? (int32_t)part[NS - 1] : -1;
}

__global__ __launch_bounds__(NT) void k_scan_apply(int64_t n, const int32_t* __restrict__ a,
                                                   const int32_t* __restrict__ bsum, int32_t* __restrict__ out,
                                                   const int32_t* __restrict__ total) {
  __shared__ int32_t s[NT];
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  s[threadIdx.x] = i < n ? a[i] : 0;
  __syncthreads();
  for (int o = 1; o < NT; o <<= 1) {
    const int32_t v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  if (i < n) out[i] = bsum[blockIdx.x] + s[threadIdx.x] - a[i];
  if (i == 0) out[n] = *total;
}

// One wave per row: candidates in LDS, first-occurrence marks, ranks.
// FILL = false: row lengths; FILL = true: columns at row_ptr + rank.
template <bool FILL>
__global__ __launch_bounds__(NT) void k_pat_rows(int64_t N, int nb, const int64_t* __restrict__ conn,
                                                 PatWork w, const int32_t* __restrict__ row_ptr,
                                                 int32_t* __restrict__ col_idx) {
  __shared__ int32_t cand[WPB][kMaxCand];
  __shared__ uint8_t first[WPB][kMaxCand];
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  // every wave runs the same number of loop trips (rows past N are empty), so
  // the barriers below are reached by the whole workgroup
  for (int64_t base = (int64_t)blockIdx.x * WPB; base < N; base += (int64_t)gridDim.x * WPB) {
    const int64_t row = base + wv;
    int K = 0;
    int32_t p0 = 0;
    if (row < N) {
      p0 = w.inc_ptr[row];
      K = (w.inc_ptr[row + 1] - p0) * nb;
      if (K > kMaxCand) {
        if (lane == 0) atomicOr(w.flag, 1);
        K = 0;
      }
    }
    for (int t = lane; t < K; t += 64) cand[wv][t] = (int32_t)conn[(int64_t)w.inc[p0 + t / nb] * nb + t % nb];
    __syncthreads();
    for (int t = lane; t < K; t += 64) {
      const int32_t v = cand[wv][t];
      bool f = true;
      for (int j = 0; j < t; ++j) f = f && cand[wv][j] != v;
      first[wv][t] = f ? 1 : 0;
    }
    __syncthreads();
    int cnt = 0;
    for (int t = lane; t < K; t += 64) {
      if (!first[wv][t]) continue;
      ++cnt;
      if (FILL) {
        const int32_t v = cand[wv][t];
        int rank = 0;
        for (int j = 0; j < K; ++j) rank += (first[wv][j] && cand[wv][j] < v) ? 1 : 0;
        col_idx[row_ptr[row] + rank] = v;
      }
    }
    if (!FILL) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
      if (lane == 0 && row < N) w.len[row] = cnt;
    }
    __syncthreads();
  }
}

// slots[(e nb + i) nb + j] = position of conn[e][j] in row conn[e][i]
__global__ __launch_bounds__(NT) void k_pat_slots(int64_t E, int nb, int64_t N, const int64_t* __restrict__ conn,
                                                  const int32_t* __restrict__ row_ptr,
                                                  const int32_t* __restrict__ col_idx, int64_t* __restrict__ slots,
                                                  int32_t* __restrict__ flag) {
  const int64_t tot = E * nb * nb;
  for (int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x; t < tot; t += (int64_t)gridDim.x * NT) {
    const int64_t e = t / (nb * nb);
    const int i = (int)(t / nb % nb), j = (int)(t % nb);
    const int64_t r = conn[e * nb + i], c = conn[e * nb + j];
    if (r < 0 || r >= N) {
      atomicOr(flag + 1, 1);
      slots[t] = 0;
      continue;
    }
    int32_t lo = row_ptr[r], hi = row_ptr[r + 1];
    while (lo < hi) {
      const int32_t mid = lo + (hi - lo) / 2;
      if (col_idx[mid] < c)
        lo = mid + 1;
      else
        hi = mid;
    }
    const bool found = lo < row_ptr[r + 1] && col_idx[lo] == c;
    if (!found) atomicOr(flag + 2, 1);
    slots[t] = found ? (int64_t)lo : 0;
  }
}

int grid_for(int64_t n) {
  const int64_t g = (n + NT - 1) / NT;
  return (int)(g < 16384 ? (g > 0 ? g : 1) : 16384);
}

int scan(int64_t n, const int32_t* a, int32_t* bsum, int32_t* out, int32_t* total, hipStream_t st) {
  const int64_t nb = scan_blocks(n);
  TM_REQUIRE(nb < (int64_t(1) << 31), TM_ERR_LIMIT, "scan too long");
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(NT), 0, st, n, a, bsum);
  hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(NS), 0, st, nb, bsum, total);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(NT), 0, st, n, a, bsum, out, total);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

}  // namespace

extern "C" int64_t tm_pattern_dev_work_size(int64_t E, int nb, int64_t N) {
  if (E < 0 || nb < 1 || nb > 4 || N < 1) return -1;
  return (int64_t)(a256(4 * (size_t)(N + 1)) + a256(4 * (size_t)N) + a256(4 * (size_t)E * nb) +
                   a256(4 * (size_t)N) + a256(4 * (size_t)scan_blocks(N + 1)) + 256);
}

extern "C" int tm_pattern_dev_rows(const int64_t* conn, int64_t E, int nb, int64_t N, void* work,
                                   int32_t* row_ptr, void* stream) {
  TM_REQUIRE(conn && work && row_ptr && E >= 0 && nb >= 1 && nb <= 4 && N >= 1, TM_ERR_ARG, "bad arguments");
  TM_REQUIRE(E * nb < (int64_t(1) << 31) && N < (int64_t(1) << 31), TM_ERR_LIMIT, "mesh too large for int32");
  hipStream_t st = (hipStream_t)stream;
  PatWork w = pat_carve(work, E, nb, N);
  hipLaunchKernelGGL(k_pat_zero, dim3(grid_for(N)), dim3(NT), 0, st, N, w.len, w.flag);
  hipLaunchKernelGGL(k_pat_zero, dim3(grid_for(N)), dim3(NT), 0, st, N, w.cursor, w.flag);
  hipLaunchKernelGGL(k_pat_count, dim3(grid_for(E * nb)), dim3(NT), 0, st, E, nb, N, conn, w.len, w.flag);
  TM_HIP(hipGetLastError());
  // incidence pointers from the counts (w.len holds the counts here)
  int rc = scan(N, w.len, w.bsum, w.inc_ptr, w.flag + 3, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pat_fill_inc, dim3(grid_for(E * nb)), dim3(NT), 0, st, E, nb, N, conn, w.inc_ptr, w.cursor,
                     w.inc);
  const int64_t g = (N + WPB - 1) / WPB;
  hipLaunchKernelGGL((k_pat_rows<false>), dim3((unsigned)(g < 65536 ? g : 65536)), dim3(NT), 0, st, N, nb, conn, w,
                     nullptr, nullptr);
  TM_HIP(hipGetLastError());
  rc = scan(N, w.len, w.bsum, row_ptr, w.flag + 3, st);
  if (rc) return rc;
  int32_t flag[4];
  TM_HIP(hipMemcpyAsync(flag, w.flag, sizeof(flag), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  TM_REQUIRE(!flag[1], TM_ERR_ARG, "connectivity holds node ids outside [0, %lld)", (long long)N);
  TM_REQUIRE(!flag[0], TM_ERR_LIMIT, "a node has more than %d element-vertex candidates", kMaxCand);
  TM_REQUIRE(flag[3] >= 0, TM_ERR_LIMIT, "nnz exceeds int32 range");
  return TM_OK;
}

extern "C" int tm_pattern_dev_cols(const int64_t* conn, int64_t E, int nb, int64_t N, const void* work,
                                   const int32_t* row_ptr, int32_t* col_idx, void* stream) {
  TM_REQUIRE(conn && work && row_ptr && col_idx && E >= 0 && nb >= 1 && nb <= 4 && N >= 1, TM_ERR_ARG,
             "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  PatWork w = pat_carve(const_cast<void*>(work), E, nb, N);
  const int64_t g = (N + WPB - 1) / WPB;
  hipLaunchKernelGGL((k_pat_rows<true>), dim3((unsigned)(g < 65536 ? g : 65536)), dim3(NT), 0, st, N, nb, conn, w,
                     row_ptr, col_idx);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

extern "C" int tm_element_slots_dev(const int64_t* conn, int64_t E, int nb, const int32_t* row_ptr,
                                    const int32_t* col_idx, void* work, int64_t N, int64_t* slots, void* stream) {
  TM_REQUIRE(conn && row_ptr && col_idx && slots && work && E >= 0 && nb >= 1 && nb <= 4 && N >= 1, TM_ERR_ARG,
             "bad arguments");
  if (E == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  PatWork w = pat_carve(work, E, nb, N);
  hipLaunchKernelGGL(k_pat_zero, dim3(1), dim3(NT), 0, st, (int64_t)0, w.len, w.flag);  // clears the flags
  hipLaunchKernelGGL(k_pat_slots, dim3(grid_for(E * nb * nb)), dim3(NT), 0, st, E, nb, N, conn, row_ptr, col_idx,
                     slots, w.flag);
  TM_HIP(hipGetLastError());
  int32_t flag[4];
  TM_HIP(hipMemcpyAsync(flag, w.flag, sizeof(flag), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  TM_REQUIRE(!flag[1], TM_ERR_ARG, "connectivity holds node ids outside [0, %lld)", (long long)N);
  TM_REQUIRE(!flag[2], TM_ERR_ARG, "pattern does not contain every element pair");
  return TM_OK;
}
