// Batched Jacobi-PCG for many independent SPD systems in one persistent launch.
//
// The systems are stacked into one block-diagonal CSR (rows of system k are a
// contiguous range).  Rows are cut into tiles that never straddle two systems;
// each workgroup walks tiles in a grid-stride loop.  Every system keeps its own
// alpha, beta and convergence flag: the dot products are per-tile partial sums,
// and every tile of a system re-sums the partials of its system in tile order,
// so all tiles obtain bit-identical scalars without extra synchronisation.
// Three grid barriers per iteration separate q = A p, the x / r update and the
// p update.  The barrier spins with a wall-clock bound: if the grid were not
// fully resident the spin times out, an error flag is raised and every
// workgroup leaves the loop, so the kernel always terminates.  The launch is
// asynchronous (no host synchronisation) and can be captured in a hipGraph.
#include "tm_device.h"
#include "tm_spmv.h"

namespace {

using tmx::kWave;
// One workgroup of 16 waves per CU: the barrier keeps one arrival per CU while
// the memory-bound phases get enough waves to hide their gather latency.
constexpr int kThreads = 1024;

constexpr int kMaxBarBlocks = 1024;  // arrival slots of the grid barrier
constexpr int kSlot = 16;            // uint32 stride of an arrival slot (64 B: one line each)

struct BatchWork {
  double* r;
  double* p;
  double* q;
  double* dinv;
  double* pq;      // [n_tiles]
  double* rz;      // [2][n_tiles] ping-pong by iteration parity
  double* rr;      // [n_tiles]
  double* bb;      // [n_tiles]
  int* status;     // [n_mesh] 0 running, 1 converged, 2 breakdown, 3 max iterations (caller buffer)
  int* iters;      // [n_mesh] (caller buffer)
  unsigned* go;    // [kSlot]: [0] epoch released by workgroup 0
  unsigned* arrive;  // [kMaxBarBlocks][kSlot]: epoch each workgroup has reached
  int* ctl;        // [0] active systems, [1] barrier timeout flag
};

// Grid barrier.  Every wave orders its stores before the workgroup barrier
// (workgroup-scope release); one wave per workgroup then publishes them at
// agent scope (the release writes back the XCD's L2, which holds every wave's
// stores) and stores the barrier's epoch in the workgroup's own arrival slot.
// Workgroup 0 waits until every slot holds the epoch, then releases the epoch
// in `go`; every other workgroup waits for `go`.  After the wait one wave per
// workgroup drops stale cached lines (agent-scope acquire: the CU's L1 and
// the XCD's L2) before the workgroup barrier lets the other waves read.
// Measured on MI355X with 256 workgroups of 16 waves: 5.7 us per barrier,
// against 80 us for one atomic counter with fences in every wave.  Every wait
// is bounded: after 2 s the timeout flag is raised and every workgroup exits.
__device__ __forceinline__ bool timed_out(const uint64_t t0, int* ctl) {
  if (__hip_atomic_load(ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return true;
  if (wall_clock64() - t0 > 200000000ull) {  // wall_clock64 runs at 100 MHz on CDNA: 2 s
    __hip_atomic_store(ctl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
  }
  return false;
}

__device__ __forceinline__ bool grid_barrier(const BatchWork& w, unsigned nblocks, unsigned& epoch) {
  ++epoch;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(w.arrive + (size_t)blockIdx.x * kSlot, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (blockIdx.x == 0) {
    const uint64_t t0 = wall_clock64();
    for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x)
      while (__hip_atomic_load(w.arrive + (size_t)b * kSlot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch)
        if (timed_out(t0, w.ctl)) break;
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      __hip_atomic_store(w.go, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else if (threadIdx.x == 0) {
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(w.go, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch)
      if (timed_out(t0, w.ctl)) break;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return __hip_atomic_load(w.ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
}

template <int S>
__device__ __forceinline__ double sub_sum(double v) {
#pragma unroll
  for (int o = S / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ double tile_sum(const double* part, int t0, int t1) {
  double s = 0.0;
  for (int t = t0; t < t1; ++t) s += tmx::load_agent(part + t);
  return s;
}

// Sums NS partial arrays over the tiles [t0, t1) of one system with the first
// wave (fixed lane-strided order and shuffle tree, so every workgroup obtains
// the same bits) and broadcasts the results to the workgroup through LDS.
template <int NS>
__device__ __forceinline__ void system_sums(const double* const (&parts)[NS], int t0, int t1, double (&out)[NS],
                                            double* sh) {
  __syncthreads();
  if (threadIdx.x < kWave) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      double s = 0.0;
      for (int t = t0 + (int)threadIdx.x; t < t1; t += kWave) s += tmx::load_agent(parts[i] + t);
      s = tmx::wave_sum(s);
      if (threadIdx.x == 0) sh[i] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NS; ++i) out[i] = sh[i];
}

template <typename VT, int S>
__global__ __launch_bounds__(kThreads) void k_batched_pcg(int64_t n, int n_tiles, int n_mesh,
                                                         const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci, const VT* __restrict__ v,
                                                         const double* __restrict__ b, double* __restrict__ x,
                                                         const int32_t* __restrict__ tile_row0,
                                                         const int32_t* __restrict__ tile_mesh,
                                                         const int32_t* __restrict__ mesh_tile0, double rtol2,
                                                         long long maxiter, BatchWork w,
                                                         double* __restrict__ mesh_relres) {
  __shared__ double red[kThreads / kWave];
  __shared__ double sh[8];
  const unsigned G = gridDim.x;
  unsigned epoch = 0;
  const int sl = threadIdx.x % S;
  const int per = kThreads / S;

  // ---- init: x = 0, r = b, p = D^-1 r, partials (r,z) (r,r) (b,b)
  for (int t = blockIdx.x; t < n_tiles; t += G) {
    const int r0 = tile_row0[t], r1 = tile_row0[t + 1];
    double lrz = 0.0, lrr = 0.0;
    for (int i = r0 + threadIdx.x; i < r1; i += kThreads) {
      double d = 1.0;
      for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) d = (double)v[k];
      const double di = 1.0 / d;
      const double bi = b[i];
      w.dinv[i] = di;
      x[i] = 0.0;
      w.r[i] = bi;
      w.p[i] = di * bi;
      lrz += bi * di * bi;
      lrr += bi * bi;
    }
    const double s_rz = tmx::block_sum<kThreads>(lrz, red);
    const double s_rr = tmx::block_sum<kThreads>(lrr, red);
    if (threadIdx.x == 0) {
      w.rz[t] = s_rz;
      w.bb[t] = s_rr;
      w.rr[t] = s_rr;
    }
  }
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < n_mesh; k += kThreads) {
      w.status[k] = 0;
      w.iters[k] = 0;
    }
  if (!grid_barrier(w, G, epoch)) return;
  // initial convergence test (designated tile = first tile of each system)
  for (int t = blockIdx.x; t < n_tiles; t += G) {
    const int k = tile_mesh[t];
    if (t != mesh_tile0[k] || threadIdx.x != 0) continue;
    const double bbk = tile_sum(w.bb, mesh_tile0[k], mesh_tile0[k + 1]);
    const double rrk = tile_sum(w.rr, mesh_tile0[k], mesh_tile0[k + 1]);
    mesh_relres[k] = bbk > 0.0 ? sqrt(rrk / bbk) : 0.0;
    if (rrk <= rtol2 * bbk) {
      w.status[k] = 1;
      atomicSub(w.ctl, 1);
    }
  }
  if (!grid_barrier(w, G, epoch)) return;

  for (long long it = 0; it < maxiter; ++it) {
    if (__hip_atomic_load(w.ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= 0) break;
    const int cur = (int)(it & 1), nxt = cur ^ 1;
    // ---- phase A: q = A p, (p, q)
    for (int t = blockIdx.x; t < n_tiles; t += G) {
      const int k = tile_mesh[t];
      if (w.status[k] != 0) continue;
      const int r0 = tile_row0[t], r1 = tile_row0[t + 1];
      const int lim = rp[r1];
      double loc = 0.0;
      for (int i = r0 + threadIdx.x / S; i < r1; i += per) {
        double o[1] = {0.0};
        tmx::row_accum<VT, S, 1>(rp, ci, v, w.p, i, sl, lim, o);  // loads of a lane's entries issued together
        const double s = sub_sum<S>(o[0]);
        if (sl == 0) {
          w.q[i] = s;
          loc += s * w.p[i];
        }
      }
      const double tot = tmx::block_sum<kThreads>(loc, red);
      if (threadIdx.x == 0) w.pq[t] = tot;
    }
    if (!grid_barrier(w, G, epoch)) return;
    // ---- phase B: alpha, x += alpha p, r -= alpha q, (r, z), (r, r)
    for (int t = blockIdx.x; t < n_tiles; t += G) {
      const int k = tile_mesh[t];
      if (w.status[k] != 0) continue;
      const int t0 = mesh_tile0[k], t1 = mesh_tile0[k + 1];
      const double* const parts[2] = {w.pq, w.rz + cur * n_tiles};
      double sums[2];
      system_sums<2>(parts, t0, t1, sums, sh);
      const double pq = sums[0], rz = sums[1];
      const double alpha = pq > 0.0 ? rz / pq : 0.0;
      const int r0 = tile_row0[t], r1 = tile_row0[t + 1];
      double lrz = 0.0, lrr = 0.0;
      for (int i = r0 + threadIdx.x; i < r1; i += kThreads) {
        x[i] += alpha * w.p[i];
        const double ri = w.r[i] - alpha * w.q[i];
        w.r[i] = ri;
        lrz += ri * w.dinv[i] * ri;
        lrr += ri * ri;
      }
      const double s_rz = tmx::block_sum<kThreads>(lrz, red);
      const double s_rr = tmx::block_sum<kThreads>(lrr, red);
      if (threadIdx.x == 0) {
        w.rz[nxt * n_tiles + t] = s_rz;
        w.rr[t] = s_rr;
      }
    }
    if (!grid_barrier(w, G, epoch)) return;
    // ---- phase C: convergence per system, p = z + beta p
    for (int t = blockIdx.x; t < n_tiles; t += G) {
      const int k = tile_mesh[t];
      if (w.status[k] != 0) continue;
      const int t0 = mesh_tile0[k], t1 = mesh_tile0[k + 1];
      const double* const parts[5] = {w.rz + nxt * n_tiles, w.rz + cur * n_tiles, w.rr, w.bb, w.pq};
      double sums[5];
      system_sums<5>(parts, t0, t1, sums, sh);
      const double rz_new = sums[0], rz_old = sums[1], rr = sums[2], bbk = sums[3], pq = sums[4];
      const bool conv = rr <= rtol2 * bbk;
      const bool brk = !(pq > 0.0);
      const bool last = it + 1 >= maxiter;
      if (t == t0 && threadIdx.x == 0) {
        // every tile of system k reaches this same decision from the same sums,
        // so the designated tile may publish it while the others still run
        w.iters[k] = (int)(it + 1);
        mesh_relres[k] = bbk > 0.0 ? sqrt(rr / bbk) : 0.0;
        const int st = conv ? 1 : (brk ? 2 : (last ? 3 : 0));
        if (st) {
          w.status[k] = st;
          atomicSub(w.ctl, 1);
        }
      }
      if (conv || brk || last) continue;
      const double beta = rz_new / rz_old;
      const int r0 = tile_row0[t], r1 = tile_row0[t + 1];
      for (int i = r0 + threadIdx.x; i < r1; i += kThreads) w.p[i] = w.dinv[i] * w.r[i] + beta * w.p[i];
    }
    if (!grid_barrier(w, G, epoch)) return;
  }
}

// Clears the barrier state (a kernel, so a captured graph holds no memset).
__global__ void k_batched_reset(BatchWork w, int n_mesh) {
  for (int b = threadIdx.x; b < kMaxBarBlocks; b += blockDim.x) w.arrive[(size_t)b * kSlot] = 0u;
  if (threadIdx.x == 0) {
    w.go[0] = 0u;
    w.ctl[0] = n_mesh;
    w.ctl[1] = 0;
  }
}

BatchWork carve(const tm_batched_args* a) {
  const int64_t n = a->n_rows;
  const int n_tiles = a->n_tiles;
  BatchWork w;
  void* work = a->work;
  double* d = (double*)work;
  w.r = d;
  w.p = d + n;
  w.q = d + 2 * n;
  w.dinv = d + 3 * n;
  w.pq = d + 4 * n;
  w.rz = w.pq + n_tiles;
  w.rr = w.rz + 2 * (int64_t)n_tiles;
  w.bb = w.rr + n_tiles;
  w.status = a->mesh_status;
  w.iters = a->mesh_iters;
  w.go = (unsigned*)(w.bb + n_tiles);
  w.arrive = w.go + kSlot;
  w.ctl = (int*)(w.arrive + (size_t)kMaxBarBlocks * kSlot);
  return w;
}

template <typename VT, int S>
int launch_batched(const tm_batched_args* a, int blocks, hipStream_t st) {
  auto kern = k_batched_pcg<VT, S>;
  int per_cu = 0;
  TM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, 0));
  int dev = 0, ncu = 0;
  TM_HIP(hipGetDevice(&dev));
  TM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int cap = per_cu * ncu;
  TM_REQUIRE(cap > 0, TM_ERR_HIP, "occupancy query returned no resident blocks");
  if (blocks <= 0 || blocks > cap) blocks = cap;
  if (blocks > a->n_tiles) blocks = a->n_tiles;
  if (blocks > kMaxBarBlocks) blocks = kMaxBarBlocks;
  BatchWork w = carve(a);
  hipLaunchKernelGGL(k_batched_reset, dim3(1), dim3(256), 0, st, w, a->n_mesh);
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), 0, st, a->n_rows, a->n_tiles, a->n_mesh, a->row_ptr,
                     a->col_idx, (const VT*)a->values, a->b, a->x, a->tile_row0, a->tile_mesh, a->mesh_tile0,
                     a->rtol * a->rtol, (long long)a->maxiter, w, a->mesh_relres);
  TM_HIP(hipGetLastError());
  return TM_OK;
}

template <typename VT>
int dispatch_lanes(const tm_batched_args* a, int lanes, int blocks, hipStream_t st) {
  switch (lanes) {
    case 2: return launch_batched<VT, 2>(a, blocks, st);
    case 4: return launch_batched<VT, 4>(a, blocks, st);
    case 8: return launch_batched<VT, 8>(a, blocks, st);
    case 16: return launch_batched<VT, 16>(a, blocks, st);
    default: TM_REQUIRE(false, TM_ERR_ARG, "lanes per row must be 2, 4, 8 or 16 (got %d)", lanes);
  }
}

}  // namespace

extern "C" int64_t tm_batched_pcg_work_size(int64_t n_rows, int32_t n_tiles) {
  return 8 * (4 * n_rows + 5 * (int64_t)n_tiles) + 4 * (int64_t)kSlot * (kMaxBarBlocks + 1) + 64;
}

extern "C" int tm_batched_pcg(const tm_batched_args* a, int lanes, int blocks, void* stream) {
  TM_REQUIRE(a && a->row_ptr && a->col_idx && a->values && a->b && a->x && a->work && a->tile_row0 &&
                 a->tile_mesh && a->mesh_tile0 && a->mesh_iters && a->mesh_status && a->mesh_relres,
             TM_ERR_ARG, "null argument");
  TM_REQUIRE(a->n_tiles > 0 && a->n_mesh > 0 && a->maxiter > 0, TM_ERR_ARG, "empty batch");
  hipStream_t st = (hipStream_t)stream;
  if (a->value_dtype == TM_F32) return dispatch_lanes<float>(a, lanes, blocks, st);
  return dispatch_lanes<double>(a, lanes, blocks, st);
}

extern "C" int tm_batched_timeout_flag(const tm_batched_args* a, int32_t* flag, void* stream) {
  TM_REQUIRE(a && flag, TM_ERR_ARG, "null argument");
  BatchWork w = carve(a);
  TM_HIP(hipMemcpyAsync(flag, w.ctl + 1, 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return TM_OK;
}
