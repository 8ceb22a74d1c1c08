// P1 owner-computes assembly kernels for gfx950.
//
// One workgroup owns one plan block: a compact set of CSR rows (mesh nodes)
// plus every element touching them.  The workgroup
//   1. gathers the coordinates / nodal coefficient of the block's halo nodes
//      into LDS and zeroes an FP64 accumulator holding the owned rows' CSR
//      entries,
//   2. walks the block's element records (one element per lane, 16-byte
//      records carrying 10-bit LDS node slots and the 6-bit in-row rank of
//      every (owned row, column) pair), computes the element geometry in FP64
//      (cofactor vectors c_b = |J| grad phi_b and det J) and adds the owned
//      rows of the element matrix into LDS with ds_add_f64,
//   3. writes each owned row once, in runs of consecutive rows whose CSR
//      ranges are contiguous (coalesced stores, optionally narrowed to FP32).
// No global atomics, no sort, every CSR entry written exactly once.
//
// Symmetric scalar forms (Laplace, kappa-Laplace with an FP32 coefficient,
// mass) read the plan's edge records instead when it has them: 8 bytes (12
// for rows of more than 16 entries) carrying one rank per vertex pair, one
// LDS atomic per element edge, and a per-block mirror list that fills the
// transposed entries of edges whose two ends are both owned.
#include <type_traits>

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

// LDS regions of one block, identical in every kernel (nh node slots; slots
// [0, nown) hold the owned nodes):
//   acc   [nacc]   owned-row CSR accumulators (FP64)
//   p0    [nh]     node words {x, y}
//   p1    [nh]     node words 3D {z, coef}, 2D {coef, -}
//   ms    [nh]     row meta of owned slots: acc offset | diag rank << 16 |
//                  row length << 24 (0 for the other slots)
//   accb  [nown]   owned-node vector accumulator (FP64)
// Node data are read as 16-byte words: a quarter-wave's reads are
// conflict-free when the slots differ mod 16, which the plan's slot
// assignment aims for.
struct Lds {
  double* acc;
  double2* p0;
  double2* p1;
  uint32_t* ms;
  double* accb;
};

__device__ __forceinline__ Lds carve(char* s, int nacc, int nh, int nown) {
  Lds L;
  size_t o = 0;
  L.acc = (double*)(s + o);
  o = a16(o + 8 * (size_t)nacc);
  L.p0 = (double2*)(s + o);
  o = a16(o + 16 * (size_t)nh);
  L.p1 = (double2*)(s + o);
  o = a16(o + 16 * (size_t)nh);
  L.ms = (uint32_t*)(s + o);
  o = a16(o + 4 * (size_t)nh);
  L.accb = (double*)(s + o);
  return L;
}

template <typename T>
__device__ __forceinline__ double ld(const T* p, int64_t i) {
  return (double)p[i];
}

__device__ __forceinline__ int rec_local(uint64_t lo, int a) { return (int)((lo >> (10 * a)) & 1023u); }

// In-row rank of vertex b in the row of owned vertex a (a != b): the 6-bit
// field at bit 40 + 6 (a (NB-1) + j) of the record (a, b are unrolled
// constants, so each field is one bit-field extract, or an align + extract
// for the field that straddles two words).
template <int NB>
__device__ __forceinline__ int rec_rank(const uint4& rc, int a, int b) {
  const int j = b - (b > a ? 1 : 0);
  const int bit = 40 + 6 * (a * (NB - 1) + j);
  const int w = bit / 32, off = bit % 32;
  const uint32_t lo = w == 1 ? rc.y : (w == 2 ? rc.z : rc.w);
  if (off <= 26) return (int)__builtin_amdgcn_ubfe(lo, off, 6);
  const uint32_t hi = w == 1 ? rc.z : rc.w;
  return (int)__builtin_amdgcn_ubfe(__builtin_amdgcn_alignbit(hi, lo, off), 0, 6);
}

template <int DIM>
__device__ __forceinline__ void node_xyz(const double2& n0, const double2& n1, double (&x)[DIM]);
template <>
__device__ __forceinline__ void node_xyz<3>(const double2& n0, const double2& n1, double (&x)[3]) {
  x[0] = n0.x;
  x[1] = n0.y;
  x[2] = n1.x;
}
template <>
__device__ __forceinline__ void node_xyz<2>(const double2& n0, const double2&, double (&x)[2]) {
  x[0] = n0.x;
  x[1] = n0.y;
}
template <int DIM>
__device__ __forceinline__ double node_coef(const double2& n1) {
  return DIM == 3 ? n1.y : n1.x;
}
// LDS address of the per-node FP64 scratch word (the coefficient word)
template <int DIM>
__device__ __forceinline__ double* node_coef_ptr(const Lds& L, int h) {
  return (double*)(L.p1 + h) + (DIM == 3 ? 1 : 0);
}

// Packed node word of the matrix kernels whose coefficient is FP32 or absent:
// the second double of p1 holds the coefficient as FP32 (low word) and the
// row meta (high word), so one 16-byte read brings both.
__device__ __forceinline__ double pack_coef_meta(float cf, uint32_t m) {
  return __hiloint2double((int)m, __float_as_int(cf));
}
__device__ __forceinline__ uint32_t packed_meta(const double2& n1) { return (uint32_t)__double2hiint(n1.y); }
__device__ __forceinline__ double packed_coef(const double2& n1) { return (double)__int_as_float(__double2loint(n1.y)); }
template <bool PK>
__device__ __forceinline__ uint32_t row_meta(const Lds& L, int h) {
  return PK ? ((const uint32_t*)(L.p1 + h))[3] : L.ms[h];
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
template <int STEPS = 2>
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int i = 0; i < STEPS; ++i) r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

__device__ __forceinline__ int meta_off(uint32_t m) { return (int)(m & 0xffffu); }
__device__ __forceinline__ int meta_diag(uint32_t m) { return (int)((m >> 16) & 0xffu); }
__device__ __forceinline__ int meta_len(uint32_t m) { return (int)(m >> 24); }

// Stages the block's halo nodes (and optionally a nodal coefficient) into LDS
// and zeroes the accumulators.  A slot with halo id -1 is empty.
// PK: packed node words (pack_coef_meta; CT float or no coefficient), L.ms unused.
template <int NT, int DIM, typename CT, bool COEF, bool PK = false>
__device__ __forceinline__ void stage(const Lds& L, const int32_t* __restrict__ halo, const uint32_t* __restrict__ meta,
                                      int nacc, int nh, int nown, const double* __restrict__ X,
                                      const CT* __restrict__ coef) {
  // A lane stages nodes tid, tid + NT, ...: the first K of them with all halo
  // ids loaded first, then all node gathers, so a block waits for two
  // dependent memory round trips, not two per NT nodes.  The loads are
  // unconditional (indices clamped into the block's halo range, empty slots
  // pointed at node 0, then masked): a branch between them would make the
  // compiler wait for each before the next is issued.  K covers 512 nodes;
  // larger halos (rare: at most 1024) stage the rest node by node.
  constexpr int K = (512 + NT - 1) / NT;
  int32_t g[K];
  uint32_t mh[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int h = min((int)threadIdx.x + k * NT, nh - 1);
    g[k] = halo[h];
    mh[k] = meta[h];
  }
  double2 n0[K], n1[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool ok = g[k] >= 0;
    const int64_t gi = ok ? g[k] : 0;
    const double cf = COEF ? ld(coef, gi) : 0.0;
    if (PK) {  // an empty slot keeps node 0's data: no record or row reads it
      const double pc = pack_coef_meta((float)cf, mh[k]);
      if (DIM == 3) {
        n0[k] = make_double2(X[gi * 3], X[gi * 3 + 1]);
        n1[k] = make_double2(X[gi * 3 + 2], pc);
      } else {
        n0[k] = make_double2(X[gi * 2], X[gi * 2 + 1]);
        n1[k] = make_double2(0.0, pc);
      }
    } else if (DIM == 3) {
      const double x = X[gi * 3], y = X[gi * 3 + 1], z = X[gi * 3 + 2];
      n0[k] = make_double2(ok ? x : 0.0, ok ? y : 0.0);
      n1[k] = make_double2(ok ? z : 0.0, ok ? cf : 0.0);
    } else {
      const double x = X[gi * 2], y = X[gi * 2 + 1];
      n0[k] = make_double2(ok ? x : 0.0, ok ? y : 0.0);
      n1[k] = make_double2(ok ? cf : 0.0, 0.0);
    }
  }
  // 16-byte zero stores (L.acc is 16-byte aligned; an odd last entry alone)
  for (int i = threadIdx.x; i < nacc / 2; i += NT) ((double2*)L.acc)[i] = make_double2(0.0, 0.0);
  if ((nacc & 1) && threadIdx.x == 0) L.acc[nacc - 1] = 0.0;
  for (int a = threadIdx.x; a < nown; a += NT) L.accb[a] = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int h = threadIdx.x + k * NT;
    if (h < nh) {
      L.p0[h] = n0[k];
      L.p1[h] = n1[k];
      if (!PK) L.ms[h] = mh[k];
    }
  }
  for (int h = threadIdx.x + K * NT; h < nh; h += NT) {
    const int64_t gh = halo[h];
    double2 a = make_double2(0.0, 0.0), c = make_double2(0.0, 0.0);
    double cf = 0.0;
    if (gh >= 0) {
      cf = COEF ? ld(coef, gh) : 0.0;
      if (DIM == 3) {
        a = make_double2(X[gh * 3], X[gh * 3 + 1]);
        c = make_double2(X[gh * 3 + 2], cf);
      } else {
        a = make_double2(X[gh * 2], X[gh * 2 + 1]);
        c = make_double2(cf, 0.0);
      }
    }
    if (PK) c = make_double2(DIM == 3 ? c.x : 0.0, pack_coef_meta((float)cf, meta[h]));
    L.p0[h] = a;
    L.p1[h] = c;
    if (!PK) L.ms[h] = meta[h];
  }
}

// Sets the diagonal of every owned row to minus the sum of its off-diagonal
// entries (exact for forms whose element rows sum to zero: sum_b c_b = 0).
// A row's entries are read 16 at a time, all reads issued before the first
// add (one LDS round trip per 16 entries instead of one per entry); the adds
// run in entry order, so the sum is the same as a plain loop's.
template <int NT, bool PK = false>
__device__ __forceinline__ void rowsum_diagonal(const Lds& L, int nown) {
  constexpr int C = 16;
  for (int a = threadIdx.x; a < nown; a += NT) {
    const uint32_t m = row_meta<PK>(L, a);
    const int b0 = meta_off(m), len = meta_len(m);
    if (len == 0) continue;
    const int d = meta_diag(m);
    double s = 0.0;
    for (int k0 = 0; k0 < len; k0 += C) {
      double v[C];
#pragma unroll
      for (int k = 0; k < C; ++k) v[k] = L.acc[b0 + min(k0 + k, len - 1)];
#pragma unroll
      for (int k = 0; k < C; ++k) s += (k0 + k < len && k0 + k != d) ? v[k] : 0.0;
    }
    L.acc[b0 + d] = -s;
  }
}

// How an element-matrix entry of owned row `row` (its LDS slot) reaches its
// LDS accumulator.
struct AddEmit {  // ds_add_f64 (the order of the adds varies from run to run)
  __device__ __forceinline__ void operator()(const Lds&, int, double* a, double v) const { atomicAdd(a, v); }
};
// Deterministic mode.  Every entry that an accumulator of row a receives is
// rounded to a multiple of 2^k_a, with 2^(k_a + 44) above the largest |entry|
// M_a of that row; sums of up to 512 such terms are integers times 2^k_a below
// 2^(k_a + 53), so every FP64 partial sum is exact and the result does not
// depend on the order of the LDS atomics.  An entry moves by at most
// 2^(k_a - 1) <= 2^-44 M_a.  A first pass over the block's records (same
// arithmetic) finds M_a; the per-row word L.accb (unused by the matrix
// kernels) holds it, then k_a.
struct MaxEmit {  // first pass: row maxima (non-negative doubles order like their bit patterns)
  __device__ __forceinline__ void operator()(const Lds& L, int row, double*, double v) const {
    atomicMax((unsigned long long*)(L.accb + row), (unsigned long long)__double_as_longlong(fabs(v)));
  }
};
struct GridEmit {  // second pass: entries rounded to the row's grid
  __device__ __forceinline__ void operator()(const Lds& L, int row, double* a, double v) const {
    const int k = *(const int*)(L.accb + row);
    atomicAdd(a, ldexp(rint(ldexp(v, -k)), k));
  }
};

// Row maxima -> grid exponents k_a (in place), between the two passes.
template <int NT>
__device__ __forceinline__ void row_grids(const Lds& L, int nown) {
  for (int a = threadIdx.x; a < nown; a += NT) {
    const unsigned long long bits = *(const unsigned long long*)(L.accb + a);
    const int e = (int)((bits >> 52) & 0x7ffull) - 1023;  // M_a in [2^e, 2^(e+1))
    *(int*)(L.accb + a) = bits ? e + 1 - 44 : 0;
  }
}

// Load vectors, deterministic mode: the row maxima (bit patterns in L.accb,
// first pass) become grid exponents in L.ms, which no load kernel reads
// otherwise, and L.accb is cleared for the second pass.
template <int NT>
__device__ __forceinline__ void load_grids(const Lds& L, int nown) {
  for (int a = threadIdx.x; a < nown; a += NT) {
    const unsigned long long bits = *(const unsigned long long*)(L.accb + a);
    const int e = (int)((bits >> 52) & 0x7ffull) - 1023;
    L.ms[a] = (uint32_t)(bits ? e + 1 - 44 : 0);
    L.accb[a] = 0.0;
  }
}

// Row stores of a block: wave w writes runs w, w + NT/64, ... of consecutive
// CSR values from the LDS accumulators.  Lane j of wave w fetches the
// descriptor of the wave's j-th run (run_descriptors, one vector load issued
// ahead of use); the store loop takes it from that lane (readlane: a
// wave-uniform value), so no run waits for a load of its own.  Runs beyond
// the wave's 64th (blocks of more than 64 NT/64 runs) load their descriptor
// in the loop.
template <int NT>
__device__ __forceinline__ int4 run_descriptors(const PlanArgs& p, int runoff, int nrun) {
  const int k = (int)(threadIdx.x / kWave) + (NT / kWave) * (int)(threadIdx.x % kWave);
  return k < nrun ? p.runs[runoff + k] : make_int4(0, 0, 0, 0);
}
template <int NT, typename OT>
__device__ __forceinline__ void store_runs(const Lds& L, const PlanArgs& p, int runoff, int nrun, const int4& dsc,
                                           OT* __restrict__ values) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int j = 0;; ++j) {
    const int k = w + (NT / kWave) * j;
    if (k >= nrun) break;
    int4 ru;
    if (j < kWave) {
      ru.x = __builtin_amdgcn_readlane(dsc.x, j);
      ru.y = __builtin_amdgcn_readlane(dsc.y, j);
      ru.z = __builtin_amdgcn_readlane(dsc.z, j);
    } else {
      ru = p.runs[runoff + k];
    }
    for (int t = lane; t < ru.z; t += kWave) values[(int64_t)ru.x + t] = (OT)L.acc[ru.y + t];
  }
}

__device__ __forceinline__ uint4 first_record(const PlanArgs& p, int roff, int nrec, int off) {
  const int r = (int)threadIdx.x + off;
  return r < nrec ? p.recs[roff + r] : make_uint4(0, 0, 0, 0);
}

// A block's records read through a raw buffer resource whose range ends after
// the block's last record: a load beyond it returns zeros (the buffer unit's
// range check), so loads ahead need no branch, and a record's address is a
// 32-bit offset from a wave-uniform base.
constexpr int kRawBufferWord3 = 0x00020000;  // gfx9 descriptor word 3: 32-bit data format, no swizzle
struct RecBuf {
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ RecBuf(const PlanArgs& p, int roff, int nrec)
      : rs(__builtin_amdgcn_make_buffer_rsrc((void*)(p.recs + roff), 0, nrec * 16, kRawBufferWord3)) {}
  __device__ __forceinline__ uint4 operator[](int r) const {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, r * 16, 0, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
  }
};

// Owned rows of one element matrix when record vertices [0, K) are owned and
// the others are not (records list their owned vertices first): each pair
// (a, b), a < b, with a owned is formed once and added to row a, and to row b
// when b is owned too.
template <int NB, int K, int FORM, typename Emit>
__device__ __forceinline__ void emit_owned(const Lds& L, const uint4& rc, const int (&l)[NB], const uint32_t (&m)[NB],
                                           const double (&c)[NB][NB - 1], double s, const QuadArg& q,
                                           const Emit& em) {
  constexpr int DIM = NB - 1;
  if (FORM == 2) {
#pragma unroll
    for (int a = 0; a < K; ++a) em(L, l[a], L.acc + meta_off(m[a]) + meta_diag(m[a]), s * q.M[a * NB + a]);
  }
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int bb = a + 1; bb < NB; ++bb) {
      const double kab = FORM == 2 ? s * q.M[a * NB + bb] : s * dotc<DIM>(c[a], c[bb]);
      em(L, l[a], L.acc + meta_off(m[a]) + rec_rank<NB>(rc, a, bb), kab);
      if (bb < K) em(L, l[bb], L.acc + meta_off(m[bb]) + rec_rank<NB>(rc, bb, a), kab);
    }
}

// The element loop of one block: each lane evaluates records r = tid, tid + NT,
// ... (nxt = the lane's first record, already loaded) and adds the owned rows
// of each element matrix into L.acc through `em` (ds_add_f64).
template <int NB, int FORM, int NT, bool PK, typename Emit>
__device__ __forceinline__ void asm_records(const Lds& L, const RecBuf& rb, int nrec, int nown, uint4 nxt,
                                            uint4 nxt2, double scale, const QuadArg& q, const Emit& em) {
  constexpr int DIM = NB - 1;
  constexpr bool ROWSUM = FORM != 2;
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = nxt;
    nxt = nxt2;
    nxt2 = rb[r + 2 * NT];  // records two iterations ahead (zeros past the block's end)
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    int l[NB];
    double x[NB][DIM], cf[NB];
    uint32_t m[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      const double2 n0 = L.p0[l[a]], n1 = L.p1[l[a]];
      m[a] = PK ? packed_meta(n1) : L.ms[l[a]];
      node_xyz<DIM>(n0, n1, x[a]);
      cf[a] = PK ? packed_coef(n1) : node_coef<DIM>(n1);
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
    int nk = 0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      own[a] = l[a] < nown;
      nk += own[a] ? 1 : 0;
    }
    // The plan groups a block's records by owned count, so most waves share
    // one count: they run a branch-free instance that forms only the pairs
    // with an owned vertex.
    const int k0 = __builtin_amdgcn_readfirstlane(nk);
    if (__all(nk == k0)) {
      if (k0 == NB) {
        emit_owned<NB, NB, FORM>(L, rc, l, m, c, s, q, em);
      } else if (k0 == 1) {
        emit_owned<NB, 1, FORM>(L, rc, l, m, c, s, q, em);
      } else if (k0 == 2) {
        emit_owned<NB, 2, FORM>(L, rc, l, m, c, s, q, em);
      } else if (NB == 4 && k0 == 3) {
        emit_owned<NB, (NB == 4 ? 3 : 1), FORM>(L, rc, l, m, c, s, q, em);
      }
      continue;
    }
    if (!ROWSUM) {
#pragma unroll
      for (int a = 0; a < NB; ++a)
        if (own[a]) em(L, l[a], L.acc + meta_off(m[a]) + meta_diag(m[a]), s * q.M[a * NB + a]);
    }
    // each symmetric pair once: K_ab = K_ba
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int bb = a + 1; bb < NB; ++bb) {
        if (!(own[a] || own[bb])) continue;
        const double kab = FORM == 2 ? s * q.M[a * NB + bb] : s * dotc<DIM>(c[a], c[bb]);
        if (own[a]) em(L, l[a], L.acc + meta_off(m[a]) + rec_rank<NB>(rc, a, bb), kab);
        if (own[bb]) em(L, l[bb], L.acc + meta_off(m[bb]) + rec_rank<NB>(rc, bb, a), kab);
      }
  }
}

// FORM: 0 laplace, 1 kappa-laplace (nodal coef), 2 mass; all multiplied by scale.
// DET: deterministic mode (entries rounded to per-row grids, see MaxEmit / GridEmit).
// At most 64 VGPRs (8 waves per SIMD): eight 20 KiB blocks per CU overlap their
// phases only if all of their waves are resident.
template <int NB, int FORM, typename CT, typename OT, int NT, bool DET>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_asm_matrix(PlanArgs p, const double* __restrict__ X,
                                                  const CT* __restrict__ coef, double scale, QuadArg q,
                                                  OT* __restrict__ values) {
  constexpr int DIM = NB - 1;
  constexpr bool ROWSUM = FORM != 2;  // diagonal from the zero row sum
  constexpr bool PK = FORM != 1 || std::is_same<CT, float>::value;  // packed node words
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, runoff = bi1.y, nrun = bi1.z, nacc = bi1.w;
  const Lds L = carve(smem, nacc, nh, nown);
  // a lane's first two records are independent of the staged nodes: load them first
  const RecBuf rb(p, roff, nrec);
  const int t = threadIdx.x;
  uint4 nxt = rb[t], nxt2 = rb[t + NT];
  stage<NT, DIM, CT, FORM == 1, PK>(L, p.halo + hoff, p.meta + hoff, nacc, nh, nown, X, coef);
  __syncthreads();

  if (DET) {
    asm_records<NB, FORM, NT, PK>(L, rb, nrec, nown, nxt, nxt2, scale, q, MaxEmit{});
    __syncthreads();
    row_grids<NT>(L, nown);
    __syncthreads();
    asm_records<NB, FORM, NT, PK>(L, rb, nrec, nown, rb[t], rb[t + NT], scale, q, GridEmit{});
  } else {
    asm_records<NB, FORM, NT, PK>(L, rb, nrec, nown, nxt, nxt2, scale, q, AddEmit{});
  }
  const int4 dsc = run_descriptors<NT>(p, runoff, nrun);  // in flight during the row sums
  __syncthreads();
  if (ROWSUM) {
    rowsum_diagonal<NT, PK>(L, nown);
    __syncthreads();
  }
  store_runs<NT>(L, p, runoff, nrun, dsc, values);
}

// ---- edge records (symmetric scalar forms) ---------------------------------
//
// A record of W 32-bit words: NB LDS node words (owned vertices first, by
// ascending slot), then one rank per vertex pair (i, j), i < j, in the order
// (0,1) (0,2) (0,3) (1,2) (1,3) (2,3): the rank of vertex j in the row of
// vertex i, which accumulates the pair's value when it is owned.  Tetrahedra:
// 10-bit node words and 4-bit ranks from bit 40 (W = 2), or 11-bit words and
// 6-bit ranks at word 1 bits 12, 18, 24 and word 2 bits 0, 6, 12 (W = 3);
// triangles: 11-bit words, 6-bit ranks at word 1 bits 12, 18, 24 (W = 2).  An
// edge with both ends owned is accumulated once, in the row of the smaller
// slot; the transposed entry is copied from it after the element loop (the
// block's mirror list, {src | dst << 16} accumulator indices).
//
// LDS of a block: the accumulators, then the nodes in groups of 16 slots,
// 512 bytes each: 16 first words {x, y} then 16 second words {z | coef,
// packed coefficient and row meta}.  Slot s sits at node word w(s) = s +
// (s & ~15) (its second word at w(s) + 16), which is what records carry: one
// address serves both reads, and a quarter-wave's reads of slots that differ
// mod 16 stay conflict-free.  The row meta's offset field holds the row's byte
// offset in the accumulators (accumulators < 8192).  The block's records come
// grouped by descending owned count; eblk gives the group ends.  After the
// nodes, the block's row table {first value index, byte offset | length << 16}
// in accumulator order drives the stores (16 lanes per row).
struct EdgeArgs {
  const uint32_t* recs;
  const int32_t* eblk;
  const uint32_t* mir;
  const int2* rows;
};

__device__ __forceinline__ int node_word(int s) { return s + (s & ~15); }

template <int W>
struct ERec {
  uint32_t w[W];
};

template <int W>
struct ERecBuf {
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ ERecBuf(const uint32_t* recs, int roff, int nrec)
      : rs(__builtin_amdgcn_make_buffer_rsrc((void*)(recs + (int64_t)roff * W), 0, nrec * W * 4,
                                             kRawBufferWord3)) {}
  __device__ __forceinline__ ERec<W> operator[](int r) const {
    ERec<W> e;
    if constexpr (W == 2) {
      const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, r * 8, 0, 0);
      e.w[0] = v[0];
      e.w[1] = v[1];
    } else {
      const auto v = __builtin_amdgcn_raw_buffer_load_b96(rs, r * 12, 0, 0);
      e.w[0] = v[0];
      e.w[1] = v[1];
      e.w[2] = v[2];
    }
    return e;
  }
};

// LDS node word of record vertex a
template <int NB, int W>
__device__ __forceinline__ uint32_t erec_word(const ERec<W>& e, int a) {
  constexpr int SB = (NB == 4 && W == 2) ? 10 : 11;
  const int bit = SB * a;
  if (bit + SB <= 32) return __builtin_amdgcn_ubfe(e.w[0], bit, SB);
  if (bit >= 32) return __builtin_amdgcn_ubfe(e.w[1], bit - 32, SB);
  return __builtin_amdgcn_alignbit(e.w[1], e.w[0], bit) & ((1u << SB) - 1);
}

// byte offset (rank * 8) of pair k's rank field
template <int NB, int W>
__device__ __forceinline__ uint32_t erec_rank8(const ERec<W>& e, int k) {
  constexpr bool R4 = NB == 4 && W == 2;
  const int word = (R4 || k < 3) ? 1 : 2;
  const int bit = R4 ? 8 + 4 * k : (k < 3 ? 12 + 6 * k : 6 * (k - 3));
  return __builtin_amdgcn_ubfe(e.w[word], bit, R4 ? 4 : 6) << 3;
}

__device__ __forceinline__ double* acc_at(const Lds& L, uint32_t byte_off) {
  return (double*)((char*)L.acc + byte_off);
}

// Pairs (i, j), i < j, of an edge record with vertices [0, K) owned: each
// pair with an owned first vertex adds its value once, to the row of i
// (rb[i]: the row's byte offset in the accumulators).
template <int NB, int K, int FORM, int W>
__device__ __forceinline__ void emit_edges(const Lds& L, const ERec<W>& e, const uint32_t (&rb)[NB],
                                           const uint32_t (&m)[NB], const double (&c)[NB][NB - 1], double s,
                                           const QuadArg& q) {
  constexpr int DIM = NB - 1;
  if (FORM == 2) {
#pragma unroll
    for (int a = 0; a < K; ++a) atomicAdd(acc_at(L, rb[a] + 8 * meta_diag(m[a])), s * q.M[a * NB + a]);
  }
  int k = 0;
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = i + 1; j < NB; ++j, ++k) {
      if (i >= K) continue;
      const double kij = FORM == 2 ? s * q.M[i * NB + j] : s * dotc<DIM>(c[i], c[j]);
      atomicAdd(acc_at(L, rb[i] + erec_rank8<NB>(e, k)), kij);
    }
}

// Stages the block's nodes in the edge layout (node word w(s) = {x, y}, w(s) +
// 16 = {z | coef, packed coef and row meta with its offset in bytes}) and
// zeroes the accumulators.  A lane stages slots tid, tid + NT, ... with all
// of its first K halo ids and meta words loaded before any gather; empty
// slots and lanes past the halo keep node 0's data or stay unwritten.
template <int NT, int DIM, bool COEF>
__device__ __forceinline__ void stage_edge(const Lds& L, double2* nodes, int2* rt, const int32_t* __restrict__ halo,
                                           const uint32_t* __restrict__ meta, const int2* __restrict__ rows,
                                           int nacc, int nh, int nown, const double* __restrict__ X,
                                           const float* __restrict__ coef) {
  constexpr int K = (256 + NT - 1) / NT;
  const int2 rv = (int)threadIdx.x < nown ? rows[threadIdx.x] : make_int2(0, 0);
  int32_t g[K];
  uint32_t mh[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int h = min((int)threadIdx.x + k * NT, nh - 1);
    g[k] = halo[h];
    mh[k] = meta[h];
  }
  double2 n0[K], n1[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t gi = g[k] >= 0 ? g[k] : 0;
    const float cf = COEF ? coef[gi] : 0.0f;
    const uint32_t mb = (mh[k] & 0xffff0000u) | ((mh[k] & 0xffffu) << 3);
    const double pc = pack_coef_meta(cf, mb);
    if (DIM == 3) {
      n0[k] = make_double2(X[gi * 3], X[gi * 3 + 1]);
      n1[k] = make_double2(X[gi * 3 + 2], pc);
    } else {
      n0[k] = make_double2(X[gi * 2], X[gi * 2 + 1]);
      n1[k] = make_double2(0.0, pc);
    }
  }
  for (int i = threadIdx.x; i < nacc / 2; i += NT) ((double2*)L.acc)[i] = make_double2(0.0, 0.0);
  if ((nacc & 1) && threadIdx.x == 0) L.acc[nacc - 1] = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int h = threadIdx.x + k * NT;
    if (h < nh) {
      nodes[node_word(h)] = n0[k];
      nodes[node_word(h) + 16] = n1[k];
    }
  }
  if ((int)threadIdx.x < nown) rt[threadIdx.x] = rv;
  for (int a = threadIdx.x + NT; a < nown; a += NT) rt[a] = rows[a];
  for (int h = threadIdx.x + K * NT; h < nh; h += NT) {
    const int64_t gh = halo[h] >= 0 ? halo[h] : 0;
    const uint32_t m = meta[h];
    const double pc = pack_coef_meta(COEF ? coef[gh] : 0.0f, (m & 0xffff0000u) | ((m & 0xffffu) << 3));
    nodes[node_word(h)] = DIM == 3 ? make_double2(X[gh * 3], X[gh * 3 + 1]) : make_double2(X[gh * 2], X[gh * 2 + 1]);
    nodes[node_word(h) + 16] = make_double2(DIM == 3 ? X[gh * 3 + 2] : 0.0, pc);
  }
}

// Owned count of the records of group g (0: NB owned, 1: NB-1, ...).
__device__ __forceinline__ int rec_group(int r, int e0, int e1, int e2) {
  return (r >= e0 ? 1 : 0) + (r >= e1 ? 1 : 0) + (r >= e2 ? 1 : 0);
}

// LO: FP32 output values.  1/|det| then takes one Newton step (relative error
// below 2^-40 from the hardware estimate) and the nodal coefficients are
// summed in FP32 (the relabelling-invariant weights are all W[0]); both stay
// far below the FP32 rounding of the stored values.  FP64 outputs keep two
// Newton steps and the FP64 weighted sum.
template <int NB, int FORM, int NT, int W, bool LO>
__device__ __forceinline__ void edge_records(const Lds& L, const double2* nodes, const ERecBuf<W>& rb, int nrec,
                                             int e0, int e1, int e2, ERec<W> nxt, double scale, const QuadArg& q) {
  constexpr int DIM = NB - 1;
  const double sw = FORM == 0 ? scale * q.wsum : scale * q.W[0];
  const int lane = threadIdx.x % kWave;
  // r0: the wave's first record of the iteration (wave-uniform, kept in SGPRs)
  const int w0 = __builtin_amdgcn_readfirstlane((int)threadIdx.x - lane);
  for (int r0 = w0; r0 < nrec; r0 += NT) {
    const int r = r0 + lane;
    const ERec<W> e = nxt;
    nxt = rb[r + NT];
    if (r >= nrec) continue;
    double x[NB][DIM];
    float cf[NB];
    uint32_t m[NB], rw[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      const double2* nd = nodes + erec_word<NB>(e, a);
      const double2 n0 = nd[0], n1 = nd[16];
      m[a] = packed_meta(n1);
      rw[a] = m[a] & 0xffffu;
      node_xyz<DIM>(n0, n1, x[a]);
      cf[a] = __int_as_float(__double2loint(n1.y));
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    const double adet = fabs(det);
    double s;
    if (FORM == 0) {
      s = sw * recip<LO ? 1 : 2>(adet);
    } else if (FORM == 1) {
      double k;
      if (LO) {
        float ks = cf[0];
#pragma unroll
        for (int a = 1; a < NB; ++a) ks += cf[a];
        k = (double)ks;
      } else {
        k = 0.0;
#pragma unroll
        for (int a = 0; a < NB; ++a) k += (double)cf[a];
      }
      s = sw * k * recip<LO ? 1 : 2>(adet);
    } else {
      s = scale * adet;
    }
    // the wave's records [r0, last] lie in one owned-count group (no group end
    // in (r0, last]): a branch-free instance, chosen on scalar registers
    const int last = min(r0 + kWave - 1, nrec - 1);
    const bool split = (r0 < e0 && e0 <= last) || (r0 < e1 && e1 <= last) || (r0 < e2 && e2 <= last);
    if (!split) {
      if (r0 < e0) {
        emit_edges<NB, NB, FORM>(L, e, rw, m, c, s, q);
      } else if (r0 < e1) {
        emit_edges<NB, NB - 1, FORM>(L, e, rw, m, c, s, q);
      } else if (r0 < e2) {
        emit_edges<NB, NB - 2, FORM>(L, e, rw, m, c, s, q);
      } else if (NB == 4) {
        emit_edges<NB, (NB == 4 ? 1 : 0), FORM>(L, e, rw, m, c, s, q);
      }
      continue;
    }
    const int nk = NB - rec_group(r, e0, e1, e2);
    if (FORM == 2) {
#pragma unroll
      for (int a = 0; a < NB; ++a)
        if (a < nk) atomicAdd(acc_at(L, rw[a] + 8 * meta_diag(m[a])), s * q.M[a * NB + a]);
    }
    int k = 0;
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = i + 1; j < NB; ++j, ++k) {
        if (i >= nk) continue;
        const double kij = FORM == 2 ? s * q.M[i * NB + j] : s * dotc<DIM>(c[i], c[j]);
        atomicAdd(acc_at(L, rw[i] + erec_rank8<NB>(e, k)), kij);
      }
  }
}

// Row sums of the edge layout (the diagonal accumulators hold 0): four lanes
// per row (a quad) each sum every fourth entry, then two DPP quad swaps add
// the four partial sums; the quad's first lane writes minus the sum to the
// diagonal.  Every lane of a quad takes the same rows, so the swaps always
// read active lanes.
template <int CTRL>
__device__ __forceinline__ double quad_swap(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
template <int NT>
__device__ __forceinline__ void rowsum_edge(const Lds& L, const double2* nodes, int nown) {
  const int q = threadIdx.x & 3;
  for (int a = threadIdx.x >> 2; a < nown; a += NT / 4) {
    const uint32_t m = (uint32_t)__double2hiint(nodes[node_word(a) + 16].y);
    const int len = meta_len(m);
    const double* r = acc_at(L, m & 0xffffu);
    double s = 0.0;
    for (int k = q; k < len; k += 4) s += r[k];
    s += quad_swap<0xB1>(s);  // quad_perm [1,0,3,2]: lanes xor 1
    s += quad_swap<0x4E>(s);  // quad_perm [2,3,0,1]: lanes xor 2
    if (q == 0 && len > 0) *acc_at(L, (m & 0xffffu) + 8 * meta_diag(m)) = -s;
  }
}

// Row stores of the edge kernel: 16 lanes per row (rows of any length),
// rows in accumulator order, so a wave's store covers four consecutive rows.
template <int NT, typename OT>
__device__ __forceinline__ void store_rows(const Lds& L, const int2* rt, int nown, OT* __restrict__ values) {
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  for (int i = g; i < nown; i += NT / 16) {
    const int2 r = rt[i];
    const int len = (int)((uint32_t)r.y >> 16);
    const double* a = acc_at(L, (uint32_t)r.y & 0xffffu);
    for (int k = j; k < len; k += 16) values[(int64_t)r.x + k] = (OT)a[k];
  }
}

// Symmetric scalar forms on edge records: staging, element loop, mirror
// copies, row sums (Laplace-type forms), row stores.
template <int NB, int FORM, typename OT, int NT, int W>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_asm_matrix_edge(
    PlanArgs p, EdgeArgs ea, const double* __restrict__ X, const float* __restrict__ coef, double scale, QuadArg q,
    OT* __restrict__ values) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int4 eb = *(const int4*)(ea.eblk + 4 * b);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, nacc = bi1.w;
  const int moff = eb.x, nmir = eb.y & 0xffff, e2 = (int)((uint32_t)eb.y >> 16);
  const int e0 = eb.z & 0xffff, e1 = (int)((uint32_t)eb.z >> 16), rowoff = eb.w;
  Lds L;
  L.acc = (double*)smem;
  double2* nodes = (double2*)(smem + a16(8 * (size_t)nacc));
  int2* rt = (int2*)((char*)nodes + 512 * (size_t)((nh + 15) / 16));
  const ERecBuf<W> rb(ea.recs, roff, nrec);
  const int t = threadIdx.x;
  ERec<W> nxt = rb[t];  // a lane's first record loads before the staging
  stage_edge<NT, DIM, FORM == 1>(L, nodes, rt, p.halo + hoff, p.meta + hoff, ea.rows + rowoff, nacc, nh, nown, X,
                                 coef);
  __syncthreads();
  edge_records<NB, FORM, NT, W, std::is_same<OT, float>::value>(L, nodes, rb, nrec, e0, e1, e2, nxt, scale, q);
  const uint32_t mv = t < nmir ? ea.mir[moff + t] : 0u;  // in flight during the barrier
  __syncthreads();
  if (t < nmir) L.acc[mv >> 16] = L.acc[mv & 0xffffu];
  for (int i = t + NT; i < nmir; i += NT) {
    const uint32_t v = ea.mir[moff + i];
    L.acc[v >> 16] = L.acc[v & 0xffffu];
  }
  __syncthreads();
  if (FORM != 2) {
    rowsum_edge<NT>(L, nodes, nown);
    __syncthreads();
  }
  store_rows<NT>(L, rt, nown, values);
}

// Integrand U_a^T D V_b with U = (phi, grad phi) of the row (u) basis
// function a and V of the column (v) basis function b, D constant
// ((DIM+1)^2 row-major; index 0 = value, 1.. = gradient components).  With
// g = c / det (c = cofactor vectors) and the quadrature moments wsum, W, M:
//   K_ab = |det| D00 M_ab + sgn(det) W (c_a . D_g0 + D_0g . c_b) + wsum/|det| c_a^T D_gg c_b.
// No symmetry and no zero row sum are assumed: every (owned a, b) entry is
// accumulated.
struct GenD {
  double d[16];
};

// Records of one block for k_asm_general: every (owned a, b) entry through `em`.
template <int NB, int NT, typename Emit>
__device__ __forceinline__ void general_records(const Lds& L, const PlanArgs& p, int roff, int nrec, int nown,
                                                const GenD& D, const QuadArg& q, const Emit& em) {
  constexpr int DIM = NB - 1, DD = DIM + 1;
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    int l[NB];
    double x[NB][DIM];
    uint32_t m[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      node_xyz<DIM>(L.p0[l[a]], L.p1[l[a]], x[a]);
      m[a] = L.ms[l[a]];
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    const double adet = fabs(det);
    const double sgg = q.wsum / adet, sw = det < 0.0 ? -q.W[0] : q.W[0];
    const double mdiag = D.d[0] * q.M[0] * adet, moff = D.d[0] * q.M[1] * adet;
    double vb[NB];  // sgn W (D_0g . c_b)
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < DIM; ++j) t += D.d[1 + j] * c[bb][j];
      vb[bb] = sw * t;
    }
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      if (l[a] >= nown) continue;
      double ta[DIM];  // (c_a^T D_gg)_j
      double ua = 0.0; // c_a . D_g0
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) t += c[a][i] * D.d[(1 + i) * DD + 1 + j];
        ta[j] = t;
        ua += c[a][j] * D.d[(1 + j) * DD];
      }
      ua *= sw;
      const int base = meta_off(m[a]);
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        double gg = 0.0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) gg += ta[j] * c[bb][j];
        const double v = (a == bb ? mdiag : moff) + ua + vb[bb] + sgg * gg;
        const int slot = bb == a ? meta_diag(m[a]) : rec_rank<NB>(rc, a, bb);
        em(L, l[a], L.acc + base + slot, v);
      }
    }
  }
}

template <int NB, typename OT, int NT, bool DET>
__global__ __launch_bounds__(NT) void k_asm_general(PlanArgs p, const double* __restrict__ X, GenD D, QuadArg q,
                                                   OT* __restrict__ values) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, runoff = bi1.y, nrun = bi1.z, nacc = bi1.w;
  const Lds L = carve(smem, nacc, nh, nown);
  stage<NT, DIM, float, false>(L, p.halo + hoff, p.meta + hoff, nacc, nh, nown, X, (const float*)nullptr);
  __syncthreads();
  if (DET) {
    general_records<NB, NT>(L, p, roff, nrec, nown, D, q, MaxEmit{});
    __syncthreads();
    row_grids<NT>(L, nown);
    __syncthreads();
    general_records<NB, NT>(L, p, roff, nrec, nown, D, q, GridEmit{});
  } else {
    general_records<NB, NT>(L, p, roff, nrec, nown, D, q, AddEmit{});
  }
  const int4 dsc = run_descriptors<NT>(p, runoff, nrun);
  __syncthreads();
  store_runs<NT>(L, p, runoff, nrun, dsc, values);
}

// Element-local vertex index of record vertex a (record bits 112-119).
__device__ __forceinline__ int rec_orig(const uint4& rc, int a) { return (int)((rc.w >> (16 + 2 * a)) & 3u); }

// Spatially varying bilinear forms U_a^T D(q) V_b (U = (phi, grad phi) of the
// row basis function a, V of the column basis function b; D depends on the
// element and the quadrature point).  The caller passes per-element moments
// of the integrand in the element's own vertex order (w_q reference weights),
// NM = (NB + DIM)^2 doubles per element:
//   MM[NB][NB] = sum_q w_q phi_a(q) phi_b(q) D00(q)
//   P [NB][DIM] = sum_q w_q phi_a(q) D_0g(q)
//   Qm[NB][DIM] = sum_q w_q D_g0(q) phi_b(q)
//   S [DIM][DIM] = sum_q w_q D_gg(q)
// With grad phi = c / det (c = cofactor vectors) the element matrix is
//   K_ab = |det| MM_ab + sgn(det) (P_a . c_b + c_a . Qm_b) + c_a^T S c_b / |det|.
// The record's vertex order may differ from the element's (rec_orig maps it).
template <int NB, int NT, typename Emit>
__device__ __forceinline__ void moments_records(const Lds& L, const PlanArgs& p, const int32_t* __restrict__ rec_elem,
                                                const double* __restrict__ mom, int roff, int nrec, int nown,
                                                const Emit& em) {
  constexpr int DIM = NB - 1, NM = (NB + DIM) * (NB + DIM);
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    const double* m = mom + (int64_t)rec_elem[roff + r] * NM;
    const double* MM = m;
    const double* P = m + NB * NB;
    const double* Qm = P + NB * DIM;
    const double* S = Qm + NB * DIM;
    int l[NB], o[NB];
    double x[NB][DIM];
    uint32_t ms[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      o[a] = rec_orig(rc, a);
      node_xyz<DIM>(L.p0[l[a]], L.p1[l[a]], x[a]);
      ms[a] = L.ms[l[a]];
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    const double adet = fabs(det), inv = recip(adet), sg = det < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      if (l[a] >= nown) continue;
      double ta[DIM];  // (c_a^T S)_j
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) t += c[a][i] * S[i * DIM + j];
        ta[j] = t;
      }
      const double* Pa = P + o[a] * DIM;
      const int base = meta_off(ms[a]);
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        const double* Qb = Qm + o[bb] * DIM;
        double gg = 0.0, pc = 0.0, cq = 0.0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
          gg += ta[j] * c[bb][j];
          pc += Pa[j] * c[bb][j];
          cq += c[a][j] * Qb[j];
        }
        const double v = adet * MM[o[a] * NB + o[bb]] + sg * (pc + cq) + inv * gg;
        const int slot = bb == a ? meta_diag(ms[a]) : rec_rank<NB>(rc, a, bb);
        em(L, l[a], L.acc + base + slot, v);
      }
    }
  }
}

// DET: deterministic mode (MaxEmit / GridEmit, as k_asm_matrix).
template <int NB, typename OT, int NT, bool DET>
__global__ __launch_bounds__(NT) void k_asm_moments(PlanArgs p, const int32_t* __restrict__ rec_elem,
                                                   const double* __restrict__ X, const double* __restrict__ mom,
                                                   OT* __restrict__ values) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w;
  const int nrec = bi1.x, runoff = bi1.y, nrun = bi1.z, nacc = bi1.w;
  const Lds L = carve(smem, nacc, nh, nown);
  stage<NT, DIM, float, false>(L, p.halo + hoff, p.meta + hoff, nacc, nh, nown, X, (const float*)nullptr);
  __syncthreads();
  if (DET) {
    moments_records<NB, NT>(L, p, rec_elem, mom, roff, nrec, nown, MaxEmit{});
    __syncthreads();
    row_grids<NT>(L, nown);
    __syncthreads();
    moments_records<NB, NT>(L, p, rec_elem, mom, roff, nrec, nown, GridEmit{});
  } else {
    moments_records<NB, NT>(L, p, rec_elem, mom, roff, nrec, nown, AddEmit{});
  }
  const int4 dsc = run_descriptors<NT>(p, runoff, nrun);
  __syncthreads();
  store_runs<NT>(L, p, runoff, nrun, dsc, values);
}

// Spatially varying linear forms F0(q) v + F1(q) . grad v with per-element
// moments in the element's vertex order, NB + DIM doubles per element:
//   L0[NB] = sum_q w_q phi_a(q) F0(q),  L1[DIM] = sum_q w_q F1(q)
// so that b_a += |det| L0_a + sgn(det) c_a . L1.
template <int NB, int NT, typename Emit>
__device__ __forceinline__ void load_moment_records(const Lds& L, const PlanArgs& p,
                                                    const int32_t* __restrict__ rec_elem,
                                                    const double* __restrict__ mom, int roff, int nrec, int nown,
                                                    const Emit& em) {
  constexpr int DIM = NB - 1, NM = NB + DIM;
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    const double* m = mom + (int64_t)rec_elem[roff + r] * NM;
    int l[NB];
    double x[NB][DIM];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      node_xyz<DIM>(L.p0[l[a]], L.p1[l[a]], x[a]);
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    const double adet = fabs(det), sg = det < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      if (l[a] >= nown) continue;
      double g = 0.0;
#pragma unroll
      for (int j = 0; j < DIM; ++j) g += c[a][j] * m[NB + j];
      em(l[a], adet * m[rec_orig(rc, a)] + sg * g);
    }
  }
}

// DET: deterministic mode, as k_asm_load.
template <int NB, int NT, bool DET>
__global__ __launch_bounds__(NT) void k_asm_load_moments(PlanArgs p, const int32_t* __restrict__ rec_elem,
                                                        const double* __restrict__ X,
                                                        const double* __restrict__ mom, double* __restrict__ out) {
  constexpr int DIM = NB - 1;
  extern __shared__ __align__(16) char smem[];
  const int b = tmx::xcd_remap(blockIdx.x, p.nblk);
  const int4 bi0 = *(const int4*)(p.blk + 8 * b);
  const int4 bi1 = *(const int4*)(p.blk + 8 * b + 4);
  const int hoff = bi0.x, nh = bi0.y, nown = bi0.z, roff = bi0.w, nrec = bi1.x;
  const Lds L = carve(smem, 0, nh, nown);
  stage<NT, DIM, float, false>(L, p.halo + hoff, p.meta + hoff, 0, nh, nown, X, (const float*)nullptr);
  __syncthreads();
  if (DET) {
    load_moment_records<NB, NT>(L, p, rec_elem, mom, roff, nrec, nown, [&](int row, double v) {
      atomicMax((unsigned long long*)(L.accb + row), (unsigned long long)__double_as_longlong(fabs(v)));
    });
    __syncthreads();
    load_grids<NT>(L, nown);
    __syncthreads();
    load_moment_records<NB, NT>(L, p, rec_elem, mom, roff, nrec, nown, [&](int row, double v) {
      const int k = (int)L.ms[row];
      atomicAdd(L.accb + row, ldexp(rint(ldexp(v, -k)), k));
    });
  } else {
    load_moment_records<NB, NT>(L, p, rec_elem, mom, roff, nrec, nown,
                                [&](int row, double v) { atomicAdd(L.accb + row, v); });
  }
  __syncthreads();
  for (int a = threadIdx.x; a < nown; a += NT) out[p.halo[hoff + a]] = L.accb[a];
}

// FORM: 0 constant load c, 1 nodal f.
// The records of one block for k_asm_load: c * int phi_a (FORM 0) or
// int (sum_c f_c phi_c) phi_a (FORM 1) of every owned vertex through
// em(row slot, value).
template <int NB, int FORM, int NT, typename Emit>
__device__ __forceinline__ void load_records(const Lds& L, const PlanArgs& p, int roff, int nrec, int nown,
                                             double cval, const QuadArg& q, const Emit& em) {
  constexpr int DIM = NB - 1;
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    int l[NB];
    double x[NB][DIM], cf[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      const double2 n0 = L.p0[l[a]], n1 = L.p1[l[a]];
      node_xyz<DIM>(n0, n1, x[a]);
      cf[a] = node_coef<DIM>(n1);
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
      em(l[a], adet * v);
    }
  }
}

// Load vectors.  DET: deterministic mode, the arithmetic of the matrix kernels'
// MaxEmit / GridEmit: a first pass takes each owned row's largest |entry| in
// L.accb, its grid exponent then replaces the row meta in L.ms (which no load
// kernel reads) and the second pass sums the rounded entries in L.accb.
template <int NB, int FORM, typename CT, int NT, bool DET>
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
  if (DET) {
    load_records<NB, FORM, NT>(L, p, roff, nrec, nown, cval, q, [&](int row, double v) {
      atomicMax((unsigned long long*)(L.accb + row), (unsigned long long)__double_as_longlong(fabs(v)));
    });
    __syncthreads();
    load_grids<NT>(L, nown);
    __syncthreads();
    load_records<NB, FORM, NT>(L, p, roff, nrec, nown, cval, q, [&](int row, double v) {
      const int k = (int)L.ms[row];
      atomicAdd(L.accb + row, ldexp(rint(ldexp(v, -k)), k));
    });
  } else {
    load_records<NB, FORM, NT>(L, p, roff, nrec, nown, cval, q,
                               [&](int row, double v) { atomicAdd(L.accb + row, v); });
  }
  __syncthreads();
  for (int a = threadIdx.x; a < nown; a += NT) out[p.halo[hoff + a]] = L.accb[a];
}

// dkappa_c += sum_e W_c / |det| * sum_{a owned, b} g[slot(a, b)] c_a . c_b
// (per-node accumulator: the coefficient word of each LDS node, unused by this kernel)
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
  for (int r = threadIdx.x; r < nrec; r += NT) {
    const uint4 rc = p.recs[roff + r];
    const uint64_t lo = (uint64_t)rc.x | ((uint64_t)rc.y << 32);
    int l[NB];
    double x[NB][DIM];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      l[a] = rec_local(lo, a);
      node_xyz<DIM>(L.p0[l[a]], L.p1[l[a]], x[a]);
    }
    double c[NB][DIM], det;
    p1_geometry<DIM>(x, c, det);
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      if (l[a] >= nown) continue;
      const uint32_t m = L.ms[l[a]];
      const double* row = L.acc + meta_off(m);
      t += row[meta_diag(m)] * dotc<DIM>(c[a], c[a]);
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        if (bb == a) continue;
        t += row[rec_rank<NB>(rc, a, bb)] * dotc<DIM>(c[a], c[bb]);
      }
    }
    t *= recip(fabs(det));
#pragma unroll
    for (int a = 0; a < NB; ++a) atomicAdd(node_coef_ptr<DIM>(L, l[a]), q.W[a] * t);
  }
  __syncthreads();
  for (int h = threadIdx.x; h < nh; h += NT) {
    const double v = *node_coef_ptr<DIM>(L, h);
    const int32_t g = p.halo[hoff + h];
    if (v != 0.0 && g >= 0) atomicAdd(dkappa + g, v);
  }
}

PlanArgs plan_args(const tm_plan_dev* d) {
  return PlanArgs{d->blk, d->halo, d->meta, (const uint4*)d->recs, (const int4*)d->runs, d->n_blocks};
}

// The plan may relabel an element's vertices, so the moments must not depend
// on the labels: W equal, M's diagonal equal, M's off-diagonal equal.
bool quad_is_symmetric(const tm_quad* q, int nb) {
  double scale = fabs(q->wsum);
  for (int i = 0; i < nb * nb; ++i) scale = fmax(scale, fabs(q->M[i]));
  const double tol = 1e-12 * scale;
  for (int a = 0; a < nb; ++a) {
    if (fabs(q->W[a] - q->W[0]) > tol || fabs(q->M[a * nb + a] - q->M[0]) > tol) return false;
    for (int b = 0; b < nb; ++b)
      if (a != b && fabs(q->M[a * nb + b] - q->M[1]) > tol) return false;
  }
  return true;
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

// Edge-record instances of the symmetric forms (coefficient absent or FP32).
template <int NB, typename OT, int W>
int dispatch_matrix_edge(int form, const tm_plan_dev* plan, const double* X, const float* coef, double scale,
                         QuadArg q, OT* values, hipStream_t st) {
  const PlanArgs p = plan_args(plan);
  const EdgeArgs e{plan->erecs, plan->eblk, plan->emir, (const int2*)plan->erows};
  TM_REQUIRE(plan->elds_bytes > 0 && plan->elds_bytes <= 65536, TM_ERR_ARG, "edge LDS size %d", plan->elds_bytes);
  tm_plan_dev pe = *plan;  // the edge layout's LDS size
  pe.lds_bytes = plan->elds_bytes;
  if (form == TM_FORM_LAPLACE)
    TM_BY_THREADS(&pe, (k_asm_matrix_edge<NB, 0, OT, NT, W>), st, p, e, X, coef, scale, q, values);
  if (form == TM_FORM_MASS)
    TM_BY_THREADS(&pe, (k_asm_matrix_edge<NB, 2, OT, NT, W>), st, p, e, X, coef, scale, q, values);
  TM_BY_THREADS(&pe, (k_asm_matrix_edge<NB, 1, OT, NT, W>), st, p, e, X, coef, scale, q, values);
}

template <int NB, typename OT, bool DET>
int dispatch_matrix(int form, const tm_plan_dev* plan, const double* X, const void* coef, int cdt,
                    double scale, QuadArg q, OT* values, hipStream_t st) {
  PlanArgs p = plan_args(plan);
  const bool edge_ok = !DET && plan->erecs != nullptr && plan->eblk != nullptr && plan->emir != nullptr &&
                       plan->erows != nullptr &&
                       (form == TM_FORM_LAPLACE || form == TM_FORM_MASS ||
                        (form == TM_FORM_KAPPA_LAPLACE && coef != nullptr && cdt == TM_F32));
  if (edge_ok) {
    TM_REQUIRE(plan->erec_words == 2 || (NB == 4 && plan->erec_words == 3), TM_ERR_ARG,
               "edge records of %d words", plan->erec_words);
    if (NB == 4 && plan->erec_words == 3)
      return dispatch_matrix_edge<NB, OT, 3>(form, plan, X, (const float*)coef, scale, q, values, st);
    return dispatch_matrix_edge<NB, OT, 2>(form, plan, X, (const float*)coef, scale, q, values, st);
  }
  if (form == TM_FORM_LAPLACE)
    TM_BY_THREADS(plan, (k_asm_matrix<NB, 0, float, OT, NT, DET>), st, p, X, (const float*)nullptr, scale, q,
                  values);
  if (form == TM_FORM_MASS)
    TM_BY_THREADS(plan, (k_asm_matrix<NB, 2, float, OT, NT, DET>), st, p, X, (const float*)nullptr, scale, q,
                  values);
  if (form == TM_FORM_KAPPA_LAPLACE) {
    TM_REQUIRE(coef != nullptr, TM_ERR_ARG, "kappa-Laplace needs a nodal coefficient");
    if (cdt == TM_F32)
      TM_BY_THREADS(plan, (k_asm_matrix<NB, 1, float, OT, NT, DET>), st, p, X, (const float*)coef, scale, q,
                    values);
    TM_BY_THREADS(plan, (k_asm_matrix<NB, 1, double, OT, NT, DET>), st, p, X, (const double*)coef, scale, q,
                  values);
  }
  TM_REQUIRE(false, TM_ERR_ARG, "unknown form %d", form);
}

template <int NB, typename OT>
int dispatch_matrix_det(int form, bool det, const tm_plan_dev* plan, const double* X, const void* coef, int cdt,
                        double scale, QuadArg q, OT* values, hipStream_t st) {
  if (det) return dispatch_matrix<NB, OT, true>(form, plan, X, coef, cdt, scale, q, values, st);
  return dispatch_matrix<NB, OT, false>(form, plan, X, coef, cdt, scale, q, values, st);
}

}  // namespace

extern "C" int tm_asm_p1_matrix(int form, int nb, int dim, const tm_plan_dev* plan,
                                const double* coords, const void* coef, int coef_dtype,
                                double scale, const tm_quad* quad, void* values, int value_dtype,
                                void* stream) {
  TM_REQUIRE(plan && quad && coords && values, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG,
             "only P1 triangles (nb=3, dim=2) and tetrahedra (nb=4, dim=3) are supported");
  TM_REQUIRE(quad_is_symmetric(quad, nb), TM_ERR_ARG,
             "quadrature moments must be invariant under vertex relabelling");
  hipStream_t st = (hipStream_t)stream;
  QuadArg q = quad_arg(quad);
  const bool det = (form & TM_ASM_DETERMINISTIC) != 0;
  form &= ~TM_ASM_DETERMINISTIC;
  if (value_dtype == TM_F32) {
    if (nb == 3)
      return dispatch_matrix_det<3>(form, det, plan, coords, coef, coef_dtype, scale, q, (float*)values, st);
    return dispatch_matrix_det<4>(form, det, plan, coords, coef, coef_dtype, scale, q, (float*)values, st);
  }
  if (nb == 3)
    return dispatch_matrix_det<3>(form, det, plan, coords, coef, coef_dtype, scale, q, (double*)values, st);
  return dispatch_matrix_det<4>(form, det, plan, coords, coef, coef_dtype, scale, q, (double*)values, st);
}

extern "C" int tm_asm_p1_general(int nb, int dim, const tm_plan_dev* plan, const double* coords, const double* D,
                                 const tm_quad* quad, void* values, int value_dtype, int flags, void* stream) {
  TM_REQUIRE(plan && quad && coords && D && values, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG,
             "only P1 triangles (nb=3, dim=2) and tetrahedra (nb=4, dim=3) are supported");
  TM_REQUIRE(quad_is_symmetric(quad, nb), TM_ERR_ARG,
             "quadrature moments must be invariant under vertex relabelling");
  TM_REQUIRE(value_dtype == TM_F32 || value_dtype == TM_F64, TM_ERR_ARG, "values must be f32 or f64");
  hipStream_t st = (hipStream_t)stream;
  const QuadArg q = quad_arg(quad);
  GenD g{};
  for (int i = 0; i < (dim + 1) * (dim + 1); ++i) g.d[i] = D[i];
  const PlanArgs p = plan_args(plan);
  if (flags & TM_ASM_DETERMINISTIC) {
    if (value_dtype == TM_F32) {
      if (nb == 3) TM_BY_THREADS(plan, (k_asm_general<3, float, NT, true>), st, p, coords, g, q, (float*)values);
      TM_BY_THREADS(plan, (k_asm_general<4, float, NT, true>), st, p, coords, g, q, (float*)values);
    }
    if (nb == 3) TM_BY_THREADS(plan, (k_asm_general<3, double, NT, true>), st, p, coords, g, q, (double*)values);
    TM_BY_THREADS(plan, (k_asm_general<4, double, NT, true>), st, p, coords, g, q, (double*)values);
  }
  if (value_dtype == TM_F32) {
    if (nb == 3) TM_BY_THREADS(plan, (k_asm_general<3, float, NT, false>), st, p, coords, g, q, (float*)values);
    TM_BY_THREADS(plan, (k_asm_general<4, float, NT, false>), st, p, coords, g, q, (float*)values);
  }
  if (nb == 3) TM_BY_THREADS(plan, (k_asm_general<3, double, NT, false>), st, p, coords, g, q, (double*)values);
  TM_BY_THREADS(plan, (k_asm_general<4, double, NT, false>), st, p, coords, g, q, (double*)values);
}

template <bool DET>
int dispatch_moments(int nb, int value_dtype, const tm_plan_dev* plan, hipStream_t st, const PlanArgs& p,
                     const double* coords, const double* moments, void* values) {
  const int32_t* re = plan->rec_elem;
  if (value_dtype == TM_F32) {
    if (nb == 3) TM_BY_THREADS(plan, (k_asm_moments<3, float, NT, DET>), st, p, re, coords, moments, (float*)values);
    TM_BY_THREADS(plan, (k_asm_moments<4, float, NT, DET>), st, p, re, coords, moments, (float*)values);
  }
  if (nb == 3) TM_BY_THREADS(plan, (k_asm_moments<3, double, NT, DET>), st, p, re, coords, moments, (double*)values);
  TM_BY_THREADS(plan, (k_asm_moments<4, double, NT, DET>), st, p, re, coords, moments, (double*)values);
}

extern "C" int tm_asm_p1_moments(int nb, int dim, const tm_plan_dev* plan, const double* coords,
                                 const double* moments, void* values, int value_dtype, int flags, void* stream) {
  TM_REQUIRE(plan && coords && moments && values && plan->rec_elem, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG,
             "only P1 triangles (nb=3, dim=2) and tetrahedra (nb=4, dim=3) are supported");
  TM_REQUIRE(value_dtype == TM_F32 || value_dtype == TM_F64, TM_ERR_ARG, "values must be f32 or f64");
  TM_REQUIRE((flags & ~TM_ASM_DETERMINISTIC) == 0, TM_ERR_ARG, "unknown flags");
  hipStream_t st = (hipStream_t)stream;
  const PlanArgs p = plan_args(plan);
  if (flags & TM_ASM_DETERMINISTIC) return dispatch_moments<true>(nb, value_dtype, plan, st, p, coords, moments, values);
  return dispatch_moments<false>(nb, value_dtype, plan, st, p, coords, moments, values);
}

extern "C" int tm_asm_p1_load_moments(int nb, int dim, const tm_plan_dev* plan, const double* coords,
                                      const double* moments, double* b, int flags, void* stream) {
  TM_REQUIRE(plan && coords && moments && b && plan->rec_elem, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG, "only P1 simplices");
  TM_REQUIRE((flags & ~TM_ASM_DETERMINISTIC) == 0, TM_ERR_ARG, "unknown flags");
  hipStream_t st = (hipStream_t)stream;
  const PlanArgs p = plan_args(plan);
  const int32_t* re = plan->rec_elem;
  if (flags & TM_ASM_DETERMINISTIC) {
    if (nb == 3) return launch_plan_kernel(k_asm_load_moments<3, 256, true>, 256, plan, st, p, re, coords, moments, b);
    return launch_plan_kernel(k_asm_load_moments<4, 256, true>, 256, plan, st, p, re, coords, moments, b);
  }
  if (nb == 3) return launch_plan_kernel(k_asm_load_moments<3, 256, false>, 256, plan, st, p, re, coords, moments, b);
  return launch_plan_kernel(k_asm_load_moments<4, 256, false>, 256, plan, st, p, re, coords, moments, b);
}

template <bool DET>
int dispatch_load(int form, int nb, const tm_plan_dev* plan, hipStream_t st, const PlanArgs& p,
                  const double* coords, const void* f, int f_dtype, double c, const QuadArg& q, double* b) {
  if (form == TM_LOAD_CONST) {
    if (nb == 3)
      return launch_plan_kernel(k_asm_load<3, 0, float, 256, DET>, 256, plan, st, p, coords, (const float*)nullptr, c,
                                q, b);
    return launch_plan_kernel(k_asm_load<4, 0, float, 256, DET>, 256, plan, st, p, coords, (const float*)nullptr, c, q,
                              b);
  }
  if (f_dtype == TM_F32) {
    if (nb == 3)
      return launch_plan_kernel(k_asm_load<3, 1, float, 256, DET>, 256, plan, st, p, coords, (const float*)f, c, q, b);
    return launch_plan_kernel(k_asm_load<4, 1, float, 256, DET>, 256, plan, st, p, coords, (const float*)f, c, q, b);
  }
  if (nb == 3)
    return launch_plan_kernel(k_asm_load<3, 1, double, 256, DET>, 256, plan, st, p, coords, (const double*)f, c, q, b);
  return launch_plan_kernel(k_asm_load<4, 1, double, 256, DET>, 256, plan, st, p, coords, (const double*)f, c, q, b);
}

extern "C" int tm_asm_p1_load(int form, int nb, int dim, const tm_plan_dev* plan, const double* coords,
                              const void* f, int f_dtype, double c, const tm_quad* quad, double* b,
                              void* stream) {
  TM_REQUIRE(plan && quad && coords && b, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG, "only P1 simplices");
  TM_REQUIRE(quad_is_symmetric(quad, nb), TM_ERR_ARG,
             "quadrature moments must be invariant under vertex relabelling");
  hipStream_t st = (hipStream_t)stream;
  const bool det = (form & TM_ASM_DETERMINISTIC) != 0;
  form &= ~TM_ASM_DETERMINISTIC;
  TM_REQUIRE(form == TM_LOAD_CONST || (form == TM_LOAD_NODAL && f != nullptr), TM_ERR_ARG, "bad load form");
  QuadArg q = quad_arg(quad);
  PlanArgs p = plan_args(plan);
  if (det) return dispatch_load<true>(form, nb, plan, st, p, coords, f, f_dtype, c, q, b);
  return dispatch_load<false>(form, nb, plan, st, p, coords, f, f_dtype, c, q, b);
}

extern "C" int tm_asm_p1_kappa_adjoint(int nb, int dim, const tm_plan_dev* plan, const double* coords,
                                       const void* g, int g_dtype, const tm_quad* quad, double* dkappa,
                                       void* stream) {
  TM_REQUIRE(plan && quad && coords && g && dkappa, TM_ERR_ARG, "null argument");
  TM_REQUIRE((nb == 3 && dim == 2) || (nb == 4 && dim == 3), TM_ERR_ARG, "only P1 simplices");
  TM_REQUIRE(quad_is_symmetric(quad, nb), TM_ERR_ARG,
             "quadrature moments must be invariant under vertex relabelling");
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
