// Host-side setup for the P1 assembly path: canonical CSR pattern, element->slot
// map and the owner-computes block plan.  Pure C++ (std::thread), no HIP.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "tm_common.h"

namespace tmx {
static thread_local std::string g_err;
void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
}  // namespace tmx

extern "C" const char* tm_last_error(void) { return tmx::g_err.c_str(); }
extern "C" int tm_version(void) { return 1; }

namespace {

int pick_threads(int req) {
  if (req > 0) return req;
  unsigned h = std::thread::hardware_concurrency();
  return h == 0 ? 1 : (int)std::min(h, 16u);
}

// Runs f(begin, end) over [0, n) split into contiguous chunks on nthreads threads.
template <class F>
void parallel_for(int64_t n, int nthreads, F f) {
  if (nthreads <= 1 || n < 8192) {
    f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=, &f] { f(b, e); });
  }
  for (auto& x : th) x.join();
}

// Node -> incident elements (ascending element ids per node).
struct Incidence {
  std::vector<int64_t> ptr;
  std::vector<int32_t> elem;
};

int build_incidence(const int64_t* conn, int64_t E, int nb, int64_t N, Incidence& inc) {
  TM_REQUIRE(E >= 0 && N >= 0 && nb >= 1 && nb <= 4, TM_ERR_ARG,
             "bad sizes: n_elem=%lld nb=%d n_points=%lld", (long long)E, nb, (long long)N);
  TM_REQUIRE(E < (int64_t(1) << 31) && N < (int64_t(1) << 31), TM_ERR_LIMIT,
             "mesh too large for int32 ids");
  inc.ptr.assign(N + 1, 0);
  for (int64_t k = 0; k < E * nb; ++k) {
    int64_t g = conn[k];
    TM_REQUIRE(g >= 0 && g < N, TM_ERR_ARG, "connectivity entry %lld = %lld out of [0, %lld)",
               (long long)k, (long long)g, (long long)N);
    inc.ptr[g + 1]++;
  }
  for (int64_t i = 0; i < N; ++i) inc.ptr[i + 1] += inc.ptr[i];
  inc.elem.resize(inc.ptr[N]);
  std::vector<int64_t> cur(inc.ptr.begin(), inc.ptr.end() - 1);
  for (int64_t e = 0; e < E; ++e)
    for (int a = 0; a < nb; ++a) inc.elem[cur[conn[e * nb + a]]++] = (int32_t)e;
  return TM_OK;
}

// Sorted unique neighbours (including itself) of node i.
inline void row_of(const Incidence& inc, const int64_t* conn, int nb, int64_t i,
                   std::vector<int32_t>& buf) {
  buf.clear();
  for (int64_t k = inc.ptr[i]; k < inc.ptr[i + 1]; ++k) {
    const int64_t* c = conn + (int64_t)inc.elem[k] * nb;
    for (int a = 0; a < nb; ++a) buf.push_back((int32_t)c[a]);
  }
  std::sort(buf.begin(), buf.end());
  buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
}

inline int32_t find_in_row(const int32_t* row_ptr, const int32_t* col_idx, int64_t r, int64_t c) {
  const int32_t* b = col_idx + row_ptr[r];
  const int32_t* e = col_idx + row_ptr[r + 1];
  const int32_t* p = std::lower_bound(b, e, (int32_t)c);
  if (p == e || *p != c) return -1;
  return (int32_t)(p - b);
}

}  // namespace

extern "C" int tm_pattern_rowptr(const int64_t* conn, int64_t E, int nb, int64_t N,
                                 int32_t* row_ptr) {
  Incidence inc;
  int st = build_incidence(conn, E, nb, N, inc);
  if (st) return st;
  std::vector<int64_t> len(N);
  parallel_for(N, pick_threads(0), [&](int64_t b, int64_t e) {
    std::vector<int32_t> buf;
    for (int64_t i = b; i < e; ++i) {
      row_of(inc, conn, nb, i, buf);
      len[i] = (int64_t)buf.size();
    }
  });
  int64_t acc = 0;
  row_ptr[0] = 0;
  for (int64_t i = 0; i < N; ++i) {
    acc += len[i];
    TM_REQUIRE(acc < (int64_t(1) << 31), TM_ERR_LIMIT, "nnz exceeds int32 range");
    row_ptr[i + 1] = (int32_t)acc;
  }
  return TM_OK;
}

extern "C" int tm_pattern_colidx(const int64_t* conn, int64_t E, int nb, int64_t N,
                                 const int32_t* row_ptr, int32_t* col_idx) {
  Incidence inc;
  int st = build_incidence(conn, E, nb, N, inc);
  if (st) return st;
  std::atomic<int> bad{0};
  parallel_for(N, pick_threads(0), [&](int64_t b, int64_t e) {
    std::vector<int32_t> buf;
    for (int64_t i = b; i < e; ++i) {
      row_of(inc, conn, nb, i, buf);
      if ((int64_t)buf.size() != (int64_t)(row_ptr[i + 1] - row_ptr[i])) {
        bad = 1;
        continue;
      }
      std::memcpy(col_idx + row_ptr[i], buf.data(), buf.size() * sizeof(int32_t));
    }
  });
  TM_REQUIRE(!bad, TM_ERR_ARG, "row_ptr does not match this connectivity");
  return TM_OK;
}

extern "C" int tm_element_slots(const int64_t* conn, int64_t E, int nb, const int32_t* row_ptr,
                                const int32_t* col_idx, int64_t* slots) {
  TM_REQUIRE(E >= 0 && nb >= 1 && nb <= 4, TM_ERR_ARG, "bad sizes");
  std::atomic<int> bad{0};
  parallel_for(E, pick_threads(0), [&](int64_t b, int64_t e) {
    for (int64_t el = b; el < e; ++el) {
      const int64_t* c = conn + el * nb;
      for (int i = 0; i < nb; ++i) {
        for (int j = 0; j < nb; ++j) {
          int32_t p = find_in_row(row_ptr, col_idx, c[i], c[j]);
          if (p < 0) {
            bad = 1;
            p = 0;
          }
          slots[(el * nb + i) * nb + j] = (int64_t)row_ptr[c[i]] + p;
        }
      }
    }
  });
  TM_REQUIRE(!bad, TM_ERR_ARG, "pattern does not contain every element pair");
  return TM_OK;
}

// ------------------------------------------------------------------ plan ----

struct PlanBlock {
  std::vector<int32_t> halo;
  std::vector<uint32_t> meta;
  std::vector<uint32_t> recs;
  std::vector<int32_t> rec_elem;
  std::vector<int32_t> runs;
  int32_t n_own = 0, n_acc = 0, lds = 0;
};

struct tm_plan {
  std::vector<int32_t> blk, halo, rec_elem, runs;
  std::vector<uint32_t> meta, recs;
  int64_t info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

namespace {

constexpr int kMaxHalo = 1024;  // 10-bit local ids in the element records
constexpr int kMaxRank = 64;    // 6-bit in-row ranks
constexpr int kSlack = 64;      // alignment padding of the kernel's LDS regions

struct PlanCtx {
  const double* pts;
  int dim;
  const int64_t* conn;
  int64_t E, N;
  int nb;
  const int32_t* row_ptr;
  const int32_t* col_idx;
  const Incidence* inc;
  int max_rows, budget, entry_bytes, node_bytes, own_bytes;
};

struct Scratch {
  std::vector<uint32_t> estamp, nstamp;
  std::vector<int32_t> local;
  uint32_t tag = 0;
  explicit Scratch(const PlanCtx& c) : estamp(c.E, 0), nstamp(c.N, 0), local(c.N, -1) {}
};

inline int64_t align16(int64_t x) { return (x + 15) & ~int64_t(15); }

// Splits ids in place into two halves along the longest bounding-box axis;
// returns the split position (left part gets ceil(parts/2)/parts of the ids).
size_t rcb_split(const PlanCtx& c, std::vector<int32_t>& ids, size_t lo, size_t hi, int max_rows) {
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (size_t k = lo; k < hi; ++k)
    for (int d = 0; d < c.dim; ++d) {
      double v = c.pts[(int64_t)ids[k] * c.dim + d];
      mn[d] = std::min(mn[d], v);
      mx[d] = std::max(mx[d], v);
    }
  int ax = 0;
  for (int d = 1; d < c.dim; ++d)
    if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
  size_t n = hi - lo;
  size_t parts = (n + max_rows - 1) / max_rows;
  if (parts < 2) parts = 2;
  size_t mid = lo + (size_t)((n * (parts / 2)) / parts);
  if (mid == lo) mid = lo + 1;
  const double* P = c.pts;
  int dim = c.dim;
  std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi,
                   [P, dim, ax](int32_t a, int32_t b) {
                     double va = P[(int64_t)a * dim + ax], vb = P[(int64_t)b * dim + ax];
                     return va < vb || (va == vb && a < b);
                   });
  return mid;
}

// Builds the block for the owned node set; on a capacity violation splits the
// set and recurses.  Returns TM_OK or an error.
int make_blocks(const PlanCtx& c, Scratch& s, std::vector<int32_t> owned,
                std::vector<PlanBlock>& out) {
  std::sort(owned.begin(), owned.end());
  const int nown = (int)owned.size();
  if (++s.tag == 0) {
    std::fill(s.estamp.begin(), s.estamp.end(), 0);
    std::fill(s.nstamp.begin(), s.nstamp.end(), 0);
    s.tag = 1;
  }
  const uint32_t tag = s.tag;
  std::vector<int32_t> elems;
  int64_t n_acc = 0;
  for (int a = 0; a < nown; ++a) {
    int32_t g = owned[a];
    s.nstamp[g] = tag;
    s.local[g] = a;
    n_acc += c.row_ptr[g + 1] - c.row_ptr[g];
    for (int64_t k = c.inc->ptr[g]; k < c.inc->ptr[g + 1]; ++k) {
      int32_t e = c.inc->elem[k];
      if (s.estamp[e] != tag) {
        s.estamp[e] = tag;
        elems.push_back(e);
      }
    }
  }
  std::sort(elems.begin(), elems.end());
  std::vector<int32_t> others;
  for (int32_t e : elems)
    for (int a = 0; a < c.nb; ++a) {
      int32_t g = (int32_t)c.conn[(int64_t)e * c.nb + a];
      if (s.nstamp[g] != tag) {
        s.nstamp[g] = tag;
        others.push_back(g);
      }
    }
  std::sort(others.begin(), others.end());
  const int64_t nh = nown + (int64_t)others.size();
  const int64_t lds = align16(n_acc * c.entry_bytes) + align16(nh * c.node_bytes) +
                      align16((int64_t)nown * c.own_bytes) + kSlack;
  if (nh > kMaxHalo || n_acc > 65535 || lds > c.budget) {
    TM_REQUIRE(nown > 1, TM_ERR_LIMIT,
               "a single row does not fit the LDS budget (halo %lld, entries %lld, %lld bytes)",
               (long long)nh, (long long)n_acc, (long long)lds);
    std::vector<int32_t> ids(owned);
    size_t mid = rcb_split(c, ids, 0, ids.size(), std::max(1, nown / 2));
    std::vector<int32_t> left(ids.begin(), ids.begin() + mid), right(ids.begin() + mid, ids.end());
    int st = make_blocks(c, s, std::move(left), out);
    if (st) return st;
    return make_blocks(c, s, std::move(right), out);
  }
  for (size_t k = 0; k < others.size(); ++k) s.local[others[k]] = nown + (int32_t)k;

  PlanBlock B;
  B.n_own = nown;
  B.n_acc = (int32_t)n_acc;
  B.lds = (int32_t)lds;
  // row meta of owned node a (halo order): accumulator offset | diagonal rank
  // << 16 | row length << 24
  std::vector<uint32_t> meta0(nh, 0u);
  uint32_t lrp = 0;
  for (int a = 0; a < nown; ++a) {
    int32_t g = owned[a];
    const int32_t len = c.row_ptr[g + 1] - c.row_ptr[g];
    int32_t dr = 0;  // a node in no element has an empty row
    if (len > 0) {
      dr = find_in_row(c.row_ptr, c.col_idx, g, g);
      TM_REQUIRE(dr >= 0 && dr < 256, TM_ERR_LIMIT, "row %d has no diagonal or rank >= 256", g);
    }
    TM_REQUIRE(len < 256, TM_ERR_LIMIT, "row %d has %d entries (at most 255 per row)", g, len);
    meta0[a] = lrp | ((uint32_t)dr << 16) | ((uint32_t)len << 24);
    lrp += (uint32_t)len;
  }
  // Record order.  Elements whose vertices are all owned come first (a wave
  // of them runs a branch-free path), then the rest grouped by owned mask.
  // Inside a group, element j*q + k (q = ceil(n/64)) goes to position 64*k + j,
  // so the 64 lanes of a wave take elements q apart: neighbouring elements
  // share nodes, and spreading them keeps a wave's LDS atomics on distinct rows.
  {
    std::vector<std::pair<uint32_t, int32_t>> bym(elems.size());
    for (size_t i = 0; i < elems.size(); ++i) {
      uint32_t mask = 0;
      for (int a = 0; a < c.nb; ++a)
        if (s.local[c.conn[(int64_t)elems[i] * c.nb + a]] < nown) mask |= 1u << a;
      const uint32_t full = (1u << c.nb) - 1;
      // records list their owned vertices first (below), so the owned count
      // alone fixes a record's branch pattern in the kernel
      bym[i] = {mask == full ? 0u : 16u - (uint32_t)__builtin_popcount(mask) * 4u, elems[i]};
    }
    std::stable_sort(bym.begin(), bym.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    size_t g0 = 0;
    while (g0 < bym.size()) {
      size_t g1 = g0;
      while (g1 < bym.size() && bym[g1].first == bym[g0].first) ++g1;
      const size_t ne = g1 - g0, qq = (ne + 63) / 64;
      std::vector<std::pair<size_t, int32_t>> key(ne);
      for (size_t i = 0; i < ne; ++i) key[i] = {(i % qq) * 64 + i / qq, bym[g0 + i].second};
      std::sort(key.begin(), key.end());
      for (size_t i = 0; i < ne; ++i) elems[g0 + i] = key[i].second;
      g0 = g1;
    }
  }
  // Element records (16 bytes): nb x 10-bit LDS slots, then for each owned
  // vertex a the in-row ranks of the other vertices (6 bits each, a-major, b
  // ascending, from bit 40), and at bit 112 the element-local index of each
  // record vertex (2 bits each).  Two layout choices cut LDS bank conflicts of the kernels; both
  // only relabel, so the assembled values do not change:
  //  * vertex order per record: a wave issues one LDS atomic per ordered pair
  //    (a, b) for all 64 lanes; the order is picked greedily, lane by lane, so
  //    that each atomic's 8-byte addresses of a half-wave spread over the 32
  //    bank pairs (P1 element matrices are symmetric under relabelling);
  //  * LDS slots of the halo nodes: node data are read as 16-byte words, so a
  //    quarter-wave's reads of one vertex position are conflict-free when
  //    their slots differ mod 16; slots are coloured greedily (owned nodes stay
  //    in [0, nown), the others in [nown, nh)).
  const int nrec = (int)elems.size();
  const int NB = c.nb;
  std::vector<int32_t> vl((size_t)nrec * NB);  // halo-order local ids, final vertex order
  std::vector<uint8_t> vr((size_t)nrec * 16, 0);  // ranks in final order, (a, j) a-major
  std::vector<uint8_t> vo((size_t)nrec, 0);       // element-local index of each record vertex, 2 bits each
  {
    int perms[24][4];
    int np = 0;
    int pv[4] = {0, 1, 2, 3};
    do {
      for (int k = 0; k < 4; ++k) perms[np][k] = pv[k];
      ++np;
    } while (std::next_permutation(pv, pv + NB));
    // instruction order of the kernels' atomics: (a, b), (b, a) for a < b
    int sx[12], sy[12], ns = 0;
    for (int a = 0; a < NB; ++a)
      for (int b = a + 1; b < NB; ++b) {
        sx[ns] = a; sy[ns] = b; ++ns;
        sx[ns] = b; sy[ns] = a; ++ns;
      }
    std::vector<int32_t> hist((size_t)ns * 2 * 32);
    int rk[4][4];
    int32_t l0[4];
    for (int w0 = 0; w0 < nrec; w0 += 64) {
      std::fill(hist.begin(), hist.end(), 0);
      for (int r = w0; r < std::min(nrec, w0 + 64); ++r) {
        const int64_t* ce = c.conn + (int64_t)elems[r] * NB;
        for (int a = 0; a < NB; ++a) l0[a] = s.local[ce[a]];
        for (int a = 0; a < NB; ++a) {
          if (l0[a] >= nown) continue;
          for (int b = 0; b < NB; ++b) {
            if (b == a) continue;
            int32_t rr = find_in_row(c.row_ptr, c.col_idx, ce[a], ce[b]);
            TM_REQUIRE(rr >= 0 && rr < kMaxRank, TM_ERR_LIMIT,
                       "row %lld: column %lld missing or rank >= %d", (long long)ce[a],
                       (long long)ce[b], kMaxRank);
            rk[a][b] = rr;
          }
        }
        const int half = ((r - w0) >> 5) & 1;
        int n_own_v = 0;
        for (int a = 0; a < NB; ++a) n_own_v += l0[a] < nown ? 1 : 0;
        int best = -1, best_sc = 1 << 30;
        for (int q = 0; q < np; ++q) {
          // owned vertices take the first positions: every lane of a wave of
          // such records then follows the same branches in the kernel
          bool owned_first = true;
          for (int x = 0; x < NB; ++x) owned_first = owned_first && ((l0[perms[q][x]] < nown) == (x < n_own_v));
          if (!owned_first) continue;
          int sc = 0;
          for (int k = 0; k < ns; ++k) {
            const int a = perms[q][sx[k]], b = perms[q][sy[k]];
            if (l0[a] >= nown) continue;
            const int bank = (int)(((meta0[l0[a]] & 0xffffu) + (uint32_t)rk[a][b]) & 31u);
            sc += hist[((size_t)k * 2 + half) * 32 + bank];
          }
          if (sc < best_sc) {
            best_sc = sc;
            best = q;
          }
        }
        const int* pq = perms[best];
        uint8_t code = 0;
        for (int a = 0; a < NB; ++a) code |= (uint8_t)(pq[a] << (2 * a));
        vo[(size_t)r] = code;
        for (int k = 0; k < ns; ++k) {
          const int a = pq[sx[k]], b = pq[sy[k]];
          if (l0[a] >= nown) continue;
          const int bank = (int)(((meta0[l0[a]] & 0xffffu) + (uint32_t)rk[a][b]) & 31u);
          ++hist[((size_t)k * 2 + half) * 32 + bank];
        }
        for (int a = 0; a < NB; ++a) {
          vl[(size_t)r * NB + a] = l0[pq[a]];
          if (l0[pq[a]] >= nown) continue;
          int j = 0;
          for (int b = 0; b < NB; ++b)
            if (b != a) vr[(size_t)r * 16 + a * (NB - 1) + j++] = (uint8_t)rk[pq[a]][pq[b]];
        }
      }
    }
  }
  std::vector<int32_t> slot(nh);
  {
    // groups: (quarter-wave, vertex position); membership lists per node
    const int ngrp = ((nrec + 15) / 16) * NB;
    std::vector<std::vector<int32_t>> memb(nh);
    for (int r = 0; r < nrec; ++r)
      for (int a = 0; a < NB; ++a) {
        const int g = (r / 16) * NB + a;
        auto& m = memb[vl[(size_t)r * NB + a]];
        if (m.empty() || m.back() != g) m.push_back(g);
      }
    std::vector<int32_t> gcol((size_t)ngrp * 16, 0);
    auto colour = [&](int lo, int hi) {
      int cap[16];
      for (int k = 0; k < 16; ++k) cap[k] = 0;
      for (int p = lo; p < hi; ++p) ++cap[p & 15];
      std::vector<int32_t> order(hi - lo);
      std::iota(order.begin(), order.end(), lo);
      std::stable_sort(order.begin(), order.end(),
                       [&](int32_t x, int32_t y) { return memb[x].size() > memb[y].size(); });
      std::vector<int32_t> col(hi - lo);
      for (int32_t x : order) {
        int best = -1, best_sc = 1 << 30;
        for (int k = 0; k < 16; ++k) {
          if (cap[k] == 0) continue;
          int sc = 0;
          for (int32_t g : memb[x]) sc += gcol[(size_t)g * 16 + k];
          if (sc < best_sc) {
            best_sc = sc;
            best = k;
          }
        }
        --cap[best];
        col[x - lo] = best;
        for (int32_t g : memb[x]) ++gcol[(size_t)g * 16 + best];
      }
      // positions p in [lo, hi) with p % 16 == colour, in increasing order
      int nextp[16];
      for (int k = 0; k < 16; ++k) {
        int p = lo;
        while (p < hi && (p & 15) != k) ++p;
        nextp[k] = p;
      }
      for (int x = lo; x < hi; ++x) {
        const int k = col[x - lo];
        slot[x] = nextp[k];
        nextp[k] += 16;
      }
    };
    colour(0, nown);
    colour(nown, (int)nh);
  }
  B.halo.assign(nh, -1);
  B.meta.assign(nh, 0u);
  for (int h = 0; h < nh; ++h) {
    B.halo[slot[h]] = h < nown ? owned[h] : others[h - nown];
    B.meta[slot[h]] = meta0[h];
  }
  B.recs.reserve((size_t)nrec * 4);
  B.rec_elem.reserve(nrec);
  for (int r = 0; r < nrec; ++r) {
    uint64_t lo = 0, hi = 0;
    for (int a = 0; a < NB; ++a) lo |= (uint64_t)slot[vl[(size_t)r * NB + a]] << (10 * a);
    for (int a = 0; a < NB; ++a) {
      if (vl[(size_t)r * NB + a] >= nown) continue;
      for (int j = 0; j < NB - 1; ++j) {
        const uint64_t v = vr[(size_t)r * 16 + a * (NB - 1) + j];
        const int bit = 40 + 6 * (a * (NB - 1) + j);
        if (bit < 64)
          lo |= v << bit;
        else
          hi |= v << (bit - 64);
      }
    }
    // bits 112-119: record vertex a is vertex (code >> 2a) & 3 of the element
    // (kernels that read per-element data in the element's own vertex order)
    hi |= (uint64_t)vo[(size_t)r] << (112 - 64);
    B.recs.push_back((uint32_t)lo);
    B.recs.push_back((uint32_t)(lo >> 32));
    B.recs.push_back((uint32_t)hi);
    B.recs.push_back((uint32_t)(hi >> 32));
    B.rec_elem.push_back(elems[r]);
  }
  // runs of consecutive owned rows -> contiguous value ranges
  int a = 0;
  uint32_t acc0 = 0;
  while (a < nown) {
    int b = a;
    uint32_t len = (uint32_t)(c.row_ptr[owned[a] + 1] - c.row_ptr[owned[a]]);
    while (b + 1 < nown && owned[b + 1] == owned[b] + 1) {
      ++b;
      len += (uint32_t)(c.row_ptr[owned[b] + 1] - c.row_ptr[owned[b]]);
    }
    B.runs.push_back(c.row_ptr[owned[a]]);
    B.runs.push_back((int32_t)acc0);
    B.runs.push_back((int32_t)len);
    B.runs.push_back(owned[a]);
    acc0 += len;
    a = b + 1;
  }
  out.push_back(std::move(B));
  return TM_OK;
}

}  // namespace

extern "C" int tm_plan_build(const double* points, int dim, const int64_t* conn, int64_t E, int nb,
                             int64_t N, const int32_t* row_ptr, const int32_t* col_idx,
                             int max_rows, int lds_budget, int entry_bytes, int node_bytes,
                             int own_bytes, int n_threads, tm_plan** out) {
  TM_REQUIRE(out != nullptr, TM_ERR_ARG, "null output");
  *out = nullptr;
  TM_REQUIRE(dim >= 1 && dim <= 3 && nb >= 2 && nb <= 4 && max_rows >= 1, TM_ERR_ARG,
             "bad plan parameters dim=%d nb=%d max_rows=%d", dim, nb, max_rows);
  Incidence inc;
  int st = build_incidence(conn, E, nb, N, inc);
  if (st) return st;
  PlanCtx c{points, dim, conn, E, N, nb, row_ptr, col_idx, &inc,
            max_rows, lds_budget, entry_bytes, node_bytes, own_bytes};

  // recursive coordinate bisection into leaves of <= max_rows nodes (DFS order)
  std::vector<int32_t> ids(N);
  std::iota(ids.begin(), ids.end(), 0);
  std::vector<std::pair<size_t, size_t>> leaves, stack;
  if (N > 0) stack.push_back({0, (size_t)N});
  while (!stack.empty()) {
    auto [lo, hi] = stack.back();
    stack.pop_back();
    if (hi - lo <= (size_t)max_rows) {
      leaves.push_back({lo, hi});
      continue;
    }
    size_t mid = rcb_split(c, ids, lo, hi, max_rows);
    stack.push_back({mid, hi});
    stack.push_back({lo, mid});
  }

  const int nt = std::max(1, std::min(pick_threads(n_threads), (int)std::max<size_t>(1, leaves.size())));
  std::vector<std::vector<PlanBlock>> res(leaves.size());
  std::atomic<size_t> next{0};
  std::atomic<int> err{0};
  std::string err_msg;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    th.emplace_back([&] {
      Scratch s(c);
      for (;;) {
        size_t k = next++;
        if (k >= leaves.size() || err) break;
        std::vector<int32_t> owned(ids.begin() + leaves[k].first, ids.begin() + leaves[k].second);
        int r = make_blocks(c, s, std::move(owned), res[k]);
        if (r && !err.exchange(r)) err_msg = tm_last_error();
      }
    });
  }
  for (auto& x : th) x.join();
  if (err) {
    tmx::set_error("%s", err_msg.c_str());
    return err;
  }

  auto* P = new tm_plan();
  int64_t nblk = 0, nh = 0, nr = 0, nrun = 0;
  for (auto& v : res)
    for (auto& b : v) {
      ++nblk;
      nh += b.halo.size();
      nr += b.rec_elem.size();
      nrun += b.runs.size() / 4;
    }
  if (nh >= (int64_t(1) << 31) || nr >= (int64_t(1) << 31)) {
    delete P;
    TM_REQUIRE(false, TM_ERR_LIMIT, "plan too large for int32 offsets");
  }
  P->blk.reserve(nblk * 8);
  P->halo.reserve(nh);
  P->meta.reserve(nh);
  P->recs.reserve(nr * 4);
  P->rec_elem.reserve(nr);
  P->runs.reserve(nrun * 4);
  int64_t max_lds = 0, max_halo = 0, max_rec = 0, max_acc = 0;
  for (auto& v : res)
    for (auto& b : v) {
      int32_t n_rec = (int32_t)b.rec_elem.size();
      int32_t n_run = (int32_t)(b.runs.size() / 4);
      P->blk.insert(P->blk.end(),
                    {(int32_t)P->halo.size(), (int32_t)b.halo.size(), b.n_own,
                     (int32_t)P->rec_elem.size(), n_rec, (int32_t)(P->runs.size() / 4), n_run,
                     b.n_acc});
      P->halo.insert(P->halo.end(), b.halo.begin(), b.halo.end());
      P->meta.insert(P->meta.end(), b.meta.begin(), b.meta.end());
      P->recs.insert(P->recs.end(), b.recs.begin(), b.recs.end());
      P->rec_elem.insert(P->rec_elem.end(), b.rec_elem.begin(), b.rec_elem.end());
      P->runs.insert(P->runs.end(), b.runs.begin(), b.runs.end());
      max_lds = std::max<int64_t>(max_lds, b.lds);
      max_halo = std::max<int64_t>(max_halo, b.halo.size());
      max_rec = std::max<int64_t>(max_rec, n_rec);
      max_acc = std::max<int64_t>(max_acc, b.n_acc);
    }
  P->info[0] = nblk;
  P->info[1] = nh;
  P->info[2] = nr;
  P->info[3] = nrun;
  P->info[4] = max_lds;
  P->info[5] = max_halo;
  P->info[6] = max_rec;
  P->info[7] = max_acc;
  *out = P;
  return TM_OK;
}

extern "C" int tm_plan_info(const tm_plan* p, int64_t* info) {
  TM_REQUIRE(p && info, TM_ERR_ARG, "null plan");
  std::memcpy(info, p->info, sizeof(p->info));
  return TM_OK;
}

extern "C" int tm_plan_export(const tm_plan* p, int32_t* blk, int32_t* halo, uint32_t* meta,
                              uint32_t* recs, int32_t* rec_elem, int32_t* runs) {
  TM_REQUIRE(p, TM_ERR_ARG, "null plan");
  if (blk) std::memcpy(blk, p->blk.data(), p->blk.size() * 4);
  if (halo) std::memcpy(halo, p->halo.data(), p->halo.size() * 4);
  if (meta) std::memcpy(meta, p->meta.data(), p->meta.size() * 4);
  if (recs) std::memcpy(recs, p->recs.data(), p->recs.size() * 4);
  if (rec_elem) std::memcpy(rec_elem, p->rec_elem.data(), p->rec_elem.size() * 4);
  if (runs) std::memcpy(runs, p->runs.data(), p->runs.size() * 4);
  return TM_OK;
}

extern "C" void tm_plan_destroy(tm_plan* p) { delete p; }

// ------------------------------------------------------------- edge plan ----
//
// Edge records for the symmetric scalar matrix kernels.  A symmetric element
// matrix needs one value per vertex pair, so a block can accumulate each
// undirected edge once: in the row of its owned end with the smaller LDS slot
// (records list owned vertices first, by ascending slot), at the rank of the
// other end in that row.  Each pair (i, j), i < j, then needs one rank field
// instead of two, and the record shrinks from 16 to 8 bytes when every rank
// fits 4 bits and every halo 512 nodes (12 bytes otherwise).  The transposed
// entry of an edge whose two ends are both owned is filled after the element
// loop from a per-block mirror list {src | dst << 16} of accumulator indices.
//
// Layout the edge kernel stages (and the records address): node slot s keeps
// its two 16-byte words at LDS words w(s) and w(s) + 16, w(s) = s + (s & ~15),
// i.e. groups of 16 slots hold 16 first words then 16 second words, so a
// record carries w(s) and both reads share one address.  Per block:
// eblk = {mirror offset, mirror count | e2 << 16, e0 | e1 << 16, row offset}:
// e0, e1, e2 end the record groups with nb, nb-1 and nb-2 owned vertices (the
// plan sorts a block's records by descending owned count); the block's owned
// rows, in accumulator order, are rows[row offset + i] = {first value index,
// accumulator byte offset | row length << 16}, for stores of 16 lanes per row.

struct tm_plan_edge {
  std::vector<uint32_t> recs;
  std::vector<int32_t> eblk;
  std::vector<uint32_t> mir;
  std::vector<int32_t> rows;
  int64_t info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

namespace {

// Rank of record vertex b in the row of owned record vertex a (16-byte record).
inline uint32_t rec16_rank(const uint32_t* r, int nb, int a, int b) {
  const int j = b - (b > a ? 1 : 0);
  const int bit = 40 + 6 * (a * (nb - 1) + j);
  const uint64_t lo = (uint64_t)r[0] | ((uint64_t)r[1] << 32), hi = (uint64_t)r[2] | ((uint64_t)r[3] << 32);
  const uint64_t v = bit < 64 ? (lo >> bit) | (bit > 58 ? hi << (64 - bit) : 0) : hi >> (bit - 64);
  return (uint32_t)(v & 63u);
}

inline uint32_t node_word(uint32_t s) { return s + (s & ~15u); }

}  // namespace

extern "C" int tm_plan_edge_build(int nb, int64_t n_blocks, const int32_t* blk, const uint32_t* meta,
                                  const uint32_t* recs, const int32_t* runs, int n_threads, tm_plan_edge** out) {
  TM_REQUIRE(out != nullptr, TM_ERR_ARG, "null output");
  *out = nullptr;
  TM_REQUIRE(nb == 3 || nb == 4, TM_ERR_ARG, "edge records cover P1 triangles and tetrahedra (nb=%d)", nb);
  TM_REQUIRE(n_blocks >= 0 && (n_blocks == 0 || (blk && meta && recs && runs)), TM_ERR_ARG, "null plan array");
  const int npair = nb * (nb - 1) / 2;
  int pi[6], pj[6];
  for (int i = 0, k = 0; i < nb; ++i)
    for (int j = i + 1; j < nb; ++j, ++k) {
      pi[k] = i;
      pj[k] = j;
    }
  // record width: tetrahedra take 10-bit node words and 4-bit ranks (8 bytes)
  // when every rank an owned row needs is below 16 and every halo has at most
  // 512 nodes, 11-bit words and 6-bit ranks (12 bytes) otherwise; triangles
  // 11-bit words and 6-bit ranks (8 bytes)
  uint32_t max_rank = 0;
  int64_t max_nh = 0, max_lds = 0;
  for (int64_t b = 0; b < n_blocks; ++b) {
    const int32_t* bi = blk + 8 * b;
    const int nown = bi[2];
    max_nh = std::max<int64_t>(max_nh, bi[1]);
    TM_REQUIRE(bi[7] < 8192 && bi[4] < 65536, TM_ERR_LIMIT,
               "block %lld: %d accumulators, %d records (edge records take < 8192, < 65536)", (long long)b, bi[7],
               bi[4]);
    max_lds = std::max<int64_t>(max_lds, ((8 * (int64_t)bi[7] + 15) & ~int64_t(15)) + 512 * (((int64_t)bi[1] + 15) / 16) +
                                             8 * (int64_t)bi[2]);
    int prev = nb;
    for (int64_t r = bi[3]; r < (int64_t)bi[3] + bi[4]; ++r) {
      const uint32_t* rc = recs + 4 * r;
      const uint64_t lo = (uint64_t)rc[0] | ((uint64_t)rc[1] << 32);
      int no = 0;
      for (int a = 0; a < nb; ++a) {
        if ((int)((lo >> (10 * a)) & 1023u) >= nown) continue;
        ++no;
        for (int c = 0; c < nb; ++c)
          if (c != a) max_rank = std::max(max_rank, rec16_rank(rc, nb, a, c));
      }
      TM_REQUIRE(no >= 1 && no <= prev, TM_ERR_ARG,
                 "block %lld: records not grouped by descending owned count", (long long)b);
      prev = no;
    }
  }
  TM_REQUIRE(max_lds <= 65536, TM_ERR_LIMIT, "edge-record blocks need %lld bytes of LDS", (long long)max_lds);
  int64_t all_rows = 0;
  for (int64_t b = 0; b < n_blocks; ++b) all_rows += blk[8 * b + 2];
  TM_REQUIRE(all_rows < (int64_t(1) << 30), TM_ERR_LIMIT, "too many owned rows");
  const int W = (nb == 4 && (max_rank > 15 || max_nh > 512)) ? 3 : 2;
  const int SB = (nb == 4 && W == 2) ? 10 : 11;
  const int R = (nb == 4 && W == 2) ? 4 : 6;
  auto rank_pos = [&](int k, int& word, int& bit) {
    if (nb == 4 && W == 2) {
      word = 1;
      bit = 8 + 4 * k;
    } else if (k < 3) {
      word = 1;
      bit = 12 + 6 * k;
    } else {
      word = 2;
      bit = 6 * (k - 3);
    }
  };

  auto* P = new tm_plan_edge();
  int64_t nrec = n_blocks > 0 ? (int64_t)blk[8 * (n_blocks - 1) + 3] + blk[8 * (n_blocks - 1) + 4] : 0;
  P->recs.assign((size_t)nrec * W, 0u);
  P->eblk.assign((size_t)n_blocks * 4, 0);
  // owned-row table: blocks' rows in accumulator order (the runs' order)
  int64_t nrows = 0;
  for (int64_t b = 0; b < n_blocks; ++b) {
    P->eblk[(size_t)b * 4 + 3] = (int32_t)nrows;
    nrows += blk[8 * b + 2];
  }
  P->rows.assign((size_t)nrows * 2, 0);
  std::vector<std::vector<uint32_t>> mirs((size_t)n_blocks);
  std::atomic<int64_t> next{0};
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(pick_threads(n_threads), std::max<int64_t>(1, n_blocks / 64)));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    th.emplace_back([&] {
      std::vector<uint64_t> keys;
      for (;;) {
        const int64_t b0 = next.fetch_add(64);
        if (b0 >= n_blocks) break;
        for (int64_t b = b0; b < std::min(n_blocks, b0 + 64); ++b) {
          const int32_t* bi = blk + 8 * b;
          const int hoff = bi[0], nown = bi[2];
          int ge[3] = {0, 0, 0};  // records with at least nb, nb-1, nb-2 owned vertices
          keys.clear();
          for (int64_t r = bi[3]; r < (int64_t)bi[3] + bi[4]; ++r) {
            const uint32_t* rc = recs + 4 * r;
            const uint64_t lo = (uint64_t)rc[0] | ((uint64_t)rc[1] << 32);
            int sl[4], ord[4], no = 0;
            for (int a = 0; a < nb; ++a) sl[a] = (int)((lo >> (10 * a)) & 1023u);
            // owned vertices by ascending slot, then the others in record order
            for (int a = 0; a < nb; ++a)
              if (sl[a] < nown) ord[no++] = a;
            std::sort(ord, ord + no, [&](int x, int y) { return sl[x] < sl[y]; });
            for (int a = 0, k = no; a < nb; ++a)
              if (sl[a] >= nown) ord[k++] = a;
            for (int g = 0; g < 3; ++g) ge[g] += no >= nb - g ? 1 : 0;
            uint32_t w[3] = {0, 0, 0};
            uint64_t words = 0;
            for (int i = 0; i < nb; ++i) words |= (uint64_t)node_word((uint32_t)sl[ord[i]]) << (SB * i);
            w[0] = (uint32_t)words;
            w[1] = (uint32_t)(words >> 32);
            for (int k = 0; k < npair; ++k) {
              const int a = ord[pi[k]], c = ord[pj[k]];
              if (pi[k] >= no) continue;  // neither end owned
              const uint32_t rk = rec16_rank(rc, nb, a, c);
              int word, bit;
              rank_pos(k, word, bit);
              w[word] |= rk << bit;
              if (pj[k] < no) {  // both ends owned: the transposed entry is mirrored
                const uint32_t src = (meta[hoff + sl[a]] & 0xffffu) + rk;
                const uint32_t dst = (meta[hoff + sl[c]] & 0xffffu) + rec16_rank(rc, nb, c, a);
                keys.push_back(((uint64_t)dst << 32) | src);
              }
            }
            for (int x = 0; x < W; ++x) P->recs[(size_t)r * W + x] = w[x];
          }
          std::sort(keys.begin(), keys.end());
          keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
          auto& m = mirs[(size_t)b];
          m.reserve(keys.size());
          for (uint64_t k : keys) m.push_back((uint32_t)(k & 0xffffu) | ((uint32_t)(k >> 32) << 16));
          int32_t* eb = P->eblk.data() + 4 * b;
          eb[1] = (int32_t)keys.size() | (ge[2] << 16);
          eb[2] = ge[0] | (ge[1] << 16);
          // rows by accumulator offset: each run's rows are consecutive in both
          // the accumulators and the values
          std::vector<std::pair<uint32_t, uint32_t>> ro;  // (acc offset, length) of owned slots
          for (int a = 0; a < nown; ++a) {
            const uint32_t m = meta[hoff + a];
            ro.push_back({m & 0xffffu, m >> 24});
          }
          std::sort(ro.begin(), ro.end());
          int32_t* rw = P->rows.data() + 2 * (size_t)eb[3];
          const int32_t* ru = runs + 4 * (int64_t)bi[5];
          int q = 0;
          for (int i = 0; i < nown; ++i) {
            const uint32_t off = ro[i].first, len = ro[i].second;
            while (q < bi[6] && (uint32_t)(ru[4 * q + 1] + ru[4 * q + 2]) <= off && len > 0) ++q;
            const int32_t vs = (len > 0 && q < bi[6]) ? ru[4 * q] + (int32_t)(off - (uint32_t)ru[4 * q + 1]) : 0;
            rw[2 * i] = vs;
            rw[2 * i + 1] = (int32_t)((off * 8) | (len << 16));
          }
        }
      }
    });
  }
  for (auto& x : th) x.join();
  int64_t tot = 0, mx = 0;
  for (int64_t b = 0; b < n_blocks; ++b) {
    if (tot + (int64_t)mirs[(size_t)b].size() >= (int64_t(1) << 31)) {
      delete P;
      TM_REQUIRE(false, TM_ERR_LIMIT, "mirror list too large for int32 offsets");
    }
    P->eblk[(size_t)b * 4] = (int32_t)tot;
    tot += (int64_t)mirs[(size_t)b].size();
    mx = std::max<int64_t>(mx, (int64_t)mirs[(size_t)b].size());
  }
  P->mir.reserve((size_t)tot);
  for (auto& m : mirs) P->mir.insert(P->mir.end(), m.begin(), m.end());
  P->info[0] = W;
  P->info[1] = tot;
  P->info[2] = mx;
  P->info[3] = max_rank;
  P->info[4] = max_lds;
  P->info[5] = SB;
  P->info[6] = nrows;
  *out = P;
  return TM_OK;
}

extern "C" int tm_plan_edge_info(const tm_plan_edge* p, int64_t* info) {
  TM_REQUIRE(p && info, TM_ERR_ARG, "null edge plan");
  std::memcpy(info, p->info, sizeof(p->info));
  return TM_OK;
}

extern "C" int tm_plan_edge_export(const tm_plan_edge* p, uint32_t* recs, int32_t* eblk, uint32_t* mir,
                                   int32_t* rows) {
  TM_REQUIRE(p, TM_ERR_ARG, "null edge plan");
  if (recs) std::memcpy(recs, p->recs.data(), p->recs.size() * 4);
  if (eblk) std::memcpy(eblk, p->eblk.data(), p->eblk.size() * 4);
  if (mir) std::memcpy(mir, p->mir.data(), p->mir.size() * 4);
  if (rows) std::memcpy(rows, p->rows.data(), p->rows.size() * 4);
  return TM_OK;
}

extern "C" void tm_plan_edge_destroy(tm_plan_edge* p) { delete p; }
