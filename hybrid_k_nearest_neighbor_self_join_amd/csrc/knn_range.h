// knn_range.h -- device helpers shared by the range kernels (select.hip, range.hip).
//
// Per-query search geometry on the grid: a query with radius r scans the cells
// [coord(x - r, rounded down), coord(x + r, rounded up)] per indexed dim (kernel.cu:61-110 scans
// the 3^NUMINDEXEDDIM adjacent cells of a width-eps grid; here r may span several cells).
// Within one row of a block (indexed dims 1.. fixed) the cells of dim 0 are adjacent linear
// ids, i.e. one contiguous range of cell-sorted points.
//
// ND is the number of indexed dims when known at compile time (1..4), or 0 for the grid's
// runtime value (up to KNN_MAX_INDEXED).  Every loop over indexed dims runs over a constant
// bound (ND, or KNN_MAX_INDEXED with an exit at the runtime count) and is fully unrolled, so
// the per-dim arrays live in registers rather than in private scratch memory.
#pragma once

#include <type_traits>

#include "knn_launch.h"

namespace knn {

// Select kernel layout (select.hip): SEL_NB buckets of the FP64 key (the last one for keys
// beyond the query's threshold), two 16-bit counters per LDS word; after the histogram the
// same LDS region H holds each query's bucket cursors as a row of SEL_RS words (one word of
// padding against bank conflicts).  Sub-blocks of at most SEL_RMAX rows have their row ranges
// looked up for the whole tile at once.
#ifndef KNN_SELECT_BUCKETS
#define KNN_SELECT_BUCKETS KNN_SELECT_BUCKETS_OTHER
#endif
#ifndef KNN_SELECT_RMAX
#define KNN_SELECT_RMAX 5
#endif
constexpr int SEL_NB = KNN_SELECT_BUCKETS;
static_assert(SEL_NB % 2 == 0 && SEL_NB >= 8 && SEL_NB <= 64, "select buckets: even, 8..64");
constexpr int SEL_NBIN = SEL_NB - 1;
constexpr uint32_t SEL_RMAX = KNN_SELECT_RMAX;
constexpr uint32_t SEL_RS = SEL_NB / 2 + 1;
// LDS words of the select kernel with nb buckets before its staging arrays: H (the histogram's
// nb x 32 words, or 64 rows of nb / 2 + 1 cursor words), rowB/rowX[64], prB/prX[SEL_RMAX][64],
// prT/prR[64], the collecting query's 32-bit cursors qc[64]
constexpr uint32_t sel_hwords(uint32_t nb) { return nb * 32 > 64 * (nb / 2 + 1) ? nb * 32 : 64 * (nb / 2 + 1); }
constexpr uint32_t sel_fixed_words(uint32_t nb) { return sel_hwords(nb) + 128 + SEL_RMAX * 128 + 128 + 64; }
constexpr uint32_t SEL_HWORDS = sel_hwords(SEL_NB);
constexpr uint32_t SEL_FIXED_WORDS = sel_fixed_words(SEL_NB);

// persistent grid of the select kernel when the tile count is known on the host: 11 one-wave
// workgroups per CU on 256 CUs
constexpr uint32_t SELECT_MAX_GRID = 2816;

// Select kernel launch for D data dims and nd indexed dims, built with NB buckets (select.hip,
// compiled once per bucket count).
template <typename T, int D, int NB>
void select_launch(const JoinArgs &a, const typename GridOf<T>::type &g, uint32_t nd, hipStream_t s);

namespace {

template <int ND>
__device__ __forceinline__ constexpr uint32_t nd_cap() {
  return ND > 0 ? (uint32_t)ND : (uint32_t)KNN_MAX_INDEXED;
}
template <int ND, typename G>
__device__ __forceinline__ uint32_t nd_of(const G &g) {
  return ND > 0 ? (uint32_t)ND : g.nd;
}

__device__ __forceinline__ bool kless(double ka, uint32_t ia, double kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

__device__ __forceinline__ uint32_t lower_bound_u64(const uint64_t *a, uint32_t n, uint64_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t m = (lo + hi) >> 1;
    if (a[m] < v) lo = m + 1; else hi = m;
  }
  return lo;
}
__device__ __forceinline__ uint32_t upper_bound_u64(const uint64_t *a, uint32_t n, uint64_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t m = (lo + hi) >> 1;
    if (a[m] <= v) lo = m + 1; else hi = m;
  }
  return lo;
}

// Reported-distance arithmetic of the input type and the filter margin that bounds it
// against the FP64 key: the T sum s = sum_j (c_j - q_j)^2 (reported distance sqrt(s)) is within
// 2^-19 of the FP64 key relative, so a candidate whose key could be <= X has s <= X (1 + 2^-16).
template <typename T> struct Num;
template <> struct Num<float> {
  template <int D>
  __device__ static __forceinline__ float sq(const float *c, const float *q) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < D; j++) {
      float df = c[j] - q[j];
      float t = df * df;
      s = s + t;
    }
    return s;
  }
  __device__ static __forceinline__ float bound(double key) { return (float)(key * (1.0 + 0x1p-16)); }
};
template <> struct Num<double> {
  template <int D>
  __device__ static __forceinline__ double sq(const double *c, const double *q) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < D; j++) {
      double df = c[j] - q[j];
      double t = df * df;
      s = s + t;
    }
    return s;
  }
  __device__ static __forceinline__ double bound(double key) { return key; }
};

// The ranking key: sum_j ((double)c_j - (double)q_j)^2, each product and sum rounded to FP64.
template <int D, typename T>
__device__ __forceinline__ double fp64_key(const T *c, const double *qd) {
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < D; j++) {
    double dx = (double)c[j] - qd[j];
    double t = dx * dx;
    s = s + t;
  }
  return s;
}

template <int D>
__device__ __forceinline__ double pick(const double *v, uint32_t j) {
  double r = v[0];
#pragma unroll
  for (int jj = 1; jj < D; jj++)
    if ((uint32_t)jj == j) r = v[jj];
  return r;
}
template <typename T, int D>
__device__ __forceinline__ T pick_t(const T *v, uint32_t j) {
  T r = v[0];
#pragma unroll
  for (int jj = 1; jj < D; jj++)
    if ((uint32_t)jj == j) r = v[jj];
  return r;
}

// Per-lane state of one tile.
template <typename T, int D>
struct Lane {
  bool active;
  uint32_t q;
  T qc[D];
  double qd[D];
  double Tq;  // success threshold on the key
  T F;        // filter bound on the T sum
};

// Query i < tl.z of a tile: the first run act[tl.y ..) holds tl.w of them, the second run
// act[tl.x ..) the rest.
__device__ __forceinline__ uint32_t tile_query(const JoinArgs &a, const uint4 &tl, uint32_t i) {
  return a.act[i < tl.w ? tl.y + i : tl.x + (i - tl.w)];
}

// Lane = query of the tile: its coordinates, radius r (its own, or the pass's one radius) and
// Tq = r^2 (1 - 2^-20); lo/hi = the wave's union of the lanes' cell ranges per indexed dim.
template <typename T, int D, int ND, typename G>
__device__ __forceinline__ void tile_setup(const JoinArgs &a, const G &g, const uint4 &tl, uint32_t lane,
                                           Lane<T, D> &L, uint32_t *lo, uint32_t *hi) {
  const T *__restrict__ x = reinterpret_cast<const T *>(a.x);
  const size_t n = a.n;
  L.active = lane < tl.z;
  L.q = tile_query(a, tl, L.active ? lane : 0u);
#pragma unroll
  for (int j = 0; j < D; j++) {
    L.qc[j] = x[(size_t)j * n + L.q];
    L.qd[j] = (double)L.qc[j];
  }
  const T r = a.radius ? reinterpret_cast<const T *>(a.radius)[L.q] : (T)a.radiusAll;
  L.Tq = (double)r * (double)r * (1.0 - 0x1p-20);
  L.F = Num<T>::bound(L.Tq);
  const T eps = (T)g.eps;
  const uint32_t nd = nd_of<ND>(g);
#pragma unroll
  for (uint32_t j = 0; j < nd_cap<ND>(); j++) {
    if (j >= nd) break;
    const T xj = pick_t<T, D>(L.qc, j);
    uint32_t l = cell_coord<T>(next_down(xj - r), (T)g.minArr[j], eps, g.nCells[j]);
    uint32_t u = cell_coord<T>(next_up(xj + r), (T)g.minArr[j], eps, g.nCells[j]);
    if (!L.active) {
      l = 0xffffffffu;
      u = 0u;
    }
    lo[j] = wave_min_u32(l);
    hi[j] = wave_max_u32(u);
  }
}

// Point range [pBeg, pEnd) of dim-0 cells [l0, h0] of the row at rowBase (dense start table,
// or binary search of the compacted cell table).
__device__ __forceinline__ void row_range(const JoinArgs &a, uint64_t rowBase, uint32_t l0, uint32_t h0,
                                          uint32_t &pBeg, uint32_t &pEnd) {
  if (a.denseStart) {
    pBeg = a.denseStart[rowBase + l0];
    pEnd = a.denseStart[rowBase + h0 + 1];
  } else {
    const uint32_t cbeg = lower_bound_u64(a.cellKeys, a.ncell, rowBase + l0);
    const uint32_t cend = upper_bound_u64(a.cellKeys, a.ncell, rowBase + h0);
    pBeg = a.cellStart[cbeg];
    pEnd = a.cellStart[cend];
  }
}

// Advances the odometer cur over indexed dims 1..nd-1 within [lo, hi]; false when done.
// Written as selects on every digit (no early exit), so each element keeps a fixed index and
// the array stays in registers.
template <int ND>
__device__ __forceinline__ bool next_row(uint32_t *cur, const uint32_t *lo, const uint32_t *hi, uint32_t nd) {
  bool carry = true;
#pragma unroll
  for (uint32_t j = 1; j < nd_cap<ND>(); j++) {
    if (j >= nd) break;
    const bool wrap = cur[j] >= hi[j];
    const uint32_t nxt = wrap ? lo[j] : cur[j] + 1;
    cur[j] = carry ? nxt : cur[j];
    carry = carry && wrap;
  }
  return !carry;
}

template <int ND, typename G>
__device__ __forceinline__ uint64_t row_base(const G &g, const uint32_t *cur) {
  uint64_t rowBase = g.keyBase;
  const uint32_t nd = nd_of<ND>(g);
#pragma unroll
  for (uint32_t j = 1; j < nd_cap<ND>(); j++) {
    if (j >= nd) break;
    rowBase += (uint64_t)cur[j] * g.mult[j];
  }
  return rowBase;
}

// Calls f(pBeg, pEnd) for each row of the block [lo, hi] holding points (the arguments and the
// calls are uniform across the wave when lo/hi are).  narrow(cur, l0, h0) may shrink a row's
// dim-0 range [l0, h0] (initially [lo[0], hi[0]]) or drop the row (returns false); it must
// return wave-uniform results.
template <int ND, typename G, typename R, typename F>
__device__ __forceinline__ void for_each_row(const JoinArgs &a, const G &g, const uint32_t *lo,
                                             const uint32_t *hi, R &&narrow, F &&f) {
  uint32_t cur[nd_cap<ND>()];
  const uint32_t nd = nd_of<ND>(g);
#pragma unroll
  for (uint32_t j = 0; j < nd_cap<ND>(); j++) cur[j] = lo[j < nd ? j : 0];
  do {
    uint32_t l0 = lo[0], h0 = hi[0];
    if (narrow(cur, l0, h0)) {
      uint32_t pBeg, pEnd;
      row_range(a, row_base<ND>(g, cur), l0, h0, pBeg, pEnd);
      // the range is wave-uniform (so are lo/hi): scalar registers, scalar loop control
      pBeg = (uint32_t)__builtin_amdgcn_readfirstlane((int)pBeg);
      pEnd = (uint32_t)__builtin_amdgcn_readfirstlane((int)pEnd);
      if (pEnd > pBeg) f(pBeg, pEnd);
    }
  } while (next_row<ND>(cur, lo, hi, nd));
}

// for_each_row without the point ranges: f(rowBase, l0, h0) for each row kept by narrow
// (uniform across the wave when lo/hi are).
template <int ND, typename G, typename R, typename F>
__device__ __forceinline__ void for_each_row_span(const G &g, const uint32_t *lo, const uint32_t *hi,
                                                  R &&narrow, F &&f) {
  uint32_t cur[nd_cap<ND>()];
  const uint32_t nd = nd_of<ND>(g);
#pragma unroll
  for (uint32_t j = 0; j < nd_cap<ND>(); j++) cur[j] = lo[j < nd ? j : 0];
  do {
    uint32_t l0 = lo[0], h0 = hi[0];
    if (narrow(cur, l0, h0)) f(row_base<ND>(g, cur), l0, h0);
  } while (next_row<ND>(cur, lo, hi, nd));
}

// Distance^2 from the query to the slab of a row's cells in dims 1..nd-1 (each slab widened
// by 2^-18 of its coordinates for the rounding of cell_coord).
template <typename T, int D, int ND, typename G>
__device__ __forceinline__ double row_gap2(const G &g, const uint32_t *cur, const double *qd) {
  double g2 = 0.0;
  const double e = (double)g.eps;
  const uint32_t nd = nd_of<ND>(g);
#pragma unroll
  for (uint32_t j = 1; j < nd_cap<ND>(); j++) {
    if (j >= nd) break;
    const double mn = (double)g.minArr[j];
    const double clo = mn + (double)cur[j] * e, chi = clo + e;
    const double pad = 0x1p-18 * (fabs(mn) + fabs(chi) + e);
    const double q = pick<D>(qd, j);
    const double gap = fmax(0.0, fmax((clo - pad) - q, q - (chi + pad)));
    g2 += gap * gap;
  }
  return g2;
}

// Chord of a ball in one row of a block (used with 3 or more indexed dims, where it drops the
// many rows of a block's corners; with 2, the per-row work costs more than the cells saved).
// A point of the row's cells lies at least sqrt(gap2) from the query in dims 1..nd-1, so a
// point within sqrt(r2) of the query has |x0 - q0| <= sqrt(r2 - gap2): only the dim-0 cells of
// that chord can hold one, and no cell of the row when gap2 > r2.  Per lane; narrows [l0, h0].
template <typename T, int D, int ND, typename G>
__device__ __forceinline__ bool chord_cells(const G &g, double g2, const double *qd, double r2, bool active,
                                            uint32_t &l, uint32_t &u) {
  l = 0xffffffffu;
  u = 0u;
  if (!(active && g2 <= r2)) return false;
  const double h = sqrt(r2 - g2) * (1.0 + 0x1p-20);
  const T xa = next_down((T)(qd[0] - h)), xb = next_up((T)(qd[0] + h));
  l = cell_coord<T>(xa, (T)g.minArr[0], (T)g.eps, g.nCells[0]);
  u = cell_coord<T>(xb, (T)g.minArr[0], (T)g.eps, g.nCells[0]);
  return true;
}

// The wave's union of its lanes' chords (inactive lanes contribute nothing), clipped to [l0, h0].
template <typename T, int D, int ND, typename G>
__device__ __forceinline__ bool row_chord(const G &g, const uint32_t *cur, const double *qd, double r2,
                                          bool active, uint32_t &l0, uint32_t &h0) {
  uint32_t l, u;
  chord_cells<T, D, ND, G>(g, row_gap2<T, D, ND, G>(g, cur, qd), qd, r2, active, l, u);
  l = wave_min_u32(l);
  u = wave_max_u32(u);
  l0 = l > l0 ? l : l0;
  h0 = u < h0 ? u : h0;
  return l0 <= h0;
}

// One lane's own chord (no wave reduction); false when the row lies outside the ball.
template <typename T, int D, int ND, typename G>
__device__ __forceinline__ bool row_chord_lane(const G &g, const uint32_t *cur, const double *qd, double r2,
                                               uint32_t &l0, uint32_t &h0) {
  uint32_t l, u;
  if (!chord_cells<T, D, ND, G>(g, row_gap2<T, D, ND, G>(g, cur, qd), qd, r2, true, l, u)) return false;
  l0 = l > l0 ? l : l0;
  h0 = u < h0 ? u : h0;
  return l0 <= h0;
}

// XCD-aware tile queue of a persistent kernel.  Workgroups are dispatched to the 8 XCDs in
// turn (workgroup b on XCD b % 8), each XCD with its own L2, and neighbouring tiles (cell
// order) read mostly the same candidates: the tiles are dealt in TILE_QUEUES contiguous ranges,
// a workgroup takes tiles from its XCD's range first (its L2 then serves neighbouring tiles)
// and, once that is empty, from the other ranges in turn.  Wave-uniform.
static_assert(TILE_QUEUES == 8, "tile ranges are split by a shift of 3");
// Next tile index of the workgroup, NT when every range is empty.  ranges: 1 or TILE_QUEUES.
// No state is kept between calls: a workgroup whose ranges are empty re-reads their heads (at
// most TILE_QUEUES atomics per call once its own range is empty).
__device__ __forceinline__ uint32_t next_tile(uint32_t *heads, uint32_t NT, uint32_t ranges, uint32_t lane) {
  const uint32_t sh = ranges == TILE_QUEUES ? 3u : 0u;
  const uint32_t mask = (1u << sh) - 1u;
  for (uint32_t j = 0; j <= mask; j++) {
    const uint32_t r = (blockIdx.x + j) & mask;
    const uint32_t b = (uint32_t)(((uint64_t)NT * r) >> sh);
    const uint32_t e = (uint32_t)(((uint64_t)NT * (r + 1)) >> sh);
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(&heads[r], 1u);
    t = (uint32_t)__shfl((int)t, 0, 64);
    if (t < e - b) return b + t;
  }
  return NT;
}

// The grid, K+1 and row addressing of one tile: the kernel's own for a single set; the
// tile's set (all queries of a tile share one cell, hence one set) in a batched join.
template <typename G>
struct TileCtx {
  G g;
  uint32_t k1;
  uint32_t idBase;  // reported id = point index - idBase (the set's first point in a batch)
  int64_t rowAdj;   // row of query q starts at element (q * k1 + rowAdj)
  __device__ __forceinline__ size_t row(uint32_t q) const { return (size_t)((int64_t)q * k1 + rowAdj); }
};
template <typename G>
__device__ __forceinline__ TileCtx<G> tile_ctx(const JoinArgs &a, const G &g, const uint4 &tl) {
  TileCtx<G> c;
  c.g = g;
  c.k1 = a.k1;
  c.idBase = 0;
  c.rowAdj = 0;
  if constexpr (std::is_same<G, GridParams>::value) {
    if (a.sets) {
      const SetDesc &sd = a.sets[a.setOf[tile_query(a, tl, 0u)]];
      c.g = sd.g;
      c.k1 = sd.k1;
      c.idBase = sd.pStart;
      c.rowAdj = (int64_t)sd.outOff - (int64_t)sd.pStart * sd.k1;
    }
  }
  return c;
}

__device__ __forceinline__ uint32_t tile_count(const JoinArgs &a) {
  return a.ntilesDev ? *a.ntilesDev : a.ntiles;
}

// Per-wave counters of a persistent kernel (wave-uniform), added to the global counters once
// when the wave runs out of tiles: per-tile atomics on the same three addresses serialise.
struct WaveTally {
  uint32_t fails = 0;
  unsigned long long bytes = 0, cand = 0;
};

template <typename T>
__device__ __forceinline__ void tile_accounting(WaveTally &w, const uint4 &tl, bool ok, bool failedLane,
                                                uint64_t tileCand, uint32_t K1) {
  const uint64_t fm = __ballot(failedLane);
  const uint64_t om = __ballot(ok);
  w.fails += (uint32_t)__popcll(fm);
  // compulsory HBM bytes of the tile's queries: ids in / failure flags out and the finished
  // rows (the query cells' candidate blocks are counted by the pre-filter)
  w.bytes += (unsigned long long)tl.z * 8ull + (unsigned long long)__popcll(om) * K1 * (4ull + sizeof(T));
  w.cand += (unsigned long long)tileCand * tl.z;
}

__device__ __forceinline__ void flush_tally(const JoinArgs &a, const WaveTally &w, uint32_t lane) {
  if (lane != 0) return;
  if (w.fails) atomicAdd(a.failCount, w.fails);
  if (w.bytes) atomicAdd(a.candCount + 1, w.bytes);
  if (w.cand) atomicAdd(a.candCount, w.cand);
}

// LDS hand-off between the lanes of a wave.  The range kernels run one wave per workgroup,
// so this is all the synchronisation they need; __syncthreads would also wait for every
// outstanding global load and store (the row writes included) at each call.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Row ranges of the block [lo, hi] (rows with nothing within sqrt(r2) of qd dropped when chord
// is set), looked up in parallel, lane = row, for rows rb .. rb+63: pB = first sorted position,
// len = its point count (0 for rows past the end or dropped).
template <typename T, int D, int ND, typename G>
__device__ __forceinline__ void row_batch(const JoinArgs &a, const G &g, const uint32_t *lo, const uint32_t *hi,
                                          bool chord, const double *qd, double r2, uint64_t R, uint64_t rr,
                                          uint32_t &pB, uint32_t &len) {
  pB = 0;
  len = 0;
  if (rr >= R) return;
  uint32_t cur[nd_cap<ND>()];
  const uint32_t nd = nd_of<ND>(g);
  uint64_t t = rr;
#pragma unroll
  for (uint32_t j = 1; j < nd_cap<ND>(); j++) {
    if (j >= nd) break;
    const uint32_t span = hi[j] - lo[j] + 1;
    if (t < span) {
      cur[j] = lo[j] + (uint32_t)t;
      t = 0;
    } else {
      cur[j] = lo[j] + (uint32_t)(t % span);
      t /= span;
    }
  }
  uint32_t l0 = lo[0], h0 = hi[0];
  const bool keep = chord ? row_chord_lane<T, D, ND, G>(g, cur, qd, r2, l0, h0) : true;
  if (keep) {
    uint32_t pE;
    row_range(a, row_base<ND>(g, cur), l0, h0, pB, pE);
    len = pE - pB;
  }
}

// Blocks at most this many cells wide in dim 0 take the slab cut (a cell lookup and two
// binary searches per cell instead of one lookup per row: wide blocks of a large radius over
// sparse cells would pay more for the lookups than the cut saves).
constexpr uint32_t SLAB_MAX_SPAN = 64;

// Narrows the sorted positions [pB, pE) of one cell to its points whose slab code lies in
// [qlo, qhi]: the cell's points are in slab order, so two binary searches on the coordinates
// xsd of the first non-indexed dim (both probes of a step loaded together).
template <typename T, typename G>
__device__ __forceinline__ void slab_narrow(const T *__restrict__ xsd, const G &g, uint32_t qlo, uint32_t qhi,
                                            uint32_t &pB, uint32_t &pE) {
  uint32_t l1 = pB, h1 = pE, l2 = pB, h2 = pE;
  while (l1 < h1 || l2 < h2) {
    const uint32_t m1 = (l1 + h1) >> 1, m2 = (l2 + h2) >> 1;
    // a finished search probes pB again (pB < pE here), so every probe stays in the cell
    const uint32_t c1 = slab_q<T>(xsd[l1 < h1 ? m1 : pB], g);
    const uint32_t c2 = slab_q<T>(xsd[l2 < h2 ? m2 : pB], g);
    if (l1 < h1) {
      if (c1 < qlo) l1 = m1 + 1; else h1 = m1;
    }
    if (l2 < h2) {
      if (c2 <= qhi) l2 = m2 + 1; else h2 = m2;
    }
  }
  pB = l1;
  pE = l2 > l1 ? l2 : l1;
}

// Segments of the block [lo, hi] with slab order: segment rr (of R rows x S0 = hi[0]-lo[0]+1
// cells, row-major) is the cell lo[0] + rr % S0 of row rr / S0, narrowed to the slab codes
// [qlo, qhi]; pB / len as row_batch (len 0 for cells outside the row's chord or past RS).
template <typename T, int D, int ND, typename G>
__device__ __forceinline__ void seg_batch(const JoinArgs &a, const G &g, const uint32_t *lo, const uint32_t *hi,
                                          bool chord, const double *qd, double r2, uint64_t RS, uint64_t rr,
                                          uint32_t qlo, uint32_t qhi, const T *__restrict__ xsd,
                                          uint32_t &pB, uint32_t &len) {
  pB = 0;
  len = 0;
  if (rr >= RS) return;
  const uint32_t S0 = hi[0] - lo[0] + 1;
  const uint64_t row = rr / S0;
  const uint32_t x0 = lo[0] + (uint32_t)(rr - row * S0);
  uint32_t cur[nd_cap<ND>()];
  const uint32_t nd = nd_of<ND>(g);
  uint64_t t = row;
#pragma unroll
  for (uint32_t j = 1; j < nd_cap<ND>(); j++) {
    if (j >= nd) break;
    const uint32_t span = hi[j] - lo[j] + 1;
    cur[j] = lo[j] + (uint32_t)(t % span);
    t /= span;
  }
  uint32_t l0 = lo[0], h0 = hi[0];
  const bool keep = chord ? row_chord_lane<T, D, ND, G>(g, cur, qd, r2, l0, h0) : true;
  if (!keep || x0 < l0 || x0 > h0) return;
  uint32_t pE;
  row_range(a, row_base<ND>(g, cur), x0, x0, pB, pE);
  if (pE > pB) slab_narrow<T>(xsd, g, qlo, qhi, pB, pE);
  len = pE - pB;
}

template <int ND, typename G>
__device__ __forceinline__ uint64_t block_rows(const G &g, const uint32_t *lo, const uint32_t *hi) {
  uint64_t R = 1;  // rows of the block (spans are <= the grid's cells, so R <= its cell count)
  const uint32_t nd = nd_of<ND>(g);
#pragma unroll
  for (uint32_t j = 1; j < nd_cap<ND>(); j++) {
    if (j >= nd) break;
    R *= (uint64_t)(hi[j] - lo[j] + 1);
  }
  return R;
}

// Inclusive wave prefix sum.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)v, o, 64);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

// Segments whose lengths average this many positions or more are streamed one at a time (the
// partly empty last round of each then costs less than looking up every candidate's segment).
constexpr uint32_t SEG_LONG = 128;

// One run of len sorted positions from pB, PR rounds of 64 at a time, lane = candidate:
// group(pv, val, f0, len) gets the PR rounds' positions (lanes past the end repeat the last
// position with val false; round u counts when f0 + 64 u < len).  Wave-uniform arguments.
template <int PR, typename GF>
__device__ __forceinline__ void stream_row(uint32_t pB, uint32_t len, uint32_t lane, GF &&group) {
  for (uint32_t f0 = 0; f0 < len; f0 += 64u * PR) {
    uint32_t pv[PR];
    bool val[PR];
#pragma unroll
    for (int u = 0; u < PR; u++) {
      const uint32_t fi = f0 + 64u * u + lane;
      val[u] = fi < len;
      pv[u] = pB + (val[u] ? fi : len - 1u);
    }
    group(pv, val, f0, len);
  }
}

// One batch of 64 segments of sorted positions (lane = segment: first position pB, length len;
// empty segments anywhere) streamed through group as stream_row does, the segments
// concatenated.  Long segments go one at a time; otherwise a candidate finds its segment from
// a bitmap of the segment starts (rowX: 2048 positions per chunk, one bit each): the segments
// begun before its window of 64 positions, plus `mbcnt` of the window's starts at or below its
// lane, index the segments' position offsets (rowB, in the order of the non-empty segments).
// rowB / rowX: 64-word LDS arrays of the wave, free for group to ignore.  capLog2: log2 of the
// positions per chunk, 6..11 (11: the whole bitmap; smaller only to exercise the chunking).
template <int PR, typename GF>
__device__ __forceinline__ void stream_segments(uint32_t pB, uint32_t len, uint32_t lane, uint32_t *rowB,
                                                uint32_t *rowX, uint32_t capLog2, GF &&group) {
  const uint32_t inc = wave_incl_sum(len, lane);
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
  if (total == 0) return;
  const uint32_t excl = inc - len;
  uint64_t nz = __ballot(len > 0);
  if (total >= (uint32_t)__popcll(nz) * SEG_LONG) {
    while (nz) {
      const int k = __builtin_ctzll(nz);
      nz &= nz - 1;
      stream_row<PR>((uint32_t)__builtin_amdgcn_readlane((int)pB, k),
                     (uint32_t)__builtin_amdgcn_readlane((int)len, k), lane, group);
    }
    return;
  }
  const uint32_t rk = __builtin_amdgcn_mbcnt_hi((uint32_t)(nz >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nz, 0u));
  if (len > 0) rowB[rk] = pB - excl;
  const uint32_t CAP = 1u << capLog2;
  for (uint32_t c0 = 0; c0 < total; c0 += CAP) {
    rowX[lane] = 0u;
    wave_lds_sync();
    if (len > 0 && excl >= c0 && excl - c0 < CAP) atomicOr(&rowX[(excl - c0) >> 5], 1u << ((excl - c0) & 31u));
    // segments with points begun before the chunk (>= 1 past the first chunk)
    uint32_t run = (uint32_t)__popcll(__ballot(len > 0 && excl < c0));
    wave_lds_sync();
    const uint32_t ctot = total - c0 < CAP ? total - c0 : CAP;
    const uint32_t nwin = (ctot + 63u) >> 6;
    for (uint32_t f0 = 0; f0 < ctot; f0 += 64u * PR) {
      uint32_t pv[PR];
      bool val[PR];
#pragma unroll
      for (int u = 0; u < PR; u++) {
        const uint32_t wdx = (f0 >> 6) + (uint32_t)u;
        // windows past the chunk's end have no starts: their lanes take the last segment
        uint64_t msk = 0;
        if (wdx < nwin) {
          const uint32_t lo32 = (uint32_t)__builtin_amdgcn_readfirstlane((int)rowX[2u * wdx]);
          const uint32_t hi32 = (uint32_t)__builtin_amdgcn_readfirstlane((int)rowX[2u * wdx + 1u]);
          msk = ((uint64_t)hi32 << 32) | lo32;
        }
        const uint32_t fi = f0 + 64u * u + lane;
        val[u] = fi < ctot;
        // a lane past the end takes the chunk's last position, whose segment is the last one
        // begun at or below it: no start lies past ctot - 1 in the chunk
        const uint32_t fc = val[u] ? fi : ctot - 1u;
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(msk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)msk, 0u));
        const uint32_t row = run + below + (uint32_t)((msk >> lane) & 1ull) - 1u;
        run += (uint32_t)__popcll(msk);
        pv[u] = c0 + fc + rowB[row];
      }
      group(pv, val, f0, ctot);
    }
  }
  wave_lds_sync();
}

}  // namespace
}  // namespace knn
