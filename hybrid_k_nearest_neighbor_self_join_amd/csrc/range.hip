// range.hip -- the epsilon-range top-(K+1) kernels for gfx950.
//
// Replace kernel.cu:61-225 (range query over the adjacent cells), the pair buffer and the two
// thrust::stable_sort_by_key of GPU.cu:1629-1634, and the host top-k of GPU.cu:3610-3704.
//
// Work unit: one wavefront per tile of <= 64 consecutive active queries of one tile block (a
// window of cells along dim 0 in one row of the grid); lane = query.  Each lane has a search
// radius r (the grid's eps, or its own radius, which may span several cells) and scans the
// cells [coord(x - r, rounded down), coord(x + r, rounded up)] per indexed dim; the tile scans
// the union of its lanes' blocks.  Within one row of that block (dims 1.. fixed) the cells of
// dim 0 are adjacent linear ids, i.e. one contiguous range of cell-sorted points, streamed
// 64 at a time (one coalesced load per lane) and broadcast to all lanes.
//
// Ranking: FP64 key = sum_j ((double)c_j - (double)q_j)^2, ties by point id.  The FP32 sum
// s = sum_j (c_j - q_j)^2 (reported distance sqrtf(s)) is within 2^-19 of the key relative,
// so it can filter: a candidate whose key could be <= X has s <= X (1 + 2^-16).
//
// Success (per query): at least K+1 candidates with key <= Tq = r^2 (1 - 2^-20).  Every point
// with key <= Tq lies within r of the query in each indexed dim, hence inside the scanned
// block, so a success is the exact global top-(K+1).  The reference's FP32 test
// (kernel.cu:130-147) differs from this only for candidates within 2^-20 of eps.
//
// k_range_select (main kernel), per tile:
//   pass 1  (lane = query) per-lane histogram of s over [0, F] (F = Tq (1+2^-16)) in NB=64
//           linear buckets, LDS ds_add without return;
//   scan    b* = first bucket whose cumulative count reaches K+1; each histogram column
//           becomes that query's bucket slot offsets (a counting sort);
//   pass 2  (wave = query, for each query of the tile) the 64 lanes scan the query's own
//           sub-block, radius sqrt(edge of b*), coalesced; candidates in buckets <= b* (plus a
//           2^-16 margin past b*'s edge) are placed in LDS staging in bucket order, ranked
//           exactly by (key, id) against buckets b-1..b+1 only, and the first K+1 are written
//           to the row (coalesced) when the (K+1)-th key is <= Tq.  The sub-blocks' row ranges
//           are looked up for the whole tile at once (lane = query) and a query's candidates
//           are loaded 4 rounds at a time, so a query waits for one memory round trip.
//   A query whose staged list exceeds `smax` LDS slots (e.g. thousands of duplicates in one
//   bucket) is flagged for k_range_heap.
// k_range_collect (5 or more indexed dims): wave = query, one pass over the query's own block
//   into the LDS top-(K+1) collector of knn_collect.h, FP32-prefiltered against its threshold.
// k_range_heap: per-lane max-heap of (key, id), in LDS when K+1 fits, else in global scratch.
#include <type_traits>

#include "knn_collect.h"
#include "knn_launch.h"

namespace knn {

namespace {

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
// against the FP64 key.
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
  static constexpr float margin = 1.0f + 0x1p-16f;
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
  static constexpr double margin = 1.0;
};

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

// Per-lane state of one tile.
template <typename T, int D>
struct Lane {
  bool active;
  uint32_t q;
  T qc[D];
  double qd[D];
  double Tq;  // success threshold on the key
  T F;        // filter bound on the FP32/T sum
};

template <typename T, int D, typename G>
__device__ __forceinline__ void tile_setup(const JoinArgs &a, const G &g, const uint4 &tl,
                                           uint32_t lane, Lane<T, D> &L, uint32_t *lo,
                                           uint32_t *hi, uint32_t *ownLo, uint32_t *ownHi,
                                           uint64_t &cellId) {
  const T *__restrict__ x = reinterpret_cast<const T *>(a.x);
  const size_t n = a.n;
  L.active = lane < tl.z;
  L.q = a.act[tl.y + (L.active ? lane : 0u)];
#pragma unroll
  for (int j = 0; j < D; j++) {
    L.qc[j] = x[(size_t)j * n + L.q];
    L.qd[j] = (double)L.qc[j];
  }
  const T r = a.radius ? reinterpret_cast<const T *>(a.radius)[L.q] : (T)a.radiusAll;
  L.Tq = (double)r * (double)r * (1.0 - 0x1p-20);
  L.F = Num<T>::bound(L.Tq);
  const T eps = (T)g.eps;
  cellId = g.keyBase;
  for (uint32_t j = 0; j < g.nd; j++) {
    const T mn = (T)g.minArr[j];
    const uint32_t nc = g.nCells[j];
    T xj = L.qc[0];
#pragma unroll
    for (int jj = 1; jj < D; jj++)
      if ((uint32_t)jj == j) xj = L.qc[jj];
    cellId += (uint64_t)cell_coord<T>(xj, mn, eps, nc) * g.mult[j];
    uint32_t l = cell_coord<T>(next_down(xj - r), mn, eps, nc);
    uint32_t u = cell_coord<T>(next_up(xj + r), mn, eps, nc);
    ownLo[j] = l;
    ownHi[j] = u;
    if (!L.active) {
      l = 0xffffffffu;
      u = 0u;
    }
    lo[j] = wave_min_u32(l);
    hi[j] = wave_max_u32(u);
  }
}

// Calls f(pBeg, pEnd) for each row of the block [lo, hi] holding points (the arguments and the
// calls are uniform across the wave when lo/hi are).  narrow(cur, l0, h0) may shrink a row's
// dim-0 range [l0, h0] (initially [lo[0], hi[0]]) or drop the row (returns false); it must
// return wave-uniform results.
template <typename G, typename R, typename F>
__device__ __forceinline__ void for_each_row(const JoinArgs &a, const G &g, const uint32_t *lo,
                                             const uint32_t *hi, R &&narrow, F &&f) {
  uint32_t cur[KNN_MAX_INDEXED];
  for (uint32_t j = 0; j < g.nd; j++) cur[j] = lo[j];
  for (;;) {
    uint32_t l0 = lo[0], h0 = hi[0];
    if (narrow(cur, l0, h0)) {
      uint64_t rowBase = g.keyBase;
      for (uint32_t j = 1; j < g.nd; j++) rowBase += (uint64_t)cur[j] * g.mult[j];
      uint32_t pBeg, pEnd;
      if (a.denseStart) {
        pBeg = a.denseStart[rowBase + l0];
        pEnd = a.denseStart[rowBase + h0 + 1];
      } else {
        const uint32_t cbeg = lower_bound_u64(a.cellKeys, a.ncell, rowBase + l0);
        const uint32_t cend = upper_bound_u64(a.cellKeys, a.ncell, rowBase + h0);
        pBeg = a.cellStart[cbeg];
        pEnd = a.cellStart[cend];
      }
      // the range is wave-uniform (so are lo/hi): scalar registers, scalar loop control
      pBeg = (uint32_t)__builtin_amdgcn_readfirstlane((int)pBeg);
      pEnd = (uint32_t)__builtin_amdgcn_readfirstlane((int)pEnd);
      if (pEnd > pBeg) f(pBeg, pEnd);
    }
    uint32_t j = 1;
    while (j < g.nd) {
      if (cur[j] < hi[j]) {
        cur[j]++;
        break;
      }
      cur[j] = lo[j];
      j++;
    }
    if (j >= g.nd) break;
  }
}

template <typename G, typename F>
__device__ __forceinline__ void for_each_range(const JoinArgs &a, const G &g, const uint32_t *lo,
                                               const uint32_t *hi, F &&f) {
  for_each_row(a, g, lo, hi, [](const uint32_t *, uint32_t &, uint32_t &) { return true; }, f);
}

template <int D>
__device__ __forceinline__ double pick(const double *v, uint32_t j) {
  double r = v[0];
#pragma unroll
  for (int jj = 1; jj < D; jj++)
    if ((uint32_t)jj == j) r = v[jj];
  return r;
}

// Chord of a ball in one row of a block (used with 3 or more indexed dims, where it drops the
// many rows of a block's corners; with 2, the per-row work costs more than the cells saved).  A point of the row's cells lies at least `gap` from
// the query in dims 1..nd-1 (gap = distance to the row's cell slab, each slab widened by
// 2^-18 of its coordinates for the rounding of cell_coord), so a point within sqrt(r2) of the
// query has |x0 - q0| <= sqrt(r2 - gap^2): only the dim-0 cells of that chord can hold one,
// and no cell of the row when gap^2 > r2.  Per lane (inactive lanes contribute nothing); the
// row range becomes the wave's union, clipped to [l0, h0].
template <typename T, int D, typename G>
__device__ __forceinline__ bool row_chord(const G &g, const uint32_t *cur, const double *qd, double r2,
                                          bool active, uint32_t &l0, uint32_t &h0) {
  double g2 = 0.0;
  const double e = (double)g.eps;
  for (uint32_t j = 1; j < g.nd; j++) {
    const double mn = (double)g.minArr[j];
    const double clo = mn + (double)cur[j] * e, chi = clo + e;
    const double pad = 0x1p-18 * (fabs(mn) + fabs(chi) + e);
    const double q = pick<D>(qd, j);
    const double gap = fmax(0.0, fmax((clo - pad) - q, q - (chi + pad)));
    g2 += gap * gap;
  }
  uint32_t l = 0xffffffffu, u = 0u;
  if (active && g2 <= r2) {
    const double h = sqrt(r2 - g2) * (1.0 + 0x1p-20);
    const T xa = next_down((T)(qd[0] - h)), xb = next_up((T)(qd[0] + h));
    l = cell_coord<T>(xa, (T)g.minArr[0], (T)g.eps, g.nCells[0]);
    u = cell_coord<T>(xb, (T)g.minArr[0], (T)g.eps, g.nCells[0]);
  }
  l = wave_min_u32(l);
  u = wave_max_u32(u);
  l0 = l > l0 ? l : l0;
  h0 = u < h0 ? u : h0;
  return l0 <= h0;
}

__device__ __forceinline__ uint32_t next_tile(uint32_t *counter, uint32_t lane) {
  uint32_t t = 0;
  if (lane == 0) t = atomicAdd(counter, 1u);
  return (uint32_t)__shfl((int)t, 0, 64);
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
      const SetDesc &sd = a.sets[a.setOf[a.act[tl.y]]];
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

template <typename T, int D>
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

// row_chord for one lane's own row (no wave reduction): narrows [l0, h0] to the chord of the
// ball of squared radius r2 around qd; false when the row lies outside the ball
template <typename T, int D, typename G>
__device__ __forceinline__ bool row_chord_lane(const G &g, const uint32_t *cur, const double *qd, double r2,
                                               uint32_t &l0, uint32_t &h0) {
  double g2 = 0.0;
  const double e = (double)g.eps;
  for (uint32_t j = 1; j < g.nd; j++) {
    const double mn = (double)g.minArr[j];
    const double clo = mn + (double)cur[j] * e, chi = clo + e;
    const double pad = 0x1p-18 * (fabs(mn) + fabs(chi) + e);
    const double q = pick<D>(qd, j);
    const double gap = fmax(0.0, fmax((clo - pad) - q, q - (chi + pad)));
    g2 += gap * gap;
  }
  if (!(g2 <= r2)) return false;
  const double h = sqrt(r2 - g2) * (1.0 + 0x1p-20);
  const T xa = next_down((T)(qd[0] - h)), xb = next_up((T)(qd[0] + h));
  const uint32_t l = cell_coord<T>(xa, (T)g.minArr[0], (T)g.eps, g.nCells[0]);
  const uint32_t u = cell_coord<T>(xb, (T)g.minArr[0], (T)g.eps, g.nCells[0]);
  l0 = l > l0 ? l : l0;
  h0 = u < h0 ? u : h0;
  return l0 <= h0;
}

// LDS hand-off between the lanes of a wave.  The range kernels run one wave per workgroup,
// so this is all the synchronisation they need; __syncthreads would also wait for every
// outstanding global load and store (the row writes included) at each call.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Calls f(p) for every sorted position p of the block [lo, hi] (rows with nothing within
// sqrt(r2) of qd dropped when chord is set), 64 positions per round, lane = position.  The
// rows' ranges are looked up in parallel (lane = row, 64 rows per batch) and concatenated
// through a wave prefix sum, so a round mixes rows and every lane has a candidate.
// rowB / rowX: 64-entry LDS arrays of the wave.  Wave-uniform arguments.
template <typename T, int D, typename G, typename F>
__device__ __forceinline__ void for_each_point_flat(const JoinArgs &a, const G &g, const uint32_t *lo,
                                                    const uint32_t *hi, bool chord, const double *qd,
                                                    double r2, uint32_t lane, uint32_t *rowB,
                                                    uint32_t *rowX, F &&f) {
  uint64_t R = 1;  // rows of the block (spans are <= the grid's cells, so R <= its cell count)
  for (uint32_t j = 1; j < g.nd; j++) R *= (uint64_t)(hi[j] - lo[j] + 1);
  for (uint64_t rb = 0; rb < R; rb += 64) {
    const uint64_t rr = rb + lane;
    uint32_t pB = 0, len = 0;
    if (rr < R) {
      uint32_t cur[KNN_MAX_INDEXED];
      uint64_t t = rr;
      uint64_t rowBase = g.keyBase;
      for (uint32_t j = 1; j < g.nd; j++) {
        const uint32_t span = hi[j] - lo[j] + 1;
        if (t < span) {
          cur[j] = lo[j] + (uint32_t)t;
          t = 0;
        } else {
          cur[j] = lo[j] + (uint32_t)(t % span);
          t /= span;
        }
        rowBase += (uint64_t)cur[j] * g.mult[j];
      }
      uint32_t l0 = lo[0], h0 = hi[0];
      const bool keep = chord ? row_chord_lane<T, D, G>(g, cur, qd, r2, l0, h0) : true;
      if (keep) {
        uint32_t pE;
        if (a.denseStart) {
          pB = a.denseStart[rowBase + l0];
          pE = a.denseStart[rowBase + h0 + 1];
        } else {
          const uint32_t cbeg = lower_bound_u64(a.cellKeys, a.ncell, rowBase + l0);
          const uint32_t cend = upper_bound_u64(a.cellKeys, a.ncell, rowBase + h0);
          pB = a.cellStart[cbeg];
          pE = a.cellStart[cend];
        }
        len = pE - pB;
      }
    }
    uint32_t inc = len;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)inc, o, 64);
      if (lane >= (uint32_t)o) inc += t;
    }
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    const uint32_t excl = inc - len;  // exclusive offset, non-decreasing over the lanes
    const uint64_t left = R - rb;
    if (left <= 8) {
      // few rows (the usual case with <= 2 indexed dims): offsets and starts in scalar
      // registers, a candidate's row found by compares
      uint32_t X[8], Bs[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        X[k] = (uint32_t)__builtin_amdgcn_readlane((int)excl, k);
        Bs[k] = (uint32_t)__builtin_amdgcn_readlane((int)pB, k);
      }
      const uint32_t nr = (uint32_t)left;
      for (uint32_t f0 = 0; f0 < total; f0 += 64) {
        const uint32_t fi = f0 + lane;
        if (fi < total) {
          uint32_t b = Bs[0], x = X[0];
#pragma unroll
          for (int k = 1; k < 8; k++)
            if ((uint32_t)k < nr && fi >= X[k]) {
              b = Bs[k];
              x = X[k];
            }
          f(b + (fi - x));
        }
      }
      continue;
    }
    rowB[lane] = pB;
    rowX[lane] = excl;
    wave_lds_sync();
    for (uint32_t f0 = 0; f0 < total; f0 += 64) {
      const uint32_t fi = f0 + lane;
      if (fi < total) {
        // last row whose exclusive offset is <= fi (rows of length 0 share the next offset)
        uint32_t l = 0, h = 64;
        while (h - l > 1) {
          const uint32_t mid = (l + h) >> 1;
          if (rowX[mid] <= fi) l = mid; else h = mid;
        }
        f(rowB[l] + (fi - rowX[l]));
      }
    }
    wave_lds_sync();
  }
}

// ====================================================================== select kernel
// Buckets of the select kernel: 63 equal buckets over [0, Tq] of the FP64 key, and bucket 63
// for every key beyond (or a NaN).  floor(min(key * scale, 63.5)) with scale = 63 / Tq is
// monotone in the key (a product by a positive constant, a minimum and a truncation, each
// monotone under round-to-nearest), so every key of bucket b is <= every key of bucket b + 1,
// and a key of a bucket <= 62 is below Tq (1 + 2^-51) < r^2: such a point lies inside the
// block of radius r.
constexpr int NB = 64;
constexpr int NBIN = NB - 1;
constexpr uint32_t SELECT_MAX_GRID = 2048;
// pass 2: sub-blocks of at most RMAX rows have their row ranges looked up for the whole tile
// at once (lane = query) and their candidates loaded PR rounds of 64 at a time
constexpr uint32_t RMAX = 6;
}  // namespace
template <typename T>
size_t range_select_lds_bytes(uint32_t smax);
namespace {

__device__ __forceinline__ uint32_t kbucket(double key, double scale) {
  const double t = fmin(key * scale, (double)NBIN + 0.5);
  return (uint32_t)t;
}

// LDS: hist[NB/2][64] u32, two 16-bit bucket counters per word (per-lane histogram
// columns, then slot offsets), the collecting query's cursors qh[NB/2], rowB/rowX[64], the
// tile's pass-2 row table prB/prX[RMAX][64] (first sorted position and exclusive offset of
// each row of each query's sub-block), prT/prR[64] (candidates and rows per query), then the
// per-query staging arrays key[SM] (double), id[SM] (u32), s[SM] (T).  A lane with 65536 or
// more candidates within Tq goes to the heap kernel, so no 16-bit counter of buckets 0..62
// can carry into its neighbour (bucket 63, the high half of the last word, may wrap: it is
// never read).
__device__ __forceinline__ uint32_t hword(uint32_t b, uint32_t col) { return (b >> 1) * 64 + col; }
__device__ __forceinline__ uint32_t hshift(uint32_t b) { return (b & 1u) << 4; }
__device__ __forceinline__ uint32_t hget(const uint32_t *w, uint32_t b) { return (w[b >> 1] >> hshift(b)) & 0xffffu; }

// Per tile (lane = query; radius r, Tq = r^2 (1 - 2^-20), FP64 key of each candidate):
//   pass 1  histogram of the keys of the tile's union of blocks over the 64 buckets;
//   scan    b* = first bucket whose cumulative count reaches K+1 (b* <= 62: success; the K+1
//           nearest are exactly the items of buckets < b* plus the smallest of bucket b*, and
//           every item of buckets <= b* lies in the block); each column becomes that query's
//           bucket slot offsets (a counting sort);
//   pass 2  (wave = query) the query's own sub-block of radius sqrt((b* + 1) / scale) is
//           streamed, lane = candidate; the items of buckets <= b* go to LDS staging slots in
//           bucket order (the same FP64 keys as pass 1, so exactly cum(b*) of them), each is
//           ranked against the items of its own bucket only (bucket order is key order), and
//           the first K+1 are written to the row.
template <typename T, int D, typename G>
__global__ __launch_bounds__(64) void k_range_select(JoinArgs a, G g0) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t *hist = reinterpret_cast<uint32_t *>(smem);
  const uint32_t SM = a.smax;
  // the collecting query's slot cursors, copied out of its histogram column: in the column
  // (word stride 64) every lane's atomic would hit the same LDS bank
  uint32_t *qh = hist + (NB / 2) * 64;
  uint32_t *rowB = qh + NB / 2;
  uint32_t *rowX = rowB + 64;
  uint32_t *prB = rowX + 64;
  uint32_t *prX = prB + RMAX * 64;
  uint32_t *prT = prX + RMAX * 64;
  uint32_t *prR = prT + 64;
  double *skey = reinterpret_cast<double *>(smem + ((NB / 2) * 64 + NB / 2 + 128 + RMAX * 128 + 128) * sizeof(uint32_t));
  uint32_t *sid = reinterpret_cast<uint32_t *>(skey + SM);
  T *ss = reinterpret_cast<T *>(sid + SM + (SM & 1));
  const uint32_t lane = threadIdx.x;
  const T *__restrict__ xs = reinterpret_cast<const T *>(a.xs);
  const size_t n = a.n;
  const uint32_t NT = tile_count(a);

  WaveTally tally;
  for (;;) {
    const uint32_t t = next_tile(a.tileCounter, lane);
    if (t >= NT) break;
    const uint4 tl = a.tiles[t];
    const TileCtx<G> C = tile_ctx<G>(a, g0, tl);
    const G &g = C.g;
    const uint32_t K1 = C.k1;
    const T eps = (T)g.eps;
    Lane<T, D> L;
    uint32_t lo[KNN_MAX_INDEXED], hi[KNN_MAX_INDEXED], ownLo[KNN_MAX_INDEXED], ownHi[KNN_MAX_INDEXED];
    uint64_t cellId;
    tile_setup<T, D, G>(a, g, tl, lane, L, lo, hi, ownLo, ownHi, cellId);
    const double scale = L.Tq > 0.0 ? (double)NBIN / L.Tq : __builtin_inf();
    wave_lds_sync();  // the previous tile's readers of hist/staging are done
#pragma unroll 8
    for (int w = 0; w < NB / 2; w++) hist[w * 64 + lane] = 0u;
    uint64_t tileCand = 0;
    uint32_t inF = 0;

    // ---- pass 1 (lane = query): histogram of the FP64 keys of the union of the tile's blocks
    // (rows narrowed to the chords of radius sqrt(Tq) with 3 or more indexed dims)
    const double rF2 = L.Tq * (1.0 + 0x1p-15);
    auto narrow1 = [&](const uint32_t *cur, uint32_t &l0, uint32_t &h0) {
      return g.nd < 3 ? true : row_chord<T, D, G>(g, cur, L.qd, rF2, L.active, l0, h0);
    };
    for_each_row(a, g, lo, hi, narrow1, [&](uint32_t pBeg, uint32_t pEnd) {
      tileCand += pEnd - pBeg;
      for (uint32_t base = pBeg; base < pEnd; base += 64) {
        const uint32_t pp = base + lane < pEnd ? base + lane : pBeg;
        T cv[D];
#pragma unroll
        for (int j = 0; j < D; j++) cv[j] = xs[(size_t)j * n + pp];
        const uint32_t m = (uint32_t)__builtin_amdgcn_readfirstlane((int)((pEnd - base) < 64u ? (pEnd - base) : 64u));
        for (uint32_t jj = 0; jj < m; jj++) {
          double key = 0.0;
#pragma unroll
          for (int j = 0; j < D; j++) {
            const double dx = (double)readlane<T>(cv[j], (int)jj) - L.qd[j];
            key = key + dx * dx;
          }
          // branch-free: every lane adds to its own column (an inactive lane's is never read)
          const uint32_t b = kbucket(key, scale);
          inF += b < (uint32_t)NBIN ? 1u : 0u;
          atomicAdd(&hist[hword(b, lane)], 1u << hshift(b));
        }
      }
    });

    // ---- scan: b* per query; counts become exclusive slot offsets (saturated past b*)
    const bool big = inF >= 65536u;
    uint32_t cum = 0, bstar = NB, cumB = 0;
    for (int w = 0; w < NB / 2; w++) {
      const uint32_t hw = hist[w * 64 + lane];
      const uint32_t h0 = hw & 0xffffu, h1 = hw >> 16;
      const uint32_t o0 = cum < 0xffffu ? cum : 0xffffu;
      if (bstar == NB && cum + h0 >= K1) {
        bstar = (uint32_t)(2 * w);
        cumB = cum + h0;
      }
      cum += h0;
      const uint32_t o1 = cum < 0xffffu ? cum : 0xffffu;
      if (bstar == NB && 2 * w + 1 < NBIN && cum + h1 >= K1) {
        bstar = (uint32_t)(2 * w + 1);
        cumB = cum + h1;
      }
      cum += h1;
      hist[w * 64 + lane] = o0 | (o1 << 16);
    }
    // buckets <= b* alone exceed the staging slots: the heap kernel takes the query
    const bool preOvf = L.active && !big && bstar < (uint32_t)NBIN && cumB > SM;
    const bool go = L.active && !big && bstar < (uint32_t)NBIN && !preOvf;
    // every staged key is below (b* + 1) / scale (1 + 2^-51): the sub-block's reach
    const double edge = go ? (double)(bstar + 1) / scale * (1.0 + 0x1p-40) : 0.0;
    bool okMine = false, ovfMine = preOvf;
    const uint64_t goMask = __ballot(go);
    wave_lds_sync();

    // ---- pass 2 set-up (lane = query): the rows of each query's collect sub-block and their
    // point ranges, for the whole tile at once: one memory round trip per tile instead of two
    // per query.  Queries with more than RMAX rows, or grids without the dense start table,
    // take the per-query path.
    const T riL = go ? (T)next_up((T)sqrt(edge)) : (T)0;
    bool fastL = false;
    if (go && a.denseStart) {
      uint32_t sl[KNN_MAX_INDEXED], sh[KNN_MAX_INDEXED];
      uint32_t R = 1;
      for (uint32_t j = 0; j < g.nd; j++) {
        T xj = L.qc[0];
#pragma unroll
        for (int jj = 1; jj < D; jj++)
          if ((uint32_t)jj == j) xj = L.qc[jj];
        sl[j] = cell_coord<T>(next_down(xj - riL), (T)g.minArr[j], eps, g.nCells[j]);
        sh[j] = cell_coord<T>(next_up(xj + riL), (T)g.minArr[j], eps, g.nCells[j]);
        if (j >= 1) {
          const uint32_t sp = sh[j] - sl[j] + 1;
          R = (R <= RMAX && sp <= RMAX) ? R * sp : RMAX + 1;
        }
      }
      if (R <= RMAX) {
        fastL = true;
        uint32_t cur[KNN_MAX_INDEXED];
        for (uint32_t j = 1; j < g.nd; j++) cur[j] = sl[j];
        const double ri2L = (double)riL * (double)riL * (1.0 + 0x1p-15);
        uint32_t acc = 0;
#pragma unroll
        for (uint32_t r = 0; r < RMAX; r++) {
          if (r < R) {
            uint64_t rowBase = g.keyBase;
            for (uint32_t j = 1; j < g.nd; j++) rowBase += (uint64_t)cur[j] * g.mult[j];
            uint32_t l0 = sl[0], h0 = sh[0];
            const bool keep = g.nd >= 3 ? row_chord_lane<T, D, G>(g, cur, L.qd, ri2L, l0, h0) : true;
            uint32_t pB = 0, len = 0;
            if (keep) {
              pB = a.denseStart[rowBase + l0];
              len = a.denseStart[rowBase + h0 + 1] - pB;
            }
            prB[r * 64 + lane] = pB;
            prX[r * 64 + lane] = acc;
            acc += len;
            for (uint32_t j = 1; j < g.nd; j++) {  // next row: dims 1.. as an odometer
              if (cur[j] < sh[j]) {
                cur[j]++;
                break;
              }
              cur[j] = sl[j];
            }
          }
        }
        prT[lane] = acc;
        prR[lane] = R;
      }
    }
    const uint64_t fastMask = __ballot(fastL);
    wave_lds_sync();

    // ---- pass 2 (wave = query): for each query, stream its own sub-block, stage the items of
    // buckets <= b* in LDS in bucket order, rank each within its bucket, write the row
    for (uint32_t i = 0; i < tl.z; i++) {
      if (!((goMask >> i) & 1ull)) continue;
      T qi[D];
      double qdi[D];
#pragma unroll
      for (int j = 0; j < D; j++) {
        qi[j] = readlane<T>(L.qc[j], (int)i);
        qdi[j] = (double)qi[j];
      }
      const double sci = readlane<double>(scale, (int)i);
      const uint32_t bsi = (uint32_t)__builtin_amdgcn_readlane((int)bstar, (int)i);
      const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)cumB, (int)i);
      const uint32_t qid = (uint32_t)__builtin_amdgcn_readlane((int)L.q, (int)i);
      if (lane < (uint32_t)(NB / 2)) qh[lane] = hist[lane * 64 + i];
      wave_lds_sync();
      bool ovf = false;
      auto stage = [&](const T *c, uint32_t cid) {
        const double key = fp64_key<D, T>(c, qdi);
        const uint32_t b = kbucket(key, sci);
        if (b <= bsi) {
          const uint32_t slot = (atomicAdd(&qh[b >> 1], 1u << hshift(b)) >> hshift(b)) & 0xffffu;
          if (slot < SM) {
            skey[slot] = key;
            sid[slot] = cid;
            ss[slot] = Num<T>::template sq<D>(c, qi);
          } else {
            ovf = true;
          }
        }
      };
      if ((fastMask >> i) & 1ull) {
        // rows from the tile's table; PR rounds of candidates loaded before any is staged
        constexpr int PR = D <= 4 ? 4 : (D <= 8 ? 2 : 1);
        const uint32_t Ri = prR[i], tot = prT[i];
        uint32_t X[RMAX], Bs[RMAX];
#pragma unroll
        for (uint32_t k = 0; k < RMAX; k++) {
          X[k] = prX[k * 64 + i];
          Bs[k] = prB[k * 64 + i];
        }
        for (uint32_t f0 = 0; f0 < tot; f0 += 64u * PR) {
          T cv[PR][D];
          uint32_t cidv[PR];
          bool val[PR];
#pragma unroll
          for (int u = 0; u < PR; u++) {
            const uint32_t fi = f0 + 64u * u + lane;
            val[u] = fi < tot;
            if (val[u]) {
              uint32_t b = Bs[0], x = X[0];
#pragma unroll
              for (uint32_t k = 1; k < RMAX; k++)
                if (k < Ri && fi >= X[k]) {
                  b = Bs[k];
                  x = X[k];
                }
              const uint32_t p = b + (fi - x);
#pragma unroll
              for (int j = 0; j < D; j++) cv[u][j] = xs[(size_t)j * n + p];
              cidv[u] = a.sortedIds[p];
            }
          }
#pragma unroll
          for (int u = 0; u < PR; u++)
            if (val[u]) stage(cv[u], cidv[u]);
        }
      } else {
        const T ri = readlane<T>(riL, (int)i);
        uint32_t slo[KNN_MAX_INDEXED], shi[KNN_MAX_INDEXED];
        for (uint32_t j = 0; j < g.nd; j++) {
          T xj = qi[0];
#pragma unroll
          for (int jj = 1; jj < D; jj++)
            if ((uint32_t)jj == j) xj = qi[jj];
          slo[j] = cell_coord<T>(next_down(xj - ri), (T)g.minArr[j], eps, g.nCells[j]);
          shi[j] = cell_coord<T>(next_up(xj + ri), (T)g.minArr[j], eps, g.nCells[j]);
        }
        const double ri2 = (double)ri * (double)ri * (1.0 + 0x1p-15);
        for_each_point_flat<T, D, G>(a, g, slo, shi, g.nd >= 3, qdi, ri2, lane, rowB, rowX, [&](uint32_t p) {
          T c[D];
#pragma unroll
          for (int j = 0; j < D; j++) c[j] = xs[(size_t)j * n + p];
          const uint32_t cid = a.sortedIds[p];  // with the coordinates: one memory round trip
          stage(c, cid);
        });
      }
      wave_lds_sync();
      // the staged count equals pass 1's cum(b*) (same keys, same buckets); anything else
      // (an overflowing slot) sends the query to the heap kernel
      if (__ballot(ovf) || hget(qh, bsi) != m) {
        if (lane == i) ovfMine = true;
        wave_lds_sync();
        continue;
      }
      // rank: an item of bucket b follows every item of the buckets below and precedes every
      // item of the buckets above, so it is compared with its own bucket only
      T *od = reinterpret_cast<T *>(a.outDists);
      for (uint32_t t0 = 0; t0 < m; t0 += 64) {
        const uint32_t tI = t0 + lane;
        if (tI < m) {
          const double kt = skey[tI];
          const uint32_t it = sid[tI];
          const uint32_t b = kbucket(kt, sci);
          const uint32_t w0 = b ? hget(qh, b - 1) : 0u;
          const uint32_t w1 = hget(qh, b);
          uint32_t r = w0;
          for (uint32_t v = w0; v < w1; v++) r += kless(skey[v], sid[v], kt, it) ? 1u : 0u;
          if (r < K1) {
            a.outIds[C.row(qid) + r] = (int32_t)(it - C.idBase);
            od[C.row(qid) + r] = sqrt(ss[tI]);
          }
        }
      }
      if (lane == i) okMine = true;
      wave_lds_sync();
    }

    ovfMine = ovfMine || (L.active && big);
    if (L.active) {
      if (ovfMine) {
        a.overflow[L.q] = 1;
      } else {
        a.failed[L.q] = okMine ? 0 : 1;
      }
    }
    const uint64_t om = __ballot(L.active && ovfMine);
    if (lane == 0 && om) atomicAdd(a.overflowCount, (uint32_t)__popcll(om));
    tile_accounting<T, D>(tally, tl, okMine, L.active && !okMine && !ovfMine, tileCand, K1);
  }
  flush_tally(a, tally, lane);
}

// ====================================================================== collect kernel
// Streams the points of the block [lo, hi] 64 per round like for_each_point_flat, but every
// lane calls f(c, p, valid) in every round (c: the coordinates of sorted position p, valid or
// not), so f may use wave-wide operations; the coordinates of PR rounds are loaded before the
// first of them is processed.
template <typename T, int D, typename G, typename F>
__device__ __forceinline__ void stream_block(const JoinArgs &a, const G &g, const uint32_t *lo,
                                             const uint32_t *hi, bool chord, const double *qd,
                                             double r2, uint32_t lane, uint32_t *rowB, uint32_t *rowX,
                                             F &&f) {
  uint64_t R = 1;
  for (uint32_t j = 1; j < g.nd; j++) R *= (uint64_t)(hi[j] - lo[j] + 1);
  for (uint64_t rb = 0; rb < R; rb += 64) {
    const uint64_t rr = rb + lane;
    uint32_t pB = 0, len = 0;
    if (rr < R) {
      uint32_t cur[KNN_MAX_INDEXED];
      uint64_t t = rr;
      uint64_t rowBase = g.keyBase;
      for (uint32_t j = 1; j < g.nd; j++) {
        const uint32_t span = hi[j] - lo[j] + 1;
        cur[j] = lo[j] + (uint32_t)(t % span);
        t /= span;
        rowBase += (uint64_t)cur[j] * g.mult[j];
      }
      uint32_t l0 = lo[0], h0 = hi[0];
      const bool keep = chord ? row_chord_lane<T, D, G>(g, cur, qd, r2, l0, h0) : true;
      if (keep) {
        uint32_t pE;
        if (a.denseStart) {
          pB = a.denseStart[rowBase + l0];
          pE = a.denseStart[rowBase + h0 + 1];
        } else {
          const uint32_t cbeg = lower_bound_u64(a.cellKeys, a.ncell, rowBase + l0);
          const uint32_t cend = upper_bound_u64(a.cellKeys, a.ncell, rowBase + h0);
          pB = a.cellStart[cbeg];
          pE = a.cellStart[cend];
        }
        len = pE - pB;
      }
    }
    uint32_t inc = len;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)inc, o, 64);
      if (lane >= (uint32_t)o) inc += t;
    }
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    rowB[lane] = pB;
    rowX[lane] = inc - len;
    wave_lds_sync();
    constexpr int PR = D <= 8 ? 4 : 2;
    const T *__restrict__ xs = reinterpret_cast<const T *>(a.xs);
    for (uint32_t f0 = 0; f0 < total; f0 += 64u * PR) {
      T cv[PR][D];
      uint32_t pv[PR];
      bool val[PR];
#pragma unroll
      for (int u = 0; u < PR; u++) {
        const uint32_t fi = f0 + 64u * u + lane;
        val[u] = fi < total;
        uint32_t p = 0;
        if (val[u]) {
          uint32_t l = 0, h = 64;
          while (h - l > 1) {
            const uint32_t mid = (l + h) >> 1;
            if (rowX[mid] <= fi) l = mid; else h = mid;
          }
          p = rowB[l] + (fi - rowX[l]);
        }
        pv[u] = p;
#pragma unroll
        for (int j = 0; j < D; j++) cv[u][j] = val[u] ? xs[(size_t)j * a.n + p] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < PR; u++)
        if (f0 + 64u * u < total) f(cv[u], pv[u], val[u]);
    }
    wave_lds_sync();
  }
}

// Wave = query, for the queries of a tile in turn: the query's own block (radius r; rows
// narrowed to the chords of its ball with 3 or more indexed dims) is streamed 64 candidates at
// a time; the FP32/T sum is compared with the T bound of the collector's threshold (first
// Tq = r^2 (1 - 2^-20), then the (K+1)-th key held), and only candidates that pass get their
// FP64 key and id and enter the LDS collector.  The query succeeds when K+1 pairs with key
// <= Tq were collected: they are then its exact top-(K+1), since every point with key <= Tq
// lies in the block.  One pass over the candidates, against the select kernel's two (the
// tile's union, then the query's sub-block): the better choice when a query has many
// candidates per neighbour, as with 3 or more indexed dims.
template <typename T, int D, typename G>
__global__ __launch_bounds__(64) void k_range_collect(JoinArgs a, G g0) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t *rowB = reinterpret_cast<uint32_t *>(smem);
  uint32_t *rowX = rowB + 64;
  unsigned char *cbase = smem + 128 * sizeof(uint32_t);
  const uint32_t lane = threadIdx.x;
  const T *__restrict__ xs = reinterpret_cast<const T *>(a.xs);
  const T *__restrict__ x = reinterpret_cast<const T *>(a.x);
  const size_t n = a.n;
  const uint32_t NT = tile_count(a);
  WaveTally tally;
  for (;;) {
    const uint32_t t = next_tile(a.tileCounter, lane);
    if (t >= NT) break;
    const uint4 tl = a.tiles[t];
    const TileCtx<G> C = tile_ctx<G>(a, g0, tl);
    const G &g = C.g;
    const uint32_t K1 = C.k1;
    const T eps = (T)g.eps;
    uint32_t nok = 0;
    unsigned long long cand = 0;
    for (uint32_t i = 0; i < tl.z; i++) {
      const uint32_t q = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.act[tl.y + i]);
      T qc[D];
      double qd[D];
#pragma unroll
      for (int j = 0; j < D; j++) {
        qc[j] = x[(size_t)j * n + q];
        qd[j] = (double)qc[j];
      }
      const T r = a.radius ? reinterpret_cast<const T *>(a.radius)[q] : (T)a.radiusAll;
      const double Tq = (double)r * (double)r * (1.0 - 0x1p-20);
      const T F = Num<T>::bound(Tq);
      uint32_t lo[KNN_MAX_INDEXED], hi[KNN_MAX_INDEXED];
      for (uint32_t j = 0; j < g.nd; j++) {
        T xj = qc[0];
#pragma unroll
        for (int jj = 1; jj < D; jj++)
          if ((uint32_t)jj == j) xj = qc[jj];
        lo[j] = cell_coord<T>(next_down(xj - r), (T)g.minArr[j], eps, g.nCells[j]);
        hi[j] = cell_coord<T>(next_up(xj + r), (T)g.minArr[j], eps, g.nCells[j]);
      }
      Collector col;
      col.init(cbase, K1);
      col.limit(Tq);
      T fb = F;
      const double rF2 = (double)F * (1.0 + 0x1p-15);
      stream_block<T, D, G>(a, g, lo, hi, g.nd >= 3, qd, rF2, lane, rowB, rowX, [&](const T *c, uint32_t p, bool valid) {
        cand += (unsigned long long)__popcll(__ballot(valid));
        const bool near = valid && Num<T>::template sq<D>(c, qc) <= fb;
        if (__ballot(near) == 0) return;
        const uint32_t cid = near ? a.sortedIds[p] : 0u;
        const double key = near ? fp64_key<D, T>(c, qd) : 0.0;
        col.offer(near, key, cid, lane);
        fb = col.nb == K1 ? Num<T>::bound(col.tk) : F;
      });
      col.merge(lane);
      const bool ok = col.nb == K1;
      if (ok) {
        T *od = reinterpret_cast<T *>(a.outDists);
        for (uint32_t e = lane; e < K1; e += 64) {
          const uint32_t id = col.best_id(e);
          T c[D];
#pragma unroll
          for (int j = 0; j < D; j++) c[j] = x[(size_t)j * n + id];
          a.outIds[C.row(q) + e] = (int32_t)(id - C.idBase);
          od[C.row(q) + e] = sqrt(Num<T>::template sq<D>(c, qc));
        }
      }
      if (lane == 0) a.failed[q] = ok ? 0 : 1;
      nok += ok ? 1u : 0u;
      br_sync();  // the collector's LDS is reused by the next query
    }
    tally.fails += tl.z - nok;
    tally.bytes += (unsigned long long)tl.z * 8ull + (unsigned long long)nok * K1 * (4ull + sizeof(T));
    tally.cand += cand;
  }
  flush_tally(a, tally, lane);
}

// ====================================================================== heap kernel
// Per-lane heap with lane-interleaved slots: key[s*64 + lane], id[s*64 + lane].
struct Heap {
  double *key;
  uint32_t *id;
  uint32_t lane;
  __device__ __forceinline__ double &K(uint32_t s) const { return key[(size_t)s * 64 + lane]; }
  __device__ __forceinline__ uint32_t &I(uint32_t s) const { return id[(size_t)s * 64 + lane]; }

  __device__ void push(uint32_t &size, double k, uint32_t kid) const {
    uint32_t i = size++;
    while (i > 0) {
      uint32_t p = (i - 1) >> 1;
      double pk = K(p);
      uint32_t pid = I(p);
      if (!kless(pk, pid, k, kid)) break;
      K(i) = pk;
      I(i) = pid;
      i = p;
    }
    K(i) = k;
    I(i) = kid;
  }
  // replace the maximum of a heap of `size` entries by (k, kid) and restore the order
  __device__ void replace_top(uint32_t size, double k, uint32_t kid) const {
    uint32_t i = 0;
    for (;;) {
      uint32_t l = 2 * i + 1;
      if (l >= size) break;
      uint32_t c = l;
      double ck = K(l);
      uint32_t cid = I(l);
      if (l + 1 < size) {
        double rk = K(l + 1);
        uint32_t rid = I(l + 1);
        if (kless(ck, cid, rk, rid)) {
          c = l + 1;
          ck = rk;
          cid = rid;
        }
      }
      if (!kless(k, kid, ck, cid)) break;
      K(i) = ck;
      I(i) = cid;
      i = c;
    }
    K(i) = k;
    I(i) = kid;
  }
  // heap sort: ascending (key, id)
  __device__ void sort(uint32_t size) const {
    for (uint32_t end = size; end > 1; end--) {
      double mk = K(0);
      uint32_t mid = I(0);
      double lk = K(end - 1);
      uint32_t lid = I(end - 1);
      K(end - 1) = mk;
      I(end - 1) = mid;
      replace_top(end - 1, lk, lid);
    }
  }
};

template <typename T, int D, bool LDS, typename G>
__global__ __launch_bounds__(64) void k_range_heap(JoinArgs a, G g0) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = threadIdx.x;
  const uint32_t KMAX = a.k1;  // heap capacity (the largest K+1 of a batch)
  Heap h;
  h.lane = lane;
  if (LDS) {
    h.key = reinterpret_cast<double *>(smem);
    h.id = reinterpret_cast<uint32_t *>(smem + (size_t)KMAX * 64 * sizeof(double));
  } else {
    unsigned char *base =
        reinterpret_cast<unsigned char *>(a.heapScratch) + (size_t)blockIdx.x * KMAX * 64 * 12;
    h.key = reinterpret_cast<double *>(base);
    h.id = reinterpret_cast<uint32_t *>(base + (size_t)KMAX * 64 * sizeof(double));
  }
  const T *__restrict__ x = reinterpret_cast<const T *>(a.x);
  const T *__restrict__ xs = reinterpret_cast<const T *>(a.xs);
  const size_t n = a.n;
  const uint32_t NT = tile_count(a);

  WaveTally tally;
  for (;;) {
    const uint32_t t = next_tile(a.tileCounter, lane);
    if (t >= NT) break;
    const uint4 tl = a.tiles[t];
    const TileCtx<G> C = tile_ctx<G>(a, g0, tl);
    const G &g = C.g;
    const uint32_t K1 = C.k1;
    Lane<T, D> L;
    uint32_t lo[KNN_MAX_INDEXED], hi[KNN_MAX_INDEXED], ownLo[KNN_MAX_INDEXED], ownHi[KNN_MAX_INDEXED];
    uint64_t cellId;
    tile_setup<T, D, G>(a, g, tl, lane, L, lo, hi, ownLo, ownHi, cellId);
    uint32_t size = 0;
    double topKey = L.Tq;
    uint32_t topId = 0xffffffffu;
    T filt = L.F;
    uint64_t tileCand = 0;

    const double rF2 = (double)L.F * (1.0 + 0x1p-15);
    auto narrowH = [&](const uint32_t *cur, uint32_t &l0, uint32_t &h0) {
      return g.nd < 3 ? true : row_chord<T, D, G>(g, cur, L.qd, rF2, L.active, l0, h0);
    };
    for_each_row(a, g, lo, hi, narrowH, [&](uint32_t pBeg, uint32_t pEnd) {
      tileCand += pEnd - pBeg;
      for (uint32_t base = pBeg; base < pEnd; base += 64) {
        const uint32_t pp = base + lane < pEnd ? base + lane : pBeg;
        T cv[D];
#pragma unroll
        for (int j = 0; j < D; j++) cv[j] = xs[(size_t)j * n + pp];
        const uint32_t cidv = a.sortedIds[pp];
        const uint32_t m = (pEnd - base) < 64u ? (pEnd - base) : 64u;
        for (uint32_t jj = 0; jj < m; jj++) {
          T c[D];
#pragma unroll
          for (int j = 0; j < D; j++) c[j] = readlane<T>(cv[j], (int)jj);
          const uint32_t cid = readlane<uint32_t>(cidv, (int)jj);
          if (!L.active) continue;
          const T s = Num<T>::template sq<D>(c, L.qc);
          if (s > filt) continue;
          const double key = fp64_key<D, T>(c, L.qd);
          if (size < K1) {
            if (key <= L.Tq) {
              h.push(size, key, cid);
              if (size == K1) {
                topKey = h.K(0);
                topId = h.I(0);
                filt = Num<T>::bound(topKey);
              }
            }
          } else if (kless(key, cid, topKey, topId)) {
            h.replace_top(K1, key, cid);
            topKey = h.K(0);
            topId = h.I(0);
            filt = Num<T>::bound(topKey);
          }
        }
      }
    });

    const bool ok = L.active && size == K1;
    if (ok) {
      h.sort(K1);
      T *od = reinterpret_cast<T *>(a.outDists);
      for (uint32_t e = 0; e < K1; e++) {
        const uint32_t id = h.I(e);
        T c[D];
#pragma unroll
        for (int j = 0; j < D; j++) c[j] = x[(size_t)j * n + id];
        a.outIds[C.row(L.q) + e] = (int32_t)(id - C.idBase);
        od[C.row(L.q) + e] = sqrt(Num<T>::template sq<D>(c, L.qc));
      }
      a.failed[L.q] = 0;
    } else if (L.active) {
      a.failed[L.q] = 1;
    }
    tile_accounting<T, D>(tally, tl, ok, L.active && !ok, tileCand, K1);
  }
  flush_tally(a, tally, lane);
}

constexpr uint32_t MAX_LDS_BYTES = 65536;
constexpr uint32_t HEAP_MAX_GRID = 4096;

template <typename T, int D, typename G>
void launch_heap(const JoinArgs &a, const G &g, hipStream_t s) {
  if (a.ntiles == 0) return;
  const size_t heapBytes = (size_t)a.k1 * 64 * 12;
  // with ntilesDev, ntiles is the persistent grid size (the tile count is on the device)
  const uint32_t grid = a.ntilesDev ? a.ntiles : (a.ntiles < HEAP_MAX_GRID ? a.ntiles : HEAP_MAX_GRID);
  if (heapBytes <= MAX_LDS_BYTES)
    k_range_heap<T, D, true, G><<<grid, 64, heapBytes, s>>>(a, g);
  else
    k_range_heap<T, D, false, G><<<grid, 64, 0, s>>>(a, g);
}

template <typename T, int D, typename G>
void launch_collect(const JoinArgs &a, const G &g, hipStream_t s) {
  if (a.ntiles == 0) return;
  const uint32_t grid = a.ntilesDev ? a.ntiles : range_select_grid(a.ntiles);
  k_range_collect<T, D, G><<<grid, 64, range_collect_lds_bytes(a.k1), s>>>(a, g);
}

template <typename T, int D, typename G>
void launch_select(const JoinArgs &a, const G &g, hipStream_t s) {
  if (a.ntiles == 0) return;
  const uint32_t grid = a.ntilesDev ? a.ntiles : range_select_grid(a.ntiles);
  k_range_select<T, D, G><<<grid, 64, range_select_lds_bytes<T>(a.smax), s>>>(a, g);
}

}  // namespace

uint32_t range_select_grid(uint32_t ntiles) {
  return ntiles < SELECT_MAX_GRID ? (ntiles ? ntiles : 1u) : SELECT_MAX_GRID;
}

template <typename T>
size_t range_select_lds_bytes(uint32_t smax) {
  return ((size_t)(NB / 2) * 64 + NB / 2 + 128 + RMAX * 128 + 128) * sizeof(uint32_t) + (size_t)smax * 8 +
         (size_t)(smax + (smax & 1)) * 4 + (size_t)smax * sizeof(T);
}

#define DISPATCH_D(FN)                              \
  switch (a.d) {                                    \
    case 1: FN<T, 1, G>(a, g, s); break;            \
    case 2: FN<T, 2, G>(a, g, s); break;            \
    case 3: FN<T, 3, G>(a, g, s); break;            \
    case 4: FN<T, 4, G>(a, g, s); break;            \
    case 5: FN<T, 5, G>(a, g, s); break;            \
    case 6: FN<T, 6, G>(a, g, s); break;            \
    case 7: FN<T, 7, G>(a, g, s); break;            \
    case 8: FN<T, 8, G>(a, g, s); break;            \
    case 9: FN<T, 9, G>(a, g, s); break;            \
    case 10: FN<T, 10, G>(a, g, s); break;          \
    case 11: FN<T, 11, G>(a, g, s); break;          \
    case 12: FN<T, 12, G>(a, g, s); break;          \
    case 13: FN<T, 13, G>(a, g, s); break;          \
    case 14: FN<T, 14, G>(a, g, s); break;          \
    case 15: FN<T, 15, G>(a, g, s); break;          \
    case 16: FN<T, 16, G>(a, g, s); break;          \
    default: break;                                 \
  }

template <typename T>
void range_select(const JoinArgs &a, const typename GridOf<T>::type &g, hipStream_t s) {
  using G = typename GridOf<T>::type;
  DISPATCH_D(launch_select)
}

template <typename T>
void range_collect(const JoinArgs &a, const typename GridOf<T>::type &g, hipStream_t s) {
  using G = typename GridOf<T>::type;
  DISPATCH_D(launch_collect)
}

size_t range_collect_lds_bytes(uint32_t k1) {
  return 128 * sizeof(uint32_t) + (size_t)k1 * 2 * (sizeof(double) + sizeof(uint32_t)) +
         (size_t)BR_BUF * (sizeof(double) + sizeof(uint32_t));
}

template <typename T>
void range_topk(const JoinArgs &a, const typename GridOf<T>::type &g, hipStream_t s) {
  using G = typename GridOf<T>::type;
  DISPATCH_D(launch_heap)
}

template <typename T>
size_t range_heap_scratch_bytes(uint32_t ntiles, uint32_t k1) {
  const size_t heapBytes = (size_t)k1 * 64 * 12;
  if (heapBytes <= MAX_LDS_BYTES) return 0;
  const uint32_t grid = ntiles < HEAP_MAX_GRID ? ntiles : HEAP_MAX_GRID;
  return heapBytes * grid;
}

#define INST(T)                                                                          \
  template void range_select<T>(const JoinArgs &, const GridOf<T>::type &, hipStream_t); \
  template size_t range_select_lds_bytes<T>(uint32_t);                                   \
  template void range_topk<T>(const JoinArgs &, const GridOf<T>::type &, hipStream_t);   \
  template void range_collect<T>(const JoinArgs &, const GridOf<T>::type &, hipStream_t); \
  template size_t range_heap_scratch_bytes<T>(uint32_t, uint32_t);
INST(float)
INST(double)
#undef INST

}  // namespace knn
