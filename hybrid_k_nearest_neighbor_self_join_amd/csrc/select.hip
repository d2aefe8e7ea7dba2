// select.hip -- the select range kernel for gfx950: the exact top-(K+1) of each query of a tile
// from the grid cells around it.
//
// Replaces kernel.cu:61-225 (range query over the adjacent cells), the pair buffer and the two
// thrust::stable_sort_by_key of GPU.cu:1629-1634, and the host top-k of GPU.cu:3610-3704.
//
// Work unit: one wavefront per tile of <= 64 consecutive active queries of one tile block (a
// window of cells along dim 0 in one row of the grid); lane = query.  Each lane has a search
// radius r (the pass's one radius, or its own) and the tile scans the union of its lanes'
// blocks (knn_range.h).
//
// Ranking: FP64 key = sum_j ((double)c_j - (double)q_j)^2, ties by point id.
// Success (per query): at least K+1 candidates with key <= Tq = r^2 (1 - 2^-20).  Every point
// with key <= Tq lies within r of the query in each indexed dim, hence inside the scanned
// block, so a success is the exact global top-(K+1).  The reference's FP32 test
// (kernel.cu:130-147) differs from this only for candidates within 2^-20 of eps.
//
// Buckets: M = SEL_NB - 1 equal buckets over [0, Tq] of the FP64 key (M = 39 or 55: the unit is
// compiled for the learned schedule's bucket count and for the other passes'), and bucket M
// for every key beyond (or a NaN).  floor(min(key * scale, M + 0.5)) with scale = M / Tq is
// monotone in the key (a product by a positive constant, a minimum and a truncation, each
// monotone under round-to-nearest), so every key of bucket b is <= every key of bucket b + 1,
// and a key of a bucket <= M - 1 is below Tq (1 + 2^-51) < r^2: such a point lies inside the
// block of radius r.
//
// Per tile:
//   pass 1  (lane = query) histogram of the keys of the tile's union of blocks over the SEL_NB
//           buckets: every candidate of a round is broadcast to all lanes (v_readlane), each
//           lane adds one to its own counter of the candidate's bucket (LDS, no return);
//   scan    b* = first bucket whose cumulative count reaches K+1 (b* <= 62: success; the K+1
//           nearest are exactly the items of buckets < b* plus the smallest of bucket b*, and
//           every item of buckets <= b* lies in the block); each lane's counts become its slot
//           cursors (a counting sort), stored as the query's own LDS row;
//   pass 2  (wave = query) the query's own sub-block of radius sqrt((b* + 1) / scale) is
//           streamed, lane = candidate; the items of buckets <= b* go to LDS staging slots in
//           bucket order (the same FP64 keys as pass 1, so exactly cum(b*) of them), each is
//           ranked against the items of its own bucket only (bucket order is key order), and
//           the first K+1 are written to the row.  The sub-blocks' row ranges are looked up for
//           the whole tile at once (lane = query).
//   A query whose staged list exceeds `smax` LDS slots (e.g. thousands of duplicates in one
//   bucket), or with 65536 or more candidates within Tq (16-bit counters), is flagged for the
//   overflow kernel (range.hip: the collect kernel, or the heap kernel above its K+1).
//
// This file is compiled once per KNN_SELECT_PART (0..5): float / double x three ranges of D,
// each instantiating the kernel for every number of indexed dims 1..4 (all per-dim arrays in
// registers) and for the grid's runtime count (5..8 indexed dims).
#include "knn_range.h"

#ifndef KNN_SELECT_PART
#define KNN_SELECT_PART 0
#endif

namespace knn {
namespace {

__device__ __forceinline__ uint32_t kbucket(double key, double scale) {
  const double t = fmin(key * scale, (double)SEL_NBIN + 0.5);
  return (uint32_t)t;
}

// Streams `total` candidates in rounds of 64, lane = candidate (fi = round base + lane; lanes
// past the end load the last candidate again): f(cv, m) gets the coordinates of a round and
// its candidate count m (wave-uniform).  The next round's loads are issued before f runs on
// the current one.  pos_of(fi) = sorted position of the fi-th candidate.
template <typename T, int D, typename P, typename F>
__device__ __forceinline__ void stream_rounds(const T *__restrict__ xs, size_t n, uint32_t total, uint32_t lane,
                                              P &&pos_of, F &&f) {
  if (total == 0) return;
  auto load = [&](T (&v)[D], uint32_t f0) {
    uint32_t fi = f0 + lane;
    fi = fi < total ? fi : total - 1;
    const uint32_t p = pos_of(fi);
#pragma unroll
    for (int j = 0; j < D; j++) v[j] = xs[(size_t)j * n + p];
  };
  T cv[D];
  load(cv, 0);
  for (uint32_t f0 = 0; f0 < total; f0 += 64) {
    T nx[D];
    const bool more = f0 + 64 < total;
    if (more) load(nx, f0 + 64);
    f(cv, total - f0 < 64u ? total - f0 : 64u);
    if (more) {
#pragma unroll
      for (int j = 0; j < D; j++) cv[j] = nx[j];
    }
  }
}

template <typename T, int D, int ND, typename G>
__global__ __launch_bounds__(64) void k_select(JoinArgs a, G g0) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t *H = reinterpret_cast<uint32_t *>(smem);
  uint32_t *rowB = H + SEL_HWORDS;
  uint32_t *rowX = rowB + 64;
  uint32_t *prB = rowX + 64;
  uint32_t *prX = prB + SEL_RMAX * 64;
  uint32_t *prT = prX + SEL_RMAX * 64;
  uint32_t *prR = prT + 64;
  uint32_t *qc = prR + 64;  // the collecting query's slot cursors, one 32-bit word per bucket
  double *cbuf = reinterpret_cast<double *>(prB);  // pass 1: the round's candidates (64 x D)
  float *cbufF = reinterpret_cast<float *>(prB);
  const uint32_t SM = a.smax;
  double *skey = reinterpret_cast<double *>(smem + SEL_FIXED_WORDS * sizeof(uint32_t));
  uint32_t *sid = reinterpret_cast<uint32_t *>(skey + SM);
  T *ss = reinterpret_cast<T *>(sid + SM + (SM & 1));
  const uint32_t lane = threadIdx.x;
  const T *__restrict__ xs = reinterpret_cast<const T *>(a.xs);
  const size_t n = a.n;
  const uint32_t NT = tile_count(a);
  // pass-2 loads: with ND > 0 a scalar base per array plus a 32-bit byte offset per lane (one
  // shift per candidate instead of 64-bit address arithmetic per array); the host launches
  // those builds only when every array fits in 4 GiB, and the ND = 0 build otherwise
  const char *xsb[D];
#pragma unroll
  for (int j = 0; j < D; j++) xsb[j] = reinterpret_cast<const char *>(xs + (size_t)j * n);
  const char *idb = reinterpret_cast<const char *>(a.sortedIds);
  // With more data dims than indexed dims (float), the grid prunes only a projection and most
  // candidates of a block lie outside a query's ball: their FP32 sum decides that before any
  // FP64 work (5M points, K = 32, plain schedule, with the LDS broadcast below: 2 of 4 dims
  // indexed, uniform 855 -> 597 ms, skewed 2658 -> 2044 ms; 4 of 6 dims indexed, uniform
  // 1261 -> 1074 ms).
  constexpr bool kPrefilter = std::is_same<T, float>::value && (ND == 0 || D > ND);
  // Pass 1 broadcasts each round's candidates as FP64 coordinates through LDS (one uniform-
  // address read per candidate instead of D readlanes and D conversions per lane); the round
  // buffer uses the pass-2 row tables (prB/prX, 2.5 KB), which are free until the scan.
  constexpr bool kLdsBcast = !kPrefilter && D <= 4;
  static_assert(!kLdsBcast || SEL_RMAX * 2 * 64 * 4 >= 64 * D * 8, "round buffer exceeds the row tables");
  // with the pre-filter, the round's FP32 coordinates go through LDS the same way
  constexpr bool kLdsBcastF = kPrefilter && D <= 8;
  static_assert(!kLdsBcastF || SEL_RMAX * 2 * 64 * 4 >= 64 * D * 4, "round buffer exceeds the row tables");
  // pass-1 histogram, lane-pair layout: the counter of bucket b of lane L is the (L & 1) half
  // of word b * 32 + (L >> 1), so a lane's update address is one shift-add of its bucket
  uint32_t *hcol = H + (lane >> 1);
  const uint32_t hval = (lane & 1u) ? 0x10000u : 1u;
  const uint32_t hsh = (lane & 1u) << 4;

  WaveTally tally;
  for (;;) {
    const uint32_t t = next_tile(a.tileCounter, NT, a.tileRanges, lane);
    if (t >= NT) break;
    const uint4 tl = a.tiles[t];
    const TileCtx<G> C = tile_ctx<G>(a, g0, tl);
    const G &g = C.g;
    const uint32_t nd = nd_of<ND>(g);
    const uint32_t K1 = C.k1;
    const T eps = (T)g.eps;
    Lane<T, D> L;
    uint32_t lo[nd_cap<ND>()], hi[nd_cap<ND>()];
    tile_setup<T, D, ND, G>(a, g, tl, lane, L, lo, hi);
    const double scale = L.Tq > 0.0 ? (double)SEL_NBIN / L.Tq : __builtin_inf();
    // slab order: the union of the lanes' slab-code ranges (radius r, as the cell ranges)
    // (wave-uniform; the runtime-dims build streams whole cells: narrowing is optional, every
    // cut point lies outside the radius)
    const uint32_t sb = ND > 0 ? g.sb : 0u;
    const T *__restrict__ xsd = xs + (size_t)nd * n;  // read only when sb > 0 (nd < D then)
    uint32_t uql = 0, uqh = 0;
    if (sb) {
      const T r = a.radius ? reinterpret_cast<const T *>(a.radius)[L.q] : (T)a.radiusAll;
      const T xq = pick_t<T, D>(L.qc, nd);
      uql = wave_min_u32(L.active ? slab_q<T>(next_down(xq - r), g) : 0xffffffffu);
      uqh = wave_max_u32(L.active ? slab_q<T>(next_up(xq + r), g) : 0u);
    }
    wave_lds_sync();  // the previous tile's readers of H and the staging arrays are done
#pragma unroll 8
    for (uint32_t w = 0; w < SEL_NB / 2; w++) H[w * 64 + lane] = 0u;

    // ---- pass 1 (lane = query): histogram of the FP64 keys of the union of the tile's blocks.
    // While the tile has seen fewer than 65536 candidates no 16-bit counter can carry; past
    // that each lane counts its candidates within Tq (from its counters at that point on).
    uint64_t seen = 0;
    bool counting = false;
    uint32_t inF = 0;
    auto add1 = [&](const T (&cv)[D], uint32_t jj) -> uint32_t {
      if constexpr (kPrefilter) {
        // FP32 sum first: a candidate with s > F has key > Tq, i.e. bucket 63, which is never
        // read, so only the lanes within F take the FP64 key and the histogram update (and
        // the wave skips both when none is)
        T c[D];
#pragma unroll
        for (int j = 0; j < D; j++) c[j] = readlane<T>(cv[j], (int)jj);
        uint32_t b = (uint32_t)SEL_NBIN;
        if (Num<T>::template sq<D>(c, L.qc) <= L.F) {
          b = kbucket(fp64_key<D, T>(c, L.qd), scale);
          atomicAdd(hcol + b * 32u, hval);
        }
        return b;
      } else {
        double key = 0.0;
#pragma unroll
        for (int j = 0; j < D; j++) {
          double cj;
          if constexpr (kLdsBcast)
            cj = cbuf[jj * D + j];
          else
            cj = (double)readlane<T>(cv[j], (int)jj);
          const double dx = cj - L.qd[j];
          key = key + dx * dx;
        }
        const uint32_t b = kbucket(key, scale);
        atomicAdd(hcol + b * 32u, hval);
        return b;
      }
    };
    // a full round with the LDS broadcast: the coordinates of the next GRP candidates are read
    // before the current GRP's histogram updates are issued (LDS reads and atomics keep their
    // program order, so a read issued after an update would wait for it)
    auto add64 = [&]() {
      constexpr int GRP = D <= 2 ? 4 : 2;
      double nx[GRP][D];
#pragma unroll
      for (int k = 0; k < GRP; k++)
#pragma unroll
        for (int j = 0; j < D; j++) nx[k][j] = cbuf[k * D + j];
#pragma unroll
      for (int g = 0; g < 64 / GRP; g++) {
        double cu[GRP][D];
#pragma unroll
        for (int k = 0; k < GRP; k++)
#pragma unroll
          for (int j = 0; j < D; j++) cu[k][j] = nx[k][j];
        if (g + 1 < 64 / GRP) {
#pragma unroll
          for (int k = 0; k < GRP; k++)
#pragma unroll
            for (int j = 0; j < D; j++) nx[k][j] = cbuf[((g + 1) * GRP + k) * D + j];
        }
#pragma unroll
        for (int k = 0; k < GRP; k++) {
          double key = 0.0;
#pragma unroll
          for (int j = 0; j < D; j++) {
            const double dx = cu[k][j] - L.qd[j];
            key = key + dx * dx;
          }
          atomicAdd(hcol + kbucket(key, scale) * 32u, hval);
        }
      }
    };
    // the same with the pre-filter: the FP32 sum (fused multiply-adds: within 2^-19 of the
    // key, like the reported-distance sum) decides whether the FP64 key is needed
    auto add64f = [&]() {
      constexpr int GRP = 2;
      float qf[D];
#pragma unroll
      for (int j = 0; j < D; j++) qf[j] = (float)L.qc[j];
      float nx[GRP][D];
#pragma unroll
      for (int k = 0; k < GRP; k++)
#pragma unroll
        for (int j = 0; j < D; j++) nx[k][j] = cbufF[k * D + j];
#pragma unroll
      for (int g = 0; g < 64 / GRP; g++) {
        float cu[GRP][D];
#pragma unroll
        for (int k = 0; k < GRP; k++)
#pragma unroll
          for (int j = 0; j < D; j++) cu[k][j] = nx[k][j];
        if (g + 1 < 64 / GRP) {
#pragma unroll
          for (int k = 0; k < GRP; k++)
#pragma unroll
            for (int j = 0; j < D; j++) nx[k][j] = cbufF[((g + 1) * GRP + k) * D + j];
        }
#pragma unroll
        for (int k = 0; k < GRP; k++) {
          float sf = 0.0f;
#pragma unroll
          for (int j = 0; j < D; j++) {
            const float dj = cu[k][j] - qf[j];
            sf = __builtin_fmaf(dj, dj, sf);
          }
          if (sf <= (float)L.F) {
            double key = 0.0;
#pragma unroll
            for (int j = 0; j < D; j++) {
              const double dx = (double)cu[k][j] - L.qd[j];
              key = key + dx * dx;
            }
            atomicAdd(hcol + kbucket(key, scale) * 32u, hval);
          }
        }
      }
    };
    auto hist_round = [&](const T (&cv)[D], uint32_t m) {
      if (!counting && seen + m >= 65536u) {
        counting = true;
        inF = 0;
#pragma unroll 8
        for (uint32_t b = 0; b < (uint32_t)SEL_NBIN; b++) inF += (hcol[b * 32u] >> hsh) & 0xffffu;
      }
      seen += m;
      if constexpr (kLdsBcast) {
        wave_lds_sync();  // the previous round's readers are done
#pragma unroll
        for (int j = 0; j < D; j++) cbuf[lane * D + j] = (double)cv[j];
        wave_lds_sync();
      } else if constexpr (kLdsBcastF) {
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < D; j++) cbufF[lane * D + j] = (float)cv[j];
        wave_lds_sync();
      }
      if (!counting) {
        if (m == 64u) {
          if constexpr (kLdsBcast) {
            add64();
          } else if constexpr (kLdsBcastF) {
            add64f();
          } else {
#pragma unroll 16
            for (uint32_t jj = 0; jj < 64u; jj++) add1(cv, jj);
          }
        } else {
          for (uint32_t jj = 0; jj < m; jj++) add1(cv, jj);
        }
      } else {
        for (uint32_t jj = 0; jj < m; jj++) inF += add1(cv, jj) < (uint32_t)SEL_NBIN ? 1u : 0u;
      }
    };
    // 64 segments of sorted positions (lane = segment: start pB, length len; lanes from
    // `left` on empty) concatenated and streamed
    auto stream_batch = [&](uint32_t pB, uint32_t len, uint64_t left) {
      const uint32_t inc = wave_incl_sum(len, lane);
      const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
      const uint32_t excl = inc - len;
      if (left <= 8) {
        uint32_t X[8], Bs[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          X[k] = (uint32_t)__builtin_amdgcn_readlane((int)excl, k);
          Bs[k] = (uint32_t)__builtin_amdgcn_readlane((int)pB, k);
        }
        const uint32_t nr = (uint32_t)left;
        stream_rounds<T, D>(xs, n, total, lane, [&](uint32_t fi) {
          uint32_t b = Bs[0], x = X[0];
#pragma unroll
          for (int k = 1; k < 8; k++)
            if ((uint32_t)k < nr && fi >= X[k]) {
              b = Bs[k];
              x = X[k];
            }
          return b + (fi - x);
        }, hist_round);
      } else {
        rowB[lane] = pB;
        rowX[lane] = excl;
        wave_lds_sync();
        stream_rounds<T, D>(xs, n, total, lane, [&](uint32_t fi) {
          uint32_t l = 0, h = 64;
          while (h - l > 1) {
            const uint32_t mid = (l + h) >> 1;
            if (rowX[mid] <= fi) l = mid; else h = mid;
          }
          return rowB[l] + (fi - rowX[l]);
        }, hist_round);
        wave_lds_sync();
      }
    };
    if constexpr (ND == 1 || ND == 2) {
      // the union's rows (no chords below 3 indexed dims) concatenated: rounds mix rows; with
      // slab order, the rows' cells, each cut to the union's slab-code range
      const bool cut = sb && hi[0] - lo[0] < SLAB_MAX_SPAN;
      const uint64_t R = block_rows<ND>(g, lo, hi);
      const uint64_t RS = cut ? R * (uint64_t)(hi[0] - lo[0] + 1) : R;
      for (uint64_t rb = 0; rb < RS; rb += 64) {
        uint32_t pB, len;
        if (cut)
          seg_batch<T, D, ND, G>(a, g, lo, hi, false, L.qd, 0.0, RS, rb + lane, uql, uqh, xsd, pB, len);
        else
          row_batch<T, D, ND, G>(a, g, lo, hi, false, L.qd, 0.0, R, rb + lane, pB, len);
        stream_batch(pB, len, RS - rb);
      }
    } else {
      // rows one at a time, narrowed to the wave's union of chords of radius sqrt(Tq)
      const double rF2 = L.Tq * (1.0 + 0x1p-15);
      auto narrow1 = [&](const uint32_t *cur, uint32_t &l0, uint32_t &h0) {
        return nd < 3 ? true : row_chord<T, D, ND, G>(g, cur, L.qd, rF2, L.active, l0, h0);
      };
      if (sb && hi[0] - lo[0] < SLAB_MAX_SPAN) {
        // with slab order: the row's cells (lane = cell), each cut to the union's slab range
        for_each_row_span<ND>(g, lo, hi, narrow1, [&](uint64_t rowBase, uint32_t l0, uint32_t h0) {
          for (uint32_t c0 = l0;; c0 += 64) {
            const uint32_t x0 = c0 + lane;
            uint32_t pB = 0, len = 0;
            if (x0 <= h0) {
              uint32_t pE;
              row_range(a, rowBase, x0, x0, pB, pE);
              if (pE > pB) slab_narrow<T>(xsd, g, uql, uqh, pB, pE);
              len = pE - pB;
            }
            stream_batch(pB, len, (uint64_t)(h0 - c0) + 1);
            if (h0 - c0 < 64) break;
          }
        });
      } else {
        for_each_row<ND>(a, g, lo, hi, narrow1, [&](uint32_t pBeg, uint32_t pEnd) {
          stream_rounds<T, D>(xs, n, pEnd - pBeg, lane, [&](uint32_t fi) { return pBeg + fi; }, hist_round);
        });
      }
    }

    // ---- scan: b* per query; each lane's counts become exclusive slot offsets (saturated past
    // b*), written as the query's row of cursors: lane L's row at word L * SEL_RS
    // a lane with 65536 or more candidates within Tq may have carried into its pair partner's
    // counters (the high half of the same words): both go to the overflow kernel
    const bool bigOwn = counting && inF >= 65536u;
    const bool big = bigOwn || __shfl_xor((int)bigOwn, 1, 64) != 0;
    uint32_t cum = 0, bstar = SEL_NB, cumB = 0;
    uint32_t col[SEL_NB / 2];
#pragma unroll
    for (uint32_t b = 0; b < (uint32_t)SEL_NB; b++) {
      const uint32_t h = (hcol[b * 32u] >> hsh) & 0xffffu;
      const uint32_t o = cum < 0xffffu ? cum : 0xffffu;
      if (b & 1u) col[b >> 1] |= o << 16; else col[b >> 1] = o;
      if (b < (uint32_t)SEL_NBIN && bstar == (uint32_t)SEL_NB && cum + h >= K1) {
        bstar = b;
        cumB = cum + h;
      }
      cum += h;
    }
    wave_lds_sync();  // every lane has read its counters
    {
      uint32_t *myRow = H + lane * SEL_RS;
#pragma unroll
      for (uint32_t w = 0; w < (uint32_t)SEL_NB / 2; w++) myRow[w] = col[w];
    }
    // buckets <= b* alone exceed the staging slots: the overflow kernel takes the query
    const bool preOvf = L.active && !big && bstar < (uint32_t)SEL_NBIN && cumB > SM;
    const bool go = L.active && !big && bstar < (uint32_t)SEL_NBIN && !preOvf;
    // every staged key is below (b* + 1) / scale (1 + 2^-51): the sub-block's reach
    const double edge = go ? (double)(bstar + 1) / scale * (1.0 + 0x1p-40) : 0.0;
    bool okMine = false, ovfMine = preOvf;
    const uint64_t goMask = __ballot(go);

    // ---- pass 2 set-up (lane = query): the rows of each query's sub-block and their point
    // ranges, for the whole tile at once (one memory round trip per tile instead of two per
    // query).  Queries with more than RMAX rows, grids without the dense start table and grids
    // in slab order take the per-query path.
    const T riL = go ? (T)next_up((T)sqrt(edge)) : (T)0;
    bool fastL = false;
    if (go && a.denseStart && !sb) {
      uint32_t sl[nd_cap<ND>()], sh[nd_cap<ND>()];
      uint32_t R = 1;
#pragma unroll
      for (uint32_t j = 0; j < nd_cap<ND>(); j++) {
        if (j >= nd) break;
        const T xj = pick_t<T, D>(L.qc, j);
        sl[j] = cell_coord<T>(next_down(xj - riL), (T)g.minArr[j], eps, g.nCells[j]);
        sh[j] = cell_coord<T>(next_up(xj + riL), (T)g.minArr[j], eps, g.nCells[j]);
        if (j >= 1) {
          const uint32_t sp = sh[j] - sl[j] + 1;
          R = (R <= SEL_RMAX && sp <= SEL_RMAX) ? R * sp : SEL_RMAX + 1;
        }
      }
      if (R <= SEL_RMAX) {
        fastL = true;
        uint32_t cur[nd_cap<ND>()];
#pragma unroll
        for (uint32_t j = 0; j < nd_cap<ND>(); j++) cur[j] = sl[j < nd ? j : 0];
        const double ri2L = (double)riL * (double)riL * (1.0 + 0x1p-15);
        uint32_t acc = 0;
#pragma unroll
        for (uint32_t r = 0; r < SEL_RMAX; r++) {
          if (r < R) {
            uint32_t l0 = sl[0], h0 = sh[0];
            // each row narrowed to the chord of the query's ball (one query per lane here, so
            // the chords cost a few operations per row of a tile, and drop a quarter or more of
            // the sub-block's cells with 2 indexed dims)
            const bool keep = nd >= 2 ? row_chord_lane<T, D, ND, G>(g, cur, L.qd, ri2L, l0, h0) : true;
            uint32_t pB = 0, len = 0;
            if (keep) {
              const uint64_t rowBase = row_base<ND>(g, cur);
              pB = a.denseStart[rowBase + l0];
              len = a.denseStart[rowBase + h0 + 1] - pB;
            }
            prB[r * 64 + lane] = pB;
            prX[r * 64 + lane] = acc;
            acc += len;
            next_row<ND>(cur, sl, sh, nd);
          }
        }
        prT[lane] = acc;
        prR[lane] = R;
      }
    }
    const uint64_t fastMask = __ballot(fastL);
    wave_lds_sync();

    // ---- pass 2 (wave = query): for each query, stream its own sub-block, stage the items of
    // buckets <= b* in LDS in bucket order, rank each within its bucket, write the row.
    // (Loading the next query's first rounds before the current query's rank was measured
    // slower: the kernel is bound by instruction issue, not by that memory round trip.)
#ifndef KNN_SELECT_PR
#define KNN_SELECT_PR 3
#endif
    constexpr int PR = D <= 4 ? KNN_SELECT_PR : (D <= 8 ? 2 : 1);
    // a fast-path query's sub-block rows: candidate fi of the concatenated rows sits at
    // fi + dlt0 + the steps stp[k] of the rows k it has passed (X[k] <= fi; rows past the
    // query's count never start)
    struct Rows {
      uint32_t tot, dlt0, X[SEL_RMAX], stp[SEL_RMAX];
    };
    auto rows_of = [&](uint32_t q) {
      Rows w;
      const uint32_t Rq = prR[q];
      w.tot = prT[q];
      w.dlt0 = prB[q];
      w.X[0] = 0;
      w.stp[0] = 0;
      uint32_t prev = w.dlt0;
#pragma unroll
      for (uint32_t k = 1; k < SEL_RMAX; k++) {
        w.X[k] = k < Rq ? prX[k * 64 + q] : 0xffffffffu;
        const uint32_t dk = prB[k * 64 + q] - w.X[k];
        w.stp[k] = dk - prev;
        prev = dk;
      }
      return w;
    };
    // a fast-path query's positions of rounds f0 .. f0 + 64 PR - 1 (lanes past the end take the
    // last candidate again; rounds past the end are loaded too: guarding them costs more than
    // the redundant loads)
    auto fast_pos = [&](const Rows &w, uint32_t f0, uint32_t (&pv)[PR], bool (&val)[PR]) {
#pragma unroll
      for (int u = 0; u < PR; u++) {
        uint32_t fi = f0 + 64u * u + lane;
        val[u] = fi < w.tot;
        fi = val[u] ? fi : w.tot - 1;
        uint32_t p = fi + w.dlt0;
#pragma unroll
        for (uint32_t k = 1; k < SEL_RMAX; k++) p += fi >= w.X[k] ? w.stp[k] : 0u;
        pv[u] = p;
      }
    };
    auto load_group = [&](const uint32_t (&pv)[PR], T (&cv)[PR][D], uint32_t (&cid)[PR]) {
      if constexpr (ND > 0) {
#pragma unroll
        for (int u = 0; u < PR; u++) {
          const uint32_t off = pv[u] * (uint32_t)sizeof(T);
#pragma unroll
          for (int j = 0; j < D; j++) cv[u][j] = *reinterpret_cast<const T *>(xsb[j] + off);
          cid[u] = *reinterpret_cast<const uint32_t *>(idb + pv[u] * 4u);
        }
      } else {
#pragma unroll
        for (int u = 0; u < PR; u++) {
#pragma unroll
          for (int j = 0; j < D; j++) cv[u][j] = xs[(size_t)j * n + pv[u]];
          cid[u] = a.sortedIds[pv[u]];
        }
      }
    };
    // every loaded value is consumed here, unconditionally: the compiler may not sink a load
    // into the staging branch that uses it (a memory round trip per round)
    auto consume = [&](T (&cv)[PR][D], uint32_t (&cid)[PR]) {
#pragma unroll
      for (int u = 0; u < PR; u++) {
#pragma unroll
        for (int j = 0; j < D; j++) asm volatile("" ::"v"(cv[u][j]));
        asm volatile("" ::"v"(cid[u]));
      }
    };
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
      // FP32 image of the sub-block's key bound: a candidate with s > Fe has a bucket above b*
      const T Fe = Num<T>::bound(readlane<double>(edge, (int)i));
      {
        // the query's row of 16-bit cursors (lane i's exclusive offsets) widened to 32 bits,
        // lane = bucket: a slot claim is then one atomic add of 1 on its bucket's word
        const uint32_t *qh = H + i * SEL_RS;
        qc[lane] = (qh[lane >> 1] >> ((lane & 1u) << 4)) & 0xffffu;
        wave_lds_sync();
      }
      bool ovf = false;
      uint32_t nStaged = 0;  // wave-uniform: staging runs on every lane, valid or not
      // stages the candidates at positions pv of PR rounds (round u counts when
      // f0 + 64 u < tot): loads, then the rounds' keys and buckets, then every slot claim, then
      // the staging writes, so the PR rounds' loads and LDS atomics are in flight together
      auto stage_group = [&](const uint32_t (&pv)[PR], const bool (&val)[PR], uint32_t f0, uint32_t tot) {
        T cv[PR][D];
        uint32_t cidv[PR];
        load_group(pv, cv, cidv);
        consume(cv, cidv);
        double keyv[PR];
        uint32_t bv[PR], slot[PR];
        bool take[PR];
        T sv[PR];
#pragma unroll
        for (int u = 0; u < PR; u++) {
          take[u] = false;
          if (f0 + 64u * u < tot) {  // wave-uniform
            if constexpr (kPrefilter) {
              sv[u] = Num<T>::template sq<D>(cv[u], qi);
              keyv[u] = 0.0;
              bv[u] = (uint32_t)SEL_NB;
              if (val[u] && sv[u] <= Fe) {
                keyv[u] = fp64_key<D, T>(cv[u], qdi);
                bv[u] = kbucket(keyv[u], sci);
              }
            } else {
              keyv[u] = fp64_key<D, T>(cv[u], qdi);
              bv[u] = kbucket(keyv[u], sci);
            }
            take[u] = val[u] && bv[u] <= bsi;
            nStaged += (uint32_t)__popcll(__ballot(take[u]));
          }
        }
#pragma unroll
        for (int u = 0; u < PR; u++) {
          slot[u] = 0;
          if (take[u]) slot[u] = atomicAdd(&qc[bv[u]], 1u);
        }
#pragma unroll
        for (int u = 0; u < PR; u++) {
          if (take[u]) {
            if (slot[u] < SM) {
              skey[slot[u]] = keyv[u];
              sid[slot[u]] = cidv[u];
              if constexpr (kPrefilter)
                ss[slot[u]] = sv[u];
              else
                ss[slot[u]] = Num<T>::template sq<D>(cv[u], qi);
            } else {
              ovf = true;
            }
          }
        }
      };
      if ((fastMask >> i) & 1ull) {
        const Rows w = rows_of(i);
        const uint32_t Rq = prR[i];
        if (w.tot >= Rq * SEG_LONG) {
          // long rows: row by row
          for (uint32_t k = 0; k < Rq; k++) {
            const uint32_t xk = prX[k * 64 + i];
            const uint32_t xe = k + 1 < Rq ? prX[(k + 1) * 64 + i] : w.tot;
            if (xe > xk) stream_row<PR>(prB[k * 64 + i], xe - xk, lane, stage_group);
          }
        } else {
          for (uint32_t f0 = 0; f0 < w.tot; f0 += 64u * PR) {
            uint32_t pv[PR];
            bool val[PR];
            fast_pos(w, f0, pv, val);
            stage_group(pv, val, f0, w.tot);
          }
        }
      } else {
        // the sub-block's rows (with slab order: their cells, each cut to the query's slab
        // range), 64 per batch (lane = row or cell), concatenated
        const T ri = readlane<T>(riL, (int)i);
        uint32_t slo[nd_cap<ND>()], shi[nd_cap<ND>()];
#pragma unroll
        for (uint32_t j = 0; j < nd_cap<ND>(); j++) {
          if (j >= nd) break;
          const T xj = pick_t<T, D>(qi, j);
          slo[j] = cell_coord<T>(next_down(xj - ri), (T)g.minArr[j], eps, g.nCells[j]);
          shi[j] = cell_coord<T>(next_up(xj + ri), (T)g.minArr[j], eps, g.nCells[j]);
        }
        const double ri2 = (double)ri * (double)ri * (1.0 + 0x1p-15);
        const uint64_t R = block_rows<ND>(g, slo, shi);
        const bool cut = sb && shi[0] - slo[0] < SLAB_MAX_SPAN;
        const uint64_t RS = cut ? R * (uint64_t)(shi[0] - slo[0] + 1) : R;
        uint32_t sql = 0, sqh = 0;
        if (cut) {
          const T xq = pick_t<T, D>(qi, nd);
          sql = slab_q<T>(next_down(xq - ri), g);
          sqh = slab_q<T>(next_up(xq + ri), g);
        }
        for (uint64_t rb = 0; rb < RS; rb += 64) {
          uint32_t pB, len;
          if (cut)
            seg_batch<T, D, ND, G>(a, g, slo, shi, nd >= 2, qdi, ri2, RS, rb + lane, sql, sqh, xsd, pB, len);
          else
            row_batch<T, D, ND, G>(a, g, slo, shi, nd >= 2, qdi, ri2, R, rb + lane, pB, len);
          stream_segments<PR>(pB, len, lane, rowB, rowX, a.segCapLog2, stage_group);
        }
      }
      wave_lds_sync();
      // the staged count equals pass 1's cum(b*) (same keys, same buckets); anything else
      // (an overflowing slot) sends the query to the overflow kernel
      if (__ballot(ovf) || nStaged != m) {
        if (lane == i) ovfMine = true;
        wave_lds_sync();
        continue;
      }
      // rank: an item of bucket b follows every item of the buckets below and precedes every
      // item of the buckets above, so it is compared with its own bucket only (whose slots are
      // [end of bucket b-1, end of bucket b), the cursors after staging).  The row is written
      // through a scalar base and the rank's byte offset (r < K+1 <= 2^31).
      char *rowI = reinterpret_cast<char *>(a.outIds + C.row(qid));
      char *rowD = reinterpret_cast<char *>(reinterpret_cast<T *>(a.outDists) + C.row(qid));
      for (uint32_t t0 = 0; t0 < m; t0 += 64) {
        const uint32_t tI = t0 + lane;
        if (tI < m) {
          const double kt = skey[tI];
          const uint32_t it = sid[tI];
          const T st = ss[tI];
          const uint32_t b = kbucket(kt, sci);
          // both cursors read at once (the empty asm consumes both loads where they are issued,
          // so neither is deferred into a branch of its own)
          const uint32_t wa = qc[b ? b - 1 : 0u], w1 = qc[b];
          asm volatile("" ::"v"(wa), "v"(w1));
          const uint32_t w0 = b ? wa : 0u;
          uint32_t r = w0;
          for (uint32_t v = w0; v < w1; v += 2) {
            // two slots per step, loaded together; slot v + 1 <= w1 <= smax stays inside the
            // staging arrays and counts only when it belongs to the bucket
            const double k0 = skey[v], k1 = skey[v + 1];
            const uint32_t i0 = sid[v], i1 = sid[v + 1];
            const bool l0 = (k0 < kt) | ((k0 == kt) & (i0 < it));
            const bool l1 = (k1 < kt) | ((k1 == kt) & (i1 < it));
            r += (l0 ? 1u : 0u) + ((l1 & (v + 1 < w1)) ? 1u : 0u);
          }
          if (r < K1) {
            *reinterpret_cast<int32_t *>(rowI + r * 4u) = (int32_t)(it - C.idBase);
            *reinterpret_cast<T *>(rowD + r * (uint32_t)sizeof(T)) = sqrt(st);
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
    tile_accounting<T>(tally, tl, okMine, L.active && !okMine && !ovfMine, seen, K1);
  }
  flush_tally(a, tally, lane);
}

template <typename T, int D, int ND, typename G>
void launch_nd(const JoinArgs &a, const G &g, hipStream_t s) {
  const uint32_t grid = a.ntilesDev ? a.ntiles : range_select_grid(a.ntiles);
  k_select<T, D, ND, G><<<grid, 64, range_select_lds_bytes<T>(a.smax, (uint32_t)SEL_NB), s>>>(a, g);
}

}  // namespace

// Launches the select kernel for D data dims and nd indexed dims (nd <= 4 on a kernel built
// for that count, else the runtime-count kernel).
template <typename T, int D, int NB>
void select_launch(const JoinArgs &a, const typename GridOf<T>::type &g, uint32_t nd, hipStream_t s) {
  static_assert(NB == SEL_NB, "select_launch is built for this unit's bucket count");
  using G = typename GridOf<T>::type;
  if (a.ntiles == 0) return;
  if ((uint64_t)a.n * sizeof(T) > 0xffffffffull) {
    launch_nd<T, D, 0, G>(a, g, s);  // 64-bit addressing throughout
  } else if (nd == 1) {
    launch_nd<T, D, 1, G>(a, g, s);
  } else if (nd == 2) {
    if constexpr (D >= 2) launch_nd<T, D, 2, G>(a, g, s);
  } else if (nd == 3) {
    if constexpr (D >= 3) launch_nd<T, D, 3, G>(a, g, s);
  } else if (nd == 4) {
    if constexpr (D >= 4) launch_nd<T, D, 4, G>(a, g, s);
  } else {
    launch_nd<T, D, 0, G>(a, g, s);
  }
}

#define SEL_INST(T, D) \
  template void select_launch<T, D, SEL_NB>(const JoinArgs &, const GridOf<T>::type &, uint32_t, hipStream_t);
#if KNN_SELECT_PART < 3
#define KNN_SEL_T float
#else
#define KNN_SEL_T double
#endif
#if KNN_SELECT_PART % 3 == 0
SEL_INST(KNN_SEL_T, 1)
SEL_INST(KNN_SEL_T, 2)
SEL_INST(KNN_SEL_T, 3)
SEL_INST(KNN_SEL_T, 4)
SEL_INST(KNN_SEL_T, 5)
#elif KNN_SELECT_PART % 3 == 1
SEL_INST(KNN_SEL_T, 6)
SEL_INST(KNN_SEL_T, 7)
SEL_INST(KNN_SEL_T, 8)
SEL_INST(KNN_SEL_T, 9)
SEL_INST(KNN_SEL_T, 10)
#else
SEL_INST(KNN_SEL_T, 11)
SEL_INST(KNN_SEL_T, 12)
SEL_INST(KNN_SEL_T, 13)
SEL_INST(KNN_SEL_T, 14)
SEL_INST(KNN_SEL_T, 15)
SEL_INST(KNN_SEL_T, 16)
#endif

}  // namespace knn
