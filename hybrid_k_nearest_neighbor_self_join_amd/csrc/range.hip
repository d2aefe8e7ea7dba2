// range.hip -- the collect and heap range kernels for gfx950, and the range-kernel dispatch.
//
// Replace kernel.cu:61-225 (range query over the adjacent cells), the pair buffer and the two
// thrust::stable_sort_by_key of GPU.cu:1629-1634, and the host top-k of GPU.cu:3610-3704, with
// the select kernel (select.hip):
// k_range_collect (5 or more indexed dims): wave = query, one pass over the query's own block
//   into the LDS top-(K+1) collector of knn_collect.h, FP32-prefiltered against its threshold.
// k_range_heap: per-lane max-heap of (key, id), in LDS when K+1 fits, else in global scratch;
//   takes the select kernel's overflow queries above the collector's K+1 (the collect kernel
//   takes them below).
#include "knn_collect.h"
#include "knn_range.h"

namespace knn {

namespace {

// ====================================================================== collect kernel
// Wave = query, for the queries of a tile in turn: the query's own block (radius r; rows
// narrowed to the chords of its ball with 3 or more indexed dims) is streamed 64 candidates at
// a time, PR rounds of coordinates and ids loaded together; the FP32/T sum is compared with the
// T bound of the collector's threshold (first Tq = r^2 (1 - 2^-20), then the (K+1)-th key
// held), and only candidates that pass get their FP64 key and enter the LDS collector.  The
// query succeeds when K+1 pairs with key <= Tq were collected: they are then its exact
// top-(K+1), since every point with key <= Tq lies in the block.  One pass over the
// candidates, against the select kernel's two (the tile's union, then the query's sub-block):
// the better choice when a query has many candidates per neighbour, as with 5 or more indexed
// dims.  Built per number of indexed dims ND where that equals D or is 6 (every per-dim array
// then lives in registers), else for the runtime count (ND = 0).  On a grid in slab order the
// fixed-dims builds cut each cell of the block to the query's slab, as the select kernel does.
template <typename T, int D, int ND, typename G>
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
  constexpr int PR = D <= 8 ? 4 : 2;
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
    uint32_t nok = 0;
    unsigned long long cand = 0;
    for (uint32_t i = 0; i < tl.z; i++) {
      const uint32_t q = (uint32_t)__builtin_amdgcn_readfirstlane((int)tile_query(a, tl, i));
      T qc[D];
      double qd[D];
#pragma unroll
      // the query's values are wave-uniform: scalar registers (the vector registers are what
      // bounds this kernel's residency)
      for (int j = 0; j < D; j++) {
        qc[j] = readfirst<T>(x[(size_t)j * n + q]);
        qd[j] = readfirst<double>((double)qc[j]);
      }
      const T r = readfirst<T>(a.radius ? reinterpret_cast<const T *>(a.radius)[q] : (T)a.radiusAll);
      const double Tq = (double)r * (double)r * (1.0 - 0x1p-20);
      const T F = Num<T>::bound(Tq);
      uint32_t lo[nd_cap<ND>()], hi[nd_cap<ND>()];
#pragma unroll
      for (uint32_t j = 0; j < nd_cap<ND>(); j++) {
        if (j >= nd) break;
        const T xj = pick_t<T, D>(qc, j);
        lo[j] = readfirst<uint32_t>(cell_coord<T>(next_down(xj - r), (T)g.minArr[j], eps, g.nCells[j]));
        hi[j] = readfirst<uint32_t>(cell_coord<T>(next_up(xj + r), (T)g.minArr[j], eps, g.nCells[j]));
      }
      Collector col;
      col.init(cbase, K1);
      col.limit(Tq);
      T fb = F;
      const double rF2 = (double)F * (1.0 + 0x1p-15);
      // PR rounds: every load first (consumed at once, so none is deferred into the filter's
      // branch), then the rounds in order through the filter and the collector
      auto group = [&](const uint32_t (&pv)[PR], const bool (&val)[PR], uint32_t f0, uint32_t tot) {
        T cv[PR][D];
        uint32_t cid[PR];
#pragma unroll
        for (int u = 0; u < PR; u++) {
#pragma unroll
          for (int j = 0; j < D; j++) cv[u][j] = xs[(size_t)j * n + pv[u]];
          cid[u] = a.sortedIds[pv[u]];
        }
#pragma unroll
        for (int u = 0; u < PR; u++) {
#pragma unroll
          for (int j = 0; j < D; j++) asm volatile("" ::"v"(cv[u][j]));
          asm volatile("" ::"v"(cid[u]));
        }
#pragma unroll
        for (int u = 0; u < PR; u++) {
          if (f0 + 64u * u >= tot) break;  // wave-uniform
          cand += (unsigned long long)__popcll(__ballot(val[u]));
          const bool near = val[u] && Num<T>::template sq<D>(cv[u], qc) <= fb;
          if (__ballot(near) == 0) continue;
          const double key = near ? fp64_key<D, T>(cv[u], qd) : 0.0;
          col.offer(near, key, cid[u], lane);
          fb = col.nb == K1 ? Num<T>::bound(col.tk) : F;
        }
      };
      // with slab order (projected grids, fixed-dims builds): the block's cells, each cut to
      // the query's slab range in the first non-indexed dim
      const bool cut = (ND > 0 && ND < D) && g.sb && hi[0] - lo[0] < SLAB_MAX_SPAN;
      const uint64_t R = block_rows<ND>(g, lo, hi);
      const uint64_t RS = cut ? R * (uint64_t)(hi[0] - lo[0] + 1) : R;
      uint32_t sql = 0, sqh = 0;
      if (cut) {
        const T xq = pick_t<T, D>(qc, nd);
        sql = readfirst<uint32_t>(slab_q<T>(next_down(xq - r), g));
        sqh = readfirst<uint32_t>(slab_q<T>(next_up(xq + r), g));
      }
      for (uint64_t rb = 0; rb < RS; rb += 64) {
        uint32_t pB, len;
        if (cut)
          seg_batch<T, D, ND, G>(a, g, lo, hi, nd >= 3, qd, rF2, RS, rb + lane, sql, sqh, xs + (size_t)nd * n, pB, len);
        else
          row_batch<T, D, ND, G>(a, g, lo, hi, nd >= 3, qd, rF2, R, rb + lane, pB, len);
        stream_segments<PR>(pB, len, lane, rowB, rowX, a.segCapLog2, group);
      }
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
    const uint32_t t = next_tile(a.tileCounter, NT, a.tileRanges, lane);
    if (t >= NT) break;
    const uint4 tl = a.tiles[t];
    const TileCtx<G> C = tile_ctx<G>(a, g0, tl);
    const G &g = C.g;
    const uint32_t K1 = C.k1;
    Lane<T, D> L;
    uint32_t lo[KNN_MAX_INDEXED], hi[KNN_MAX_INDEXED];
    tile_setup<T, D, 0, G>(a, g, tl, lane, L, lo, hi);
    uint32_t size = 0;
    double topKey = L.Tq;
    uint32_t topId = 0xffffffffu;
    T filt = L.F;
    uint64_t tileCand = 0;

    const double rF2 = (double)L.F * (1.0 + 0x1p-15);
    auto narrowH = [&](const uint32_t *cur, uint32_t &l0, uint32_t &h0) {
      return g.nd < 3 ? true : row_chord<T, D, 0, G>(g, cur, L.qd, rF2, L.active, l0, h0);
    };
    for_each_row<0>(a, g, lo, hi, narrowH, [&](uint32_t pBeg, uint32_t pEnd) {
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
    tile_accounting<T>(tally, tl, ok, L.active && !ok, tileCand, K1);
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
  const size_t lds = range_collect_lds_bytes(a.k1);
  if constexpr (D <= KNN_MAX_INDEXED) {
    if (g.nd == (uint32_t)D) {
      k_range_collect<T, D, D, G><<<grid, 64, lds, s>>>(a, g);
      return;
    }
  }
  // 6 indexed dims of more data dims: the reference's default NUMINDEXEDDIM above 6 dims
  if constexpr (D > 6) {
    if (g.nd == 6u) {
      k_range_collect<T, D, 6, G><<<grid, 64, lds, s>>>(a, g);
      return;
    }
  }
  k_range_collect<T, D, 0, G><<<grid, 64, lds, s>>>(a, g);
}

}  // namespace

uint32_t range_select_grid(uint32_t ntiles) {
  return ntiles < SELECT_MAX_GRID ? (ntiles ? ntiles : 1u) : SELECT_MAX_GRID;
}

template <typename T>
size_t range_select_lds_bytes(uint32_t smax, uint32_t nb) {
  return (size_t)sel_fixed_words(nb) * sizeof(uint32_t) + (size_t)smax * 8 +
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

// the learned schedule's bucket count when a.selBuckets names it, else the other one
template <typename T>
void range_select(const JoinArgs &a, const typename GridOf<T>::type &g, hipStream_t s) {
  const bool learned = a.selBuckets == (uint32_t)SEL_NB_LEARNED;
#define SEL_CASE(DD)                                                        \
  case DD:                                                                  \
    if (learned)                                                            \
      select_launch<T, DD, SEL_NB_LEARNED>(a, g, g.nd, s);                  \
    else                                                                    \
      select_launch<T, DD, SEL_NB_OTHER>(a, g, g.nd, s);                    \
    break;
  switch (a.d) {
    SEL_CASE(1) SEL_CASE(2) SEL_CASE(3) SEL_CASE(4) SEL_CASE(5) SEL_CASE(6) SEL_CASE(7) SEL_CASE(8)
    SEL_CASE(9) SEL_CASE(10) SEL_CASE(11) SEL_CASE(12) SEL_CASE(13) SEL_CASE(14) SEL_CASE(15) SEL_CASE(16)
    default: break;
  }
#undef SEL_CASE
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
  template size_t range_select_lds_bytes<T>(uint32_t, uint32_t);                                   \
  template void range_topk<T>(const JoinArgs &, const GridOf<T>::type &, hipStream_t);   \
  template void range_collect<T>(const JoinArgs &, const GridOf<T>::type &, hipStream_t); \
  template size_t range_heap_scratch_bytes<T>(uint32_t, uint32_t);
INST(float)
INST(double)
#undef INST

}  // namespace knn
