// join.hip -- grid index and the epsilon-range top-(K+1) kernel for gfx950.
//
// Index (replaces GPU.cu:3715-4067): per-point cell key (kernel.cu:1673-1697) -> stable LSD
// radix sort of (key, point id) (prims.hip) -> boundary flags + scan -> compacted non-empty
// cell table {cellKeys[c], cellStart[c]} in HBM, plus a cell-sorted dim-major copy of the
// coordinates (xs) so that every candidate range is one contiguous, coalesced stream.
//
// Range kernel (replaces kernel.cu:61-225 + the pair buffer, the two stable_sort_by_key of
// GPU.cu:1629-1634 and the host top-k of GPU.cu:3610-3704): one wavefront per tile of <= 64
// queries that share a cell; lane = query.  The tile scans the union of its queries' cell
// blocks; for 2-D+ grids the cells x-1..x+1 of one row are adjacent linear ids, hence one
// contiguous range of sorted points.  Candidates are loaded 64 at a time (one per lane,
// coalesced) and broadcast lane by lane with v_readlane; each lane keeps its query's
// top-(K+1) as a max-heap on (FP64 key, id), in LDS when it fits, else in a per-block
// global scratch.  A cheap FP32 distance filters candidates first: a candidate whose FP64
// key can beat the heap top has an FP32 sum within (1+2^-16) of it.
//
// Success rule (per query): at least K+1 candidates with key <= Teps = eps^2 (1-2^-20).
// The scanned block holds every point with key <= Teps, so success gives the exact
// top-(K+1) by (key, id); the reference's FP32 test (kernel.cu:130-147) differs from it only
// for candidates within 2^-20 of eps.
#include "knn_launch.h"

namespace knn {

namespace {

// ------------------------------------------------------------------ index kernels
template <typename T, typename K, typename G>
__global__ void k_cell_keys(const T *__restrict__ x, uint32_t n, G g, K *keys, uint32_t *ids) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t id = 0;
  for (uint32_t j = 0; j < g.nd; j++) {
    uint32_t c = cell_coord<T>(x[(size_t)j * n + i], (T)g.minArr[j], (T)g.eps, g.nCells[j]);
    id += (uint64_t)c * g.mult[j];
  }
  keys[i] = (K)id;
  ids[i] = i;
}

template <typename K>
__global__ void k_boundary_flags(const K *keys, uint32_t n, uint32_t *flags) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  flags[p] = (p == 0 || keys[p] != keys[p - 1]) ? 1u : 0u;
}

template <typename K>
__global__ void k_compact_cells(const K *keys, const uint32_t *flags, const uint32_t *pos,
                                uint32_t n, uint64_t *cellKeys, uint32_t *cellStart) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (flags[p]) {
    cellKeys[pos[p]] = (uint64_t)keys[p];
    cellStart[pos[p]] = p;
  }
  if (p == n - 1) cellStart[pos[n]] = n;
}

template <typename T>
__global__ void k_gather_soa(const T *__restrict__ x, const uint32_t *__restrict__ ids, uint32_t n,
                             uint32_t d, T *__restrict__ xs) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t i = ids[p];
  for (uint32_t j = 0; j < d; j++) xs[(size_t)j * n + p] = x[(size_t)j * n + i];
}

__global__ void k_active_flags(const uint32_t *sortedIds, const uint8_t *failed, uint32_t n,
                               uint32_t *flags) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  flags[p] = failed ? (uint32_t)failed[sortedIds[p]] : 1u;
}

__global__ void k_compact_active(const uint32_t *sortedIds, const uint32_t *flags,
                                 const uint32_t *pos, uint32_t n, uint32_t *act) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (flags[p]) act[pos[p]] = sortedIds[p];
}

__global__ void k_cell_tile_counts(const uint32_t *cellStart, const uint32_t *pos, uint32_t ncell,
                                   uint32_t *ntiles) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const uint32_t a0 = pos[cellStart[c]], a1 = pos[cellStart[c + 1]];
  ntiles[c] = (a1 - a0 + 63) / 64;
}

__global__ void k_write_tiles(const uint32_t *cellStart, const uint32_t *pos,
                              const uint32_t *tileOff, uint32_t ncell, uint4 *tiles) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const uint32_t a0 = pos[cellStart[c]], a1 = pos[cellStart[c + 1]];
  uint32_t t = tileOff[c];
  for (uint32_t a = a0; a < a1; a += 64, t++)
    tiles[t] = make_uint4(c, a, (a1 - a < 64u) ? a1 - a : 64u, a == a0 ? 1u : 0u);
}

// ------------------------------------------------------------------ range kernel
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

template <typename T> struct Acc;
template <> struct Acc<float> {
  // filter value: FP32 fixed-order sum of squares
  template <int D>
  __device__ static __forceinline__ float filt(const float *c, const float *q) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < D; j++) {
      float df = c[j] - q[j];
      float sq = df * df;
      s = s + sq;
    }
    return s;
  }
  __device__ static __forceinline__ float filt_bound(double key) {
    return (float)(key * (1.0 + 0x1p-16));
  }
};
template <> struct Acc<double> {
  template <int D>
  __device__ static __forceinline__ double filt(const double *c, const double *q) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < D; j++) {
      double df = c[j] - q[j];
      double sq = df * df;
      s = s + sq;
    }
    return s;
  }
  __device__ static __forceinline__ double filt_bound(double key) { return key; }
};

template <int D, typename T>
__device__ __forceinline__ double fp64_key(const T *c, const double *qd) {
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < D; j++) {
    double dx = (double)c[j] - qd[j];
    double sq = dx * dx;
    s = s + sq;
  }
  return s;
}

template <typename T, int D, bool LDS, typename G>
__global__ __launch_bounds__(64) void k_range_topk(JoinArgs a, G g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = threadIdx.x;
  const uint32_t K1 = a.k1;
  Heap h;
  h.lane = lane;
  if (LDS) {
    h.key = reinterpret_cast<double *>(smem);
    h.id = reinterpret_cast<uint32_t *>(smem + (size_t)K1 * 64 * sizeof(double));
  } else {
    unsigned char *base =
        reinterpret_cast<unsigned char *>(a.heapScratch) + (size_t)blockIdx.x * K1 * 64 * 12;
    h.key = reinterpret_cast<double *>(base);
    h.id = reinterpret_cast<uint32_t *>(base + (size_t)K1 * 64 * sizeof(double));
  }
  const T *__restrict__ x = reinterpret_cast<const T *>(a.x);
  const T *__restrict__ xs = reinterpret_cast<const T *>(a.xs);
  const size_t n = a.n;
  const T eps = (T)g.eps;
  const double E = (double)eps * (double)eps;
  const double Teps = E * (1.0 - 0x1p-20);
  const T filtEps = Acc<T>::filt_bound(Teps);
  unsigned long long myCand = 0;

  for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const uint4 tl = a.tiles[t];
    const uint32_t cnt = tl.z;
    const bool active = lane < cnt;
    const uint32_t q = a.act[tl.y + (active ? lane : 0u)];
    T qc[D];
    double qd[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
      qc[j] = x[(size_t)j * n + q];
      qd[j] = (double)qc[j];
    }
    // per indexed dim: [min(c-1, coord(x-eps, down)), max(c+1, coord(x+eps, up))], union
    uint32_t lo[KNN_MAX_INDEXED], hi[KNN_MAX_INDEXED];
    for (uint32_t j = 0; j < g.nd; j++) {
      const T mn = (T)g.minArr[j];
      const uint32_t nc = g.nCells[j];
      T xj = qc[0];
#pragma unroll
      for (int jj = 0; jj < D; jj++)
        if ((uint32_t)jj == j) xj = qc[jj];
      const uint32_t c = cell_coord<T>(xj, mn, eps, nc);
      const uint32_t ca = cell_coord<T>(next_down(xj - eps), mn, eps, nc);
      const uint32_t cb = cell_coord<T>(next_up(xj + eps), mn, eps, nc);
      const uint32_t cl = c > 0 ? c - 1 : 0u;
      const uint32_t ch = c + 1 < nc ? c + 1 : nc - 1;
      uint32_t l = ca < cl ? ca : cl;
      uint32_t u = cb > ch ? cb : ch;
      if (!active) {
        l = 0xffffffffu;
        u = 0u;
      }
      lo[j] = wave_min_u32(l);
      hi[j] = wave_max_u32(u);
    }

    uint32_t size = 0;
    double topKey = Teps;
    uint32_t topId = 0xffffffffu;
    T filt = filtEps;
    uint64_t tileCand = 0, tileCells = 0;

    uint32_t cur[KNN_MAX_INDEXED];
    for (uint32_t j = 0; j < g.nd; j++) cur[j] = lo[j];
    for (;;) {
      uint64_t rowBase = 0;
      for (uint32_t j = 1; j < g.nd; j++) rowBase += (uint64_t)cur[j] * g.mult[j];
      const uint64_t idLo = rowBase + lo[0], idHi = rowBase + hi[0];
      const uint32_t cbeg = lower_bound_u64(a.cellKeys, a.ncell, idLo);
      const uint32_t cend = upper_bound_u64(a.cellKeys, a.ncell, idHi);
      tileCells += cend - cbeg;
      const uint32_t pBeg = a.cellStart[cbeg], pEnd = a.cellStart[cend];
      tileCand += pEnd - pBeg;
      for (uint32_t base = pBeg; base < pEnd; base += 64) {
        const uint32_t p = base + lane;
        const bool pv = p < pEnd;
        const uint32_t pp = pv ? p : pBeg;
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
          if (!active) continue;
          const T s = Acc<T>::template filt<D>(c, qc);
          if (s > filt) continue;
          const double key = fp64_key<D, T>(c, qd);
          if (size < K1) {
            if (key <= Teps) {
              h.push(size, key, cid);
              if (size == K1) {
                topKey = h.K(0);
                topId = h.I(0);
                filt = Acc<T>::filt_bound(topKey);
              }
            }
          } else if (kless(key, cid, topKey, topId)) {
            h.replace_top(K1, key, cid);
            topKey = h.K(0);
            topId = h.I(0);
            filt = Acc<T>::filt_bound(topKey);
          }
        }
      }
      // next row (odometer over dims 1..nd-1)
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
    if (active) myCand += tileCand;

    // write rows / failures
    bool ok = active && size == K1;
    if (ok) {
      h.sort(K1);
      T *od = reinterpret_cast<T *>(a.outDists);
      double rs = 0.0;
      for (uint32_t e = 0; e < K1; e++) {
        const uint32_t id = h.I(e);
        T c[D];
#pragma unroll
        for (int j = 0; j < D; j++) c[j] = x[(size_t)j * n + id];
        const T s = Acc<T>::template filt<D>(c, qc);
        const T dist = sqrt(s);
        a.outIds[(size_t)q * K1 + e] = (int32_t)id;
        od[(size_t)q * K1 + e] = dist;
        const T sq = dist * dist;
        rs += (double)sq;
      }
      a.rowSum[q] = rs;
      a.failed[q] = 0;
    } else if (active) {
      a.failed[q] = 1;
    }
    const uint64_t fm = __ballot(active && !ok);
    const uint64_t om = __ballot(ok);
    if (lane == 0) {
      if (fm) atomicAdd(a.failCount, (uint32_t)__popcll(fm));
      // compulsory HBM bytes: the candidates of a query cell once (first tile of the cell),
      // cell lookups, query ids in / failure flags out, and the finished rows
      unsigned long long b = (unsigned long long)cnt * 8ull +
                             (unsigned long long)__popcll(om) * K1 * (4ull + sizeof(T));
      if (tl.w) b += tileCand * (sizeof(T) * D + 4ull) + tileCells * 8ull;
      atomicAdd(a.candCount + 1, b);
    }
  }
  // aggregate candidate counts
  for (int o = 32; o > 0; o >>= 1) myCand += __shfl_xor(myCand, o, 64);
  if (lane == 0 && myCand) atomicAdd(a.candCount, myCand);
}

constexpr uint32_t MAX_LDS_BYTES = 65536;

template <typename T, int D, typename G>
void launch_range(const JoinArgs &a, const G &g, hipStream_t s) {
  if (a.ntiles == 0) return;
  const size_t heapBytes = (size_t)a.k1 * 64 * 12;
  uint32_t grid = a.ntiles < 8192u ? a.ntiles : 8192u;
  if (heapBytes <= MAX_LDS_BYTES) {
    k_range_topk<T, D, true, G><<<grid, 64, heapBytes, s>>>(a, g);
  } else {
    k_range_topk<T, D, false, G><<<grid, 64, 0, s>>>(a, g);
  }
}

// ------------------------------------------------------------------ checksum
constexpr uint32_t SUM_CHUNK = 4096;
__global__ void k_sum_chunks(const double *rowSum, uint32_t n, double *partial) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t lo = (uint64_t)c * SUM_CHUNK;
  if (lo >= n) return;
  const uint64_t hi = lo + SUM_CHUNK < n ? lo + SUM_CHUNK : n;
  double s = 0.0;
  for (uint64_t r = lo; r < hi; r++) s += rowSum[r];
  partial[c] = s;
}
__global__ void k_sum_final(const double *partial, uint32_t m, double *out) {
  if (threadIdx.x || blockIdx.x) return;
  double s = 0.0;
  for (uint32_t c = 0; c < m; c++) s += partial[c];
  *out = s;
}

}  // namespace

// ------------------------------------------------------------------ host launchers
template <typename T>
void cell_keys_u32(const T *x, uint32_t n, const typename GridOf<T>::type &g, uint32_t *keys,
                   uint32_t *ids, hipStream_t s) {
  if (n) k_cell_keys<T, uint32_t><<<(n + 255) / 256, 256, 0, s>>>(x, n, g, keys, ids);
}
template <typename T>
void cell_keys_u64(const T *x, uint32_t n, const typename GridOf<T>::type &g, uint64_t *keys,
                   uint32_t *ids, hipStream_t s) {
  if (n) k_cell_keys<T, uint64_t><<<(n + 255) / 256, 256, 0, s>>>(x, n, g, keys, ids);
}
void boundary_flags_u32(const uint32_t *keys, uint32_t n, uint32_t *flags, hipStream_t s) {
  if (n) k_boundary_flags<uint32_t><<<(n + 255) / 256, 256, 0, s>>>(keys, n, flags);
}
void boundary_flags_u64(const uint64_t *keys, uint32_t n, uint32_t *flags, hipStream_t s) {
  if (n) k_boundary_flags<uint64_t><<<(n + 255) / 256, 256, 0, s>>>(keys, n, flags);
}
void compact_cells_u32(const uint32_t *keys, const uint32_t *flags, const uint32_t *pos,
                       uint32_t n, uint64_t *cellKeys, uint32_t *cellStart, hipStream_t s) {
  if (n) k_compact_cells<uint32_t><<<(n + 255) / 256, 256, 0, s>>>(keys, flags, pos, n, cellKeys, cellStart);
}
void compact_cells_u64(const uint64_t *keys, const uint32_t *flags, const uint32_t *pos,
                       uint32_t n, uint64_t *cellKeys, uint32_t *cellStart, hipStream_t s) {
  if (n) k_compact_cells<uint64_t><<<(n + 255) / 256, 256, 0, s>>>(keys, flags, pos, n, cellKeys, cellStart);
}
template <typename T>
void gather_soa(const T *x, const uint32_t *ids, uint32_t n, uint32_t d, T *xs, hipStream_t s) {
  if (n) k_gather_soa<T><<<(n + 255) / 256, 256, 0, s>>>(x, ids, n, d, xs);
}
void active_flags(const uint32_t *sortedIds, const uint8_t *failed, uint32_t n, uint32_t *flags,
                  hipStream_t s) {
  if (n) k_active_flags<<<(n + 255) / 256, 256, 0, s>>>(sortedIds, failed, n, flags);
}
void compact_active(const uint32_t *sortedIds, const uint32_t *flags, const uint32_t *pos,
                    uint32_t n, uint32_t *act, hipStream_t s) {
  if (n) k_compact_active<<<(n + 255) / 256, 256, 0, s>>>(sortedIds, flags, pos, n, act);
}
void cell_tile_counts(const uint32_t *cellStart, const uint32_t *pos, uint32_t ncell,
                      uint32_t *ntiles, hipStream_t s) {
  if (ncell) k_cell_tile_counts<<<(ncell + 255) / 256, 256, 0, s>>>(cellStart, pos, ncell, ntiles);
}
void write_tiles(const uint32_t *cellStart, const uint32_t *pos, const uint32_t *tileOff,
                 uint32_t ncell, uint4 *tiles, hipStream_t s) {
  if (ncell) k_write_tiles<<<(ncell + 255) / 256, 256, 0, s>>>(cellStart, pos, tileOff, ncell, tiles);
}

template <typename T>
void range_topk(const JoinArgs &a, const typename GridOf<T>::type &g, hipStream_t s) {
  using G = typename GridOf<T>::type;
  switch (a.d) {
    case 1: launch_range<T, 1, G>(a, g, s); break;
    case 2: launch_range<T, 2, G>(a, g, s); break;
    case 3: launch_range<T, 3, G>(a, g, s); break;
    case 4: launch_range<T, 4, G>(a, g, s); break;
    case 5: launch_range<T, 5, G>(a, g, s); break;
    case 6: launch_range<T, 6, G>(a, g, s); break;
    case 7: launch_range<T, 7, G>(a, g, s); break;
    case 8: launch_range<T, 8, G>(a, g, s); break;
    case 9: launch_range<T, 9, G>(a, g, s); break;
    case 10: launch_range<T, 10, G>(a, g, s); break;
    case 11: launch_range<T, 11, G>(a, g, s); break;
    case 12: launch_range<T, 12, G>(a, g, s); break;
    case 13: launch_range<T, 13, G>(a, g, s); break;
    case 14: launch_range<T, 14, G>(a, g, s); break;
    case 15: launch_range<T, 15, G>(a, g, s); break;
    case 16: launch_range<T, 16, G>(a, g, s); break;
    default: break;
  }
}

template <typename T>
size_t range_heap_scratch_bytes(uint32_t ntiles, uint32_t k1) {
  const size_t heapBytes = (size_t)k1 * 64 * 12;
  if (heapBytes <= MAX_LDS_BYTES) return 0;
  const uint32_t grid = ntiles < 8192u ? ntiles : 8192u;
  return heapBytes * grid;
}

void sum_rows(const double *rowSum, uint32_t n, double *partial, double *out, hipStream_t s) {
  const uint32_t m = (n + SUM_CHUNK - 1) / SUM_CHUNK;
  if (m) k_sum_chunks<<<(m + 63) / 64, 64, 0, s>>>(rowSum, n, partial);
  k_sum_final<<<1, 64, 0, s>>>(partial, m, out);
}

#define INST(T)                                                                                \
  template void cell_keys_u32<T>(const T *, uint32_t, const GridOf<T>::type &, uint32_t *,      \
                                 uint32_t *, hipStream_t);                                     \
  template void cell_keys_u64<T>(const T *, uint32_t, const GridOf<T>::type &, uint64_t *,      \
                                 uint32_t *, hipStream_t);                                     \
  template void gather_soa<T>(const T *, const uint32_t *, uint32_t, uint32_t, T *, hipStream_t); \
  template void range_topk<T>(const JoinArgs &, const GridOf<T>::type &, hipStream_t);          \
  template size_t range_heap_scratch_bytes<T>(uint32_t, uint32_t);
INST(float)
INST(double)
#undef INST

}  // namespace knn
