// join.hip -- grid index, active-query tiles and the checksum reduction for gfx950.
//
// Index (replaces GPU.cu:3715-4067): per-point cell key (kernel.cu:1673-1697) -> stable LSD
// radix sort of (key, point id) (prims.hip) -> boundary flags + scan -> compacted non-empty
// cell table {cellKeys[c], cellStart[c]} in HBM, plus a cell-sorted dim-major copy of the
// coordinates (xs) so that every candidate range is one contiguous, coalesced stream.
//
#include "knn_range.h"

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
  // slab order: the cell id above the slab code of the first non-indexed dim (one sort gives
  // cell order, and slab order within each cell)
  if (g.sb) id = (id << g.sb) | slab_q<T>(x[(size_t)g.nd * n + i], g);
  keys[i] = (K)id;
  ids[i] = i;
}

// sb: the slab code's bits below the cell id in the keys (0 without slab order)
template <typename K>
__global__ void k_boundary_flags(const K *keys, uint32_t n, uint32_t sb, uint32_t *flags) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  flags[p] = (p == 0 || (keys[p] >> sb) != (keys[p - 1] >> sb)) ? 1u : 0u;
}

template <typename K>
__global__ void k_compact_cells(const K *keys, const uint32_t *flags, const uint32_t *pos,
                                uint32_t n, uint32_t sb, uint64_t *cellKeys, uint32_t *cellStart) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (flags[p]) {
    cellKeys[pos[p]] = (uint64_t)(keys[p] >> sb);
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

__global__ void k_gather_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ idx, uint32_t n,
                             uint32_t *__restrict__ dst) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[p] = src[idx[p]];
}

// xs[j*n + p] = in[ids[p]*d + perm[j]]: the same values as k_gather_soa from the permuted
// SoA copy, read from the row-major input, whose d coordinates of a point share a cache line
template <typename T>
__global__ void k_gather_rows(const T *__restrict__ in, const uint32_t *__restrict__ ids, uint32_t n,
                              uint32_t d, DimPerm perm, T *__restrict__ xs) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const T *row = in + (size_t)ids[p] * d;
  for (uint32_t j = 0; j < d; j++) xs[(size_t)j * n + p] = row[perm.v[j]];
}

__global__ void k_dense_counts(const uint64_t *cellKeys, const uint32_t *cellStart,
                               uint32_t ncell, uint32_t *dense) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  dense[cellKeys[c]] = cellStart[c + 1] - cellStart[c];
}

__device__ __forceinline__ uint64_t block_key(uint64_t id, uint32_t ncell0, uint32_t span) {
  const uint64_t nb0 = (ncell0 + span - 1) / span;
  return (id / ncell0) * nb0 + (id % ncell0) / span;
}

// bits[w] bit l = (src[64 w + l] != 0), for the ids below n (the last word's bits past n are 0)
__global__ void k_pack_bits(const uint8_t *__restrict__ src, uint32_t n, uint64_t *__restrict__ bits) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t b = __ballot(i < n && src[i] != 0);
  if ((threadIdx.x & 63u) == 0 && i < n) bits[i >> 6] = b;
}

// Per active query (sorted position p): the number of points in its own cell block (radius
// r = its own, or rAll for every query, cells [coord(x-r) rounded down, coord(x+r) rounded up]) from
// the dense cell-start table.  Fewer than K+1 points there means the query cannot succeed
// at this radius: it is marked failed without a distance evaluation.  The per-cell maximum of
// those counts feeds the algorithmic-bytes accounting (each query cell's block once).
// The active flags come as a bitmap by point id (bit q & 63 of word q >> 6): its 1/8 of a byte
// per point stays in the L2 for these reads in sorted order, where a byte per point (10 MB at
// 10M points) made each read a line from beyond it.
template <typename T, typename G, int ND>
__global__ void k_prefilter(const T *__restrict__ xs, uint32_t n, G g, const T *radius, T rAll,
                            const uint32_t *__restrict__ dense, const uint32_t *__restrict__ sortedIds,
                            const uint64_t *activeBits, uint8_t *failed, uint32_t k1,
                            const uint32_t *cellOf, uint32_t *cellMax, uint32_t *flags) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63;
  const bool inRange = p < n;
  const uint32_t q = inRange ? sortedIds[p] : 0u;
  const bool active = inRange && !(activeBits && !((activeBits[q >> 6] >> (q & 63u)) & 1ull));
  uint32_t cnt = 0;
  if (active) {
    const T r = radius ? radius[q] : rAll;
    const T eps = (T)g.eps;
    const uint32_t nd = nd_of<ND>(g);
    // per-dim arrays over a constant bound (registers, not private scratch)
    uint32_t lo[nd_cap<ND>()], hi[nd_cap<ND>()], cur[nd_cap<ND>()];
#pragma unroll
    for (uint32_t j = 0; j < nd_cap<ND>(); j++) {
      if (j >= nd) break;
      const T xj = xs[(size_t)j * n + p];  // cell-sorted copy: the query's own coordinates, coalesced
      lo[j] = cell_coord<T>(next_down(xj - r), (T)g.minArr[j], eps, g.nCells[j]);
      hi[j] = cell_coord<T>(next_up(xj + r), (T)g.minArr[j], eps, g.nCells[j]);
      cur[j] = lo[j];
    }
    do {
      const uint64_t rowBase = row_base<ND>(g, cur);
      cnt += dense[rowBase + hi[0] + 1] - dense[rowBase + lo[0]];
    } while (next_row<ND>(cur, lo, hi, nd));
  }
  if (cellMax) {
    // cellOf is non-decreasing in p, so a cell's lanes are one run of the wave: the run's
    // maximum is gathered into its first lane (each lane takes the maximum over the next
    // 1, 2, 4, ... lanes of its own cell), which issues the run's only atomic
    const uint32_t c = inRange ? cellOf[p] : 0xFFFFFFFFu;
    uint32_t m = active ? cnt : 0u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t mo = __shfl_down(m, o, 64);
      const uint32_t co = __shfl_down(c, o, 64);
      if (lane + o < 64 && co == c && mo > m) m = mo;
    }
    const uint32_t cPrev = __shfl_up(c, 1, 64);
    if (inRange && m && (lane == 0 || cPrev != c)) atomicMax(&cellMax[c], m);
  }
  if (!inRange) return;
  if (!active) {
    flags[p] = 0;
  } else if (cnt < k1) {
    failed[q] = 1;
    flags[p] = 0;
  } else {
    flags[p] = 1;
  }
}

// sum over cells of cellMax (candidate points of each active query cell's block) and the
// number of active query cells; out[0] += points, out[1] += cells.  Grid-stride per thread
// over 16-byte vectors (cellMax comes from an aligned allocation), then a wave reduction and
// one pair of atomics per block.
__global__ __launch_bounds__(256) void k_cell_bytes(const uint32_t *cellMax, uint32_t ncell,
                                                    unsigned long long *out) {
  unsigned long long sp = 0, sc = 0;
  const uint32_t stride = gridDim.x * blockDim.x, gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nv = ncell / 4;
  const uint4 *v4 = reinterpret_cast<const uint4 *>(cellMax);
  for (uint32_t i = gid; i < nv; i += stride) {
    const uint4 v = v4[i];
    sp += (unsigned long long)v.x + (unsigned long long)v.y + (unsigned long long)v.z +
          (unsigned long long)v.w;
    sc += (v.x ? 1ull : 0ull) + (v.y ? 1ull : 0ull) + (v.z ? 1ull : 0ull) + (v.w ? 1ull : 0ull);
  }
  for (uint32_t c = 4 * nv + gid; c < ncell; c += stride) {
    const uint32_t v = cellMax[c];
    sp += v;
    sc += v ? 1ull : 0ull;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sp += __shfl_xor(sp, o, 64);
    sc += __shfl_xor(sc, o, 64);
  }
  // the block's four wave sums, then one pair of atomics per block (every wave adding to the
  // same two counters serialised them)
  __shared__ unsigned long long ws[2][4];
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    ws[0][w] = sp;
    ws[1][w] = sc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long bp = ws[0][0] + ws[0][1] + ws[0][2] + ws[0][3];
    const unsigned long long bc = ws[1][0] + ws[1][1] + ws[1][2] + ws[1][3];
    if (bp || bc) {
      atomicAdd(out, bp);
      atomicAdd(out + 1, bc);
    }
  }
}

// cellOf[p] = compacted cell index of sorted position p (inclusive scan of boundary flags - 1)
__global__ void k_cell_of(const uint32_t *flags, const uint32_t *pos, uint32_t n, uint32_t *cellOf) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  cellOf[p] = pos[p] + flags[p] - 1;
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

// Tile-block shape of a pass from the active count A = *activeCount (read on the device): a
// tile of `rows` rows (1, or 2 when rowsMax = 2) scans the union of its lanes' blocks, about
// (span + 2 reach) x (rows + 2 reach) cells, for min(64, span * rows * rho) queries (rho =
// active queries per cell holding any, *activeCells when given, else per non-empty cell), so
// the cost per query is about that area over the query count; a pass with fewer than minTiles
// tiles leaves the persistent waves idle at its end, which costs in proportion.  shape[0] =
// span, shape[1] = rows.
__global__ void k_choose_span(const uint32_t *activeCount, const unsigned long long *activeCells,
                              uint32_t ncell, double avgReach, uint32_t spanMax, uint32_t rowsMax,
                              double minTiles, uint32_t *shape) {
  if (threadIdx.x || blockIdx.x) return;
  const uint32_t A = *activeCount;
  uint32_t best = 1, bestRows = 1;
  if (A && ncell) {
    const unsigned long long ac = activeCells ? *activeCells : 0ull;
    const double rhoA = (double)A / (ac ? (double)ac : (double)ncell);
    double bc = 1e300;
    for (uint32_t rows = 1; rows <= rowsMax; rows++) {
      for (uint32_t w = 1; w <= spanMax; w *= 2) {
        const double q = fmax(1.0, fmin(64.0, w * rows * rhoA));
        const double tilesEst = (double)A / q;
        const double cost = (w + 2.0 * avgReach) * (rows + 2.0 * avgReach) / q * fmax(1.0, minTiles / tilesEst);
        if (cost < bc * 0.97) {
          bc = cost;
          best = w;
          bestRows = rows;
        }
      }
    }
  }
  shape[0] = best;
  shape[1] = bestRows;
}

__global__ void k_block_flags_dev(const uint64_t *cellKeys, uint32_t ncell, uint32_t ncell0,
                                  const uint32_t *span, uint32_t *flags) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const uint32_t w = *span;
  flags[c] = (c == 0 || block_key(cellKeys[c], ncell0, w) != block_key(cellKeys[c - 1], ncell0, w)) ? 1u : 0u;
}

__global__ void k_compact_blocks_keyed(const uint64_t *cellKeys, const uint32_t *cellStart,
                                       const uint32_t *flags, const uint32_t *pos, uint32_t ncell,
                                       uint32_t ncell0, const uint32_t *shape, uint32_t n,
                                       uint32_t *blockStart, uint64_t *blockKey) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  if (flags[c]) {
    blockStart[pos[c]] = cellStart[c];
    blockKey[pos[c]] = block_key(cellKeys[c], ncell0, shape[0]);
  }
  if (c == ncell - 1) blockStart[pos[ncell]] = n;
}

constexpr uint32_t NO_PARTNER = 0xffffffffu;

// first index in [lo, hi) of the increasing keys with keys[i] >= v
__device__ __forceinline__ uint32_t lower_bound_keys(const uint64_t *keys, uint32_t lo, uint32_t hi, uint64_t v) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// per tile block b < *nblocks (0 for b up to ncell): tiles of <= 64 active queries.  With row
// pairs (shape[1] = 2) a block of an even dim-1 coordinate takes its partner's queries (the
// block one row up, same window; partner[b]), and a block whose partner one row down exists
// has no tiles of its own.  Block keys are distinct and increase with b, so the partner of key
// +- nw (nw = windows per row) lies within nw places of b: one binary search over those.
__global__ void k_pair_tile_counts(const uint32_t *blockStart, const uint64_t *blockKey, const uint32_t *pos,
                                   const uint32_t *nblocks, uint32_t ncell, uint32_t ncell0, uint32_t ncell1,
                                   const uint32_t *shape, uint32_t *partner, uint32_t *ntiles) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ncell) return;
  const uint32_t nb = *nblocks;
  if (b >= nb) {
    ntiles[b] = 0;
    return;
  }
  uint32_t cnt = pos[blockStart[b + 1]] - pos[blockStart[b]];
  uint32_t pj = NO_PARTNER;
  if (shape[1] == 2u && ncell1 > 1u) {
    const uint64_t key = blockKey[b];
    const uint64_t nw = (ncell0 + shape[0] - 1) / shape[0];  // windows per row
    const uint32_t r1 = (uint32_t)((key / nw) % ncell1);       // the block's dim-1 coordinate
    if ((r1 & 1u) == 0u) {
      if (r1 + 1u < ncell1) {
        const uint32_t hiB = (uint64_t)nb - b - 1 > nw ? b + 1 + (uint32_t)nw : nb;
        const uint32_t j = lower_bound_keys(blockKey, b + 1, hiB, key + nw);
        if (j < nb && blockKey[j] == key + nw) {
          pj = j;
          cnt += pos[blockStart[j + 1]] - pos[blockStart[j]];
        }
      }
    } else {
      const uint32_t loB = (uint64_t)b > nw ? b - (uint32_t)nw : 0u;
      const uint32_t j = lower_bound_keys(blockKey, loB, b, key - nw);
      if (j < b && blockKey[j] == key - nw) cnt = 0;
    }
  }
  partner[b] = pj;
  ntiles[b] = (cnt + 63) / 64;
}

// tiles of block b (and its partner): runs of 64 over the block's active queries followed by
// the partner's, each tile as {second run start, first run start, count, first run length}
__global__ void k_write_pair_tiles(const uint32_t *blockStart, const uint32_t *pos, const uint32_t *tileOff,
                                   const uint32_t *nblocks, const uint32_t *partner, uint32_t ncell,
                                   uint4 *tiles) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ncell || b >= *nblocks) return;
  uint32_t t = tileOff[b];
  if (tileOff[b + 1] == t) return;
  const uint32_t a0 = pos[blockStart[b]], n0 = pos[blockStart[b + 1]] - a0;
  const uint32_t j = partner[b];
  uint32_t a1 = 0, n1 = 0;
  if (j != NO_PARTNER) {
    a1 = pos[blockStart[j]];
    n1 = pos[blockStart[j + 1]] - a1;
  }
  const uint32_t tot = n0 + n1;
  for (uint32_t base = 0; base < tot; base += 64, t++) {
    const uint32_t m = tot - base < 64u ? tot - base : 64u;
    const uint32_t m0 = base < n0 ? (n0 - base < m ? n0 - base : m) : 0u;
    const uint32_t s0 = a0 + (base < n0 ? base : n0);
    const uint32_t s1 = a1 + (base > n0 ? base - n0 : 0u);
    tiles[t] = make_uint4(s1, s0, m, m0);
  }
}

// ------------------------------------------------------------------ checksum
constexpr uint32_t SUM_CHUNK = 4096;
// one block per chunk of SUM_CHUNK rows: coalesced copy to LDS, then lane 0 adds in row order
__global__ __launch_bounds__(256) void k_sum_chunks(const double *rowSum, uint32_t n, double *partial) {
  __shared__ double buf[SUM_CHUNK];
  const uint64_t lo = (uint64_t)blockIdx.x * SUM_CHUNK;
  if (lo >= n) return;
  const uint32_t cnt = (uint32_t)(lo + SUM_CHUNK < n ? SUM_CHUNK : n - lo);
  for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) buf[i] = rowSum[lo + i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    uint32_t r = 0;
    for (; r + 8 <= cnt; r += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = buf[r + u];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; r < cnt; r++) s += buf[r];
    partial[blockIdx.x] = s;
  }
}
// Per-row sums in entry order.  A block stages ROWS consecutive rows (one contiguous range of
// the table) in LDS with coalesced loads, then thread r sums row r from LDS.
// 16 KB of rows per block: 8 blocks (32 waves) per CU overlap their loads and sums (32 KB:
// 0.57 ms per 10M x 65 table, 16 KB: 0.45 ms)
constexpr uint32_t SUM_LDS_BYTES = 16384;
template <typename T>
__global__ __launch_bounds__(256) void k_row_sums(const T *__restrict__ dists, uint32_t n, uint32_t k1,
                                                  uint32_t rowsPerBlock, double *rowSum) {
  __shared__ T buf[SUM_LDS_BYTES / sizeof(T)];
  const uint64_t r0 = (uint64_t)blockIdx.x * rowsPerBlock;
  if (r0 >= n) return;
  const uint32_t rows = (uint32_t)((n - r0) < rowsPerBlock ? (n - r0) : rowsPerBlock);
  const T *src = dists + r0 * k1;
  const uint32_t cnt = rows * k1;
  // eight independent loads in flight per thread before their LDS stores
  constexpr uint32_t U = 8;
  for (uint32_t i0 = threadIdx.x; i0 < cnt; i0 += U * 256u) {
    T v[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t i = i0 + u * 256u;
      v[u] = i < cnt ? src[i] : (T)0;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t i = i0 + u * 256u;
      if (i < cnt) buf[i] = v[u];
    }
  }
  __syncthreads();
  for (uint32_t r = threadIdx.x; r < rows; r += blockDim.x) {
    const T *row = buf + (size_t)r * k1;
    double s = 0.0;
    for (uint32_t e = 0; e < k1; e++) {
      const T d = row[e];
      const T sq = d * d;
      s += (double)sq;
    }
    rowSum[r0 + r] = s;
  }
}

// fallback for rows longer than the LDS stage
template <typename T>
__global__ void k_row_sums_direct(const T *__restrict__ dists, uint32_t n, uint32_t k1, double *rowSum) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const T *row = dists + (size_t)r * k1;
  double s = 0.0;
  for (uint32_t e = 0; e < k1; e++) {
    const T d = row[e];
    const T sq = d * d;
    s += (double)sq;
  }
  rowSum[r] = s;
}

// the chunk sums in order by one thread; sixteen loads in flight ahead of the adds
__global__ void k_sum_final(const double *partial, uint32_t m, double *out) {
  if (threadIdx.x || blockIdx.x) return;
  double s = 0.0;
  uint32_t c = 0;
  for (; c + 16 <= m; c += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = partial[c + u];
#pragma unroll
    for (int u = 0; u < 16; u++) s += v[u];
  }
  for (; c < m; c++) s += partial[c];
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
void boundary_flags_u32(const uint32_t *keys, uint32_t n, uint32_t sb, uint32_t *flags, hipStream_t s) {
  if (n) k_boundary_flags<uint32_t><<<(n + 255) / 256, 256, 0, s>>>(keys, n, sb, flags);
}
void boundary_flags_u64(const uint64_t *keys, uint32_t n, uint32_t *flags, hipStream_t s) {
  if (n) k_boundary_flags<uint64_t><<<(n + 255) / 256, 256, 0, s>>>(keys, n, 0u, flags);
}
void compact_cells_u32(const uint32_t *keys, const uint32_t *flags, const uint32_t *pos,
                       uint32_t n, uint32_t sb, uint64_t *cellKeys, uint32_t *cellStart, hipStream_t s) {
  if (n) k_compact_cells<uint32_t><<<(n + 255) / 256, 256, 0, s>>>(keys, flags, pos, n, sb, cellKeys, cellStart);
}
void compact_cells_u64(const uint64_t *keys, const uint32_t *flags, const uint32_t *pos,
                       uint32_t n, uint64_t *cellKeys, uint32_t *cellStart, hipStream_t s) {
  if (n) k_compact_cells<uint64_t><<<(n + 255) / 256, 256, 0, s>>>(keys, flags, pos, n, 0u, cellKeys, cellStart);
}
template <typename T>
void gather_soa(const T *x, const uint32_t *ids, uint32_t n, uint32_t d, T *xs, hipStream_t s) {
  if (n) k_gather_soa<T><<<(n + 255) / 256, 256, 0, s>>>(x, ids, n, d, xs);
}
template <typename T>
void gather_rows(const T *in, const uint32_t *ids, uint32_t n, uint32_t d, const DimPerm &perm, T *xs,
                 hipStream_t s) {
  if (n) k_gather_rows<T><<<(n + 255) / 256, 256, 0, s>>>(in, ids, n, d, perm, xs);
}
void gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t n, uint32_t *dst, hipStream_t s) {
  if (n) k_gather_u32<<<(n + 255) / 256, 256, 0, s>>>(src, idx, n, dst);
}
void dense_cell_counts(const uint64_t *cellKeys, const uint32_t *cellStart, uint32_t ncell,
                       uint32_t *dense, hipStream_t s) {
  if (ncell) k_dense_counts<<<(ncell + 255) / 256, 256, 0, s>>>(cellKeys, cellStart, ncell, dense);
}
void pack_bits(const uint8_t *src, uint32_t n, uint64_t *bits, hipStream_t s) {
  if (n) k_pack_bits<<<(n + 255) / 256, 256, 0, s>>>(src, n, bits);
}
template <typename T>
void prefilter(const T *xs, uint32_t n, const typename GridOf<T>::type &g, const T *radius,
               T rAll, const uint32_t *dense, const uint32_t *sortedIds, const uint64_t *activeBits,
               uint8_t *failed, uint32_t k1, const uint32_t *cellOf, uint32_t *cellMax,
               uint32_t *flags, hipStream_t s) {
  if (!n) return;
  using G = typename GridOf<T>::type;
  const dim3 grid((n + 255) / 256), block(256);
  switch (g.nd) {
    case 1: k_prefilter<T, G, 1><<<grid, block, 0, s>>>(xs, n, g, radius, rAll, dense, sortedIds, activeBits, failed, k1, cellOf, cellMax, flags); break;
    case 2: k_prefilter<T, G, 2><<<grid, block, 0, s>>>(xs, n, g, radius, rAll, dense, sortedIds, activeBits, failed, k1, cellOf, cellMax, flags); break;
    case 3: k_prefilter<T, G, 3><<<grid, block, 0, s>>>(xs, n, g, radius, rAll, dense, sortedIds, activeBits, failed, k1, cellOf, cellMax, flags); break;
    case 4: k_prefilter<T, G, 4><<<grid, block, 0, s>>>(xs, n, g, radius, rAll, dense, sortedIds, activeBits, failed, k1, cellOf, cellMax, flags); break;
    default: k_prefilter<T, G, 0><<<grid, block, 0, s>>>(xs, n, g, radius, rAll, dense, sortedIds, activeBits, failed, k1, cellOf, cellMax, flags); break;
  }
}
void cell_bytes(const uint32_t *cellMax, uint32_t ncell, unsigned long long *out, hipStream_t s) {
  if (ncell) {
    const uint32_t blocks = (ncell / 4 + 255) / 256 + 1;
    k_cell_bytes<<<blocks < 2048 ? blocks : 2048, 256, 0, s>>>(cellMax, ncell, out);
  }
}
void cell_of(const uint32_t *flags, const uint32_t *pos, uint32_t n, uint32_t *cellOf, hipStream_t s) {
  if (n) k_cell_of<<<(n + 255) / 256, 256, 0, s>>>(flags, pos, n, cellOf);
}

void active_flags(const uint32_t *sortedIds, const uint8_t *failed, uint32_t n, uint32_t *flags,
                  hipStream_t s) {
  if (n) k_active_flags<<<(n + 255) / 256, 256, 0, s>>>(sortedIds, failed, n, flags);
}
void compact_active(const uint32_t *sortedIds, const uint32_t *flags, const uint32_t *pos,
                    uint32_t n, uint32_t *act, hipStream_t s) {
  if (n) k_compact_active<<<(n + 255) / 256, 256, 0, s>>>(sortedIds, flags, pos, n, act);
}
void choose_span(const uint32_t *activeCount, const unsigned long long *activeCells, uint32_t ncell,
                 double avgReach, uint32_t spanMax, uint32_t rowsMax, double minTiles, uint32_t *shape,
                 hipStream_t s) {
  k_choose_span<<<1, 64, 0, s>>>(activeCount, activeCells, ncell, avgReach, spanMax, rowsMax, minTiles, shape);
}
void block_flags_dev(const uint64_t *cellKeys, uint32_t ncell, uint32_t ncell0, const uint32_t *span,
                     uint32_t *flags, hipStream_t s) {
  if (ncell) k_block_flags_dev<<<(ncell + 255) / 256, 256, 0, s>>>(cellKeys, ncell, ncell0, span, flags);
}
void compact_blocks_keyed(const uint64_t *cellKeys, const uint32_t *cellStart, const uint32_t *flags,
                          const uint32_t *pos, uint32_t ncell, uint32_t ncell0, const uint32_t *shape,
                          uint32_t n, uint32_t *blockStart, uint64_t *blockKey, hipStream_t s) {
  if (ncell)
    k_compact_blocks_keyed<<<(ncell + 255) / 256, 256, 0, s>>>(cellKeys, cellStart, flags, pos, ncell, ncell0,
                                                                shape, n, blockStart, blockKey);
}
void pair_tiles_dev(const uint32_t *blockStart, const uint64_t *blockKey, const uint32_t *pos,
                    const uint32_t *nblocks, uint32_t ncell, uint32_t ncell0, uint32_t ncell1,
                    const uint32_t *shape, uint32_t *partner, uint32_t *ntiles, uint32_t *tileOff,
                    uint32_t *scratch, uint4 *tiles, hipStream_t s) {
  if (!ncell) return;
  k_pair_tile_counts<<<(ncell + 255) / 256, 256, 0, s>>>(blockStart, blockKey, pos, nblocks, ncell, ncell0, ncell1,
                                                         shape, partner, ntiles);
  scan_exclusive_u32(ntiles, tileOff, ncell, scratch, s);
  k_write_pair_tiles<<<(ncell + 255) / 256, 256, 0, s>>>(blockStart, pos, tileOff, nblocks, partner, ncell, tiles);
}

template <typename T>
void row_sums(const T *dists, uint32_t n, uint32_t k1, double *rowSum, hipStream_t s) {
  if (!n) return;
  const uint32_t rpb = (uint32_t)(SUM_LDS_BYTES / sizeof(T) / k1);
  if (rpb >= 1) {
    const uint32_t rows = rpb < 256 ? rpb : 256;
    k_row_sums<T><<<(n + rows - 1) / rows, 256, 0, s>>>(dists, n, k1, rows, rowSum);
  } else {
    k_row_sums_direct<T><<<(n + 255) / 256, 256, 0, s>>>(dists, n, k1, rowSum);
  }
}

void sum_rows(const double *rowSum, uint32_t n, double *partial, double *out, hipStream_t s) {
  const uint32_t m = (n + SUM_CHUNK - 1) / SUM_CHUNK;
  if (m) k_sum_chunks<<<m, 256, 0, s>>>(rowSum, n, partial);
  k_sum_final<<<1, 64, 0, s>>>(partial, m, out);
}

#define INST(T)                                                                                \
  template void cell_keys_u32<T>(const T *, uint32_t, const GridOf<T>::type &, uint32_t *,      \
                                 uint32_t *, hipStream_t);                                     \
  template void cell_keys_u64<T>(const T *, uint32_t, const GridOf<T>::type &, uint64_t *,      \
                                 uint32_t *, hipStream_t);                                     \
  template void gather_soa<T>(const T *, const uint32_t *, uint32_t, uint32_t, T *, hipStream_t); \
  template void gather_rows<T>(const T *, const uint32_t *, uint32_t, uint32_t, const DimPerm &, T *, \
                               hipStream_t);                                                   \
  template void row_sums<T>(const T *, uint32_t, uint32_t, double *, hipStream_t);          \
  template void prefilter<T>(const T *, uint32_t, const GridOf<T>::type &, const T *, T,      \
                             const uint32_t *, const uint32_t *, const uint64_t *, uint8_t *, \
                             uint32_t, const uint32_t *, uint32_t *, uint32_t *, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace knn
