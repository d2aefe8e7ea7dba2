// batched.hip -- many independent point sets joined together (float32).
//
// No counterpart in the reference, which joins one set per process (main.cu:109-510).  The
// sets of one dimensionality form a group; the group's points are concatenated and every
// stage runs once per group over all of them:
//   per-set variance reorder (main.cu:2521-2615, same FP32 sums as a single-set join),
//   per-set min/max, per-set grid (own width, own key range [keyBase, keyBase + cells)),
//   one radix sort / cell table / dense start table for the whole group,
//   tiles of <= TQ queries per cell (a cell belongs to one set), and the range kernels of
//   range.hip with per-tile set descriptors (grid, K+1, row offset).
// Radius schedule: every query starts at its set's r0 (the radius whose ball holds about
// 3 (K+1) points at the set's mean density); failed queries double their radius and the
// set's grid is rebuilt with cell width = its largest failed radius, so a query's block is
// always at most 3 cells per indexed dim.  Rows are exact whatever the schedule (a query is
// accepted only when its K+1 nearest are provably inside its block) and equal the rows of a
// single-set join of the same set.
//
// Every kernel reads its sizes (cells, tiles) from device memory, so a pass is a fixed
// sequence of launches: the whole join (input copy, reorder, grids, indexes, P passes) is
// captured once in a hipGraph and replayed for later calls with the same set shapes.
// Passes left after the graph (rare) run eagerly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "knn_launch.h"

namespace knn {

namespace {

constexpr uint32_t BLK = 256;
inline uint32_t nblk(size_t n) { return (uint32_t)((n + BLK - 1) / BLK); }

// main.cu:2537-2562 per set: thread (set, dim) sums rows 0,100,... of its set in order
__global__ void k_b_variance(const float *__restrict__ xin, const SetDesc *sets, uint32_t nsets,
                             uint32_t d, float *var) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsets * d) return;
  const uint32_t si = t / d, i = t % d;
  const uint32_t n = sets[si].n;
  const float *pts = xin + (size_t)sets[si].pStart * d;
  const float inv = (float)(1.0 / (100 * 1.0));
  const float denom = (float)n * inv;
  float sum = 0;
  for (uint64_t j = 0; j < n; j += 100) sum += pts[j * d + i];
  const float avg = sum / denom;
  sum = 0;
  for (uint64_t j = 0; j < n; j += 100) {
    float df = pts[j * d + i] - avg;
    float sq = df * df;
    sum += sq;
  }
  var[t] = sum / denom;
}

// stable descending insertion sort of each set's variances (main.cu:2580)
__global__ void k_b_perm(const float *var, uint32_t nsets, uint32_t d, int reorder, uint32_t *perm) {
  const uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nsets) return;
  uint32_t p[KNN_MAXDIM];
  for (uint32_t i = 0; i < d; i++) p[i] = i;
  if (reorder) {
    const float *v = var + (size_t)si * d;
    for (uint32_t i = 1; i < d; i++) {
      const uint32_t pi = p[i];
      const float key = v[pi];
      int j = (int)i - 1;
      while (j >= 0 && key > v[p[j]]) {
        p[j + 1] = p[j];
        j--;
      }
      p[j + 1] = pi;
    }
  }
  for (uint32_t i = 0; i < KNN_MAXDIM; i++) perm[(size_t)si * KNN_MAXDIM + i] = i < d ? p[i] : 0u;
}

__global__ void k_b_permute(const float *__restrict__ xin, uint32_t N, uint32_t d,
                            const uint32_t *__restrict__ setOf, const uint32_t *__restrict__ perm,
                            float *__restrict__ x) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint32_t *pm = perm + (size_t)setOf[p] * KNN_MAXDIM;
  for (uint32_t j = 0; j < d; j++) x[(size_t)j * N + p] = xin[(size_t)p * d + pm[j]];
}

// one block per set: min / max of its indexed dims -> mnmx[si][0..7] min, [8..15] max
__global__ __launch_bounds__(256) void k_b_minmax(const float *__restrict__ x, uint32_t N,
                                                  const SetDesc *sets, float *mnmx) {
  __shared__ float smn[256], smx[256];
  const SetDesc &sd = sets[blockIdx.x];
  for (uint32_t j = 0; j < sd.nd; j++) {
    float mn = __builtin_inff(), mx = -__builtin_inff();
    for (uint32_t i = threadIdx.x; i < sd.n; i += 256) {
      const float v = x[(size_t)j * N + sd.pStart + i];
      mn = fminf(mn, v);
      mx = fmaxf(mx, v);
    }
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + o]);
        smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + o]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      mnmx[(size_t)blockIdx.x * 16 + j] = smn[0];
      mnmx[(size_t)blockIdx.x * 16 + 8 + j] = smx[0];
    }
    __syncthreads();
  }
}

__device__ double unit_ball_volume(uint32_t nd) {
  // V_1 = 2, V_2 = pi, V_n = V_{n-2} * 2 pi / n
  double v = (nd & 1) ? 2.0 : 3.14159265358979323846;
  for (uint32_t m = (nd & 1) ? 3 : 4; m <= nd; m += 2) v *= 2.0 * 3.14159265358979323846 / m;
  return v;
}

// Initial radius per set: the ball holding about 3 (K+1) points at the set's mean density
// over its bounding box; degenerate boxes fall back to the largest extent / 8 (or 1).
__global__ void k_b_init_radius(SetDesc *sets, uint32_t nsets, const float *mnmx, float *width) {
  const uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nsets) return;
  SetDesc &sd = sets[si];
  double vol = 1.0, ext = 0.0;
  for (uint32_t j = 0; j < sd.nd; j++) {
    const double e = (double)mnmx[(size_t)si * 16 + 8 + j] - (double)mnmx[(size_t)si * 16 + j];
    vol *= e;
    ext = e > ext ? e : ext;
  }
  double r = 0.0;
  if (vol > 0.0 && sd.n > 0)
    r = pow(3.0 * sd.k1 * vol / ((double)sd.n * unit_ball_volume(sd.nd)), 1.0 / sd.nd);
  if (!(r > 0.0) || !(r < 3.0e38)) r = ext > 0.0 ? ext / 8.0 : 1.0;
  float rf = (float)r;
  if (!(rf > 0.0f)) rf = 1.0f;
  sd.r0 = rf;
  width[si] = rf;
}

__global__ void k_b_fill_radius(const SetDesc *sets, const uint32_t *setOf, uint32_t N, float *radius,
                                uint8_t *failed) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  radius[p] = sets[setOf[p]].r0;
  failed[p] = 1;
}

// Grid of every set at its requested width (main.cu:2465-2486 arithmetic: lo = min - w,
// nCells = ceil((max + w - lo) / w)); the key ranges are laid end to end.  When the total
// would exceed `bound` cells, every width grows by 1/4 until it fits.  One thread.
__global__ void k_b_grids(SetDesc *sets, uint32_t nsets, const float *mnmx, float *width,
                          uint64_t bound, uint32_t *status) {
  if (blockIdx.x || threadIdx.x) return;
  for (int attempt = 0; attempt < 256; attempt++) {
    uint64_t base = 0;
    bool fits = true;
    for (uint32_t si = 0; si < nsets && fits; si++) {
      SetDesc &sd = sets[si];
      const float w = width[si];
      GridParams &g = sd.g;
      g.eps = w;
      g.nd = sd.nd;
      g.keyBase = base;
      uint64_t tot = 1;
      for (uint32_t j = 0; j < KNN_MAX_INDEXED; j++) {
        if (j < sd.nd) {
          const float lo = mnmx[(size_t)si * 16 + j] - w;
          const float hi = mnmx[(size_t)si * 16 + 8 + j] + w;
          const float q = (hi - lo) / w;
          float c = ceilf(q);
          if (!(c >= 1.0f)) c = 1.0f;
          if (c > 2147483647.0f) c = 2147483647.0f;
          g.minArr[j] = lo;
          g.nCells[j] = (uint32_t)c;
          g.mult[j] = tot;
          tot *= (uint64_t)g.nCells[j];
          if (tot > bound) {
            fits = false;
            break;
          }
        } else {
          g.minArr[j] = 0.0f;
          g.nCells[j] = 1;
          g.mult[j] = tot;
        }
      }
      base += tot;
      if (base > bound) fits = false;
    }
    if (fits) {
      status[0] = 0;
      status[1] = (uint32_t)base;
      return;
    }
    for (uint32_t si = 0; si < nsets; si++) width[si] *= 1.25f;
  }
  // no fit (only with very many sets of many indexed dims): one cell per set, keys 0..nsets-1
  // (bound >= 8 N >= nsets); every query then scans its whole set, which is slow but exact
  for (uint32_t si = 0; si < nsets; si++) {
    GridParams &g = sets[si].g;
    g.keyBase = si;
    g.nd = sets[si].nd;
    for (uint32_t j = 0; j < KNN_MAX_INDEXED; j++) {
      g.minArr[j] = 0.0f;
      g.nCells[j] = 1;
      g.mult[j] = 1;
    }
  }
  status[0] = 1;
  status[1] = nsets;
}

__global__ void k_b_cell_keys(const float *__restrict__ x, uint32_t N, const SetDesc *sets,
                              const uint32_t *__restrict__ setOf, uint32_t *keys, uint32_t *ids) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const GridParams &g = sets[setOf[p]].g;
  uint64_t id = g.keyBase;
  for (uint32_t j = 0; j < g.nd; j++)
    id += (uint64_t)cell_coord<float>(x[(size_t)j * N + p], g.minArr[j], g.eps, g.nCells[j]) * g.mult[j];
  keys[p] = (uint32_t)id;
  ids[p] = p;
}

// dense[cellKeys[c]] = points of cell c, for c < *ncell
__global__ void k_b_dense_counts(const uint64_t *cellKeys, const uint32_t *cellStart,
                                 const uint32_t *ncell, uint32_t N, uint32_t *dense) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N || c >= *ncell) return;
  dense[cellKeys[c]] = cellStart[c + 1] - cellStart[c];
}

__global__ void k_b_flags_u8(const uint32_t *sortedIds, const uint8_t *f, uint32_t N, uint32_t *flags) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  flags[p] = f[sortedIds[p]] ? 1u : 0u;
}

// tiles of <= tq active queries per non-empty cell (cells c < *ncell; entries up to N + 1)
__global__ void k_b_tile_counts(const uint32_t *cellStart, const uint32_t *pos, const uint32_t *ncell,
                                uint32_t N, uint32_t tq, uint32_t *ntiles) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > N) return;
  if (c >= *ncell) {
    ntiles[c] = 0;
    return;
  }
  const uint32_t a0 = pos[cellStart[c]], a1 = pos[cellStart[c + 1]];
  ntiles[c] = (a1 - a0 + tq - 1) / tq;
}

__global__ void k_b_write_tiles(const uint32_t *cellStart, const uint32_t *pos, const uint32_t *tileOff,
                                const uint32_t *ncell, uint32_t N, uint32_t tq, uint4 *tiles) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N || c >= *ncell) return;
  const uint32_t a0 = pos[cellStart[c]], a1 = pos[cellStart[c + 1]];
  uint32_t t = tileOff[c];
  for (uint32_t a = a0; a < a1; a += tq, t++)
    tiles[t] = make_uint4(c, a, (a1 - a < tq) ? a1 - a : tq, a == a0 ? 1u : 0u);
}

// failed queries: radius *= 2 and the set's largest failed radius (float bits, atomicMax)
__global__ void k_b_grow(const uint32_t *setOf, const uint8_t *failed, uint32_t N, float *radius,
                         uint32_t *setMaxBits) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N || !failed[p]) return;
  const float r = radius[p] * 2.0f;
  radius[p] = r;
  atomicMax(&setMaxBits[setOf[p]], __float_as_uint(r));
}

// next width of each set: its largest failed radius when it has failures (else unchanged)
__global__ void k_b_next_width(const uint32_t *setMaxBits, uint32_t nsets, float *width) {
  const uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nsets) return;
  const uint32_t b = setMaxBits[si];
  if (b) {
    const float r = __uint_as_float(b);
    width[si] = r > width[si] ? r : width[si];
  }
}

}  // namespace

// ------------------------------------------------------------------ host launchers
void batch_variance_perm(const float *xin, const SetDesc *sets, uint32_t nsets, uint32_t d, int reorder,
                         float *var, uint32_t *perm, hipStream_t s) {
  k_b_variance<<<nblk((size_t)nsets * d), BLK, 0, s>>>(xin, sets, nsets, d, var);
  k_b_perm<<<nblk(nsets), BLK, 0, s>>>(var, nsets, d, reorder, perm);
}
void batch_permute(const float *xin, uint32_t N, uint32_t d, const uint32_t *setOf, const uint32_t *perm,
                   float *x, hipStream_t s) {
  if (N) k_b_permute<<<nblk(N), BLK, 0, s>>>(xin, N, d, setOf, perm, x);
}
void batch_minmax(const float *x, uint32_t N, const SetDesc *sets, uint32_t nsets, float *mnmx,
                  hipStream_t s) {
  if (nsets) k_b_minmax<<<nsets, 256, 0, s>>>(x, N, sets, mnmx);
}
void batch_init_radius(SetDesc *sets, uint32_t nsets, const float *mnmx, float *width,
                       const uint32_t *setOf, uint32_t N, float *radius, uint8_t *failed, hipStream_t s) {
  k_b_init_radius<<<nblk(nsets), BLK, 0, s>>>(sets, nsets, mnmx, width);
  if (N) k_b_fill_radius<<<nblk(N), BLK, 0, s>>>(sets, setOf, N, radius, failed);
}
void batch_grids(SetDesc *sets, uint32_t nsets, const float *mnmx, float *width, uint64_t bound,
                 uint32_t *status, hipStream_t s) {
  k_b_grids<<<1, 1, 0, s>>>(sets, nsets, mnmx, width, bound, status);
}
void batch_cell_keys(const float *x, uint32_t N, const SetDesc *sets, const uint32_t *setOf,
                     uint32_t *keys, uint32_t *ids, hipStream_t s) {
  if (N) k_b_cell_keys<<<nblk(N), BLK, 0, s>>>(x, N, sets, setOf, keys, ids);
}
void batch_dense_counts(const uint64_t *cellKeys, const uint32_t *cellStart, const uint32_t *ncell,
                        uint32_t N, uint32_t *dense, hipStream_t s) {
  if (N) k_b_dense_counts<<<nblk(N), BLK, 0, s>>>(cellKeys, cellStart, ncell, N, dense);
}
void batch_flags_u8(const uint32_t *sortedIds, const uint8_t *f, uint32_t N, uint32_t *flags, hipStream_t s) {
  if (N) k_b_flags_u8<<<nblk(N), BLK, 0, s>>>(sortedIds, f, N, flags);
}
void batch_tiles(const uint32_t *cellStart, const uint32_t *pos, const uint32_t *ncell, uint32_t N,
                 uint32_t tq, uint32_t *ntiles, uint32_t *tileOff, uint32_t *scratch, uint4 *tiles,
                 hipStream_t s) {
  k_b_tile_counts<<<nblk((size_t)N + 1), BLK, 0, s>>>(cellStart, pos, ncell, N, tq, ntiles);
  scan_exclusive_u32(ntiles, tileOff, N + 1, scratch, s);
  if (N) k_b_write_tiles<<<nblk(N), BLK, 0, s>>>(cellStart, pos, tileOff, ncell, N, tq, tiles);
}
void batch_grow(const uint32_t *setOf, const uint8_t *failed, uint32_t N, float *radius,
                uint32_t *setMaxBits, float *width, uint32_t nsets, hipStream_t s) {
  fill_u32(setMaxBits, 0u, nsets, s);
  if (N) k_b_grow<<<nblk(N), BLK, 0, s>>>(setOf, failed, N, radius, setMaxBits);
  k_b_next_width<<<nblk(nsets), BLK, 0, s>>>(setMaxBits, nsets, width);
}

}  // namespace knn
