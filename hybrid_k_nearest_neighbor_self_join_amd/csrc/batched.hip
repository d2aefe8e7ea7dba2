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
// Radius schedule (per query, any schedule gives the same rows): each set first gets a grid of
// width wf (about K+1 points per cell at its mean density); a query's initial radius is the
// one whose ball holds about 2 (K+1) points at the density of its own and the two adjacent
// cells along dim 0.  Then passes over radius octaves: in each pass every set joins its
// pending queries of its lowest occupied radius class (an octave [R/2, R) with <= 2 indexed
// dims, half an octave above) on a grid matched to R (width R/2 with <= 2 indexed dims, R
// above); failed queries move their radius one class up.
// Rows are exact whatever the radii (a query is accepted only when its K+1 nearest are
// provably inside its block) and equal the rows of a single-set join of the same set.
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

// Per set: wf = (V (K+1) / n)^(1/nd) over its bounding box (degenerate boxes: largest extent / 8,
// or 1); the first grid uses it.  All queries pending.
__global__ void k_b_init_width(SetDesc *sets, uint32_t nsets, const float *mnmx, float *width) {
  const uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nsets) return;
  SetDesc &sd = sets[si];
  double vol = 1.0, ext = 0.0;
  for (uint32_t j = 0; j < sd.nd; j++) {
    const double e = (double)mnmx[(size_t)si * 16 + 8 + j] - (double)mnmx[(size_t)si * 16 + j];
    vol *= e;
    ext = e > ext ? e : ext;
  }
  double w = 0.0;
  if (vol > 0.0 && sd.n > 0) w = pow((double)sd.k1 * vol / (double)sd.n, 1.0 / sd.nd);
  if (!(w > 0.0) || !(w < 3.0e38)) w = ext > 0.0 ? ext / 8.0 : 1.0;
  float wf = (float)w;
  if (!(wf > 0.0f)) wf = 1.0f;
  sd.r0 = wf;
  width[si] = wf;
}

__global__ void k_b_pending(uint32_t N, uint8_t *failed) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) failed[p] = 1;
}

// quarter-octave bin of r relative to wf: floor(4 log2(r / wf)) + 32 in [0, 63]
__device__ __forceinline__ int bin_of(float r, float wf) {
  const float x = 4.0f * log2f(r / wf);
  if (!(x >= -32.0f)) return 0;
  if (x >= 31.0f) return 63;
  const int b = (int)floorf(x) + 32;
  return b < 0 ? 0 : (b > 63 ? 63 : b);
}

// Initial radius of every query from the points of its own and the two adjacent cells along
// dim 0 on the wf grid: density rho = count / volume, r = (2 (K+1) / (V_nd rho))^(1/nd),
// kept within [wf / 64, 64 wf].
__global__ void k_b_density_radius(const float *__restrict__ x, uint32_t N, const SetDesc *sets,
                                   const uint32_t *__restrict__ setOf, const uint32_t *__restrict__ dense,
                                   float *radius, uint8_t *bin) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  const SetDesc &sd = sets[setOf[q]];
  const GridParams &g = sd.g;
  uint64_t row = g.keyBase;
  uint32_t c0 = 0;
  for (uint32_t j = 0; j < g.nd; j++) {
    const uint32_t c = cell_coord<float>(x[(size_t)j * N + q], g.minArr[j], g.eps, g.nCells[j]);
    if (j == 0) c0 = c; else row += (uint64_t)c * g.mult[j];
  }
  const uint32_t lo = c0 > 0 ? c0 - 1 : 0;
  const uint32_t hi = c0 + 1 < g.nCells[0] ? c0 + 1 : g.nCells[0] - 1;
  const uint32_t cnt = dense[row + hi + 1] - dense[row + lo];
  const double wf = (double)sd.r0;
  const double vol = (double)(hi - lo + 1) * pow((double)g.eps, (double)g.nd);
  const double rho = (double)(cnt > 0 ? cnt : 1u) / vol;
  double r = pow(2.0 * sd.k1 / (unit_ball_volume(g.nd) * rho), 1.0 / g.nd);
  r = fmin(fmax(r, wf / 64.0), wf * 64.0);
  radius[q] = (float)r;
  bin[q] = (uint8_t)bin_of((float)r, sd.r0);
}

// per set: the lowest radius bin among its pending queries (wave-aggregated; 64 = none)
__global__ void k_b_min_bin(const uint32_t *__restrict__ setOf, const uint8_t *failed, const uint8_t *bin,
                            uint32_t N, uint32_t *setMinBin) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool pend = q < N && failed[q];
  uint32_t key = pend ? (setOf[q] << 8) | bin[q] : 0xffffffffu;
  bool mine = pend;
  for (;;) {
    const uint64_t live = __ballot(mine);
    if (!live) break;
    const int leader = __ffsll((long long)live) - 1;
    const uint32_t ls = (uint32_t)__shfl((int)(key >> 8), leader, 64);
    uint32_t b = (mine && (key >> 8) == ls) ? (key & 0xffu) : 0xffu;
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t t = (uint32_t)__shfl_xor((int)b, o, 64);
      b = t < b ? t : b;
    }
    if ((int)(threadIdx.x & 63) == leader) atomicMin(&setMinBin[ls], b);
    if ((key >> 8) == ls) mine = false;
  }
}

// per set: the pass's radius class (octave [c, c + 3] with <= 2 indexed dims, half octave
// [c, c + 1] above)
// and grid width R / div (R = the class's upper radius); sets without pending queries keep
// their grid
__global__ void k_b_class_width(const SetDesc *sets, uint32_t nsets, const uint32_t *setMinBin, float *width) {
  const uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nsets) return;
  const uint32_t c = setMinBin[si];
  if (c >= 64) return;
  const double span = sets[si].nd <= 2 ? 4.0 : 2.0;  // bins per class
  const double R = (double)sets[si].r0 * exp2(((double)c + span - 32.0) / 4.0);
  const double div = sets[si].nd <= 2 ? 2.0 : 1.0;
  width[si] = (float)(R / div);
}

// act[q] = pending with a bin in its set's class
__global__ void k_b_class_flags(const SetDesc *sets, const uint32_t *__restrict__ setOf, const uint8_t *failed,
                                const uint8_t *bin, const uint32_t *setMinBin, uint32_t N, uint8_t *act) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  bool a = false;
  if (failed[q]) {
    const uint32_t si = setOf[q];
    const uint32_t c = setMinBin[si];
    a = bin[q] >= c && bin[q] <= c + (sets[si].nd <= 2 ? 3u : 1u);
  }
  act[q] = a ? 1 : 0;
}

// queries of the pass that failed: radius into the next class; pending count into *pending
__global__ void k_b_grow_class(const SetDesc *sets, const uint32_t *__restrict__ setOf, const uint8_t *act,
                               const uint8_t *failed, uint32_t N, float *radius, uint8_t *bin,
                               uint32_t *pending) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  bool pend = false;
  if (q < N) {
    pend = failed[q] != 0;
    if (pend && act[q]) {
      // one class further: an octave, half an octave with >= 3 indexed dims
      const float r = radius[q] * (sets[setOf[q]].nd <= 2 ? 2.0f : 1.41421356f);
      radius[q] = r;
      const int b = bin_of(r, sets[setOf[q]].r0);
      // at least one bin further, so that every failure moves the query to a later class
      bin[q] = (uint8_t)(b > (int)bin[q] ? b : ((int)bin[q] + 1 < 64 ? (int)bin[q] + 1 : 63));
    }
  }
  const uint64_t m = __ballot(pend);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(pending, (uint32_t)__popcll(m));
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
  for (uint32_t a = a0; a < a1; a += tq, t++) {
    const uint32_t m = (a1 - a < tq) ? a1 - a : tq;
    tiles[t] = make_uint4(0u, a, m, m);
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
void batch_init_width(SetDesc *sets, uint32_t nsets, const float *mnmx, float *width, uint32_t N,
                      uint8_t *failed, hipStream_t s) {
  k_b_init_width<<<nblk(nsets), BLK, 0, s>>>(sets, nsets, mnmx, width);
  if (N) k_b_pending<<<nblk(N), BLK, 0, s>>>(N, failed);
}
void batch_density_radius(const float *x, uint32_t N, const SetDesc *sets, const uint32_t *setOf,
                          const uint32_t *dense, float *radius, uint8_t *bin, hipStream_t s) {
  if (N) k_b_density_radius<<<nblk(N), BLK, 0, s>>>(x, N, sets, setOf, dense, radius, bin);
}
void batch_class(const SetDesc *sets, uint32_t nsets, const uint32_t *setOf, const uint8_t *failed,
                 const uint8_t *bin, uint32_t N, uint32_t *setMinBin, float *width, uint8_t *act, hipStream_t s) {
  fill_u32(setMinBin, 64u, nsets, s);
  if (N) k_b_min_bin<<<nblk(N), BLK, 0, s>>>(setOf, failed, bin, N, setMinBin);
  k_b_class_width<<<nblk(nsets), BLK, 0, s>>>(sets, nsets, setMinBin, width);
  if (N) k_b_class_flags<<<nblk(N), BLK, 0, s>>>(sets, setOf, failed, bin, setMinBin, N, act);
}
void batch_grow_class(const SetDesc *sets, const uint32_t *setOf, const uint8_t *act, const uint8_t *failed,
                      uint32_t N, float *radius, uint8_t *bin, uint32_t *pending, hipStream_t s) {
  fill_u32(pending, 0u, 1, s);
  if (N) k_b_grow_class<<<nblk(N), BLK, 0, s>>>(sets, setOf, act, failed, N, radius, bin, pending);
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
}  // namespace knn
