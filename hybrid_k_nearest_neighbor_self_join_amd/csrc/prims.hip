// prims.hip -- scan, stable LSD radix sort, min/max reductions and fills for gfx950.
//
// These replace the Thrust calls of the reference's index build (thrust::sort + unique
// GPU.cu:3884-3886, thrust::sort_by_key GPU.cu:3959).  Wave64 throughout: a block is four
// wavefronts; per-wave work uses 64-bit ballots and shuffles.
#include "knn_launch.h"

namespace knn {

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr uint32_t SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;  // 4096
// Up to this many tiles, each block of the down-sweep sums the tile totals before its own
// (at most 32 KB, read from L2) instead of a recursive scan of the totals: two launches per
// scan and no inter-block waiting.
constexpr uint32_t SCAN_DIRECT_TILES = 8192;

// inclusive scan of v over the 256 threads of a block; returns the exclusive prefix and
// sets *total.  Uses __shfl_up inside a wave and 4 LDS slots across waves.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t s_w[SCAN_BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_BLOCK / 64; i++) {
    uint32_t x = s_w[i];
    woff += (i < w) ? x : 0u;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return woff + inc - v;
}

// The thread's SCAN_ITEMS consecutive inputs of the tile (16-byte loads when the tile is
// whole), zero past n.
__device__ __forceinline__ void load_items(const uint32_t *in, uint32_t n, uint32_t base, uint32_t (&v)[SCAN_ITEMS]) {
  if (base + SCAN_ITEMS <= n && (reinterpret_cast<uintptr_t>(in + base) & 15u) == 0) {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i += 4) {
      const uint4 q = *reinterpret_cast<const uint4 *>(in + base + i);
      v[i] = q.x;
      v[i + 1] = q.y;
      v[i + 2] = q.z;
      v[i + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) v[i] = base + i < n ? in[base + i] : 0u;
  }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t *in, uint32_t n,
                                                             uint32_t *sums) {
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  load_items(in, n, base, v);
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) acc += v[i];
  uint32_t tot;
  (void)block_exclusive_scan(acc, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Down-sweep of one tile.  The tile's offset is offs[blockIdx.x] (a scanned array of tile
// totals), or, with direct set, the sum of sums[0 .. blockIdx.x) (raw tile totals).
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const uint32_t *in, uint32_t *out,
                                                           uint32_t n, const uint32_t *offs, int direct) {
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  load_items(in, n, base, v);
  uint32_t tileOff = 0;
  if (direct) {
    uint32_t a = 0;
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += SCAN_BLOCK) a += offs[i];
    uint32_t tot;
    (void)block_exclusive_scan(a, &tot);
    tileOff = tot;
  } else if (offs) {
    tileOff = offs[blockIdx.x];
  }
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) acc += v[i];
  uint32_t tot;
  uint32_t off = tileOff + block_exclusive_scan(acc, &tot);
  uint32_t o[SCAN_ITEMS];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    o[i] = off;
    off += v[i];
  }
  if (base + SCAN_ITEMS <= n && (reinterpret_cast<uintptr_t>(out + base) & 15u) == 0) {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i += 4)
      *reinterpret_cast<uint4 *>(out + base + i) = make_uint4(o[i], o[i + 1], o[i + 2], o[i + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
      if (base + i < n) out[base + i] = o[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_BLOCK - 1) out[n] = off;
}

}  // namespace

size_t scan_scratch_words(uint32_t n) {
  size_t words = 0;
  uint32_t m = n;
  while (m > SCAN_TILE) {
    uint32_t nb = (m + SCAN_TILE - 1) / SCAN_TILE;
    words += (size_t)nb + 1;
    m = nb;
  }
  return words + 16;
}

void scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *scratch,
                        hipStream_t s) {
  if (n == 0) {
    fill_u32(out, 0u, 1, s);
    return;
  }
  uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    k_scan_down<<<1, SCAN_BLOCK, 0, s>>>(in, out, n, nullptr, 0);
    return;
  }
  uint32_t *sums = scratch;  // nb + 1 words
  k_scan_reduce<<<nb, SCAN_BLOCK, 0, s>>>(in, n, sums);
  if (nb <= SCAN_DIRECT_TILES) {
    k_scan_down<<<nb, SCAN_BLOCK, 0, s>>>(in, out, n, sums, 1);
    return;
  }
  scan_exclusive_u32(sums, sums, nb, scratch + nb + 1, s);
  k_scan_down<<<nb, SCAN_BLOCK, 0, s>>>(in, out, n, sums, 0);
}

// ------------------------------------------------------------------------- radix sort
namespace {

constexpr int RADIX_BLOCK = 256;
constexpr int RADIX_ROUNDS = 16;
constexpr uint32_t RADIX_TILE = RADIX_BLOCK * RADIX_ROUNDS;  // 4096 keys per block

template <typename K>
__global__ __launch_bounds__(RADIX_BLOCK) void k_radix_hist(const K *keys, uint32_t n,
                                                             uint32_t ntiles, int shift,
                                                             uint32_t *counts) {
  __shared__ uint32_t h[256];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * RADIX_TILE;
#pragma unroll 4
  for (int r = 0; r < RADIX_ROUNDS; r++) {
    uint32_t idx = base + r * RADIX_BLOCK + t;
    if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[(size_t)t * ntiles + blockIdx.x] = h[t];
}

// Stable scatter in two steps.  Ranking: round r of a tile covers keys base + r*256 + [0,256),
// thread t takes key t of the round, so tile order = key order; inside a round each wave
// ranks its lanes by the 8-bit match mask (8 ballots), earlier waves' counts come from LDS,
// and each key goes to its digit-ordered slot of the tile in LDS.  Writing: the tile's keys
// leave LDS in slot order, so each digit's run (about 16 keys) is written to consecutive
// addresses.  The tile's count of digit t is offsets[t][b+1] - offsets[t][b] of the exclusive
// scan over [digit][tile] (offsets[256 * ntiles] = n).
template <typename K>
__global__ __launch_bounds__(RADIX_BLOCK) void k_radix_scatter(
    const K *__restrict__ kin, const uint32_t *__restrict__ vin, K *__restrict__ kout,
    uint32_t *__restrict__ vout, const uint32_t *__restrict__ offsets, uint32_t n,
    uint32_t ntiles, int shift) {
  __shared__ uint32_t s_off[256];   // global start of digit t for this tile
  __shared__ uint32_t s_loc[256];   // tile-local start of digit t
  __shared__ uint32_t s_run[256];   // next free tile-local slot of digit t
  __shared__ uint32_t s_wc[4][256];
  __shared__ K s_key[RADIX_TILE];
  __shared__ uint32_t s_val[RADIX_TILE];
  const int t = threadIdx.x, w = t >> 6;
  const size_t oi = (size_t)t * ntiles + blockIdx.x;
  const uint32_t g0 = offsets[oi];
  const uint32_t cnt = offsets[oi + 1] - g0;
  uint32_t total;
  const uint32_t loc = block_exclusive_scan(cnt, &total);
  s_off[t] = g0;
  s_loc[t] = loc;
  s_run[t] = loc;
  s_wc[0][t] = 0;
  s_wc[1][t] = 0;
  s_wc[2][t] = 0;
  s_wc[3][t] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * RADIX_TILE;
  const uint64_t lt = lanemask_lt();
  for (int r = 0; r < RADIX_ROUNDS; r++) {
    const uint32_t rb = base + r * RADIX_BLOCK;
    if (rb >= n) break;  // uniform across the block
    const uint32_t idx = rb + t;
    const bool valid = idx < n;
    K key = valid ? kin[idx] : (K)0;
    uint32_t val = valid ? vin[idx] : 0u;
    uint32_t digit = (uint32_t)(key >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      bool bit = (digit >> b) & 1u;
      uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    uint32_t rank = __popcll(m & lt);
    if (valid && rank == 0) s_wc[w][digit] = __popcll(m);
    __syncthreads();
    if (valid) {
      uint32_t off = s_run[digit] + rank;
      for (int ww = 0; ww < w; ww++) off += s_wc[ww][digit];
      s_key[off] = key;
      s_val[off] = val;
    }
    __syncthreads();
    s_run[t] += s_wc[0][t] + s_wc[1][t] + s_wc[2][t] + s_wc[3][t];
    s_wc[0][t] = 0;
    s_wc[1][t] = 0;
    s_wc[2][t] = 0;
    s_wc[3][t] = 0;
    __syncthreads();
  }
  for (uint32_t i = t; i < total; i += RADIX_BLOCK) {
    const K key = s_key[i];
    const uint32_t digit = (uint32_t)(key >> shift) & 255u;
    const uint32_t g = s_off[digit] + (i - s_loc[digit]);
    kout[g] = key;
    vout[g] = s_val[i];
  }
}

template <typename K>
void radix_sort_impl(K *keys, uint32_t *vals, K *keys_tmp, uint32_t *vals_tmp, uint32_t n,
                     int bits, uint32_t *scratch, hipStream_t s) {
  if (n <= 1 || bits <= 0) return;
  const uint32_t ntiles = (n + RADIX_TILE - 1) / RADIX_TILE;
  const uint32_t ncounts = 256u * ntiles;
  uint32_t *counts = scratch;                   // ncounts + 1
  uint32_t *scan_scr = scratch + ncounts + 1;   // scan scratch
  K *ka = keys, *kb = keys_tmp;
  uint32_t *va = vals, *vb = vals_tmp;
  const int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; p++) {
    const int shift = 8 * p;
    k_radix_hist<K><<<ntiles, RADIX_BLOCK, 0, s>>>(ka, n, ntiles, shift, counts);
    scan_exclusive_u32(counts, counts, ncounts, scan_scr, s);
    k_radix_scatter<K><<<ntiles, RADIX_BLOCK, 0, s>>>(ka, va, kb, vb, counts, n, ntiles, shift);
    K *tk = ka;
    ka = kb;
    kb = tk;
    uint32_t *tv = va;
    va = vb;
    vb = tv;
  }
  if (ka != keys) {
    hipMemcpyAsync(keys, ka, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(vals, va, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
  }
}

}  // namespace

size_t radix_scratch_words(uint32_t n) {
  const uint32_t ntiles = (n + RADIX_TILE - 1) / RADIX_TILE;
  const uint32_t ncounts = 256u * (ntiles ? ntiles : 1);
  return (size_t)ncounts + 1 + scan_scratch_words(ncounts);
}

void radix_sort_pairs_u32(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                          uint32_t n, int bits, uint32_t *scratch, hipStream_t s) {
  radix_sort_impl<uint32_t>(keys, vals, keys_tmp, vals_tmp, n, bits, scratch, s);
}

void radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_tmp, uint32_t *vals_tmp,
                          uint32_t n, int bits, uint32_t *scratch, hipStream_t s) {
  radix_sort_impl<uint64_t>(keys, vals, keys_tmp, vals_tmp, n, bits, scratch, s);
}

// ------------------------------------------------------------------------- min / max
namespace {
constexpr int MM_BLOCK = 256;
constexpr uint32_t MM_MAXBLOCKS = 1024;

template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return b < a ? b : a; }
template <typename T>
__device__ __forceinline__ T tmax(T a, T b) { return b > a ? b : a; }

// block b reduces rows [b*chunk, (b+1)*chunk) of dim blockIdx.y
template <typename T>
__global__ __launch_bounds__(MM_BLOCK) void k_minmax_part(const T *x, uint32_t n, uint32_t chunk,
                                                          T *pmin, T *pmax) {
  __shared__ T smin[MM_BLOCK], smax[MM_BLOCK];
  const uint32_t j = blockIdx.y;
  const T *xj = x + (size_t)j * n;
  uint32_t lo = blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  T mn = xj[lo < n ? lo : 0], mx = mn;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += MM_BLOCK) {
    T v = xj[i];
    mn = tmin(mn, v);
    mx = tmax(mx, v);
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int o = MM_BLOCK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      smin[threadIdx.x] = tmin(smin[threadIdx.x], smin[threadIdx.x + o]);
      smax[threadIdx.x] = tmax(smax[threadIdx.x], smax[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmin[(size_t)j * gridDim.x + blockIdx.x] = smin[0];
    pmax[(size_t)j * gridDim.x + blockIdx.x] = smax[0];
  }
}

template <typename T>
// one wave per dim (block j): lane l folds partials l, l+64, ...; then a wave reduction (min
// and max are exact in any order)
__global__ __launch_bounds__(64) void k_minmax_final(const T *pmin, const T *pmax, uint32_t nb, uint32_t nd,
                                                     T *mn, T *mx) {
  const uint32_t j = blockIdx.x, l = threadIdx.x;
  if (j >= nd) return;
  T a = pmin[(size_t)j * nb], b = pmax[(size_t)j * nb];
  for (uint32_t i = l; i < nb; i += 64) {
    a = tmin(a, pmin[(size_t)j * nb + i]);
    b = tmax(b, pmax[(size_t)j * nb + i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = tmin(a, (T)__shfl_xor(a, o, 64));
    b = tmax(b, (T)__shfl_xor(b, o, 64));
  }
  if (l == 0) {
    mn[j] = a;
    mx[j] = b;
  }
}
}  // namespace

size_t minmax_scratch_elems(uint32_t n, uint32_t nd) { return 2ull * MM_MAXBLOCKS * nd; }

template <typename T>
void minmax_soa(const T *x, uint32_t n, uint32_t nd, T *mn, T *mx, T *scratch, hipStream_t s) {
  uint32_t nb = (n + 8191) / 8192;
  if (nb > MM_MAXBLOCKS) nb = MM_MAXBLOCKS;
  if (nb == 0) nb = 1;
  uint32_t chunk = (n + nb - 1) / nb;
  T *pmin = scratch, *pmax = scratch + (size_t)MM_MAXBLOCKS * nd;
  k_minmax_part<T><<<dim3(nb, nd), MM_BLOCK, 0, s>>>(x, n, chunk, pmin, pmax);
  if (nd) k_minmax_final<T><<<nd, 64, 0, s>>>(pmin, pmax, nb, nd, mn, mx);
}
template void minmax_soa<float>(const float *, uint32_t, uint32_t, float *, float *, float *,
                                hipStream_t);
template void minmax_soa<double>(const double *, uint32_t, uint32_t, double *, double *, double *,
                                 hipStream_t);

// ------------------------------------------------------------------------- fills
namespace {
__global__ void k_iota(uint32_t *a, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}
__global__ void k_fill_u8(uint8_t *a, uint8_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_fill_u32(uint32_t *a, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
}  // namespace

void iota_u32(uint32_t *a, uint32_t n, hipStream_t s) {
  if (n) k_iota<<<(n + 255) / 256, 256, 0, s>>>(a, n);
}
void fill_u8(uint8_t *a, uint8_t v, size_t n, hipStream_t s) {
  if (n) k_fill_u8<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, v, n);
}
void fill_u32(uint32_t *a, uint32_t v, size_t n, hipStream_t s) {
  if (n) k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, v, n);
}

}  // namespace knn
