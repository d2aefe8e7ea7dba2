// brute.hip -- exact top-(K+1) of a list of queries against every point, without the index.
//
// Counterpart of the reference's no-selectivity path: distanceTableNDBruteForce
// (GPU.cu:1903-2930) with kernelNDBruteForce (kernel.cu:902-975), which emits every
// (query, point, distance) triple into the pair buffer and sorts it, and the criterion
// checkIndexSelectivity (main.cu:2687-2716) that would select it (both unused by the
// reference's main loop).  Here no pair is stored: each query keeps its K+1 best in LDS.
//
// Ranking is that of the range kernels: FP64 key = sum_j ((double)c_j - (double)q_j)^2 (dims
// in the reordered order, no FMA), ties by point id; the reported distance is sqrt of the
// input-type sum sum_j (c_j - q_j)^2.  A row is therefore bit-identical to the one the range
// kernels write for the same query.
//
// k_brute_partial: one wave per (tile of up to 4 queries, chunk of points); lane = point,
//   64 points per round, each point's key computed for every query of the tile.  Candidates
//   below the query's current (K+1)-th (key, id) are appended to an LDS buffer through a
//   ballot; a full buffer is bitonic-sorted and merged into the sorted best list by rank
//   (each element's output position = own index + elements of the other list before it).
//   With one chunk the kernel writes the rows; otherwise each chunk's K+1 best go to a
//   partial table.
// k_brute_merge: one wave per query streams its partial lists through the same collector
//   and writes the row.
#include <type_traits>

#include "knn_collect.h"
#include "knn_launch.h"

namespace knn {

namespace {

constexpr uint32_t BR_QW = 4;  // queries per wave in k_brute_partial (upper bound)

template <int D, typename T>
__device__ __forceinline__ double br_key(const T *c, const T *q) {
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < D; j++) {
    const double dx = (double)c[j] - (double)q[j];
    const double t = dx * dx;
    s = s + t;
  }
  return s;
}

template <int D, typename T>
__device__ __forceinline__ T br_dist_sq(const T *c, const T *q) {
  T s = (T)0;
#pragma unroll
  for (int j = 0; j < D; j++) {
    const T df = c[j] - q[j];
    const T t = df * df;
    s = s + t;
  }
  return s;
}

// writes query q's row from a finished collector (entries past nb, possible only when fewer
// than k1 points exist, are left as id -1 / distance 0)
template <typename T, int D>
__device__ void br_write_row(const BruteArgs &a, const Collector &c, uint32_t q, uint32_t lane) {
  const T *x = static_cast<const T *>(a.x);
  T qc[D];
#pragma unroll
  for (int j = 0; j < D; j++) qc[j] = x[(size_t)j * a.n + q];
  const size_t row = (size_t)q * a.k1;
  T *od = static_cast<T *>(a.outDists);
  for (uint32_t e = lane; e < a.k1; e += 64) {
    if (e < c.nb) {
      const uint32_t id = c.best_id(e);
      T cc[D];
#pragma unroll
      for (int j = 0; j < D; j++) cc[j] = x[(size_t)j * a.n + id];
      a.outIds[row + e] = (int32_t)id;
      od[row + e] = sqrt(br_dist_sq<D, T>(cc, qc));
    } else {
      a.outIds[row + e] = -1;
      od[row + e] = (T)0;
    }
  }
  if (lane == 0 && a.failed) a.failed[q] = 0;
}

template <typename T, int D>
__global__ __launch_bounds__(64) void k_brute_partial(BruteArgs a) {
  extern __shared__ __align__(16) unsigned char br_lds[];
  const uint32_t lane = threadIdx.x;
  const uint32_t tile = blockIdx.x / a.nchunks;
  const uint32_t chunk = blockIdx.x - tile * a.nchunks;
  const uint32_t q0 = tile * a.qw;
  const uint32_t qw = (a.nq - q0) < a.qw ? (a.nq - q0) : a.qw;
  const T *x = static_cast<const T *>(a.x);

  Collector col[BR_QW];
  T qc[BR_QW][D];
  uint32_t qid[BR_QW];
  // float input: the FP32 sum s = sum_j (c_j - q_j)^2 is within 2^-19 of the FP64 key
  // relative, so a point whose key can reach the threshold tk has s <= tk (1 + 2^-16); the
  // key is computed only for such points
  T fb[BR_QW];
#pragma unroll
  for (uint32_t t = 0; t < BR_QW; t++) {
    col[t].init(br_lds + Collector::bytes(a.k1) * t, a.k1);
    qid[t] = t < qw ? (a.queries ? a.queries[q0 + t] : q0 + t) : 0u;
#pragma unroll
    for (int j = 0; j < D; j++) qc[t][j] = t < qw ? x[(size_t)j * a.n + qid[t]] : (T)0;
    fb[t] = (T)__builtin_inf();
  }
  const uint64_t c0 = (uint64_t)chunk * a.chunk;
  const uint32_t p0 = (uint32_t)c0;
  const uint32_t p1 = c0 + a.chunk < a.n ? (uint32_t)(c0 + a.chunk) : a.n;
  for (uint32_t base = p0; base < p1; base += 64) {
    const uint32_t p = base + lane;
    const bool valid = p < p1;
    T c[D];
#pragma unroll
    for (int j = 0; j < D; j++) c[j] = valid ? x[(size_t)j * a.n + p] : (T)0;
#pragma unroll
    for (uint32_t t = 0; t < BR_QW; t++) {
      if (t >= qw) continue;
      if (std::is_same<T, float>::value) {
        const bool near = valid && br_dist_sq<D, T>(c, qc[t]) <= fb[t];
        if (__ballot(near) == 0) continue;
        col[t].offer(near, br_key<D, T>(c, qc[t]), p, lane);
        fb[t] = (T)(col[t].tk * (1.0 + 0x1p-16));  // +inf until k1 entries are held
      } else {
        col[t].offer(valid, br_key<D, T>(c, qc[t]), p, lane);
      }
    }
  }
#pragma unroll
  for (uint32_t t = 0; t < BR_QW; t++) {
    if (t >= qw) continue;
    col[t].merge(lane);
    if (a.nchunks == 1) {
      br_write_row<T, D>(a, col[t], qid[t], lane);
    } else {
      const size_t off = ((size_t)(q0 + t) * a.nchunks + chunk) * a.k1;
      for (uint32_t e = lane; e < a.k1; e += 64) {
        const bool have = e < col[t].nb;
        a.partKey[off + e] = have ? col[t].best_key(e) : __builtin_inf();
        a.partId[off + e] = have ? col[t].best_id(e) : BR_NONE;
      }
    }
  }
}

template <typename T, int D>
__global__ __launch_bounds__(64) void k_brute_merge(BruteArgs a) {
  extern __shared__ __align__(16) unsigned char br_lds[];
  const uint32_t lane = threadIdx.x;
  const uint32_t qi = blockIdx.x;
  if (qi >= a.nq) return;
  const uint32_t q = a.queries ? a.queries[qi] : qi;
  Collector col;
  col.init(br_lds, a.k1);
  const size_t base = (size_t)qi * a.nchunks * a.k1;
  const size_t total = (size_t)a.nchunks * a.k1;
  for (size_t e0 = 0; e0 < total; e0 += 64) {
    const size_t e = e0 + lane;
    const bool valid = e < total;
    const double k = valid ? a.partKey[base + e] : __builtin_inf();
    const uint32_t id = valid ? a.partId[base + e] : BR_NONE;
    col.offer(valid && id != BR_NONE, k, id, lane);
  }
  col.merge(lane);
  br_write_row<T, D>(a, col, q, lane);
}

size_t br_lds_bytes(uint32_t k1, uint32_t qw) {
  return ((size_t)k1 * 2 * 12 + (size_t)BR_BUF * 12) * qw;
}

template <typename T, int D>
void launch_brute(const BruteArgs &a, hipStream_t s) {
  const uint32_t tiles = (a.nq + a.qw - 1) / a.qw;
  const uint64_t blocks = (uint64_t)tiles * a.nchunks;
  k_brute_partial<T, D><<<(uint32_t)blocks, 64, br_lds_bytes(a.k1, a.qw), s>>>(a);
  if (a.nchunks > 1) k_brute_merge<T, D><<<a.nq, 64, br_lds_bytes(a.k1, 1), s>>>(a);
}

__global__ void k_sel_flags(const uint8_t *sel, uint32_t n, uint32_t *flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = sel[i] != 0 ? 1u : 0u;
}

__global__ void k_sel_compact(const uint32_t *flags, const uint32_t *pos, uint32_t n, uint32_t *list) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) list[pos[i]] = i;
}

}  // namespace

void select_list(const uint8_t *sel, uint32_t n, uint32_t *flags, uint32_t *pos, uint32_t *list,
                 uint32_t *scratch, hipStream_t s) {
  if (n == 0) return;
  const uint32_t g = (n + 255) / 256;
  k_sel_flags<<<g, 256, 0, s>>>(sel, n, flags);
  scan_exclusive_u32(flags, pos, n, scratch, s);
  k_sel_compact<<<g, 256, 0, s>>>(flags, pos, n, list);
}

BrutePlan brute_plan(uint32_t n, uint32_t nq, uint32_t k1, uint64_t maxPartialEntries) {
  BrutePlan p{};
  if (k1 == 0 || k1 > BRUTE_MAX_K1 || n == 0) return p;
  // queries per wave: as many as the 64 KB LDS budget allows, up to 4
  uint32_t qw = BR_QW;
  while (qw > 1 && br_lds_bytes(k1, qw) > 65536) qw--;
  p.qw = qw;
  const uint32_t tiles = (nq + qw - 1) / qw;
  // enough waves to fill 256 CUs several times over, chunks of at least 16 (K+1) points so
  // the partial lists stay a small fraction of the points scanned
  const uint64_t wantBlocks = 8192;
  uint64_t nchunks = tiles >= wantBlocks ? 1 : (wantBlocks + tiles - 1) / tiles;
  uint64_t minChunk = (uint64_t)k1 * 16 > 1024 ? (uint64_t)k1 * 16 : 1024;
  uint64_t maxChunks = (n + minChunk - 1) / minChunk;
  if (nchunks > maxChunks) nchunks = maxChunks;
  if (nchunks == 0) nchunks = 1;
  // partial tables within maxPartialEntries entries (12 bytes each)
  const uint64_t memChunks = (maxPartialEntries / ((uint64_t)nq * k1)) > 0 ? (maxPartialEntries / ((uint64_t)nq * k1)) : 1;
  if (nchunks > memChunks) nchunks = memChunks;
  uint64_t chunk = (n + nchunks - 1) / nchunks;
  chunk = (chunk + 63) & ~63ull;
  nchunks = (n + chunk - 1) / chunk;
  if (tiles * nchunks > 0x7fffffffull) return BrutePlan{};
  p.chunk = (uint32_t)chunk;
  p.nchunks = (uint32_t)nchunks;
  p.partialEntries = nchunks > 1 ? (uint64_t)nq * nchunks * k1 : 0;
  return p;
}

template <typename T>
void brute_join(const BruteArgs &a, hipStream_t s) {
  if (a.nq == 0) return;
  switch (a.d) {
    case 1: launch_brute<T, 1>(a, s); break;
    case 2: launch_brute<T, 2>(a, s); break;
    case 3: launch_brute<T, 3>(a, s); break;
    case 4: launch_brute<T, 4>(a, s); break;
    case 5: launch_brute<T, 5>(a, s); break;
    case 6: launch_brute<T, 6>(a, s); break;
    case 7: launch_brute<T, 7>(a, s); break;
    case 8: launch_brute<T, 8>(a, s); break;
    case 9: launch_brute<T, 9>(a, s); break;
    case 10: launch_brute<T, 10>(a, s); break;
    case 11: launch_brute<T, 11>(a, s); break;
    case 12: launch_brute<T, 12>(a, s); break;
    case 13: launch_brute<T, 13>(a, s); break;
    case 14: launch_brute<T, 14>(a, s); break;
    case 15: launch_brute<T, 15>(a, s); break;
    case 16: launch_brute<T, 16>(a, s); break;
    default: break;
  }
}

template void brute_join<float>(const BruteArgs &, hipStream_t);
template void brute_join<double>(const BruteArgs &, hipStream_t);

}  // namespace knn
