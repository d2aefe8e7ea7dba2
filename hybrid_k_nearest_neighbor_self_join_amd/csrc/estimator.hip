// estimator.hip -- dimension reorder by variance and the BETA-driven epsilon estimate.
//
//  * variance_rows: main.cu:2521-2615 in T arithmetic with sequential sums over rows
//    0,100,200,...  The order of those float additions decides the variances' last bits, so
//    each dim is summed by one thread, in row order.
//  * avg_dist_soa: kernel.cu:343-381.  The reference adds every sampled distance to one
//    double with atomicAdd (order not reproducible); here each thread sums its own pairs
//    in order and the per-thread sums are added in thread order by one lane.
//  * kdist_hist_soa: kernel.cu:450-523 with a per-block LDS histogram (10000 x u32) merged
//    into the global one by integer atomics (order-free).  Points past N are skipped where
//    the reference reads beyond the array (kernel.cu:486-492); bucket == 10000 is dropped
//    where the reference writes one past the histogram.
#include <algorithm>

#include "knn_launch.h"

namespace knn {

namespace {

// One block per dim.  Waves 1..3 stage the sampled values (rows 0,100,...) of the next chunk
// in LDS while lane 0 of wave 0 adds the current chunk in row order, so the sequential sum
// (whose rounding decides the dimension order) runs at LDS rather than HBM latency.
constexpr uint32_t VAR_CHUNK = 2048;
template <typename T>
__device__ __forceinline__ void stage_chunk(const T *__restrict__ pts, uint32_t d, uint32_t i, uint64_t base,
                                            uint32_t cnt, T *buf) {
  if (threadIdx.x < 64) return;
  for (uint32_t t = threadIdx.x - 64; t < cnt; t += blockDim.x - 64)
    buf[t] = pts[(base + t) * 100 * (uint64_t)d + i];
}
template <typename T, bool SQ>
__device__ T sequential_pass(const T *__restrict__ pts, uint32_t n, uint32_t d, uint32_t i, T avg,
                             T (*buf)[VAR_CHUNK]) {
  const uint64_t m = ((uint64_t)n + 99) / 100;
  T sum = 0;
  uint32_t cur = 0;
  if (m) stage_chunk<T>(pts, d, i, 0, (uint32_t)(m < VAR_CHUNK ? m : VAR_CHUNK), buf[0]);
  __syncthreads();
  for (uint64_t base = 0; base < m; base += VAR_CHUNK) {
    const uint32_t cnt = (uint32_t)(m - base < VAR_CHUNK ? m - base : VAR_CHUNK);
    const uint64_t nb = base + VAR_CHUNK;
    if (nb < m) stage_chunk<T>(pts, d, i, nb, (uint32_t)(m - nb < VAR_CHUNK ? m - nb : VAR_CHUNK), buf[cur ^ 1]);
    if (threadIdx.x == 0) {
      const T *b = buf[cur];
      // 8 LDS reads issued ahead of 8 dependent adds (same order of additions)
      uint32_t t = 0;
      for (; t + 8 <= cnt; t += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = b[t + u];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          if (SQ) {
            T df = v[u] - avg;
            T sq = df * df;
            sum += sq;
          } else {
            sum += v[u];
          }
        }
      }
      for (; t < cnt; t++) {
        if (SQ) {
          T df = b[t] - avg;
          T sq = df * df;
          sum += sq;
        } else {
          sum += b[t];
        }
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  return sum;
}

template <typename T>
__global__ __launch_bounds__(256) void k_variance(const T *__restrict__ pts, uint32_t n, uint32_t d, T *var) {
  __shared__ T buf[2][VAR_CHUNK];
  __shared__ T shAvg;
  const uint32_t i = blockIdx.x;
  const T inv = (T)(1.0 / (100 * 1.0));
  const T denom = (T)n * inv;
  const T sum = sequential_pass<T, false>(pts, n, d, i, (T)0, buf);
  if (threadIdx.x == 0) shAvg = sum / denom;
  __syncthreads();
  const T avg = shAvg;
  const T sum2 = sequential_pass<T, true>(pts, n, d, i, avg, buf);
  if (threadIdx.x == 0) var[i] = sum2 / denom;
}

// Moments of the sampled rows (0, 100, ...) per dim in FP64, for the error intervals of the
// sequential T-arithmetic variance: mom[4*i + {0,1,2,3}] = {sum x, sum |x|, sum (x-a)^2,
// sum (x-a)} with a = sum x / (n/100) taken from a first launch (second launch, phase 1).
template <typename T>
__global__ __launch_bounds__(256) void k_var_moments(const T *__restrict__ pts, uint32_t n, uint32_t d,
                                                     int phase, double *mom) {
  const uint32_t i = blockIdx.y;
  const uint64_t m = ((uint64_t)n + 99) / 100;
  double a0 = 0.0, a1 = 0.0;
  const double avg = phase ? mom[4 * i] / ((double)n / 100.0) : 0.0;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m; r += (uint64_t)gridDim.x * blockDim.x) {
    const double x = (double)pts[r * 100 * (uint64_t)d + i];
    if (phase == 0) {
      a0 += x;
      a1 += fabs(x);
    } else {
      const double y = x - avg;
      a0 += y * y;
      a1 += y;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, 64);
    a1 += __shfl_xor(a1, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&mom[4 * i + 2 * phase], a0);
    atomicAdd(&mom[4 * i + 2 * phase + 1], a1);
  }
}

struct Perm {
  uint32_t p[KNN_MAXDIM];
};

template <typename T>
__global__ void k_permute_to_soa(const T *__restrict__ pts, uint32_t n, uint32_t d, Perm perm,
                                 T *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T *row = pts + (size_t)i * d;
  for (uint32_t j = 0; j < d; j++) out[(size_t)j * n + i] = row[perm.p[j]];
}

// kernel.cu:343-381: sample point tid = g*1000 (g < G = ceil(0.001N/256)*256) against points
// i = g%256 + 1000k, per pair sqrt(sum (double)(T)(diff*diff)).  The reference adds with
// double atomics (no reproducible order); here the sum has a fixed order over a fixed
// partition, spread over many threads.  With Kmax = ceil(N/1000) pairs for the longest
// thread, S = min(ceil(Kmax/64), max(1, floor(2^26 / G))) segments per thread of
// L = ceil(Kmax/S) pairs (L = 64 up to about 50M points; the cap bounds the scratch):
//   seg(g, s)  = sum of pairs k in [L s, L s + L) in k order
//   part(g)    = sum over s of seg(g, s) in s order
//   chunk(c)   = sum over g in [256c, 256c + 256) of part(g) in g order
//   total      = sum over c of chunk(c) in c order
struct AvgLayout {
  uint32_t G, S, L;
};

template <typename T>
__global__ __launch_bounds__(256) void k_avg_seg(const T *__restrict__ x, uint32_t n, uint32_t d,
                                                 uint32_t G, uint32_t S, uint32_t L, double *seg) {
  const uint64_t item = (uint64_t)blockIdx.x * 256 + threadIdx.x;  // item = s * G + g
  if (item >= (uint64_t)G * S) return;
  const uint32_t g = (uint32_t)(item % G), sg = (uint32_t)(item / G);
  const uint64_t tid = (uint64_t)g * 1000;
  double sum = 0;
  if (tid < n) {
    T q[KNN_MAXDIM];
#pragma unroll
    for (uint32_t j = 0; j < KNN_MAXDIM; j++) q[j] = j < d ? x[(size_t)j * n + tid] : (T)0;
    const uint64_t i0 = (uint64_t)(g % 256) + 1000ull * L * sg;
    for (uint32_t k = 0; k < L; k++) {
      const uint64_t i = i0 + 1000ull * k;
      if (i >= n) break;
      double run = 0;
#pragma unroll
      for (uint32_t j = 0; j < KNN_MAXDIM; j++) {
        if (j < d) {
          T df = x[(size_t)j * n + i] - q[j];
          T sq = df * df;
          run += (double)sq;
        }
      }
      sum += sqrt(run);
    }
  }
  seg[item] = sum;
}

__global__ void k_avg_part(const double *seg, uint32_t G, uint32_t S, double *part) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  double t = 0;
  for (uint32_t sg = 0; sg < S; sg++) t += seg[(size_t)sg * G + g];
  part[g] = t;
}

__global__ void k_avg_chunks(const double *part, uint32_t G, double *chunk) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G / 256) return;
  double t = 0;
  for (uint32_t g = 256 * c; g < 256 * c + 256; g++) t += part[g];
  chunk[c] = t;
}

__global__ void k_avg_total(const double *chunk, uint32_t C, double *total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t = 0;
  for (uint32_t c = 0; c < C; c++) t += chunk[c];
  *total = t;
}

constexpr uint32_t NBUCKETS = 10000;

// kernel.cu:450-523: sample s = blockIdx.x / KD_PARTS (point s * offset) against the first
// 256 points of every 2048, +8 per hit in bucket floor(d / bw) for d < avg (self excluded).
// The counts are integers, so the partition is free: each sample's 2048-point strides are
// dealt to KD_PARTS blocks, and a thread loads KD_UNROLL strides before it bins any.  Built
// per dimensionality D (the coordinates in registers, no per-dim predication); a pair whose
// squared distance is at or above avg^2 (1 + 2^-40) has a rounded root >= avg and skips the
// root and the division.
constexpr uint32_t KD_PARTS = 2, KD_UNROLL = 4;
template <typename T, int D>
__global__ __launch_bounds__(256) void k_kdist(const T *__restrict__ x, uint32_t n, uint32_t offset,
                                               double avg, double bw, uint32_t *hist) {
  __shared__ uint32_t lh[NBUCKETS];
  for (uint32_t b = threadIdx.x; b < NBUCKETS; b += 256) lh[b] = 0;
  __syncthreads();
  const uint64_t tid = (uint64_t)(blockIdx.x / KD_PARTS) * offset;
  const uint32_t part = blockIdx.x % KD_PARTS;
  const double cut2 = avg * avg * (1.0 + 0x1p-40);
  if (tid < n) {
    T q[D];
#pragma unroll
    for (int j = 0; j < D; j++) q[j] = x[(size_t)j * n + tid];
    const uint64_t step = 2048ull * KD_PARTS;
    for (uint64_t i0 = 2048ull * part; i0 < n; i0 += step * KD_UNROLL) {
      T c[KD_UNROLL][D];
      bool ok[KD_UNROLL];
#pragma unroll
      for (uint32_t u = 0; u < KD_UNROLL; u++) {
        const uint64_t p = i0 + step * u + threadIdx.x;
        ok[u] = p < n;
        const uint64_t pc = ok[u] ? p : (uint64_t)n - 1;
#pragma unroll
        for (int j = 0; j < D; j++) c[u][j] = x[(size_t)j * n + pc];
      }
#pragma unroll
      for (uint32_t u = 0; u < KD_UNROLL; u++) {
        const uint64_t p = i0 + step * u + threadIdx.x;
        double run = 0;
#pragma unroll
        for (int j = 0; j < D; j++) {
          T df = c[u][j] - q[j];
          T sq = df * df;
          run += (double)sq;
        }
        if (!ok[u] || !(run < cut2) || p == tid) continue;
        run = sqrt(run);
        if (run < avg) {
          uint32_t bucket = (uint32_t)(run / bw);
          if (bucket < NBUCKETS) atomicAdd(&lh[bucket], 8u);
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < NBUCKETS; b += 256)
    if (lh[b]) atomicAdd(&hist[b], lh[b]);
}

}  // namespace

template <typename T>
void variance_rows(const T *pts, uint32_t n, uint32_t d, T *var, hipStream_t s) {
  k_variance<T><<<d, 256, 0, s>>>(pts, n, d, var);
}

template <typename T>
void variance_moments(const T *pts, uint32_t n, uint32_t d, double *mom, hipStream_t s) {
  fill_u32(reinterpret_cast<uint32_t *>(mom), 0u, (size_t)8 * d, s);
  const uint64_t m = ((uint64_t)n + 99) / 100;
  uint32_t bx = (uint32_t)((m + 255) / 256);
  bx = bx < 64 ? (bx ? bx : 1) : 64;
  k_var_moments<T><<<dim3(bx, d), 256, 0, s>>>(pts, n, d, 0, mom);
  k_var_moments<T><<<dim3(bx, d), 256, 0, s>>>(pts, n, d, 1, mom);
}

template <typename T>
void permute_to_soa(const T *pts, uint32_t n, uint32_t d, const uint32_t *perm_host, T *out,
                    hipStream_t s) {
  Perm p;
  for (uint32_t j = 0; j < KNN_MAXDIM; j++) p.p[j] = j < d ? perm_host[j] : 0;
  if (n) k_permute_to_soa<T><<<(n + 255) / 256, 256, 0, s>>>(pts, n, d, p, out);
}

size_t avgdist_threads(uint32_t n) {
  int total_blocks = (int)ceil((0.001 * n) / (1.0 * 256));  // GPU.cu:3276
  if (total_blocks < 1) total_blocks = 1;
  return (size_t)total_blocks * 256;
}

static AvgLayout avg_layout(uint32_t n) {
  AvgLayout a;
  a.G = (uint32_t)avgdist_threads(n);
  const uint64_t kmax = ((uint64_t)n + 999) / 1000;
  uint64_t S = (kmax + 63) / 64;
  const uint64_t cap = std::max<uint64_t>(1, (1ull << 26) / a.G);
  if (S > cap) S = cap;
  if (S == 0) S = 1;
  a.S = (uint32_t)S;
  a.L = (uint32_t)((kmax + S - 1) / S);
  if (a.L == 0) a.L = 1;
  return a;
}

size_t avgdist_scratch_doubles(uint32_t n) {
  const AvgLayout a = avg_layout(n);
  return (size_t)a.G * a.S + a.G + a.G / 256 + 8;
}

template <typename T>
void avg_dist_soa(const T *x, uint32_t n, uint32_t d, double *scratch, double *total, hipStream_t s) {
  const AvgLayout a = avg_layout(n);
  const uint32_t G = a.G, S = a.S;
  double *seg = scratch, *part = seg + (size_t)G * S, *chunk = part + G;
  const uint64_t items = (uint64_t)G * S;
  k_avg_seg<T><<<(uint32_t)((items + 255) / 256), 256, 0, s>>>(x, n, d, G, S, a.L, seg);
  k_avg_part<<<(G + 255) / 256, 256, 0, s>>>(seg, G, S, part);
  k_avg_chunks<<<(G / 256 + 255) / 256, 256, 0, s>>>(part, G, chunk);
  k_avg_total<<<1, 64, 0, s>>>(chunk, G / 256, total);
}

uint64_t avgdist_pairs(uint32_t n) {
  const uint64_t G = avgdist_threads(n);
  uint64_t c = 0;
  for (uint64_t g = 0; g < G && g * 1000 < n; g++) c += ((uint64_t)n - g % 256 + 999) / 1000;
  return c;
}

template <typename T>
void kdist_hist_soa(const T *x, uint32_t n, uint32_t d, uint32_t offset, uint32_t nsample,
                    double avg, double bw, uint32_t *hist, hipStream_t s) {
  fill_u32(hist, 0u, NBUCKETS, s);
  if (!nsample) return;
  const uint32_t grid = nsample * KD_PARTS;
  switch (d) {
#define KD_CASE(DD) \
    case DD: k_kdist<T, DD><<<grid, 256, 0, s>>>(x, n, offset, avg, bw, hist); break;
    KD_CASE(1) KD_CASE(2) KD_CASE(3) KD_CASE(4) KD_CASE(5) KD_CASE(6) KD_CASE(7) KD_CASE(8)
    KD_CASE(9) KD_CASE(10) KD_CASE(11) KD_CASE(12) KD_CASE(13) KD_CASE(14) KD_CASE(15) KD_CASE(16)
#undef KD_CASE
    default: break;
  }
}

#define INST(T)                                                                            \
  template void variance_rows<T>(const T *, uint32_t, uint32_t, T *, hipStream_t);          \
  template void variance_moments<T>(const T *, uint32_t, uint32_t, double *, hipStream_t);  \
  template void permute_to_soa<T>(const T *, uint32_t, uint32_t, const uint32_t *, T *,     \
                                  hipStream_t);                                             \
  template void avg_dist_soa<T>(const T *, uint32_t, uint32_t, double *, double *, hipStream_t); \
  template void kdist_hist_soa<T>(const T *, uint32_t, uint32_t, uint32_t, uint32_t, double, \
                                  double, uint32_t *, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace knn
