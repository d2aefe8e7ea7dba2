// outlier.hip -- outlier scores from a K-nearest-neighbour table (printOutlierScores,
// main.cu:952-1067), a consumer of the join's output.
//
// Per point i of a table with k1 entries per row:
//   dist_sum[i]  = sum over the row of the distances, in entry order, in FP32 (main.cu:966-969)
//   dist_rank[i] = position of i when the points are ordered by dist_sum descending
//   indegree[i]  = number of rows (its own included) in which i appears (main.cu:1011-1020)
//   indeg_rank[i]= position of i when the points are ordered by indegree ascending
// The reference orders with std::sort, which leaves the order of equal values unspecified;
// here equal values are ordered by point id (stable radix sorts over ids 0..N-1), so the ranks
// are deterministic and equal the reference's whenever the values are distinct.
#include "knn_launch.h"

namespace knn {

namespace {

// row sums in entry order; key = bits of the sum arranged so that an ascending unsigned sort
// orders the sums descending (non-negative floats: larger value, larger bits)
__global__ void k_outlier_rows(const int32_t *__restrict__ ids, const float *__restrict__ dists, uint32_t n,
                               uint32_t k1, float *distSum, uint32_t *key, uint32_t *val,
                               uint32_t *indeg) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *row = dists + (size_t)i * k1;
  float s = 0.0f;
  for (uint32_t e = 0; e < k1; e++) s += row[e];
  distSum[i] = s;
  const uint32_t b = __float_as_uint(s);
  // -0.0 and +0.0 compare equal; NaN sums (absent entries) sort last
  const uint32_t mag = (s == 0.0f) ? 0u : ((b & 0x80000000u) ? 0u : b);
  key[i] = (s != s) ? 0xffffffffu : 0x7fffffffu - (mag & 0x7fffffffu);
  val[i] = i;
  const int32_t *r = ids + (size_t)i * k1;
  for (uint32_t e = 0; e < k1; e++) {
    const int32_t id = r[e];
    if (id >= 0 && (uint32_t)id < n) atomicAdd(&indeg[id], 1u);
  }
}

__global__ void k_indeg_keys(const uint32_t *indeg, uint32_t n, uint32_t *key, uint32_t *val) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = indeg[i];
  val[i] = i;
}

// rank[val[p]] = p for the sorted order
__global__ void k_scatter_rank(const uint32_t *val, uint32_t n, uint32_t *rank) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  rank[val[p]] = p;
}

}  // namespace

void outlier_scores(const int32_t *ids, const float *dists, uint32_t n, uint32_t k1, float *distSum,
                    uint32_t *distRank, uint32_t *indeg, uint32_t *indegRank, uint32_t *keys,
                    uint32_t *vals, uint32_t *keysTmp, uint32_t *valsTmp, uint32_t *scratch, hipStream_t s) {
  if (!n) return;
  const uint32_t g = (n + 255) / 256;
  fill_u32(indeg, 0u, n, s);
  k_outlier_rows<<<g, 256, 0, s>>>(ids, dists, n, k1, distSum, keys, vals, indeg);
  radix_sort_pairs_u32(keys, vals, keysTmp, valsTmp, n, 32, scratch, s);
  k_scatter_rank<<<g, 256, 0, s>>>(vals, n, distRank);
  k_indeg_keys<<<g, 256, 0, s>>>(indeg, n, keys, vals);
  radix_sort_pairs_u32(keys, vals, keysTmp, valsTmp, n, 32, scratch, s);
  k_scatter_rank<<<g, 256, 0, s>>>(vals, n, indegRank);
}

}  // namespace knn
