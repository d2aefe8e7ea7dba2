// knn_device.h -- shared device helpers of the gfx950 KNN self-join.
//
// Numerics (all device code is built with -ffp-contract=off and correctly rounded FP32
// divide/sqrt, so each expression below is one IEEE operation):
//   cell coordinate  (unsigned)((x - min_j) / eps)                    kernel.cu:1685-1687
//   linear cell id   sum_j c_j * prod_{i<j} nCells_i, dim 0 fastest   kernel.cu:31-47
//   FP64 key         sum_j ((double)c_j - (double)q_j)^2              ranks neighbours
//   FP32 distance    sqrtf(sum_j (c_j - q_j)^2)                       kernel.cu:127-144
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KNN_MAXDIM 16
#define KNN_MAX_INDEXED 8
#define KNN_WAVE 64
// learned-radius MLP 12 -> 64 -> 64 -> 1
#define MLP_IN 12
#define MLP_H 64
#define MLP_NW (MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1)

namespace knn {

struct GridParams {
  float minArr[KNN_MAX_INDEXED];
  uint32_t nCells[KNN_MAX_INDEXED];
  uint64_t mult[KNN_MAX_INDEXED];  // prod_{i<j} nCells_i
  uint64_t keyBase;                 // added to every linear id (one key range per set)
  float eps;
  uint32_t nd;                      // indexed dims
  // slab order (sb > 0, single-set grids with fewer indexed than data dims): within a cell the
  // points are sorted by slab_q of dim nd, the first non-indexed dim, quantised to sb bits
  float smin, sscale;
  uint32_t sb;
};

struct GridParamsD {
  double minArr[KNN_MAX_INDEXED];
  uint32_t nCells[KNN_MAX_INDEXED];
  uint64_t mult[KNN_MAX_INDEXED];
  uint64_t keyBase;
  double eps;
  uint32_t nd;
  double smin, sscale;
  uint32_t sb;
};

// One point set of a batched join: its grid, its points [pStart, pStart + n) of the group
// arrays, its K+1 and the first element of its rows in the group's output table.
struct SetDesc {
  GridParams g;
  uint32_t pStart, n;
  uint32_t k1, nd;
  uint64_t outOff;
  float r0;          // initial search radius
  uint32_t pad;
};

template <typename T> struct GridOf;
template <> struct GridOf<float> { using type = GridParams; };
template <> struct GridOf<double> { using type = GridParamsD; };

// kernel.cu:1685-1687 with clamping to [0, ncell-1]; the clamp never binds for data points
// (min = datamin - eps, nCells = ceil((max-min)/eps)), only for query range bounds.
template <typename T>
__device__ __forceinline__ uint32_t cell_coord(T x, T mn, T eps, uint32_t ncell) {
  T t = (x - mn) / eps;
  if (!(t > (T)0)) return 0u;
  if (t >= (T)4294967295.0) return ncell - 1;
  uint32_t c = (uint32_t)t;
  return c >= ncell ? ncell - 1 : c;
}

// Slab code of a coordinate of the first non-indexed dim: floor((x - smin) * sscale) clamped to
// [0, 2^sb - 1], in T arithmetic; non-decreasing in x, so the points of a cell whose codes lie
// in [code(a), code(b)] include every point with a <= x <= b.
template <typename T, typename G>
__device__ __forceinline__ uint32_t slab_q(T x, const G &g) {
  const T t = (x - (T)g.smin) * (T)g.sscale;
  const uint32_t top = (1u << g.sb) - 1u;
  if (!(t > (T)0)) return 0u;
  if (t >= (T)top) return top;
  return (uint32_t)t;
}

// IEEE successor / predecessor of a finite value (nextafter toward +-inf); used to round the
// query's [x-eps, x+eps] outward before mapping it to cell coordinates.
__device__ __forceinline__ float next_up(float x) {
  if (x != x || x == __builtin_inff()) return x;
  if (x == 0.0f) return __int_as_float(1);
  int b = __float_as_int(x);
  return __int_as_float(x > 0.0f ? b + 1 : b - 1);
}
__device__ __forceinline__ float next_down(float x) { return -next_up(-x); }
__device__ __forceinline__ double next_up(double x) {
  if (x != x || x == __builtin_inf()) return x;
  if (x == 0.0) return __longlong_as_double(1ll);
  long long b = __double_as_longlong(x);
  return __longlong_as_double(x > 0.0 ? b + 1 : b - 1);
}
__device__ __forceinline__ double next_down(double x) { return -next_up(-x); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <typename T>
__device__ __forceinline__ T readlane(T v, int lane);

template <>
__device__ __forceinline__ float readlane<float>(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
template <>
__device__ __forceinline__ uint32_t readlane<uint32_t>(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
template <>
__device__ __forceinline__ double readlane<double>(double v, int lane) {
  uint64_t b = (uint64_t)__double_as_longlong(v);
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// A wave-uniform value moved to scalar registers (the first active lane's copy), so that it
// holds no vector register.
template <typename T>
__device__ __forceinline__ T readfirst(T v);
template <>
__device__ __forceinline__ float readfirst<float>(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
template <>
__device__ __forceinline__ uint32_t readfirst<uint32_t>(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
template <>
__device__ __forceinline__ double readfirst<double>(double v) {
  uint64_t b = (uint64_t)__double_as_longlong(v);
  uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

}  // namespace knn
