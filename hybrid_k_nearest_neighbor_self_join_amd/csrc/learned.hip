// learned.hip -- learned per-query initial radius: features, MLP 12->64->64->1, training.
//
// The radius only changes the epsilon schedule of the join (which radius each query is first
// tried at); every query's row is still decided by the exact range kernel, and a query that
// fails at its radius goes through the expansion loop, so no result depends on the model.
//
// Features of a query (12), on the "feature grid" of width wf = (V * (K+1) / N)^(1/nd) over
// the data's bounding box (V = its volume in the indexed dims: a cell of the grid would hold
// K+1 points of a uniform set):
//   f0      log1p(count of the query's own cell)
//   f1..f9  log1p(counts of the 3x3 cells around it in indexed dims 0 and 1, dim 1 outer,
//           dim 0 inner, other indexed dims at the query's own coordinate; for nd == 1 the 9
//           cells c-4..c+4 of dim 0); cells outside the grid count 0
//   f10     log(K+1)
//   f11     log(N / occupied feature cells): the mean population of a non-empty cell, a
//           global clustering measure that, like every other feature, does not change when
//           the same distribution is sampled with more points
// Target: log(r_K1 / wf), r_K1 = distance to the (K+1)-th nearest point (self included).
// Radius: r = wf * exp(mlp(f) + margin), margin calibrated on the training residuals.
//
// MLP: h1 = relu(W1 f + b1), h2 = relu(W2 h1 + b2), y = w3 . h2 + b3.  Training, the
// gradient and the forward entry point run in FP32 with fmaf in index order (so the CPU
// oracle, using C's fmaf in the same order, reproduces them bitwise); the join's inference
// (k_radius_mfma) runs on the matrix cores with FP16 operands and FP32 accumulation.
// Weights, flat: W1[64][12], b1[64], W2[64][64], b2[64], w3[64], b3.
#include <algorithm>

#include "knn_launch.h"

namespace knn {

namespace {

__device__ __forceinline__ float relu(float v) { return v > 0.0f ? v : 0.0f; }

// quarter-octave bin of a radius relative to the feature-grid width: floor(4 log2(r / wf)) + 32,
// clamped to [0, 63]; zero, NaN and tiny radii fall in bin 0, huge ones in bin 63
template <typename T>
__device__ __forceinline__ int radius_bin(T r, float wf) {
  const float x = 4.0f * log2f((float)r / wf);
  if (!(x >= -32.0f)) return 0;
  if (x >= 31.0f) return 63;
  const int b = (int)floorf(x) + 32;
  return b < 0 ? 0 : (b > 63 ? 63 : b);
}

// Forward pass of one sample; W = flat weights (LDS).  Layer 1 unit j is computed and
// immediately folded into the 64 layer-2 accumulators, so no register array is indexed by a
// loop variable (the summation order per unit is unchanged: bias first, inputs ascending).
// H1/H2 (optional) receive the activations.
__device__ __forceinline__ float mlp_forward(const float *W, const float *f, float *H1, float *H2) {
  const float *W1 = W, *b1 = W + 768, *W2 = W + 832, *b2 = W + 4928, *w3 = W + 4992;
  float acc[MLP_H];
#pragma unroll
  for (int i = 0; i < MLP_H; i++) acc[i] = b2[i];
  for (int j = 0; j < MLP_H; j++) {
    float a = b1[j];
#pragma unroll
    for (int t = 0; t < MLP_IN; t++) a = fmaf(W1[j * MLP_IN + t], f[t], a);
    const float h = relu(a);
    if (H1) H1[j] = h;
#pragma unroll
    for (int i = 0; i < MLP_H; i++) acc[i] = fmaf(W2[i * MLP_H + j], h, acc[i]);
  }
  float y = W[5056];
#pragma unroll
  for (int i = 0; i < MLP_H; i++) {
    const float h = relu(acc[i]);
    if (H2) H2[i] = h;
    y = fmaf(w3[i], h, y);
  }
  return y;
}

// mlp_forward with W2 stored transposed (W2t[j][i] = W2[i][j]): the 64 weights a layer-1
// unit feeds are contiguous, so they load 4 at a time.  Same operations in the same order.
__device__ __forceinline__ float mlp_forward_t(const float *W, const float *f) {
  const float *W1 = W, *b1 = W + 768, *W2t = W + 832, *b2 = W + 4928, *w3 = W + 4992;
  float acc[MLP_H];
#pragma unroll
  for (int i = 0; i < MLP_H; i++) acc[i] = b2[i];
  for (int j = 0; j < MLP_H; j++) {
    float a = b1[j];
#pragma unroll
    for (int t = 0; t < MLP_IN; t++) a = fmaf(W1[j * MLP_IN + t], f[t], a);
    const float h = relu(a);
    const float4 *w = reinterpret_cast<const float4 *>(W2t + j * MLP_H);
#pragma unroll
    for (int i4 = 0; i4 < MLP_H / 4; i4++) {
      const float4 v = w[i4];
      acc[4 * i4 + 0] = fmaf(v.x, h, acc[4 * i4 + 0]);
      acc[4 * i4 + 1] = fmaf(v.y, h, acc[4 * i4 + 1]);
      acc[4 * i4 + 2] = fmaf(v.z, h, acc[4 * i4 + 2]);
      acc[4 * i4 + 3] = fmaf(v.w, h, acc[4 * i4 + 3]);
    }
  }
  float y = W[5056];
#pragma unroll
  for (int i = 0; i < MLP_H; i++) y = fmaf(w3[i], relu(acc[i]), y);
  return y;
}

template <typename T, typename G>
__global__ void k_feature_counts(const T *__restrict__ x, uint32_t n, G g, uint32_t *cnt, uint64_t cells,
                                 uint32_t reps) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t id = 0;
  for (uint32_t j = 0; j < g.nd; j++)
    id += (uint64_t)cell_coord<T>(x[(size_t)j * n + i], (T)g.minArr[j], (T)g.eps, g.nCells[j]) * g.mult[j];
  atomicAdd(&cnt[(uint64_t)(blockIdx.x % reps) * cells + id], 1u);
}

// cnt[c] = the sum of the reps replica grids' counts of cell c (integers: exact in any order)
__global__ void k_sum_reps(uint32_t *cnt, uint64_t cells, uint32_t reps) {
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  uint32_t t = cnt[c];
  for (uint32_t r = 1; r < reps; r++) t += cnt[(uint64_t)r * cells + c];
  cnt[c] = t;
}

template <typename T, typename G>
__device__ __forceinline__ void features_of(const T *__restrict__ x, uint32_t n, uint32_t q,
                                            const G &g, const uint32_t *__restrict__ cnt,
                                            float logk1, float logdens, float *f) {
  int64_t c[KNN_MAX_INDEXED];
  uint64_t own = 0;
  for (uint32_t j = 0; j < g.nd; j++) {
    c[j] = cell_coord<T>(x[(size_t)j * n + q], (T)g.minArr[j], (T)g.eps, g.nCells[j]);
    own += (uint64_t)c[j] * g.mult[j];
  }
  f[0] = log1pf((float)cnt[own]);
  int k = 1;
  for (int dy = -1; dy <= 1; dy++) {
    for (int dx = -1; dx <= 1; dx++) {
      int64_t c0, c1 = 0;
      if (g.nd == 1) {
        c0 = c[0] + (dy + 1) * 3 + dx - 4;  // c-4 .. c+4
      } else {
        c0 = c[0] + dx;
        c1 = c[1] + dy;
      }
      float v = 0.0f;
      const bool in0 = c0 >= 0 && c0 < (int64_t)g.nCells[0];
      const bool in1 = g.nd == 1 || (c1 >= 0 && c1 < (int64_t)g.nCells[1]);
      if (in0 && in1) {
        uint64_t id = own - (uint64_t)c[0] * g.mult[0] + (uint64_t)c0 * g.mult[0];
        if (g.nd > 1) id = id - (uint64_t)c[1] * g.mult[1] + (uint64_t)c1 * g.mult[1];
        v = log1pf((float)cnt[id]);
      }
      f[k++] = v;
    }
  }
  f[10] = logk1;
  f[11] = logdens;
}

// features + MLP + radius for every query; optional feature output (training) and a
// 64-bin histogram of 4*log2(r/wf) + 32 (grid width choice)
template <typename T, typename G>
__global__ __launch_bounds__(256) void k_radius(const T *__restrict__ x, uint32_t n, G g,
                                                const uint32_t *__restrict__ cnt, float logk1,
                                                float logdens, const float *__restrict__ Wg,
                                                float margin, float wf, T *radius,
                                                float *featOut, uint32_t *rhist) {
  __shared__ __align__(16) float W[MLP_NW + 3];
  __shared__ uint32_t lh[64];
  // W2 (64 x 64 at 832) stored transposed for mlp_forward_t
  for (uint32_t i = threadIdx.x; i < MLP_NW; i += blockDim.x) {
    uint32_t dst = i;
    if (i >= 832 && i < 4928) {
      const uint32_t r = (i - 832) / MLP_H, c = (i - 832) % MLP_H;
      dst = 832 + c * MLP_H + r;
    }
    W[dst] = Wg ? Wg[i] : 0.0f;
  }
  if (threadIdx.x < 64) lh[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) {
    float f[MLP_IN];
    features_of<T, G>(x, n, q, g, cnt, logk1, logdens, f);
    if (featOut)
      for (int j = 0; j < MLP_IN; j++) featOut[(size_t)q * MLP_IN + j] = f[j];
    if (radius) {
      const float y = mlp_forward_t(W, f);
      // r = wf e^(y + margin), the exponent kept within [-8, 14] (about wf/3000 .. 1.2e6 wf)
      // so that any weights give a finite radius the expansion loop can start from
      float e = y + margin;
      if (!(e == e)) e = 0.0f;
      e = fminf(fmaxf(e, -8.0f), 14.0f);
      const float r = wf * expf(e);
      radius[q] = (T)r;
      if (rhist) atomicAdd(&lh[radius_bin<float>(r, wf)], 1u);
    }
  }
  __syncthreads();
  if (rhist && threadIdx.x < 64 && lh[threadIdx.x]) atomicAdd(&rhist[threadIdx.x], lh[threadIdx.x]);
}

// The join's radius inference on the matrix cores: the MLP in FP16 operands with FP32
// accumulation (v_mfma_f32_16x16x16f16), in the transposed form H^T = W X^T so that a layer's
// output fragment (lane: column = query lane%16, rows 4*(lane/16)+i of a 16-row tile) is the
// next layer's B fragment (rows = k) as it stands.  Per group of 16 queries: 4 MFMAs for
// layer 1 (12 features padded to 16), 16 for layer 2, then layer 3 as a 16-term dot product
// per lane and two lane exchanges.  The radius only steers the schedule (rows are exact for
// any radius), so FP16 operands (y within about 1e-2 of the FP32 forward) are enough; the
// features are computed in FP32 exactly as k_radius does.  Persistent grid; each wave loads its
// 20 weight fragments once.
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename T, typename G>
__global__ __launch_bounds__(256, 4) void k_radius_mfma(const T *__restrict__ x, uint32_t n, G g,
                                                     const uint32_t *__restrict__ cnt, float logk1,
                                                     float logdens, const float *__restrict__ Wg,
                                                     float margin, float wf, T *radius, uint32_t *rhist) {
  __shared__ _Float16 fs[4][64][16];
  __shared__ uint32_t lh[64];
  // the hidden biases and the output weights (64 each), read from LDS where they are used: in
  // registers they took 48 VGPRs a lane and held the kernel to 3 waves per SIMD
  __shared__ __align__(16) float sb1[64], sb2[64], sw3[64];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t r16 = lane & 15, kq = lane >> 4;
  const float *W1 = Wg, *b1 = Wg + 768, *W2 = Wg + 832, *b2 = Wg + 4928, *w3 = Wg + 4992;
  half4_t a1[4], a2[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint32_t k = 4 * kq + i;
      a1[mt][i] = (_Float16)(k < (uint32_t)MLP_IN ? W1[(16 * mt + r16) * MLP_IN + k] : 0.0f);
#pragma unroll
      for (int t = 0; t < 4; t++) a2[mt][t][i] = (_Float16)W2[(16 * mt + r16) * MLP_H + 16 * t + 4 * kq + i];
    }
  }
  const float b3 = Wg[5056];
  if (threadIdx.x < 64) {
    lh[threadIdx.x] = 0;
    sb1[threadIdx.x] = b1[threadIdx.x];
    sb2[threadIdx.x] = b2[threadIdx.x];
    sw3[threadIdx.x] = w3[threadIdx.x];
  }
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t base = (uint64_t)blockIdx.x * 256 + wv * 64; base < n; base += stride) {
    const uint64_t q = base + lane;
    const bool valid = q < n;
    float f[MLP_IN];
    if (valid) {
      features_of<T, G>(x, n, (uint32_t)q, g, cnt, logk1, logdens, f);
    } else {
#pragma unroll
      for (int j = 0; j < MLP_IN; j++) f[j] = 0.0f;
    }
#pragma unroll
    for (int j = 0; j < 16; j++) fs[wv][lane][j] = (_Float16)(j < MLP_IN ? f[j] : 0.0f);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    float yMine = 0.0f;
#pragma unroll
    for (uint32_t b = 0; b < 4; b++) {
      half4_t bx;
#pragma unroll
      for (int i = 0; i < 4; i++) bx[i] = fs[wv][16 * b + r16][4 * kq + i];
      half4_t h1[4];
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a1[mt], bx, acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) h1[mt][i] = (_Float16)relu(acc[i] + sb1[16 * mt + 4 * kq + i]);
      }
      float part = 0.0f;
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 4; t++) acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a2[mt][t], h1[t], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) part = fmaf(sw3[16 * mt + 4 * kq + i], relu(acc[i] + sb2[16 * mt + 4 * kq + i]), part);
      }
      part += __shfl_xor(part, 16, 64);
      part += __shfl_xor(part, 32, 64);
      if (kq == b) yMine = part + b3;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();  // fs is rewritten by the next round
    if (valid) {
      float e = yMine + margin;
      if (!(e == e)) e = 0.0f;
      e = fminf(fmaxf(e, -8.0f), 14.0f);
      const float r = wf * expf(e);
      radius[q] = (T)r;
      if (rhist) atomicAdd(&lh[radius_bin<float>(r, wf)], 1u);
    }
  }
  __syncthreads();
  if (rhist && threadIdx.x < 64 && lh[threadIdx.x]) atomicAdd(&rhist[threadIdx.x], lh[threadIdx.x]);
}

// 64-bin histogram of the pending queries' radius bins
template <typename T>
__global__ __launch_bounds__(256) void k_radius_bins(const T *radius, const uint8_t *pending, uint32_t n,
                                                     float wf, uint32_t *hist) {
  __shared__ uint32_t lh[64];
  if (threadIdx.x < 64) lh[threadIdx.x] = 0;
  __syncthreads();
  // grid-stride (a bounded number of blocks, so few global atomics per bin); within a wave,
  // one LDS add per distinct bin (a few), not one per query
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t rounds = (uint32_t)(((uint64_t)n + stride - 1) / stride);
  for (uint32_t it = 0; it < rounds; it++) {
    const uint64_t q = (uint64_t)it * stride + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = (q < n && pending[q]) ? radius_bin<T>(radius[q], wf) : -1;
    bool mine = b >= 0;
    for (;;) {
      const uint64_t live = __ballot(mine);
      if (!live) break;
      const int leader = __ffsll((long long)live) - 1;
      const int bb = __shfl(b, leader, 64);
      const uint64_t same = __ballot(mine && b == bb);
      if ((int)(threadIdx.x & 63) == leader) atomicAdd(&lh[bb], (uint32_t)__popcll(same));
      if (b == bb) mine = false;
    }
  }
  __syncthreads();
  if (threadIdx.x < 64 && lh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

// act[q] = pending with a radius bin in [lo, hi]
template <typename T>
__global__ void k_radius_class(const T *radius, const uint8_t *pending, uint32_t n, float wf, uint32_t lo,
                               uint32_t hi, uint8_t *act) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  bool a = false;
  if (pending[q]) {
    const uint32_t b = (uint32_t)radius_bin<T>(radius[q], wf);
    a = b >= lo && b <= hi;
  }
  act[q] = a ? 1 : 0;
}

// radius *= factor for the queries of the last pass (act) that failed it
template <typename T>
__global__ void k_grow_class(T *radius, const uint8_t *act, const uint8_t *failed, uint32_t n, T factor) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && act[q] && failed[q]) radius[q] = radius[q] * factor;
}

// number of non-zero counts (occupied feature cells) into *out
__global__ void k_count_nonzero(const uint32_t *cnt, uint64_t cells, unsigned long long *out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool nz = i < cells && cnt[i] != 0u;
  const uint64_t m = __ballot(nz);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(out, (unsigned long long)__popcll(m));
}

// radius *= factor for the queries that failed the last pass
template <typename T>
__global__ void k_grow_radius(T *radius, const uint8_t *failed, uint32_t n, T factor) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && failed[q]) radius[q] = radius[q] * factor;
}

// largest radius among the failed queries (bit pattern of a positive value orders as an
// unsigned integer)
__global__ void k_max_radius_f(const float *radius, const uint8_t *failed, uint32_t n, uint32_t *out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && failed[q]) atomicMax(out, __float_as_uint(radius[q]));
}
__global__ void k_max_radius_d(const double *radius, const uint8_t *failed, uint32_t n,
                               unsigned long long *out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && failed[q]) atomicMax(out, (unsigned long long)__double_as_longlong(radius[q]));
}

// gather rows of a (n x w) float matrix: out[b] = in[idx[b]]
__global__ void k_gather_rows(const float *in, const uint32_t *idx, uint32_t B, uint32_t w, float *out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * w) return;
  const uint32_t b = t / w, j = t % w;
  out[t] = in[(size_t)idx[b] * w + j];
}

// ---------------------------------------------------------------- training kernels
__global__ void k_mlp_fwd(const float *__restrict__ Wg, const float *__restrict__ X, uint32_t B,
                          float *H1, float *H2, float *out) {
  __shared__ float W[MLP_NW];
  for (uint32_t i = threadIdx.x; i < MLP_NW; i += blockDim.x) W[i] = Wg[i];
  __syncthreads();
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float f[MLP_IN];
#pragma unroll
  for (int j = 0; j < MLP_IN; j++) f[j] = X[(size_t)b * MLP_IN + j];
  out[b] = mlp_forward(W, f, H1 + (size_t)b * MLP_H, H2 + (size_t)b * MLP_H);
}

// per sample: dout = 2 (out - y) / B, dH2 = dout w3 [h2 > 0], dH1 = (W2^T dH2) [h1 > 0]
__global__ void k_mlp_bwd_samples(const float *__restrict__ Wg, const float *__restrict__ Y,
                                  uint32_t B, const float *H1, const float *H2, const float *out,
                                  float *dout, float *dH1, float *dH2) {
  __shared__ float W[MLP_NW];
  for (uint32_t i = threadIdx.x; i < MLP_NW; i += blockDim.x) W[i] = Wg[i];
  __syncthreads();
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float *W2 = W + 832, *w3 = W + 4992;
  const float d = 2.0f * (out[b] - Y[b]) / (float)B;
  dout[b] = d;
  float g2[MLP_H];
#pragma unroll
  for (int i = 0; i < MLP_H; i++) {
    g2[i] = H2[(size_t)b * MLP_H + i] > 0.0f ? d * w3[i] : 0.0f;
    dH2[(size_t)b * MLP_H + i] = g2[i];
  }
  for (int j = 0; j < MLP_H; j++) {
    float a = 0.0f;
#pragma unroll
    for (int i = 0; i < MLP_H; i++) a = fmaf(W2[i * MLP_H + j], g2[i], a);
    dH1[(size_t)b * MLP_H + j] = H1[(size_t)b * MLP_H + j] > 0.0f ? a : 0.0f;
  }
}

// one thread per parameter: gradients summed over the batch in sample order
__global__ void k_mlp_wgrad(const float *__restrict__ X, uint32_t B, const float *H1, const float *H2,
                            const float *dout, const float *dH1, const float *dH2, float *G) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NW) return;
  float a = 0.0f;
  if (p < 768) {  // W1[i][j] = sum dH1[b][i] X[b][j]
    const uint32_t i = p / MLP_IN, j = p % MLP_IN;
    for (uint32_t b = 0; b < B; b++) a = fmaf(dH1[(size_t)b * MLP_H + i], X[(size_t)b * MLP_IN + j], a);
  } else if (p < 832) {  // b1
    const uint32_t i = p - 768;
    for (uint32_t b = 0; b < B; b++) a += dH1[(size_t)b * MLP_H + i];
  } else if (p < 4928) {  // W2[i][j] = sum dH2[b][i] H1[b][j]
    const uint32_t i = (p - 832) / MLP_H, j = (p - 832) % MLP_H;
    for (uint32_t b = 0; b < B; b++) a = fmaf(dH2[(size_t)b * MLP_H + i], H1[(size_t)b * MLP_H + j], a);
  } else if (p < 4992) {  // b2
    const uint32_t i = p - 4928;
    for (uint32_t b = 0; b < B; b++) a += dH2[(size_t)b * MLP_H + i];
  } else if (p < 5056) {  // w3[i] = sum dout[b] H2[b][i]
    const uint32_t i = p - 4992;
    for (uint32_t b = 0; b < B; b++) a = fmaf(dout[b], H2[(size_t)b * MLP_H + i], a);
  } else {  // b3
    for (uint32_t b = 0; b < B; b++) a += dout[b];
  }
  G[p] = a;
}

// mean squared error of the batch, summed in sample order
__global__ void k_mlp_loss(const float *out, const float *Y, uint32_t B, float *loss) {
  if (threadIdx.x || blockIdx.x) return;
  float s = 0.0f;
  for (uint32_t b = 0; b < B; b++) {
    const float e = out[b] - Y[b];
    s = fmaf(e, e, s);
  }
  *loss = s / (float)B;
}

// Adam: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2, w -= lr * (m / c1) / (sqrt(v / c2) + eps)
__global__ void k_adam(float *W, const float *G, float *M, float *V, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NW) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.0f - b1) * g;
  const float v = b2 * V[p] + (1.0f - b2) * (g * g);
  M[p] = m;
  V[p] = v;
  W[p] = W[p] - lr * ((m / c1) / (sqrtf(v / c2) + eps));
}

}  // namespace

template <typename T>
void feature_counts(const T *x, uint32_t n, const typename GridOf<T>::type &g, uint32_t *cnt,
                    uint64_t cells, uint32_t reps, hipStream_t s) {
  if (!n || !cells) return;
  if (reps < 1) reps = 1;
  k_feature_counts<T, typename GridOf<T>::type><<<(n + 255) / 256, 256, 0, s>>>(x, n, g, cnt, cells, reps);
  if (reps > 1) k_sum_reps<<<(uint32_t)((cells + 255) / 256), 256, 0, s>>>(cnt, cells, reps);
}

template <typename T>
void radius_kernel(const T *x, uint32_t n, const typename GridOf<T>::type &g, const uint32_t *cnt,
                   float logk1, float logdens, const float *W, float margin, float wf, T *radius,
                   float *featOut, uint32_t *rhist, hipStream_t s) {
  if (!n) return;
  if (radius && !featOut && W) {
    // inference only (the join): matrix cores, persistent grid
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)n + 255) / 256, 2048);
    k_radius_mfma<T, typename GridOf<T>::type><<<blocks, 256, 0, s>>>(x, n, g, cnt, logk1, logdens, W,
                                                                        margin, wf, radius, rhist);
    return;
  }
  k_radius<T, typename GridOf<T>::type><<<(n + 255) / 256, 256, 0, s>>>(
      x, n, g, cnt, logk1, logdens, W, margin, wf, radius, featOut, rhist);
}

template <typename T>
void radius_bins(const T *radius, const uint8_t *pending, uint32_t n, float wf, uint32_t *hist, hipStream_t s) {
  if (n) {
    const uint32_t blocks = (n + 255) / 256;
    k_radius_bins<T><<<blocks < 1024 ? blocks : 1024, 256, 0, s>>>(radius, pending, n, wf, hist);
  }
}
template <typename T>
void radius_class_flags(const T *radius, const uint8_t *pending, uint32_t n, float wf, uint32_t lo,
                        uint32_t hi, uint8_t *act, hipStream_t s) {
  if (n) k_radius_class<T><<<(n + 255) / 256, 256, 0, s>>>(radius, pending, n, wf, lo, hi, act);
}
template <typename T>
void grow_class_radius(T *radius, const uint8_t *act, const uint8_t *failed, uint32_t n, T factor,
                       hipStream_t s) {
  if (n) k_grow_class<T><<<(n + 255) / 256, 256, 0, s>>>(radius, act, failed, n, factor);
}

void count_nonzero(const uint32_t *cnt, uint64_t cells, unsigned long long *out, hipStream_t s) {
  if (cells) k_count_nonzero<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(cnt, cells, out);
}

template <typename T>
void grow_radius(T *radius, const uint8_t *failed, uint32_t n, T factor, hipStream_t s) {
  if (n) k_grow_radius<T><<<(n + 255) / 256, 256, 0, s>>>(radius, failed, n, factor);
}

void max_failed_radius(const float *radius, const uint8_t *failed, uint32_t n, void *out, hipStream_t s) {
  fill_u32(reinterpret_cast<uint32_t *>(out), 0u, 2, s);
  if (n) k_max_radius_f<<<(n + 255) / 256, 256, 0, s>>>(radius, failed, n, reinterpret_cast<uint32_t *>(out));
}
void max_failed_radius(const double *radius, const uint8_t *failed, uint32_t n, void *out, hipStream_t s) {
  fill_u32(reinterpret_cast<uint32_t *>(out), 0u, 2, s);
  if (n) k_max_radius_d<<<(n + 255) / 256, 256, 0, s>>>(radius, failed, n, reinterpret_cast<unsigned long long *>(out));
}
void gather_rows(const float *in, const uint32_t *idx, uint32_t B, uint32_t w, float *out, hipStream_t s) {
  if (B) k_gather_rows<<<(B * w + 255) / 256, 256, 0, s>>>(in, idx, B, w, out);
}

void mlp_forward_batch(const float *W, const float *X, uint32_t B, float *H1, float *H2, float *out,
                       hipStream_t s) {
  if (B) k_mlp_fwd<<<(B + 127) / 128, 128, 0, s>>>(W, X, B, H1, H2, out);
}

void mlp_grad_batch(const float *W, const float *X, const float *Y, uint32_t B, float *H1,
                    float *H2, float *out, float *dout, float *dH1, float *dH2, float *G,
                    float *loss, hipStream_t s) {
  if (!B) return;
  k_mlp_fwd<<<(B + 127) / 128, 128, 0, s>>>(W, X, B, H1, H2, out);
  k_mlp_bwd_samples<<<(B + 127) / 128, 128, 0, s>>>(W, Y, B, H1, H2, out, dout, dH1, dH2);
  k_mlp_wgrad<<<(MLP_NW + 63) / 64, 64, 0, s>>>(X, B, H1, H2, dout, dH1, dH2, G);
  if (loss) k_mlp_loss<<<1, 64, 0, s>>>(out, Y, B, loss);
}

void adam_step(float *W, const float *G, float *M, float *V, float lr, float b1, float b2, float c1,
               float c2, float eps, hipStream_t s) {
  k_adam<<<(MLP_NW + 255) / 256, 256, 0, s>>>(W, G, M, V, lr, b1, b2, c1, c2, eps);
}

#define INST(T)                                                                                   \
  template void feature_counts<T>(const T *, uint32_t, const GridOf<T>::type &, uint32_t *,        \
                                  uint64_t, uint32_t, hipStream_t);                               \
  template void radius_kernel<T>(const T *, uint32_t, const GridOf<T>::type &, const uint32_t *,   \
                                 float, float, const float *, float, float, T *, float *,         \
                                 uint32_t *, hipStream_t);                                        \
  template void grow_radius<T>(T *, const uint8_t *, uint32_t, T, hipStream_t);                   \
  template void radius_bins<T>(const T *, const uint8_t *, uint32_t, float, uint32_t *, hipStream_t); \
  template void radius_class_flags<T>(const T *, const uint8_t *, uint32_t, float, uint32_t, uint32_t, \
                                      uint8_t *, hipStream_t);                                     \
  template void grow_class_radius<T>(T *, const uint8_t *, const uint8_t *, uint32_t, T, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace knn
