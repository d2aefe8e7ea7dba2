// knn_collect.h -- per-query exact top-(K+1) collector in LDS, shared by the brute-force
// kernels and the range collect kernel.
//
// One wave owns a collector: the K+1 best (key, id) pairs seen so far, sorted, and an
// unsorted buffer of newer candidates.  Candidates below the current (K+1)-th pair enter the
// buffer through a ballot; a buffer more than half full is bitonic-sorted and merged into the
// best list by rank (output position = own index + number of entries of the other list that
// order before it), so the threshold tightens as candidates stream by.  Keys are compared as
// (FP64 key, point id) pairs; every point is offered at most once, so pairs are distinct.
#pragma once

#include "knn_device.h"

namespace knn {

constexpr uint32_t BR_BUF = 128;  // candidate buffer entries per query
constexpr uint32_t BR_NONE = 0xffffffffu;

__device__ __forceinline__ bool br_less(double ka, uint32_t ia, double kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

__device__ __forceinline__ void br_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// number of entries of the sorted list (k[0..n), id[0..n)) that order before (kv, iv)
__device__ __forceinline__ uint32_t br_count_less(const double *k, const uint32_t *id, uint32_t n,
                                                  double kv, uint32_t iv) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t m = (lo + hi) >> 1;
    if (br_less(k[m], id[m], kv, iv)) lo = m + 1; else hi = m;
  }
  return lo;
}

// The K+1 best (key, id) of one query seen so far, with an unsorted buffer of newer
// candidates.  LDS layout per query: best keys [2][k1], best ids [2][k1] (double-buffered
// for the merge), buffer keys [BR_BUF], buffer ids [BR_BUF].
struct Collector {
  double *kb;    // best keys, two halves of k1
  uint32_t *ib;  // best ids, two halves of k1
  double *uk;
  uint32_t *ui;
  uint32_t k1;
  uint32_t nb = 0, nu = 0, cur = 0;  // cur: offset (0 or k1) of the current best half
  double tk = __builtin_inf();  // threshold: the k1-th best once nb == k1
  uint32_t ti = BR_NONE;

  __device__ static size_t bytes(uint32_t k1) {
    return (size_t)k1 * 2 * (sizeof(double) + sizeof(uint32_t)) + BR_BUF * (sizeof(double) + sizeof(uint32_t));
  }
  __device__ void init(unsigned char *base, uint32_t k1_) {
    k1 = k1_;
    kb = reinterpret_cast<double *>(base);
    uk = kb + 2 * k1;
    ib = reinterpret_cast<uint32_t *>(uk + BR_BUF);
    ui = ib + 2 * k1;
  }
  // initial threshold: only pairs ordering before (tk0, BR_NONE), i.e. keys <= tk0, are kept
  __device__ __forceinline__ void limit(double tk0) {
    tk = tk0;
    ti = BR_NONE;
  }
  __device__ __forceinline__ double best_key(uint32_t e) const { return kb[cur + e]; }
  __device__ __forceinline__ uint32_t best_id(uint32_t e) const { return ib[cur + e]; }

  // offers one candidate per lane (wave-uniform call); merges when the buffer could overflow
  __device__ __forceinline__ void offer(bool valid, double key, uint32_t id, uint32_t lane) {
    const bool pass = valid && br_less(key, id, tk, ti);
    const uint64_t m = __ballot(pass);
    if (m == 0) return;
    if (pass) {
      const uint32_t slot = nu + (uint32_t)__popcll(m & lanemask_lt());
      uk[slot] = key;
      ui[slot] = id;
    }
    nu += (uint32_t)__popcll(m);
    if (nu > BR_BUF - 64) merge(lane);
  }

  __device__ __forceinline__ void merge(uint32_t lane) {
    if (nu == 0) return;
    br_sync();
    uint32_t P = 64;
    while (P < nu) P <<= 1;
    for (uint32_t e = nu + lane; e < P; e += 64) {
      uk[e] = __builtin_inf();
      ui[e] = BR_NONE;
    }
    br_sync();
    // bitonic sort of the buffer, ascending by (key, id)
    for (uint32_t size = 2; size <= P; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t t = lane; t < P / 2; t += 64) {
          const uint32_t i = 2 * t - (t & (stride - 1));
          const uint32_t j = i + stride;
          const bool up = (i & size) == 0;
          const double ka = uk[i], kb = uk[j];
          const uint32_t ia = ui[i], ib = ui[j];
          if (br_less(kb, ib, ka, ia) == up) {
            uk[i] = kb;
            ui[i] = ib;
            uk[j] = ka;
            ui[j] = ia;
          }
        }
        br_sync();
      }
    }
    // merge by rank into the other best half (keys are distinct pairs: each point is
    // offered once, so the ranks form a permutation)
    const double *ok = kb + cur;
    const uint32_t *oi = ib + cur;
    const uint32_t nxt = cur ^ k1;
    double *nk = kb + nxt;
    uint32_t *ni = ib + nxt;
    for (uint32_t e = lane; e < nb; e += 64) {
      const double kv = ok[e];
      const uint32_t iv = oi[e];
      const uint32_t r = e + br_count_less(uk, ui, nu, kv, iv);
      if (r < k1) {
        nk[r] = kv;
        ni[r] = iv;
      }
    }
    for (uint32_t e = lane; e < nu; e += 64) {
      const double kv = uk[e];
      const uint32_t iv = ui[e];
      const uint32_t r = e + br_count_less(ok, oi, nb, kv, iv);
      if (r < k1) {
        nk[r] = kv;
        ni[r] = iv;
      }
    }
    br_sync();
    cur = nxt;
    nb = nb + nu < k1 ? nb + nu : k1;
    nu = 0;
    if (nb == k1) {
      tk = kb[cur + k1 - 1];
      ti = ib[cur + k1 - 1];
    }
  }
};

}  // namespace knn
