/*
 * knn_oracle.c -- CPU oracle for the exact KNN self-join (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the hot path of mgowanlock/hybrid_k_nearest_neighbor_self_join in plain C and
 * adds an FP64 brute-force KNN.  Loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py; the product library never links it.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp; no -ffast-math, so float and
 * double arithmetic is IEEE round-to-nearest, as in the reference's host code).
 */
#include "knn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- bounded max-heap on (key, id), used as the per-query top-(K+1) ---- */
typedef struct {
  double *key;
  uint32_t *id;
  uint32_t size, cap;
} orc_heap;

static inline int orc_less(double ka, uint32_t ia, double kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

static void orc_sift_down(orc_heap *h, uint32_t i, uint32_t size) {
  for (;;) {
    uint32_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < size && orc_less(h->key[m], h->id[m], h->key[l], h->id[l])) m = l;
    if (r < size && orc_less(h->key[m], h->id[m], h->key[r], h->id[r])) m = r;
    if (m == i) return;
    double tk = h->key[i];
    uint32_t ti = h->id[i];
    h->key[i] = h->key[m];
    h->id[i] = h->id[m];
    h->key[m] = tk;
    h->id[m] = ti;
    i = m;
  }
}

static inline void orc_heap_push(orc_heap *h, double key, uint32_t id) {
  if (h->size < h->cap) {
    uint32_t i = h->size++;
    h->key[i] = key;
    h->id[i] = id;
    while (i > 0) {
      uint32_t p = (i - 1) / 2;
      if (!orc_less(h->key[p], h->id[p], h->key[i], h->id[i])) break;
      double tk = h->key[i];
      uint32_t ti = h->id[i];
      h->key[i] = h->key[p];
      h->id[i] = h->id[p];
      h->key[p] = tk;
      h->id[p] = ti;
      i = p;
    }
  } else if (orc_less(key, id, h->key[0], h->id[0])) {
    h->key[0] = key;
    h->id[0] = id;
    orc_sift_down(h, 0, h->size);
  }
}

/* heap sort in place: ascending (key, id) */
static void orc_heap_sort(orc_heap *h) {
  for (uint32_t end = h->size; end > 1; end--) {
    double tk = h->key[0];
    uint32_t ti = h->id[0];
    h->key[0] = h->key[end - 1];
    h->id[0] = h->id[end - 1];
    h->key[end - 1] = tk;
    h->id[end - 1] = ti;
    orc_sift_down(h, 0, end - 1);
  }
}

/* Stable LSD radix sort of ids 0..n-1 by 64-bit key (the stable sort_by_key of
 * GPU.cu:3959 on a value array initialised to 0..n-1, GPU.cu:3952). */
static void orc_radix_sort_pairs(const uint64_t *key, uint32_t *ids, uint64_t *sorted_keys,
                                 uint32_t n) {
  uint64_t maxk = 0;
  for (uint32_t i = 0; i < n; i++)
    if (key[i] > maxk) maxk = key[i];
  uint64_t *ka = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  uint64_t *kb = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  uint32_t *va = ids;
  uint32_t *vb = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
  for (uint32_t i = 0; i < n; i++) {
    ka[i] = key[i];
    va[i] = i;
  }
  int passes = 0;
  while (passes < 8 && (maxk >> (8 * passes)) != 0) passes++;
  for (int p = 0; p < passes; p++) {
    uint32_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    int sh = 8 * p;
    for (uint32_t i = 0; i < n; i++) cnt[((ka[i] >> sh) & 255) + 1]++;
    for (int b = 0; b < 256; b++) cnt[b + 1] += cnt[b];
    for (uint32_t i = 0; i < n; i++) {
      uint32_t pos = cnt[(ka[i] >> sh) & 255]++;
      kb[pos] = ka[i];
      vb[pos] = va[i];
    }
    uint64_t *tk = ka;
    ka = kb;
    kb = tk;
    uint32_t *tv = va;
    va = vb;
    vb = tv;
  }
  if (va != ids) {
    memcpy(ids, va, sizeof(uint32_t) * (size_t)n);
    free(va);
  } else {
    free(vb);
  }
  if (sorted_keys) memcpy(sorted_keys, ka, sizeof(uint64_t) * (size_t)n);
  free(ka);
  free(kb);
}

/* main.cu:2385-2401 / kernel.cu:31-47: dim 0 fastest mixed radix */
static inline uint64_t orc_linear_id_impl(const uint32_t *idx, const uint32_t *nCells,
                                          uint32_t nd) {
  uint64_t id = 0, mult = 1;
  for (uint32_t i = 0; i < nd; i++) {
    id += (uint64_t)idx[i] * mult;
    mult *= nCells[i];
  }
  return id;
}

/* main.cu:330-339: eps_est += (eps_est_initial*0.5) -- float + double product, stored
 * back into the float DTYPE variable. */
float orc_grow_eps_f32(float eps, float eps0) { return (float)((double)eps + (double)eps0 * 0.5); }
static double orc_grow_eps_f64(double eps, double eps0) { return eps + eps0 * 0.5; }

/* ---- instantiate for float ---- */
#define T float
#define SFX f32
#define NEXTAFTER nextafterf
#define ORC_GROW orc_grow_eps_f32
#include "knn_oracle_impl.inc"
#undef T
#undef SFX
#undef NEXTAFTER
#undef ORC_GROW

/* ---- instantiate for double ---- */
#define T double
#define SFX f64
#define NEXTAFTER nextafter
#define ORC_GROW orc_grow_eps_f64
#include "knn_oracle_impl.inc"
#undef T
#undef SFX
#undef NEXTAFTER
#undef ORC_GROW

/* ---- public wrappers ---- */
uint32_t orc_cell_coord_f32(float x, float mn, float eps, uint32_t ncell) {
  return orc_coord_f32(x, mn, eps, ncell);
}

void orc_avg_dist_f32(const float *pts, uint32_t n, uint32_t d, double *avg, uint64_t *cnt) {
  orc_avg_dist_impl_f32(pts, n, d, avg, cnt);
}

void orc_minmax_f32(const float *pts, uint32_t n, uint32_t d, uint32_t nd, float *mn, float *mx) {
  orc_minmax_impl_f32(pts, n, d, nd, mn, mx);
}

void orc_grid_dims_f32(const float *mn, const float *mx, uint32_t nd, float eps, float *minArr,
                       uint32_t *nCells, uint64_t *totalCells) {
  orc_grid_dims_impl_f32(mn, mx, nd, eps, minArr, nCells, totalCells);
}

uint32_t orc_build_index_f32(const float *pts, uint32_t n, uint32_t d, uint32_t nd, float eps,
                             const float *minArr, const uint32_t *nCells, uint64_t *pointKey,
                             uint32_t *sortedIds, uint64_t *cellKeys, uint32_t *cellStart) {
  return orc_build_index_impl_f32(pts, n, d, nd, eps, minArr, nCells, pointKey, sortedIds,
                                  cellKeys, cellStart);
}

int orc_knnjoin_f32(const float *pts, uint32_t n, uint32_t d, uint32_t k, uint32_t nd, double beta,
                    int nthreads, int32_t *outIds, float *outDists, orc_stats *st) {
  return orc_knnjoin_impl_f32(pts, n, d, k, nd, beta, nthreads, NULL, 0, outIds, outDists, st);
}

/* the plain method restricted to a query subset (rows of other points are untouched);
 * the epsilon schedule then follows the subset's failure fractions, which leaves every
 * row unchanged because a success is exact at any epsilon. */
int orc_knnjoin_subset_f32(const float *pts, uint32_t n, uint32_t d, uint32_t k, uint32_t nd,
                           double beta, int nthreads, const uint32_t *subset, uint32_t nsub,
                           int32_t *outIds, float *outDists, orc_stats *st) {
  return orc_knnjoin_impl_f32(pts, n, d, k, nd, beta, nthreads, subset, nsub, outIds, outDists, st);
}

int orc_knnjoin_f64(const double *pts, uint32_t n, uint32_t d, uint32_t k, uint32_t nd, double beta,
                    int nthreads, int32_t *outIds, double *outDists, orc_stats *st) {
  return orc_knnjoin_impl_f64(pts, n, d, k, nd, beta, nthreads, NULL, 0, outIds, outDists, st);
}

/* main.cu:452 / GPU.cu:3690-3693: sum of (float)(d*d); each row's K+1 entries are summed in
 * entry order, rows in chunks of 4096 (row order), then the chunk sums in order (the
 * reference's OpenMP/atomic order is not fixed). */
double orc_checksum_f32(const float *dists, uint64_t rows, uint32_t k1) {
  double s = 0, chunk = 0;
  for (uint64_t r = 0; r < rows; r++) {
    double rs = 0;
    for (uint32_t e = 0; e < k1; e++) {
      float dd = dists[r * k1 + e];
      float sq = dd * dd;
      rs += (double)sq;
    }
    chunk += rs;
    if ((r + 1) % 4096 == 0 || r + 1 == rows) {
      s += chunk;
      chunk = 0;
    }
  }
  return s;
}
