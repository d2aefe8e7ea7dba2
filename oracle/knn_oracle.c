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
#include <omp.h>
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
 * GPU.cu:3959 on a value array initialised to 0..n-1, GPU.cu:3952).  8-bit digits; each
 * OpenMP thread counts and then scatters its own contiguous slice, and a slice's
 * destinations follow the slices before it for every digit, so the order of equal keys is
 * the input order (stable) with any thread count. */
static void orc_radix_sort_pairs(const uint64_t *key, uint32_t *ids, uint64_t *sorted_keys,
                                 uint32_t n) {
  uint64_t maxk = 0;
  for (uint32_t i = 0; i < n; i++)
    if (key[i] > maxk) maxk = key[i];
  uint64_t *ka = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
  uint64_t *kb = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
  uint32_t *va = ids;
  uint32_t *vb = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  for (uint32_t i = 0; i < n; i++) {
    ka[i] = key[i];
    va[i] = i;
  }
  int passes = 0;
  while (passes < 8 && (maxk >> (8 * passes)) != 0) passes++;
  const int nt = n >= (1u << 16) ? omp_get_max_threads() : 1;
  size_t *cnt = (size_t *)malloc(sizeof(size_t) * 256 * (size_t)nt);
  for (int p = 0; p < passes; p++) {
    const int sh = 8 * p;
#pragma omp parallel num_threads(nt)
    {
      const int t = omp_get_thread_num();
      const size_t lo = (size_t)n * t / nt, hi = (size_t)n * (t + 1) / nt;
      size_t *c = cnt + 256 * (size_t)t;
      memset(c, 0, sizeof(size_t) * 256);
      for (size_t i = lo; i < hi; i++) c[(ka[i] >> sh) & 255]++;
#pragma omp barrier
#pragma omp single
      {
        size_t run = 0;
        for (int b = 0; b < 256; b++)
          for (int u = 0; u < nt; u++) {
            const size_t v = cnt[256 * (size_t)u + b];
            cnt[256 * (size_t)u + b] = run;
            run += v;
          }
      }
      for (size_t i = lo; i < hi; i++) {
        const size_t pos = c[(ka[i] >> sh) & 255]++;
        kb[pos] = ka[i];
        vb[pos] = va[i];
      }
    }
    uint64_t *tk = ka;
    ka = kb;
    kb = tk;
    uint32_t *tv = va;
    va = vb;
    vb = tv;
  }
  free(cnt);
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

/* ======================================================================================
 * Learned-radius MLP (12 -> 64 -> 64 -> 1, ReLU) -- CPU restatement of the product's
 * forward / gradient / Adam arithmetic, in the same fmaf and summation order, so that the
 * GPU results must agree bit for bit.  Weights: W1[64][12], b1[64], W2[64][64], b2[64],
 * w3[64], b3.
 * ====================================================================================== */
#define MLP_IN 12
#define MLP_H 64
#define MLP_NW 5057

static float orc_relu(float v) { return v > 0.0f ? v : 0.0f; }

static float orc_mlp_fwd1(const float *W, const float *f, float *H1, float *H2) {
  const float *W1 = W, *b1 = W + 768, *W2 = W + 832, *b2 = W + 4928, *w3 = W + 4992;
  float acc[MLP_H];
  for (int i = 0; i < MLP_H; i++) acc[i] = b2[i];
  for (int j = 0; j < MLP_H; j++) {
    float a = b1[j];
    for (int t = 0; t < MLP_IN; t++) a = fmaf(W1[j * MLP_IN + t], f[t], a);
    float h = orc_relu(a);
    if (H1) H1[j] = h;
    for (int i = 0; i < MLP_H; i++) acc[i] = fmaf(W2[i * MLP_H + j], h, acc[i]);
  }
  float y = W[5056];
  for (int i = 0; i < MLP_H; i++) {
    float h = orc_relu(acc[i]);
    if (H2) H2[i] = h;
    y = fmaf(w3[i], h, y);
  }
  return y;
}

void orc_mlp_forward(const float *W, const float *X, uint32_t B, float *out) {
  for (uint32_t b = 0; b < B; b++) out[b] = orc_mlp_fwd1(W, X + (size_t)b * MLP_IN, NULL, NULL);
}

/* gradient of mean((out - y)^2) over the batch, in the product's order */
void orc_mlp_grad(const float *W, const float *X, const float *Y, uint32_t B, float *G, float *loss) {
  float *H1 = (float *)malloc(sizeof(float) * B * MLP_H);
  float *H2 = (float *)malloc(sizeof(float) * B * MLP_H);
  float *dH1 = (float *)malloc(sizeof(float) * B * MLP_H);
  float *dH2 = (float *)malloc(sizeof(float) * B * MLP_H);
  float *out = (float *)malloc(sizeof(float) * B);
  float *dout = (float *)malloc(sizeof(float) * B);
  const float *W2 = W + 832, *w3 = W + 4992;
  for (uint32_t b = 0; b < B; b++) out[b] = orc_mlp_fwd1(W, X + (size_t)b * MLP_IN, H1 + (size_t)b * MLP_H, H2 + (size_t)b * MLP_H);
  for (uint32_t b = 0; b < B; b++) {
    float d = 2.0f * (out[b] - Y[b]) / (float)B;
    dout[b] = d;
    float g2[MLP_H];
    for (int i = 0; i < MLP_H; i++) {
      g2[i] = H2[(size_t)b * MLP_H + i] > 0.0f ? d * w3[i] : 0.0f;
      dH2[(size_t)b * MLP_H + i] = g2[i];
    }
    for (int j = 0; j < MLP_H; j++) {
      float a = 0.0f;
      for (int i = 0; i < MLP_H; i++) a = fmaf(W2[i * MLP_H + j], g2[i], a);
      dH1[(size_t)b * MLP_H + j] = H1[(size_t)b * MLP_H + j] > 0.0f ? a : 0.0f;
    }
  }
  for (uint32_t p = 0; p < MLP_NW; p++) {
    float a = 0.0f;
    if (p < 768) {
      uint32_t i = p / MLP_IN, j = p % MLP_IN;
      for (uint32_t b = 0; b < B; b++) a = fmaf(dH1[(size_t)b * MLP_H + i], X[(size_t)b * MLP_IN + j], a);
    } else if (p < 832) {
      for (uint32_t b = 0; b < B; b++) a += dH1[(size_t)b * MLP_H + (p - 768)];
    } else if (p < 4928) {
      uint32_t i = (p - 832) / MLP_H, j = (p - 832) % MLP_H;
      for (uint32_t b = 0; b < B; b++) a = fmaf(dH2[(size_t)b * MLP_H + i], H1[(size_t)b * MLP_H + j], a);
    } else if (p < 4992) {
      for (uint32_t b = 0; b < B; b++) a += dH2[(size_t)b * MLP_H + (p - 4928)];
    } else if (p < 5056) {
      for (uint32_t b = 0; b < B; b++) a = fmaf(dout[b], H2[(size_t)b * MLP_H + (p - 4992)], a);
    } else {
      for (uint32_t b = 0; b < B; b++) a += dout[b];
    }
    G[p] = a;
  }
  if (loss) {
    float s = 0.0f;
    for (uint32_t b = 0; b < B; b++) {
      float e = out[b] - Y[b];
      s = fmaf(e, e, s);
    }
    *loss = s / (float)B;
  }
  free(H1); free(H2); free(dH1); free(dH2); free(out); free(dout);
}

/* Adam training restated (splitmix64 Fisher-Yates per epoch, batches of the shuffled order) */
void orc_mlp_train(float *W, const float *X, const float *Y, uint32_t n, uint32_t epochs,
                   uint32_t batch, float lr, uint64_t seed) {
  uint32_t B = batch < n ? batch : n;
  float *M = (float *)calloc(MLP_NW, sizeof(float));
  float *V = (float *)calloc(MLP_NW, sizeof(float));
  float *G = (float *)malloc(sizeof(float) * MLP_NW);
  float *Xb = (float *)malloc(sizeof(float) * B * MLP_IN);
  float *Yb = (float *)malloc(sizeof(float) * B);
  uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * n);
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  uint64_t st = seed, step = 0;
  const float b1 = 0.9f, b2 = 0.999f, epsA = 1e-8f;
  for (uint32_t e = 0; e < epochs; e++) {
    for (uint32_t i = n - 1; i > 0; i--) {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z = z ^ (z >> 31);
      uint32_t j = (uint32_t)(z % (uint64_t)(i + 1));
      uint32_t t = order[i]; order[i] = order[j]; order[j] = t;
    }
    for (uint32_t bi = 0; bi < n / B; bi++) {
      for (uint32_t b = 0; b < B; b++) {
        memcpy(Xb + (size_t)b * MLP_IN, X + (size_t)order[bi * B + b] * MLP_IN, sizeof(float) * MLP_IN);
        Yb[b] = Y[order[bi * B + b]];
      }
      orc_mlp_grad(W, Xb, Yb, B, G, NULL);
      step++;
      float c1 = (float)(1.0 - pow(0.9, (double)step));
      float c2 = (float)(1.0 - pow(0.999, (double)step));
      for (uint32_t p = 0; p < MLP_NW; p++) {
        float g = G[p];
        float m = b1 * M[p] + (1.0f - b1) * g;
        float v = b2 * V[p] + (1.0f - b2) * (g * g);
        M[p] = m;
        V[p] = v;
        W[p] = W[p] - lr * ((m / c1) / (sqrtf(v / c2) + epsA));
      }
    }
  }
  free(M); free(V); free(G); free(Xb); free(Yb); free(order);
}

/* Features of the learned radius (see the product's learned.hip header for the definition),
 * on points already in the reordered dim order (row-major n x d). */
void orc_radius_features(const float *pts, uint32_t n, uint32_t d, uint32_t k, uint32_t nd,
                         float *feat, float *wf_out) {
  float mn[ORC_MAXDIM], mx[ORC_MAXDIM], minArr[ORC_MAXDIM];
  uint32_t nCells[ORC_MAXDIM];
  uint64_t tot;
  orc_minmax_impl_f32(pts, n, d, nd, mn, mx);
  double V = 1.0;
  for (uint32_t j = 0; j < nd; j++) {
    double e = (double)(mx[j] - mn[j]);
    V *= e > 0.0 ? e : 1.0;
  }
  float wf = (float)pow(V * (double)(k + 1) / (double)n, 1.0 / (double)nd);
  if (!(wf > 0.0f)) wf = 1.0f;
  for (;;) {
    orc_grid_dims_impl_f32(mn, mx, nd, wf, minArr, nCells, &tot);
    if (tot <= (1ull << 27)) break;
    wf = wf * 2.0f;
  }
  uint64_t mult[ORC_MAXDIM];
  uint64_t m = 1;
  for (uint32_t j = 0; j < nd; j++) { mult[j] = m; m *= nCells[j]; }
  uint32_t *cnt = (uint32_t *)calloc(tot, sizeof(uint32_t));
  int64_t *cc = (int64_t *)malloc(sizeof(int64_t) * (size_t)n * nd);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t id = 0;
    for (uint32_t j = 0; j < nd; j++) {
      cc[i * nd + j] = orc_coord_f32(pts[i * d + j], minArr[j], wf, nCells[j]);
      id += (uint64_t)cc[i * nd + j] * mult[j];
    }
    cnt[id]++;
  }
  float logk1 = logf((float)(k + 1));
  uint64_t occ = 0;
  for (uint64_t c = 0; c < tot; c++) occ += cnt[c] != 0;
  float logdens = (float)log((double)n / (double)occ);
  for (uint64_t q = 0; q < n; q++) {
    const int64_t *c = cc + q * nd;
    uint64_t own = 0;
    for (uint32_t j = 0; j < nd; j++) own += (uint64_t)c[j] * mult[j];
    float *f = feat + q * 12;
    f[0] = log1pf((float)cnt[own]);
    int kk = 1;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int64_t c0, c1 = 0;
        if (nd == 1) c0 = c[0] + (dy + 1) * 3 + dx - 4;
        else { c0 = c[0] + dx; c1 = c[1] + dy; }
        float v = 0.0f;
        int in0 = c0 >= 0 && c0 < (int64_t)nCells[0];
        int in1 = nd == 1 || (c1 >= 0 && c1 < (int64_t)nCells[1]);
        if (in0 && in1) {
          uint64_t id = own - (uint64_t)c[0] * mult[0] + (uint64_t)c0 * mult[0];
          if (nd > 1) id = id - (uint64_t)c[1] * mult[1] + (uint64_t)c1 * mult[1];
          v = log1pf((float)cnt[id]);
        }
        f[kk++] = v;
      }
    f[10] = logk1;
    f[11] = logdens;
  }
  if (wf_out) *wf_out = wf;
  free(cnt);
  free(cc);
}


/* ---- outlier scores (main.cu:952-1067) ---- */
static const float *g_sort_sum;
static const uint32_t *g_sort_deg;
static int cmp_sum_desc(const void *a, const void *b) {
  const uint32_t i = *(const uint32_t *)a, j = *(const uint32_t *)b;
  const float x = g_sort_sum[i], y = g_sort_sum[j];
  const int xn = x != x, yn = y != y;
  if (xn != yn) return xn - yn; /* NaN last */
  if (!xn && x > y) return -1;
  if (!xn && x < y) return 1;
  return (i > j) - (i < j);
}
static int cmp_deg_asc(const void *a, const void *b) {
  const uint32_t i = *(const uint32_t *)a, j = *(const uint32_t *)b;
  if (g_sort_deg[i] != g_sort_deg[j]) return g_sort_deg[i] < g_sort_deg[j] ? -1 : 1;
  return (i > j) - (i < j);
}

void orc_outlier_scores(const int32_t *ids, const float *dists, uint32_t n, uint32_t k1,
                        float *dist_sum, uint32_t *dist_rank, uint32_t *indegree,
                        uint32_t *indegree_rank) {
  uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; i++) {
    float s = 0.0f;
    for (uint32_t e = 0; e < k1; e++) s += dists[(size_t)i * k1 + e];
    dist_sum[i] = s;
    indegree[i] = 0;
  }
  for (uint64_t t = 0; t < (uint64_t)n * k1; t++) {
    const int32_t id = ids[t];
    if (id >= 0 && (uint32_t)id < n) indegree[id]++;
  }
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  g_sort_sum = dist_sum;
  qsort(order, n, sizeof(uint32_t), cmp_sum_desc);
  for (uint32_t p = 0; p < n; p++) dist_rank[order[p]] = p;
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  g_sort_deg = indegree;
  qsort(order, n, sizeof(uint32_t), cmp_deg_asc);
  for (uint32_t p = 0; p < n; p++) indegree_rank[order[p]] = p;
  free(order);
}
