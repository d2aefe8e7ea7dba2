/*
 * knn_oracle.h -- CPU oracle for the exact KNN self-join hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the algorithm of
 * mgowanlock/hybrid_k_nearest_neighbor_self_join on its hot path (dimension reorder,
 * BETA-driven epsilon estimate, grid build, epsilon-range top-(K+1), failed-query loop)
 * and adds an FP64 brute-force KNN.  It is loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker or the
 * reported CPU baseline.  The product library (libgpuknnselfjoin.so) never links or calls it.
 *
 * Every function cites the reference file:line it follows.
 */
#ifndef KNN_ORACLE_H
#define KNN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXDIM 16
#define ORC_NBUCKETS 10000u

enum {
  ORC_OK = 0,
  ORC_ERR_ARG = 1,        /* bad sizes (N==0 handled as OK) */
  ORC_ERR_TOO_FEW = 2,    /* N < K+1: the reference loops forever */
  ORC_ERR_EPS_ZERO = 3,   /* epsilon estimate is 0 (all points identical) */
  ORC_ERR_NO_PROGRESS = 4 /* epsilon growth stalled (eps + eps0/2 == eps) */
};

typedef struct {
  double avg;          /* mean sampled pairwise distance */
  uint64_t cnt;        /* number of sampled pairs */
  double bucket_width; /* avg / 10000 */
  uint32_t offset;     /* k-dist sample stride */
  uint32_t nsample;    /* number of k-dist sample points (TOTALBLOCKS2) */
  uint32_t bucket;     /* chosen bucket */
  uint32_t threshold;  /* k + (100k-k)*beta, truncated */
  float eps;           /* epsilon_0 */
} orc_eps_info;

typedef struct {
  uint32_t perm[ORC_MAXDIM];
  float eps0;
  float eps_final;
  uint32_t iterations;      /* numIter of main.cu:312-430 */
  uint32_t reindexes;       /* epsilon growths */
  uint64_t candidates;      /* distance evaluations in the join passes */
  double checksum;          /* sum over rows of sum (float)(d*d), main.cu:452 */
  double ms_prep;           /* wall ms: dimension reorder, epsilon estimate, min/max */
  double ms_index;          /* wall ms: every index build (the first and each re-index) */
  double ms_passes;         /* wall ms: the join passes and the failed-list updates */
} orc_stats;

/* ---- float (DTYPE float) ---- */
void orc_reorder_f32(const float *pts, uint32_t n, uint32_t d, uint32_t *perm, float *var);
void orc_apply_perm_f32(const float *pts, uint32_t n, uint32_t d, const uint32_t *perm, float *out);
void orc_avg_dist_f32(const float *pts, uint32_t n, uint32_t d, double *avg, uint64_t *cnt);
int orc_estimate_eps_f32(const float *pts, uint32_t n, uint32_t d, uint32_t k, double beta,
                         uint32_t *hist, orc_eps_info *info);
void orc_minmax_f32(const float *pts, uint32_t n, uint32_t d, uint32_t nd, float *mn, float *mx);
void orc_grid_dims_f32(const float *mn, const float *mx, uint32_t nd, float eps, float *minArr,
                       uint32_t *nCells, uint64_t *totalCells);
uint32_t orc_build_index_f32(const float *pts, uint32_t n, uint32_t d, uint32_t nd, float eps,
                             const float *minArr, const uint32_t *nCells, uint64_t *pointKey,
                             uint32_t *sortedIds, uint64_t *cellKeys, uint32_t *cellStart);
uint32_t orc_join_pass_f32(const float *pts, uint32_t n, uint32_t d, uint32_t nd, uint32_t k,
                           float eps, const float *minArr, const uint32_t *nCells, uint32_t ncell,
                           const uint64_t *cellKeys, const uint32_t *cellStart,
                           const uint32_t *sortedIds, const uint32_t *queries, uint32_t nq,
                           int nthreads, int32_t *outIds, float *outDists, uint8_t *failed,
                           uint64_t *candidates);
int orc_knnjoin_f32(const float *pts, uint32_t n, uint32_t d, uint32_t k, uint32_t nd, double beta,
                    int nthreads, int32_t *outIds, float *outDists, orc_stats *st);
void orc_bruteforce_f32(const float *pts, uint32_t n, uint32_t d, uint32_t k,
                        const uint32_t *queries, uint32_t nq, int nthreads, int32_t *outIds,
                        float *outDists);
double orc_checksum_f32(const float *dists, uint64_t rows, uint32_t k1);

/* printOutlierScores, main.cu:952-1067: FP32 row sums of the distances in entry order and
 * their rank by sum descending; in-degree of every point (its own row included) and its
 * rank ascending.  Equal values are ranked by point id (the reference's std::sort leaves
 * them in unspecified order). */
void orc_outlier_scores(const int32_t *ids, const float *dists, uint32_t n, uint32_t k1,
                        float *dist_sum, uint32_t *dist_rank, uint32_t *indegree,
                        uint32_t *indegree_rank);
int orc_knnjoin_subset_f32(const float *pts, uint32_t n, uint32_t d, uint32_t k, uint32_t nd,
                           double beta, int nthreads, const uint32_t *subset, uint32_t nsub,
                           int32_t *outIds, float *outDists, orc_stats *st);

/* ---- double (DTYPE double) ---- */
void orc_reorder_f64(const double *pts, uint32_t n, uint32_t d, uint32_t *perm, double *var);
void orc_apply_perm_f64(const double *pts, uint32_t n, uint32_t d, const uint32_t *perm, double *out);
int orc_estimate_eps_f64(const double *pts, uint32_t n, uint32_t d, uint32_t k, double beta,
                         uint32_t *hist, orc_eps_info *info);
int orc_knnjoin_f64(const double *pts, uint32_t n, uint32_t d, uint32_t k, uint32_t nd, double beta,
                    int nthreads, int32_t *outIds, double *outDists, orc_stats *st);
void orc_bruteforce_f64(const double *pts, uint32_t n, uint32_t d, uint32_t k,
                        const uint32_t *queries, uint32_t nq, int nthreads, int32_t *outIds,
                        double *outDists);

/* eps0 used when the BETA estimate is 0 (bounding-box estimate; see knn_oracle_impl.inc) */
float orc_fallback_eps_f32(const float *mn, const float *mx, uint32_t nd, uint32_t n, uint32_t k1);

/* the epsilon growth rule of main.cu:330-339 (float DTYPE) */
float orc_grow_eps_f32(float eps, float eps0);
/* the FP32 cell coordinate of kernel.cu:1685-1687 (with clamping to [0, ncell-1]) */
uint32_t orc_cell_coord_f32(float x, float mn, float eps, uint32_t ncell);

/* ---- learned radius: MLP 12 -> 64 -> 64 -> 1 and its features ---- */
void orc_mlp_forward(const float *W, const float *X, uint32_t B, float *out);
void orc_mlp_grad(const float *W, const float *X, const float *Y, uint32_t B, float *G, float *loss);
void orc_mlp_train(float *W, const float *X, const float *Y, uint32_t n, uint32_t epochs,
                   uint32_t batch, float lr, uint64_t seed);
void orc_radius_features(const float *pts, uint32_t n, uint32_t d, uint32_t k, uint32_t nd,
                         float *feat, float *wf_out);

#ifdef __cplusplus
}
#endif
#endif
