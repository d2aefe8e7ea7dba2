/*
 * gpuknnselfjoin.h -- C ABI of libgpuknnselfjoin.so, an exact K-nearest-neighbour
 * self-join for one AMD MI355X (gfx950), written in HIP.
 *
 * Contract of every join entry point (same as the reference's KNNJoinPy):
 *   input   row-major N x NDIM points (never modified);
 *   output  row i = point i's K+1 nearest points, self included, ascending by distance;
 *           ids int32, Euclidean distances (not squared) in the input's type.
 * Ordering: neighbours are ranked by the FP64 squared distance
 *   key = sum_j ((double)c_j - (double)q_j)^2  (no FMA, dims in descending-variance order)
 * and ties by point id, so rows equal an FP64 brute-force KNN.  Reported float distances are
 *   sqrtf(sum_j (c_j - q_j)^2) in FP32, same dim order, no FMA contraction.
 * Calls are serialised internally (one call at a time per process); buffers reused across
 * calls stay allocated until knnjoin_release() or process exit.
 */
#ifndef GPUKNNSELFJOIN_H
#define GPUKNNSELFJOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference returns void and prints; see knnjoin_last_status) ---- */
enum {
  KNN_OK = 0,
  KNN_ERR_ARG = 1,         /* bad pointer / size / dims (NDIM != expected, NUMINDEXEDDIM > NDIM) */
  KNN_ERR_TOO_FEW = 2,     /* N < K+1: no row can be filled (the reference never terminates) */
  KNN_ERR_EPS_ZERO = 3,    /* epsilon estimate or growth gives 0 (main.cu:334-338) */
  KNN_ERR_NO_PROGRESS = 4, /* eps + eps0/2 == eps in floating point: the loop cannot advance */
  KNN_ERR_HIP = 5,         /* a HIP runtime call failed (message in knnjoin_last_error) */
  KNN_ERR_NOMEM = 6,       /* device allocation failed */
  KNN_ERR_NO_DEVICE = 7,   /* no HIP device visible */
  KNN_ERR_MODEL = 8        /* learned-radius model missing or of the wrong shape */
};

/* dtype tags */
enum { KNN_FLOAT32 = 0, KNN_FLOAT64 = 1 };

/* radius schedule */
enum {
  KNN_MODE_BETA = 0,    /* plain method: BETA-driven eps0 + failed-query loop (main.cu:305-430) */
  KNN_MODE_LEARNED = 1, /* learned per-query initial radius, same loop as fallback */
  KNN_MODE_BRUTE = 2    /* no index: every query against every point (distanceTableNDBruteForce,
                           GPU.cu:1903-2930); K+1 <= 2048 */
};

typedef struct {
  uint32_t num_indexed_dims; /* NUMINDEXEDDIM (params.h:9); 0 = min(NDIM, 6) */
  double beta;               /* BETA (params.h:58); < 0 = 0.05, the reference's default */
  int32_t mode;              /* KNN_MODE_* */
  int32_t reorder;           /* REORDER (params.h:35); 1 = reorder dims by variance */
  int32_t verbose;           /* 1 = print phase timings to stdout */
  int32_t device_ptrs;       /* 1 = points/outputs are device pointers (HBM-resident) */
  void *stream;              /* hipStream_t for device_ptrs mode; NULL = library stream */
  int32_t brute_fallback;    /* brute-force the queries the index cannot serve cheaply: when no
                                indexed dim keeps more than 5 cells at the next radius
                                (checkIndexSelectivity, main.cu:2687-2716), when the failed
                                queries x N x w is at most 7.5e8 + 1000 N, or the whole join
                                when N^2 w <= 2^30, w = max(1, (K+1)/9).  1 on, 0 off (the reference's
                                loop), -1 (default) = on for KNN_MODE_LEARNED, off for
                                KNN_MODE_BETA.  Rows are identical either way. */
  uint32_t hybrid_threads;   /* hybrid CPU/GPU split (splitWork, main.cu:2733-2933), BETA schedule
                                with host pointers: the queries whose own cell of the first grid
                                holds at most ceil((K+1) 2^nd / V_nd) points are joined by this
                                many host threads while the GPU runs the loop over the others;
                                0 (default) = GPU only.  Rows are identical either way. */
  uint32_t hybrid_max_queries; /* at most this many host queries, those of the sparsest cells
                                first (0 = every query the criterion selects) */
} knnjoin_params;

typedef struct {
  uint32_t perm[16];        /* dim order used (column j = input column perm[j]) */
  float eps0;               /* initial epsilon (BETA estimate; 0 in learned mode) */
  float eps_final;          /* epsilon of the last iteration */
  uint32_t iterations;      /* numIter of main.cu:312-430, counting the same-eps reruns */
  uint32_t reindexes;       /* epsilon growths (grid rebuilds) */
  uint32_t kernel_launches; /* range-kernel launches */
  uint32_t failed_final;    /* queries left unfilled (0 on success) */
  uint64_t candidates;      /* (query, candidate) pairs evaluated by the range kernel */
  uint64_t algo_bytes;      /* compulsory HBM bytes of the range kernel launches */
  double checksum;          /* sum over rows of sum_j (float)(d_j*d_j): main.cu:452 */
  double ms_total;          /* device pipeline: points resident -> table resident */
  double ms_reorder, ms_eps, ms_index, ms_join, ms_other;
  double ms_range_kernel;   /* sum of range-kernel durations (HIP events) */
  double ms_range_kernel_max;
  uint64_t algo_bytes_max;  /* algorithmic bytes of the longest range-kernel launch */
  uint64_t overflow_queries; /* queries re-run by the heap kernel (staging overflow) */
  uint32_t eps_bucket;      /* k-distance histogram bucket of eps0 (GPU.cu:3484-3527) */
  uint32_t brute_queries;   /* queries whose rows the brute-force kernel wrote */
  double ms_index_first;    /* the first index build (inside ms_index) */
  double ms_reference_total; /* the reference's totalTime (main.cu:440-445): everything but the
                               first index build, i.e. ms_total - ms_index_first (reorder,
                               epsilon / radius, every pass and re-index, checksum) */
  uint32_t host_syncs;      /* blocking host waits on the device from the first index build on */
  uint32_t hybrid_queries;  /* queries the host threads joined (hybrid split) */
  double ms_hybrid_host;    /* wall time of the host share, concurrent with the GPU loop */
} knnjoin_stats;

/*
 * Replaces main.cu:514 `extern "C" void KNNJoinPy(DTYPE*, unsigned, unsigned, unsigned, int*, DTYPE*)`
 * with GPUNUMDIM=NDIM and NUMINDEXEDDIM=min(NDIM, 6) as runtime values (the reference fixes
 * them at compile time, params.h:5-9, Makefile:61-71), BETA=0.05 (params.h:58), REORDER=1.
 * DTYPE is fixed when the library is built, as in the reference (params.h:5): float in
 * libgpuknnselfjoin.so, double in double/libgpuknnselfjoin.so (`build.py --dtype double`,
 * compiled with KNN_PY_DTYPE_DOUBLE), so the reference wrapper's DTYPE="double" binding
 * (knnjoingpu.py:213-217) loads the double build under the same symbol.
 * Host pointers.  Void like the reference; errors are reported by knnjoin_last_status().
 * Rows of points never resolved (error exits) are left as -1 (main.cu:260-268).
 */
#ifdef KNN_PY_DTYPE_DOUBLE
#define KNN_PY_DTYPE double
#else
#define KNN_PY_DTYPE float
#endif
void KNNJoinPy(KNN_PY_DTYPE *dataset, unsigned int NUMPOINTS, unsigned int kNN, unsigned int NDIM,
               int *outNearestNeighborTable, KNN_PY_DTYPE *outNearestNeighborTableDistances);

/* DTYPE of this build's KNNJoinPy: KNN_FLOAT32 or KNN_FLOAT64. */
int knnjoin_py_dtype(void);

/* KNNJoinPy with DTYPE=double in either build (the wrapper's DTYPE="double",
 * knnjoingpu.py:213-217, against a library whose KNNJoinPy is float). */
void KNNJoinPyDouble(double *dataset, unsigned int NUMPOINTS, unsigned int kNN, unsigned int NDIM,
                     int *outNearestNeighborTable, double *outNearestNeighborTableDistances);

/*
 * Extended entry point: the whole pipeline of main.cu:109-510 / 513-908 with runtime
 * NUMINDEXEDDIM, BETA, REORDER and radius mode, host or device pointers, and a status code.
 * dtype: KNN_FLOAT32 / KNN_FLOAT64.  params/stats may be NULL.
 */
int knnjoin_ex(const void *points, int dtype, uint32_t n, uint32_t ndim, uint32_t k,
               const knnjoin_params *params, int32_t *out_ids, void *out_dists,
               knnjoin_stats *stats);

/*
 * knnjoin_ex without any device allocation: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, rows and status codes as knnjoin_ex, with
 * params->device_ptrs = 1 required (points and outputs in device memory, the caller's).  Every
 * working buffer lives in one static device array of KNN_NOALLOC_SCRATCH_BYTES (256 MiB),
 * which each device the library runs on holds until the process ends.  Limits: K+1 <= 85 and
 * N * (8 NDIM + 129) bytes (float; 16 NDIM + 133 for double) plus about 3 MiB must fit in the
 * array (about 1.8M points of 2 dims); larger inputs return KNN_ERR_ARG / KNN_ERR_NOMEM before
 * any work.  The dense cell-start table and the learned schedule's feature grid are laid out
 * in the same array when they fit (without the dense table the join searches the cell table
 * instead).  The first call on a device looks up the array's address and creates the
 * library's timing events; later calls allocate nothing.  Calls share the device's array:
 * the library serialises its own calls, and two processes each hold their own copy.
 */
#define KNN_NOALLOC_SCRATCH_BYTES (256u << 20)
int knnjoin_ex_noalloc(const void *points, int dtype, uint32_t n, uint32_t ndim, uint32_t k,
                       const knnjoin_params *params, int32_t *out_ids, void *out_dists,
                       knnjoin_stats *stats);

/* Default parameters (reference defaults: NUMINDEXEDDIM=min(NDIM,6), BETA=0.05, REORDER=1). */
void knnjoin_default_params(knnjoin_params *p);

/* Status / message of the last call on this process (KNNJoinPy and friends return void). */
int knnjoin_last_status(void);
const char *knnjoin_last_error(void);

/* Frees the cached device buffers (the reference keeps its pinned buffers for the life of
 * the process, GPU.cu:51-59; cleanUpPinned GPU.cu:3570). */
void knnjoin_release(void);

/* ---------------------------------------------------------------------------------------
 * Stage entry points: the hot path's intermediate results, exposed so that each stage can be
 * checked against the CPU oracle.  All take host pointers, float32 points, row-major.
 * ------------------------------------------------------------------------------------- */

/* ReorderByDimension, main.cu:2521-2615: perm[j] = input column placed at column j. */
int knnjoin_stage_reorder(const float *points, uint32_t n, uint32_t ndim, uint32_t *perm,
                          float *variance);

/* sampleNeighborsBruteForce, GPU.cu:3088-3566 (points already reordered).
 * hist: 10000 buckets or NULL.  Returns eps0 through *eps. */
int knnjoin_stage_estimate_eps(const float *points, uint32_t n, uint32_t ndim, uint32_t k,
                               double beta, float *eps, double *avg_dist, uint64_t *pair_count,
                               uint32_t *bucket, uint32_t *hist);

/* generateNDGridDimensions main.cu:2425-2488 + populateNDGridIndexAndLookupArrayGPU
 * GPU.cu:3715-4067 (points already reordered).  Outputs (caller-sized: n, n, n+1):
 *   sorted_ids  point ids grouped by cell, ascending id within a cell (indexLookupArr),
 *   cell_keys   linear ids of the non-empty cells, ascending (gridCellLookup.gridLinearID),
 *   cell_start  first sorted position of each cell, plus n at the end (grid[c].indexmin).
 * min_arr/n_cells: NUMINDEXEDDIM values.  Returns the non-empty cell count via *num_cells. */
int knnjoin_stage_index(const float *points, uint32_t n, uint32_t ndim, uint32_t nd, float eps,
                        float *min_arr, uint32_t *n_cells, uint64_t *total_cells,
                        uint32_t *sorted_ids, uint64_t *cell_keys, uint32_t *cell_start,
                        uint32_t *num_cells);

/* One range-kernel pass (kernel.cu:61-225 + GPU.cu:3610-3704 + main.cu:1192-1223) at a fixed
 * eps over the query ids given (points already reordered): rows of succeeding queries are
 * written, failed[q] set to 1 for the others.  Returns the failure count via *num_failed. */
int knnjoin_stage_join_pass(const float *points, uint32_t n, uint32_t ndim, uint32_t nd,
                            uint32_t k, float eps, const uint32_t *queries, uint32_t nq,
                            int32_t *out_ids, float *out_dists, uint8_t *failed,
                            uint32_t *num_failed);

/* ---------------------------------------------------------------------------------------
 * Outlier scores of a neighbour table (printOutlierScores, main.cu:952-1067; written by the
 * reference's CLI when PRINTOUTLIERSCORES is set, params.h:22).  ids/dists: n rows of k1
 * entries (host, or device with device_ptrs).  Per point: dist_sum = FP32 sum of its row's
 * distances in entry order (the reference prints dist_sum / (k1 - 1)), dist_rank = its position
 * ordered by dist_sum descending, indegree = rows (its own included) it appears in,
 * indegree_rank = its position ordered by indegree ascending.  The reference's std::sort
 * leaves equal values in unspecified order; here they are ordered by point id.
 * ------------------------------------------------------------------------------------- */
int knnjoin_outlier_scores(const int32_t *ids, const float *dists, uint32_t n, uint32_t k1,
                           int device_ptrs, float *dist_sum, uint32_t *dist_rank, uint32_t *indegree,
                           uint32_t *indegree_rank);

/* ---------------------------------------------------------------------------------------
 * Learned per-query initial radius (no counterpart in the reference, whose only schedule is
 * the BETA estimate + growth loop of main.cu:305-430; used by KNN_MODE_LEARNED).
 * MLP 12 -> 64 -> 64 -> 1, ReLU, FP32; flat weights W1[64][12], b1[64], W2[64][64], b2[64],
 * w3[64], b3 (5057 floats).  Radius = wf * exp(mlp(features) + margin).
 * ------------------------------------------------------------------------------------- */

/* Registers the model used for data with `nd` indexed dims (any NDIM). */
int knnjoin_set_radius_model(uint32_t nd, const float *weights, uint32_t nweights, float margin);

/* Registers the model used for NDIM-dimensional data with `nd` indexed dims; it takes
 * precedence over a model registered for `nd` alone (ndim 0 = that any-NDIM model). */
int knnjoin_set_radius_model_dims(uint32_t ndim, uint32_t nd, const float *weights, uint32_t nweights,
                                  float margin);

/* The 12 features of every point (n x 12, row-major) as the learned schedule computes them
 * (after the dimension reorder when `reorder`), the feature-grid width wf, and (when a model
 * is registered for nd and radius_out != NULL) the radius of every point. */
int knnjoin_radius_features(const float *points, uint32_t n, uint32_t ndim, uint32_t k, uint32_t nd,
                            int reorder, float *features, float *wf_out, float *radius_out);

/* MLP forward over B samples (X: B x 12). */
int knnjoin_mlp_forward(const float *weights, const float *X, uint32_t B, float *out);

/* Gradient of the mean squared error over B samples: grads[5057] in the weight layout. */
int knnjoin_mlp_grad(const float *weights, const float *X, const float *Y, uint32_t B, float *grads,
                     float *loss);

/* Adam training on the GPU (lr, beta1 0.9, beta2 0.999, eps 1e-8), `epochs` passes over a
 * seeded shuffle in mini-batches of `batch`; weights updated in place. */
int knnjoin_mlp_train(float *weights, const float *X, const float *Y, uint32_t n, uint32_t epochs,
                      uint32_t batch, float lr, uint64_t seed, float *loss_per_epoch);

/* ---------------------------------------------------------------------------------------
 * Batched joins: many independent point sets (float32), each joined as knnjoin_ex would
 * join it alone (same rows, same distance bits), processed together: one index and one
 * persistent range-kernel launch per radius pass for all sets of a dimensionality, the whole
 * join captured in a hipGraph and replayed for later calls with the same set shapes.
 * No counterpart in the reference (one set per process, main.cu:109-510); for callers with
 * many small sets, where per-set launches and host round trips dominate.
 * ------------------------------------------------------------------------------------- */
typedef struct {
  uint32_t num_indexed_dims; /* per set: 0 = min(NDIM, 6) */
  int32_t reorder;           /* 1 = reorder each set's dims by variance */
  int32_t device_ptrs;       /* 1 = set points and outputs are device pointers */
  int32_t use_graph;         /* 1 = capture the join in a hipGraph, replay for equal shapes */
  uint32_t graph_passes;     /* radius passes inside the graph (0 = 8); later passes run eagerly */
  uint32_t tile_queries;     /* queries per tile, <= 64 (0 = autotuned value or 16) */
  uint32_t grid_blocks;      /* persistent range-kernel blocks, <= 8192 (0 = autotuned or 4096) */
  int32_t verbose;
} knnjoin_batch_params;

typedef struct {
  double ms_device;          /* device time of the join (graph + eager passes), HIP events */
  double ms_setup;           /* host time before the launch: plan build, capture, input copies */
  uint32_t groups;           /* distinct dimensionalities */
  uint32_t graph_replayed;   /* 1 = an existing graph was replayed */
  uint32_t passes_graph, passes_eager;
  uint32_t tile_queries, grid_blocks;
  uint64_t points, candidates, algo_bytes;
} knnjoin_batch_stats;

void knnjoin_batch_default_params(knnjoin_batch_params *p);

/* Set i: points[i] (row-major n[i] x ndim[i]), k[i]; outputs out_ids[i] (int32, n[i] x (k[i]+1))
 * and out_dists[i] (float32, same shape), the rows knnjoin_ex gives for that set alone. */
int knnjoin_batched(uint32_t nsets, const float *const *points, const uint32_t *n, const uint32_t *ndim,
                    const uint32_t *k, const knnjoin_batch_params *params, int32_t *const *out_ids,
                    float *const *out_dists, knnjoin_batch_stats *stats);

/* Times knnjoin_batched over tile_queries x grid_blocks candidates ({16,32,64} x
 * {1024,2048,4096,8192}) on these sets, keeps the fastest for later calls with the same
 * shapes (parameters left 0), and returns it; ms_table (12 entries) gets each candidate's
 * best device time.  The outputs hold the last run's rows. */
int knnjoin_batch_autotune(uint32_t nsets, const float *const *points, const uint32_t *n,
                           const uint32_t *ndim, const uint32_t *k, const knnjoin_batch_params *params,
                           int32_t *const *out_ids, float *const *out_dists, uint32_t *best_tq,
                           uint32_t *best_grid, double *ms_table, uint32_t *nconfigs);

#ifdef __cplusplus
}
#endif
#endif
