// knn_launch.h -- host-side launchers of the device kernels (all asynchronous on `s`).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "knn_device.h"

namespace knn {

// Select-kernel bucket counts (knn_range.h, select.hip).
// The select kernel is built for two bucket counts (select.hip is compiled once per count): fewer
// buckets take less LDS, so more workgroups per CU, but stage and rank more items per query.
// The learned schedule's passes run with SEL_NB_LEARNED (40: 14 workgroups per CU), every other
// pass with SEL_NB_OTHER (56: 12 per CU).  Measured on one MI355X, 10M mixed 2-D points (ms per
// join, learned K=16 / K=64 / K=128 / plain K=64): 64 buckets 15.1 / 19.4 / 29.4 / 31.7, 56
// 14.6 / 18.6 / 27.9 / 30.4, 48 14.6 / 18.2 / 27.3 / 31.9, 40 14.1 / 17.5 / 26.9 / 35.3, 32
// 13.8 / 17.8 / 26.2 / 35.6.
#ifndef KNN_SELECT_BUCKETS_LEARNED
#define KNN_SELECT_BUCKETS_LEARNED 40
#endif
#ifndef KNN_SELECT_BUCKETS_OTHER
#define KNN_SELECT_BUCKETS_OTHER 56
#endif
constexpr int SEL_NB_LEARNED = KNN_SELECT_BUCKETS_LEARNED;
constexpr int SEL_NB_OTHER = KNN_SELECT_BUCKETS_OTHER;
// persistent grid of the select kernel with nb buckets: as many one-wave workgroups per CU (256
// CUs) as its LDS admits at K = 64 (11 at 64 buckets, 12 at 56, 14 at 40)
constexpr uint32_t sel_pass_grid(uint32_t nb) { return nb <= 40 ? 3584u : (nb <= 48 ? 3328u : (nb <= 56 ? 3072u : 2816u)); }


// ---------------------------------------------------------------- primitives (prims.hip)
// Exclusive scan of n u32 values into out[0..n); out[n] = total.  `scratch` must hold
// scan_scratch_words(n) u32 words.  in may alias out.
size_t scan_scratch_words(uint32_t n);
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *scratch,
                        hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the key.  Result
// ends in (keys, vals); (keys_tmp, vals_tmp) are ping-pong buffers of n elements; `scratch`
// must hold radix_scratch_words(n) u32 words.
size_t radix_scratch_words(uint32_t n);
void radix_sort_pairs_u32(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                          uint32_t n, int bits, uint32_t *scratch, hipStream_t s);
void radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_tmp, uint32_t *vals_tmp,
                          uint32_t n, int bits, uint32_t *scratch, hipStream_t s);

// Per-dimension min/max of the first nd rows of a dim-major (SoA) array x[j*n + i].
// scratch: minmax_scratch_elems(n, nd) elements of T.  Results (device) in mn[nd], mx[nd].
size_t minmax_scratch_elems(uint32_t n, uint32_t nd);
template <typename T>
void minmax_soa(const T *x, uint32_t n, uint32_t nd, T *mn, T *mx, T *scratch, hipStream_t s);

void iota_u32(uint32_t *a, uint32_t n, hipStream_t s);
void fill_u8(uint8_t *a, uint8_t v, size_t n, hipStream_t s);
void fill_u32(uint32_t *a, uint32_t v, size_t n, hipStream_t s);

// --------------------------------------------------------------- estimator (estimator.hip)
// main.cu:2521-2615: exact sequential variance (T arithmetic) per dim over rows 0,100,...
// of the row-major input.  var[d] (device).
template <typename T>
void variance_rows(const T *pts, uint32_t n, uint32_t d, T *var, hipStream_t s);
// FP64 moments of the same sampled rows, 4 per dim: sum x, sum |x|, sum (x-a)^2, sum (x-a)
// with a = sum x / (n/100); mom holds 4 d doubles (zeroed here)
template <typename T>
void variance_moments(const T *pts, uint32_t n, uint32_t d, double *mom, hipStream_t s);
// out[j*n + i] = pts[i*d + perm[j]]  (row-major -> reordered dim-major)
template <typename T>
void permute_to_soa(const T *pts, uint32_t n, uint32_t d, const uint32_t *perm_host, T *out,
                    hipStream_t s);
// kernel.cu:343-381: the sampled pairwise distance sum in a fixed hierarchical order (64-pair
// segments, per-thread sums, 256-thread chunks, total) into total[0]; scratch holds
// avgdist_scratch_doubles(n) doubles.  avgdist_pairs(n): the number of pairs summed.
size_t avgdist_threads(uint32_t n);
size_t avgdist_scratch_doubles(uint32_t n);
uint64_t avgdist_pairs(uint32_t n);
template <typename T>
void avg_dist_soa(const T *x, uint32_t n, uint32_t d, double *scratch, double *total, hipStream_t s);
// kernel.cu:450-523: k-distance histogram (10000 buckets, +8 per hit) of nsample sample
// points (stride offset) against every 8th 256-point block.  avg/bw are host values.
template <typename T>
void kdist_hist_soa(const T *x, uint32_t n, uint32_t d, uint32_t offset, uint32_t nsample,
                    double avg, double bw, uint32_t *hist, hipStream_t s);

// ------------------------------------------------------------------- index (join.hip)
template <typename T>
void cell_keys_u32(const T *x, uint32_t n, const typename GridOf<T>::type &g, uint32_t *keys,
                   uint32_t *ids, hipStream_t s);
template <typename T>
void cell_keys_u64(const T *x, uint32_t n, const typename GridOf<T>::type &g, uint64_t *keys,
                   uint32_t *ids, hipStream_t s);
// flag[p] = (p == 0 || cell[p] != cell[p-1]), cell = key >> sb (sb: slab-code bits, else 0)
void boundary_flags_u32(const uint32_t *keys, uint32_t n, uint32_t sb, uint32_t *flags, hipStream_t s);
void boundary_flags_u64(const uint64_t *keys, uint32_t n, uint32_t *flags, hipStream_t s);
// cellKeys[pos[p]] = key[p] >> sb, cellStart[pos[p]] = p for boundary p; cellStart[C] = n.
// Keys are widened to u64 in cellKeys.
void compact_cells_u32(const uint32_t *keys, const uint32_t *flags, const uint32_t *pos,
                       uint32_t n, uint32_t sb, uint64_t *cellKeys, uint32_t *cellStart, hipStream_t s);
void compact_cells_u64(const uint64_t *keys, const uint32_t *flags, const uint32_t *pos,
                       uint32_t n, uint64_t *cellKeys, uint32_t *cellStart, hipStream_t s);
// xs[j*n + p] = x[j*n + ids[p]]
template <typename T>
void gather_soa(const T *x, const uint32_t *ids, uint32_t n, uint32_t d, T *xs, hipStream_t s);
// dst[p] = src[idx[p]]
void gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t n, uint32_t *dst, hipStream_t s);
// xs[j*n + p] = in[ids[p]*d + perm[j]] from the row-major input (one cache line per point)
struct DimPerm {
  uint32_t v[KNN_MAXDIM];
};
template <typename T>
void gather_rows(const T *in, const uint32_t *ids, uint32_t n, uint32_t d, const DimPerm &perm, T *xs,
                 hipStream_t s);

// Dense cell-start table (totalCells + 1 entries): denseStart[id] = first sorted position of
// linear cell id (number of points in cells < id).  counts are scattered from the compacted
// cell table, then scanned in place.
void dense_cell_counts(const uint64_t *cellKeys, const uint32_t *cellStart, uint32_t ncell,
                       uint32_t *dense, hipStream_t s);

// Active flags per sorted position with the own-block pre-check: queries whose own cell
// block holds fewer than k1 points get failed[q] = 1 and flag 0.  activeBits (pack_bits of the
// per-id active bytes) null = all active.  cellOf/cellMax (optional) collect per-cell
// candidate counts for accounting.
void pack_bits(const uint8_t *src, uint32_t n, uint64_t *bits, hipStream_t s);
template <typename T>
void prefilter(const T *xs, uint32_t n, const typename GridOf<T>::type &g, const T *radius,
               T rAll, const uint32_t *dense, const uint32_t *sortedIds, const uint64_t *activeBits,
               uint8_t *failed, uint32_t k1, const uint32_t *cellOf, uint32_t *cellMax,
               uint32_t *flags, hipStream_t s);
void cell_bytes(const uint32_t *cellMax, uint32_t ncell, unsigned long long *out, hipStream_t s);
void cell_of(const uint32_t *flags, const uint32_t *pos, uint32_t n, uint32_t *cellOf, hipStream_t s);

// active-query flags per sorted position (all = 1 when failed == NULL)
void active_flags(const uint32_t *sortedIds, const uint8_t *failed, uint32_t n, uint32_t *flags,
                  hipStream_t s);
void compact_active(const uint32_t *sortedIds, const uint32_t *flags, const uint32_t *pos,
                    uint32_t n, uint32_t *act, hipStream_t s);

// Tile blocks: runs of cells in one row (dims 1.. equal) whose dim-0 coordinate falls in the
// same window of `span` cells (blockStart[b] = first sorted position of block b, blockStart[B]
// = n; flags/pos are ncell + 1 scratch words), built with every count read on the device (no
// host round trip; pos is the exclusive scan of the active flags, pos[n] = total): the block
// shape from the active count *activeCount (k_choose_span's cost model: shape[0] = span,
// shape[1] = rows), block flags at that span, and tiles of the *nblocks blocks (ntiles /
// tileOff hold ncell + 1 words; tileOff[ncell] = tile count, read by the range kernels through
// JoinArgs::ntilesDev).
void choose_span(const uint32_t *activeCount, const unsigned long long *activeCells, uint32_t ncell,
                 double avgReach, uint32_t spanMax, uint32_t rowsMax, double minTiles, uint32_t *shape,
                 hipStream_t s);
void block_flags_dev(const uint64_t *cellKeys, uint32_t ncell, uint32_t ncell0, const uint32_t *span,
                     uint32_t *flags, hipStream_t s);
// Row pairs: with shape[1] = 2 (k_choose_span's choice when rowsMax = 2), the tiles of the
// block at (row r, window w) of an even dim-1 coordinate r also take the active queries of the
// block at (r + 1, w), so a tile's queries fill a window of two rows; blockKey[b] = row *
// windows per row + window of block b (compact_blocks_keyed).  partner holds ncell words.
void compact_blocks_keyed(const uint64_t *cellKeys, const uint32_t *cellStart, const uint32_t *flags,
                          const uint32_t *pos, uint32_t ncell, uint32_t ncell0, const uint32_t *shape,
                          uint32_t n, uint32_t *blockStart, uint64_t *blockKey, hipStream_t s);
void pair_tiles_dev(const uint32_t *blockStart, const uint64_t *blockKey, const uint32_t *pos,
                    const uint32_t *nblocks, uint32_t ncell, uint32_t ncell0, uint32_t ncell1,
                    const uint32_t *shape, uint32_t *partner, uint32_t *ntiles, uint32_t *tileOff,
                    uint32_t *scratch, uint4 *tiles, hipStream_t s);

// Tile queues of the persistent range kernels: one per XCD (the chip's 8 L2 domains), each
// over a contiguous eighth of the tiles.
constexpr uint32_t TILE_QUEUES = 8;

struct JoinArgs {
  const void *x;         // reordered dim-major points x[j*n + i]
  const void *xs;        // cell-sorted copy xs[j*n + p]
  const uint32_t *sortedIds;
  const uint64_t *cellKeys;
  const uint32_t *cellStart;
  uint32_t ncell;
  const uint32_t *denseStart;  // dense cell-start table, or null (binary search of cellKeys)
  const uint32_t *act;   // active query ids, grouped by cell
  // tile = up to 64 active queries in at most two runs of act: {start of the second run, start
  // of the first, query count, length of the first run (= count for a single run)}
  const uint4 *tiles;
  uint32_t ntiles;       // tile count; with ntilesDev: the persistent grid size (blocks)
  uint32_t n, d, k1;
  const void *radius;    // per-query search radius (T); null = radiusAll for all
  double radiusAll;      // the one search radius when radius is null (a T value)
  int32_t *outIds;
  void *outDists;
  uint8_t *failed;       // per point: 1 = failed this pass
  uint32_t *failCount;   // device counter (added to)
  unsigned long long *candCount;  // [0] (query, candidate) pairs, [1] algorithmic bytes (added to)
  uint32_t *tileCounter; // TILE_QUEUES dynamic tile queue heads (zeroed before the launch)
  uint32_t tileRanges;   // 1: one queue over all tiles; TILE_QUEUES: one range per XCD
  uint32_t segCapLog2;   // positions per chunk of the segment-start bitmap, log2 (6..11; 11 by default)
  uint32_t selBuckets;   // select kernel: SEL_NB_LEARNED for the learned schedule's passes, else SEL_NB_OTHER (0 = other)
  void *heapScratch;     // heap kernel: global heap storage when K+1 does not fit in LDS
  uint32_t smax;         // select kernel: LDS staging slots per query
  uint8_t *overflow;     // select kernel: per point, 1 = needs the overflow (collect or heap) kernel
  uint32_t *overflowCount;
  // batched joins: per-set grids / K+1 / row offsets (sets[setOf[q]]), and the tile count
  // read from device memory at kernel start (null: single set, `ntiles`)
  const SetDesc *sets;
  const uint32_t *setOf;
  const uint32_t *ntilesDev;
};
// Two-pass select kernel (histogram, then bucket-ordered collect); queries whose candidates
// do not fit in `smax` staging slots are flagged in `overflow` for the overflow kernel.
template <typename T>
void range_select(const JoinArgs &a, const typename GridOf<T>::type &g, hipStream_t s);
template <typename T>
size_t range_select_lds_bytes(uint32_t smax, uint32_t nb);
uint32_t range_select_grid(uint32_t ntiles);
// Collect kernel (wave = query, one pass over its block into an LDS top-(K+1) collector): for
// queries with many candidates per neighbour.  K+1 <= BRUTE_MAX_K1.
template <typename T>
void range_collect(const JoinArgs &a, const typename GridOf<T>::type &g, hipStream_t s);
size_t range_collect_lds_bytes(uint32_t k1);
// Heap kernel (per-lane max-heap on (key, id)): handles any input; the overflow kernel above
// the collector's K+1.
template <typename T>
void range_topk(const JoinArgs &a, const typename GridOf<T>::type &g, hipStream_t s);
// bytes of heap scratch needed in global memory for `ntiles` tiles (0 if LDS is used)
template <typename T>
size_t range_heap_scratch_bytes(uint32_t ntiles, uint32_t k1);

// per-row sums of (T)(d*d) over the K+1 entries of each row in entry order (main.cu:452)
template <typename T>
void row_sums(const T *dists, uint32_t n, uint32_t k1, double *rowSum, hipStream_t s);
// ------------------------------------------------------------- learned radius (learned.hip)
// per-cell point counts of the feature grid: cnt holds reps replica grids of `cells` entries,
// zeroed by the caller; blocks count into the replica of their index mod reps (a clustered set's
// densest cells then take reps shorter chains of atomics), summed into the first grid
template <typename T>
void feature_counts(const T *x, uint32_t n, const typename GridOf<T>::type &g, uint32_t *cnt,
                    uint64_t cells, uint32_t reps, hipStream_t s);
// features (featOut, n x 12, optional) and radius = wf * exp(mlp + margin) (radius, optional),
// with a 64-bin histogram of 4*log2(radius/wf)+32 (rhist, optional, zeroed by the caller)
template <typename T>
void radius_kernel(const T *x, uint32_t n, const typename GridOf<T>::type &g, const uint32_t *cnt,
                   float logk1, float logdens, const float *W, float margin, float wf, T *radius,
                   float *featOut, uint32_t *rhist, hipStream_t s);
// radius octaves of the learned schedule: histogram of the pending queries' quarter-octave
// bins floor(4 log2(r / wf)) + 32 (hist zeroed by the caller, 64 bins); flags of the pending
// queries with a bin in [lo, hi]; radius *= factor for flagged queries that failed
template <typename T>
void radius_bins(const T *radius, const uint8_t *pending, uint32_t n, float wf, uint32_t *hist, hipStream_t s);
template <typename T>
void radius_class_flags(const T *radius, const uint8_t *pending, uint32_t n, float wf, uint32_t lo,
                        uint32_t hi, uint8_t *act, hipStream_t s);
template <typename T>
void grow_class_radius(T *radius, const uint8_t *act, const uint8_t *failed, uint32_t n, T factor,
                       hipStream_t s);
// occupied cells of a count table: *out += number of non-zero entries
void count_nonzero(const uint32_t *cnt, uint64_t cells, unsigned long long *out, hipStream_t s);
template <typename T>
void grow_radius(T *radius, const uint8_t *failed, uint32_t n, T factor, hipStream_t s);
// largest radius among failed queries into out (8 bytes: float bits or double bits)
void max_failed_radius(const float *radius, const uint8_t *failed, uint32_t n, void *out, hipStream_t s);
void max_failed_radius(const double *radius, const uint8_t *failed, uint32_t n, void *out, hipStream_t s);
void gather_rows(const float *in, const uint32_t *idx, uint32_t B, uint32_t w, float *out, hipStream_t s);
void mlp_forward_batch(const float *W, const float *X, uint32_t B, float *H1, float *H2, float *out,
                       hipStream_t s);
void mlp_grad_batch(const float *W, const float *X, const float *Y, uint32_t B, float *H1,
                    float *H2, float *out, float *dout, float *dH1, float *dH2, float *G,
                    float *loss, hipStream_t s);
void adam_step(float *W, const float *G, float *M, float *V, float lr, float b1, float b2, float c1,
               float c2, float eps, hipStream_t s);

// ------------------------------------------------------------- batched sets (batched.hip)
// Group arrays hold the concatenated points of sets with one dimensionality d; sets[si]
// describes set si, setOf[p] the set of point p.  Sizes are read from device memory where a
// count is produced on the device (cells, tiles), so every launch below has a fixed shape.
void batch_variance_perm(const float *xin, const SetDesc *sets, uint32_t nsets, uint32_t d, int reorder,
                         float *var, uint32_t *perm, hipStream_t s);
void batch_permute(const float *xin, uint32_t N, uint32_t d, const uint32_t *setOf, const uint32_t *perm,
                   float *x, hipStream_t s);
void batch_minmax(const float *x, uint32_t N, const SetDesc *sets, uint32_t nsets, float *mnmx,
                  hipStream_t s);
void batch_init_width(SetDesc *sets, uint32_t nsets, const float *mnmx, float *width, uint32_t N,
                      uint8_t *failed, hipStream_t s);
// initial per-query radius from the density around each query on the wf grid (dense table of
// the current index); bin = quarter-octave of radius / wf
void batch_density_radius(const float *x, uint32_t N, const SetDesc *sets, const uint32_t *setOf,
                          const uint32_t *dense, float *radius, uint8_t *bin, hipStream_t s);
// per set: lowest pending octave, the pass's grid width, and the pass's active flags
void batch_class(const SetDesc *sets, uint32_t nsets, const uint32_t *setOf, const uint8_t *failed,
                 const uint8_t *bin, uint32_t N, uint32_t *setMinBin, float *width, uint8_t *act, hipStream_t s);
// radius * 2 for the pass's failed queries; *pending = queries still pending
void batch_grow_class(const SetDesc *sets, const uint32_t *setOf, const uint8_t *act, const uint8_t *failed,
                      uint32_t N, float *radius, uint8_t *bin, uint32_t *pending, hipStream_t s);
void batch_grids(SetDesc *sets, uint32_t nsets, const float *mnmx, float *width, uint64_t bound,
                 uint32_t *status, hipStream_t s);
void batch_cell_keys(const float *x, uint32_t N, const SetDesc *sets, const uint32_t *setOf,
                     uint32_t *keys, uint32_t *ids, hipStream_t s);
void batch_dense_counts(const uint64_t *cellKeys, const uint32_t *cellStart, const uint32_t *ncell,
                        uint32_t N, uint32_t *dense, hipStream_t s);
void batch_flags_u8(const uint32_t *sortedIds, const uint8_t *f, uint32_t N, uint32_t *flags, hipStream_t s);
// tiles of <= tq active queries per cell: ntiles/tileOff hold N + 2 words, tileOff[N + 1] =
// tile count
void batch_tiles(const uint32_t *cellStart, const uint32_t *pos, const uint32_t *ncell, uint32_t N,
                 uint32_t tq, uint32_t *ntiles, uint32_t *tileOff, uint32_t *scratch, uint4 *tiles,
                 hipStream_t s);


// ------------------------------------------------------------- outlier scores (outlier.hip)
// main.cu:952-1067 over an N x k1 table: FP32 row sums of the distances (entry order) and
// their rank (descending), in-degree of every point and its rank (ascending); equal values
// ranked by point id.  keys/vals/keysTmp/valsTmp: n words each; scratch: radix_scratch_words(n).
void outlier_scores(const int32_t *ids, const float *dists, uint32_t n, uint32_t k1, float *distSum,
                    uint32_t *distRank, uint32_t *indeg, uint32_t *indegRank, uint32_t *keys,
                    uint32_t *vals, uint32_t *keysTmp, uint32_t *valsTmp, uint32_t *scratch, hipStream_t s);

// ------------------------------------------------------------- brute force (brute.hip)
// Exact top-(K+1) of a query list against every point, no index (the reference's
// no-selectivity path, GPU.cu:1903-2930 / kernel.cu:902-975).  Rows are bit-identical to the
// range kernels' rows for the same query.
constexpr uint32_t BRUTE_MAX_K1 = 2048;
struct BruteArgs {
  const void *x;            // reordered dim-major points x[j*n + i]
  uint32_t n, d, k1;
  const uint32_t *queries;  // query ids, or null = 0..nq-1
  uint32_t nq;
  uint32_t qw, chunk, nchunks;  // from brute_plan
  double *partKey;          // nq * nchunks * k1 entries each when nchunks > 1
  uint32_t *partId;
  int32_t *outIds;
  void *outDists;
  uint8_t *failed;          // optional: failed[q] = 0 for every row written
};
// qw == 0: not possible (K+1 above BRUTE_MAX_K1 or too many blocks)
struct BrutePlan {
  uint32_t qw, chunk, nchunks;
  uint64_t partialEntries;
};
// partial tables of at most maxPartialEntries entries (a single chunk always fits: no table)
BrutePlan brute_plan(uint32_t n, uint32_t nq, uint32_t k1, uint64_t maxPartialEntries = 1ull << 26);
// list[0..count) = ids i with sel[i] != 0, ascending; count in pos[n].  flags/pos: n + 1 words,
// scratch: scan_scratch_words(n)
void select_list(const uint8_t *sel, uint32_t n, uint32_t *flags, uint32_t *pos, uint32_t *list,
                 uint32_t *scratch, hipStream_t s);
template <typename T>
void brute_join(const BruteArgs &a, hipStream_t s);

// ------------------------------------------------------------- hybrid split (hybrid.hip)
// Host copy of a grid index: cell-sorted coordinates xs[j*n + p], the point id of each sorted
// position, the non-empty cell table.
template <typename T>
struct HybridIndex {
  const T *xs;
  const uint32_t *sortedIds;
  const uint64_t *cellKeys;
  const uint32_t *cellStart;
  uint32_t ncell, n, d;
};
// splitWork's threshold (main.cu:2745-2770): queries of cells with at most this many points
// go to the host
uint32_t hybrid_point_threshold(uint32_t k1, uint32_t nd);
// cpu[p] = 1 for the sorted positions whose own cell holds at most thr points; gpu[q] = 1 for
// the points of the other cells
void hybrid_flags(const uint32_t *cellStart, const uint32_t *cellOf, const uint32_t *sortedIds, uint32_t n,
                  uint32_t thr, uint8_t *cpu, uint8_t *gpu, hipStream_t s);
// exact rows of the points at sorted positions qpos[0..nq) on `threads` host threads (packed:
// row i = the point at qpos[i])
template <typename T>
void hybrid_host_join(const HybridIndex<T> &ix, const typename GridOf<T>::type &g, uint32_t k1,
                      const uint32_t *qpos, uint32_t nq, uint32_t threads, int32_t *outIds, T *outDists);
// packed rows -> table rows of their queries
template <typename T>
void hybrid_scatter(const int32_t *rowIds, const T *rowDists, const uint32_t *list, uint32_t nq, uint32_t k1,
                    int32_t *ids, T *dists, hipStream_t s);

// deterministic sum of rowSum over the rows (chunks of 4096 rows, then chunk sums): out[0]
void sum_rows(const double *rowSum, uint32_t n, double *partial, double *out, hipStream_t s);

}  // namespace knn
