// api.hip -- host orchestration and the C ABI of libgpuknnselfjoin.so.
//
// Pipeline of one join (main.cu:109-510 / KNNJoinPy main.cu:513-908), all on one HIP stream:
//   points (host -> HBM, or already in HBM)
//   -> variance per dim, dim permutation (main.cu:2521-2615)       estimator.hip
//   -> BETA epsilon estimate (GPU.cu:3088-3566)                     estimator.hip
//   -> min/max, grid dims (main.cu:2425-2488)                       prims.hip + host
//   -> index: cell keys, radix sort, cell table (GPU.cu:3715-4067)  join.hip + prims.hip
//   -> loop (main.cu:305-430): active tiles -> range top-(K+1) kernel -> failure count;
//      eps += eps0/2 and re-index when more than 25% failed.
// The host only reads back scalars (variances, the 40 KB k-distance histogram, cell and tile
// counts, the failure count); the table stays in HBM until the caller's copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "gpuknnselfjoin.h"
#include "knn_launch.h"

namespace knn {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::knn::Error(e_ == hipErrorOutOfMemory ? KNN_ERR_NOMEM : KNN_ERR_HIP,           \
                         std::string(#expr) + ": " + hipGetErrorString(e_));                \
  } while (0)

// A kernel launch that could not start (launch configuration, LDS size, a prior fault) is
// reported by the runtime's last-error state, not by the launch itself.
#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

// ------------------------------------------------------------------ buffer cache
enum Slot {
  S_PTS, S_X, S_XS, S_KEYS, S_KEYS_TMP, S_IDS, S_IDS_TMP, S_CELLKEYS, S_CELLSTART, S_FLAGS, S_POS,
  S_ACT, S_NTILES, S_TILEOFF, S_TILES, S_FAILED, S_ROWSUM, S_OUTIDS, S_OUTDISTS, S_SCRATCH,
  S_HEAP, S_SMALL, S_PART, S_CTR, S_CSUM, S_HOSTFLAGS, S_OVERFLOW, S_STAGE, S_DENSE, S_BFLAGS,
  S_BPOS, S_BSTART, S_CELLOF, S_CELLMAX, S_FCNT, S_RADIUS, S_RHIST, S_MODEL, S_TRAIN, S_OCC, S_ACTQ, S_BRUTE_PK, S_BRUTE_PI, S_BLIST, S_TILEMETA, S_HCPU, S_HGPU, S_HLIST, S_HROWI, S_HROWD, S_XS2, S_SIDS, S_SIDS2, S_BKEY, S_PARTNER, S_ABITS, S_NSLOTS
};

// Slots of one no-allocation call: regions of the static device scratch array handed out in
// order (a slot asked for more than its region holds gets a new region; the call's regions
// are all released when it ends).
struct Arena {
  unsigned char *base = nullptr;
  size_t cap = 0, used = 0;
  void *p[S_NSLOTS] = {};
  size_t sz[S_NSLOTS] = {};
  static size_t round(size_t b) { return ((b ? b : 16) + 255) & ~size_t(255); }
  // bytes of free scratch
  size_t room() const { return cap - used; }
  // scratch bytes that slot s would take to hold `bytes` (0 when its region holds them)
  size_t extra(Slot s, size_t bytes) const { return sz[s] >= bytes && p[s] ? 0 : round(bytes); }
  void *get(Slot s, size_t bytes) {
    if (sz[s] >= bytes && p[s]) return p[s];
    const size_t b = round(bytes);
    if (b > room())
      throw Error(KNN_ERR_NOMEM, "no-allocation scratch exhausted: " + std::to_string(used) + " of " +
                                     std::to_string(cap) + " bytes in use, " + std::to_string(b) + " more needed");
    p[s] = base + used;
    sz[s] = b;
    used += b;
    return p[s];
  }
};
// set for the duration of a no-allocation call (calls are serialised by g_mu)
static Arena *g_arena = nullptr;

struct Cache {
  void *p[S_NSLOTS] = {};
  size_t cap[S_NSLOTS] = {};
  void *get(Slot s, size_t bytes) {
    if (g_arena) return g_arena->get(s, bytes);
    if (bytes == 0) bytes = 16;
    if (cap[s] < bytes) {
      if (p[s]) HIP_CHECK(hipFree(p[s]));
      p[s] = nullptr;
      cap[s] = 0;
      size_t b = (bytes + 4095) & ~size_t(4095);
      HIP_CHECK(hipMalloc(&p[s], b));
      cap[s] = b;
    }
    return p[s];
  }
  void release() {
    for (int i = 0; i < S_NSLOTS; i++) {
      if (p[i]) (void)hipFree(p[i]);
      p[i] = nullptr;
      cap[i] = 0;
    }
  }
};

static std::mutex g_mu;
static Cache g_cache;
static void release_batch();
static hipStream_t g_stream = nullptr;
static int g_status = KNN_OK;
static std::string g_err;

static hipStream_t lib_stream() {
  if (!g_stream) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      throw Error(KNN_ERR_NO_DEVICE, "no HIP device visible");
    HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  }
  return g_stream;
}

// ------------------------------------------------------------------ host-side numerics
// main.cu:330-339: eps_est += (eps_est_initial*0.5), DTYPE eps, double literal.
static inline float grow_eps(float eps, float eps0) { return (float)((double)eps + (double)eps0 * 0.5); }
static inline double grow_eps(double eps, double eps0) { return eps + eps0 * 0.5; }

template <typename T>
static void grid_dims(const T *mn, const T *mx, uint32_t nd, T eps, typename GridOf<T>::type &g,
                      uint64_t &totalCells) {
  // main.cu:2465-2486
  g.eps = eps;
  g.nd = nd;
  uint64_t tot = 1;
  for (uint32_t j = 0; j < nd; j++) {
    T lo = mn[j] - eps;
    T hi = mx[j] + eps;
    g.minArr[j] = lo;
    T q = (hi - lo) / eps;
    T c = std::ceil(q);
    if (!(c >= (T)1) || c > (T)4294967295.0)
      throw Error(KNN_ERR_ARG, "grid: cell count per dimension out of range (eps too small)");
    g.nCells[j] = (uint32_t)c;
    g.mult[j] = tot;
    if (tot > UINT64_MAX / g.nCells[j]) throw Error(KNN_ERR_ARG, "grid: more than 2^64 cells");
    tot *= g.nCells[j];
  }
  for (uint32_t j = nd; j < KNN_MAX_INDEXED; j++) {
    g.minArr[j] = 0;
    g.nCells[j] = 1;
    g.mult[j] = tot;
  }
  g.keyBase = 0;
  g.smin = 0;
  g.sscale = 0;
  g.sb = 0;
  totalCells = tot;
}

// eps0 when the BETA estimate is 0.  The reference divides by zero then (N <= 1000: its
// distance sample is point 0 against itself; or all sampled pairs coincide).  Any eps > 0
// gives the same exact rows; take the radius holding K+1 points of a uniform set over the
// bounding box: largest indexed extent * (K1/N)^(1/nd), or 1 when all points coincide.
template <typename T>
static T fallback_eps(const T *mn, const T *mx, uint32_t nd, uint32_t n, uint32_t k1) {
  T ext = 0;
  for (uint32_t j = 0; j < nd; j++) {
    T e = mx[j] - mn[j];
    if (e > ext) ext = e;
  }
  T eps = (T)((double)ext * std::pow((double)k1 / (double)n, 1.0 / (double)nd));
  return eps > (T)0 ? eps : (T)1;
}

static int bits_for(uint64_t totalCells) {
  uint64_t m = totalCells ? totalCells - 1 : 0;
  int b = 0;
  while (m) {
    b++;
    m >>= 1;
  }
  return b;
}

// ------------------------------------------------------------------ learned-radius model
struct RadiusModel {
  std::vector<float> w;
  float margin = 0.0f;
  bool ok = false;
};
// by (NDIM, indexed dims); NDIM 0 = a model for any NDIM with that many indexed dims
static RadiusModel g_models[KNN_MAXDIM + 1][KNN_MAX_INDEXED + 1];
static const RadiusModel &radius_model(uint32_t d, uint32_t nd) {
  const RadiusModel &m = g_models[d <= KNN_MAXDIM ? d : 0][nd];
  return m.ok ? m : g_models[0][nd];
}

// Feature grid of the learned radius: width wf = (V (K+1) / N)^(1/nd), V = product of the
// indexed extents (an extent of 0 counts as 1), doubled until the grid has <= 2^27 cells.
template <typename T>
struct FeatureGrid {
  typename GridOf<T>::type g;
  uint64_t cells;
  T wf;
  float logk1, logdens;
};

template <typename T>
static void feature_grid(const T *mn, const T *mx, uint32_t nd, uint32_t n, uint32_t k1,
                         FeatureGrid<T> &F) {
  double V = 1.0;
  for (uint32_t j = 0; j < nd; j++) {
    const double e = (double)(mx[j] - mn[j]);
    V *= e > 0.0 ? e : 1.0;
  }
  T wf = (T)std::pow(V * (double)k1 / (double)n, 1.0 / (double)nd);
  if (!(wf > (T)0)) wf = (T)1;
  for (;;) {
    grid_dims<T>(mn, mx, nd, wf, F.g, F.cells);
    if (F.cells <= (1ull << 27)) break;
    wf = wf * (T)2;
  }
  F.wf = wf;
  F.logk1 = logf((float)k1);
  F.logdens = 0.0f;  // set by feature_occupancy once the cells are counted
}

// f11 = log(N / occupied feature cells) from the per-cell counts (one host read)
static void feature_occupancy(const uint32_t *cnt, uint64_t cells, uint32_t n, float &logdens,
                              hipStream_t s) {
  unsigned long long *d = reinterpret_cast<unsigned long long *>(g_cache.get(S_OCC, 64));
  fill_u32(reinterpret_cast<uint32_t *>(d), 0u, 2, s);
  count_nonzero(cnt, cells, d, s);
  unsigned long long occ = 0;
  HIP_CHECK(hipMemcpyAsync(&occ, d, sizeof(occ), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  logdens = (float)std::log((double)n / (double)(occ ? occ : 1ull));
}

// ------------------------------------------------------------------ pipeline
struct PassResult {
  uint32_t nfail, ntiles, noverflow;
  double kernel_ms, pass_ms;
  unsigned long long bytes_total, pairs;
};

// Event pairs created once (the first call that times anything) and reused by every call:
// pair i serves the i-th timer of a call (calls are serialised by g_mu).
constexpr int kTimerPairs = 8;
static hipEvent_t g_events[kTimerPairs][2];
static bool g_eventsReady = false;

struct Timer {
  hipEvent_t a, b;
  hipStream_t s;
  Timer(hipStream_t st, int pair) : s(st) {
    if (!g_eventsReady) {
      for (int i = 0; i < kTimerPairs; i++)
        for (int j = 0; j < 2; j++) HIP_CHECK(hipEventCreate(&g_events[i][j]));
      g_eventsReady = true;
    }
    a = g_events[pair][0];
    b = g_events[pair][1];
  }
  void start() { HIP_CHECK(hipEventRecord(a, s)); }
  double stop() {
    HIP_CHECK(hipEventRecord(b, s));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
  // records the end without waiting; elapsed() once a later wait has covered it
  void stop_async() { HIP_CHECK(hipEventRecord(b, s)); }
  double elapsed() const {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
};

// Pinned host words for the values the host reads back during a join (cell, active, failure
// and overflow counts, accounting counters, radius histograms): allocated once, so that those
// copies are asynchronous and one stream wait per pass collects them all.
enum PinWord { PW_NCELL = 0, PW_KFAIL = 1, PW_OVF = 2, PW_ACTIVE = 3, PW_SPAN = 4, PW_CTR = 8, PW_HIST = 64, PW_MAXR = 128 };
static uint32_t *g_pin = nullptr;
static uint32_t *pinned() {
  if (!g_pin) HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&g_pin), 4096, hipHostMallocDefault));
  return g_pin;
}

template <typename T>
struct Pipeline {
  using G = typename GridOf<T>::type;
  hipStream_t s;
  uint32_t n, d, nd, k, K1;
  T *x = nullptr;  // reordered dim-major
  T *xs = nullptr; // cell-sorted dim-major
  uint32_t *sortedIds = nullptr;
  uint64_t *cellKeys = nullptr;
  uint32_t *cellStart = nullptr;
  uint32_t ncell = 0;
  G g{};
  uint64_t totalCells = 0;
  uint32_t *scratch = nullptr;  // u32 words
  size_t scratchWords = 0;
  bool verbose = false;
  bool brute = false;  // brute-force fallback enabled
  const T *rowMajor = nullptr;  // the row-major input, when x is its column permutation rowPerm
  DimPerm rowPerm{};

  template <typename V>
  V *buf(Slot sl, size_t count) { return reinterpret_cast<V *>(g_cache.get(sl, sizeof(V) * count)); }

  // populateNDGridIndexAndLookupArrayGPU (GPU.cu:3715-4067), all in HBM.
  // The first build of a join sorts the points (original ids) by cell; a later build (a
  // re-index at another cell width) computes its keys from the previous build's cell-sorted
  // copy and sorts those positions, so its gathers of coordinates and ids read a spatially
  // ordered array (nearby positions, cached lines) instead of the input's random order.
  // Within a cell the points then follow the previous order; every row and count is the same
  // (rows are ranked by (key, id)).  The no-allocation path keeps one copy (original order).
  void build_index(const T *mn, const T *mx, T eps) {
    grid_dims<T>(mn, mx, nd, eps, g, totalCells);
    const int bits = bits_for(totalCells);
    // slab order within cells (u32 keys only: the slab code takes the key bits the cell id
    // leaves, up to 16, and at least 8)
    if (slab && bits <= 24 && slabMax > slabMin) {
      g.sb = (uint32_t)std::min(16, 32 - bits);
      g.smin = slabMin;
      g.sscale = (T)(std::ldexp(1.0, (int)g.sb) / ((double)slabMax - (double)slabMin));
    }
    const bool fromPrev = xs != nullptr && sortedIds != nullptr && !g_arena;
    const T *keySrc = fromPrev ? xs : x;
    // the sort's payload: original ids (first build) or previous positions (later builds)
    uint32_t *ids = buf<uint32_t>(fromPrev ? S_IDS : S_SIDS, n);
    uint32_t *ids_tmp = buf<uint32_t>(S_IDS_TMP, n);
    if (bits <= 32) {
      uint32_t *keys = buf<uint32_t>(S_KEYS, n);
      uint32_t *kt = buf<uint32_t>(S_KEYS_TMP, n);
      cell_keys_u32<T>(keySrc, n, g, keys, ids, s);
      radix_sort_pairs_u32(keys, ids, kt, ids_tmp, n, bits + (int)g.sb, scratch, s);
      uint32_t *flags = buf<uint32_t>(S_FLAGS, n + 1);
      uint32_t *pos = buf<uint32_t>(S_POS, n + 1);
      boundary_flags_u32(keys, n, g.sb, flags, s);
      scan_exclusive_u32(flags, pos, n, scratch, s);
      cellKeys = buf<uint64_t>(S_CELLKEYS, n);
      cellStart = buf<uint32_t>(S_CELLSTART, n + 1);
      compact_cells_u32(keys, flags, pos, n, g.sb, cellKeys, cellStart, s);
      cellOf = buf<uint32_t>(S_CELLOF, n);
      cell_of(flags, pos, n, cellOf, s);
      ncellDev = pos + n;
    } else {
      uint64_t *keys = buf<uint64_t>(S_KEYS, n);
      uint64_t *kt = buf<uint64_t>(S_KEYS_TMP, n);
      cell_keys_u64<T>(keySrc, n, g, keys, ids, s);
      radix_sort_pairs_u64(keys, ids, kt, ids_tmp, n, bits, scratch, s);
      uint32_t *flags = buf<uint32_t>(S_FLAGS, n + 1);
      uint32_t *pos = buf<uint32_t>(S_POS, n + 1);
      boundary_flags_u64(keys, n, flags, s);
      scan_exclusive_u32(flags, pos, n, scratch, s);
      cellKeys = buf<uint64_t>(S_CELLKEYS, n);
      cellStart = buf<uint32_t>(S_CELLSTART, n + 1);
      compact_cells_u64(keys, flags, pos, n, cellKeys, cellStart, s);
      cellOf = buf<uint32_t>(S_CELLOF, n);
      cell_of(flags, pos, n, cellOf, s);
      ncellDev = pos + n;
    }
    if (fromPrev) {
      const bool second = xsSlot == S_XS;
      T *xsNew = buf<T>(second ? S_XS2 : S_XS, (size_t)n * d);
      uint32_t *sidNew = buf<uint32_t>(second ? S_SIDS2 : S_SIDS, n);
      gather_soa<T>(xs, ids, n, d, xsNew, s);
      gather_u32(sortedIds, ids, n, sidNew, s);
      xs = xsNew;
      sortedIds = sidNew;
      xsSlot = second ? S_XS2 : S_XS;
    } else {
      sortedIds = ids;
      xs = buf<T>(S_XS, (size_t)n * d);
      xsSlot = S_XS;
      if (rowMajor)
        gather_rows<T>(rowMajor, sortedIds, n, d, rowPerm, xs, s);
      else
        gather_soa<T>(x, sortedIds, n, d, xs, s);
    }
    LAUNCH_CHECK();
    // dense cell-start table: O(1) row lookups when the grid is small enough (its counts are
    // scattered for the cells below the device-side cell count, so no host round trip here)
    denseStart = nullptr;
    // with the no-allocation scratch: only when the table, its scan scratch and the pass
    // buffers still to come (active list, tiles, block tables, flags) fit
    bool denseFits = totalCells + 1 <= kDenseMaxCells;
    if (denseFits && g_arena) {
      const size_t tcb = (size_t)totalCells + 1;
      const size_t more = g_arena->extra(S_DENSE, tcb * 4) +
                          g_arena->extra(S_SCRATCH, (scan_scratch_words((uint32_t)tcb) + 1024) * 4);
      denseFits = more + 48ull * n + (64u << 10) <= g_arena->room();
    }
    if (denseFits) {
      const uint32_t tc = (uint32_t)totalCells;
      const size_t need = scan_scratch_words(tc + 1) + 1024;
      if (need > scratchWords) {
        scratchWords = need;
        scratch = buf<uint32_t>(S_SCRATCH, scratchWords);
      }
      uint32_t *dense = buf<uint32_t>(S_DENSE, (size_t)tc + 1);
      fill_u32(dense, 0u, (size_t)tc + 1, s);
      batch_dense_counts(cellKeys, cellStart, ncellDev, n, dense, s);
      scan_exclusive_u32(dense, dense, tc, scratch, s);
      LAUNCH_CHECK();
      denseStart = dense;
    }
    // the build's one wait: the cell count, for the launch sizes of the passes
    uint32_t *pin = pinned();
    HIP_CHECK(hipMemcpyAsync(pin + PW_NCELL, ncellDev, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    if (tBuild) tBuild->stop_async();
    sync();
    ncell = pin[PW_NCELL];
  }
  // slab order (fewer indexed than data dims): on when index_minmax has found the range of
  // the first non-indexed dim; the stage entry points leave it off
  bool slab = false;
  T slabMin = 0, slabMax = 0;
  // min/max of the indexed dims into hmn/hmx (nd values each), and of the first non-indexed
  // dim for the slab order (KNN_SLAB=0 turns the slab order off)
  void index_minmax(T *hmn, T *hmx) {
    const char *ev = std::getenv("KNN_SLAB");
    const bool want = nd < d && nd < KNN_MAX_INDEXED && !(ev && ev[0] == '0');
    const uint32_t nde = want ? nd + 1 : nd;
    T *mm = buf<T>(S_PART, 2 * KNN_MAX_INDEXED + minmax_scratch_elems(n, nde));
    minmax_soa<T>(x, n, nde, mm, mm + KNN_MAX_INDEXED, mm + 2 * KNN_MAX_INDEXED, s);
    T lo[KNN_MAX_INDEXED], hi[KNN_MAX_INDEXED];
    HIP_CHECK(hipMemcpyAsync(lo, mm, sizeof(T) * nde, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(hi, mm + KNN_MAX_INDEXED, sizeof(T) * nde, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(hmn, lo, sizeof(T) * nd);
    std::memcpy(hmx, hi, sizeof(T) * nd);
    slab = want;
    slabMin = want ? lo[nd] : (T)0;
    slabMax = want ? hi[nd] : (T)0;
  }
  const uint32_t *ncellDev = nullptr;  // device copy of ncell (the scan total of the build)
  Slot xsSlot = S_XS;                   // the slot holding the current cell-sorted copy
  Timer *tBuild = nullptr;              // stopped before the build's wait when set
  uint32_t syncs = 0;                   // stream waits of the passes and builds
  void sync() {
    HIP_CHECK(hipStreamSynchronize(s));
    syncs++;
  }

  // 1 GiB of cell starts at most: filled and scanned once per build (~0.6 ms at 2^28 cells),
  // against a binary search of the cell table for every row of every block without it
  static constexpr uint64_t kDenseMaxCells = 1ull << 28;
  // tiles a pass should have before the cost model widens its blocks (2816 persistent waves;
  // KNN_MIN_TILES overrides it for tuning: 2816, 4096, 8192 and 16384 are within 1 % of each
  // other on the 10M mixed 2-D joins, K = 16 to 128)
  static double min_tiles() {
    const char *v = std::getenv("KNN_MIN_TILES");
    const double t = v ? std::atof(v) : 8192.0;
    return t >= 1.0 ? t : 8192.0;
  }
  const uint32_t *denseStart = nullptr;
  uint32_t *cellOf = nullptr;  // compacted cell index of each sorted position

  // One range pass (kernel.cu:61-225 + GPU.cu:3610-3704 + main.cu:1192-1223): queries with
  // activeFailed[q] != 0 (all when null), grouped by cell into tiles of <= 64.
  // `beforeWait` (optional) is queued after the range kernel and before the pass's wait, to
  // have work whose results the host reads at that same wait; when staging overflows send
  // queries to the heap kernel (a second wait), `afterHeap` is queued after that kernel.
  PassResult pass(const uint8_t *activeFailed, uint32_t nq, uint8_t *failed, int32_t *dIds,
                  T *dDists, unsigned long long *ctr, Timer &tK, Timer &tPass,
                  const std::function<void()> &beforeWait = {}, const std::function<void()> &afterHeap = {});
  T uniform_radius() const { return radiusAll > (T)0 ? radiusAll : (T)g.eps; }
  // tiles of the active queries, every count kept on the device; the tile count ends in
  // *ntilesDev (the range kernels read it there)
  void make_tiles(const uint8_t *activeFailed, uint32_t *&act, uint4 *&tiles, const uint32_t *&ntilesDev,
                  uint8_t *failed, unsigned long long *cellAcc);
  unsigned long long accBytes = 0;  // accumulated pre-check accounting bytes
  const T *radius = nullptr;  // per-query radii (learned schedule), null = radiusAll
  T radiusAll = (T)0;         // one radius for every query (plain schedule); 0 = the grid's eps
  uint32_t lastSpan = 1;
  uint32_t lastRows = 1;
  uint32_t selBuckets = (uint32_t)SEL_NB_OTHER;  // the select kernel's bucket count for the passes
  uint32_t lastActive = 0;   // queries left in tiles after the pre-check
  double avgReach = 1.0;     // typical search radius in cells (tile-span cost model)
  // persistent range-kernel workgroups: 11 per CU on 256 CUs (the select kernel's LDS,
  // about 14 KB per one-wave workgroup, admits 11 per CU).  The kernel is latency-bound, so
  // residency pays: 10M mixed 2-D points, K = 64, learned / plain schedule: 2048 -> 34.7 /
  // 60.6 ms per join, 2560 -> 31.1 / 54.9 ms, 2816 -> 30.5 / 53.6 ms; 32 buckets with up to
  // 16 waves per CU: 30.0 / 57.5 ms (coarser buckets stage and rank more in the plain schedule).
#ifndef KNN_PASS_GRID
#define KNN_PASS_GRID 2816
#endif
  static constexpr uint32_t kPassGrid = KNN_PASS_GRID;
  static constexpr uint32_t kHeapGrid = 512;   // persistent heap-kernel workgroups
  const uint32_t *activeDev = nullptr;         // active count of the last tile build (device)
  const uint32_t *spanDev = nullptr;           // its tile-block span (device)
};

template <typename T>
static void sort_perm_desc(const T *var, uint32_t d, uint32_t *perm) {
  // std::sort with compareByDimVariance (main.cu:44-47, 2580) on <= 16 elements is
  // libstdc++'s insertion sort: stable, descending
  for (uint32_t i = 0; i < d; i++) perm[i] = i;
  for (uint32_t i = 1; i < d; i++) {
    uint32_t p = perm[i];
    T key = var[p];
    int j = (int)i - 1;
    while (j >= 0 && key > var[perm[j]]) {
      perm[j + 1] = perm[j];
      j--;
    }
    perm[j + 1] = p;
  }
}

// Interval of the variance that main.cu:2537-2562 computes in T arithmetic (sequential sums
// over the m sampled rows, then / (n * T(1/100))), from exact-ish FP64 moments: with u the
// unit roundoff of T and g = (m + 4) u (error bound of recursive summation), a = S / (n/100):
//   |avg_T - a| <= da = (g + 4u) * sum|x| / (n/100),
//   sum (x - avg_T)^2 = Q + 2 (a - avg_T) C + m (a - avg_T)^2, Q = sum (x - a)^2, C = sum (x - a),
//   var_T in [(Q - 2 da |C|)(1 - g)(1 - 4u), (Q + 2 da |C| + m da^2)(1 + g)(1 + 4u)] / (n/100).
// When the intervals of all dims are pairwise disjoint, sorting the FP64 centres descending
// gives the order the T values give (strictly ordered, so the stable sort's tie rule is moot).
template <typename T>
static bool order_from_intervals(const double *mom, uint32_t n, uint32_t d, uint32_t *perm) {
  const double u = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
  const double m = (double)(((uint64_t)n + 99) / 100);
  const double g = (m + 4.0) * u * 1.01;
  const double den = (double)n / 100.0;
  double lo[KNN_MAXDIM], hi[KNN_MAXDIM], mid[KNN_MAXDIM];
  for (uint32_t i = 0; i < d; i++) {
    const double A = mom[4 * i + 1], Q = mom[4 * i + 2], C = mom[4 * i + 3];
    const double da = (g + 4.0 * u) * A / den;
    const double slack = 1e-9 * (Q + A * A / m);  // FP64 rounding of the moments themselves
    lo[i] = (Q - 2.0 * da * std::fabs(C) - slack) * (1.0 - g) * (1.0 - 4.0 * u) / den;
    hi[i] = (Q + 2.0 * da * std::fabs(C) + m * da * da + slack) * (1.0 + g) * (1.0 + 4.0 * u) / den;
    mid[i] = Q / den;
    if (!(lo[i] == lo[i]) || !(hi[i] == hi[i])) return false;
  }
  for (uint32_t i = 0; i < d; i++)
    for (uint32_t j = i + 1; j < d; j++)
      if (!(lo[i] > hi[j] || lo[j] > hi[i])) return false;
  for (uint32_t i = 0; i < d; i++) perm[i] = i;
  std::stable_sort(perm, perm + d, [&](uint32_t a, uint32_t b) { return mid[a] > mid[b]; });
  return true;
}

struct EpsResult {
  double avg;
  uint64_t cnt;
  double bw;
  uint32_t offset, nsample, bucket;
  double best;
};

// sampleNeighborsBruteForce GPU.cu:3088-3566 on the reordered dim-major points
template <typename T>
static EpsResult estimate_eps(const T *x, uint32_t n, uint32_t d, uint32_t k, double beta,
                              uint32_t *hist_out, hipStream_t s) {
  EpsResult r{};
  const size_t m = avgdist_scratch_doubles(n);
  double *scr = reinterpret_cast<double *>(g_cache.get(S_PART, m * sizeof(double) + 64));
  double *tot = scr + m;
  avg_dist_soa<T>(x, n, d, scr, tot, s);
  double h_tot = 0;
  const uint64_t h_cnt = avgdist_pairs(n);
  HIP_CHECK(hipMemcpyAsync(&h_tot, tot, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  r.avg = h_cnt ? h_tot / (h_cnt * 1.0) : 0.0;  // GPU.cu:3334
  r.cnt = h_cnt;
  const uint32_t nb = 10000;
  r.bw = r.avg / (nb * 1.0);
  double offsetRate = 0.01;
  if (n >= 50000) {
    double frac = (50000.0 / n) * offsetRate;
    r.offset = (uint32_t)(1.0 / frac);
  } else {
    r.offset = (uint32_t)(1.0 / offsetRate);
  }
  r.nsample = n / r.offset;
  uint32_t *hist = reinterpret_cast<uint32_t *>(g_cache.get(S_SMALL, 64 * 1024));
  kdist_hist_soa<T>(x, n, d, r.offset, r.nsample, r.avg, r.bw, hist, s);
  std::vector<uint32_t> h(nb);
  HIP_CHECK(hipMemcpyAsync(h.data(), hist, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  // GPU.cu:3484-3527
  uint32_t cumulative = 0;
  int found = -1;
  double best = 0;
  unsigned long min_pts = (unsigned long)(k * 1.0);
  unsigned long max_pts = (unsigned long)(k * 100.0);
  unsigned long thr = (unsigned long)(min_pts + (max_pts - min_pts) * beta);
  for (uint32_t i = 0; i < nb; i++) {
    cumulative += h[i];
    double per_point = cumulative / (1.0 * r.nsample);
    if (per_point > thr && found == -1) {
      best = r.bw * (i + 1.0);
      found = (int)i;
    }
    if (i == nb - 1 && found == -1) {
      best = r.bw * (i + 1.0);
      found = (int)i;
    }
  }
  r.bucket = (uint32_t)found;
  r.best = best;
  if (hist_out) std::memcpy(hist_out, h.data(), nb * sizeof(uint32_t));
  return r;
}


// Range kernel of a pass: the collect kernel (one pass over each query's own block) with 5 or
// more indexed dims, else the two-pass select kernel, whose first pass reads each candidate
// once for all queries of a tile.  Measured on MI355X, 5M points, K = 32 (learned / plain
// schedule, ms): 6 indexed dims, uniform 225 / 264 against 317 / 302, skewed 339 / 409
// against 567 / 582; 4 indexed dims, uniform 63 / 98 against 59 / 66; 2 of 4 dims indexed
// 2475 against 1504 (learned): queries share most of their candidates with the rest of their
// tile there.  KNN_RANGE_KERNEL=select|collect overrides the choice for tuning.
static bool use_collect(uint32_t nd, uint32_t k1) {
  const char *v = std::getenv("KNN_RANGE_KERNEL");  // read per pass: tests switch it in-process
  const int env = !v ? -1 : (std::strcmp(v, "collect") == 0 ? 1 : (std::strcmp(v, "select") == 0 ? 0 : -1));
  if (k1 > BRUTE_MAX_K1) return false;
  return env >= 0 ? env == 1 : nd >= 5;
}

// Kernel for the queries the select kernel cannot stage (staging overflow, or K+1 beyond its
// LDS slots): the collect kernel (wave = query, LDS collector) up to K+1 = BRUTE_MAX_K1, the
// per-lane heap kernel above.  KNN_OVERFLOW_KERNEL=heap|collect overrides it for tuning.
static bool overflow_by_collect(uint32_t k1) {
  const char *v = std::getenv("KNN_OVERFLOW_KERNEL");
  if (k1 > BRUTE_MAX_K1) return false;
  return !(v && std::strcmp(v, "heap") == 0);
}

// Positions per chunk of the range kernels' segment-start bitmap, log2: 11 (the whole 2048-bit
// bitmap); KNN_SEG_CHUNK_LOG2=6..10 makes the chunks smaller so that the tests exercise the
// chunk boundaries on small inputs (same rows, more chunks).
static uint32_t seg_cap_log2() {
  const char *v = std::getenv("KNN_SEG_CHUNK_LOG2");
  const int c = v ? std::atoi(v) : 11;
  return (c >= 6 && c <= 11) ? (uint32_t)c : 11u;
}

// Pending-query fraction (1/div of N) at or below which the learned schedule keeps the current
// grid for a pass; KNN_REUSE_TAIL=div overrides the default 64 for tuning.  A larger fraction
// also keeps the first grid for the second pass (10M mixed 2-D points, 550k queries at K=64):
// its wider blocks then cost what the skipped build saves (ms per join, div 64 against 8:
// K=16 16.65 / 17.18, K=64 19.98 / 20.12, K=128 30.09 / 30.00).
static uint64_t reuse_tail_div() {
  const char *v = std::getenv("KNN_REUSE_TAIL");
  const long c = v ? std::atol(v) : 64;
  return c >= 1 ? (uint64_t)c : 64u;
}

// Rows of a tile block with 2 indexed dims: up to 2 (k_choose_span picks 1 or 2 per pass from
// the active density); KNN_TILE_ROWS=1 keeps single-row blocks, for tuning.  Measured on one
// MI355X (ms per join, single rows against row pairs, same rows out): 10M mixed 2-D points,
// K=64, plain 34.90 / 32.45, learned 20.00 / 20.01 (K=16 16.69 / 16.59, K=128 30.10 / 30.28);
// with 4 indexed dims (5M uniform 4-D, K=32, plain) 25.8 / 26.2 with 15 % more candidates, as
// the rows' chords narrow less for a union of two rows: above 2 indexed dims blocks keep one.
static uint32_t tile_rows_max() {
  const char *v = std::getenv("KNN_TILE_ROWS");  // read per pass: tuning runs switch it in-process
  return (v && std::atoi(v) == 1) ? 1u : 2u;
}

// Density basis of the tile-block cost model: active queries per cell holding any (default),
// or KNN_SPAN_DENSITY=cells for active queries per non-empty cell, for tuning.  Clustered
// active queries fill a block's 64 query slots over fewer cells than the average over all
// non-empty cells says, and further tiles of the block rescan the same union.  Measured on one
// MI355X with row pairs (ms per join, per non-empty cell against per active cell, same rows
// out): 10M mixed 2-D points, learned K=16 16.90 / 15.55, K=64 20.22 / 19.92, K=128 30.50 /
// 29.89, plain K=64 32.54 / 32.07; 5M points, K=32, 3 and 4 indexed dims: equal candidates.
static bool span_by_active_cells() {
  const char *v = std::getenv("KNN_SPAN_DENSITY");  // read per pass: tuning runs switch it in-process
  return !(v && std::strcmp(v, "cells") == 0);
}

// Replica grids of the learned features' point counts: up to 16, within 2^24 counters in all
// (none in a no-allocation call, whose scratch is sized for one grid); KNN_FEATURE_REPS=1..16
// overrides it for tuning.
static uint32_t feature_reps(uint64_t cells) {
  if (g_arena || cells == 0) return 1u;
  const char *v = std::getenv("KNN_FEATURE_REPS");
  const long e = v ? std::atol(v) : 0;
  if (e >= 1 && e <= 16) return (uint32_t)e;
  const uint64_t r = (1ull << 24) / cells;
  return r < 1 ? 1u : (r > 16 ? 16u : (uint32_t)r);
}

// Tile ranges of a range-kernel launch: one per XCD for single sets of 3 or more data dims,
// one queue for 2-D sets and batched groups.  Measured on one MI355X, each pair in one process
// (ms per join, one queue against one range per XCD): 10M mixed 2-D points K=64, plain 35.0 /
// 35.9, learned 21.3 / 21.5; 5M points, K=32, plain unless marked: 4-D uniform 26.7 / 25.9,
// skewed 49.3 / 52.2; 4-D on 2 indexed dims uniform 92.1 / 79.8, skewed (learned) 189 / 187;
// 6-D on 4 indexed dims uniform 233 / 221, skewed (learned) 532 / 552; 6-D (collect kernel)
// uniform 132 / 106, skewed 189 / 177; the 256-set batched join 144 / 164.
// KNN_TILE_RANGES=1|8 overrides the choice for tuning.
static uint32_t tile_ranges(uint32_t d, bool batched) {
  const char *v = std::getenv("KNN_TILE_RANGES");  // read per launch: tuning runs switch it in-process
  if (v && std::atoi(v) == 1) return 1u;
  if (v && std::atoi(v) == (int)TILE_QUEUES) return TILE_QUEUES;
  return (d >= 3 && !batched) ? TILE_QUEUES : 1u;
}

// Builds tiles of <= 64 active queries per tile block: activeFailed[q] != 0 selects q (all
// when null); act/tiles point into the cache, the tile count is left in device memory.
// Tile blocks are windows of `span` cells along dim 0 within one row; the span is chosen on the
// device from the active count (k_choose_span: wide windows pay only when active queries are
// dense, and a pass with few tiles ends with its slowest ones).
template <typename T>
void Pipeline<T>::make_tiles(const uint8_t *activeFailed, uint32_t *&act, uint4 *&tiles,
                             const uint32_t *&ntilesDev, uint8_t *failed, unsigned long long *cellAcc) {
  uint32_t *flags = buf<uint32_t>(S_FLAGS, n + 1);
  uint32_t *pos = buf<uint32_t>(S_POS, n + 1);
  bool all = !activeFailed;
  if (denseStart && failed) {
    // own-block pre-check (exact early failures) + per-cell candidate accounting
    uint32_t *cellMax = buf<uint32_t>(S_CELLMAX, (size_t)ncell + 1);
    fill_u32(cellMax, 0u, ncell, s);
    uint64_t *activeBits = nullptr;
    if (activeFailed) {
      activeBits = buf<uint64_t>(S_ABITS, ((size_t)n + 63) / 64);
      pack_bits(activeFailed, n, activeBits, s);
    }
    prefilter<T>(xs, n, g, radius, uniform_radius(), denseStart, sortedIds, activeBits, failed, K1, cellOf, cellMax,
                 flags, s);
    if (cellAcc) cell_bytes(cellMax, ncell, cellAcc, s);
    all = false;
  } else {
    active_flags(sortedIds, activeFailed, n, flags, s);
  }
  scan_exclusive_u32(flags, pos, n, scratch, s);
  if (all) {
    act = sortedIds;
  } else {
    act = buf<uint32_t>(S_ACT, n);
    compact_active(sortedIds, flags, pos, n, act, s);
  }
  activeDev = pos + n;
  uint32_t *meta = buf<uint32_t>(S_TILEMETA, 16);
  const uint32_t spanMax = nd == 1 ? 4096u : (nd == 2 ? 256u : 8u);
  // the pre-check's count of cells holding active queries (cellAcc[1]) as the density basis
  const unsigned long long *activeCells = (cellAcc && span_by_active_cells()) ? cellAcc + 1 : nullptr;
  choose_span(activeDev, activeCells, ncell, avgReach, spanMax, nd == 2 ? tile_rows_max() : 1u, min_tiles(), meta, s);
  spanDev = meta;
  uint32_t *bflags = buf<uint32_t>(S_BFLAGS, (size_t)ncell + 1);
  uint32_t *bpos = buf<uint32_t>(S_BPOS, (size_t)ncell + 1);
  uint32_t *bstart = buf<uint32_t>(S_BSTART, (size_t)ncell + 1);
  uint64_t *bkey = buf<uint64_t>(S_BKEY, (size_t)ncell + 1);
  uint32_t *partner = buf<uint32_t>(S_PARTNER, (size_t)ncell + 1);
  block_flags_dev(cellKeys, ncell, g.nCells[0], meta, bflags, s);
  scan_exclusive_u32(bflags, bpos, ncell, scratch, s);
  compact_blocks_keyed(cellKeys, cellStart, bflags, bpos, ncell, g.nCells[0], meta, n, bstart, bkey, s);
  uint32_t *ntiles = buf<uint32_t>(S_NTILES, (size_t)ncell + 1);
  uint32_t *tileOff = buf<uint32_t>(S_TILEOFF, (size_t)ncell + 1);
  tiles = buf<uint4>(S_TILES, n);
  pair_tiles_dev(bstart, bkey, pos, bpos + ncell, ncell, g.nCells[0], nd >= 2 ? g.nCells[1] : 1u, meta, partner,
                 ntiles, tileOff, scratch, tiles, s);
  ntilesDev = tileOff + ncell;
}

template <typename T>
PassResult Pipeline<T>::pass(const uint8_t *activeFailed, uint32_t nq, uint8_t *failed, int32_t *dIds,
                             T *dDists, unsigned long long *ctr, Timer &tK, Timer &tPass,
                             const std::function<void()> &beforeWait, const std::function<void()> &afterHeap) {
  // ctr layout (u64 words): [0] pairs, [1] algorithmic bytes, [2] fail|overflow counters,
  // [4..5] pre-check accounting (candidate points, query cells), [16..19] the tile queue
  // heads (TILE_QUEUES u32)
  uint32_t *counters = reinterpret_cast<uint32_t *>(ctr + 2);
  uint32_t *overflowCount = counters + 1;
  uint32_t *tileCounter = reinterpret_cast<uint32_t *>(ctr + 16);
  uint32_t *act;
  uint4 *tiles;
  const uint32_t *ntDev;
  tPass.start();
  fill_u32(reinterpret_cast<uint32_t *>(ctr + 4), 0u, 4, s);
  make_tiles(activeFailed, act, tiles, ntDev, failed, ctr + 4);
  uint8_t *overflow = buf<uint8_t>(S_OVERFLOW, n);
  fill_u8(overflow, 0, n, s);
  fill_u32(counters, 0u, 4, s);
  fill_u32(tileCounter, 0u, TILE_QUEUES, s);

  JoinArgs a{};
  a.x = x;
  a.xs = xs;
  a.sortedIds = sortedIds;
  a.cellKeys = cellKeys;
  a.cellStart = cellStart;
  a.ncell = ncell;
  a.denseStart = denseStart;
  a.act = act;
  a.tiles = tiles;
  a.ntiles = kPassGrid;  // persistent grid; the tile count is read from ntilesDev
  a.ntilesDev = ntDev;
  a.n = n;
  a.d = d;
  a.k1 = K1;
  a.radius = radius;
  a.radiusAll = (double)uniform_radius();
  a.outIds = dIds;
  a.outDists = dDists;
  a.failed = failed;
  a.failCount = counters;
  a.candCount = ctr;
  a.tileCounter = tileCounter;
  a.tileRanges = tile_ranges(d, false);
  a.segCapLog2 = seg_cap_log2();
  // LDS staging slots per query: K+1 and 64 more (the items of the K+1-th neighbour's
  // bucket beyond K+1), at most 512, within the 64 KB a workgroup may allocate
  a.selBuckets = selBuckets;
  a.smax = std::min<uint32_t>(K1 + 64, 512u);
  while (a.smax > K1 && range_select_lds_bytes<T>(a.smax, selBuckets) > 65536) a.smax -= 16;
  a.overflow = overflow;
  a.overflowCount = overflowCount;
  PassResult r{};
  bool selectK = false;
  tK.start();
  if (use_collect(nd, K1)) {
    range_collect<T>(a, g, s);
  } else if (!(a.smax > K1 && range_select_lds_bytes<T>(a.smax, selBuckets) <= 65536)) {
    // K+1 too large for LDS staging: the collect (or heap) kernel takes every query of the pass
    if (overflow_by_collect(K1)) {
      range_collect<T>(a, g, s);
    } else {
      a.ntiles = kHeapGrid;
      size_t hs = range_heap_scratch_bytes<T>(kHeapGrid, K1);
      a.heapScratch = hs ? g_cache.get(S_HEAP, hs) : nullptr;
      range_topk<T>(a, g, s);
    }
  } else {
    a.ntiles = sel_pass_grid(selBuckets);  // persistent grid: the workgroups its LDS admits
    range_select<T>(a, g, s);
    selectK = true;
  }
  LAUNCH_CHECK();
  tK.stop_async();
  // the pass's wait: failure / overflow / active counts and the accounting counters
  uint32_t *pin = pinned();
  auto readback = [&]() {
    HIP_CHECK(hipMemcpyAsync(pin + PW_KFAIL, counters, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(pin + PW_ACTIVE, activeDev, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(pin + PW_SPAN, spanDev, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(pin + PW_CTR, ctr, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  };
  if (beforeWait) beforeWait();
  readback();
  tPass.stop_async();
  sync();
  r.kernel_ms = tK.elapsed();
  lastActive = pin[PW_ACTIVE];
  lastSpan = pin[PW_SPAN];
  lastRows = pin[PW_SPAN + 1];
  r.noverflow = selectK ? pin[PW_OVF] : 0u;
  if (r.noverflow) {
    // queries whose candidates overflow the staging slots: the collect (or heap) kernel
    uint32_t *act2;
    uint4 *tiles2;
    const uint32_t *nt2;
    const uint32_t A1 = lastActive;
    make_tiles(overflow, act2, tiles2, nt2, nullptr, nullptr);
    fill_u32(tileCounter, 0u, TILE_QUEUES, s);
    a.act = act2;
    a.tiles = tiles2;
    a.ntilesDev = nt2;
    const bool byCollect = overflow_by_collect(K1);
    a.ntiles = byCollect ? kPassGrid : kHeapGrid;
    if (!byCollect) {
      size_t hs = range_heap_scratch_bytes<T>(kHeapGrid, K1);
      a.heapScratch = hs ? g_cache.get(S_HEAP, hs) : nullptr;
    }
    Timer tH(s, 5);
    tH.start();
    if (byCollect)
      range_collect<T>(a, g, s);
    else
      range_topk<T>(a, g, s);
    LAUNCH_CHECK();
    tH.stop_async();
    if (afterHeap) afterHeap();
    readback();
    tPass.stop_async();
    sync();
    r.kernel_ms += tH.elapsed();
    lastActive = A1;
  }
  r.pass_ms = tPass.elapsed();
  // failures = pre-check failures + kernel failures; algorithmic bytes += the pre-check's
  // per-query-cell candidate blocks ((4d+4) bytes per candidate, 8 bytes per scanned cell)
  const uint32_t prefailed = nq > lastActive ? nq - lastActive : 0u;
  r.nfail = pin[PW_KFAIL] + prefailed;
  unsigned long long hctr[6];
  std::memcpy(hctr, pin + PW_CTR, sizeof(hctr));
  uint64_t cellsScanned = 1;
  for (uint32_t j = 0; j < nd; j++) cellsScanned *= 3;
  accBytes += hctr[4] * (sizeof(T) * d + 4ull) + hctr[5] * cellsScanned * 8ull;
  r.bytes_total = hctr[1] + accBytes;
  r.pairs = hctr[0];
  return r;
}

namespace {
__global__ void k_fill_failed_rows(const uint8_t *failed, uint32_t n, uint32_t k1, int32_t *ids,
                                   int all) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)n * k1) return;
  const uint32_t q = (uint32_t)(i / k1);
  if (all || failed[q]) ids[i] = -1;
}
}  // namespace

template <typename T>
static void run_plain(Pipeline<T> &P, const knnjoin_params &prm, double beta, knnjoin_stats &stats,
                      uint8_t *failed, int32_t *dIds, T *dDists, Timer &tPh, Timer &tK);
template <typename T>
static void run_learned(Pipeline<T> &P, const knnjoin_params &p, knnjoin_stats &stats,
                        uint8_t *failed, int32_t *dIds, T *dDists, Timer &tPh, Timer &tK);

// ------------------------------------------------------------------ brute force
// When the brute-force kernel takes over, from measurements on MI355X (2-D, float32): it
// evaluates about 1.5e12 (query, point) pairs/s at K = 64 and 2e12-3e12 at K = 8 (less below
// about 50k points, where a launch is latency-bound), while one more index build and range
// pass costs about 0.25 ms + 0.3 ns per point (10M points: a 755-query tail took 5 ms more by
// brute force than by the learned schedule's last pass).  With w = max(1, (K+1)/9):
//   tail:  pending queries x N x w <= 7.5e8 + 1000 N  (10M points, K = 64: <= 149 queries)
//   whole: N^2 w <= 2^30                              (N <= 32k at K = 8, 12k at K = 64)
static bool brute_tail(uint64_t pending, uint64_t n, uint32_t k1) {
  const double w = std::max(1.0, (double)k1 / 9.0);
  return (double)pending * (double)n * w <= 7.5e8 + 1000.0 * (double)n;
}
static bool brute_whole(uint64_t n, uint32_t k1) {
  const double w = std::max(1.0, (double)k1 / 9.0);
  return (double)n * (double)n * w <= 1073741824.0;
}

// checkIndexSelectivity (main.cu:2687-2716): a grid of width eps prunes nothing when no
// indexed dim has more than 5 cells (the 3 a query searches plus the 2 of the eps margins).
template <typename T>
static bool no_selectivity(const T *mn, const T *mx, uint32_t nd, T eps) {
  if (!(eps > (T)0)) return false;
  for (uint32_t j = 0; j < nd; j++) {
    const T lo = mn[j] - eps, hi = mx[j] + eps;
    const T c = std::ceil((hi - lo) / eps);
    if (!(c <= (T)5)) return false;
  }
  return true;
}

// Exact rows of the queries with sel[q] != 0 (all when sel is null) by the brute-force kernel
// (distanceTableNDBruteForce, GPU.cu:1903-2930); failed[q] is cleared for each.  Returns the
// number of queries.
template <typename T>
static uint32_t run_brute(Pipeline<T> &P, const uint8_t *sel, uint8_t *failed, int32_t *dIds, T *dDists,
                          double &ms, Timer &tK) {
  hipStream_t s = P.s;
  const uint32_t n = P.n;
  const uint32_t *list = nullptr;
  uint32_t nq = n;
  if (sel) {
    uint32_t *flags = P.template buf<uint32_t>(S_FLAGS, n + 1);
    uint32_t *pos = P.template buf<uint32_t>(S_POS, n + 1);
    uint32_t *l = P.template buf<uint32_t>(S_BLIST, n);
    select_list(sel, n, flags, pos, l, P.scratch, s);
    HIP_CHECK(hipMemcpyAsync(&nq, pos + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    list = l;
  }
  if (nq == 0) return 0;
  // with the no-allocation scratch, partial tables within its free space
  const uint64_t maxPart = g_arena ? (g_arena->room() > (1u << 20) ? (g_arena->room() - (1u << 20)) / 12 : 0)
                                   : (1ull << 26);
  const BrutePlan pl = brute_plan(n, nq, P.K1, std::min<uint64_t>(maxPart, 1ull << 26));
  if (pl.qw == 0) throw Error(KNN_ERR_ARG, "brute force: K+1 must be <= 2048");
  BruteArgs a{};
  a.x = P.x;
  a.n = n;
  a.d = P.d;
  a.k1 = P.K1;
  a.queries = list;
  a.nq = nq;
  a.qw = pl.qw;
  a.chunk = pl.chunk;
  a.nchunks = pl.nchunks;
  if (pl.partialEntries) {
    a.partKey = P.template buf<double>(S_BRUTE_PK, pl.partialEntries);
    a.partId = P.template buf<uint32_t>(S_BRUTE_PI, pl.partialEntries);
  }
  a.outIds = dIds;
  a.outDists = dDists;
  a.failed = failed;
  tK.start();
  brute_join<T>(a, s);
  LAUNCH_CHECK();
  ms += tK.stop();
  if (P.verbose)
    printf("[knn] brute force: queries %u chunks %u x %u points, %u queries per wave\n", nq, pl.nchunks,
           pl.chunk, pl.qw);
  return nq;
}

template <typename T>
static int run_join(const T *pts_in, uint32_t n, uint32_t d, uint32_t k, const knnjoin_params &p,
                    int32_t *outIds, T *outDists, knnjoin_stats *st) {
  knnjoin_stats stats{};
  if (n == 0) {
    if (st) *st = stats;
    return KNN_OK;
  }
  if (!pts_in || !outIds || !outDists) throw Error(KNN_ERR_ARG, "null pointer argument");
  if (d == 0 || d > KNN_MAXDIM) throw Error(KNN_ERR_ARG, "NDIM must be in 1..16");
  uint32_t nd = p.num_indexed_dims ? p.num_indexed_dims : std::min<uint32_t>(d, 6);
  if (nd > d || nd > KNN_MAX_INDEXED) throw Error(KNN_ERR_ARG, "NUMINDEXEDDIM must be <= NDIM and <= 8");
  if ((uint64_t)n < (uint64_t)k + 1)
    throw Error(KNN_ERR_TOO_FEW, "N < K+1: no point has K+1 neighbours (the reference never terminates)");
  const double beta = p.beta < 0 ? 0.05 : p.beta;
  const uint32_t K1 = k + 1;
  const bool devptr = p.device_ptrs != 0;
  hipStream_t s = (devptr && p.stream) ? reinterpret_cast<hipStream_t>(p.stream) : lib_stream();

  Pipeline<T> P;
  P.s = s;
  P.n = n;
  P.d = d;
  P.nd = nd;
  P.k = k;
  P.K1 = K1;
  P.verbose = p.verbose != 0;
  P.brute = p.brute_fallback < 0 ? p.mode == KNN_MODE_LEARNED : p.brute_fallback != 0;
  if (p.mode != KNN_MODE_BETA && p.mode != KNN_MODE_LEARNED && p.mode != KNN_MODE_BRUTE)
    throw Error(KNN_ERR_ARG, "unknown radius mode");
  if (p.hybrid_threads && p.mode != KNN_MODE_BETA)
    throw Error(KNN_ERR_ARG, "the hybrid CPU/GPU split runs with the BETA schedule (KNN_MODE_BETA)");
  if (p.hybrid_threads && g_arena)
    throw Error(KNN_ERR_ARG, "the no-allocation entry point runs without the hybrid split");
  if ((p.mode == KNN_MODE_BRUTE || P.brute) && K1 > BRUTE_MAX_K1) {
    if (p.mode == KNN_MODE_BRUTE) throw Error(KNN_ERR_ARG, "KNN_MODE_BRUTE: K+1 must be <= 2048");
    P.brute = false;
  }
  P.scratchWords = std::max(scan_scratch_words(n + 1), radix_scratch_words(n)) + 1024;
  P.scratch = P.template buf<uint32_t>(S_SCRATCH, P.scratchWords);

  Timer tAll(s, 0), tPh(s, 1), tK(s, 2);
  auto wall0 = std::chrono::steady_clock::now();

  // ---- inputs
  const T *dpts;
  int32_t *dIds;
  T *dDists;
  if (devptr) {
    dpts = pts_in;
    dIds = outIds;
    dDists = outDists;
  } else {
    T *b = P.template buf<T>(S_PTS, (size_t)n * d);
    HIP_CHECK(hipMemcpyAsync(b, pts_in, sizeof(T) * (size_t)n * d, hipMemcpyHostToDevice, s));
    dpts = b;
    dIds = P.template buf<int32_t>(S_OUTIDS, (size_t)n * K1);
    dDists = P.template buf<T>(S_OUTDISTS, (size_t)n * K1);
  }
  uint8_t *failed = P.template buf<uint8_t>(S_FAILED, n);
  double *rowSum = P.template buf<double>(S_ROWSUM, n);
  double h2d_ms = 0;
  if (!devptr) {
    HIP_CHECK(hipStreamSynchronize(s));
    h2d_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
  }
  tAll.start();

  // ---- dimension reorder (main.cu:2521-2615)
  tPh.start();
  uint32_t perm[KNN_MAXDIM];
  for (uint32_t j = 0; j < KNN_MAXDIM; j++) perm[j] = j;
  if (p.reorder && d > 1) {
    // the order the sequential T variances give, decided from FP64 error intervals when they
    // separate every pair of dims, else from the sequential sums themselves
    double *mom = P.template buf<double>(S_SMALL, 64 * 1024 / sizeof(double));
    variance_moments<T>(dpts, n, d, mom, s);
    double hm[4 * KNN_MAXDIM];
    HIP_CHECK(hipMemcpyAsync(hm, mom, sizeof(double) * 4 * d, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (!order_from_intervals<T>(hm, n, d, perm)) {
      T *var = P.template buf<T>(S_SMALL, 64 * 1024 / sizeof(T));
      variance_rows<T>(dpts, n, d, var, s);
      T hv[KNN_MAXDIM];
      HIP_CHECK(hipMemcpyAsync(hv, var, sizeof(T) * d, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      sort_perm_desc<T>(hv, d, perm);
    }
  }
  P.x = P.template buf<T>(S_X, (size_t)n * d);
  permute_to_soa<T>(dpts, n, d, perm, P.x, s);
  P.rowMajor = dpts;
  for (uint32_t j = 0; j < KNN_MAXDIM; j++) P.rowPerm.v[j] = j < d ? perm[j] : 0u;
  stats.ms_reorder = tPh.stop();
  for (uint32_t j = 0; j < 16; j++) stats.perm[j] = j < d ? perm[j] : 0;

  if (p.mode == KNN_MODE_LEARNED && !radius_model(d, nd).ok)
    throw Error(KNN_ERR_MODEL, "no learned-radius model registered for this number of indexed dims");
  if (p.mode == KNN_MODE_BRUTE || (P.brute && brute_whole(n, K1))) {
    // the whole join without the index
    tPh.start();
    double bms = 0;
    stats.brute_queries = run_brute<T>(P, nullptr, nullptr, dIds, dDists, bms, tK);
    stats.ms_join = tPh.stop();
    stats.iterations = 1;
    stats.kernel_launches = 1;
  } else if (p.mode == KNN_MODE_LEARNED) {
    run_learned<T>(P, p, stats, failed, dIds, dDists, tPh, tK);
  } else {
    run_plain<T>(P, p, beta, stats, failed, dIds, dDists, tPh, tK);
  }
  // ---- checksum (main.cu:452)
  double *part = P.template buf<double>(S_CSUM, 64 + (n + 4095) / 4096 + 16);
  double *csum = part + (n + 4095) / 4096 + 8;
  row_sums<T>(dDists, n, K1, rowSum, s);
  sum_rows(rowSum, n, part, csum, s);
  LAUNCH_CHECK();
  double hcs = 0;
  HIP_CHECK(hipMemcpyAsync(&hcs, csum, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  stats.ms_total = tAll.stop();
  stats.ms_reference_total = stats.ms_total - stats.ms_index_first;
  stats.checksum = hcs;
  stats.failed_final = 0;

  if (!devptr) {
    auto t0 = std::chrono::steady_clock::now();
    HIP_CHECK(hipMemcpyAsync(outIds, dIds, sizeof(int32_t) * (size_t)n * K1, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(outDists, dDists, sizeof(T) * (size_t)n * K1, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    stats.ms_other = h2d_ms + std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (P.verbose)
    printf("[knn] done: iterations %u reindexes %u eps0 %.9g eps %.9g checksum %.6f total %.3f ms\n",
           stats.iterations, stats.reindexes, (double)stats.eps0, (double)stats.eps_final, hcs,
           stats.ms_total);
  if (st) *st = stats;
  return KNN_OK;
}

// Grids of the plain schedule: cell width eps / plain_grid_div at a build, every query
// searching radius eps; the index is rebuilt when eps grows past plain_max_reach cells.  The
// reference re-indexes at every growth with cells of width eps (main.cu:347); the rows, the
// eps sequence and every pass's failures depend only on eps, not on the cells.
// KNN_PLAIN_GRID_DIV (1..8) and KNN_PLAIN_MAX_REACH (1..8) override them for tuning.
static double env_knob(const char *name, double lo, double hi) {
  const char *v = std::getenv(name);
  const double x = v ? std::atof(v) : 0.0;
  return (x >= lo && x <= hi) ? x : 0.0;
}
// Measured on MI355X (plain schedule, 2 indexed dims): cells of width eps kept until eps
// doubles, against a re-index at every growth: 10M mixed points 85.4 / 98.1 ms at K = 64 and
// 83.0 / 97.7 ms at K = 16, 10M uniform 47.2 / 48.1 ms, 1M mixed K = 8 17.8 / 18.7 ms.
static double plain_grid_div(uint32_t) {
  const double e = env_knob("KNN_PLAIN_GRID_DIV", 1.0, 8.0);
  return e > 0.0 ? e : 1.0;
}
static double plain_max_reach(uint32_t nd) {
  const double e = env_knob("KNN_PLAIN_MAX_REACH", 1.0, 8.0);
  return e > 0.0 ? e : (nd <= 2 ? 2.0 : 1.0);
}

// ------------------------------------------------------------------ hybrid split
// Pinned host buffers of the hybrid split's copies (grow-only, released by knnjoin_release).
struct HostPinned {
  void *p[6] = {};
  size_t cap[6] = {};
  void *get(int i, size_t bytes) {
    if (cap[i] < bytes) {
      if (p[i]) HIP_CHECK(hipHostFree(p[i]));
      p[i] = nullptr;
      cap[i] = 0;
      HIP_CHECK(hipHostMalloc(&p[i], bytes ? bytes : 16, hipHostMallocDefault));
      cap[i] = bytes;
    }
    return p[i];
  }
  void release() {
    for (int i = 0; i < 6; i++) {
      if (p[i]) (void)hipHostFree(p[i]);
      p[i] = nullptr;
      cap[i] = 0;
    }
  }
};
static HostPinned g_hostPin;

// The host share of a hybrid join (splitWork, main.cu:2733-2933): the sparse-cell queries of
// the first index, joined by host threads from a host copy of that index while the GPU loop
// runs over the other queries; the rows are scattered into the device table at the end.
template <typename T>
struct HybridShare {
  uint32_t nq = 0;               // host queries
  const uint8_t *gpuActive = nullptr;  // device: 1 = a query of the GPU loop
  uint32_t *listIds = nullptr;   // device: the host queries' point ids, row order of the share
  HybridIndex<T> ix{};
  typename GridOf<T>::type g{};
  const uint32_t *qpos = nullptr;  // host: their sorted positions
  int32_t *rowIds = nullptr;
  T *rowDists = nullptr;
  uint32_t threads = 1, k1 = 0;
  double ms = 0;
  std::thread th;
  std::exception_ptr err;
  ~HybridShare() {
    if (th.joinable()) th.join();
  }
};

namespace {
__global__ void k_ids_of_positions(const uint32_t *pos, const uint32_t *sortedIds, uint32_t m, uint32_t *ids) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) ids[i] = sortedIds[pos[i]];
}
// the queries at the listed sorted positions go back to the GPU loop
__global__ void return_to_gpu(const uint32_t *pos, uint32_t m, const uint32_t *sortedIds, uint8_t *gpu) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) gpu[sortedIds[pos[i]]] = 1;
}
}  // namespace

template <typename T>
static void hybrid_start(Pipeline<T> &P, const knnjoin_params &prm, HybridShare<T> &H) {
  hipStream_t s = P.s;
  const uint32_t n = P.n;
  uint32_t thr = hybrid_point_threshold(P.K1, P.nd);
  // the cell table on the host (also the host index's), and with a cap on the host share the
  // largest threshold whose sparse cells hold at most that many queries
  uint32_t *cs = reinterpret_cast<uint32_t *>(g_hostPin.get(0, sizeof(uint32_t) * ((size_t)P.ncell + 1)));
  uint64_t *ck = reinterpret_cast<uint64_t *>(g_hostPin.get(1, sizeof(uint64_t) * ((size_t)P.ncell + 1)));
  HIP_CHECK(hipMemcpyAsync(cs, P.cellStart, sizeof(uint32_t) * ((size_t)P.ncell + 1), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(ck, P.cellKeys, sizeof(uint64_t) * (size_t)P.ncell, hipMemcpyDeviceToHost, s));
  P.sync();
  const uint32_t cap = prm.hybrid_max_queries;
  if (cap) {
    // the sparsest cells first: every cell of up to t points, where t is the largest
    // population whose cells hold at most `cap` queries, then cells of t + 1 points in cell
    // order until the cap (the list below is cut there)
    std::vector<uint64_t> h((size_t)thr + 1, 0);
    for (uint32_t c = 0; c < P.ncell; c++) {
      const uint32_t m = cs[c + 1] - cs[c];
      if (m <= thr) h[m] += m;
    }
    uint64_t acc = 0;
    uint32_t t = 0;
    for (uint32_t m = 1; m <= thr; m++) {
      if (acc + h[m] > cap) break;
      acc += h[m];
      t = m;
    }
    thr = std::min(thr, t + 1);
  }
  uint8_t *cpu = P.template buf<uint8_t>(S_HCPU, n);
  uint8_t *gpu = P.template buf<uint8_t>(S_HGPU, n);
  hybrid_flags(P.cellStart, P.cellOf, P.sortedIds, n, thr, cpu, gpu, s);
  uint32_t *flags = P.template buf<uint32_t>(S_FLAGS, n + 1);
  uint32_t *pos = P.template buf<uint32_t>(S_POS, n + 1);
  uint32_t *list = P.template buf<uint32_t>(S_HLIST, n + 1);
  select_list(cpu, n, flags, pos, list, P.scratch, s);
  uint32_t *pin = pinned();
  HIP_CHECK(hipMemcpyAsync(pin + PW_ACTIVE, pos + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  P.sync();
  H.nq = pin[PW_ACTIVE];
  if (cap && H.nq > cap) {
    // past the cap the listed queries stay with the GPU
    return_to_gpu<<<(H.nq - cap + 255) / 256, 256, 0, s>>>(list + cap, H.nq - cap, P.sortedIds, gpu);
    LAUNCH_CHECK();
    H.nq = cap;
  }
  if (H.nq == 0) return;
  const uint32_t d = P.d;
  T *xs = reinterpret_cast<T *>(g_hostPin.get(2, sizeof(T) * (size_t)n * d));
  uint32_t *sid = reinterpret_cast<uint32_t *>(g_hostPin.get(3, sizeof(uint32_t) * (size_t)n));
  uint32_t *qp = reinterpret_cast<uint32_t *>(g_hostPin.get(4, sizeof(uint32_t) * (size_t)H.nq));
  HIP_CHECK(hipMemcpyAsync(xs, P.xs, sizeof(T) * (size_t)n * d, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(sid, P.sortedIds, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(qp, list, sizeof(uint32_t) * (size_t)H.nq, hipMemcpyDeviceToHost, s));
  // the rows' destinations: the point ids of the listed positions (device, for the scatter)
  H.listIds = P.template buf<uint32_t>(S_HROWI, H.nq);
  k_ids_of_positions<<<(H.nq + 255) / 256, 256, 0, s>>>(list, P.sortedIds, H.nq, H.listIds);
  LAUNCH_CHECK();
  P.sync();
  H.ix = HybridIndex<T>{xs, sid, ck, cs, P.ncell, n, d};
  H.g = P.g;
  H.qpos = qp;
  H.k1 = P.K1;
  H.threads = prm.hybrid_threads;
  const size_t rowBytes = (size_t)H.nq * P.K1;
  H.rowIds = reinterpret_cast<int32_t *>(g_hostPin.get(5, rowBytes * (sizeof(int32_t) + sizeof(T))));
  H.rowDists = reinterpret_cast<T *>(H.rowIds + rowBytes);
  H.gpuActive = gpu;
  H.th = std::thread([&H]() {
    try {
      auto t0 = std::chrono::steady_clock::now();
      hybrid_host_join<T>(H.ix, H.g, H.k1, H.qpos, H.nq, H.threads, H.rowIds, H.rowDists);
      H.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    } catch (...) {
      H.err = std::current_exception();
    }
  });
}

// waits for the host share and scatters its rows into the device table
template <typename T>
static void hybrid_finish(Pipeline<T> &P, HybridShare<T> &H, int32_t *dIds, T *dDists) {
  if (!H.nq) return;
  if (H.th.joinable()) H.th.join();
  if (H.err) std::rethrow_exception(H.err);
  const size_t cnt = (size_t)H.nq * P.K1;
  int32_t *ri = P.template buf<int32_t>(S_HLIST, cnt);
  T *rd = P.template buf<T>(S_HROWD, cnt);
  HIP_CHECK(hipMemcpyAsync(ri, H.rowIds, sizeof(int32_t) * cnt, hipMemcpyHostToDevice, P.s));
  HIP_CHECK(hipMemcpyAsync(rd, H.rowDists, sizeof(T) * cnt, hipMemcpyHostToDevice, P.s));
  hybrid_scatter<T>(ri, rd, H.listIds, H.nq, P.K1, dIds, dDists, P.s);
  LAUNCH_CHECK();
}

// The plain method: BETA eps0 estimate and the failed-query loop of main.cu:305-430.  Each
// pass waits on the device once (twice when staging overflows send queries to the heap
// kernel), each index build once.
template <typename T>
static void run_plain(Pipeline<T> &P, const knnjoin_params &prm, double beta, knnjoin_stats &stats,
                      uint8_t *failed, int32_t *dIds, T *dDists, Timer &tPh, Timer &tK) {
  hipStream_t s = P.s;
  const uint32_t n = P.n, d = P.d, nd = P.nd, k = P.k, K1 = P.K1;
  auto fail_all = [&](int code, const std::string &msg) {
    const size_t tot = (size_t)n * K1;
    k_fill_failed_rows<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(failed, n, K1, dIds, 1);
    throw Error(code, msg);
  };
  // ---- epsilon estimate (GPU.cu:3088-3566)
  tPh.start();
  EpsResult er = estimate_eps<T>(P.x, n, d, k, beta, nullptr, s);
  T eps0 = (T)er.best;
  stats.eps_bucket = er.bucket;
  stats.ms_eps = tPh.stop();

  // ---- min/max of the indexed dims (main.cu:2439-2456)
  T hmn[KNN_MAX_INDEXED], hmx[KNN_MAX_INDEXED];
  P.index_minmax(hmn, hmx);
  if (eps0 == (T)0) eps0 = fallback_eps<T>(hmn, hmx, nd, n, K1);
  stats.eps0 = (float)eps0;

  // ---- the join loop (main.cu:305-430)
  const double pdiv = plain_grid_div(nd), preach = plain_max_reach(nd);
  T curW = (T)0;
  auto build = [&](T e) -> double {
    T w = (T)((double)e / pdiv);
    if (pdiv > 1.0) {  // a finer grid only while it keeps the dense start table
      typename GridOf<T>::type gt;
      uint64_t tc = 0;
      bool ok = w > (T)0;
      try {
        if (ok) grid_dims<T>(hmn, hmx, nd, w, gt, tc);
      } catch (const Error &) {
        ok = false;
      }
      if (!ok || tc + 1 > Pipeline<T>::kDenseMaxCells) w = e;
    }
    P.tBuild = &tPh;
    tPh.start();
    P.build_index(hmn, hmx, w);
    P.tBuild = nullptr;
    curW = w;
    return tPh.elapsed();
  };
  P.radius = nullptr;  // every query searches the same radius, passed by value
  T eps = eps0;
  const uint32_t syncs0 = P.syncs;
  stats.ms_index = build(eps);
  stats.ms_index_first = stats.ms_index;

  unsigned long long *ctr = P.template buf<unsigned long long>(S_CTR, 64);
  fill_u32(reinterpret_cast<uint32_t *>(ctr), 0u, 16, s);
  // ---- hybrid split of the first index's queries (splitWork, main.cu:2733-2933)
  HybridShare<T> H;
  if (prm.hybrid_threads) {
    hybrid_start<T>(P, prm, H);
    if (H.nq) fill_u8(failed, 0, n, s);  // the host queries are never pending on the GPU
  }
  stats.hybrid_queries = H.nq;
  uint32_t nq = n - H.nq, numIter = 0, reindexes = 0, launches = 0;
  double frac = 0;
  bool first = true;
  double joinMs = 0, kMs = 0, kMax = 0;
  uint64_t bytesAtMax = 0, prevBytes = 0, pairs = 0;
  while (nq != 0) {
    if (numIter != 0) {
      if (frac > 0.25) {  // main.cu:330-356
        T grown = grow_eps(eps, eps0);
        if (grown == (T)0) fail_all(KNN_ERR_EPS_ZERO, "epsilon grew to 0");
        if (grown == eps) {
          k_fill_failed_rows<<<(unsigned)(((size_t)n * K1 + 255) / 256), 256, 0, s>>>(failed, n, K1, dIds, 0);
          throw Error(KNN_ERR_NO_PROGRESS, "eps + eps0/2 == eps: epsilon cannot grow");
        }
        eps = grown;
        if ((double)eps > preach * (double)curW) stats.ms_index += build(eps);
        reindexes++;  // the reference's re-index count: eps growths
      } else {
        // The reference reruns the failed queries at the same eps (main.cu:330 is false);
        // the pass is deterministic, so every one of them fails again (fraction 1.0).
        numIter++;
        frac = 1.0;
        continue;
      }
    }
    P.radiusAll = eps;
    P.avgReach = (double)eps / (double)curW;
    PassResult pr = P.pass(first ? H.gpuActive : failed, nq, failed, dIds, dDists, ctr, tK, tPh);
    launches++;
    stats.overflow_queries += pr.noverflow;
    kMs += pr.kernel_ms;
    if (pr.kernel_ms > kMax) {
      kMax = pr.kernel_ms;
      bytesAtMax = pr.bytes_total - prevBytes;
    }
    prevBytes = pr.bytes_total;
    const uint32_t nfail = pr.nfail;
    joinMs += pr.pass_ms;
    if (P.verbose)
      printf("[knn] iter %u eps %.9g queries %u failed %u cells %u span %u rows %u overflow %u candidates %llu kernel %.3f ms\n",
             numIter, (double)eps, nq, nfail, P.ncell, P.lastSpan, P.lastRows, pr.noverflow,
             (unsigned long long)(pr.pairs - pairs), pr.kernel_ms);
    pairs = pr.pairs;
    frac = (nfail * 1.0) / (nq * 1.0);  // main.cu:402
    nq = nfail;
    numIter++;
    first = false;
    if (nq != 0 && P.brute &&
        (brute_tail(nq, n, K1) || no_selectivity<T>(hmn, hmx, nd, grow_eps(eps, eps0)))) {
      tPh.start();
      double bms = 0;
      stats.brute_queries += run_brute<T>(P, failed, failed, dIds, dDists, bms, tK);
      joinMs += tPh.stop();
      launches++;
      numIter++;
      nq = 0;
    }
  }
  hybrid_finish<T>(P, H, dIds, dDists);
  stats.ms_hybrid_host = H.ms;
  stats.host_syncs = P.syncs - syncs0;
  stats.ms_join = joinMs;
  stats.ms_range_kernel = kMs;
  stats.ms_range_kernel_max = kMax;
  stats.algo_bytes_max = bytesAtMax;
  stats.candidates = pairs;
  stats.algo_bytes = prevBytes;
  stats.eps_final = (float)eps;
  P.radiusAll = (T)0;
  stats.iterations = numIter;
  stats.reindexes = reindexes;
  stats.kernel_launches = launches;
}

// Cells per radius-class upper bound R in the learned schedule's grids: with <= 2 indexed
// dims 3 up to K+1 = 80 and 4 above (with the wider classes below), 1 with more indexed dims,
// 4 on a projected grid (fewer indexed than data dims; see learned_class_bins);
// KNN_LEARNED_GRID_DIV (1..8) overrides it for tuning.
static double learned_grid_div(uint32_t d, uint32_t nd, uint32_t k1) {
  static const double env = []() {
    const char *v = std::getenv("KNN_LEARNED_GRID_DIV");
    const double x = v ? std::atof(v) : 0.0;
    return (x >= 1.0 && x <= 8.0) ? x : 0.0;
  }();
  if (env > 0.0) return env;
  if (nd < d) return 4.0;
  return nd <= 2 ? (k1 <= 80 ? 3.0 : 4.0) : 1.0;
}

// Quarter-octave bins per radius class with <= 2 indexed dims: 8 (two octaves) up to
// K+1 = 80, 10 above; 2 with more indexed dims.  Measured on MI355X, 10M mixed 2-D points
// (ms per join, classes of 8 bins on cells of R/3 against 10 bins on cells of R/4): K = 16
// 28.9 / 28.2, K = 32 30.4 / 30.2, K = 64 39.4 / 40.4, K = 96 56.0 / 50.2, K = 128 66.5 /
// 59.7, K = 256 128.1 / 117.6.  Two-octave classes on cells of R/3 beat one-octave classes
// on cells of R/2 everywhere (K = 64: 41.2 against 47.3 ms), since each class costs an index
// build over all N points.  On a projected grid (fewer indexed than data dims) 6 bins on
// cells of R/4; measured on 5M points, K = 32, learned ms (bins / cells: default of the
// indexed dims, 4 / R/3, 6 / R/4): 4-D with 2 indexed, uniform 126 / 90 / 87, skewed
// 202 / 213 / 191; 6-D with 4 indexed, uniform 270 / 204 / 202, skewed 1819 / 505 / 537.
// With 3 or more dims, every one indexed: 2 bins on cells of R (1 / 2 / 3 bins, ms: 4-D
// uniform 32.2 / 31.5 / 34.4, skewed 52.8 / 49.5 / 48.2; 6-D uniform 97.9 / 97.3 / 118, skewed
// 146 / 144 / 209; 4 bins on cells of R/3: 4-D uniform 84.5).  KNN_LEARNED_CLASS_BINS (1..16)
// overrides it for tuning.
static int learned_class_bins(uint32_t d, uint32_t nd, uint32_t k1) {
  static const int env = []() {
    const char *v = std::getenv("KNN_LEARNED_CLASS_BINS");
    const int x = v ? std::atoi(v) : 0;
    return (x >= 1 && x <= 16) ? x : 0;
  }();
  if (env > 0) return env;
  if (nd < d) return 6;
  return nd <= 2 ? (k1 <= 80 ? 8 : 10) : 2;
}

// The learned schedule: per-query initial radius from the MLP, then passes over radius
// octaves (each on a grid matched to its radii) with the expansion loop (radius x2 for the
// queries that fail) until every query has its K+1 neighbours.  A pass waits on the device
// once: the next pass's radius histogram is queued before that wait.
template <typename T>
static void run_learned(Pipeline<T> &P, const knnjoin_params &p, knnjoin_stats &stats,
                        uint8_t *failed, int32_t *dIds, T *dDists, Timer &tPh, Timer &tK) {
  (void)p;
  hipStream_t s = P.s;
  const uint32_t n = P.n, nd = P.nd, K1 = P.K1;
  const RadiusModel &M = radius_model(P.d, nd);
  if (!M.ok) throw Error(KNN_ERR_MODEL, "no learned-radius model registered for this number of indexed dims");
  P.selBuckets = (uint32_t)SEL_NB_LEARNED;
  // min/max of the indexed dims
  T hmn[KNN_MAX_INDEXED], hmx[KNN_MAX_INDEXED];
  P.index_minmax(hmn, hmx);

  // ---- features + MLP -> per-query radius
  tPh.start();
  FeatureGrid<T> F;
  feature_grid<T>(hmn, hmx, nd, n, K1, F);
  const uint32_t reps = feature_reps(F.cells);
  uint32_t *cnt = P.template buf<uint32_t>(S_FCNT, (size_t)F.cells * reps);
  fill_u32(cnt, 0u, (size_t)F.cells * reps, s);
  feature_counts<T>(P.x, n, F.g, cnt, F.cells, reps, s);
  feature_occupancy(cnt, F.cells, n, F.logdens, s);
  float *Wd = P.template buf<float>(S_MODEL, MLP_NW);
  HIP_CHECK(hipMemcpyAsync(Wd, M.w.data(), sizeof(float) * MLP_NW, hipMemcpyHostToDevice, s));
  T *radius = P.template buf<T>(S_RADIUS, n);
  radius_kernel<T>(P.x, n, F.g, cnt, F.logk1, F.logdens, Wd, M.margin, (float)F.wf, radius, nullptr,
                   nullptr, s);
  LAUNCH_CHECK();
  P.radius = radius;

  // ---- radius classes: the pending queries whose radius lies in the lowest occupied class
  // of bins [c, c + learned_class_bins) (two or two and a half octaves with <= 2 indexed dims,
  // one and a half on projected grids, half an octave above) are joined on a grid matched to R
  // (cell width R / learned_grid_div);
  // queries that fail grow their radius (x2, or x2^(1/4) with 3 or more indexed dims), which
  // moves them to a later class.  One index build per occupied class.
  unsigned long long *ctr = P.template buf<unsigned long long>(S_CTR, 64);
  fill_u32(reinterpret_cast<uint32_t *>(ctr), 0u, 16, s);
  uint64_t *maxr = reinterpret_cast<uint64_t *>(ctr + 8);
  uint32_t *ph = P.template buf<uint32_t>(S_RHIST, 64);
  uint8_t *act = P.template buf<uint8_t>(S_ACTQ, n);
  fill_u8(failed, 1, n, s);  // pending
  const float wfF = (float)F.wf;
  uint32_t *pin = pinned();
  // the pending queries' radius histogram into pin[PW_HIST..] (read at the next wait)
  auto queue_bins = [&]() {
    fill_u32(ph, 0u, 64, s);
    radius_bins(radius, failed, n, wfF, ph, s);
    HIP_CHECK(hipMemcpyAsync(pin + PW_HIST, ph, 64 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  };
  queue_bins();
  tPh.stop_async();
  const uint32_t syncs0 = P.syncs;
  P.sync();
  stats.ms_eps = tPh.elapsed();
  uint32_t passes = 0, reindexes = 0;
  double joinMs = 0, kMs = 0, kMax = 0;
  uint64_t bytesAtMax = 0, prevBytes = 0, pairs = 0;
  T curW = (T)0;
  stats.eps0 = 0.0f;
  for (;;) {
    uint32_t h[64];
    std::memcpy(h, pin + PW_HIST, sizeof(h));
    uint64_t pending = 0;
    int c = -1;
    for (int b = 0; b < 64; b++) {
      pending += h[b];
      if (c < 0 && h[b]) c = b;
    }
    if (pending == 0) break;
    if (passes > 200) throw Error(KNN_ERR_NO_PROGRESS, "learned schedule: queries still failing after 200 passes");
    if (P.brute && (brute_tail(pending, n, K1) ||
                    no_selectivity<T>(hmn, hmx, nd, (T)((double)wfF * std::exp2((c - 32) / 4.0))))) {
      // few pending queries, or even the smallest pending radius (the lower edge of the lowest
      // occupied bin) leaves the index nothing to prune: brute-force them together
      tPh.start();
      double bms = 0;
      stats.brute_queries += run_brute<T>(P, failed, failed, dIds, dDists, bms, tK);
      joinMs += tPh.stop();
      passes++;
      break;
    }
    // wide classes with <= 2 indexed dims and on projected grids; half an octave above, where a
    // cell up to 2x the radius would cost up to 2^nd in block volume
    const int cHi = std::min(c + learned_class_bins(P.d, nd, P.K1) - 1, 63);
    uint32_t nq = 0;
    for (int b = c; b <= cHi; b++) nq += h[b];
    double R;
    if (cHi < 63) {
      R = (double)wfF * std::exp2((cHi + 1 - 32) / 4.0);
    } else {
      // the top bin is open-ended: the class's largest radius bounds it
      max_failed_radius(radius, failed, n, maxr, s);
      HIP_CHECK(hipMemcpyAsync(pin + PW_MAXR, maxr, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
      P.sync();
      uint64_t mb = 0;
      std::memcpy(&mb, pin + PW_MAXR, sizeof(mb));
      if (sizeof(T) == 4) {
        uint32_t b32 = (uint32_t)mb;
        float f;
        std::memcpy(&f, &b32, 4);
        R = f;
      } else {
        std::memcpy(&R, &mb, 8);
      }
      R = R * (1.0 + 1e-6);
    }
    // cell width R/3 with <= 2 indexed dims (R/4 on projected grids); above, R (<= 3 cells per
    // dim), since every extra cell of reach multiplies the rows of a block
    T w = (T)(R / learned_grid_div(P.d, nd, P.K1));
    if (!(w > (T)0)) w = F.wf;
    // a small tail of pending queries (at most N/64) keeps the current grid when its blocks
    // stay within 1024 cells: an index build over all N points costs more than the wider
    // blocks of the few queries left (10M mixed 2-D points, K=64: the fourth pass, 15.8k
    // queries, without its build)
    bool reuse = false;
    if (curW > (T)0 && (uint64_t)nq * reuse_tail_div() <= (uint64_t)n) {
      const double reach = std::ceil(R / (double)curW);
      reuse = std::pow(2.0 * reach + 1.0, (double)nd) <= 1024.0;
    }
    if (reuse) w = curW;
    for (;;) {  // keep the grid within 2^30 cells
      typename GridOf<T>::type gt;
      uint64_t tc = 0;
      bool ok = true;
      try {
        grid_dims<T>(hmn, hmx, nd, w, gt, tc);
      } catch (const Error &) {
        ok = false;
      }
      if (ok && tc <= (1ull << 30)) break;
      w = w * (T)2;
    }
    if (w != curW) {
      P.tBuild = &tPh;
      tPh.start();
      P.build_index(hmn, hmx, w);
      P.tBuild = nullptr;
      const double bms = tPh.elapsed();
      stats.ms_index += bms;
      if (curW == (T)0) stats.ms_index_first = bms;
      if (curW != (T)0) reindexes++;
      curW = w;
      if (stats.eps0 == 0.0f) stats.eps0 = (float)w;
    }
    // typical reach in cells (tile-span cost model): 1.5 on the class's own grid, the half
    // radius in cells on a kept one
    P.avgReach = reuse ? std::max(1.5, 0.5 * R / (double)curW) : 1.5;
    radius_class_flags(radius, failed, n, wfF, (uint32_t)c, (uint32_t)cHi, act, s);
    // growth on failure: one class (an octave; a quarter octave with >= 3 indexed dims), then
    // the next histogram, both before the pass's wait; after a heap-kernel pass (overflowing
    // staging) only the histogram again
    const T factor = nd <= 2 ? (T)2 : (T)1.189207115002721;
    PassResult pr = P.pass(act, nq, failed, dIds, dDists, ctr, tK, tPh,
                           [&]() {
                             grow_class_radius(radius, act, failed, n, factor, s);
                             queue_bins();
                           },
                           [&]() { queue_bins(); });
    stats.overflow_queries += pr.noverflow;
    kMs += pr.kernel_ms;
    if (pr.kernel_ms > kMax) {
      kMax = pr.kernel_ms;
      bytesAtMax = pr.bytes_total - prevBytes;
    }
    prevBytes = pr.bytes_total;
    joinMs += pr.pass_ms;
    if (P.verbose)
      printf("[knn] learned pass %u octave %d grid %.9g queries %u failed %u cells %u span %u rows %u overflow %u candidates %llu kernel %.3f ms\n",
             passes, c, (double)P.g.eps, nq, pr.nfail, P.ncell, P.lastSpan, P.lastRows, pr.noverflow,
             (unsigned long long)(pr.pairs - pairs), pr.kernel_ms);
    pairs = pr.pairs;
    passes++;
  }
  stats.host_syncs = P.syncs - syncs0;
  stats.ms_join = joinMs;
  stats.ms_range_kernel = kMs;
  stats.ms_range_kernel_max = kMax;
  stats.algo_bytes_max = bytesAtMax;
  stats.candidates = pairs;
  stats.algo_bytes = prevBytes;
  stats.eps_final = (float)P.g.eps;
  stats.iterations = passes;
  stats.reindexes = reindexes;
  stats.kernel_launches = passes;
  P.radius = nullptr;
}

// Where a no-allocation call takes its buffers: this device's copy of the static scratch array
// (set up on the first such call on each device).
static void *noalloc_scratch(size_t &bytes);
// Scratch a join of n points of d dims takes at most, besides the optional grid tables (the
// dense cell-start table and the learned schedule's feature grid, checked where they are
// laid out): two coordinate copies, 64-bit keys and ids with their ping-pong copies, flags,
// scans, cell and block tables, the active list and tiles, failure flags, radii, row sums.
static size_t noalloc_bound(uint32_t n, uint32_t d, size_t tsz) {
  const size_t perPoint = 2 * d * tsz + tsz + 16 + 8 + 8 + 8 + 4 + 4 + 4 + 16 + 3 + 8 + 24 + 4 + 12 + 1;
  return (size_t)n * perPoint + (radix_scratch_words(n) + scan_scratch_words(n + 1) + 2048) * 4 +
         avgdist_scratch_doubles(n) * 8 + (2u << 20);
}

template <typename T>
static int guarded_join(const T *pts, uint32_t n, uint32_t d, uint32_t k, const knnjoin_params &p,
                        int32_t *ids, T *dists, knnjoin_stats *st, bool noalloc = false) {
  std::lock_guard<std::mutex> lk(g_mu);
  Arena arena;
  struct ArenaScope {
    ~ArenaScope() { g_arena = nullptr; }
  } scope;
  try {
    if (noalloc) {
      if (p.device_ptrs == 0) throw Error(KNN_ERR_ARG, "the no-allocation entry point takes device pointers only");
      if (k + 1 > 85) throw Error(KNN_ERR_ARG, "the no-allocation entry point takes K+1 <= 85");
      size_t cap = 0;
      void *base = noalloc_scratch(cap);
      const size_t need = noalloc_bound(n, d, sizeof(T));
      if (need > cap)
        throw Error(KNN_ERR_NOMEM, "no-allocation join of " + std::to_string(n) + " points x " + std::to_string(d) +
                                       " dims needs up to " + std::to_string(need) + " bytes of scratch; the limit is " +
                                       std::to_string(cap));
      arena.base = reinterpret_cast<unsigned char *>(base);
      arena.cap = cap;
      g_arena = &arena;
    }
    int rc = run_join<T>(pts, n, d, k, p, ids, dists, st);
    g_status = rc;
    g_err.clear();
    return rc;
  } catch (const Error &e) {
    g_status = e.code;
    g_err = e.what();
    if (p.device_ptrs == 0 && e.code != KNN_ERR_ARG && ids && n && k + 1 > 0 &&
        e.code != KNN_ERR_HIP && e.code != KNN_ERR_NOMEM && e.code != KNN_ERR_NO_DEVICE) {
      // copy back the -1 rows of the unresolved queries (main.cu:260-268)
      int32_t *dIds = reinterpret_cast<int32_t *>(g_cache.p[S_OUTIDS]);
      if (dIds && g_stream) {
        (void)hipMemcpyAsync(ids, dIds, sizeof(int32_t) * (size_t)n * (k + 1), hipMemcpyDeviceToHost, g_stream);
        (void)hipStreamSynchronize(g_stream);
      }
    }
    if (e.code == KNN_ERR_TOO_FEW && ids && p.device_ptrs == 0) {
      for (size_t i = 0; i < (size_t)n * (k + 1); i++) ids[i] = -1;
    }
    return e.code;
  } catch (const std::exception &e) {
    g_status = KNN_ERR_HIP;
    g_err = e.what();
    return KNN_ERR_HIP;
  }
}

}  // namespace knn

// ================================================================== C ABI
extern "C" {

void knnjoin_default_params(knnjoin_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->num_indexed_dims = 0;
  p->beta = 0.05;  // params.h:58
  p->mode = KNN_MODE_BETA;
  p->reorder = 1;  // params.h:35
  p->brute_fallback = -1;
}

int knnjoin_ex(const void *points, int dtype, uint32_t n, uint32_t ndim, uint32_t k,
               const knnjoin_params *params, int32_t *out_ids, void *out_dists,
               knnjoin_stats *stats) {
  knnjoin_params p;
  if (params) p = *params; else knnjoin_default_params(&p);
  if (dtype == KNN_FLOAT32)
    return knn::guarded_join<float>(reinterpret_cast<const float *>(points), n, ndim, k, p, out_ids,
                                    reinterpret_cast<float *>(out_dists), stats);
  if (dtype == KNN_FLOAT64)
    return knn::guarded_join<double>(reinterpret_cast<const double *>(points), n, ndim, k, p, out_ids,
                                     reinterpret_cast<double *>(out_dists), stats);
  std::lock_guard<std::mutex> lk(knn::g_mu);
  knn::g_status = KNN_ERR_ARG;
  knn::g_err = "dtype must be KNN_FLOAT32 or KNN_FLOAT64";
  return KNN_ERR_ARG;
}

void KNNJoinPy(KNN_PY_DTYPE *dataset, unsigned int NUMPOINTS, unsigned int kNN, unsigned int NDIM,
               int *outNearestNeighborTable, KNN_PY_DTYPE *outNearestNeighborTableDistances) {
  knnjoin_params p;
  knnjoin_default_params(&p);
  p.num_indexed_dims = NDIM <= 6 ? NDIM : 6;
  (void)knnjoin_ex(dataset, knnjoin_py_dtype(), NUMPOINTS, NDIM, kNN, &p, outNearestNeighborTable,
                   outNearestNeighborTableDistances, nullptr);
}

int knnjoin_py_dtype(void) { return sizeof(KNN_PY_DTYPE) == 8 ? KNN_FLOAT64 : KNN_FLOAT32; }

void KNNJoinPyDouble(double *dataset, unsigned int NUMPOINTS, unsigned int kNN, unsigned int NDIM,
                     int *outNearestNeighborTable, double *outNearestNeighborTableDistances) {
  knnjoin_params p;
  knnjoin_default_params(&p);
  p.num_indexed_dims = NDIM <= 6 ? NDIM : 6;
  (void)knnjoin_ex(dataset, KNN_FLOAT64, NUMPOINTS, NDIM, kNN, &p, outNearestNeighborTable,
                   outNearestNeighborTableDistances, nullptr);
}

int knnjoin_ex_noalloc(const void *points, int dtype, uint32_t n, uint32_t ndim, uint32_t k,
                       const knnjoin_params *params, int32_t *out_ids, void *out_dists,
                       knnjoin_stats *stats) {
  knnjoin_params p;
  if (params) p = *params; else knnjoin_default_params(&p);
  if (dtype == KNN_FLOAT32)
    return knn::guarded_join<float>(reinterpret_cast<const float *>(points), n, ndim, k, p, out_ids,
                                    reinterpret_cast<float *>(out_dists), stats, true);
  if (dtype == KNN_FLOAT64)
    return knn::guarded_join<double>(reinterpret_cast<const double *>(points), n, ndim, k, p, out_ids,
                                     reinterpret_cast<double *>(out_dists), stats, true);
  std::lock_guard<std::mutex> lk(knn::g_mu);
  knn::g_status = KNN_ERR_ARG;
  knn::g_err = "dtype must be KNN_FLOAT32 or KNN_FLOAT64";
  return KNN_ERR_ARG;
}

int knnjoin_last_status(void) {
  std::lock_guard<std::mutex> lk(knn::g_mu);
  return knn::g_status;
}

const char *knnjoin_last_error(void) {
  std::lock_guard<std::mutex> lk(knn::g_mu);
  return knn::g_err.c_str();
}

void knnjoin_release(void) {
  std::lock_guard<std::mutex> lk(knn::g_mu);
  knn::g_cache.release();
  knn::g_hostPin.release();
  knn::release_batch();
}

}  // extern "C"

// ------------------------------------------------------------------ stage entry points
namespace knn {

template <typename F>
static int guarded(F &&f) {
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    int rc = f();
    g_status = rc;
    g_err.clear();
    return rc;
  } catch (const Error &e) {
    g_status = e.code;
    g_err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_status = KNN_ERR_HIP;
    g_err = e.what();
    return KNN_ERR_HIP;
  }
}

// uploads reordered row-major float points and builds the dim-major copy (identity order)
static void stage_setup(Pipeline<float> &P, const float *points, uint32_t n, uint32_t d,
                        uint32_t nd, uint32_t k) {
  if (!points || n == 0 || d == 0 || d > KNN_MAXDIM || nd == 0 || nd > d || nd > KNN_MAX_INDEXED)
    throw Error(KNN_ERR_ARG, "bad stage arguments");
  P.s = lib_stream();
  P.n = n;
  P.d = d;
  P.nd = nd;
  P.k = k;
  P.K1 = k + 1;
  P.scratchWords = std::max(scan_scratch_words(n + 1), radix_scratch_words(n)) + 1024;
  P.scratch = P.buf<uint32_t>(S_SCRATCH, P.scratchWords);
  float *b = P.buf<float>(S_PTS, (size_t)n * d);
  HIP_CHECK(hipMemcpyAsync(b, points, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice, P.s));
  uint32_t ident[KNN_MAXDIM];
  for (uint32_t j = 0; j < KNN_MAXDIM; j++) ident[j] = j;
  P.x = P.buf<float>(S_X, (size_t)n * d);
  permute_to_soa<float>(b, n, d, ident, P.x, P.s);
}

static void stage_minmax(Pipeline<float> &P, float *hmn, float *hmx) {
  float *mm = P.buf<float>(S_PART, 2 * KNN_MAX_INDEXED + minmax_scratch_elems(P.n, P.nd));
  minmax_soa<float>(P.x, P.n, P.nd, mm, mm + KNN_MAX_INDEXED, mm + 2 * KNN_MAX_INDEXED, P.s);
  HIP_CHECK(hipMemcpyAsync(hmn, mm, sizeof(float) * P.nd, hipMemcpyDeviceToHost, P.s));
  HIP_CHECK(hipMemcpyAsync(hmx, mm + KNN_MAX_INDEXED, sizeof(float) * P.nd, hipMemcpyDeviceToHost, P.s));
  HIP_CHECK(hipStreamSynchronize(P.s));
}

}  // namespace knn

extern "C" {

int knnjoin_stage_reorder(const float *points, uint32_t n, uint32_t ndim, uint32_t *perm,
                          float *variance) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (!points || !perm || n == 0 || ndim == 0 || ndim > KNN_MAXDIM)
      throw Error(KNN_ERR_ARG, "bad reorder arguments");
    hipStream_t s = lib_stream();
    float *b = reinterpret_cast<float *>(g_cache.get(S_PTS, sizeof(float) * (size_t)n * ndim));
    HIP_CHECK(hipMemcpyAsync(b, points, sizeof(float) * (size_t)n * ndim, hipMemcpyHostToDevice, s));
    float *var = reinterpret_cast<float *>(g_cache.get(S_SMALL, 64 * 1024));
    variance_rows<float>(b, n, ndim, var, s);
    float hv[KNN_MAXDIM];
    HIP_CHECK(hipMemcpyAsync(hv, var, sizeof(float) * ndim, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    sort_perm_desc<float>(hv, ndim, perm);
    if (variance)
      for (uint32_t j = 0; j < ndim; j++) variance[j] = hv[perm[j]];
    return KNN_OK;
  });
}

int knnjoin_stage_estimate_eps(const float *points, uint32_t n, uint32_t ndim, uint32_t k,
                               double beta, float *eps, double *avg_dist, uint64_t *pair_count,
                               uint32_t *bucket, uint32_t *hist) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    Pipeline<float> P;
    stage_setup(P, points, n, ndim, 1, k);
    EpsResult er = estimate_eps<float>(P.x, n, ndim, k, beta, hist, P.s);
    if (eps) *eps = (float)er.best;
    if (avg_dist) *avg_dist = er.avg;
    if (pair_count) *pair_count = er.cnt;
    if (bucket) *bucket = er.bucket;
    return ((float)er.best == 0.0f) ? KNN_ERR_EPS_ZERO : KNN_OK;
  });
}

int knnjoin_stage_index(const float *points, uint32_t n, uint32_t ndim, uint32_t nd, float eps,
                        float *min_arr, uint32_t *n_cells, uint64_t *total_cells,
                        uint32_t *sorted_ids, uint64_t *cell_keys, uint32_t *cell_start,
                        uint32_t *num_cells) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (!(eps > 0.0f)) throw Error(KNN_ERR_EPS_ZERO, "eps must be > 0");
    Pipeline<float> P;
    stage_setup(P, points, n, ndim, nd, 0);
    float hmn[KNN_MAX_INDEXED], hmx[KNN_MAX_INDEXED];
    stage_minmax(P, hmn, hmx);
    P.build_index(hmn, hmx, eps);
    if (min_arr) std::memcpy(min_arr, P.g.minArr, sizeof(float) * nd);
    if (n_cells) std::memcpy(n_cells, P.g.nCells, sizeof(uint32_t) * nd);
    if (total_cells) *total_cells = P.totalCells;
    if (num_cells) *num_cells = P.ncell;
    if (sorted_ids)
      HIP_CHECK(hipMemcpyAsync(sorted_ids, P.sortedIds, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, P.s));
    if (cell_keys)
      HIP_CHECK(hipMemcpyAsync(cell_keys, P.cellKeys, sizeof(uint64_t) * P.ncell, hipMemcpyDeviceToHost, P.s));
    if (cell_start)
      HIP_CHECK(hipMemcpyAsync(cell_start, P.cellStart, sizeof(uint32_t) * ((size_t)P.ncell + 1),
                               hipMemcpyDeviceToHost, P.s));
    HIP_CHECK(hipStreamSynchronize(P.s));
    return KNN_OK;
  });
}

int knnjoin_stage_join_pass(const float *points, uint32_t n, uint32_t ndim, uint32_t nd,
                            uint32_t k, float eps, const uint32_t *queries, uint32_t nq,
                            int32_t *out_ids, float *out_dists, uint8_t *failed,
                            uint32_t *num_failed) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (!(eps > 0.0f)) throw Error(KNN_ERR_EPS_ZERO, "eps must be > 0");
    if (!out_ids || !out_dists || !failed || (nq && !queries))
      throw Error(KNN_ERR_ARG, "null output");
    Pipeline<float> P;
    stage_setup(P, points, n, ndim, nd, k);
    float hmn[KNN_MAX_INDEXED], hmx[KNN_MAX_INDEXED];
    stage_minmax(P, hmn, hmx);
    P.build_index(hmn, hmx, eps);
    const uint32_t K1 = k + 1;
    std::vector<uint8_t> act(n, 0);
    for (uint32_t i = 0; i < nq; i++) {
      if (queries[i] >= n) throw Error(KNN_ERR_ARG, "query id out of range");
      act[queries[i]] = 1;
    }
    uint8_t *dAct = P.buf<uint8_t>(S_HOSTFLAGS, n);
    uint8_t *dFailed = P.buf<uint8_t>(S_FAILED, n);
    int32_t *dIds = P.buf<int32_t>(S_OUTIDS, (size_t)n * K1);
    float *dD = P.buf<float>(S_OUTDISTS, (size_t)n * K1);
    double *rowSum = P.buf<double>(S_ROWSUM, n);
    HIP_CHECK(hipMemcpyAsync(dAct, act.data(), n, hipMemcpyHostToDevice, P.s));
    HIP_CHECK(hipMemcpyAsync(dFailed, failed, n, hipMemcpyHostToDevice, P.s));
    HIP_CHECK(hipMemcpyAsync(dIds, out_ids, sizeof(int32_t) * (size_t)n * K1, hipMemcpyHostToDevice, P.s));
    HIP_CHECK(hipMemcpyAsync(dD, out_dists, sizeof(float) * (size_t)n * K1, hipMemcpyHostToDevice, P.s));
    unsigned long long *ctr = P.buf<unsigned long long>(S_CTR, 64);
    fill_u32(reinterpret_cast<uint32_t *>(ctr), 0u, 16, P.s);
    Timer tK(P.s, 3), tP(P.s, 6);
    PassResult pr = P.pass(dAct, nq, dFailed, dIds, dD, ctr, tK, tP);
    HIP_CHECK(hipMemcpyAsync(out_ids, dIds, sizeof(int32_t) * (size_t)n * K1, hipMemcpyDeviceToHost, P.s));
    HIP_CHECK(hipMemcpyAsync(out_dists, dD, sizeof(float) * (size_t)n * K1, hipMemcpyDeviceToHost, P.s));
    HIP_CHECK(hipMemcpyAsync(failed, dFailed, n, hipMemcpyDeviceToHost, P.s));
    HIP_CHECK(hipStreamSynchronize(P.s));
    if (num_failed) *num_failed = pr.nfail;
    return KNN_OK;
  });
}

}  // extern "C"

// ------------------------------------------------------------------ outlier scores
extern "C" {

int knnjoin_outlier_scores(const int32_t *ids, const float *dists, uint32_t n, uint32_t k1,
                           int device_ptrs, float *dist_sum, uint32_t *dist_rank, uint32_t *indegree,
                           uint32_t *indegree_rank) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (n == 0) return KNN_OK;
    if (!ids || !dists || !dist_sum || !dist_rank || !indegree || !indegree_rank || k1 == 0)
      throw Error(KNN_ERR_ARG, "bad outlier-score arguments");
    hipStream_t s = lib_stream();
    const size_t tab = (size_t)n * k1;
    const int32_t *dIds = ids;
    const float *dD = dists;
    if (!device_ptrs) {
      int32_t *a = reinterpret_cast<int32_t *>(g_cache.get(S_OUTIDS, sizeof(int32_t) * tab));
      float *b = reinterpret_cast<float *>(g_cache.get(S_OUTDISTS, sizeof(float) * tab));
      HIP_CHECK(hipMemcpyAsync(a, ids, sizeof(int32_t) * tab, hipMemcpyHostToDevice, s));
      HIP_CHECK(hipMemcpyAsync(b, dists, sizeof(float) * tab, hipMemcpyHostToDevice, s));
      dIds = a;
      dD = b;
    }
    const size_t sw = radix_scratch_words(n) + 1024;
    uint32_t *w = reinterpret_cast<uint32_t *>(g_cache.get(S_TRAIN, sizeof(uint32_t) * ((size_t)8 * n + sw)));
    float *ds = reinterpret_cast<float *>(w);
    uint32_t *dr = w + n, *dg = w + 2 * (size_t)n, *gr = w + 3 * (size_t)n;
    uint32_t *k0 = w + 4 * (size_t)n, *v0 = w + 5 * (size_t)n, *k1t = w + 6 * (size_t)n, *v1t = w + 7 * (size_t)n;
    uint32_t *scr = w + 8 * (size_t)n;
    outlier_scores(dIds, dD, n, k1, ds, dr, dg, gr, k0, v0, k1t, v1t, scr, s);
    const hipMemcpyKind kout = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_CHECK(hipMemcpyAsync(dist_sum, ds, sizeof(float) * n, kout, s));
    HIP_CHECK(hipMemcpyAsync(dist_rank, dr, sizeof(uint32_t) * n, kout, s));
    HIP_CHECK(hipMemcpyAsync(indegree, dg, sizeof(uint32_t) * n, kout, s));
    HIP_CHECK(hipMemcpyAsync(indegree_rank, gr, sizeof(uint32_t) * n, kout, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return KNN_OK;
  });
}

}  // extern "C"

// ------------------------------------------------------------------ learned-radius entries
extern "C" {

int knnjoin_set_radius_model_dims(uint32_t ndim, uint32_t nd, const float *weights, uint32_t nweights,
                                  float margin) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (nd == 0 || nd > KNN_MAX_INDEXED) throw Error(KNN_ERR_ARG, "nd must be in 1..8");
    if (ndim > KNN_MAXDIM || (ndim != 0 && ndim < nd)) throw Error(KNN_ERR_ARG, "ndim must be 0 or in nd..16");
    if (!weights || nweights != (uint32_t)MLP_NW) throw Error(KNN_ERR_MODEL, "radius model must have 5057 weights");
    RadiusModel &m = g_models[ndim][nd];
    m.w.assign(weights, weights + MLP_NW);
    m.margin = margin;
    m.ok = true;
    return KNN_OK;
  });
}

int knnjoin_set_radius_model(uint32_t nd, const float *weights, uint32_t nweights, float margin) {
  return knnjoin_set_radius_model_dims(0, nd, weights, nweights, margin);
}

int knnjoin_radius_features(const float *points, uint32_t n, uint32_t ndim, uint32_t k, uint32_t nd,
                            int reorder, float *features, float *wf_out, float *radius_out) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (!points || n == 0 || ndim == 0 || ndim > KNN_MAXDIM || nd == 0 || nd > ndim || nd > KNN_MAX_INDEXED)
      throw Error(KNN_ERR_ARG, "bad feature arguments");
    Pipeline<float> P;
    P.s = lib_stream();
    P.n = n;
    P.d = ndim;
    P.nd = nd;
    P.k = k;
    P.K1 = k + 1;
    hipStream_t s = P.s;
    float *b = P.buf<float>(S_PTS, (size_t)n * ndim);
    HIP_CHECK(hipMemcpyAsync(b, points, sizeof(float) * (size_t)n * ndim, hipMemcpyHostToDevice, s));
    uint32_t perm[KNN_MAXDIM];
    for (uint32_t j = 0; j < KNN_MAXDIM; j++) perm[j] = j;
    if (reorder) {
      float *var = P.buf<float>(S_SMALL, 16 * 1024);
      variance_rows<float>(b, n, ndim, var, s);
      float hv[KNN_MAXDIM];
      HIP_CHECK(hipMemcpyAsync(hv, var, sizeof(float) * ndim, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      sort_perm_desc<float>(hv, ndim, perm);
    }
    P.x = P.buf<float>(S_X, (size_t)n * ndim);
    permute_to_soa<float>(b, n, ndim, perm, P.x, s);
    float *mm = P.buf<float>(S_PART, 2 * KNN_MAX_INDEXED + minmax_scratch_elems(n, nd));
    minmax_soa<float>(P.x, n, nd, mm, mm + KNN_MAX_INDEXED, mm + 2 * KNN_MAX_INDEXED, s);
    float hmn[KNN_MAX_INDEXED], hmx[KNN_MAX_INDEXED];
    HIP_CHECK(hipMemcpyAsync(hmn, mm, sizeof(float) * nd, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(hmx, mm + KNN_MAX_INDEXED, sizeof(float) * nd, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    FeatureGrid<float> F;
    feature_grid<float>(hmn, hmx, nd, n, k + 1, F);
    const uint32_t reps = feature_reps(F.cells);
    uint32_t *cnt = P.buf<uint32_t>(S_FCNT, (size_t)F.cells * reps);
    fill_u32(cnt, 0u, (size_t)F.cells * reps, s);
    feature_counts<float>(P.x, n, F.g, cnt, F.cells, reps, s);
    feature_occupancy(cnt, F.cells, n, F.logdens, s);
    float *fd = features ? P.buf<float>(S_TRAIN, (size_t)n * MLP_IN) : nullptr;
    float *rd = nullptr;
    float *Wd = nullptr;
    const RadiusModel &M = radius_model(ndim, nd);
    if (radius_out) {
      if (!M.ok) throw Error(KNN_ERR_MODEL, "no radius model registered for nd");
      rd = P.buf<float>(S_RADIUS, n);
      Wd = P.buf<float>(S_MODEL, MLP_NW);
      HIP_CHECK(hipMemcpyAsync(Wd, M.w.data(), sizeof(float) * MLP_NW, hipMemcpyHostToDevice, s));
    }
    // features in FP32 (training data); the radius with the join's own inference kernel
    if (fd)
      radius_kernel<float>(P.x, n, F.g, cnt, F.logk1, F.logdens, nullptr, 0.0f, (float)F.wf, nullptr, fd,
                           nullptr, s);
    if (rd)
      radius_kernel<float>(P.x, n, F.g, cnt, F.logk1, F.logdens, Wd, M.margin, (float)F.wf, rd, nullptr,
                           nullptr, s);
    if (features)
      HIP_CHECK(hipMemcpyAsync(features, fd, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyDeviceToHost, s));
    if (radius_out)
      HIP_CHECK(hipMemcpyAsync(radius_out, rd, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (wf_out) *wf_out = (float)F.wf;
    return KNN_OK;
  });
}

int knnjoin_mlp_forward(const float *weights, const float *X, uint32_t B, float *out) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (!weights || !X || !out) throw Error(KNN_ERR_ARG, "null argument");
    hipStream_t s = lib_stream();
    float *buf = reinterpret_cast<float *>(g_cache.get(S_TRAIN, sizeof(float) * ((size_t)MLP_NW + (size_t)B * (MLP_IN + 2 * MLP_H + 1))));
    float *Wd = buf, *Xd = Wd + MLP_NW, *H1 = Xd + (size_t)B * MLP_IN, *H2 = H1 + (size_t)B * MLP_H, *o = H2 + (size_t)B * MLP_H;
    HIP_CHECK(hipMemcpyAsync(Wd, weights, sizeof(float) * MLP_NW, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(Xd, X, sizeof(float) * (size_t)B * MLP_IN, hipMemcpyHostToDevice, s));
    mlp_forward_batch(Wd, Xd, B, H1, H2, o, s);
    HIP_CHECK(hipMemcpyAsync(out, o, sizeof(float) * B, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return KNN_OK;
  });
}

int knnjoin_mlp_grad(const float *weights, const float *X, const float *Y, uint32_t B, float *grads,
                     float *loss) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (!weights || !X || !Y || !grads) throw Error(KNN_ERR_ARG, "null argument");
    hipStream_t s = lib_stream();
    const size_t words = (size_t)2 * MLP_NW + 1 + (size_t)B * (MLP_IN + 1 + 4 * MLP_H + 2);
    float *buf = reinterpret_cast<float *>(g_cache.get(S_TRAIN, sizeof(float) * words));
    float *Wd = buf, *Gd = Wd + MLP_NW, *Ld = Gd + MLP_NW, *Xd = Ld + 1, *Yd = Xd + (size_t)B * MLP_IN;
    float *H1 = Yd + B, *H2 = H1 + (size_t)B * MLP_H, *dH1 = H2 + (size_t)B * MLP_H, *dH2 = dH1 + (size_t)B * MLP_H;
    float *o = dH2 + (size_t)B * MLP_H, *dout = o + B;
    HIP_CHECK(hipMemcpyAsync(Wd, weights, sizeof(float) * MLP_NW, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(Xd, X, sizeof(float) * (size_t)B * MLP_IN, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(Yd, Y, sizeof(float) * B, hipMemcpyHostToDevice, s));
    mlp_grad_batch(Wd, Xd, Yd, B, H1, H2, o, dout, dH1, dH2, Gd, Ld, s);
    HIP_CHECK(hipMemcpyAsync(grads, Gd, sizeof(float) * MLP_NW, hipMemcpyDeviceToHost, s));
    if (loss) HIP_CHECK(hipMemcpyAsync(loss, Ld, sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return KNN_OK;
  });
}

// Adam over mini-batches of a shuffled epoch order (splitmix64 Fisher-Yates, seeded); the
// weights are updated in place, loss_per_epoch[e] = mean batch loss of epoch e.
int knnjoin_mlp_train(float *weights, const float *X, const float *Y, uint32_t n, uint32_t epochs,
                      uint32_t batch, float lr, uint64_t seed, float *loss_per_epoch) {
  return knn::guarded([&]() -> int {
    using namespace knn;
    if (!weights || !X || !Y || n == 0 || batch == 0) throw Error(KNN_ERR_ARG, "bad training arguments");
    hipStream_t s = lib_stream();
    const uint32_t B = std::min(batch, n);
    const size_t words = (size_t)4 * MLP_NW + 1 + (size_t)n * (MLP_IN + 1) +
                         (size_t)B * (MLP_IN + 1 + 4 * MLP_H + 2) + n;
    float *buf = reinterpret_cast<float *>(g_cache.get(S_TRAIN, sizeof(float) * words));
    float *Wd = buf, *Gd = Wd + MLP_NW, *Md = Gd + MLP_NW, *Vd = Md + MLP_NW, *Ld = Vd + MLP_NW;
    float *Xall = Ld + 1, *Yall = Xall + (size_t)n * MLP_IN;
    float *Xb = Yall + n, *Yb = Xb + (size_t)B * MLP_IN;
    float *H1 = Yb + B, *H2 = H1 + (size_t)B * MLP_H, *dH1 = H2 + (size_t)B * MLP_H, *dH2 = dH1 + (size_t)B * MLP_H;
    float *o = dH2 + (size_t)B * MLP_H, *dout = o + B;
    uint32_t *idx = reinterpret_cast<uint32_t *>(dout + B);
    HIP_CHECK(hipMemcpyAsync(Wd, weights, sizeof(float) * MLP_NW, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(Xall, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(Yall, Y, sizeof(float) * n, hipMemcpyHostToDevice, s));
    fill_u32(reinterpret_cast<uint32_t *>(Md), 0u, 2 * (size_t)MLP_NW, s);
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    uint64_t st = seed;
    auto next = [&]() {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    };
    const float b1 = 0.9f, b2 = 0.999f, epsA = 1e-8f;
    uint64_t step = 0;
    const uint32_t nb = n / B;
    for (uint32_t e = 0; e < epochs; e++) {
      for (uint32_t i = n - 1; i > 0; i--) {
        const uint32_t j = (uint32_t)(next() % (uint64_t)(i + 1));
        std::swap(order[i], order[j]);
      }
      HIP_CHECK(hipMemcpyAsync(idx, order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
      double lsum = 0;
      for (uint32_t bi = 0; bi < nb; bi++) {
        gather_rows(Xall, idx + (size_t)bi * B, B, MLP_IN, Xb, s);
        gather_rows(Yall, idx + (size_t)bi * B, B, 1, Yb, s);
        mlp_grad_batch(Wd, Xb, Yb, B, H1, H2, o, dout, dH1, dH2, Gd, Ld, s);
        step++;
        const float c1 = (float)(1.0 - std::pow(0.9, (double)step));
        const float c2 = (float)(1.0 - std::pow(0.999, (double)step));
        adam_step(Wd, Gd, Md, Vd, lr, b1, b2, c1, c2, epsA, s);
        float hl = 0;
        if (loss_per_epoch) {
          HIP_CHECK(hipMemcpyAsync(&hl, Ld, sizeof(float), hipMemcpyDeviceToHost, s));
          HIP_CHECK(hipStreamSynchronize(s));
          lsum += hl;
        }
      }
      if (loss_per_epoch) loss_per_epoch[e] = nb ? (float)(lsum / nb) : 0.0f;
    }
    HIP_CHECK(hipMemcpyAsync(weights, Wd, sizeof(float) * MLP_NW, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return KNN_OK;
  });
}

}  // extern "C"

// ================================================================== batched sets
// Host side of batched.hip: a plan per set-shape signature (allocations, descriptors, and the
// captured hipGraph of the whole join), replayed while the caller's shapes stay the same.
namespace knn {

struct BatchGroup {
  uint32_t d = 0, nsets = 0, N = 0, k1max = 0, smax = 0;
  uint64_t outElems = 0, bound = 0;
  int bits = 0;
  std::vector<uint32_t> members;  // caller set indices
  std::vector<SetDesc> hsets;
  float *xin = nullptr, *x = nullptr, *xs = nullptr, *radius = nullptr, *width = nullptr,
        *mnmx = nullptr, *var = nullptr, *outDists = nullptr;
  uint32_t *perm = nullptr, *setOf = nullptr, *keys = nullptr, *keysTmp = nullptr, *ids = nullptr,
           *idsTmp = nullptr, *cflags = nullptr, *cpos = nullptr, *cellStart = nullptr,
           *dense = nullptr, *aflags = nullptr, *apos = nullptr, *act = nullptr,
           *ntiles = nullptr, *tileOff = nullptr, *scratch = nullptr, *status = nullptr,
           *setMaxBits = nullptr, *ctr = nullptr;
  uint64_t *cellKeys = nullptr;
  unsigned long long *cand = nullptr;
  uint4 *tiles = nullptr;
  uint8_t *failed = nullptr, *overflow = nullptr, *bin = nullptr, *actq = nullptr;
  uint32_t *setMinBin = nullptr;
  SetDesc *sets = nullptr;
  int32_t *outIds = nullptr;
  void *heap = nullptr;
};

struct BatchPlan {
  std::vector<uint64_t> sig;
  std::vector<BatchGroup> groups;
  std::vector<void *> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t tq = 64, grid = 2048, passes = 3;
  int reorder = 1;
  bool useGraph = true;
  template <typename U>
  U *alloc(size_t count) {
    void *p = nullptr;
    HIP_CHECK(hipMalloc(&p, std::max<size_t>(count * sizeof(U), 256)));
    allocs.push_back(p);
    return reinterpret_cast<U *>(p);
  }
  // The graph's kernel nodes hold pointers into `allocs`: it is destroyed first, and only
  // once the device has finished every launch of it, then the buffers are freed.
  void drop_graph() {
    if (exec || graph) HIP_CHECK(hipDeviceSynchronize());
    hipGraphExec_t e = exec;
    hipGraph_t g = graph;
    exec = nullptr;
    graph = nullptr;
    if (e) HIP_CHECK(hipGraphExecDestroy(e));
    if (g) HIP_CHECK(hipGraphDestroy(g));
  }
  void free_all() {
    drop_graph();
    if (!allocs.empty()) HIP_CHECK(hipDeviceSynchronize());
    std::vector<void *> a;
    a.swap(allocs);
    groups.clear();
    sig.clear();
    for (void *p : a) HIP_CHECK(hipFree(p));
  }
};

// Ends a stream capture that an exception interrupts, so that the stream leaves capture mode
// and later calls can use it; the partial graph is discarded.
struct CaptureGuard {
  hipStream_t s;
  bool open = true;
  explicit CaptureGuard(hipStream_t st) : s(st) {}
  hipGraph_t end() {
    hipGraph_t g = nullptr;
    open = false;
    HIP_CHECK(hipStreamEndCapture(s, &g));
    return g;
  }
  ~CaptureGuard() {
    if (!open) return;
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
  }
};
static BatchPlan g_batch;
// tuned (tile queries, grid blocks) of the last autotuned signature
static std::vector<uint64_t> g_tunedSig;
static uint32_t g_tunedTq = 0, g_tunedGrid = 0;

constexpr uint32_t kBatchMaxGrid = 8192;

static void batch_build(BatchPlan &B, uint32_t nsets, const uint32_t *n, const uint32_t *ndim,
                        const uint32_t *k, uint32_t ndParam, hipStream_t s) {
  B.free_all();
  std::vector<int> groupOf(KNN_MAXDIM + 1, -1);
  for (uint32_t si = 0; si < nsets; si++) {
    const uint32_t d = ndim[si];
    if (d == 0 || d > KNN_MAXDIM) throw Error(KNN_ERR_ARG, "set dimensionality must be in 1..16");
    const uint32_t nd = ndParam ? ndParam : std::min<uint32_t>(d, 6);
    if (nd > d || nd > KNN_MAX_INDEXED) throw Error(KNN_ERR_ARG, "NUMINDEXEDDIM must be <= NDIM and <= 8");
    if (n[si] != 0 && (uint64_t)n[si] < (uint64_t)k[si] + 1)
      throw Error(KNN_ERR_TOO_FEW, "a set has fewer than K+1 points");
    if (n[si] == 0) continue;
    if (groupOf[d] < 0) {
      groupOf[d] = (int)B.groups.size();
      B.groups.emplace_back();
      B.groups.back().d = d;
    }
    BatchGroup &G = B.groups[groupOf[d]];
    SetDesc sd{};
    sd.pStart = G.N;
    sd.n = n[si];
    sd.k1 = k[si] + 1;
    sd.nd = nd;
    sd.outOff = G.outElems;
    if ((uint64_t)G.N + n[si] > 0x7fffffffull) throw Error(KNN_ERR_ARG, "more than 2^31 points of one dimensionality");
    G.N += n[si];
    G.outElems += (uint64_t)n[si] * sd.k1;
    G.k1max = std::max(G.k1max, sd.k1);
    G.hsets.push_back(sd);
    G.members.push_back(si);
  }
  for (BatchGroup &G : B.groups) {
    G.nsets = (uint32_t)G.hsets.size();
    const uint32_t N = G.N, d = G.d;
    G.bound = std::min<uint64_t>(1ull << 27, std::max<uint64_t>(1ull << 16, 8ull * N));
    G.bits = 1;
    while ((1ull << G.bits) < G.bound) G.bits++;
    G.smax = std::min<uint32_t>(2 * G.k1max + 32, 512);
    const size_t sw = std::max({scan_scratch_words((uint32_t)G.bound + 1), scan_scratch_words(N + 2),
                                radix_scratch_words(N)}) + 1024;
    G.xin = B.alloc<float>((size_t)N * d);
    G.x = B.alloc<float>((size_t)N * d);
    G.xs = B.alloc<float>((size_t)N * d);
    G.radius = B.alloc<float>(N);
    G.width = B.alloc<float>(G.nsets);
    G.mnmx = B.alloc<float>((size_t)G.nsets * 16);
    G.var = B.alloc<float>((size_t)G.nsets * d);
    G.perm = B.alloc<uint32_t>((size_t)G.nsets * KNN_MAXDIM);
    G.setOf = B.alloc<uint32_t>(N);
    G.keys = B.alloc<uint32_t>(N);
    G.keysTmp = B.alloc<uint32_t>(N);
    G.ids = B.alloc<uint32_t>(N);
    G.idsTmp = B.alloc<uint32_t>(N);
    G.cflags = B.alloc<uint32_t>((size_t)N + 2);
    G.cpos = B.alloc<uint32_t>((size_t)N + 2);
    G.cellKeys = B.alloc<uint64_t>((size_t)N + 1);
    G.cellStart = B.alloc<uint32_t>((size_t)N + 2);
    G.dense = B.alloc<uint32_t>(G.bound + 2);
    G.aflags = B.alloc<uint32_t>((size_t)N + 2);
    G.apos = B.alloc<uint32_t>((size_t)N + 2);
    G.act = B.alloc<uint32_t>((size_t)N + 1);
    G.ntiles = B.alloc<uint32_t>((size_t)N + 2);
    G.tileOff = B.alloc<uint32_t>((size_t)N + 3);
    G.tiles = B.alloc<uint4>((size_t)N + 1);
    G.scratch = B.alloc<uint32_t>(sw);
    G.status = B.alloc<uint32_t>(4);
    G.setMaxBits = B.alloc<uint32_t>(G.nsets);
    G.ctr = B.alloc<uint32_t>(8 + TILE_QUEUES);  // [8..]: the tile queue heads
    G.cand = B.alloc<unsigned long long>(4);
    G.failed = B.alloc<uint8_t>(N);
    G.overflow = B.alloc<uint8_t>(N);
    G.bin = B.alloc<uint8_t>(N);
    G.actq = B.alloc<uint8_t>(N);
    G.setMinBin = B.alloc<uint32_t>(G.nsets);
    G.sets = B.alloc<SetDesc>(G.nsets);
    G.outIds = B.alloc<int32_t>(G.outElems);
    G.outDists = B.alloc<float>(G.outElems);
    const size_t heapBytes = (size_t)G.k1max * 64 * 12;
    G.heap = heapBytes > 65536 ? B.alloc<unsigned char>(heapBytes * kBatchMaxGrid) : nullptr;
    std::vector<uint32_t> so(N);
    for (uint32_t i = 0; i < G.nsets; i++)
      std::fill(so.begin() + G.hsets[i].pStart, so.begin() + G.hsets[i].pStart + G.hsets[i].n, i);
    HIP_CHECK(hipMemcpyAsync(G.setOf, so.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(G.sets, G.hsets.data(), sizeof(SetDesc) * G.nsets, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
}

static JoinArgs batch_args(const BatchPlan &B, const BatchGroup &G) {
  JoinArgs a{};
  a.x = G.x;
  a.xs = G.xs;
  a.sortedIds = G.ids;
  a.cellKeys = G.cellKeys;
  a.cellStart = G.cellStart;
  a.ncell = 0;
  a.denseStart = G.dense;
  a.act = G.act;
  a.tiles = G.tiles;
  a.ntiles = B.grid;
  a.n = G.N;
  a.d = G.d;
  a.k1 = G.k1max;
  a.radius = G.radius;
  a.outIds = G.outIds;
  a.outDists = G.outDists;
  a.failed = G.failed;
  a.failCount = G.ctr + 1;
  a.candCount = G.cand;
  a.tileCounter = G.ctr + 8;
  a.tileRanges = tile_ranges(G.d, true);
  a.segCapLog2 = seg_cap_log2();
  a.heapScratch = G.heap;
  a.smax = G.smax;
  a.overflow = G.overflow;
  a.overflowCount = G.ctr + 2;
  a.sets = G.sets;
  a.setOf = G.setOf;
  a.ntilesDev = G.tileOff + G.N + 1;
  return a;
}

// Index of a group at the sets' current widths: per-set grids, cell keys, radix sort, cell
// table, cell-sorted coordinates, dense cell-start table.
static void batch_index(BatchGroup &G, hipStream_t s) {
  const uint32_t N = G.N;
  batch_grids(G.sets, G.nsets, G.mnmx, G.width, G.bound, G.status, s);
  batch_cell_keys(G.x, N, G.sets, G.setOf, G.keys, G.ids, s);
  radix_sort_pairs_u32(G.keys, G.ids, G.keysTmp, G.idsTmp, N, G.bits, G.scratch, s);
  boundary_flags_u32(G.keys, N, 0u, G.cflags, s);
  scan_exclusive_u32(G.cflags, G.cpos, N, G.scratch, s);
  compact_cells_u32(G.keys, G.cflags, G.cpos, N, 0u, G.cellKeys, G.cellStart, s);
  gather_soa<float>(G.x, G.ids, N, G.d, G.xs, s);
  fill_u32(G.dense, 0u, G.bound + 1, s);
  batch_dense_counts(G.cellKeys, G.cellStart, G.cpos + N, N, G.dense, s);
  scan_exclusive_u32(G.dense, G.dense, (uint32_t)G.bound, G.scratch, s);
}

// One radius pass of a group: every set's lowest pending octave, grids matched to it, tiles of
// those queries, the select kernel (the collect kernel with 5+ indexed dims), the heap kernel
// for staging overflows, and radius growth of the queries that failed; G.ctr[3] = queries
// still pending.
static void batch_pass(const BatchPlan &B, BatchGroup &G, hipStream_t s) {
  const uint32_t N = G.N;
  batch_class(G.sets, G.nsets, G.setOf, G.failed, G.bin, N, G.setMinBin, G.width, G.actq, s);
  batch_index(G, s);

  const JoinArgs a = batch_args(B, G);
  // every tile takes its set's grid; g0 carries only the group's number of indexed dims (one
  // per group: it depends on the dimensionality alone), which picks the kernel build
  GridParams g0{};
  g0.nd = G.hsets.empty() ? 0u : G.hsets[0].nd;
  batch_flags_u8(G.ids, G.actq, N, G.aflags, s);
  scan_exclusive_u32(G.aflags, G.apos, N, G.scratch, s);
  compact_active(G.ids, G.aflags, G.apos, N, G.act, s);
  batch_tiles(G.cellStart, G.apos, G.cpos + N, N, B.tq, G.ntiles, G.tileOff, G.scratch, G.tiles, s);
  fill_u32(G.ctr, 0u, 8 + TILE_QUEUES, s);
  if (use_collect(G.hsets.empty() ? 0u : G.hsets[0].nd, G.k1max))
    range_collect<float>(a, g0, s);
  else
    range_select<float>(a, g0, s);

  batch_flags_u8(G.ids, G.overflow, N, G.aflags, s);
  scan_exclusive_u32(G.aflags, G.apos, N, G.scratch, s);
  compact_active(G.ids, G.aflags, G.apos, N, G.act, s);
  batch_tiles(G.cellStart, G.apos, G.cpos + N, N, B.tq, G.ntiles, G.tileOff, G.scratch, G.tiles, s);
  fill_u32(G.ctr + 8, 0u, TILE_QUEUES, s);
  if (overflow_by_collect(G.k1max))
    range_collect<float>(a, g0, s);
  else
    range_topk<float>(a, g0, s);
  fill_u8(G.overflow, 0, N, s);

  batch_grow_class(G.sets, G.setOf, G.actq, G.failed, N, G.radius, G.bin, G.ctr + 3, s);
}

static void batch_body(const BatchPlan &B, hipStream_t s) {
  for (BatchGroup &G : const_cast<BatchPlan &>(B).groups) {
    batch_variance_perm(G.xin, G.sets, G.nsets, G.d, B.reorder, G.var, G.perm, s);
    batch_permute(G.xin, G.N, G.d, G.setOf, G.perm, G.x, s);
    batch_minmax(G.x, G.N, G.sets, G.nsets, G.mnmx, s);
    batch_init_width(G.sets, G.nsets, G.mnmx, G.width, G.N, G.failed, s);
    fill_u8(G.overflow, 0, G.N, s);
    fill_u32(reinterpret_cast<uint32_t *>(G.cand), 0u, 8, s);
    batch_index(G, s);
    batch_density_radius(G.x, G.N, G.sets, G.setOf, G.dense, G.radius, G.bin, s);
    for (uint32_t p = 0; p < B.passes; p++) batch_pass(B, G, s);
  }
}

static int run_batched(uint32_t nsets, const float *const *points, const uint32_t *n,
                       const uint32_t *ndim, const uint32_t *k, const knnjoin_batch_params &p,
                       int32_t *const *outIds, float *const *outDists, knnjoin_batch_stats *st) {
  if (nsets == 0) {
    if (st) *st = knnjoin_batch_stats{};
    return KNN_OK;
  }
  if (!points || !n || !ndim || !k || !outIds || !outDists) throw Error(KNN_ERR_ARG, "null argument");
  hipStream_t s = lib_stream();
  knnjoin_batch_stats stats{};
  BatchPlan &B = g_batch;
  std::vector<uint64_t> sig;
  sig.reserve((size_t)nsets * 3 + 4);
  for (uint32_t i = 0; i < nsets; i++) {
    sig.push_back(n[i]);
    sig.push_back(((uint64_t)ndim[i] << 32) | k[i]);
  }
  sig.push_back(p.num_indexed_dims);
  uint32_t tq = p.tile_queries, grid = p.grid_blocks;
  if ((!tq || !grid) && sig == g_tunedSig) {
    if (!tq) tq = g_tunedTq;
    if (!grid) grid = g_tunedGrid;
  }
  // defaults: the best pair of the autotuner's sweep on the 256-set mix of 2-6 dims (the 6-D
  // group on the collect kernel favours small tiles and a wide grid: 79 ms against 131 ms at
  // 64 x 2048; the other groups change by under 5 %)
  if (!tq) tq = 16;
  if (!grid) grid = 4096;
  if (tq > 64) throw Error(KNN_ERR_ARG, "tile_queries must be <= 64");
  if (grid > kBatchMaxGrid) throw Error(KNN_ERR_ARG, "grid_blocks must be <= 8192");
  const uint32_t passes = p.graph_passes ? p.graph_passes : 8;
  Timer tAll(s, 4);
  auto w0 = std::chrono::steady_clock::now();
  if (B.sig != sig) {
    batch_build(B, nsets, n, ndim, k, p.num_indexed_dims, s);
    B.sig = sig;
  }
  if (B.groups.empty()) {  // every set is empty
    if (st) *st = stats;
    return KNN_OK;
  }
  if (B.tq != tq || B.grid != grid || B.passes != passes || B.reorder != (p.reorder ? 1 : 0) ||
      B.useGraph != (p.use_graph != 0)) {
    B.drop_graph();
    B.tq = tq;
    B.grid = grid;
    B.passes = passes;
    B.reorder = p.reorder ? 1 : 0;
    B.useGraph = p.use_graph != 0;
  }
  // inputs
  const hipMemcpyKind kin = p.device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  for (BatchGroup &G : B.groups)
    for (uint32_t i = 0; i < G.nsets; i++) {
      const uint32_t si = G.members[i];
      if (!points[si]) throw Error(KNN_ERR_ARG, "null set pointer");
      HIP_CHECK(hipMemcpyAsync(G.xin + (size_t)G.hsets[i].pStart * G.d, points[si],
                               sizeof(float) * (size_t)n[si] * G.d, kin, s));
    }
  bool replay = false;
  if (B.useGraph && !B.exec) {
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    {
      CaptureGuard cap(s);
      batch_body(B, s);
      B.graph = cap.end();
    }
    hipGraphExec_t ex = nullptr;
    HIP_CHECK(hipGraphInstantiate(&ex, B.graph, nullptr, nullptr, 0));
    B.exec = ex;
  } else if (B.useGraph) {
    replay = true;
  }
  HIP_CHECK(hipStreamSynchronize(s));
  stats.ms_setup = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
  tAll.start();
  if (B.useGraph) {
    HIP_CHECK(hipGraphLaunch(B.exec, s));
  } else {
    batch_body(B, s);
    LAUNCH_CHECK();
  }
  // leftover failures: eager passes until none remain
  uint32_t eager = 0;
  for (;;) {
    bool any = false;
    for (BatchGroup &G : B.groups) {
      uint32_t f = 0;  // queries still pending after the last pass
      HIP_CHECK(hipMemcpyAsync(&f, G.ctr + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (f) {
        any = true;
        batch_pass(B, G, s);
        LAUNCH_CHECK();
      }
    }
    if (!any) break;
    if (++eager > 200) throw Error(KNN_ERR_NO_PROGRESS, "batched join: queries still failing after 200 extra passes");
  }
  stats.ms_device = tAll.stop();
  for (BatchGroup &G : B.groups) {
    unsigned long long c[2] = {0, 0};
    HIP_CHECK(hipMemcpyAsync(c, G.cand, sizeof(c), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    stats.candidates += c[0];
    stats.algo_bytes += c[1];
    stats.points += G.N;
  }
  // outputs
  const hipMemcpyKind kout = p.device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  for (BatchGroup &G : B.groups)
    for (uint32_t i = 0; i < G.nsets; i++) {
      const uint32_t si = G.members[i];
      const SetDesc &sd = G.hsets[i];
      const size_t cnt = (size_t)sd.n * sd.k1;
      if (!outIds[si] || !outDists[si]) throw Error(KNN_ERR_ARG, "null output pointer");
      HIP_CHECK(hipMemcpyAsync(outIds[si], G.outIds + sd.outOff, sizeof(int32_t) * cnt, kout, s));
      HIP_CHECK(hipMemcpyAsync(outDists[si], G.outDists + sd.outOff, sizeof(float) * cnt, kout, s));
    }
  HIP_CHECK(hipStreamSynchronize(s));
  stats.groups = (uint32_t)B.groups.size();
  stats.graph_replayed = replay ? 1u : 0u;
  stats.passes_graph = B.passes;
  stats.passes_eager = eager;
  stats.tile_queries = B.tq;
  stats.grid_blocks = B.grid;
  if (p.verbose)
    printf("[knn] batched: %u sets in %u groups, %llu points, device %.3f ms (%s), eager passes %u\n",
           nsets, stats.groups, (unsigned long long)stats.points, stats.ms_device,
           B.useGraph ? (replay ? "graph replay" : "graph, first launch") : "eager", eager);
  if (st) *st = stats;
  return KNN_OK;
}

}  // namespace knn

extern "C" {

void knnjoin_batch_default_params(knnjoin_batch_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->reorder = 1;
  p->use_graph = 1;
}

int knnjoin_batched(uint32_t nsets, const float *const *points, const uint32_t *n, const uint32_t *ndim,
                    const uint32_t *k, const knnjoin_batch_params *params, int32_t *const *out_ids,
                    float *const *out_dists, knnjoin_batch_stats *stats) {
  knnjoin_batch_params p;
  if (params) p = *params; else knnjoin_batch_default_params(&p);
  return knn::guarded([&]() -> int {
    return knn::run_batched(nsets, points, n, ndim, k, p, out_ids, out_dists, stats);
  });
}

int knnjoin_batch_autotune(uint32_t nsets, const float *const *points, const uint32_t *n,
                           const uint32_t *ndim, const uint32_t *k, const knnjoin_batch_params *params,
                           int32_t *const *out_ids, float *const *out_dists, uint32_t *best_tq,
                           uint32_t *best_grid, double *ms_table, uint32_t *nconfigs) {
  knnjoin_batch_params p;
  if (params) p = *params; else knnjoin_batch_default_params(&p);
  return knn::guarded([&]() -> int {
    using namespace knn;
    static const uint32_t tqs[] = {16, 32, 64};
    static const uint32_t grids[] = {1024, 2048, 4096, 8192};
    double best = 1e300;
    uint32_t bt = 64, bg = 2048, c = 0;
    for (uint32_t tq : tqs)
      for (uint32_t gr : grids) {
        knnjoin_batch_params q = p;
        q.tile_queries = tq;
        q.grid_blocks = gr;
        knnjoin_batch_stats st{};
        run_batched(nsets, points, n, ndim, k, q, out_ids, out_dists, &st);  // build / capture
        double ms = 1e300;
        for (int r = 0; r < 2; r++) {
          run_batched(nsets, points, n, ndim, k, q, out_ids, out_dists, &st);
          ms = std::min(ms, st.ms_device);
        }
        if (ms_table) ms_table[c] = ms;
        c++;
        if (ms < best) {
          best = ms;
          bt = tq;
          bg = gr;
        }
      }
    std::vector<uint64_t> sig;
    for (uint32_t i = 0; i < nsets; i++) {
      sig.push_back(n[i]);
      sig.push_back(((uint64_t)ndim[i] << 32) | k[i]);
    }
    sig.push_back(p.num_indexed_dims);
    g_tunedSig = sig;
    g_tunedTq = bt;
    g_tunedGrid = bg;
    if (best_tq) *best_tq = bt;
    if (best_grid) *best_grid = bg;
    if (nconfigs) *nconfigs = c;
    return KNN_OK;
  });
}

}  // extern "C"

namespace knn {
static void release_batch() {
  try {
    g_batch.free_all();
  } catch (const Error &) {
    g_batch = BatchPlan{};
  }
}
}  // namespace knn

// ------------------------------------------------------------------ no-allocation scratch
// The buffers of knnjoin_ex_noalloc: a static device array (each device holds its own copy
// for the life of the process).  Host code uses only the device address that
// hipGetSymbolAddress returns for the current device, never the array's own name.
static __device__ __align__(256) unsigned char scratch[KNN_NOALLOC_SCRATCH_BYTES];

namespace knn {
constexpr int kMaxDevices = 64;
static void *g_scratchPtr[kMaxDevices] = {};

static void *noalloc_scratch(size_t &bytes) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(KNN_ERR_NO_DEVICE, "device index out of range");
  if (!g_scratchPtr[dev]) {
    void *ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(KNN_ERR_HIP, "no device address for the no-allocation scratch");
    g_scratchPtr[dev] = ptr;
  }
  bytes = KNN_NOALLOC_SCRATCH_BYTES;
  return g_scratchPtr[dev];
}
}  // namespace knn
