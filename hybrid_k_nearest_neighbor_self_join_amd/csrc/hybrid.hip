// hybrid.hip -- the host share of a hybrid CPU/GPU join.
//
// splitWork (main.cu:2733-2933, with the cell-population criterion of main.cu:2740-2770):
// the queries whose own cell of the first grid (width eps0) holds at most
//   ceil((K + 1) * 2^nd / V_nd)        (V_nd the volume of the unit nd-ball; gamma = 0)
// points are joined on the host while the GPU runs its failed-query loop over the others.
// The host threads search a copy of the GPU's first index (cell-sorted coordinates, point
// ids, non-empty cell table) cell ring by cell ring around the query and keep the K+1 best
// (FP64 key, id) pairs; a query is done once its (K+1)-th key lies below the squared
// distance to the nearest face of the searched block, or the block covers the grid.  Rows
// are the device path's rows: same key, same tie rule, same FP32 distance arithmetic (this
// file is built with -ffp-contract=off like the kernels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "knn_launch.h"

namespace knn {

namespace {

template <typename T>
inline uint32_t host_cell_coord(T x, T mn, T eps, uint32_t ncell) {
  T t = (x - mn) / eps;  // kernel.cu:1685-1687, as cell_coord in knn_device.h
  if (!(t > (T)0)) return 0u;
  if (t >= (T)4294967295.0) return ncell - 1;
  uint32_t c = (uint32_t)t;
  return c >= ncell ? ncell - 1 : c;
}

inline bool hless(double ka, uint32_t ia, double kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// bounded max-heap of (key, id) pairs (with each point's sorted position), the largest on top
struct TopK {
  std::vector<double> key;
  std::vector<uint32_t> id, pos;
  uint32_t size = 0, cap = 0;
  void reset(uint32_t k1) {
    cap = k1;
    size = 0;
    key.resize(k1);
    id.resize(k1);
    pos.resize(k1);
  }
  void swap_entries(uint32_t a, uint32_t b) {
    std::swap(key[a], key[b]);
    std::swap(id[a], id[b]);
    std::swap(pos[a], pos[b]);
  }
  void sift_down(uint32_t i, uint32_t n) {
    for (;;) {
      uint32_t l = 2 * i + 1, r = l + 1, m = i;
      if (l < n && hless(key[m], id[m], key[l], id[l])) m = l;
      if (r < n && hless(key[m], id[m], key[r], id[r])) m = r;
      if (m == i) return;
      swap_entries(i, m);
      i = m;
    }
  }
  void offer(double k, uint32_t i, uint32_t p) {
    if (size < cap) {
      uint32_t c = size++;
      key[c] = k;
      id[c] = i;
      pos[c] = p;
      while (c > 0) {
        uint32_t up = (c - 1) / 2;
        if (!hless(key[up], id[up], key[c], id[c])) break;
        swap_entries(up, c);
        c = up;
      }
    } else if (hless(k, i, key[0], id[0])) {
      key[0] = k;
      id[0] = i;
      pos[0] = p;
      sift_down(0, size);
    }
  }
  void sort() {
    for (uint32_t end = size; end > 1; end--) {
      swap_entries(0, end - 1);
      sift_down(0, end - 1);
    }
  }
};

template <typename T, typename G>
struct HostSearch {
  const HybridIndex<T> &ix;
  const G &g;
  uint32_t k1;
  TopK top;
  std::vector<T> qc;
  std::vector<double> qd;

  HostSearch(const HybridIndex<T> &index, const G &grid, uint32_t k1_)
      : ix(index), g(grid), k1(k1_), qc(index.d), qd(index.d) {}

  T coord(uint32_t j, uint32_t p) const { return ix.xs[(size_t)j * ix.n + p]; }

  // sorted positions of the cells [l0, h0] (dim 0) of the row rowBase
  void scan_range(uint64_t rowBase, uint32_t l0, uint32_t h0) {
    const uint64_t *ck = ix.cellKeys;
    const uint64_t *cb = std::lower_bound(ck, ck + ix.ncell, rowBase + l0);
    const uint64_t *ce = std::upper_bound(cb, ck + ix.ncell, rowBase + h0);
    const uint32_t pb = ix.cellStart[cb - ck], pe = ix.cellStart[ce - ck];
    const uint32_t d = ix.d;
    const size_t n = ix.n;
    for (uint32_t p = pb; p < pe; p++) {
      double key = 0.0;
      for (uint32_t j = 0; j < d; j++) {
        const double dx = (double)ix.xs[(size_t)j * n + p] - qd[j];
        const double t = dx * dx;
        key = key + t;
      }
      top.offer(key, ix.sortedIds[p], p);
    }
  }

  // the exact K+1 nearest of the point at sorted position qp (K+1 output entries)
  void query(uint32_t qp, int32_t *ids, T *dists) {
    const uint32_t d = ix.d, nd = g.nd;
    for (uint32_t j = 0; j < d; j++) {
      qc[j] = coord(j, qp);
      qd[j] = (double)qc[j];
    }
    uint32_t c[KNN_MAX_INDEXED], lo[KNN_MAX_INDEXED], hi[KNN_MAX_INDEXED], cur[KNN_MAX_INDEXED];
    for (uint32_t j = 0; j < nd; j++) c[j] = host_cell_coord<T>(qc[j], (T)g.minArr[j], (T)g.eps, g.nCells[j]);
    top.reset(k1);
    for (uint32_t t = 0;; t++) {
      bool whole = true;
      for (uint32_t j = 0; j < nd; j++) {
        lo[j] = c[j] >= t ? c[j] - t : 0u;
        hi[j] = (uint64_t)c[j] + t < g.nCells[j] ? c[j] + t : g.nCells[j] - 1;
        whole = whole && lo[j] == 0 && hi[j] == g.nCells[j] - 1;
      }
      // the shell of ring t: rows of dims 1.. on the ring's faces in full, the others at
      // their two dim-0 ends
      for (uint32_t j = 1; j < nd; j++) cur[j] = lo[j];
      for (;;) {
        uint64_t rowBase = g.keyBase;
        bool face = t == 0;
        for (uint32_t j = 1; j < nd; j++) {
          rowBase += (uint64_t)cur[j] * g.mult[j];
          face = face || cur[j] + t == c[j] || cur[j] == c[j] + t;
        }
        if (face) {
          scan_range(rowBase, lo[0], hi[0]);
        } else {
          if (c[0] >= t) scan_range(rowBase, c[0] - t, c[0] - t);
          if ((uint64_t)c[0] + t < g.nCells[0]) scan_range(rowBase, c[0] + t, c[0] + t);
        }
        uint32_t j = 1;
        while (j < nd) {
          if (cur[j] < hi[j]) {
            cur[j]++;
            break;
          }
          cur[j] = lo[j];
          j++;
        }
        if (j >= nd) break;
      }
      if (whole) break;
      if (top.size == k1) {
        // every point outside the block is at least `gap` away in some indexed dim (less
        // a margin for the rounding of its cell coordinate)
        double gap = __builtin_inf();
        for (uint32_t j = 0; j < nd; j++) {
          const double e = (double)g.eps, mn = (double)g.minArr[j];
          if (lo[j] > 0) {
            const double edge = mn + (double)lo[j] * e;
            gap = std::min(gap, qd[j] - edge - (std::fabs(qd[j]) + std::fabs(edge) + e) * 0x1p-20);
          }
          if (hi[j] < g.nCells[j] - 1) {
            const double edge = mn + (double)(hi[j] + 1) * e;
            gap = std::min(gap, edge - qd[j] - (std::fabs(qd[j]) + std::fabs(edge) + e) * 0x1p-20);
          }
        }
        if (gap > 0.0 && top.key[0] < gap * gap) break;
      }
    }
    top.sort();
    for (uint32_t e = 0; e < k1; e++) {
      ids[e] = e < top.size ? (int32_t)top.id[e] : -1;
      T s = (T)0;
      if (e < top.size)
        for (uint32_t j = 0; j < d; j++) {
          const T df = coord(j, top.pos[e]) - qc[j];
          const T sq = df * df;
          s = s + sq;
        }
      dists[e] = std::sqrt(s);
    }
  }
};

// rows of the host share into the table: row of query list[i] <- packed row i
template <typename T>
__global__ void k_scatter_rows(const int32_t *rowIds, const T *rowDists, const uint32_t *list, uint32_t nq,
                               uint32_t k1, int32_t *ids, T *dists) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)nq * k1) return;
  const uint32_t r = (uint32_t)(i / k1), e = (uint32_t)(i % k1);
  const uint64_t dst = (uint64_t)list[r] * k1 + e;
  ids[dst] = rowIds[i];
  dists[dst] = rowDists[i];
}

// per sorted position p: cpu[p] = 1 when its own cell holds at most `thr` points; gpu[q] for
// its point q is the complement
__global__ void k_sparse_cell_flags(const uint32_t *cellStart, const uint32_t *cellOf, const uint32_t *sortedIds,
                                    uint32_t n, uint32_t thr, uint8_t *cpu, uint8_t *gpu) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t c = cellOf[p];
  const bool sparse = cellStart[c + 1] - cellStart[c] <= thr;
  cpu[p] = sparse ? 1 : 0;
  gpu[sortedIds[p]] = sparse ? 0 : 1;
}

}  // namespace

// splitWork's point threshold (main.cu:2745-2770): ceil((K + 1) * 2^nd / V_nd), gamma = 0
uint32_t hybrid_point_threshold(uint32_t k1, uint32_t nd) {
  const double cube = std::pow(2.0, (double)nd);
  const double pi = 3.14159265358979323846;
  const double ball = std::pow(pi, nd * 0.5) / std::tgamma(nd * 0.5 + 1.0);
  return (uint32_t)std::ceil((double)k1 * (cube / ball));
}

void hybrid_flags(const uint32_t *cellStart, const uint32_t *cellOf, const uint32_t *sortedIds, uint32_t n,
                  uint32_t thr, uint8_t *cpu, uint8_t *gpu, hipStream_t s) {
  if (n) k_sparse_cell_flags<<<(n + 255) / 256, 256, 0, s>>>(cellStart, cellOf, sortedIds, n, thr, cpu, gpu);
}

template <typename T>
void hybrid_scatter(const int32_t *rowIds, const T *rowDists, const uint32_t *list, uint32_t nq, uint32_t k1,
                    int32_t *ids, T *dists, hipStream_t s) {
  const uint64_t tot = (uint64_t)nq * k1;
  if (tot) k_scatter_rows<T><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(rowIds, rowDists, list, nq, k1, ids, dists);
}

template <typename T>
void hybrid_host_join(const HybridIndex<T> &ix, const typename GridOf<T>::type &g, uint32_t k1,
                      const uint32_t *queries, uint32_t nq, uint32_t threads, int32_t *outIds, T *outDists) {
  using G = typename GridOf<T>::type;
  std::atomic<uint32_t> next{0};
  const uint32_t chunk = 64;
  auto work = [&]() {
    HostSearch<T, G> hs(ix, g, k1);
    for (;;) {
      const uint32_t b = next.fetch_add(chunk);
      if (b >= nq) break;
      const uint32_t e = std::min(nq, b + chunk);
      for (uint32_t i = b; i < e; i++) hs.query(queries[i], outIds + (size_t)i * k1, outDists + (size_t)i * k1);
    }
  };
  std::vector<std::thread> pool;
  const uint32_t nt = std::max(1u, threads);
  for (uint32_t t = 1; t < nt; t++) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
}

template void hybrid_scatter<float>(const int32_t *, const float *, const uint32_t *, uint32_t, uint32_t,
                                    int32_t *, float *, hipStream_t);
template void hybrid_scatter<double>(const int32_t *, const double *, const uint32_t *, uint32_t, uint32_t,
                                     int32_t *, double *, hipStream_t);
template void hybrid_host_join<float>(const HybridIndex<float> &, const GridParams &, uint32_t, const uint32_t *,
                                      uint32_t, uint32_t, int32_t *, float *);
template void hybrid_host_join<double>(const HybridIndex<double> &, const GridParamsD &, uint32_t,
                                       const uint32_t *, uint32_t, uint32_t, int32_t *, double *);

}  // namespace knn
