// cli_main.cpp -- `knnjoin_main <FNAME> <DIM> <K> [options]`, the command-line front end of
// the self-join (main.cu:109-510), as host C++ over the C ABI of libgpuknnselfjoin.so.
//
// Input (importNDDataset, import_dataset.cpp:50-86): the file is split at every ',' and each
// piece is read with `stream >> DTYPE` until the first failure, so numbers may be separated by
// commas or whitespace (newlines included); a piece that starts with text yields nothing (a
// header "x,y" followed by a newline also swallows the first number of the next line, as
// in the reference).  N = floor(values / DIM); trailing values are dropped.
//
// Output: the reference's console lines ("Time to join only", "[Verification] Total distance
// (without square root): ..."), one line appended to gpu_stats.txt with the reference's
// fields and tuple (main.cu:481-489), with --print-table
// the files KNN_out.txt and KNN_out_distances.txt (printNeighborTable, main.cu:911-949), and
// with --outlier-scores the file KNN_outlier_scores.txt (printOutlierScores,
// main.cu:952-1067; float32 only).
//
// Options (the reference sets these at compile time in params.h):
//   --indexed-dims N   NUMINDEXEDDIM (default min(DIM, 6))
//   --beta B           BETA (default 0.05)
//   --double           DTYPE double (default float)
//   --learned FILE     learned per-query radius: FILE holds 5057 little-endian float32
//                      weights followed by the float32 margin (written by learned_radius.py)
//   --hybrid-threads N host threads of the hybrid CPU/GPU split (splitWork, main.cu:2733-2933)
//   --hybrid-max-queries M  at most M host queries (sparsest cells first)
//   --print-table      write KNN_out.txt / KNN_out_distances.txt
//   --outlier-scores   write KNN_outlier_scores.txt
//   --parse-only       read the file and print N and a digest of the values (no GPU use)
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "gpuknnselfjoin.h"

namespace {

template <typename T>
bool import_values(const char *fname, std::vector<T> &out) {
  std::ifstream in(fname);
  if (!in) return false;
  std::string piece;
  while (std::getline(in, piece, ',')) {
    std::istringstream ss(piece);
    T v;
    while (ss >> v) out.push_back(v);
  }
  return true;
}

// FNV-1a over the value bits: lets a test compare the parse with a restatement.
template <typename T>
uint64_t digest(const std::vector<T> &v) {
  uint64_t h = 1469598103934665603ull;
  for (const T &x : v) {
    unsigned char b[sizeof(T)];
    std::memcpy(b, &x, sizeof(T));
    for (size_t i = 0; i < sizeof(T); i++) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  }
  return h;
}

template <typename T>
void write_table(unsigned n, unsigned k1, const int32_t *ids, const T *d) {
  const char *f1 = "KNN_out.txt";
  const char *f2 = "KNN_out_distances.txt";
  std::printf("\n\nOutputting neighbors to: %s\n", f1);
  std::ofstream a(f1, std::ios::out);
  a << "#data point (line is the point id), neighbor point ids\n";
  for (size_t i = 0; i < n; i++) {
    for (unsigned j = 0; j < k1; j++) a << ids[i * k1 + j] << ", ";
    a << "\n";
  }
  std::printf("\nOutputting distances to: %s\n", f2);
  std::ofstream b(f2, std::ios::out);
  b << "#data point (line is the point id), neighbor point distances\n";
  for (size_t i = 0; i < n; i++) {
    for (unsigned j = 0; j < k1; j++) b << d[i * k1 + j] << ", ";
    b << "\n";
  }
}

// main.cu:1050-1058: per point, mean distance (row sum / (K+1 - 1)), its rank, in-degree, its rank
void write_outlier_scores(unsigned n, unsigned k1, const int32_t *ids, const float *d) {
  std::vector<float> sum(n);
  std::vector<uint32_t> dr(n), deg(n), gr(n);
  if (knnjoin_outlier_scores(ids, d, n, k1, 0, sum.data(), dr.data(), deg.data(), gr.data()) != KNN_OK) {
    std::fprintf(stderr, "outlier scores: %s\n", knnjoin_last_error());
    return;
  }
  const char *fname = "KNN_outlier_scores.txt";
  std::printf("\nOutputting outlier scores to: %s\n", fname);
  std::ofstream o(fname, std::ios::out);
  o << "#data point (line is the point id), col0: mean distance between point and its k neighbors,";
  o << "col1: outlier ranking for col0, col2: the number of times the point appears in another set (in-degree), col3: the ranking for col2 \n";
  for (size_t i = 0; i < n; i++)
    o << sum[i] / ((k1 * 1.0) - 1.0) << ", " << dr[i] << ", " << deg[i] << ", " << gr[i] << std::endl;
}

bool load_model(const char *path, uint32_t nd) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    std::fprintf(stderr, "cannot open model file %s\n", path);
    return false;
  }
  const uint32_t nw = 5057;
  std::vector<float> w(nw + 1);
  f.read(reinterpret_cast<char *>(w.data()), sizeof(float) * (nw + 1));
  if (f.gcount() != (std::streamsize)(sizeof(float) * (nw + 1))) {
    std::fprintf(stderr, "model file %s: expected %u floats\n", path, nw + 1);
    return false;
  }
  if (knnjoin_set_radius_model(nd, w.data(), nw, w[nw]) != KNN_OK) {
    std::fprintf(stderr, "model: %s\n", knnjoin_last_error());
    return false;
  }
  return true;
}

template <typename T>
int run(const char *fname, unsigned ndim, unsigned k, knnjoin_params &p, bool printTable,
        bool outlierScores, bool parseOnly, const char *dtypeName) {
  std::vector<T> vals;
  if (!import_values<T>(fname, vals)) {
    std::fprintf(stderr, "cannot open %s\n", fname);
    return 1;
  }
  const size_t n = vals.size() / ndim;
  vals.resize(n * ndim);
  if (parseOnly) {
    std::printf("points %zu values %zu digest %016" PRIx64 "\n", n, vals.size(), digest(vals));
    return 0;
  }
  std::printf("\nNumber of points: %zu, dims %u, K %u, dtype %s\n", n, ndim, k, dtypeName);
  if (n > 0xFFFFFFFFull) {
    std::fprintf(stderr, "more than 2^32-1 points\n");
    return 1;
  }
  const unsigned k1 = k + 1;
  std::vector<int32_t> ids((size_t)n * k1);
  std::vector<T> d((size_t)n * k1);
  knnjoin_stats st{};
  p.verbose = 1;
  int rc = knnjoin_ex(vals.data(), sizeof(T) == 4 ? KNN_FLOAT32 : KNN_FLOAT64, (uint32_t)n, ndim, k,
                      &p, ids.data(), d.data(), &st);
  if (rc != KNN_OK) {
    std::fprintf(stderr, "\nERROR (%d): %s\n", rc, knnjoin_last_error());
    return 1;
  }
  // main.cu:420-445: the join (every pass and re-index) and totalTime, which leaves out the
  // first index build
  std::printf("\nTime to join only: %f", (st.ms_join + st.ms_index - st.ms_index_first) / 1e3);
  std::printf("\nTotal time (reference definition, first index build excluded): %f",
              st.ms_reference_total / 1e3);
  std::printf("\nDevice pipeline (first index build included): %f", st.ms_total / 1e3);
  std::printf("\nSchedule: %s, iterations %u, re-indexes %u",
              p.mode == KNN_MODE_LEARNED ? "learned" : "beta", st.iterations, st.reindexes);
  if (p.hybrid_threads)
    std::printf("\nHybrid split: %u queries on %u host threads (%.3f s, concurrent with the GPU)",
                st.hybrid_queries, p.hybrid_threads, st.ms_hybrid_host / 1e3);
  std::printf("\n[Verification] Total distance (without square root): %f\n", st.checksum);
  if (printTable) write_table<T>((unsigned)n, k1, ids.data(), d.data());
  if (outlierScores) {
    if (sizeof(T) == 4)
      write_outlier_scores((unsigned)n, k1, ids.data(), reinterpret_cast<const float *>(d.data()));
    else
      std::fprintf(stderr, "--outlier-scores needs float32 data\n");
  }
  // main.cu:481-489 (THREADMULTI -2, params.h:38): totalTime, file, K, eps bucket, total
  // distance, then the build tuple.  SHORTCIRCUIT, THREADMULTI, MAXTHREADSPERPOINT and
  // STATICTHREADSPERPOINT configure the reference's CUDA kernel and have no counterpart here;
  // they are written with the reference's defaults (params.h:32-51) so that the line parses
  // as the reference's does.
  std::ofstream gs("gpu_stats.txt", std::ios::app);
  gs << st.ms_reference_total / 1e3 << ", " << fname << ", KNN: " << k << ", Eps bucket: " << st.eps_bucket
     << ", Total dist: " << std::setprecision(9) << st.checksum << ", "
     << "GPUNUMDIM/NUMINDEXEDDIM/REORDER/SHORTCIRCUIT/THREADMULTI/MAXTHREADSPERPOINT/DYNAMICTHRESHOLD/"
        "STATICTHREADSPERPOINT/BETA/DTYPE(float/double): "
     << ndim << ", " << (p.num_indexed_dims ? p.num_indexed_dims : (ndim < 6 ? ndim : 6)) << ", "
     << p.reorder << ", 1, -2, 1024, N/A, 8, " << p.beta << ", " << dtypeName << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<const char *> pos;
  knnjoin_params p;
  knnjoin_default_params(&p);
  bool dbl = false, printTable = false, parseOnly = false, outlierScores = false;
  const char *model = nullptr;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto need = [&](const char *what) -> const char * {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "%s needs a value\n", what);
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--indexed-dims") p.num_indexed_dims = (uint32_t)std::atoi(need("--indexed-dims"));
    else if (a == "--beta") p.beta = std::atof(need("--beta"));
    else if (a == "--double") dbl = true;
    else if (a == "--learned") model = need("--learned");
    else if (a == "--hybrid-threads") p.hybrid_threads = (uint32_t)std::atoi(need("--hybrid-threads"));
    else if (a == "--hybrid-max-queries") p.hybrid_max_queries = (uint32_t)std::atoi(need("--hybrid-max-queries"));
    else if (a == "--print-table") printTable = true;
    else if (a == "--outlier-scores") outlierScores = true;
    else if (a == "--parse-only") parseOnly = true;
    else pos.push_back(argv[i]);
  }
  if (pos.size() != 3) {
    // main.cu:121-125
    std::cout << "\n\nIncorrect number of input parameters.  \nShould be dataset file, number of dimensions, kNN\n";
    return 0;
  }
  const unsigned ndim = (unsigned)std::atoi(pos[1]);
  const unsigned k = (unsigned)std::atoi(pos[2]);
  if (ndim < 1 || ndim > 16) {
    std::fprintf(stderr, "DIM must be in 1..16\n");
    return 2;
  }
  if (model && !parseOnly) {
    const uint32_t nd = p.num_indexed_dims ? p.num_indexed_dims : (ndim < 6 ? ndim : 6);
    if (!load_model(model, nd)) return 1;
    p.mode = KNN_MODE_LEARNED;
  }
  return dbl ? run<double>(pos[0], ndim, k, p, printTable, outlierScores, parseOnly, "double")
             : run<float>(pos[0], ndim, k, p, printTable, outlierScores, parseOnly, "float");
}
