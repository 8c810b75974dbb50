// ref_shim.cpp — extern "C" entry points over the reference's own CPU DAWN.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// reference's CPU sources where they lie (src/*.cpp, src/algorithm/cpu/*.cpp)
// into oracle/_ref/libdawn_ref.so.  Used by tests to validate the C
// restatement in dawn_oracle.c and by bench.py as the "reference" CPU
// baseline.  Nothing here is part of the product.
#include <dawn/algorithm/cpu/bfs.hxx>
#include <dawn/algorithm/cpu/cc.hxx>
#include <dawn/algorithm/cpu/mssp.hxx>
#include <dawn/algorithm/cpu/sssp.hxx>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

namespace {

// Parse "src dst dist" lines written by DAWN::IO::outfile (src/io.cpp:215-250).
template <class T>
int parse_outfile(const std::string& path, int n, T* dist) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return -6;
  long s, d;
  double v;
  while (std::fscanf(f, "%ld %ld %lf", &s, &d, &v) == 3) {
    if (d >= 0 && d < n) dist[d] = (T)v;
  }
  std::fclose(f);
  return 0;
}

double wall_now() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

extern "C" {

// DAWN::BFS_CPU::BFS (dense, src/algorithm/cpu/bfs.cpp:23-77) with print on;
// distances recovered from its output file.  Reference encoding (0 =
// unvisited / source).  The caller must not pass an isolated source.
int ref_bfs(int n, int* row_ptr, int* col, int source, int sparse,
            int* dist, const char* tmp_path) {
  std::string out(tmp_path);
  std::memset(dist, 0, sizeof(int) * n);
  if (sparse)
    DAWN::BFS_CPU::BFS_kernel(row_ptr, col, n, source, true, out);
  else
    DAWN::BFS_CPU::BFS(row_ptr, col, n, source, true, out);
  return parse_outfile(out, n, dist);
}

// DAWN::SSSP_CPU::SSSP (dense, src/algorithm/cpu/sssp.cpp:23-80) with float
// weights; unreached vertices come back as FLT_MAX.  threads > 0 sets the
// OpenMP team size first: GOVMP's unsynchronised min (sssp.cpp:186-190)
// loses updates with more than one thread.
int ref_sssp(int n, int* row_ptr, int* col, float* val, int source,
             float* dist, const char* tmp_path, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  std::string out(tmp_path);
  for (int i = 0; i < n; i++) dist[i] = 0.0f;
  DAWN::SSSP_CPU::SSSP(row_ptr, col, val, n, source, true, out);
  return parse_outfile(out, n, dist);
}

// DAWN::CC_CPU::kernel / kernel_Weighted (src/algorithm/cpu/cc.cpp:37-142).
// threads > 0 sets the OpenMP team size first (the weighted kernel's GOVMP
// min is unsynchronised, so exact values need one thread).
float ref_closeness(int n, int* row_ptr, int* col, float* val, int source, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  float elapsed = 0.0f;
  if (val)
    return DAWN::CC_CPU::kernel_Weighted(row_ptr, col, val, n, source, elapsed);
  return DAWN::CC_CPU::kernel(row_ptr, col, n, source, elapsed);
}

// DAWN::MSSP_CPU::run (src/algorithm/cpu/mssp.cpp:9-32) on a borrowed CSR.
// Returns wall seconds measured around the call; *reported receives the
// value the reference itself returns (summed per-source ms / (thread*1000)).
double ref_mssp_time(int n, long long m, int* row_ptr, int* col, float* val,
                     const int* sources, int count, int threads,
                     float* reported) {
  DAWN::Graph::Graph_t g;
  g.rows = n;
  g.cols = n;
  g.nnz = (uint64_t)m;
  g.csr.row_ptr = row_ptr;
  g.csr.col = col;
  g.csr.val = val;
  g.thread = threads;
  g.interval = 100;
  g.stream = 1;
  g.block_size = 1024;
  g.source = 0;
  g.print = false;
  g.weighted = val != nullptr;
  g.directed = false;
  g.msource.assign(sources, sources + count);
  if (threads > 0) omp_set_num_threads(threads);
  std::string out("/dev/null");
  double t0 = wall_now();
  float r = DAWN::MSSP_CPU::run(g, out);
  double t1 = wall_now();
  if (reported) *reported = r;
  return t1 - t0;
}

}  // extern "C"
