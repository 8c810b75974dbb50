// Test helper: runs one of the drop-in host functions (MatrixMarket readers,
// COO helpers, source-less outfile) and dumps the result as text.  Compiled
// twice by tests/test_dropin_host.py: against include/dawn_amd.hxx +
// libdawn_amd.so, and (USE_REF) against the reference's own headers and
// sources compiled into oracle/_ref/libdawn_ref.so; the dumps must match.
#ifdef USE_REF
#include <dawn/graph.hxx>
#include <dawn/io.hxx>
#include <dawn/matrix.hxx>
#else
#include "dawn_amd.hxx"
#endif

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static void dump_csr(FILE* f, int n, uint64_t nnz, const DAWN::Matrix::Csr_t& csr, bool val) {
  std::fprintf(f, "n %d nnz %llu\n", n, (unsigned long long)nnz);
  for (int i = 0; i <= n; i++) std::fprintf(f, "%d ", csr.row_ptr[i]);
  std::fprintf(f, "\n");
  for (uint64_t k = 0; k < nnz; k++) std::fprintf(f, "%d ", csr.col[k]);
  std::fprintf(f, "\n");
  if (val)
    for (uint64_t k = 0; k < nnz; k++) std::fprintf(f, "%g ", (double)csr.val[k]);
  std::fprintf(f, "\n");
}

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const std::string mode = argv[1];
  std::string in = argv[2];
  std::string out = argv[3];
  FILE* f = std::fopen(out.c_str(), "w");
  if (!f) return 3;
  if (mode.rfind("readGraph", 0) == 0) {
    DAWN::Graph::Graph_t g{};
    // rows and the header's entry count, as createGraph sets them before
    // calling a reader in the reference (src/graph.cpp:58-62).
    g.rows = argc > 4 ? std::atoi(argv[4]) : 0;
    g.cols = g.rows;
    g.nnz = argc > 5 ? std::strtoull(argv[5], nullptr, 10) : 0;
    const bool w = mode.find("Weighted") != std::string::npos;
    if (mode == "readGraph") DAWN::IO::readGraph(in, g);
    else if (mode == "readGraph_Weighted") DAWN::IO::readGraph_Weighted(in, g);
    else if (mode == "readGraph_Directed") DAWN::IO::readGraph_Directed(in, g);
    else DAWN::IO::readGraph_Directed_Weighted(in, g);
    dump_csr(f, g.rows, g.nnz, g.csr, w);
  } else if (mode == "coo") {
    // in: "n nnz" then nnz lines "row col val"; sorted with transpose*,
    // then coo2Csr*.
    FILE* fi = std::fopen(in.c_str(), "r");
    int n = 0, nnz = 0;
    if (!fi || std::fscanf(fi, "%d %d", &n, &nnz) != 2) return 4;
    std::vector<int> r(nnz), c(nnz), r2(nnz), c2(nnz);
    std::vector<float> v(nnz), v2(nnz);
    for (int i = 0; i < nnz; i++)
      if (std::fscanf(fi, "%d %d %f", &r[i], &c[i], &v[i]) != 3) return 5;
    std::fclose(fi);
    r2 = r;
    c2 = c;
    v2 = v;
    DAWN::Matrix::Coo_t coo{};
    coo.row = r.data();
    coo.col = c.data();
    coo.val = v.data();
    DAWN::Matrix::transpose_Weighted(nnz, coo);
    DAWN::Matrix::Csr_t csr{};
    DAWN::Matrix::coo2Csr_Weighted(n, nnz, csr, coo);
    dump_csr(f, n, (uint64_t)nnz, csr, true);
    DAWN::Matrix::Coo_t coo2{};
    coo2.row = r2.data();
    coo2.col = c2.data();
    DAWN::Matrix::transpose(nnz, coo2);
    DAWN::Matrix::Csr_t csr2{};
    DAWN::Matrix::coo2Csr(n, nnz, csr2, coo2);
    dump_csr(f, n, (uint64_t)nnz, csr2, false);
  } else if (mode == "outfile") {
    // in: "n" then n ints; writes both source-less forms next to `out`.
    FILE* fi = std::fopen(in.c_str(), "r");
    int n = 0;
    if (!fi || std::fscanf(fi, "%d", &n) != 1) return 6;
    std::vector<int> d(n);
    std::vector<float> df(n);
    for (int i = 0; i < n; i++) {
      if (std::fscanf(fi, "%d", &d[i]) != 1) return 7;
      df[i] = d[i] * 0.5f;
    }
    std::fclose(fi);
    std::string oi = out + ".int", of = out + ".float";
    DAWN::IO::outfile(n, d.data(), oi);
    DAWN::IO::outfile(n, df.data(), of);
    std::fprintf(f, "ok\n");
  } else {
    return 8;
  }
  std::fclose(f);
  return 0;
}
