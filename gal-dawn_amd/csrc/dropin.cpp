// C++ drop-in host API (include/dawn_amd.hxx) implemented on the C ABI.
// Each entry point uploads the caller's CSR, runs the HIP engine through
// dawn_capi.h and reproduces the reference's return values, messages and
// output files (see the header for the reference lines).
#include "../../include/dawn_amd.hxx"

#include <getopt.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/dawn_capi.h"

namespace {

void die(int rc, const char* what) {
  std::fprintf(stderr, "%s: %s (%d) %s\n", what, dawn_status_string(rc), rc, dawn_last_error());
  std::exit(EXIT_FAILURE);
}

// Uploads graph.csr; weights (float in Graph_t, as in the reference) must be
// integral and non-negative for the int32 min-plus engine.
dawn_graph* upload(DAWN::Graph::Graph_t& graph, bool weighted) {
  dawn_csr c{};
  c.n = graph.rows;
  c.m = (int64_t)graph.nnz;
  c.directed = graph.directed ? 1 : 0;
  c.row_ptr = graph.csr.row_ptr;
  c.col = graph.csr.col;
  std::vector<int32_t> w;
  if (weighted) {
    if (!graph.csr.val) die(DAWN_ERR_NOT_WEIGHTED, "upload");
    w.resize((size_t)c.m);
    for (int64_t k = 0; k < c.m; k++) {
      const float x = graph.csr.val[k];
      if (!(x >= 0.f) || std::floor(x) != x || x > 2147483647.f) die(DAWN_ERR_FORMAT, "weights");
      w[k] = (int32_t)x;
    }
    c.weight = w.data();
    c.weighted = 1;
  }
  dawn_graph* g = nullptr;
  const int rc = dawn_graph_create(&c, &g);
  if (rc != DAWN_OK) die(rc, "dawn_graph_create");
  return g;
}

bool isolated(const DAWN::Graph::Graph_t& graph, int s) {
  return graph.csr.row_ptr[s] == graph.csr.row_ptr[s + 1];
}

int norm_source(int64_t s, int n) {
  int64_t r = s % n;
  return (int)(r < 0 ? r + n : r);
}

// Weighted output as the reference prints it (src/io.cpp:233-250): float
// values with the stream's default formatting, FLT_MAX for unreached.
void outfile_weighted(int n, const int32_t* dist, int source, const std::string& path) {
  std::vector<float> f(n);
  for (int j = 0; j < n; j++) f[j] = dist[j] == DAWN_INF ? FLT_MAX : (float)dist[j];
  std::string p(path);
  DAWN::IO::outfile(n, f.data(), source, p);
}

float multi(DAWN::Graph::Graph_t& graph, std::string& output_path, const std::vector<int>& src,
            bool weighted) {
  dawn_graph* g = upload(graph, weighted);
  std::vector<int32_t> run;
  for (int s : src) {
    if (isolated(graph, s)) {
      std::printf("Source [%d] is isolated node\n", s);  // mssp.cu:28-34
      continue;
    }
    run.push_back(s);
  }
  dawn_stats st{};
  std::vector<int32_t> row;
  int print_at = -1;
  if (graph.print)
    for (size_t i = 0; i < run.size(); i++)
      if (run[i] == graph.source) print_at = (int)i;
  if (!run.empty()) {
    const int rc = dawn_mssp(g, run.data(), (int32_t)run.size(), nullptr, nullptr,
                             DAWN_POLICY_LEARNED, &st);
    if (rc != DAWN_OK) die(rc, "dawn_mssp");
  }
  if (print_at >= 0) {
    row.resize(graph.rows);
    int32_t s = run[print_at];
    const int rc = dawn_mssp(g, &s, 1, row.data(), nullptr, DAWN_POLICY_LEARNED, nullptr);
    if (rc != DAWN_OK) die(rc, "dawn_mssp");
    std::printf("Start print\n");
    if (weighted)
      outfile_weighted(graph.rows, row.data(), s, output_path);
    else
      DAWN::IO::outfile(graph.rows, row.data(), s, output_path);
  }
  dawn_graph_destroy(g);
  return (float)(st.device_ms / 1000.0);
}

}  // namespace

namespace DAWN {

namespace {

// Reads path with the reader's semantics (as_directed: -1 from the header,
// 0 both directions, 1 as listed) into graph.csr (plain new[] arrays owned
// by Graph_t, as in the reference, src/matrix.cpp:32-34).  Returns false
// after printing the reference's message when the file cannot be used.
bool read_into(std::string& input_path, DAWN::Graph::Graph_t& graph, bool weighted,
               int as_directed) {
  dawn_csr c{};
  const int rc = dawn_read_mtx_as(input_path.c_str(), weighted ? 1 : 0, as_directed, &c);
  if (rc == DAWN_ERR_IO) {
    std::cerr << "Error opening file " << input_path << std::endl;  // graph.cpp:11-14
    return false;
  }
  if (rc == DAWN_ERR_FORMAT) {
    std::cout << "invalid file" << std::endl;  // graph.cpp:45-48
    return false;
  }
  if (rc != DAWN_OK) die(rc, "dawn_read_mtx_as");
  graph.rows = c.n;
  graph.cols = c.n;
  graph.nnz = (uint64_t)c.m;
  graph.directed = c.directed != 0;
  graph.source = c.n ? norm_source(graph.source, c.n) : 0;
  graph.csr.row_ptr = new int[(size_t)c.n + 1];
  graph.csr.col = new int[(size_t)c.m > 0 ? (size_t)c.m : 1];
  std::memcpy(graph.csr.row_ptr, c.row_ptr, sizeof(int) * ((size_t)c.n + 1));
  if (c.m) std::memcpy(graph.csr.col, c.col, sizeof(int) * (size_t)c.m);
  if (c.weight) {
    graph.csr.val = new float[(size_t)c.m > 0 ? (size_t)c.m : 1];
    for (int64_t k = 0; k < c.m; k++) graph.csr.val[k] = (float)c.weight[k];
  }
  dawn_csr_free(&c);
  return true;
}

}  // namespace

void Graph::createGraph(std::string& input_path, Graph_t& graph) {
  if (read_into(input_path, graph, graph.weighted, -1))
    std::cout << "Read Input Graph" << std::endl;
}

// The four readers with their own edge semantics, whatever the header says
// (src/io.cpp:2-192): undirected ones push both directions of each pair.
void IO::readGraph(std::string& input_path, Graph::Graph_t& graph) {
  read_into(input_path, graph, false, 0);
}

void IO::readGraph_Weighted(std::string& input_path, Graph::Graph_t& graph) {
  read_into(input_path, graph, true, 0);
}

void IO::readGraph_Directed(std::string& input_path, Graph::Graph_t& graph) {
  read_into(input_path, graph, false, 1);
}

void IO::readGraph_Directed_Weighted(std::string& input_path, Graph::Graph_t& graph) {
  read_into(input_path, graph, true, 1);
}

// Row-sorted COO -> CSR on the host, as the reference's helpers
// (src/matrix.cpp:28-82): rows counted, offsets by a prefix sum, columns
// (and values) copied in COO order, so the COO must already be sorted by
// row.  The CSR arrays are new[] arrays owned by the caller.
void Matrix::coo2Csr(int n, int nnz, Csr_t& csr, Coo_t& coo) {
  csr.row_ptr = new int[(size_t)n + 1]();
  csr.col = new int[nnz > 0 ? (size_t)nnz : 1];
  for (int i = 0; i < nnz; i++) csr.row_ptr[coo.row[i] + 1]++;
  for (int i = 0; i < n; i++) csr.row_ptr[i + 1] += csr.row_ptr[i];
  if (nnz > 0) std::memcpy(csr.col, coo.col, sizeof(int) * (size_t)nnz);
}

void Matrix::coo2Csr_Weighted(int n, int nnz, Csr_t& csr, Coo_t& coo) {
  coo2Csr(n, nnz, csr, coo);
  csr.val = new float[nnz > 0 ? (size_t)nnz : 1];
  if (nnz > 0) std::memcpy(csr.val, coo.val, sizeof(float) * (size_t)nnz);
}

// Despite the name, the reference's transpose sorts the COO entries by
// (row, column[, value]) in place (src/matrix.cpp:3-26); the readers of
// directed graphs use it to order their entries.
void Matrix::transpose(int nnz, Coo_t& coo) {
  std::vector<std::pair<int, int>> e((size_t)std::max(nnz, 0));
  for (int i = 0; i < nnz; i++) e[i] = {coo.row[i], coo.col[i]};
  std::sort(e.begin(), e.end());
  for (int i = 0; i < nnz; i++) {
    coo.row[i] = e[i].first;
    coo.col[i] = e[i].second;
  }
}

void Matrix::transpose_Weighted(int nnz, Coo_t& coo) {
  std::vector<std::tuple<int, int, float>> e((size_t)std::max(nnz, 0));
  for (int i = 0; i < nnz; i++) e[i] = std::make_tuple(coo.row[i], coo.col[i], coo.val[i]);
  std::sort(e.begin(), e.end());
  for (int i = 0; i < nnz; i++) {
    coo.row[i] = std::get<0>(e[i]);
    coo.col[i] = std::get<1>(e[i]);
    coo.val[i] = std::get<2>(e[i]);
  }
}

void IO::readList(std::string& input_path, Graph::Graph_t& graph) {
  int32_t* v = nullptr;
  int64_t n = 0;
  const int rc = dawn_read_list(input_path.c_str(), &v, &n);
  if (rc != DAWN_OK) {
    std::cerr << "Error opening file " << input_path << std::endl;
    return;
  }
  for (int64_t i = 0; i < n; i++) graph.msource.push_back(v[i]);
  dawn_free(v);
}

void IO::outfile(int n, int* result, int source, std::string& output_path) {
  const int rc = dawn_write_dist(output_path.c_str(), n, result, source);
  if (rc != DAWN_OK) std::cerr << "Error opening file " << output_path << std::endl;
}

void IO::outfile(int n, float* result, int source, std::string& output_path) {
  FILE* f = std::fopen(output_path.c_str(), "w");
  if (!f) {
    std::cerr << "Error opening file " << output_path << std::endl;
    return;
  }
  for (int j = 0; j < n; j++)
    if (source != j && result[j] != 0) std::fprintf(f, "%d %d %g\n", source, j, (double)result[j]);
  std::fclose(f);
}

// Source-less forms (src/io.cpp:252-281): "j dist" lines, int values > 0,
// float values != 0 with six decimals.
void IO::outfile(int n, int* result, std::string& output_path) {
  FILE* f = std::fopen(output_path.c_str(), "w");
  if (!f) {
    std::cerr << "Error opening file " << output_path << std::endl;
    return;
  }
  for (int j = 0; j < n; j++)
    if (result[j] > 0) std::fprintf(f, "%d %d\n", j, result[j]);
  std::fclose(f);
}

void IO::outfile(int n, float* result, std::string& output_path) {
  FILE* f = std::fopen(output_path.c_str(), "w");
  if (!f) {
    std::cerr << "Error opening file " << output_path << std::endl;
    return;
  }
  for (int j = 0; j < n; j++)
    if (result[j] != 0) std::fprintf(f, "%d %.6f\n", j, (double)result[j]);
  std::fclose(f);
}

void IO::printHelp() {
  std::printf(
      "Usage: program [options]\n"
      "Options:\n"
      "  -h, --help                    Show this help message and exit\n"
      "  -i, --input PATH              Input file path\n"
      "  -o, --output PATH             Output file path\n"
      "  -l, --sourceList PATH         List of Source nodes file path\n"
      "  -p, --print [true|false]      Print source information\n"
      "  -w, --weighted [true|false]   Set graph as weighted\n"
      "  -s, --source NUM              Source node (optional, defaults to a random value)\n");
}

IO::parameters_t IO::parameters(int argc, char* argv[]) {
  parameters_t p;
  static const struct option opts[] = {{"help", no_argument, nullptr, 'h'},
                                       {"input", required_argument, nullptr, 'i'},
                                       {"output", required_argument, nullptr, 'o'},
                                       {"sourceList", required_argument, nullptr, 'l'},
                                       {"print", required_argument, nullptr, 'p'},
                                       {"weighted", required_argument, nullptr, 'w'},
                                       {"source", required_argument, nullptr, 's'},
                                       {nullptr, 0, nullptr, 0}};
  auto boolean = [](const char* v, const char* name) {
    if (!std::strcmp(v, "true")) return true;
    if (!std::strcmp(v, "false")) return false;
    std::cerr << "Invalid value for " << name << " option. Use 'true' or 'false'.\n";
    std::exit(EXIT_FAILURE);
  };
  int c, idx = 0;
  optind = 1;
  while ((c = getopt_long(argc, argv, "hi:o:l:p:w:s:", opts, &idx)) != -1) {
    switch (c) {
      case 'h': printHelp(); std::exit(EXIT_SUCCESS);
      case 'i': p.input_path = optarg; break;
      case 'o': p.output_path = optarg; break;
      case 'l': p.sourceList_path = optarg; break;
      case 'p': p.print = boolean(optarg, "print"); break;
      case 'w': p.weighted = boolean(optarg, "weighted"); break;
      case 's': p.source = std::atoi(optarg); break;
      default: printHelp(); std::exit(EXIT_FAILURE);
    }
  }
  if (p.input_path.empty()) {
    std::cerr << "Error: Missing the input file path.\n";
    printHelp();
    std::exit(EXIT_FAILURE);
  }
  if (p.source == -1) p.source = std::rand();  // unseeded, as the reference
  return p;
}

float BFS_GPU::run(Graph::Graph_t& graph, std::string& output_path) {
  const int source = graph.source;
  if (isolated(graph, source)) {
    std::printf("Source [%d] is isolated node\n", source);  // bfs.cu:110-113
    std::exit(0);
  }
  dawn_graph* g = upload(graph, false);
  std::vector<int32_t> dist(graph.rows);
  dawn_stats st{};
  const int rc = dawn_bfs(g, source, dist.data(), DAWN_POLICY_LEARNED, &st);
  if (rc != DAWN_OK) die(rc, "dawn_bfs");
  if (graph.print) {
    std::printf("Start print\n");
    IO::outfile(graph.rows, dist.data(), source, output_path);
  }
  dawn_graph_destroy(g);
  return (float)(st.device_ms / 1000.0);
}

float SSSP_GPU::run(Graph::Graph_t& graph, std::string& output_path) {
  const int source = graph.source;
  if (isolated(graph, source)) {
    std::printf("Source [%d] is isolated node\n", source);  // sssp.cu:134-137
    std::exit(0);
  }
  dawn_graph* g = upload(graph, true);
  std::vector<int32_t> dist(graph.rows);
  dawn_stats st{};
  const int rc = dawn_sssp(g, source, dist.data(), DAWN_POLICY_FIXED, &st);
  if (rc != DAWN_OK) die(rc, "dawn_sssp");
  if (graph.print) {
    std::printf("Start print\n");
    outfile_weighted(graph.rows, dist.data(), source, output_path);
  }
  dawn_graph_destroy(g);
  return (float)(st.device_ms / 1000.0);
}

float MSSP_GPU::run(Graph::Graph_t& graph, std::string& output_path) {
  std::vector<int> src;
  for (int s : graph.msource) src.push_back(norm_source(s, graph.rows));
  return multi(graph, output_path, src, false);
}

float MSSP_GPU::run_Weighted(Graph::Graph_t& graph, std::string& output_path) {
  std::vector<int> src;
  for (int s : graph.msource) src.push_back(norm_source(s, graph.rows));
  return multi(graph, output_path, src, true);
}

float APSP_GPU::run(Graph::Graph_t& graph, std::string& output_path) {
  std::vector<int> src(graph.rows);
  for (int i = 0; i < graph.rows; i++) src[i] = i;
  return multi(graph, output_path, src, false);
}

float APSP_GPU::run_Weighted(Graph::Graph_t& graph, std::string& output_path) {
  std::vector<int> src(graph.rows);
  for (int i = 0; i < graph.rows; i++) src[i] = i;
  return multi(graph, output_path, src, true);
}

static void closeness(Graph::Graph_t& graph, int source, bool weighted, double* exact,
                      float* ref) {
  if (isolated(graph, source)) {
    std::printf("Source [%d] is isolated node\n", source);  // cc.cu:11-14
    std::exit(0);
  }
  dawn_graph* g = upload(graph, weighted);
  int32_t s = source;
  int64_t sum = 0;
  dawn_stats st{};
  if (!weighted) {
    const int rc = dawn_closeness(g, &s, 1, exact, &sum, ref, DAWN_POLICY_LEARNED, &st);
    if (rc != DAWN_OK) die(rc, "dawn_closeness");
  } else {
    // The reference's weighted value (src/algorithm/cpu/cc.cpp:132-133):
    // the float distance row, FLT_MAX for unreached vertices, accumulated
    // in order from 0.0f, then (row - 1) / sum in double, as float.  So one
    // unreached vertex gives about (n-1)/FLT_MAX and two give 0.
    std::vector<int32_t> row((size_t)graph.rows);
    const int rc = dawn_mssp(g, &s, 1, row.data(), &sum, DAWN_POLICY_LEARNED, &st);
    if (rc != DAWN_OK) die(rc, "dawn_mssp");
    float acc = 0.0f;
    for (int j = 0; j < graph.rows; j++) acc += row[j] == DAWN_INF ? FLT_MAX : (float)row[j];
    *ref = (float)((1.0 * graph.rows - 1.0f) / acc);
    *exact = sum > 0 ? (double)(graph.rows - 1) / (double)sum : 0.0;
  }
  dawn_graph_destroy(g);
  std::printf("%-21s%3.5d\n", "Node:", graph.rows);  // cc.cu:96-100
  std::printf("%-21s%3.5ld\n", "Edges:", (long)graph.nnz);
  std::printf("%-21s%3.5lf\n", "Time:", st.device_ms / 1000.0);
}

float CC_GPU::run(Graph::Graph_t& graph, int source) {
  double exact = 0;
  float ref = 0;
  closeness(graph, source, false, &exact, &ref);
  return ref;
}

float CC_GPU::run_Weighted(Graph::Graph_t& graph, int source) {
  double exact = 0;
  float ref = 0;
  closeness(graph, source, true, &exact, &ref);
  return ref;
}

double CC_GPU::run_exact(Graph::Graph_t& graph, int source) {
  double exact = 0;
  float ref = 0;
  closeness(graph, source, graph.weighted, &exact, &ref);
  return exact;
}

}  // namespace DAWN
