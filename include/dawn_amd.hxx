// dawn_amd.hxx — C++ drop-in host API over the C ABI (dawn_capi.h).
//
// Same names, signatures and behaviour as the reference's host API so that
// a caller of lxrzlyr/GAL-DAWN's GPU entry points can switch by changing the
// include and the library:
//   DAWN::Graph::Graph_t / createGraph     include/dawn/graph.hxx:14-33
//   DAWN::Matrix::Csr_t / Coo_t            include/dawn/matrix.hxx:7-19
//   DAWN::Matrix::coo2Csr* / transpose*    include/dawn/matrix.hxx:21-30
//   DAWN::IO::readGraph* / readList /
//             outfile (all four)           include/dawn/io.hxx:7-24
//   DAWN::IO::parameters                   include/dawn/parameters.hxx:8-17
//   DAWN::BFS_GPU::run                     include/dawn/algorithm/gpu/bfs.cuh:12
//   DAWN::SSSP_GPU::run                    include/dawn/algorithm/gpu/sssp.cuh:20
//   DAWN::MSSP_GPU::run / run_Weighted     include/dawn/algorithm/gpu/mssp.cuh:13-15
//   DAWN::APSP_GPU::run / run_Weighted     include/dawn/algorithm/gpu/apsp.cuh:13-15
//   DAWN::CC_GPU::run / run_Weighted       include/dawn/algorithm/gpu/cc.cuh:27-29
// The CUDA-typed kernel() overloads (thrust::device_vector parameters) have
// no counterpart: the step level is reached through the C ABI.
//
// Behaviour kept: run() returns seconds; an isolated source prints the
// reference's message and exits with status 0 (single-source entry points)
// or is skipped (MSSP/APSP); graph.print writes "src dst dist" lines.
// Differences: weights must be integral (int32 min-plus).  CC_GPU::run
// returns the reference's value (row-1)/int32-sum with integer division
// (src/algorithm/cpu/cc.cpp:77-78), CC_GPU::run_Weighted the reference's
// float value (the distance row with FLT_MAX for unreached vertices summed
// in float, cc.cpp:132-133), CC_GPU::run_exact the engine's (n-1)/int64-sum
// over reached vertices in double.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace DAWN {

namespace Matrix {
struct Csr_t {
  int* row_ptr = nullptr;
  int* col = nullptr;
  float* val = nullptr;
};
struct Coo_t {
  int* row = nullptr;
  int* col = nullptr;
  float* val = nullptr;
};
// Row-sorted COO -> CSR (new[] arrays), src/matrix.cpp:28-82.
void coo2Csr(int n, int nnz, Csr_t& csr, Coo_t& coo);
void coo2Csr_Weighted(int n, int nnz, Csr_t& csr, Coo_t& coo);
// In-place sort of the COO by (row, col[, val]), src/matrix.cpp:3-26.
void transpose(int nnz, Coo_t& coo);
void transpose_Weighted(int nnz, Coo_t& coo);
}  // namespace Matrix

namespace Graph {
class Graph_t {
 public:
  int rows = 0;
  int cols = 0;
  uint64_t nnz = 0;
  DAWN::Matrix::Csr_t csr;
  DAWN::Matrix::Coo_t coo;
  int thread = 1;
  int interval = 100;
  int stream = 1;
  int block_size = 256;
  int source = 0;
  bool print = false;
  bool weighted = false;
  bool directed = false;
  std::vector<int> msource;  // source list for multi-source algorithms
};

// Reads a MatrixMarket file with the reference's semantics
// (src/graph.cpp:9-77, src/io.cpp:2-192); graph.weighted selects whether
// weights are read.  graph.source is reduced modulo rows.
void createGraph(std::string& input_path, Graph_t& graph);
}  // namespace Graph

namespace IO {
// The MatrixMarket readers with their own edge semantics (src/io.cpp:2-192):
// the undirected ones push both directions of each listed pair, the
// directed ones keep them as listed; self-loops dropped, 1-based ids,
// duplicates kept, rows sorted.  graph.rows/cols/nnz and graph.csr are set.
void readGraph(std::string& input_path, DAWN::Graph::Graph_t& graph);
void readGraph_Weighted(std::string& input_path, DAWN::Graph::Graph_t& graph);
void readGraph_Directed(std::string& input_path, DAWN::Graph::Graph_t& graph);
void readGraph_Directed_Weighted(std::string& input_path, DAWN::Graph::Graph_t& graph);
void readList(std::string& input_path, DAWN::Graph::Graph_t& graph);
void outfile(int n, int* result, int source, std::string& output_path);
void outfile(int n, float* result, int source, std::string& output_path);
// "j dist" lines without the source column (src/io.cpp:252-281).
void outfile(int n, int* result, std::string& output_path);
void outfile(int n, float* result, std::string& output_path);

struct parameters_t {
  std::string input_path;
  std::string output_path;
  std::string sourceList_path;
  bool print = false;
  bool weighted = false;
  int source = -1;
};
// getopt_long over -i/--input -o/--output -l/--sourceList -p/--print
// -w/--weighted -s/--source (src/parameters.cpp:24-98).
parameters_t parameters(int argc, char* argv[]);
void printHelp();
}  // namespace IO

namespace BFS_GPU {
float run(Graph::Graph_t& graph, std::string& output_path);
}
namespace SSSP_GPU {
float run(Graph::Graph_t& graph, std::string& output_path);
}
namespace MSSP_GPU {
float run(Graph::Graph_t& graph, std::string& output_path);
float run_Weighted(Graph::Graph_t& graph, std::string& output_path);
}  // namespace MSSP_GPU
namespace APSP_GPU {
float run(Graph::Graph_t& graph, std::string& output_path);
float run_Weighted(Graph::Graph_t& graph, std::string& output_path);
}  // namespace APSP_GPU
namespace CC_GPU {
float run(Graph::Graph_t& graph, int source);
float run_Weighted(Graph::Graph_t& graph, int source);
double run_exact(Graph::Graph_t& graph, int source);
}  // namespace CC_GPU

}  // namespace DAWN
