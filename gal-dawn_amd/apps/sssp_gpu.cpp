// Drop-in for the reference's sssp_gpu main (algorithm/gpu/sssp_gpu/sssp_gpu.cu:9-26).
#include <cstdio>
#include "../../include/dawn_amd.hxx"

int main(int argc, char* argv[]) {
  DAWN::IO::parameters_t params = DAWN::IO::parameters(argc, argv);
  DAWN::Graph::Graph_t graph;
  graph.source = params.source;
  graph.print = params.print;
  graph.weighted = true;
  DAWN::Graph::createGraph(params.input_path, graph);
  float elapsed_time = DAWN::SSSP_GPU::run(graph, params.output_path);
  printf("%-21s%3.5d\n", "Nodes:", graph.rows);
  printf("%-21s%3.5ld\n", "Edges:", (long)graph.nnz);
  printf("%-21s%3.5lf\n", "Time:", elapsed_time);
  return 0;
}
