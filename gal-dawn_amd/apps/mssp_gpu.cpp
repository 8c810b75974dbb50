// Drop-in for the reference's mssp_gpu main (algorithm/gpu/mssp_gpu/mssp_gpu.cu:9-37).
#include <cstdio>
#include "../../include/dawn_amd.hxx"

int main(int argc, char* argv[]) {
  DAWN::IO::parameters_t params = DAWN::IO::parameters(argc, argv);
  DAWN::Graph::Graph_t graph;
  graph.print = params.print;
  graph.weighted = params.weighted;
  graph.source = params.source;
  DAWN::Graph::createGraph(params.input_path, graph);
  DAWN::IO::readList(params.sourceList_path, graph);
  float elapsed_time = graph.weighted ? DAWN::MSSP_GPU::run_Weighted(graph, params.output_path)
                                      : DAWN::MSSP_GPU::run(graph, params.output_path);
  printf("%-21s%3.5d\n", "Nodes:", graph.rows);
  printf("%-21s%3.5ld\n", "Edges:", (long)graph.nnz);
  printf("%-21s%3.5lf\n", "Time:", elapsed_time);
  return 0;
}
