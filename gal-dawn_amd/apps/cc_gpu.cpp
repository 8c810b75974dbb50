// Drop-in for the reference's cc_gpu main (algorithm/gpu/cc_gpu/cc_gpu.cu:9-33).
#include <cstdio>
#include "../../include/dawn_amd.hxx"

int main(int argc, char* argv[]) {
  DAWN::IO::parameters_t params = DAWN::IO::parameters(argc, argv);
  DAWN::Graph::Graph_t graph;
  graph.source = params.source;
  graph.print = params.print;
  graph.weighted = params.weighted;
  DAWN::Graph::createGraph(params.input_path, graph);
  float closeness = graph.weighted ? DAWN::CC_GPU::run_Weighted(graph, graph.source)
                                   : DAWN::CC_GPU::run(graph, graph.source);
  printf("%-21s%3.5d\n", "Source:", graph.source);
  printf("%-21s%3.5lf\n", "Closeness Centrality:", closeness);
  return 0;
}
