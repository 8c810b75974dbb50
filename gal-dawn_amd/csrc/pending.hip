// Entry points whose engines are not built yet: they fail loudly.
#include "common.hpp"

extern "C" {
int dawn_batch_mssp(dawn_graph* const* graphs, int32_t num_graphs, const int32_t* sources,
                    int32_t count_per_graph, int32_t* dist, const int64_t* dist_offsets,
                    dawn_stats* stats) {
  (void)graphs; (void)num_graphs; (void)sources; (void)count_per_graph; (void)dist;
  (void)dist_offsets; (void)stats;
  return DAWN_ERR_UNSUPPORTED;
}
int dawn_tune(const dawn_graph* g, int32_t verbose) { (void)g; (void)verbose; return DAWN_ERR_UNSUPPORTED; }
int dawn_tune_load(const char* path) { (void)path; return DAWN_ERR_UNSUPPORTED; }
int dawn_tune_save(const char* path) { (void)path; return DAWN_ERR_UNSUPPORTED; }
}
