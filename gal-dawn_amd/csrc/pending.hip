// Entry points whose engines are not built yet: they fail loudly.
#include "common.hpp"

extern "C" {
int dawn_tune(const dawn_graph* g, int32_t verbose) { (void)g; (void)verbose; return DAWN_ERR_UNSUPPORTED; }
int dawn_tune_load(const char* path) { (void)path; return DAWN_ERR_UNSUPPORTED; }
int dawn_tune_save(const char* path) { (void)path; return DAWN_ERR_UNSUPPORTED; }
}
