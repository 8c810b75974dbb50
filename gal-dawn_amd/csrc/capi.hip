// C ABI entry points of the frontier engine (declared in include/dawn_capi.h)
// and the process-wide scheduler network store.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "sched.hpp"

namespace dawn {

int run_single_source(dawn_graph* g, int kind, int32_t source, int32_t* dist_out, int32_t policy,
                      dawn_stats* stats, int32_t force_mode, float* trace_feat,
                      int32_t* trace_mode, int32_t trace_max, int32_t* trace_rows);
// rows: with dist null, still compute every distance row (into a scratch
// buffer) rather than only the per-source sums.
int run_msbfs(dawn_graph* g, const int32_t* sources, int32_t S, int32_t* dist, int64_t* sums,
              int64_t* reached, int64_t* edges, int32_t policy, dawn_stats* stats, bool rows);
int run_wmsbfs(dawn_graph* g, const int32_t* sources, int32_t S, int32_t* dist, int64_t* sums,
               int32_t policy, dawn_stats* stats);
int run_ms_record(dawn_graph* g, const int32_t* sources, int32_t S, int32_t repeats,
                  float* features, float* times_ms, int32_t* chosen, int32_t max_rows,
                  int32_t* rows_out, int32_t policy);
int run_ms_profile(dawn_graph* g, const int32_t* sources, int32_t S, int32_t policy,
                   double* kernel_ms, int64_t* launches, int32_t* levels_out);
int run_ss_record(dawn_graph* g, const int32_t* sources, int32_t S, int32_t repeats,
                  float* features, float* times_ms, int32_t* chosen, int32_t max_rows,
                  int32_t* rows_out, int32_t policy);

// ---- scheduler network store (one copy per device) -------------------------
namespace {
std::mutex g_sched_lock;
struct SchedNet {
  float* params = nullptr;  // device
  float* norm = nullptr;    // device mean[8], inv_std[8]
};
// Batched engine's network (g_*), single-source engine's network (g_single_*).
std::map<int, SchedNet> g_nets;
std::vector<float> g_host_params, g_host_norm;
bool g_have_net = false;
std::map<int, SchedNet> g_single_nets;
std::vector<float> g_single_params, g_single_norm;
bool g_have_single = false;

int install_net(SchedNet& net, const std::vector<float>& params, const std::vector<float>& norm) {
  if (!net.params) DAWN_HIP(hipMalloc(&net.params, sizeof(float) * DAWN_SCHED_PARAMS));
  if (!net.norm) DAWN_HIP(hipMalloc(&net.norm, sizeof(float) * 16));
  DAWN_HIP(hipMemcpy(net.params, params.data(), sizeof(float) * DAWN_SCHED_PARAMS,
                     hipMemcpyHostToDevice));
  DAWN_HIP(hipMemcpy(net.norm, norm.data(), sizeof(float) * 16, hipMemcpyHostToDevice));
  return DAWN_OK;
}

int install_on(int device, SchedNet& net) {
  (void)device;
  return install_net(net, g_host_params, g_host_norm);
}

// Copies the host parameters and normalisation into the store.
void set_host_net(const float* params, const float* mean, const float* inv_std,
                  std::vector<float>& hp, std::vector<float>& hn) {
  hp.assign(params, params + DAWN_SCHED_PARAMS);
  hn.assign(16, 0.f);
  for (int i = 0; i < 8; i++) {
    hn[i] = mean ? mean[i] : 0.f;
    hn[8 + i] = inv_std ? inv_std[i] : 1.f;
  }
}
}  // namespace

// The single-source engine's network on `device` (installed there on first
// use) and its normalisation, or null when none is set.
const float* sched_single_mlp_device(int device, const float** norm) {
  std::lock_guard<std::mutex> l(g_sched_lock);
  *norm = nullptr;
  if (!g_have_single) return nullptr;
  SchedNet& net = g_single_nets[device];
  if (!net.params) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device);
    const int rc = install_net(net, g_single_params, g_single_norm);
    (void)hipSetDevice(prev);
    if (rc != DAWN_OK) return nullptr;
  }
  *norm = net.norm;
  return net.params;
}

// The same without installing (never allocates): null unless already there.
const float* sched_single_mlp_if_present(int device, const float** norm) {
  std::lock_guard<std::mutex> l(g_sched_lock);
  *norm = nullptr;
  if (!g_have_single) return nullptr;
  auto it = g_single_nets.find(device);
  if (it == g_single_nets.end() || !it->second.params) return nullptr;
  *norm = it->second.norm;
  return it->second.params;
}

const float* sched_mlp_device(int device) {
  std::lock_guard<std::mutex> l(g_sched_lock);
  if (!g_have_net) return nullptr;
  SchedNet& net = g_nets[device];
  if (!net.params) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device);
    int rc = install_on(device, net);
    (void)hipSetDevice(prev);
    if (rc != DAWN_OK) return nullptr;
  }
  return net.params;
}

const float* sched_mlp_device_if_present(int device, const float** norm) {
  std::lock_guard<std::mutex> l(g_sched_lock);
  *norm = nullptr;
  if (!g_have_net) return nullptr;
  auto it = g_nets.find(device);
  if (it == g_nets.end() || !it->second.params) return nullptr;
  *norm = it->second.norm;
  return it->second.params;
}

const float* sched_norm_device(int device) {
  if (!sched_mlp_device(device)) return nullptr;
  std::lock_guard<std::mutex> l(g_sched_lock);
  return g_nets[device].norm;
}


}  // namespace dawn

using namespace dawn;

extern "C" {

int dawn_bfs(const dawn_graph* gc, int32_t source, int32_t* dist, int32_t policy,
             dawn_stats* stats) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g) return DAWN_ERR_INVALID_ARG;
  return run_single_source(g, 0, source, dist, policy, stats, -1, nullptr, nullptr, 0, nullptr);
}

int dawn_sssp(const dawn_graph* gc, int32_t source, int32_t* dist, int32_t policy,
              dawn_stats* stats) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g) return DAWN_ERR_INVALID_ARG;
  return run_single_source(g, 1, source, dist, policy, stats, -1, nullptr, nullptr, 0, nullptr);
}

int dawn_mssp(const dawn_graph* gc, const int32_t* sources, int32_t count, int32_t* dist,
              int64_t* sums, int32_t policy, dawn_stats* stats) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g || count < 0 || (count > 0 && !sources)) return DAWN_ERR_INVALID_ARG;
  if (count == 0) return DAWN_OK;
  if (g->weighted) return run_wmsbfs(g, sources, count, dist, sums, policy, stats);
  return run_msbfs(g, sources, count, dist, sums, nullptr, nullptr, policy, stats, true);
}

int dawn_apsp_block(const dawn_graph* gc, int32_t first, int32_t count, int32_t* dist,
                    int64_t* sums, int32_t policy, dawn_stats* stats) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g || first < 0 || count < 0 || (int64_t)first + count > g->n) return DAWN_ERR_INVALID_ARG;
  std::vector<int32_t> src(count);
  for (int32_t i = 0; i < count; i++) src[i] = first + i;
  return dawn_mssp(g, src.data(), count, dist, sums, policy, stats);
}

int dawn_closeness(const dawn_graph* gc, const int32_t* sources, int32_t count,
                   double* closeness, int64_t* sums, float* ref_int, int32_t policy,
                   dawn_stats* stats) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g || count < 0 || (count > 0 && !sources)) return DAWN_ERR_INVALID_ARG;
  std::vector<int64_t> s(count);
  if (count > 0) {
    // Sums only: the unweighted engine skips the distance rows.
    if (g->weighted) DAWN_TRY(run_wmsbfs(g, sources, count, nullptr, s.data(), policy, stats));
    else DAWN_TRY(run_msbfs(g, sources, count, nullptr, s.data(), nullptr, nullptr, policy, stats, false));
  }
  for (int32_t i = 0; i < count; i++) {
    // (n-1) / sum, sum exact in int64, one IEEE double division
    // (src/algorithm/cpu/cc.cpp:77-78 divides by an int32 sum instead).
    if (closeness) closeness[i] = s[i] > 0 ? (double)(g->n - 1) / (double)s[i] : 0.0;
    if (sums) sums[i] = s[i];
    if (ref_int) {
      const int32_t a = (int32_t)(uint32_t)(uint64_t)s[i];  // int32 accumulate wraps
      ref_int[i] = (!g->weighted && a != 0) ? (float)((g->n - 1) / a) : 0.0f;
    }
  }
  return DAWN_OK;
}

int dawn_mssp_profile(const dawn_graph* gc, const int32_t* sources, int32_t count,
                      int32_t policy, double* kernel_ms, int64_t* launches,
                      int32_t* levels) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g) return DAWN_ERR_INVALID_ARG;
  if (g->weighted) return DAWN_ERR_UNSUPPORTED;
  return run_ms_profile(g, sources, count, policy, kernel_ms, launches, levels);
}

int dawn_sched_set_mlp(const float* params, const float* feat_mean, const float* feat_inv_std) {
  std::lock_guard<std::mutex> l(g_sched_lock);
  if (!params) {
    g_have_net = false;
    return DAWN_OK;
  }
  set_host_net(params, feat_mean, feat_inv_std, g_host_params, g_host_norm);
  g_have_net = true;
  // Refresh copies already on devices.
  for (auto& kv : g_nets) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(kv.first);
    int rc = install_on(kv.first, kv.second);
    (void)hipSetDevice(prev);
    if (rc != DAWN_OK) return rc;
  }
  return DAWN_OK;
}

int dawn_sched_set_mlp_single(const float* params, const float* feat_mean,
                              const float* feat_inv_std) {
  std::lock_guard<std::mutex> l(g_sched_lock);
  if (!params) {
    g_have_single = false;
    return DAWN_OK;
  }
  set_host_net(params, feat_mean, feat_inv_std, g_single_params, g_single_norm);
  g_have_single = true;
  for (auto& kv : g_single_nets) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(kv.first);
    const int rc = install_net(kv.second, g_single_params, g_single_norm);
    (void)hipSetDevice(prev);
    if (rc != DAWN_OK) return rc;
  }
  return DAWN_OK;
}

int dawn_sched_record(const dawn_graph* gc, const int32_t* sources, int32_t count,
                      int32_t repeats, float* features, float* times_ms, int32_t* fixed_modes,
                      int32_t max_rows, int32_t* rows) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g || !features || !times_ms || !rows || max_rows < 0) return DAWN_ERR_INVALID_ARG;
  if (g->weighted) return DAWN_ERR_UNSUPPORTED;
  return run_ms_record(g, sources, count, repeats, features, times_ms, fixed_modes, max_rows,
                       rows, DAWN_POLICY_FIXED);
}

int dawn_sched_trace(const dawn_graph* gc, const int32_t* sources, int32_t count, int32_t policy,
                     int32_t* modes, float* features, float* times_ms, int32_t max_rows,
                     int32_t* rows) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g || !rows || max_rows < 0) return DAWN_ERR_INVALID_ARG;
  if (g->weighted) return DAWN_ERR_UNSUPPORTED;
  // The per-step modes of the policy, with every mode timed at each step.
  std::vector<float> f((size_t)max_rows * 8), t((size_t)max_rows * DAWN_NUM_MODES);
  std::vector<int32_t> chosen(max_rows);
  DAWN_TRY(run_ms_record(g, sources, count, 1, f.data(), t.data(), chosen.data(), max_rows, rows,
                         policy));
  for (int32_t i = 0; i < *rows; i++) {
    if (modes) modes[i] = chosen[i];
    if (features) std::memcpy(features + (size_t)i * 8, f.data() + (size_t)i * 8, 8 * sizeof(float));
    if (times_ms)
      std::memcpy(times_ms + (size_t)i * DAWN_NUM_MODES, t.data() + (size_t)i * DAWN_NUM_MODES,
                  DAWN_NUM_MODES * sizeof(float));
  }
  return DAWN_OK;
}

int dawn_sched_record_single(const dawn_graph* gc, const int32_t* sources, int32_t count,
                             int32_t repeats, float* features, float* times_ms,
                             int32_t* fixed_modes, int32_t max_rows, int32_t* rows) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g || !features || !times_ms || !rows || max_rows < 0) return DAWN_ERR_INVALID_ARG;
  return run_ss_record(g, sources, count, repeats, features, times_ms, fixed_modes, max_rows,
                       rows, DAWN_POLICY_FIXED);
}

int dawn_sched_trace_single(const dawn_graph* gc, const int32_t* sources, int32_t count,
                            int32_t policy, int32_t* modes, float* features, float* times_ms,
                            int32_t max_rows, int32_t* rows) {
  dawn_graph* g = const_cast<dawn_graph*>(gc);
  if (!g || !rows || max_rows < 0) return DAWN_ERR_INVALID_ARG;
  std::vector<float> f((size_t)max_rows * 8), t((size_t)max_rows * DAWN_NUM_MODES);
  std::vector<int32_t> chosen(max_rows);
  DAWN_TRY(run_ss_record(g, sources, count, 1, f.data(), t.data(), chosen.data(), max_rows, rows,
                         policy));
  for (int32_t i = 0; i < *rows; i++) {
    if (modes) modes[i] = chosen[i];
    if (features) std::memcpy(features + (size_t)i * 8, f.data() + (size_t)i * 8, 8 * sizeof(float));
    if (times_ms)
      std::memcpy(times_ms + (size_t)i * DAWN_NUM_MODES, t.data() + (size_t)i * DAWN_NUM_MODES,
                  DAWN_NUM_MODES * sizeof(float));
  }
  return DAWN_OK;
}

}  // extern "C"
