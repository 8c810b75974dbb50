// Step-level entry point: one frontier step on caller-owned device state.
//
// The reference exposes its steps as public functions — SOVMP, the dense
// BOVM step (src/algorithm/cpu/bfs.cpp:159-185), GOVMP, the dense min-plus
// step (src/algorithm/cpu/sssp.cpp:169-199) — and CUDA-typed GPU kernel()
// overloads (include/dawn/algorithm/gpu/bfs.cuh:14-19, sssp.cuh:12-18).
// dawn_step is their device equivalent with plain pointers and an opaque
// stream: the frontier is a bitmap (alpha), the next frontier is OR-ed into
// a second bitmap (beta), alpha is cleared afterwards as SOVMP clears
// alpha[j], and the number of vertices that changed is counted on the
// device (0 = converged).
//
// Kernels: PUSH scans the frontier bitmap one lane per vertex; vertices with
// out-degree >= knobs.heavy_degree are expanded by wave-sized edge chunks
// (the graph's out-chunk list), so a hub does not serialise one lane.  PULL
// (unweighted and weighted) lets each unreached / improvable vertex scan its
// in-edges; heavy in-lists by chunks.  A vertex is counted once per step:
// by the lane whose atomicOr first sets its beta bit.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace dawn {
namespace {

struct StepArgs {
  int32_t n;
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* weight;      // null: unweighted
  const int32_t* in_row_ptr;
  const int32_t* in_col;
  const int32_t* in_weight;
  const int32_t* oc_v;        // heavy out-edge chunks
  const int32_t* oc_beg;
  const int32_t* oc_end;
  int64_t num_oc;
  const int32_t* hc_v;        // heavy in-edge chunks
  const int32_t* hc_beg;
  const int32_t* hc_end;
  int64_t num_hc;
  int32_t heavy;
  int32_t level;
  const uint32_t* alpha;
  uint32_t* beta;
  int32_t* dist;
  unsigned long long* changed;
};

__device__ __forceinline__ bool in_frontier(const uint32_t* bm, int32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Marks v in beta; true for the first marking of v in this step.
__device__ __forceinline__ bool mark(uint32_t* beta, int32_t v) {
  const uint32_t bit = 1u << (v & 31);
  return !(atomicOr(&beta[v >> 5], bit) & bit);
}

// Candidate distance of v through frontier vertex u over edge e.
__device__ __forceinline__ int32_t through(const StepArgs& a, int32_t u, const int32_t* w,
                                           int64_t e) {
  if (!a.weight) return a.level;
  return minplus(a.dist[u], w[e]);
}

// Relaxes v with candidate d (atomicMin); counts v on its first marking.
__device__ __forceinline__ void relax(const StepArgs& a, int32_t v, int32_t d, unsigned& cnt) {
  if (d >= a.dist[v]) return;
  if (d < atomicMin(&a.dist[v], d) && mark(a.beta, v)) cnt++;
}

__device__ __forceinline__ void flush_count(const StepArgs& a, unsigned cnt) {
  const unsigned long long t = wave_sum_u64(cnt);
  if (lane_id() == 0 && t) atomicAdd(a.changed, t);
}

__global__ void step_reset_kernel(unsigned long long* changed) { *changed = 0; }

__global__ void __launch_bounds__(kBlock) step_push_kernel(StepArgs a) {
  unsigned cnt = 0;
  const int64_t gs = (int64_t)gridDim.x * kBlock;
  for (int64_t u0 = (int64_t)blockIdx.x * kBlock + threadIdx.x; u0 < a.n; u0 += gs) {
    const int32_t u = (int32_t)u0;
    if (!in_frontier(a.alpha, u)) continue;
    const int beg = a.row_ptr[u], end = a.row_ptr[u + 1];
    if (end - beg >= a.heavy) continue;  // chunked below
    for (int e = beg; e < end; e++) relax(a, a.col[e], through(a, u, a.weight, e), cnt);
  }
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t c = wave; c < a.num_oc; c += nwaves) {
    const int32_t u = a.oc_v[c];
    if (!in_frontier(a.alpha, u)) continue;
    const int end = a.oc_end[c];
    for (int e = a.oc_beg[c] + lane; e < end; e += kWave)
      relax(a, a.col[e], through(a, u, a.weight, e), cnt);
  }
  flush_count(a, cnt);
}

__global__ void __launch_bounds__(kBlock) step_pull_kernel(StepArgs a) {
  unsigned cnt = 0;
  const int64_t gs = (int64_t)gridDim.x * kBlock;
  for (int64_t v0 = (int64_t)blockIdx.x * kBlock + threadIdx.x; v0 < a.n; v0 += gs) {
    const int32_t v = (int32_t)v0;
    if (!a.weight && a.dist[v] != kInf) continue;
    const int beg = a.in_row_ptr[v], end = a.in_row_ptr[v + 1];
    if (end - beg >= a.heavy) continue;  // chunked below
    int32_t best = kInf;
    for (int e = beg; e < end; e++) {
      const int32_t u = a.in_col[e];
      if (!in_frontier(a.alpha, u)) continue;
      best = min(best, through(a, u, a.in_weight, e));
      if (!a.weight) break;  // any frontier in-neighbour gives `level`
    }
    if (best != kInf) relax(a, v, best, cnt);
  }
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t c = wave; c < a.num_hc; c += nwaves) {
    const int32_t v = a.hc_v[c];
    if (!a.weight && a.dist[v] != kInf) continue;
    int32_t best = kInf;
    const int end = a.hc_end[c];
    for (int e = a.hc_beg[c] + lane; e < end; e += kWave) {
      const int32_t u = a.in_col[e];
      if (in_frontier(a.alpha, u)) best = min(best, through(a, u, a.in_weight, e));
    }
    best = wave_min_i32(best);
    if (lane == 0 && best != kInf) relax(a, v, best, cnt);
  }
  flush_count(a, cnt);
}

__global__ void step_clear_kernel(uint32_t* alpha, int64_t words) {
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += gs) alpha[w] = 0u;
}

}  // namespace
}  // namespace dawn

using namespace dawn;

extern "C" int dawn_step(const dawn_graph* g, const dawn_frontier* f, int32_t mode, void* stream) {
  if (!g || !f || !f->alpha || !f->beta || !f->dist || !f->changed) return DAWN_ERR_INVALID_ARG;
  if (mode != DAWN_MODE_PUSH && mode != DAWN_MODE_PULL) return DAWN_ERR_INVALID_ARG;
  for (const void* p : {(const void*)f->alpha, (const void*)f->beta, (const void*)f->dist,
                        (const void*)f->changed})
    if (!is_device_pointer(p)) return DAWN_ERR_INVALID_ARG;
  DAWN_HIP(hipSetDevice(g->device));
  hipStream_t s = stream ? (hipStream_t)stream : g->stream;
  StepArgs a{};
  a.n = g->n;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.weight = g->weighted ? g->weight : nullptr;
  a.in_row_ptr = g->in_row_ptr;
  a.in_col = g->in_col;
  a.in_weight = g->weighted ? g->in_weight : nullptr;
  a.oc_v = g->oc_v;
  a.oc_beg = g->oc_beg;
  a.oc_end = g->oc_end;
  a.num_oc = g->num_oc;
  a.hc_v = g->hc_v;
  a.hc_beg = g->hc_beg;
  a.hc_end = g->hc_end;
  a.num_hc = g->num_hc;
  a.heavy = g->knobs.heavy_degree;
  a.level = f->level;
  a.alpha = f->alpha;
  a.beta = f->beta;
  a.dist = f->dist;
  a.changed = reinterpret_cast<unsigned long long*>(f->changed);
  const unsigned grid = grid_for(std::max<int64_t>(g->n, 1), kBlock, 2048);
  step_reset_kernel<<<1, 1, 0, s>>>(a.changed);
  if (mode == DAWN_MODE_PUSH) step_push_kernel<<<grid, kBlock, 0, s>>>(a);
  else step_pull_kernel<<<grid, kBlock, 0, s>>>(a);
  const int64_t words = ((int64_t)g->n + 31) / 32;
  step_clear_kernel<<<grid_for(words), kBlock, 0, s>>>(f->alpha, words);
  DAWN_HIP(hipGetLastError());
  return DAWN_OK;
}
