// Step scheduler shared by the frontier engines: per-level counters, the
// per-step feature vector, the GAPBS alpha/beta rule and MLP inference.
// The decision runs on the device (one wave) so that step loops can be
// captured in a hipGraph without host round trips.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/dawn_capi.h"

namespace dawn {

// Counters of one completed level, accumulated by the step kernels.
struct LevelCounters {
  unsigned long long fv;      // frontier vertices (new this level)
  unsigned long long fe;      // sum of out-degree over them
  unsigned long long scout;   // sum of max(out-degree, 1) (GAPBS scout_count)
  unsigned long long flanes;  // sum of popcount of frontier words (batched)
  unsigned long long uv;      // vertices still missing a live lane after this level
  unsigned long long ue;      // sum of in-degree over those (pull's edges to check)
  unsigned long long alive[16];  // batched: OR of the level's frontier words, one
                                 // per 64 lanes (lanes that can still reach new vertices)
  unsigned int qlen;          // queue entries written for this level
  unsigned int pad;
};

constexpr int kRing = 4;          // ring of LevelCounters indexed by level % 4
constexpr int kTraceMax = 8192;   // steps recorded in the device trace

// Device-resident scheduler state of one traversal.
struct SchedState {
  int32_t step;      // last completed level (the decide kernel advances it)
  int32_t done;      // set when a level comes out empty
  int32_t mode;      // mode of the step in flight
  int32_t phase;     // fixed rule: 0 top-down, 1 bottom-up
  long long e2c;     // fixed rule: edges_to_check
  long long scout;   // fixed rule: scout_count
  long long awake;   // fixed rule: awake_count of the previous bottom-up step
  int32_t last_mode;
  int32_t levels;    // number of non-empty levels (set when done)
  int32_t force_mode;// >= 0: every step uses this mode
  int32_t source;    // single-source engines: the traversal's source (read by
                     // captured step kernels; never a baked launch argument)
  unsigned long long work_fv;  // frontier vertices expanded, summed over the steps
  unsigned long long work_fe;  // their out-edges (the relaxations of a push schedule)
};

// Constant per traversal.
struct SchedParams {
  int32_t n;
  int32_t policy;        // DAWN_POLICY_*
  long long m;
  int32_t alpha, beta;   // GAPBS defaults 15, 18 (test/gabps/bfs.cc:125-126)
  int32_t batched;       // 1: 64-lane rule variant
  int32_t allow_pull;    // 0 for engines without a pull mode
  float avg_degree;
  float max_degree;
  const float* mlp;      // device params (DAWN_SCHED_PARAMS) or null
  const float* mlp_norm; // device mean[8], inv_std[8] or null
  long long uv_base;     // unvisited vertices the counters leave out (the
                         // batched engine never folds vertices of in-degree 0)
};

struct SchedTrace {
  float* feat;   // kTraceMax * 8
  int32_t* mode; // kTraceMax
};

__host__ __device__ inline float log2p1(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __log2f((float)x + 1.0f);
#else
  return (float)__builtin_log2(x + 1.0);
#endif
}

// Feature vector of the step that expands the level with counters c.
// Vertices still to reach and their edges: the batched engine counts them
// per level (vertices missing a live lane, their in-degree sum); the
// single-source engine's levels are disjoint, so they are n and m minus
// the levels found so far (the expanded ones, s.work_*, and this one).
__host__ __device__ inline void sched_features(const LevelCounters& c, int step,
                                               const SchedParams& p, const SchedState& s,
                                               float* x) {
  double uv, ue;
  if (p.batched) {
    uv = (double)c.uv + (double)p.uv_base;
    ue = (double)c.ue;
  } else {
    uv = (double)p.n - (double)s.work_fv - (double)c.fv;
    ue = (double)p.m - (double)s.work_fe - (double)c.fe;
    uv = uv > 0 ? uv : 0;
    ue = ue > 0 ? ue : 0;
  }
  x[0] = log2p1((double)c.fv);
  x[1] = log2p1((double)c.fe);
  x[2] = log2p1(uv);
  x[3] = log2p1(ue);
  x[4] = (float)step;
  x[5] = log2p1((double)p.n);
  x[6] = log2p1((double)p.avg_degree);
  x[7] = log2p1((double)p.max_degree);
}

// One-wave MLP inference (8 -> 32 -> 32 -> 3, ReLU); returns argmin of the
// predicted log step times over the allowed modes.  Called by all 64 lanes.
__device__ inline int sched_mlp_mode(const float* x, const SchedParams& p) {
  __shared__ float xin[8], h1[32], h2[32], out[3];
  const int t = threadIdx.x;
  const float* W1 = p.mlp;
  const float* b1 = W1 + 32 * 8;
  const float* W2 = b1 + 32;
  const float* b2 = W2 + 32 * 32;
  const float* W3 = b2 + 32;
  const float* b3 = W3 + 3 * 32;
  if (t < 8) {
    float v = x[t];
    if (p.mlp_norm) v = (v - p.mlp_norm[t]) * p.mlp_norm[8 + t];
    xin[t] = v;
  }
  __syncthreads();
  if (t < 32) {
    float a = b1[t];
    for (int j = 0; j < 8; j++) a += W1[t * 8 + j] * xin[j];
    h1[t] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (t < 32) {
    float a = b2[t];
    for (int j = 0; j < 32; j++) a += W2[t * 32 + j] * h1[j];
    h2[t] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (t < 3) {
    float a = b3[t];
    for (int j = 0; j < 32; j++) a += W3[t * 32 + j] * h2[j];
    out[t] = a;
  }
  __syncthreads();
  int best = DAWN_MODE_PUSH;
  float bv = out[0];
  if (p.allow_pull && out[1] < bv) { best = DAWN_MODE_PULL; bv = out[1]; }
  if (out[2] < bv) best = DAWN_MODE_BITMAP;
  return best;
}

// GAPBS direction-optimizing rule, test/gabps/bfs.cc:143-176, as a state
// machine over completed levels.  Single source: edges_to_check starts at m
// and drops by scout on every top-down step; scout_count is the
// max(deg,1)-sum of the level just discovered.  Batched (64 lanes): the
// same rule with edges_to_check measured directly as the in-degree sum of
// vertices that still miss a lane.  Returns DAWN_MODE_PUSH or _PULL.
__device__ inline int sched_fixed_mode(SchedState& s, const LevelCounters& c,
                                       const SchedParams& p) {
  if (s.step >= 1 && s.last_mode != DAWN_MODE_PULL) s.scout = (long long)c.scout;
  if (p.batched) s.e2c = (long long)c.ue;
  if (s.phase == 1) {
    long long old = s.awake;
    s.awake = (long long)c.fv;
    if (s.awake >= old || s.awake > (long long)p.n / p.beta) return DAWN_MODE_PULL;
    s.phase = 0;
    s.scout = 1;
  }
  if (s.scout > s.e2c / p.alpha) {
    s.phase = 1;
    s.awake = (long long)c.fv;
    return DAWN_MODE_PULL;
  }
  if (!p.batched) s.e2c -= s.scout;
  return DAWN_MODE_PUSH;
}

// Decide the mode of the next step.  Launched as one wave (64 threads).
// Advances s.step, sets s.done when the last level was empty, records the
// features and mode in the trace.
__device__ inline void sched_decide(SchedState* sp, const LevelCounters* ring,
                                    const SchedParams& p, SchedTrace tr) {
  __shared__ int s_mode, s_done;
  __shared__ float s_x[8];
  const int t = threadIdx.x;
  if (t == 0) {
    SchedState s = *sp;
    s_done = s.done;
    if (!s.done) {
      const LevelCounters c = ring[s.step % kRing];
      if (c.fv == 0) {
        s.done = 1;
        s.levels = s.step;  // levels 0..step-1 are non-empty
        s_done = 1;
      } else {
        float x[8];
        sched_features(c, s.step + 1, p, s, x);
        for (int i = 0; i < 8; i++) s_x[i] = x[i];
        int mode;
        if (s.force_mode >= 0) {
          mode = s.force_mode;
        } else {
          mode = sched_fixed_mode(s, c, p);  // also keeps the rule state current
          if (p.policy == DAWN_POLICY_PUSH) mode = DAWN_MODE_PUSH;
          else if (p.policy == DAWN_POLICY_PULL) mode = DAWN_MODE_PULL;
          else if (p.policy == DAWN_POLICY_BITMAP) mode = DAWN_MODE_BITMAP;
        }
        if (!p.allow_pull && mode == DAWN_MODE_PULL) mode = DAWN_MODE_BITMAP;
        s.step += 1;
        s.mode = mode;
        s.work_fv += c.fv;  // the level this step expands
        s.work_fe += c.fe;
      }
    }
    *sp = s;
    s_mode = sp->mode;
  }
  __syncthreads();
  if (s_done) return;
  if (p.policy == DAWN_POLICY_LEARNED && p.mlp && sp->force_mode < 0) {
    int m2 = sched_mlp_mode(s_x, p);
    if (t == 0) {
      sp->mode = m2;
      s_mode = m2;
    }
    __syncthreads();
  }
  if (t == 0) {
    sp->last_mode = s_mode;
    int st = sp->step;
    if (st < kTraceMax && tr.mode) {
      tr.mode[st] = s_mode;
      for (int i = 0; i < 8; i++) tr.feat[st * 8 + i] = s_x[i];
    }
  }
}

// The single-source engine's network on `device` (installed there on first
// use) and its normalisation, or null when none is set; the variant that
// never allocates returns null unless it is already there.
const float* sched_single_mlp_device(int device, const float** norm);
const float* sched_single_mlp_if_present(int device, const float** norm);
// Host-side view of the installed scheduler network (sched.cpp).
const float* sched_mlp_device(int device);
const float* sched_norm_device(int device);
// The network and normalisation already on `device`, or null: never
// allocates (the device copy is made on first use by sched_mlp_device).
const float* sched_mlp_device_if_present(int device, const float** norm);

}  // namespace dawn
