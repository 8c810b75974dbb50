// Single-source frontier engine: unweighted DAWN (BOVM) and int32 min-plus
// DAWN (SOVM) on one HIP device.
//
// Reference: DAWN::BFS_GPU (src/algorithm/gpu/bfs.cu:9-129) and
// DAWN::SSSP_GPU (src/algorithm/gpu/sssp.cu:18-157); CPU semantics
// DAWN::BFS_CPU (src/algorithm/cpu/bfs.cpp:23-185) and DAWN::SSSP_CPU
// (src/algorithm/cpu/sssp.cpp:23-199).
//
// Layout in HBM (per graph, reused across calls):
//   dist   int32[n]                      distances (INF = unreached)
//   q      int32[3][n]                   frontier queue of level l in slot l%3
//   bm     uint32[3][words]              frontier bitmap of level l in slot l%3
//   ring   LevelCounters[4]              per-level counters, level l in l%4
//   state  SchedState                    step / done / mode / rule state
// Every step writes BOTH the queue and the bitmap of the new level, so the
// three modes are interchangeable from step to step with no conversion:
//   PUSH    queue in: workgroup-balanced edge expansion (LDS degree scan)
//   PULL    bitmap in: unvisited vertices scan in-edges (CSC) with early exit
//   BITMAP  bitmap in: dense scan of the frontier bitmap, top-down expansion
// The step loop is decide(1 wave) + expand(grid), captured K steps at a
// time into a hipGraph; the host reads one flag per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "sched.hpp"

namespace dawn {

struct SsArgs {
  int32_t n;
  int32_t source;    // read only by the (uncaptured) seed kernels
  int64_t words;  // bitmap words per level slot
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* w;
  const int32_t* in_row_ptr;
  const int32_t* in_col;
  const int32_t* in_w;
  int32_t* dist;
  int32_t* q;        // 3 * n
  uint32_t* bm;      // 3 * words
  LevelCounters* ring;
  unsigned long long* part;  // kRing x kStripes counter stripes
  SchedState* st;
};

// Level counters (fv, fe, scout) are added per workgroup into one of
// kStripes stripes of the level, each on its own 128-byte line, so the
// atomics of thousands of workgroups spread over 64 addresses; the decide
// kernel sums the stripes of the level it reads.
constexpr int kStripes = 64;
constexpr int kStripeU64 = 16;  // 128 bytes per stripe
constexpr int kPartU64 = kRing * kStripes * kStripeU64;
// Queue entries a workgroup stages in LDS before reserving global space.
constexpr int kStageCap = 4096;
// Queue length of a level whose queue was not written (a pull step writes
// the bitmap only): a top-down step that follows reads the bitmap instead.
constexpr unsigned kNoQueue = 0xFFFFFFFFu;
static_assert(kStripes == kWave, "the decide kernel sums one stripe per lane");

struct SsWorkspace {
  int32_t n = 0;
  int64_t words = 0;
  DevBuf<int32_t> dist, q;
  DevBuf<uint32_t> bm;
  DevBuf<LevelCounters> ring;
  DevBuf<unsigned long long> part;
  DevBuf<SchedState> st;
  DevBuf<float> trace_feat;
  DevBuf<int32_t> trace_mode;
  DevBuf<unsigned long long> reach;  // [2]
  SchedState* h_state = nullptr;     // pinned
  int32_t last_pairs[2] = {0, 0};     // decide+expand pairs the last traversal (per kind) took
  // Captured K-step chunks keyed by (kind, policy, K).
  std::map<std::pair<long long, const void*>, hipGraphExec_t> graphs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

void ss_invalidate_graphs(SsWorkspace* ws) {
  if (!ws) return;
  for (auto& kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
  ws->graphs.clear();
}

void ss_workspace_free(SsWorkspace* ws) {
  if (!ws) return;
  for (auto& kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
  if (ws->h_state) (void)hipHostFree(ws->h_state);
  if (ws->ev0) (void)hipEventDestroy(ws->ev0);
  if (ws->ev1) (void)hipEventDestroy(ws->ev1);
  delete ws;
}

namespace {

__device__ __forceinline__ bool test_bit(const uint32_t* bm, int32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Largest j in [0, kBlock) with off[j] <= e (off non-decreasing).
__device__ __forceinline__ int owner_of(const int* off, int e) {
  int lo = 0, hi = kBlock - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct Acc {
  unsigned long long fv = 0, fe = 0, scout = 0;
  __device__ void add(const int32_t* row_ptr, int32_t v) {
    int d = row_ptr[v + 1] - row_ptr[v];
    fv += 1;
    fe += (unsigned)d;
    scout += (unsigned)(d > 0 ? d : 1);
  }
  // Wave sums into the workgroup's LDS totals (all 64 lanes).
  __device__ void reduce(unsigned long long* s_tot) {
    unsigned long long a = wave_sum_u64(fv), b = wave_sum_u64(fe), s = wave_sum_u64(scout);
    if (lane_id() == 0 && a) {
      atomicAdd(&s_tot[0], a);
      atomicAdd(&s_tot[1], b);
      atomicAdd(&s_tot[2], s);
    }
  }
};

// Workgroup-level staging of a step's outputs: the new level's queue
// entries and its counters.
struct SsStage {
  LdsQueue<kStageCap> q;
  unsigned long long tot[3];
  __device__ void init() {
    if (threadIdx.x == 0) {
      q.init();
      tot[0] = tot[1] = tot[2] = 0ull;
    }
    __syncthreads();
  }
  // Every thread, once, at the end of the step kernel.
  __device__ void flush(const SsArgs& a, Acc& acc, int lv, unsigned* qlen, int32_t* q_out) {
    acc.reduce(tot);
    q.flush(qlen, q_out);  // synchronises the workgroup first
    if (threadIdx.x < 3 && tot[threadIdx.x]) {
      unsigned long long* stripe =
          a.part + ((size_t)(lv % kRing) * kStripes + blockIdx.x % kStripes) * kStripeU64;
      atomicAdd(&stripe[threadIdx.x], tot[threadIdx.x]);
    }
  }
};

// Records a newly discovered vertex of the level: staged queue slot and
// counters.  Called by every lane of the wave.
__device__ __forceinline__ void emit(bool disc, int32_t v, SsStage& sg, unsigned* qlen,
                                     int32_t* q_out, Acc& acc, const int32_t* row_ptr) {
  sg.q.append(disc, v, qlen, q_out);
  if (disc) acc.add(row_ptr, v);
}

// Clears the counter stripes of level lv (grid-stride).
__device__ __forceinline__ void clear_stripes(const SsArgs& a, int lv, int64_t gt, int64_t gs) {
  unsigned long long* p = a.part + (size_t)(lv % kRing) * kStripes * kStripeU64;
  for (int64_t i = gt; i < kStripes * kStripeU64; i += gs) p[i] = 0ull;
}

// Level counters of a seed kernel: stripe 0 of the level (the others are
// zero after the fill kernel) and the queue length in the ring.
__device__ __forceinline__ void seed_counters(const SsArgs& a, int lv, const LevelCounters& c) {
  a.ring[lv % kRing] = c;
  unsigned long long* p = a.part + (size_t)(lv % kRing) * kStripes * kStripeU64;
  p[0] = c.fv;
  p[1] = c.fe;
  p[2] = c.scout;
}

// ---------------------------------------------------------------- kernels ---

__global__ void ss_fill_kernel(SsArgs a, int weighted) {
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = t0; i < a.n; i += gs) a.dist[i] = kInf;
  for (int64_t i = t0; i < 3 * a.words; i += gs) a.bm[i] = 0u;
  if (t0 < kRing) a.ring[t0] = LevelCounters{};
  for (int64_t i = t0; i < kPartU64; i += gs) a.part[i] = 0ull;
  (void)weighted;
}

// Level 0 = {source}.  GAPBS starts scout_count at out_degree(source)
// (test/gabps/bfs.cc:140).
__global__ void bfs_seed_kernel(SsArgs a, int32_t policy) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int32_t s = a.source;
  const int d = a.row_ptr[s + 1] - a.row_ptr[s];
  a.dist[s] = 0;
  a.bm[s >> 5] |= 1u << (s & 31);
  a.q[0] = s;
  LevelCounters c{};
  c.fv = 1;
  c.fe = (unsigned)d;
  c.scout = (unsigned)d;
  c.flanes = 1;
  c.qlen = 1;
  seed_counters(a, 0, c);
  SchedState st{};
  st.step = 0;
  st.e2c = a.row_ptr[a.n];
  st.scout = d;
  st.force_mode = -1;
  st.source = s;
  st.last_mode = DAWN_MODE_PUSH;
  *a.st = st;
  (void)policy;
}

// Level 1 for min-plus, as SSSP_CPU::SSSP initialises it
// (src/algorithm/cpu/sssp.cpp:40-46): every listed out-edge of the source
// writes its weight in row order (a later duplicate overwrites an earlier
// one), the distinct neighbours form the first frontier, distance[source]
// = 0, and the source is never relaxed afterwards.
__global__ void sssp_seed_kernel(SsArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int32_t s = a.source;
  const int32_t b = a.row_ptr[s], e = a.row_ptr[s + 1];
  for (int32_t k = b; k < e; k++) a.dist[a.col[k]] = a.w[k];
  a.dist[s] = 0;
  uint32_t* bm1 = a.bm + a.words;  // level 1 slot
  int32_t* q1 = a.q + a.n;
  LevelCounters c{};
  for (int32_t k = b; k < e; k++) {
    int32_t v = a.col[k];
    if (v == s) continue;
    uint32_t bit = 1u << (v & 31);
    if (bm1[v >> 5] & bit) continue;
    bm1[v >> 5] |= bit;
    q1[c.qlen++] = v;
    int d = a.row_ptr[v + 1] - a.row_ptr[v];
    c.fv += 1;
    c.fe += (unsigned)d;
    c.scout += (unsigned)(d > 0 ? d : 1);
  }
  c.flanes = c.fv;
  seed_counters(a, 1, c);
  SchedState st{};
  st.step = 1;
  st.e2c = a.row_ptr[a.n];
  st.scout = (long long)c.scout;
  st.force_mode = -1;
  st.source = s;
  st.last_mode = DAWN_MODE_PUSH;
  *a.st = st;
}

// One wave: sums the counter stripes of the last completed level into its
// ring entry (an assignment, so running it twice is harmless), then decides.
__global__ void ss_decide_kernel(SsArgs a, SchedParams p, SchedTrace tr) {
  const SchedState s = *a.st;
  if (!s.done) {
    const int slot = s.step % kRing;
    const unsigned long long* stripe =
        a.part + ((size_t)slot * kStripes + lane_id()) * kStripeU64;
    const unsigned long long fv = wave_sum_u64(stripe[0]);
    const unsigned long long fe = wave_sum_u64(stripe[1]);
    const unsigned long long sc = wave_sum_u64(stripe[2]);
    if (threadIdx.x == 0) {
      LevelCounters* c = a.ring + slot;
      c->fv = fv;
      c->fe = fe;
      c->scout = sc;
      c->flanes = fv;
    }
  }
  sched_decide(a.st, a.ring, p, tr);
  if (threadIdx.x == 0) {
    SchedState* sp = a.st;
    if (!sp->done && sp->mode == DAWN_MODE_PUSH &&
        a.ring[(sp->step + kRing - 1) % kRing].qlen == kNoQueue) {
      // The previous level has no queue: the same top-down step from its
      // bitmap.
      sp->mode = DAWN_MODE_BITMAP;
      sp->last_mode = DAWN_MODE_BITMAP;
      if (sp->step < kTraceMax && tr.mode) tr.mode[sp->step] = DAWN_MODE_BITMAP;
    }
  }
}

// One BOVM step: discovers level lv = st->step.
__global__ void __launch_bounds__(kBlock) bfs_expand_kernel(SsArgs a) {
  __shared__ int s_off[kBlock];
  __shared__ int s_beg[kBlock];
  __shared__ int s_scan[kBlock / kWave];
  const SchedState st = *a.st;
  if (st.done) return;
  const int lv = st.step;
  const int pi = (lv + 2) % 3, ci = lv % 3, ni = (lv + 1) % 3;
  const uint32_t* bm_in = a.bm + pi * a.words;
  uint32_t* bm_out = a.bm + ci * a.words;
  uint32_t* bm_nx = a.bm + ni * a.words;
  const int32_t* q_in = a.q + (int64_t)pi * a.n;
  int32_t* q_out = a.q + (int64_t)ci * a.n;
  LevelCounters* c_out = a.ring + (lv % kRing);
  unsigned* qlen = &c_out->qlen;
  const int lane = lane_id();
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = gt; i < a.words; i += gs) bm_nx[i] = 0u;
  if (gt == 0) a.ring[(lv + 1) % kRing] = LevelCounters{};
  clear_stripes(a, lv + 1, gt, gs);
  __shared__ SsStage sg;
  sg.init();
  Acc acc;

  if (st.mode == DAWN_MODE_PUSH) {
    const unsigned qn = a.ring[(lv + kRing - 1) % kRing].qlen;
    for (int64_t c0 = (int64_t)blockIdx.x * kBlock; c0 < qn; c0 += (int64_t)gridDim.x * kBlock) {
      const int64_t i = c0 + threadIdx.x;
      int beg = 0, deg = 0;
      if (i < qn) {
        int u = q_in[i];
        beg = a.row_ptr[u];
        deg = a.row_ptr[u + 1] - beg;
      }
      int total;
      int off = block_exclusive_scan(deg, s_scan, &total);
      s_off[threadIdx.x] = off;
      s_beg[threadIdx.x] = beg;
      __syncthreads();
      for (int e0 = 0; e0 < total; e0 += kBlock) {
        const int e = e0 + threadIdx.x;
        bool disc = false;
        int32_t v = 0;
        if (e < total) {
          int j = owner_of(s_off, e);
          v = a.col[s_beg[j] + (e - s_off[j])];
          if (a.dist[v] == kInf && atomicCAS(&a.dist[v], kInf, lv) == kInf) {
            disc = true;
            atomicOr(&bm_out[v >> 5], 1u << (v & 31));
          }
        }
        emit(disc, v, sg, qlen, q_out, acc, a.row_ptr);
      }
      __syncthreads();
    }
  } else if (st.mode == DAWN_MODE_PULL) {
    if (gt == 0) *qlen = kNoQueue;  // the level's bitmap only
    const int64_t wstride = (int64_t)gridDim.x * kBlock;
    for (int64_t base = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < a.n;
         base += wstride) {
      const int32_t v = (int32_t)(base + lane);
      bool cand = v < a.n && a.dist[v] == kInf;
      int beg = 0, deg = 0;
      if (cand) {
        beg = a.in_row_ptr[v];
        deg = a.in_row_ptr[v + 1] - beg;
      }
      const bool light = cand && deg < kHeavyDegree;
      const int maxd = wave_max_i32(light ? deg : 0);
      bool found = false;
      for (int j = 0; j < maxd; j++) {
        if (light && !found && j < deg) found = test_bit(bm_in, a.in_col[beg + j]);
      }
      uint64_t hm = __ballot(cand && deg >= kHeavyDegree);
      while (hm) {
        const int l = __ffsll((unsigned long long)hm) - 1;
        hm &= hm - 1;
        const int hb = __shfl(beg, l), he = hb + __shfl(deg, l);
        bool f = false;
        for (int k0 = hb; k0 < he; k0 += kWave) {
          const int k = k0 + lane;
          const bool b = k < he && test_bit(bm_in, a.in_col[k]);
          if (__ballot(b)) { f = true; break; }
        }
        if (lane == l) found = f;
      }
      if (found) a.dist[v] = lv;
      const uint64_t fm = __ballot(found);
      if (lane == 0) {
        bm_out[base >> 5] = (uint32_t)fm;
        bm_out[(base >> 5) + 1] = (uint32_t)(fm >> 32);
      }
      if (found) acc.add(a.row_ptr, v);  // bitmap only: no queue entries
    }
  } else {  // DAWN_MODE_BITMAP: dense top-down over the frontier bitmap
    const int64_t wstride = (int64_t)gridDim.x * kBlock;
    for (int64_t base = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < a.n;
         base += wstride) {
      const int32_t u = (int32_t)(base + lane);
      const bool act = u < a.n && test_bit(bm_in, u);
      if (__ballot(act) == 0) continue;
      int beg = 0, deg = 0;
      if (act) {
        beg = a.row_ptr[u];
        deg = a.row_ptr[u + 1] - beg;
      }
      const bool light = act && deg < kHeavyDegree;
      const int maxd = wave_max_i32(light ? deg : 0);
      for (int j = 0; j < maxd; j++) {
        bool disc = false;
        int32_t v = 0;
        if (light && j < deg) {
          v = a.col[beg + j];
          if (a.dist[v] == kInf && atomicCAS(&a.dist[v], kInf, lv) == kInf) {
            disc = true;
            atomicOr(&bm_out[v >> 5], 1u << (v & 31));
          }
        }
        emit(disc, v, sg, qlen, q_out, acc, a.row_ptr);
      }
      uint64_t hm = __ballot(act && deg >= kHeavyDegree);
      while (hm) {
        const int l = __ffsll((unsigned long long)hm) - 1;
        hm &= hm - 1;
        const int hb = __shfl(beg, l), he = hb + __shfl(deg, l);
        for (int k0 = hb; k0 < he; k0 += kWave) {
          const int k = k0 + lane;
          bool disc = false;
          int32_t v = 0;
          if (k < he) {
            v = a.col[k];
            if (a.dist[v] == kInf && atomicCAS(&a.dist[v], kInf, lv) == kInf) {
              disc = true;
              atomicOr(&bm_out[v >> 5], 1u << (v & 31));
            }
          }
          emit(disc, v, sg, qlen, q_out, acc, a.row_ptr);
        }
      }
    }
  }
  sg.flush(a, acc, lv, qlen, q_out);
}

// Relaxation of edge (u -> v, w) with du = current distance of u.  Returns
// true when v joins the next frontier for the first time this level.
__device__ __forceinline__ bool relax(const SsArgs& a, int32_t src, int32_t du, int32_t v,
                                      int32_t w, uint32_t* bm_out) {
  if (v == src) return false;
  const long long nd = (long long)du + (long long)w;
  if (nd >= kInf) return false;
  const int32_t ndi = (int32_t)nd;
  if (ndi >= a.dist[v]) return false;
  const int32_t old = atomicMin(&a.dist[v], ndi);
  if (ndi >= old) return false;
  const uint32_t bit = 1u << (v & 31);
  const uint32_t prev = atomicOr(&bm_out[v >> 5], bit);
  return (prev & bit) == 0;
}

// One min-plus step (GOVM/GOVMP semantics): relaxes the out-edges of the
// vertices whose distance dropped in the previous step.
__global__ void __launch_bounds__(kBlock) sssp_expand_kernel(SsArgs a) {
  __shared__ int s_off[kBlock];
  __shared__ int s_beg[kBlock];
  __shared__ int s_du[kBlock];
  __shared__ int s_scan[kBlock / kWave];
  const SchedState st = *a.st;
  if (st.done) return;
  const int32_t src = st.source;
  const int lv = st.step;
  const int pi = (lv + 2) % 3, ci = lv % 3, ni = (lv + 1) % 3;
  const uint32_t* bm_in = a.bm + pi * a.words;
  uint32_t* bm_out = a.bm + ci * a.words;
  uint32_t* bm_nx = a.bm + ni * a.words;
  const int32_t* q_in = a.q + (int64_t)pi * a.n;
  int32_t* q_out = a.q + (int64_t)ci * a.n;
  LevelCounters* c_out = a.ring + (lv % kRing);
  unsigned* qlen = &c_out->qlen;
  const int lane = lane_id();
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = gt; i < a.words; i += gs) bm_nx[i] = 0u;
  if (gt == 0) a.ring[(lv + 1) % kRing] = LevelCounters{};
  clear_stripes(a, lv + 1, gt, gs);
  __shared__ SsStage sg;
  sg.init();
  Acc acc;

  if (st.mode == DAWN_MODE_PUSH) {
    const unsigned qn = a.ring[(lv + kRing - 1) % kRing].qlen;
    for (int64_t c0 = (int64_t)blockIdx.x * kBlock; c0 < qn; c0 += (int64_t)gridDim.x * kBlock) {
      const int64_t i = c0 + threadIdx.x;
      int beg = 0, deg = 0, du = 0;
      if (i < qn) {
        int u = q_in[i];
        beg = a.row_ptr[u];
        deg = a.row_ptr[u + 1] - beg;
        du = a.dist[u];
      }
      int total;
      int off = block_exclusive_scan(deg, s_scan, &total);
      s_off[threadIdx.x] = off;
      s_beg[threadIdx.x] = beg;
      s_du[threadIdx.x] = du;
      __syncthreads();
      for (int e0 = 0; e0 < total; e0 += kBlock) {
        const int e = e0 + threadIdx.x;
        bool disc = false;
        int32_t v = 0;
        if (e < total) {
          int j = owner_of(s_off, e);
          int k = s_beg[j] + (e - s_off[j]);
          v = a.col[k];
          disc = relax(a, src, s_du[j], v, a.w[k], bm_out);
        }
        emit(disc, v, sg, qlen, q_out, acc, a.row_ptr);
      }
      __syncthreads();
    }
  } else if (st.mode == DAWN_MODE_PULL) {
    // Every vertex takes the minimum over its in-edges from frontier
    // vertices; each vertex is written only by its own lane.  Bitmap only:
    // no queue entries.
    if (gt == 0) *qlen = kNoQueue;
    const int64_t wstride = (int64_t)gridDim.x * kBlock;
    for (int64_t base = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < a.n;
         base += wstride) {
      const int32_t v = (int32_t)(base + lane);
      const bool cand = v < a.n && v != src;
      int beg = 0, deg = 0;
      int32_t cur = kInf;
      if (cand) {
        beg = a.in_row_ptr[v];
        deg = a.in_row_ptr[v + 1] - beg;
        cur = a.dist[v];
      }
      int32_t best = cur;
      const bool light = cand && deg < kHeavyDegree;
      const int maxd = wave_max_i32(light ? deg : 0);
      for (int j = 0; j < maxd; j++) {
        if (light && j < deg) {
          const int32_t u = a.in_col[beg + j];
          if (test_bit(bm_in, u)) {
            long long nd = (long long)a.dist[u] + a.in_w[beg + j];
            if (nd < best) best = (int32_t)nd;
          }
        }
      }
      uint64_t hm = __ballot(cand && deg >= kHeavyDegree);
      while (hm) {
        const int l = __ffsll((unsigned long long)hm) - 1;
        hm &= hm - 1;
        const int hb = __shfl(beg, l), he = hb + __shfl(deg, l);
        int32_t hbest = kInf;
        for (int k = hb + lane; k < he; k += kWave) {
          const int32_t u = a.in_col[k];
          if (test_bit(bm_in, u)) {
            long long nd = (long long)a.dist[u] + a.in_w[k];
            if (nd < hbest) hbest = (int32_t)nd;
          }
        }
        for (int o = 32; o >= 1; o >>= 1) hbest = min(hbest, __shfl_xor(hbest, o));
        if (lane == l && hbest < best) best = hbest;
      }
      const bool found = cand && best < cur;
      if (found) a.dist[v] = best;
      const uint64_t fm = __ballot(found);
      if (lane == 0) {
        bm_out[base >> 5] = (uint32_t)fm;
        bm_out[(base >> 5) + 1] = (uint32_t)(fm >> 32);
      }
      if (found) acc.add(a.row_ptr, v);  // bitmap only: no queue entries
    }
  } else {  // DAWN_MODE_BITMAP (GOVMP shape: dense scan of the frontier)
    const int64_t wstride = (int64_t)gridDim.x * kBlock;
    for (int64_t base = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < a.n;
         base += wstride) {
      const int32_t u = (int32_t)(base + lane);
      const bool act = u < a.n && test_bit(bm_in, u);
      if (__ballot(act) == 0) continue;
      int beg = 0, deg = 0;
      int32_t du = 0;
      if (act) {
        beg = a.row_ptr[u];
        deg = a.row_ptr[u + 1] - beg;
        du = a.dist[u];
      }
      const bool light = act && deg < kHeavyDegree;
      const int maxd = wave_max_i32(light ? deg : 0);
      for (int j = 0; j < maxd; j++) {
        bool disc = false;
        int32_t v = 0;
        if (light && j < deg) {
          v = a.col[beg + j];
          disc = relax(a, src, du, v, a.w[beg + j], bm_out);
        }
        emit(disc, v, sg, qlen, q_out, acc, a.row_ptr);
      }
      uint64_t hm = __ballot(act && deg >= kHeavyDegree);
      while (hm) {
        const int l = __ffsll((unsigned long long)hm) - 1;
        hm &= hm - 1;
        const int hb = __shfl(beg, l), he = hb + __shfl(deg, l);
        const int32_t hdu = __shfl(du, l);
        for (int k0 = hb; k0 < he; k0 += kWave) {
          const int k = k0 + lane;
          bool disc = false;
          int32_t v = 0;
          if (k < he) {
            v = a.col[k];
            disc = relax(a, src, hdu, v, a.w[k], bm_out);
          }
          emit(disc, v, sg, qlen, q_out, acc, a.row_ptr);
        }
      }
    }
  }
  sg.flush(a, acc, lv, qlen, q_out);
}

// Sum of out-degree and count of reached vertices (GTEPS numerator):
// wave sums, then one pair of atomics per workgroup (launched with at most
// kReachBlocks workgroups: the two counters are single addresses).
constexpr unsigned kReachBlocks = 512;
__global__ void reach_kernel(const int32_t* dist, const int32_t* row_ptr, int32_t n,
                             unsigned long long* out) {
  __shared__ unsigned long long s_t[2];
  if (threadIdx.x == 0) s_t[0] = s_t[1] = 0ull;
  __syncthreads();
  unsigned long long cnt = 0, edges = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (dist[v] != kInf) {
      cnt++;
      edges += (unsigned)(row_ptr[v + 1] - row_ptr[v]);
    }
  }
  cnt = wave_sum_u64(cnt);
  edges = wave_sum_u64(edges);
  if (lane_id() == 0 && cnt) {
    atomicAdd(&s_t[0], cnt);
    atomicAdd(&s_t[1], edges);
  }
  __syncthreads();
  if (threadIdx.x < 2 && s_t[threadIdx.x]) atomicAdd(&out[threadIdx.x], s_t[threadIdx.x]);
}

int ensure_ss(dawn_graph* g) {
  if (!g->ss) g->ss = new SsWorkspace();
  SsWorkspace* ws = g->ss;
  if (ws->n == g->n && ws->dist.p) return DAWN_OK;
  ws->n = g->n;
  ws->words = ((int64_t)g->n + 63) / 64 * 2;
  DAWN_TRY(ws->dist.alloc(g->n));
  DAWN_TRY(ws->q.alloc((size_t)3 * g->n));
  DAWN_TRY(ws->bm.alloc((size_t)3 * ws->words));
  DAWN_TRY(ws->ring.alloc(kRing));
  DAWN_TRY(ws->part.alloc(kPartU64));
  DAWN_TRY(ws->st.alloc(1));
  DAWN_TRY(ws->trace_feat.alloc((size_t)kTraceMax * 8));
  DAWN_TRY(ws->trace_mode.alloc(kTraceMax));
  DAWN_TRY(ws->reach.alloc(2));
  if (!ws->h_state) DAWN_HIP(hipHostMalloc((void**)&ws->h_state, sizeof(SchedState)));
  if (!ws->ev0) DAWN_HIP(hipEventCreate(&ws->ev0));
  if (!ws->ev1) DAWN_HIP(hipEventCreate(&ws->ev1));
  return DAWN_OK;
}

}  // namespace

// Runs one single-source traversal.  kind 0: BOVM, 1: min-plus.
// steps_per_chunk: decide+expand pairs per captured graph.
int run_single_source(dawn_graph* g, int kind, int32_t source, int32_t* dist_out,
                      int32_t policy, dawn_stats* stats, int32_t force_mode,
                      float* trace_feat, int32_t* trace_mode, int32_t trace_max,
                      int32_t* trace_rows) {
  if (!g || source < 0 || source >= g->n) return DAWN_ERR_INVALID_ARG;
  if (kind == 1 && !g->weighted) return DAWN_ERR_NOT_WEIGHTED;
  if (g->h_row_ptr[source] == g->h_row_ptr[source + 1]) return DAWN_ERR_ISOLATED_SOURCE;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  DAWN_TRY(ensure_ss(g));
  SsWorkspace* ws = g->ss;
  hipStream_t s = g->stream;

  SsArgs a{};
  a.n = g->n;
  a.source = source;
  a.words = ws->words;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.w = g->weight;
  a.in_row_ptr = g->in_row_ptr;
  a.in_col = g->in_col;
  a.in_w = g->in_weight;
  a.dist = ws->dist.p;
  a.q = ws->q.p;
  a.bm = ws->bm.p;
  a.ring = ws->ring.p;
  a.part = ws->part.p;
  a.st = ws->st.p;

  SchedParams p{};
  p.n = g->n;
  p.m = g->m;
  p.policy = policy;
  p.alpha = 15;
  p.beta = 18;
  p.batched = 0;
  p.allow_pull = 1;
  p.avg_degree = g->n ? (float)((double)g->m / g->n) : 0.f;
  p.max_degree = (float)g->max_out_degree;
  // The single-source network (trained on single-source steps); without
  // one the learned policy takes the fixed rule.  Min-plus: the fixed rule.
  p.mlp = kind == 0 ? sched_single_mlp_device(g->device, &p.mlp_norm) : nullptr;
  if (kind == 1 && policy == DAWN_POLICY_LEARNED) p.policy = DAWN_POLICY_FIXED;
  SchedTrace tr{ws->trace_feat.p, ws->trace_mode.p};

  const unsigned grid = grid_for(std::max<int64_t>(g->n, 1), kBlock, 4096);
  // A chunk = K decide+expand pairs, captured once per (kind, policy, mode,
  // K).  Without a knob K is what the last traversal on this graph took
  // (levels + 1, rounded up to even), so that one chunk and one read of the
  // done flag usually suffice (BFS on ER n=65,536, 9 levels: 0.20 ms with
  // K = 8, 0.156 ms with K = 10).
  int K = g->knobs.ss_chunk > 0 ? g->knobs.ss_chunk : (g->n < (1 << 20) ? 8 : 16);
  if (g->knobs.ss_chunk <= 0 && ws->last_pairs[kind] > 0)
    K = std::min(128, (ws->last_pairs[kind] + 1) & ~1);
  // Launch arguments are baked into the graph: key it by every one that
  // varies between calls (the network pointer changes when one is installed).
  const std::pair<long long, const void*> key{
      ((long long)kind << 40) | ((long long)(policy & 0xff) << 32) |
          ((long long)(force_mode + 1) << 16) | K,
      (const void*)p.mlp};
  hipGraphExec_t exec = nullptr;
  auto it = ws->graphs.find(key);
  if (it != ws->graphs.end()) {
    exec = it->second;
  } else {
    hipGraph_t graph;
    DAWN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < K; i++) {
      ss_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
      if (kind == 0)
        bfs_expand_kernel<<<grid, kBlock, 0, s>>>(a);
      else
        sssp_expand_kernel<<<grid, kBlock, 0, s>>>(a);
    }
    DAWN_HIP(hipStreamEndCapture(s, &graph));
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DAWN_HIP(e);
    ws->graphs[key] = exec;
  }

  DAWN_HIP(hipEventRecord(ws->ev0, s));
  ss_fill_kernel<<<grid_for(std::max<int64_t>(3 * ws->words, g->n)), kBlock, 0, s>>>(a, kind);
  if (kind == 0)
    bfs_seed_kernel<<<1, 64, 0, s>>>(a, policy);
  else
    sssp_seed_kernel<<<1, 64, 0, s>>>(a);
  DAWN_HIP(hipGetLastError());
  if (force_mode >= 0) {
    // The seed kernel leaves force_mode = -1; set it before stepping.
    const int32_t fm = force_mode;
    DAWN_HIP(hipMemcpyAsync(&ws->st.p->force_mode, &fm, sizeof(int32_t),
                            hipMemcpyHostToDevice, s));
    DAWN_HIP(hipStreamSynchronize(s));
  }
  int64_t launched = 0;
  const int64_t limit = (int64_t)g->n + 2;
  for (;;) {
    DAWN_HIP(hipGraphLaunch(exec, s));
    launched += K;
    DAWN_HIP(hipMemcpyAsync(ws->h_state, ws->st.p, sizeof(SchedState),
                            hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    if (ws->h_state->done || launched > limit) break;
  }
  DAWN_HIP(hipEventRecord(ws->ev1, s));
  DAWN_HIP(hipEventSynchronize(ws->ev1));
  float ms = 0.f;
  DAWN_HIP(hipEventElapsedTime(&ms, ws->ev0, ws->ev1));
  if (!ws->h_state->done) return DAWN_ERR_HIP;  // step bound exceeded

  const int levels = ws->h_state->levels;
  const int first_step = kind == 0 ? 1 : 2;
  // Decides that ran: one per expanded level after the seeded ones, plus
  // the one that found the frontier empty.
  ws->last_pairs[kind] = std::max(1, levels - first_step + 2);
  if (stats) {
    DAWN_HIP(hipMemsetAsync(ws->reach.p, 0, 2 * sizeof(unsigned long long), s));
    reach_kernel<<<grid_for(g->n, kBlock, kReachBlocks), kBlock, 0, s>>>(ws->dist.p, g->row_ptr, g->n, ws->reach.p);
    unsigned long long rr[2];
    DAWN_HIP(hipMemcpyAsync(rr, ws->reach.p, sizeof rr, hipMemcpyDeviceToHost, s));
    // Trace entries 1..levels hold the modes of the expanding steps.
    std::vector<int32_t> modes(levels + 1);
    const int nrec = std::min(levels + 1, kTraceMax);
    if (nrec > 0)
      DAWN_HIP(hipMemcpyAsync(modes.data(), ws->trace_mode.p, sizeof(int32_t) * nrec,
                              hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    stats->device_ms += ms;
    stats->vertices_reached += (int64_t)rr[0];
    stats->edges_reached += (int64_t)rr[1];
    stats->batches += 1;
    stats->frontier_edges += (int64_t)ws->h_state->work_fe;
    // Steps first_step..levels expanded a level; the final step found none.
    for (int st = first_step; st <= levels && st < kTraceMax; st++) {
      stats->steps += 1;
      int md = modes[st];
      if (md >= 0 && md < DAWN_NUM_MODES) stats->mode_steps[md] += 1;
    }
  }
  if (trace_rows) {
    int rows = 0;
    int nrec = std::min(levels + 1, kTraceMax);
    std::vector<int32_t> modes(nrec);
    std::vector<float> feat((size_t)nrec * 8);
    DAWN_HIP(hipMemcpyAsync(modes.data(), ws->trace_mode.p, sizeof(int32_t) * nrec,
                            hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipMemcpyAsync(feat.data(), ws->trace_feat.p, sizeof(float) * 8 * nrec,
                            hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    for (int st = first_step; st <= levels && st < nrec && rows < trace_max; st++, rows++) {
      if (trace_mode) trace_mode[rows] = modes[st];
      if (trace_feat)
        for (int i = 0; i < 8; i++) trace_feat[rows * 8 + i] = feat[(size_t)st * 8 + i];
    }
    *trace_rows = rows;
  }
  if (dist_out) {
    const bool dev = is_device_pointer(dist_out);
    DAWN_HIP(hipMemcpyAsync(dist_out, ws->dist.p, sizeof(int32_t) * g->n,
                            dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
  }
  return DAWN_OK;
}

// ---- per-step mode timings (scheduler training data) ------------------------
//
// Every mode discovers the same levels (the frontier of level l is the set
// of vertices at hop distance l, whatever the modes of the steps before), so
// a step starts from the same state in every run: one BOVM traversal per
// mode with that mode forced, each step launched on its own with HIP events
// around the expand kernel, gives the time of every mode at every step
// (minimum over `repeats` traversals); one traversal under `policy` gives
// the step features and the policy's choices.

namespace {

// One traversal from a.source, step by step.  force < 0: the policy's
// choices, else that mode at every step.  times: expand-kernel time of each
// step (steps 1..levels).
int ss_stepwise(dawn_graph* g, SsWorkspace* ws, const SsArgs& a, const SchedParams& p,
                SchedTrace tr, int force, unsigned grid, std::vector<float>* times,
                int* levels) {
  hipStream_t s = g->stream;
  ss_fill_kernel<<<grid_for(std::max<int64_t>(3 * ws->words, g->n)), kBlock, 0, s>>>(a, 0);
  bfs_seed_kernel<<<1, 64, 0, s>>>(a, p.policy);
  DAWN_HIP(hipGetLastError());
  if (force >= 0) {
    const int32_t fm = force;
    DAWN_HIP(hipMemcpyAsync(&ws->st.p->force_mode, &fm, sizeof(int32_t), hipMemcpyHostToDevice, s));
  }
  times->clear();
  for (int64_t k = 0; k <= (int64_t)g->n + 1; k++) {
    ss_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
    DAWN_HIP(hipMemcpyAsync(ws->h_state, ws->st.p, sizeof(SchedState), hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    if (ws->h_state->done) {
      *levels = ws->h_state->levels;
      return DAWN_OK;
    }
    DAWN_HIP(hipEventRecord(ws->ev0, s));
    bfs_expand_kernel<<<grid, kBlock, 0, s>>>(a);
    DAWN_HIP(hipGetLastError());
    DAWN_HIP(hipEventRecord(ws->ev1, s));
    DAWN_HIP(hipEventSynchronize(ws->ev1));
    float ms = 0.f;
    DAWN_HIP(hipEventElapsedTime(&ms, ws->ev0, ws->ev1));
    times->push_back(ms);
  }
  return DAWN_ERR_HIP;  // step bound exceeded
}

}  // namespace

int run_ss_record(dawn_graph* g, const int32_t* sources, int32_t S, int32_t repeats,
                  float* features, float* times_ms, int32_t* chosen, int32_t max_rows,
                  int32_t* rows_out, int32_t policy) {
  if (!g || S < 0 || (S > 0 && !sources) || repeats < 1 || !rows_out || max_rows < 0)
    return DAWN_ERR_INVALID_ARG;
  if (g->weighted) return DAWN_ERR_UNSUPPORTED;
  *rows_out = 0;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  DAWN_TRY(ensure_ss(g));
  SsWorkspace* ws = g->ss;
  const unsigned grid = grid_for(std::max<int64_t>(g->n, 1), kBlock, 4096);
  int32_t rows = 0;
  std::vector<float> t, feat;
  std::vector<int32_t> modes;
  for (int32_t i = 0; i < S && rows < max_rows; i++) {
    int64_t src = sources[i] % (int64_t)g->n;
    if (src < 0) src += g->n;
    if (g->h_row_ptr[src] == g->h_row_ptr[src + 1]) continue;  // isolated: no traversal
    SsArgs a{};
    a.n = g->n;
    a.source = (int32_t)src;
    a.words = ws->words;
    a.row_ptr = g->row_ptr;
    a.col = g->col;
    a.w = g->weight;
    a.in_row_ptr = g->in_row_ptr;
    a.in_col = g->in_col;
    a.in_w = g->in_weight;
    a.dist = ws->dist.p;
    a.q = ws->q.p;
    a.bm = ws->bm.p;
    a.ring = ws->ring.p;
    a.part = ws->part.p;
    a.st = ws->st.p;
    SchedParams p{};
    p.n = g->n;
    p.m = g->m;
    p.policy = policy;
    p.alpha = 15;
    p.beta = 18;
    p.batched = 0;
    p.allow_pull = 1;
    p.avg_degree = g->n ? (float)((double)g->m / g->n) : 0.f;
    p.max_degree = (float)g->max_out_degree;
    p.mlp = sched_single_mlp_device(g->device, &p.mlp_norm);
    SchedTrace tr{ws->trace_feat.p, ws->trace_mode.p};
    // The policy's run: features and choices of steps 1..levels.
    int levels = 0;
    DAWN_TRY(ss_stepwise(g, ws, a, p, tr, -1, grid, &t, &levels));
    const int nrec = std::min(levels + 1, kTraceMax);
    modes.assign(nrec, 0);
    feat.assign((size_t)nrec * 8, 0.f);
    DAWN_HIP(hipMemcpy(modes.data(), ws->trace_mode.p, sizeof(int32_t) * nrec, hipMemcpyDeviceToHost));
    DAWN_HIP(hipMemcpy(feat.data(), ws->trace_feat.p, sizeof(float) * 8 * nrec, hipMemcpyDeviceToHost));
    // Every mode forced, minimum over the repeats.
    std::vector<float> best((size_t)levels * DAWN_NUM_MODES, 1e30f);
    for (int m = 0; m < DAWN_NUM_MODES; m++)
      for (int r = 0; r < repeats; r++) {
        int lv = 0;
        DAWN_TRY(ss_stepwise(g, ws, a, p, tr, m, grid, &t, &lv));
        if (lv != levels || (int)t.size() != levels) return DAWN_ERR_HIP;  // modes must agree
        for (int k = 0; k < levels; k++)
          best[(size_t)k * DAWN_NUM_MODES + m] = std::min(best[(size_t)k * DAWN_NUM_MODES + m], t[k]);
      }
    for (int k = 1; k <= levels && k < nrec && rows < max_rows; k++, rows++) {
      for (int j = 0; j < 8; j++) features[(size_t)rows * 8 + j] = feat[(size_t)k * 8 + j];
      for (int m = 0; m < DAWN_NUM_MODES; m++)
        times_ms[(size_t)rows * DAWN_NUM_MODES + m] = best[(size_t)(k - 1) * DAWN_NUM_MODES + m];
      if (chosen) chosen[rows] = modes[k];
    }
  }
  *rows_out = rows;
  return DAWN_OK;
}

// ---- allocation-free BOVM traversal (dawn_bfs_noalloc) --------------------
//
// For callers that cannot allocate device memory (serving loops, harnesses
// that forbid hipMalloc): the single-source BOVM traversal on a fixed
// scratch region instead of the graph's workspace.  The region is the
// file-scope device array `scratch` below, sized for graphs of up to
// kNoAllocMaxN vertices (66 MB); the code object carries it, so each GPU
// the library runs on holds its own copy until the process ends.  Host code
// never passes the array's name to a kernel (in host code it is a stand-in,
// not a device address): once per device, with that device current, the
// first call takes its device address from hipGetSymbolAddress and keeps it
// per device; every kernel argument and copy uses that address.
//
// Every call re-initialises what it reads (fill and seed kernels, a zeroing
// kernel for the reach counters).  Calls on one GPU share the array, so the
// calls of a process are serialised per device by a lock (two calls that
// overlapped on one GPU would race).  The step loop is launched directly,
// K decide+expand pairs between reads of the done flag: instantiating a
// captured graph would allocate.  The single-source scheduler network is
// used only if it is already on the device (installing it allocates);
// otherwise the learned policy falls back to the fixed rule, as when no
// network is set.

constexpr int32_t kNoAllocMaxN = 1 << 22;
constexpr int kNoAllocMaxDevices = 64;

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr int64_t kNoAllocWords = ((int64_t)kNoAllocMaxN + 63) / 64 * 2;
constexpr size_t kNaDist = 0;
constexpr size_t kNaQ = align256(kNaDist + sizeof(int32_t) * (size_t)kNoAllocMaxN);
constexpr size_t kNaBm = align256(kNaQ + sizeof(int32_t) * 3 * (size_t)kNoAllocMaxN);
constexpr size_t kNaRing = align256(kNaBm + sizeof(uint32_t) * 3 * (size_t)kNoAllocWords);
constexpr size_t kNaPart = align256(kNaRing + sizeof(LevelCounters) * kRing);
constexpr size_t kNaState = align256(kNaPart + sizeof(unsigned long long) * kPartU64);
constexpr size_t kNaFeat = align256(kNaState + sizeof(SchedState));
constexpr size_t kNaMode = align256(kNaFeat + sizeof(float) * 8 * (size_t)kTraceMax);
constexpr size_t kNaReach = align256(kNaMode + sizeof(int32_t) * (size_t)kTraceMax);
constexpr size_t kNoAllocBytes = align256(kNaReach + 2 * sizeof(unsigned long long));

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_noalloc_lock[kNoAllocMaxDevices];
unsigned char* g_noalloc_base[kNoAllocMaxDevices] = {};  // device address of `scratch`

__global__ void zero_u64_kernel(unsigned long long* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0ull;
}

}  // namespace

int run_bfs_noalloc(dawn_graph* g, int32_t source, int32_t* dist_out, int32_t policy,
                    dawn_stats* stats) {
  if (!g || source < 0 || source >= g->n) return DAWN_ERR_INVALID_ARG;
  if (g->n > kNoAllocMaxN) {
    set_error("dawn_bfs_noalloc: the graph has more vertices than the scratch region holds "
              "(at most 4194304)");
    return DAWN_ERR_SCRATCH_LIMIT;
  }
  if (g->device < 0 || g->device >= kNoAllocMaxDevices) return DAWN_ERR_UNSUPPORTED;
  if (g->h_row_ptr[source] == g->h_row_ptr[source + 1]) return DAWN_ERR_ISOLATED_SOURCE;
  std::lock_guard<std::mutex> device_guard(g_noalloc_lock[g->device]);
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  unsigned char* base = g_noalloc_base[g->device];
  if (!base) {
    DAWN_HIP(hipGetSymbolAddress((void**)&base, HIP_SYMBOL(scratch)));
    if (!base) return DAWN_ERR_HIP;
    g_noalloc_base[g->device] = base;
  }
  hipStream_t s = g->stream;

  SsArgs a{};
  a.n = g->n;
  a.source = source;
  a.words = ((int64_t)g->n + 63) / 64 * 2;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.w = g->weight;
  a.in_row_ptr = g->in_row_ptr;
  a.in_col = g->in_col;
  a.in_w = g->in_weight;
  a.dist = (int32_t*)(base + kNaDist);
  a.q = (int32_t*)(base + kNaQ);
  a.bm = (uint32_t*)(base + kNaBm);
  a.ring = (LevelCounters*)(base + kNaRing);
  a.part = (unsigned long long*)(base + kNaPart);
  a.st = (SchedState*)(base + kNaState);
  unsigned long long* reach = (unsigned long long*)(base + kNaReach);

  SchedParams p{};
  p.n = g->n;
  p.m = g->m;
  p.policy = policy;
  p.alpha = 15;
  p.beta = 18;
  p.batched = 0;
  p.allow_pull = 1;
  p.avg_degree = g->n ? (float)((double)g->m / g->n) : 0.f;
  p.max_degree = (float)g->max_out_degree;
  p.mlp = sched_single_mlp_if_present(g->device, &p.mlp_norm);
  SchedTrace tr{(float*)(base + kNaFeat), (int32_t*)(base + kNaMode)};

  const unsigned grid = grid_for(std::max<int64_t>(g->n, 1), kBlock, 4096);
  const int K = g->knobs.ss_chunk > 0 ? g->knobs.ss_chunk : (g->n < (1 << 20) ? 8 : 16);
  const auto t0 = std::chrono::steady_clock::now();
  ss_fill_kernel<<<grid_for(std::max<int64_t>(3 * a.words, g->n)), kBlock, 0, s>>>(a, 0);
  bfs_seed_kernel<<<1, 64, 0, s>>>(a, policy);
  DAWN_HIP(hipGetLastError());
  SchedState h_state{};
  int64_t launched = 0;
  const int64_t limit = (int64_t)g->n + 2;
  for (;;) {
    for (int i = 0; i < K; i++) {
      ss_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
      bfs_expand_kernel<<<grid, kBlock, 0, s>>>(a);
    }
    DAWN_HIP(hipGetLastError());
    launched += K;
    DAWN_HIP(hipMemcpyAsync(&h_state, a.st, sizeof(SchedState), hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    if (h_state.done || launched > limit) break;
  }
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (!h_state.done) return DAWN_ERR_HIP;  // step bound exceeded

  if (stats) {
    const int levels = h_state.levels;
    zero_u64_kernel<<<1, 64, 0, s>>>(reach, 2);
    reach_kernel<<<grid_for(g->n, kBlock, kReachBlocks), kBlock, 0, s>>>(a.dist, g->row_ptr, g->n, reach);
    DAWN_HIP(hipGetLastError());
    unsigned long long rr[2];
    DAWN_HIP(hipMemcpyAsync(rr, reach, sizeof rr, hipMemcpyDeviceToHost, s));
    // Trace entries 1..levels hold the modes of the expanding steps.
    std::vector<int32_t> modes(levels + 1);
    const int nrec = std::min(levels + 1, kTraceMax);
    if (nrec > 0)
      DAWN_HIP(hipMemcpyAsync(modes.data(), tr.mode, sizeof(int32_t) * nrec,
                              hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    stats->device_ms += ms;
    stats->vertices_reached += (int64_t)rr[0];
    stats->edges_reached += (int64_t)rr[1];
    stats->batches += 1;
    for (int st = 1; st <= levels && st < kTraceMax; st++) {
      stats->steps += 1;
      const int md = modes[st];
      if (md >= 0 && md < DAWN_NUM_MODES) stats->mode_steps[md] += 1;
    }
  }
  if (dist_out) {
    DAWN_HIP(hipMemcpyAsync(dist_out, a.dist, sizeof(int32_t) * g->n, hipMemcpyDefault, s));
    DAWN_HIP(hipStreamSynchronize(s));
  }
  return DAWN_OK;
}

}  // namespace dawn

extern "C" int dawn_bfs_noalloc(const dawn_graph* g, int32_t source, int32_t* dist,
                                int32_t policy, dawn_stats* stats) {
  return dawn::run_bfs_noalloc(const_cast<dawn_graph*>(g), source, dist, policy, stats);
}

