// 64-source bit-packed unweighted DAWN (batched BOVM) on one HIP device.
//
// Replaces the reference's per-source loops DAWN::MSSP_GPU::run
// (src/algorithm/gpu/mssp.cu:9-51), DAWN::APSP_GPU::run
// (src/algorithm/gpu/apsp.cu:62-111) and DAWN::MSSP_CPU::run
// (src/algorithm/cpu/mssp.cpp:9-32): one step advances 64 frontiers at
// once, each vertex carrying one uint64 word with bit b = source lane b.
//
// Layout in HBM (per graph, reused across batches):
//   seen   u64[npad]          lanes that have reached v
//   next   u64[npad]          push/bitmap scatter target (kept all-zero
//                             between steps by the finalize pass)
//   hist   u64[H][npad]       level history: slot l%H holds the lanes that
//                             reached v exactly at level l (= frontier of l)
//   q, qb  int32[2][qcap]     frontier queue (level parity): one entry per
//                             256 out-edges of a frontier vertex (vertex,
//                             first edge), so push work per entry is bounded
//   dist   int32[64][n]       output rows of the batch (caller or workspace)
// Step modes (all produce the same next level in next[]):
//   PULL    every vertex missing a lane ORs the previous level's words of its
//           in-neighbours (CSC): light vertices one lane each with early exit
//           once all missing lanes are found; heavy vertices' in-edge lists
//           in 256-edge chunks, one wave per chunk, merged by atomicOr.
//   PUSH    queue entries, workgroup-balanced expansion (LDS degree scan),
//           64-bit atomicOr of (frontier & ~seen[v]) into next[v].
//   BITMAP  dense scan of the previous level slot, same scatter as PUSH.
//   Every step ends with a finalize pass that folds next into the level
//   slot, seen, the queue and the level counters.
// After the last level one pass turns the history into int32 distances
// (64 coalesced row stores per wave) and per-lane sums/counts (64x64 bit
// transposes + popcount), so no distance is written twice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "common.hpp"
#include "sched.hpp"

namespace dawn {

constexpr int kLanes = 64;
constexpr int kHist = 32;        // level slots
constexpr int kChunk = 24;       // steps per captured chunk (< kHist)
constexpr int kSlices = 6;       // bits of a relative level (2^6 > kHist)

struct MsBatch {
  int32_t nlanes;
  int32_t lo;          // first level not yet flushed to dist
  uint64_t lanemask;   // valid lanes
  int32_t* dist;       // output rows: lane b at dist + b * stride
  int64_t stride;
  int32_t src[kLanes];
};

struct MsArgs {
  int32_t n;
  int32_t H;
  int64_t npad;
  int64_t m;
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* in_row_ptr;
  const int32_t* in_col;
  uint64_t* seen;
  uint64_t* next;
  uint64_t* hist;
  uint32_t* fbits;     // [2][fwords] 1 bit per vertex: in the frontier of a
  int64_t fwords;      // level (any lane), slot = level parity
  int32_t* q;          // [2][qcap] frontier entries: vertex
  int32_t* qb;         // [2][qcap] first out-edge of the entry's chunk
  int64_t qcap;
  const int32_t* hc_v;    // heavy in-edge chunks (pull)
  const int32_t* hc_beg;
  const int32_t* hc_end;
  int64_t num_hc;
  const int32_t* oc_v;    // heavy out-edge chunks (dense top-down)
  const int32_t* oc_beg;
  const int32_t* oc_end;
  int64_t num_oc;
  int32_t heavy;          // degree from which a vertex's edges are chunked
  int32_t balanced;       // light pull edge-balanced over the workgroup
  int32_t deg_bits;       // bits of the largest out-degree
  LevelCounters* ring;
  SchedState* st;
  MsBatch* batch;
};

struct MsOut {
  unsigned long long* sum;  // per lane, indexed [lane]
  unsigned long long* reached;
  unsigned long long* edges;
  int32_t want_edges;
};

struct MsWorkspace {
  int32_t n = 0;
  int64_t npad = 0;
  DevBuf<uint64_t> seen, next, hist, seen_bak;
  DevBuf<uint32_t> fbits;
  DevBuf<int32_t> q, qb, distbuf;
  int64_t qcap = 0;
  DevBuf<LevelCounters> ring, ring_bak;
  DevBuf<SchedState> st;
  DevBuf<MsBatch> batch;
  DevBuf<float> trace_feat;
  DevBuf<int32_t> trace_mode;
  DevBuf<unsigned long long> lane_stats;  // 3 * 64
  SchedState* h_state = nullptr;
  MsBatch* h_batch = nullptr;
  std::map<std::pair<long long, const void*>, hipGraphExec_t> graphs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, e_a = nullptr, e_b = nullptr;
};

void ss_invalidate_graphs(SsWorkspace* ws);

// Captured step graphs bake grid sizes; the autotuner drops them when it
// changes the knobs.
void invalidate_graphs(dawn_graph* g) {
  if (g->ms) {
    for (auto& kv : g->ms->graphs) (void)hipGraphExecDestroy(kv.second);
    g->ms->graphs.clear();
  }
  ss_invalidate_graphs(g->ss);
}

void ms_workspace_free(MsWorkspace* ws) {
  if (!ws) return;
  for (auto& kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
  if (ws->h_state) (void)hipHostFree(ws->h_state);
  if (ws->h_batch) (void)hipHostFree(ws->h_batch);
  for (hipEvent_t e : {ws->ev0, ws->ev1, ws->e_a, ws->e_b})
    if (e) (void)hipEventDestroy(e);
  delete ws;
}

namespace {

__device__ __forceinline__ int out_deg(const MsArgs& a, int32_t v) {
  return a.row_ptr[v + 1] - a.row_ptr[v];
}

__device__ __forceinline__ int owner_of(const int* off, int e) {
  int lo = 0, hi = kBlock - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Per-thread counters of a level, reduced over the workgroup and flushed
// with one set of atomics per workgroup (same-address atomics from every
// wave serialise in L2).  flush() must be called by every thread.
struct MsAcc {
  unsigned long long fv = 0, fe = 0, scout = 0, lanes = 0, uv = 0, ue = 0;
  uint64_t alive = 0;
  __device__ void flush(LevelCounters* c) {
    __shared__ unsigned long long s_v[kBlock / kWave][7];
    const int w = threadIdx.x / kWave;
    const unsigned long long r[7] = {wave_sum_u64(fv),    wave_sum_u64(fe), wave_sum_u64(scout),
                                     wave_sum_u64(lanes), wave_sum_u64(uv), wave_sum_u64(ue),
                                     wave_or_u64(alive)};
    if (lane_id() == 0)
      for (int i = 0; i < 7; i++) s_v[w][i] = r[i];
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t[7] = {0, 0, 0, 0, 0, 0, 0};
      for (int j = 0; j < kBlock / kWave; j++) {
        for (int i = 0; i < 6; i++) t[i] += s_v[j][i];
        t[6] |= s_v[j][6];
      }
      if (t[0]) {
        atomicAdd(&c->fv, t[0]);
        atomicAdd(&c->fe, t[1]);
        atomicAdd(&c->scout, t[2]);
        atomicAdd(&c->flanes, t[3]);
      }
      if (t[4]) {
        atomicAdd(&c->uv, t[4]);
        atomicAdd(&c->ue, t[5]);
      }
      if (t[6]) atomicOr((unsigned long long*)&c->alive, t[6]);
    }
  }
};

// ---------------------------------------------------------------- kernels ---

__global__ void ms_fill_kernel(MsArgs a, MsOut o) {
  const uint64_t inval = ~a.batch->lanemask;
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.npad; v += gs) {
    a.seen[v] = v < a.n ? inval : ~0ull;
    a.hist[v] = 0ull;  // level-0 slot
    a.next[v] = 0ull;
  }
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < a.fwords; w += gs)
    a.fbits[w] = 0u;  // level-0 frontier bitmap
  if (blockIdx.x == 0 && threadIdx.x < kRing) a.ring[threadIdx.x] = LevelCounters{};
  if (blockIdx.x == 0 && threadIdx.x < kLanes) {
    o.sum[threadIdx.x] = 0;
    o.reached[threadIdx.x] = 0;
    o.edges[threadIdx.x] = 0;
  }
}

// Level 0: the batch's sources (duplicates share one frontier vertex).
__global__ void ms_seed_kernel(MsArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const MsBatch& b = *a.batch;
  LevelCounters c{};
  for (int i = 0; i < b.nlanes; i++) {
    const int32_t v = b.src[i];
    const uint64_t bit = 1ull << i;
    if (a.hist[v] == 0) {
      const int d = out_deg(a, v);
      a.fbits[v >> 5] |= 1u << (v & 31);
      c.fv += 1;
      c.fe += (unsigned)d;
      c.scout += (unsigned)d;
    }
    a.hist[v] |= bit;
    a.seen[v] |= bit;
    c.flanes += 1;
    c.alive |= bit;
  }
  // Vertices still missing a lane after level 0, and their in-degree sum.
  c.uv = (unsigned long long)a.n;
  c.ue = (unsigned long long)a.m;
  for (int i = 0; i < b.nlanes; i++) {
    const int32_t v = b.src[i];
    const bool first = (a.hist[v] & ((1ull << i) - 1)) == 0;  // lowest lane of v
    if (first && a.seen[v] == ~0ull) {
      c.uv -= 1;
      c.ue -= (unsigned)(a.in_row_ptr[v + 1] - a.in_row_ptr[v]);
    }
  }
  a.ring[0] = c;
  SchedState st{};
  st.step = 0;
  st.e2c = a.m;
  st.scout = (long long)c.scout;
  st.force_mode = -1;
  st.last_mode = DAWN_MODE_PUSH;
  *a.st = st;
}

__global__ void ms_decide_kernel(MsArgs a, SchedParams p, SchedTrace tr) {
  sched_decide(a.st, a.ring, p, tr);
}

__global__ void __launch_bounds__(kBlock) ms_expand_kernel(MsArgs a) {
  __shared__ int s_off[kBlock];
  __shared__ int s_beg[kBlock];
  __shared__ uint64_t s_f[kBlock];
  __shared__ int s_scan[kBlock / kWave];
  const SchedState st = *a.st;
  if (st.done) return;
  const int lv = st.step;
  const uint64_t* prev = a.hist + (int64_t)((lv - 1) % a.H) * a.npad;
  const int lane = lane_id();
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gt == 0) a.ring[(lv + 1) % kRing] = LevelCounters{};

  if (st.mode == DAWN_MODE_PUSH) {
    const int32_t* q_in = a.q;  // built by ms_compact_kernel for this step
    const int32_t* qb_in = a.qb;
    const unsigned qn = a.ring[(lv - 1) % kRing].qlen;
    for (int64_t c0 = (int64_t)blockIdx.x * kBlock; c0 < qn; c0 += (int64_t)gridDim.x * kBlock) {
      const int64_t i = c0 + threadIdx.x;
      int beg = 0, deg = 0;
      uint64_t f = 0;
      if (i < qn) {
        const int32_t u = q_in[i];
        beg = qb_in[i];
        deg = min(a.row_ptr[u + 1] - beg, kPushChunk);
        f = prev[u];
      }
      int total;
      const int off = block_exclusive_scan(deg, s_scan, &total);
      s_off[threadIdx.x] = off;
      s_beg[threadIdx.x] = beg;
      s_f[threadIdx.x] = f;
      __syncthreads();
      for (int e = threadIdx.x; e < total; e += kBlock) {
        const int j = owner_of(s_off, e);
        const int32_t v = a.col[s_beg[j] + (e - s_off[j])];
        const uint64_t nv = s_f[j] & ~a.seen[v];
        if (nv) atomicOr((unsigned long long*)&a.next[v], (unsigned long long)nv);
      }
      __syncthreads();
    }
  } else if (st.mode == DAWN_MODE_BITMAP) {
    // Dense top-down: light frontier vertices one lane each, heavy ones in
    // out-edge chunks of knobs.pull_chunk edges, one wave per chunk.
    const uint32_t* fbp = a.fbits + (int64_t)((lv - 1) & 1) * a.fwords;
    const int64_t wstride = (int64_t)gridDim.x * kBlock;
    for (int64_t base = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < a.n;
         base += wstride) {
      const int32_t u = (int32_t)(base + lane);
      if (u >= a.n || !((fbp[u >> 5] >> (u & 31)) & 1u)) continue;
      const uint64_t f = prev[u];
      const int beg = a.row_ptr[u];
      const int deg = a.row_ptr[u + 1] - beg;
      if (deg >= a.heavy) continue;
      for (int j = 0; j < deg; j++) {
        const int32_t v = a.col[beg + j];
        const uint64_t nv = f & ~a.seen[v];
        if (nv) atomicOr((unsigned long long*)&a.next[v], (unsigned long long)nv);
      }
    }
    const int64_t wave = (int64_t)(blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t c = wave; c < a.num_oc; c += nwaves) {
      const int32_t u = a.oc_v[c];
      const uint64_t f = prev[u];
      if (!f) continue;
      const int e = a.oc_end[c];
      for (int k = a.oc_beg[c] + lane; k < e; k += kWave) {
        const int32_t v = a.col[k];
        const uint64_t nv = f & ~a.seen[v];
        if (nv) atomicOr((unsigned long long*)&a.next[v], (unsigned long long)nv);
      }
    }
  } else {  // DAWN_MODE_PULL
    // Only lanes alive in the previous level can arrive now: a vertex looks
    // for miss = (lanes it lacks) & (live lanes) and stops once all are found.
    const uint64_t alive = a.ring[(lv - 1) % kRing].alive;
    // Frontier bitmap of the previous level (2 MB at n = 2^24, L2-resident):
    // a 64-bit word is gathered only for in-neighbours that are in it.
    const uint32_t* fb = a.fbits + (int64_t)((lv - 1) & 1) * a.fwords;
    const int64_t wstride = (int64_t)gridDim.x * kBlock;
    if (a.balanced) {
      // Light vertices, edge-balanced: the workgroup's 256 vertices' in-edge
      // lists are concatenated (degree scan) and read one edge per thread,
      // coalesced; hits are OR-ed into the owner's LDS word.  Gathers stop
      // for an owner once it has found all its missing lanes.
      __shared__ unsigned long long s_acc[kBlock];
      for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < a.n; b0 += wstride) {
        const int32_t v = (int32_t)(b0 + threadIdx.x);
        const uint64_t miss = v < a.n ? (~a.seen[v] & alive) : 0ull;
        int beg = 0, deg = 0;
        if (miss) {
          beg = a.in_row_ptr[v];
          const int d = a.in_row_ptr[v + 1] - beg;
          deg = d < a.heavy ? d : 0;
        }
        int total;
        const int off = block_exclusive_scan(deg, s_scan, &total);
        s_off[threadIdx.x] = off;
        s_beg[threadIdx.x] = beg;
        s_f[threadIdx.x] = miss;
        s_acc[threadIdx.x] = 0ull;
        __syncthreads();
        for (int e = threadIdx.x; e < total; e += kBlock) {
          const int j = owner_of(s_off, e);
          const int32_t u = a.in_col[s_beg[j] + (e - s_off[j])];
          if ((fb[u >> 5] >> (u & 31)) & 1u) {
            const uint64_t mj = s_f[j];
            if ((__atomic_load_n(&s_acc[j], __ATOMIC_RELAXED) & mj) != mj) {
              const uint64_t hit = prev[u] & mj;
              if (hit) atomicOr(&s_acc[j], (unsigned long long)hit);
            }
          }
        }
        __syncthreads();
        if (deg && s_acc[threadIdx.x]) a.next[v] = s_acc[threadIdx.x];
      }
    } else
    // Light vertices (in-degree < a.heavy): one lane each, four gathers
    // in flight between early-exit checks; the lane owns next[v].
    for (int64_t base = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < a.n;
         base += wstride) {
      const int32_t v = (int32_t)(base + lane);
      const uint64_t miss = v < a.n ? (~a.seen[v] & alive) : 0ull;
      if (!miss) continue;
      const int beg = a.in_row_ptr[v];
      const int deg = a.in_row_ptr[v + 1] - beg;
      if (deg >= a.heavy || deg == 0) continue;
      const int32_t* nb = a.in_col + beg;
      uint64_t acc_w = 0;
      int j = 0;
      for (; j + 4 <= deg; j += 4) {
        const int32_t u0 = nb[j], u1 = nb[j + 1], u2 = nb[j + 2], u3 = nb[j + 3];
        if ((fb[u0 >> 5] >> (u0 & 31)) & 1u) acc_w |= prev[u0];
        if ((fb[u1 >> 5] >> (u1 & 31)) & 1u) acc_w |= prev[u1];
        if ((fb[u2 >> 5] >> (u2 & 31)) & 1u) acc_w |= prev[u2];
        if ((fb[u3 >> 5] >> (u3 & 31)) & 1u) acc_w |= prev[u3];
        if ((acc_w & miss) == miss) break;
      }
      if ((acc_w & miss) != miss)
        for (; j < deg; j++) {
          const int32_t u = nb[j];
          if ((fb[u >> 5] >> (u & 31)) & 1u) acc_w |= prev[u];
        }
      if (acc_w & miss) a.next[v] = acc_w & miss;
    }
    // Heavy vertices: their in-edge lists are cut into chunks of at most
    // knobs.pull_chunk edges, one wave per chunk, merged with atomicOr.
    const int64_t wave = (int64_t)(blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t c = wave; c < a.num_hc; c += nwaves) {
      const int32_t v = a.hc_v[c];
      const uint64_t miss = ~a.seen[v] & alive;
      if (!miss) continue;
      if ((a.next[v] & miss) == miss) continue;  // other chunks already found them all
      const int he = a.hc_end[c];
      uint64_t part = 0;
      for (int k = a.hc_beg[c] + lane; k < he; k += kWave) {
        const int32_t u = a.in_col[k];
        if ((fb[u >> 5] >> (u & 31)) & 1u) part |= prev[u];
      }
      part = wave_or_u64(part) & miss;
      if (lane == 0 && part) atomicOr((unsigned long long*)&a.next[v], (unsigned long long)part);
    }
  }
}

// Folds next[] into the level slot, seen, the queue and the counters.
__global__ void __launch_bounds__(kBlock) ms_finalize_kernel(MsArgs a) {
  const SchedState st = *a.st;
  if (st.done) return;
  const int lv = st.step;
  uint64_t* cur = a.hist + (int64_t)(lv % a.H) * a.npad;
  uint32_t* fb = a.fbits + (int64_t)(lv & 1) * a.fwords;
  LevelCounters* c_out = a.ring + (lv % kRing);
  const uint64_t alive_prev = a.ring[(lv - 1) % kRing].alive;
  MsAcc acc;
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < a.npad; b0 += (int64_t)gridDim.x * kBlock) {
    const int32_t v = (int32_t)(b0 + threadIdx.x);
    uint64_t nw = 0;
    if (v < a.n) {
      const uint64_t nx = a.next[v];
      const uint64_t sv = a.seen[v];
      if (nx) a.next[v] = 0ull;
      nw = nx & ~sv;
      cur[v] = nw;
      if (nw) {
        a.seen[v] = sv | nw;
        const int d = out_deg(a, v);
        acc.fv += 1;
        acc.fe += (unsigned)d;
        acc.scout += (unsigned)(d > 0 ? d : 1);
        acc.lanes += (unsigned)__popcll(nw);
        acc.alive |= nw;
      }
      if ((~(sv | nw) & alive_prev) != 0) {  // still misses a live lane
        acc.uv += 1;
        acc.ue += (unsigned)(a.in_row_ptr[v + 1] - a.in_row_ptr[v]);
      }
    }
    // Frontier bitmap of the level: each wave owns its 64 vertices' bits.
    const uint64_t fm = __ballot(nw != 0);
    if (lane_id() == 0) {
      const int64_t w = (b0 + (threadIdx.x & ~(kWave - 1))) >> 5;
      fb[w] = (uint32_t)fm;
      fb[w + 1] = (uint32_t)(fm >> 32);
    }
  }
  acc.flush(c_out);
}

// Builds the frontier queue of the previous level from its bitmap, only
// for a PUSH step: one entry per kPushChunk out-edges of each frontier
// vertex, so the push step's work per entry is bounded whatever the degree.
__global__ void __launch_bounds__(kBlock) ms_compact_kernel(MsArgs a) {
  __shared__ int s_wtot[kBlock / kWave];
  __shared__ unsigned s_base;
  const SchedState st = *a.st;
  if (st.done || st.mode != DAWN_MODE_PUSH) return;
  const int lv = st.step;
  const uint32_t* fb = a.fbits + (int64_t)((lv - 1) & 1) * a.fwords;
  int32_t* q = a.q;
  int32_t* qb = a.qb;
  unsigned* qlen = &a.ring[(lv - 1) % kRing].qlen;
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < a.npad; b0 += (int64_t)gridDim.x * kBlock) {
    const int32_t v = (int32_t)(b0 + threadIdx.x);
    const bool isf = v < a.n && ((fb[v >> 5] >> (v & 31)) & 1u);
    if (!__syncthreads_or(isf)) continue;  // block-uniform
    const int d = isf ? out_deg(a, v) : 0;
    const int k = (d + kPushChunk - 1) / kPushChunk;
    const int pos = block_append_n(qlen, k, s_wtot, &s_base);
    if (isf) {
      const int32_t rb = a.row_ptr[v];
      for (int i = 0; i < k; i++) {
        q[pos + i] = v;
        qb[pos + i] = rb + i * kPushChunk;
      }
    }
  }
}

// Turns levels [lo, hi) of the history into distances.  from_state: lo and
// hi come from the batch / scheduler state and the pass runs only once the
// traversal is done (then also writes INF for lanes that never arrived).
__global__ void __launch_bounds__(kBlock) ms_final_kernel(MsArgs a, MsOut o, int from_state,
                                                          int lo_arg, int hi_arg, int final_arg) {
  __shared__ unsigned long long s_sum[kLanes], s_reach[kLanes], s_edges[kLanes];
  int lo = lo_arg, hi = hi_arg, fin = final_arg;
  if (from_state) {
    const SchedState st = *a.st;
    if (!st.done) return;
    lo = a.batch->lo;
    hi = st.levels;
    fin = 1;
  }
  const int nl = a.batch->nlanes;
  int32_t* const dist = a.batch->dist;
  const int64_t stride = a.batch->stride;
  if (threadIdx.x < kLanes) {
    s_sum[threadIdx.x] = 0;
    s_reach[threadIdx.x] = 0;
    s_edges[threadIdx.x] = 0;
  }
  __syncthreads();
  // Levels are stored relative to lo in K bit slices: S_k = OR of the
  // level words whose (l - lo) has bit k set.  hi - lo <= kHist.
  int K = 0;
  while ((1 << K) < hi - lo) K++;
  const int lane = lane_id();
  // Each lane owns 4 consecutive vertices (a wave 256), so a row store is
  // 16 bytes per lane, 1 KB per wave.  With one flush of every level
  // (lo == 0, final) every (lane, vertex) is written: level or INF.
  const bool vec_ok = dist && (stride % 4 == 0) && (((uintptr_t)dist & 15) == 0) && lo == 0 && fin;
  const int64_t wstride = (int64_t)gridDim.x * kBlock * 4;
  for (int64_t base = (int64_t)blockIdx.x * kBlock * 4 + (threadIdx.x & ~(kWave - 1)) * 4; base < a.n;
       base += wstride) {
    const int64_t v0 = base + 4 * lane;
    uint64_t F[4], S[4][kSlices], sv[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      F[j] = 0;
      sv[j] = 0;
#pragma unroll
      for (int k = 0; k < kSlices; k++) S[j][k] = 0;
    }
    if (v0 + 3 < a.n) {
      const ulonglong2* sp = reinterpret_cast<const ulonglong2*>(a.seen + v0);
      const ulonglong2 s01 = sp[0], s23 = sp[1];
      sv[0] = s01.x; sv[1] = s01.y; sv[2] = s23.x; sv[3] = s23.y;
      for (int l = lo; l < hi; l++) {
        const ulonglong2* hp = reinterpret_cast<const ulonglong2*>(a.hist + (int64_t)(l % a.H) * a.npad + v0);
        const ulonglong2 h01 = hp[0], h23 = hp[1];
        const uint64_t w[4] = {h01.x, h01.y, h23.x, h23.y};
        const int r = l - lo;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          F[j] |= w[j];
#pragma unroll
          for (int k = 0; k < kSlices; k++)
            if ((r >> k) & 1) S[j][k] |= w[j];
        }
      }
    } else {
      for (int j = 0; j < 4; j++) {
        const int64_t v = v0 + j;
        if (v >= a.n) continue;
        sv[j] = a.seen[v];
        for (int l = lo; l < hi; l++) {
          const uint64_t w = a.hist[(int64_t)(l % a.H) * a.npad + v];
          const int r = l - lo;
          F[j] |= w;
#pragma unroll
          for (int k = 0; k < kSlices; k++)
            if ((r >> k) & 1) S[j][k] |= w;
        }
      }
    }
    if (dist && v0 < a.n) {
      const bool full = vec_ok && v0 + 3 < a.n;
      for (int b = 0; b < nl; b++) {
        int32_t d[4];
        bool wr[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          int32_t x = lo;
#pragma unroll
          for (int k = 0; k < kSlices; k++)
            if (k < K) x += (int32_t)((S[j][k] >> b) & 1ull) << k;
          wr[j] = true;
          if (!((F[j] >> b) & 1ull)) {
            x = kInf;
            wr[j] = fin && !((sv[j] >> b) & 1ull);
          }
          d[j] = x;
        }
        int32_t* row = dist + (int64_t)b * stride + v0;
        if (full) {
          *reinterpret_cast<int4*>(row) = make_int4(d[0], d[1], d[2], d[3]);
        } else {
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (v0 + j < a.n && wr[j]) row[j] = d[j];
        }
      }
    }
    // Per-lane sums: sum_b = lo * reach_b + sum_k 2^k * |{v : bit b of S_k}|,
    // counted with 64x64 bit transposes (lane b receives lane-b bits).
    unsigned long long sum = 0, reach = 0, edges = 0;
    uint64_t tfs[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint64_t tf = wave_transpose64(F[j]);
      tfs[j] = tf;
      const unsigned long long rc = (unsigned long long)__popcll(tf);
      reach += rc;
      sum += rc * (unsigned long long)lo;
#pragma unroll
      for (int k = 0; k < kSlices; k++) {
        if (k < K) {
          const uint64_t t = wave_transpose64(S[j][k]);
          sum += (unsigned long long)__popcll(t) << k;
        }
      }
    }
    if (o.want_edges) {
      // edges_b = sum_i [bit i of tf] * deg_i, bit-sliced over the degree:
      // sum_t 2^t * popcount(tf & ballot(bit t of deg)).
      int dg[4];
#pragma unroll
      for (int j = 0; j < 4; j++) dg[j] = v0 + j < a.n ? out_deg(a, (int32_t)(v0 + j)) : 0;
      for (int t = 0; t < a.deg_bits; t++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint64_t dm = __ballot((dg[j] >> t) & 1);
          edges += (unsigned long long)__popcll(tfs[j] & dm) << t;
        }
      }
    }
    if (sum) atomicAdd(&s_sum[lane], sum);
    if (reach) atomicAdd(&s_reach[lane], reach);
    if (edges) atomicAdd(&s_edges[lane], edges);
  }
  __syncthreads();
  if (threadIdx.x < nl) {
    const int b = threadIdx.x;
    if (s_sum[b]) atomicAdd(&o.sum[b], s_sum[b]);
    if (s_reach[b]) atomicAdd(&o.reached[b], s_reach[b]);
    if (s_edges[b]) atomicAdd(&o.edges[b], s_edges[b]);
  }
}

int ensure_ms(dawn_graph* g) {
  if (!g->ms) g->ms = new MsWorkspace();
  MsWorkspace* ws = g->ms;
  if (ws->n == g->n && ws->seen.p) return DAWN_OK;
  ws->n = g->n;
  ws->npad = ((int64_t)g->n + 63) / 64 * 64;
  DAWN_TRY(ws->seen.alloc(ws->npad));
  DAWN_TRY(ws->next.alloc(ws->npad));
  DAWN_TRY(ws->fbits.alloc((size_t)2 * (ws->npad / 32)));
  DAWN_TRY(ws->hist.alloc((size_t)kHist * ws->npad));
  // Queue entries per level: one per kPushChunk out-edges of each frontier
  // vertex, at most n + m / kPushChunk.
  ws->qcap = (int64_t)g->n + g->m / kPushChunk + 64;
  DAWN_TRY(ws->q.alloc((size_t)(2 * ws->qcap)));
  DAWN_TRY(ws->qb.alloc((size_t)(2 * ws->qcap)));
  DAWN_TRY(ws->ring.alloc(kRing));
  DAWN_TRY(ws->ring_bak.alloc(kRing));
  DAWN_TRY(ws->st.alloc(1));
  DAWN_TRY(ws->batch.alloc(1));
  DAWN_TRY(ws->trace_feat.alloc((size_t)kTraceMax * 8));
  DAWN_TRY(ws->trace_mode.alloc(kTraceMax));
  DAWN_TRY(ws->lane_stats.alloc(3 * kLanes));
  DAWN_HIP(hipMemset(ws->next.p, 0, sizeof(uint64_t) * ws->npad));
  if (!ws->h_state) DAWN_HIP(hipHostMalloc((void**)&ws->h_state, sizeof(SchedState)));
  if (!ws->h_batch) DAWN_HIP(hipHostMalloc((void**)&ws->h_batch, sizeof(MsBatch)));
  for (hipEvent_t* e : {&ws->ev0, &ws->ev1, &ws->e_a, &ws->e_b})
    if (!*e) DAWN_HIP(hipEventCreate(e));
  return DAWN_OK;
}

MsArgs make_args(dawn_graph* g) {
  MsWorkspace* ws = g->ms;
  MsArgs a{};
  a.n = g->n;
  a.H = kHist;
  a.npad = ws->npad;
  a.m = g->m;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.in_row_ptr = g->in_row_ptr;
  a.in_col = g->in_col;
  a.seen = ws->seen.p;
  a.next = ws->next.p;
  a.fbits = ws->fbits.p;
  a.fwords = ws->npad / 32;
  a.hist = ws->hist.p;
  a.q = ws->q.p;
  a.qb = ws->qb.p;
  a.qcap = ws->qcap;
  a.hc_v = g->hc_v;
  a.hc_beg = g->hc_beg;
  a.hc_end = g->hc_end;
  a.num_hc = g->num_hc;
  a.oc_v = g->oc_v;
  a.oc_beg = g->oc_beg;
  a.oc_end = g->oc_end;
  a.num_oc = g->num_oc;
  a.heavy = g->knobs.heavy_degree;
  a.balanced = g->knobs.pull_balanced;
  a.deg_bits = 32 - __builtin_clz((unsigned)std::max(1, g->max_out_degree));
  a.ring = ws->ring.p;
  a.st = ws->st.p;
  a.batch = ws->batch.p;
  return a;
}

SchedParams make_params(dawn_graph* g, int32_t policy) {
  SchedParams p{};
  p.n = g->n;
  p.m = g->m;
  p.policy = policy;
  p.alpha = 15;
  p.beta = 18;
  p.batched = 1;
  p.allow_pull = 1;
  p.avg_degree = g->n ? (float)((double)g->m / g->n) : 0.f;
  p.max_degree = (float)g->max_out_degree;
  p.mlp = sched_mlp_device(g->device);
  p.mlp_norm = sched_norm_device(g->device);
  return p;
}

// Dense step kernels: at most 8 workgroups per CU (2048), grid-stride beyond.
unsigned ms_grid(const dawn_graph* g) {
  return grid_for(std::max<int64_t>(g->n, 1), kBlock, (unsigned)g->knobs.dense_blocks);
}
// Distance pass: 4 vertices per thread.
unsigned ms_final_grid(const dawn_graph* g) {
  return grid_for(std::max<int64_t>(g->n, 1), kBlock * 4, (unsigned)g->knobs.dense_blocks);
}

// Stages one batch's sources (taken mod n) into the pinned batch block.
void stage_batch(MsWorkspace* ws, const int32_t* sources, int32_t lanes) {
  MsBatch& b = *ws->h_batch;
  std::memset(&b, 0, sizeof b);
  b.nlanes = lanes;
  b.lo = 0;
  b.lanemask = lanes >= 64 ? ~0ull : ((1ull << lanes) - 1);
  for (int i = 0; i < lanes; i++) {
    int64_t s = sources[i] % (int64_t)ws->n;
    if (s < 0) s += ws->n;
    b.src[i] = (int32_t)s;
  }
}

}  // namespace

// Runs the batches of an unweighted multi-source traversal.
//   dist: caller buffer (S*n int32, host/device) or null
//   sums/reached/edges: per source (host) or null
// mode_trace (optional): receives per-step modes of every batch.
int run_msbfs(dawn_graph* g, const int32_t* sources, int32_t S, int32_t* dist,
              int64_t* sums, int64_t* reached, int64_t* edges, int32_t policy,
              dawn_stats* stats) {
  if (!g || S < 0 || (S > 0 && !sources) || g->n <= 0) return DAWN_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  DAWN_TRY(ensure_ms(g));
  MsWorkspace* ws = g->ms;
  hipStream_t s = g->stream;
  const bool dist_dev = dist && is_device_pointer(dist);
  if (!dist_dev) DAWN_TRY(ws->distbuf.alloc((size_t)kLanes * g->n));

  MsArgs a = make_args(g);
  SchedParams p = make_params(g, policy);
  SchedTrace tr{ws->trace_feat.p, ws->trace_mode.p};
  const unsigned grid = ms_grid(g);
  const unsigned fgrid = ms_final_grid(g);
  unsigned long long* ls = ws->lane_stats.p;

  // Graph of one batch: fill, seed, kChunk steps, final pass (no-op unless
  // the traversal finished within the chunk).  Sources and the output
  // pointer live in the device batch block, so one graph serves all batches.
  auto get_graph = [&](int want_edges, hipGraphExec_t* exec) -> int {
    const std::pair<long long, const void*> key{((long long)policy << 8) | want_edges,
                                                (const void*)p.mlp};
    auto it = ws->graphs.find(key);
    if (it != ws->graphs.end()) {
      *exec = it->second;
      return DAWN_OK;
    }
    MsOut o{ls, ls + kLanes, ls + 2 * kLanes, want_edges};
    hipGraph_t graph;
    DAWN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    ms_fill_kernel<<<grid_for(ws->npad), kBlock, 0, s>>>(a, o);
    ms_seed_kernel<<<1, 64, 0, s>>>(a);
    for (int i = 0; i < kChunk; i++) {
      ms_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
      ms_compact_kernel<<<grid, kBlock, 0, s>>>(a);
      ms_expand_kernel<<<grid, kBlock, 0, s>>>(a);
      ms_finalize_kernel<<<grid, kBlock, 0, s>>>(a);
    }
    ms_final_kernel<<<fgrid, kBlock, 0, s>>>(a, o, 1, 0, 0, 1);
    DAWN_HIP(hipStreamEndCapture(s, &graph));
    hipError_t e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DAWN_HIP(e);
    ws->graphs[key] = *exec;
    return DAWN_OK;
  };

  const int want_edges = (stats || edges) ? 1 : 0;
  std::vector<unsigned long long> h_ls(3 * kLanes);
  std::vector<int32_t> h_modes(kTraceMax);
  float total_ms = 0.f;
  for (int32_t b0 = 0; b0 < S; b0 += kLanes) {
    const int32_t lanes = std::min<int32_t>(kLanes, S - b0);
    int32_t* out_base = dist_dev ? dist + (int64_t)b0 * g->n : ws->distbuf.p;
    hipGraphExec_t exec;
    DAWN_TRY(get_graph(want_edges, &exec));
    stage_batch(ws, sources + b0, lanes);
    ws->h_batch->dist = out_base;
    ws->h_batch->stride = g->n;
    DAWN_HIP(hipMemcpyAsync(ws->batch.p, ws->h_batch, sizeof(MsBatch), hipMemcpyHostToDevice, s));
    DAWN_HIP(hipEventRecord(ws->ev0, s));
    DAWN_HIP(hipGraphLaunch(exec, s));
    DAWN_HIP(hipMemcpyAsync(ws->h_state, ws->st.p, sizeof(SchedState), hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    // Long traversals: more chunks, flushing all but the newest level
    // before the history ring wraps.
    int lo = 0;
    int64_t guard_steps = kChunk;
    while (!ws->h_state->done) {
      MsOut o{ls, ls + kLanes, ls + 2 * kLanes, want_edges};
      const int stp = ws->h_state->step;
      if (stp > lo) {
        ms_final_kernel<<<fgrid, kBlock, 0, s>>>(a, o, 0, lo, stp, 0);
        lo = stp;
      }
      for (int i = 0; i < kChunk; i++) {
        ms_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
        ms_compact_kernel<<<grid, kBlock, 0, s>>>(a);
        ms_expand_kernel<<<grid, kBlock, 0, s>>>(a);
        ms_finalize_kernel<<<grid, kBlock, 0, s>>>(a);
      }
      guard_steps += kChunk;
      DAWN_HIP(hipMemcpyAsync(ws->h_state, ws->st.p, sizeof(SchedState), hipMemcpyDeviceToHost, s));
      DAWN_HIP(hipStreamSynchronize(s));
      if (ws->h_state->done) {
        ms_final_kernel<<<fgrid, kBlock, 0, s>>>(a, o, 0, lo, ws->h_state->levels, 1);
        break;
      }
      if (guard_steps > (int64_t)g->n + 2 * kChunk) return DAWN_ERR_HIP;
    }
    DAWN_HIP(hipGetLastError());
    DAWN_HIP(hipEventRecord(ws->ev1, s));
    DAWN_HIP(hipMemcpyAsync(h_ls.data(), ls, sizeof(unsigned long long) * 3 * kLanes,
                            hipMemcpyDeviceToHost, s));
    const int levels_now = ws->h_state->levels;
    if (stats && levels_now > 0)
      DAWN_HIP(hipMemcpyAsync(h_modes.data(), ws->trace_mode.p,
                              sizeof(int32_t) * std::min(levels_now + 1, kTraceMax),
                              hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    DAWN_HIP(hipEventElapsedTime(&ms, ws->ev0, ws->ev1));
    total_ms += ms;
    if (dist && !dist_dev) {
      DAWN_HIP(hipMemcpyAsync(dist + (int64_t)b0 * g->n, ws->distbuf.p,
                              sizeof(int32_t) * (size_t)lanes * g->n, hipMemcpyDeviceToHost, s));
      DAWN_HIP(hipStreamSynchronize(s));
    }
    for (int i = 0; i < lanes; i++) {
      if (sums) sums[b0 + i] = (int64_t)h_ls[i];
      if (reached) reached[b0 + i] = (int64_t)h_ls[kLanes + i];
      if (edges) edges[b0 + i] = (int64_t)h_ls[2 * kLanes + i];
    }
    if (stats) {
      stats->batches += 1;
      for (int i = 0; i < lanes; i++) {
        stats->vertices_reached += (int64_t)h_ls[kLanes + i];
        stats->edges_reached += (int64_t)h_ls[2 * kLanes + i];
        const int32_t sv = ws->h_batch->src[i];
        if (g->h_row_ptr[sv] == g->h_row_ptr[sv + 1]) stats->skipped_sources += 1;
      }
      const int levels = ws->h_state->levels;
      for (int st = 1; st <= levels && st < kTraceMax; st++) {
        stats->steps += 1;
        const int md = h_modes[st];
        if (md >= 0 && md < DAWN_NUM_MODES) stats->mode_steps[md] += 1;
      }
    }
  }
  if (stats) stats->device_ms += total_ms;
  return DAWN_OK;
}

// Per-step mode timings from identical state (scheduler training data).
// Each step: decide (features), snapshot, then every mode is run `repeats`
// times from the snapshot and timed with HIP events; the minimum is kept.
int run_ms_record(dawn_graph* g, const int32_t* sources, int32_t S, int32_t repeats,
                  float* features, float* times_ms, int32_t* chosen, int32_t max_rows,
                  int32_t* rows_out, int32_t policy) {
  if (!g || S <= 0 || !sources || repeats < 1) return DAWN_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  DAWN_TRY(ensure_ms(g));
  MsWorkspace* ws = g->ms;
  DAWN_TRY(ws->seen_bak.alloc(ws->npad));
  hipStream_t s = g->stream;
  MsArgs a = make_args(g);
  SchedParams p = make_params(g, policy);
  SchedTrace tr{ws->trace_feat.p, ws->trace_mode.p};
  const unsigned grid = ms_grid(g);
  unsigned long long* ls = ws->lane_stats.p;
  MsOut o{ls, ls + kLanes, ls + 2 * kLanes, 0};
  int32_t rows = 0;
  for (int32_t b0 = 0; b0 < S && rows < max_rows; b0 += kLanes) {
    const int32_t lanes = std::min<int32_t>(kLanes, S - b0);
    stage_batch(ws, sources + b0, lanes);
    DAWN_HIP(hipMemcpyAsync(ws->batch.p, ws->h_batch, sizeof(MsBatch), hipMemcpyHostToDevice, s));
    ms_fill_kernel<<<grid_for(ws->npad), kBlock, 0, s>>>(a, o);
    ms_seed_kernel<<<1, 64, 0, s>>>(a);
    for (int step = 1; rows < max_rows && step < kHist; step++) {
      ms_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
      DAWN_HIP(hipMemcpyAsync(ws->h_state, ws->st.p, sizeof(SchedState), hipMemcpyDeviceToHost, s));
      DAWN_HIP(hipStreamSynchronize(s));
      if (ws->h_state->done) break;
      const SchedState snap = *ws->h_state;
      if (chosen) chosen[rows] = snap.mode;
      DAWN_HIP(hipMemcpyAsync(features + (size_t)rows * 8, ws->trace_feat.p + (size_t)snap.step * 8,
                              sizeof(float) * 8, hipMemcpyDeviceToHost, s));
      DAWN_HIP(hipMemcpyAsync(ws->seen_bak.p, ws->seen.p, sizeof(uint64_t) * ws->npad,
                              hipMemcpyDeviceToDevice, s));
      DAWN_HIP(hipMemcpyAsync(ws->ring_bak.p, ws->ring.p, sizeof(LevelCounters) * kRing,
                              hipMemcpyDeviceToDevice, s));
      for (int mode = 0; mode < DAWN_NUM_MODES; mode++) {
        float best = 1e30f;
        for (int r = 0; r < repeats; r++) {
          SchedState st2 = snap;
          st2.mode = mode;
          *ws->h_state = st2;
          DAWN_HIP(hipMemcpyAsync(ws->st.p, ws->h_state, sizeof(SchedState), hipMemcpyHostToDevice, s));
          DAWN_HIP(hipMemcpyAsync(ws->seen.p, ws->seen_bak.p, sizeof(uint64_t) * ws->npad,
                                  hipMemcpyDeviceToDevice, s));
          DAWN_HIP(hipMemcpyAsync(ws->ring.p, ws->ring_bak.p, sizeof(LevelCounters) * kRing,
                                  hipMemcpyDeviceToDevice, s));
          DAWN_HIP(hipEventRecord(ws->e_a, s));
          ms_compact_kernel<<<grid, kBlock, 0, s>>>(a);
          ms_expand_kernel<<<grid, kBlock, 0, s>>>(a);
          ms_finalize_kernel<<<grid, kBlock, 0, s>>>(a);
          DAWN_HIP(hipEventRecord(ws->e_b, s));
          DAWN_HIP(hipEventSynchronize(ws->e_b));
          float ms = 0.f;
          DAWN_HIP(hipEventElapsedTime(&ms, ws->e_a, ws->e_b));
          best = std::min(best, ms);
        }
        times_ms[(size_t)rows * DAWN_NUM_MODES + mode] = best;
      }
      rows++;
    }
    DAWN_HIP(hipStreamSynchronize(s));
  }
  *rows_out = rows;
  return DAWN_OK;
}

}  // namespace dawn

namespace dawn {

// Same kernels as run_msbfs, launched one by one with HIP events around each
// launch on the graph's stream.  kernel_ms / launches are indexed by
// DAWN_PROF_* (fill, seed, decide, expand push, expand pull, expand bitmap,
// finalize, final).  levels_out: per batch number of levels (may be null).
int run_ms_profile(dawn_graph* g, const int32_t* sources, int32_t S, int32_t policy,
                   double* kernel_ms, int64_t* launches, int32_t* levels_out) {
  if (!g || S <= 0 || !sources || !kernel_ms || !launches) return DAWN_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  DAWN_TRY(ensure_ms(g));
  MsWorkspace* ws = g->ms;
  DAWN_TRY(ws->distbuf.alloc((size_t)kLanes * g->n));
  hipStream_t s = g->stream;
  MsArgs a = make_args(g);
  SchedParams p = make_params(g, policy);
  SchedTrace tr{ws->trace_feat.p, ws->trace_mode.p};
  const unsigned grid = ms_grid(g);
  const unsigned fgrid = ms_final_grid(g);
  unsigned long long* ls = ws->lane_stats.p;
  // Same work as a traversal that writes distance rows and statistics.
  MsOut o{ls, ls + kLanes, ls + 2 * kLanes, 1};
  for (int i = 0; i < 8; i++) {
    kernel_ms[i] = 0;
    launches[i] = 0;
  }
  auto timed = [&](int slot, auto&& launch) -> int {
    DAWN_HIP(hipEventRecord(ws->e_a, s));
    launch();
    DAWN_HIP(hipGetLastError());
    DAWN_HIP(hipEventRecord(ws->e_b, s));
    DAWN_HIP(hipEventSynchronize(ws->e_b));
    float ms = 0.f;
    DAWN_HIP(hipEventElapsedTime(&ms, ws->e_a, ws->e_b));
    kernel_ms[slot] += ms;
    launches[slot] += 1;
    return DAWN_OK;
  };
  int batch = 0;
  for (int32_t b0 = 0; b0 < S; b0 += kLanes, batch++) {
    const int32_t lanes = std::min<int32_t>(kLanes, S - b0);
    stage_batch(ws, sources + b0, lanes);
    ws->h_batch->dist = ws->distbuf.p;
    ws->h_batch->stride = g->n;
    DAWN_HIP(hipMemcpyAsync(ws->batch.p, ws->h_batch, sizeof(MsBatch), hipMemcpyHostToDevice, s));
    DAWN_TRY(timed(0, [&] { ms_fill_kernel<<<grid_for(ws->npad), kBlock, 0, s>>>(a, o); }));
    DAWN_TRY(timed(1, [&] { ms_seed_kernel<<<1, 64, 0, s>>>(a); }));
    int lo = 0;
    for (int64_t guard_steps = 0;; guard_steps++) {
      if (guard_steps > (int64_t)g->n + 2) return DAWN_ERR_HIP;
      DAWN_TRY(timed(2, [&] { ms_decide_kernel<<<1, 64, 0, s>>>(a, p, tr); }));
      DAWN_HIP(hipMemcpy(ws->h_state, ws->st.p, sizeof(SchedState), hipMemcpyDeviceToHost));
      const SchedState st = *ws->h_state;
      if (st.done) {
        DAWN_TRY(timed(7, [&] { ms_final_kernel<<<fgrid, kBlock, 0, s>>>(a, o, 0, lo, st.levels, 1); }));
        if (levels_out) levels_out[batch] = st.levels;
        break;
      }
      const int slot = st.mode == DAWN_MODE_PUSH ? 3 : st.mode == DAWN_MODE_PULL ? 4 : 5;
      DAWN_TRY(timed(slot, [&] {
        ms_compact_kernel<<<grid, kBlock, 0, s>>>(a);
        ms_expand_kernel<<<grid, kBlock, 0, s>>>(a);
      }));
      DAWN_TRY(timed(6, [&] { ms_finalize_kernel<<<grid, kBlock, 0, s>>>(a); }));
      if (st.step - lo >= kHist - 2) {  // flush before the history ring wraps
        DAWN_TRY(timed(7, [&] { ms_final_kernel<<<fgrid, kBlock, 0, s>>>(a, o, 0, lo, st.step, 0); }));
        lo = st.step;
      }
    }
  }
  DAWN_HIP(hipStreamSynchronize(s));
  return DAWN_OK;
}

}  // namespace dawn
