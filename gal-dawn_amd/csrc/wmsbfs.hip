// Batched weighted multi-source shortest paths: int32 min-plus over up to 64
// sources at once on one HIP device.
//
// Replaces the per-source weighted loops of the reference,
// DAWN::MSSP_GPU::run_Weighted (src/algorithm/gpu/mssp.cu:53-102) and
// DAWN::APSP_GPU::run_Weighted (src/algorithm/gpu/apsp.cu:9-60) (and the
// CPU MSSP, whose weighted sparse step GOVM reads distance[j] for the queue
// index, src/algorithm/cpu/sssp.cpp:157; not reproduced): the sources of a
// batch share every frontier step.  Per source the result is the reference
// SSSP's (SSSP_CPU::SSSP, sssp.cpp:23-80): the source's listed out-edges
// seed their weight in row order (a later duplicate overwrites an earlier
// one), the source itself is never relaxed, and the distances are the
// fixed point of min-plus relaxation from that state.
//
// Layout in HBM (per graph, reused across batches):
//   dist   int32[n][L]   L = 8, 16, 32 or 64 lanes (the smallest power of
//                        two >= the batch's sources, at least 8), so one
//                        vertex's distances of all lanes share cache lines
//   act    u64[2][n]     lanes of v improved in the step (step parity)
//   q      int32[2][n]   vertices with any improved lane (step parity)
// Step (push, chaotic like GOVMP): a frontier vertex u of out-degree < 16
// gets one lane, which relaxes dist[v][b] = min(dist[v][b], dist[u][b] + w)
// for each lane b active at u and each edge (edge by edge, 16 lanes at a
// time from 16-byte record loads); a heavier u gets one
// wave, split into 64/L groups of L lanes, group g taking edges g, g+64/L,
// ... and lane b of a group relaxing source lane b.  Improved lanes are
// OR-ed into act_next[v]; the first improvement appends v to the next
// queue (staged per workgroup in LDS, one global reservation per
// workgroup).  A K-step loop is captured once into a
// hipGraph; every kernel reads the step and the done flag from device
// state, so the graph is replayed until the frontier empties.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "common.hpp"

namespace dawn {

namespace {

constexpr int kWChunk = 32;  // steps per captured graph before a traversal's depth is known
constexpr int kWLight = 16;  // out-degree below which a queued vertex gets one lane, not a wave
constexpr int kWStageCap = 4096;  // next-queue entries a workgroup stages in LDS

struct WState {
  int32_t step;      // parity of the current queue = step & 1
  int32_t done;
  unsigned qlen[2];
  int32_t mode;        // DAWN_MODE_PUSH or DAWN_MODE_PULL for the current step
  int32_t pull_steps;  // steps taken in pull mode so far
};

// Pull when the frontier holds at least n / kWPullDiv vertices (policy
// FIXED / LEARNED); PUSH and PULL policies force one mode.
constexpr int64_t kWPullDiv = 16;

struct WBatch {
  int32_t nlanes;
  int32_t pad;
  int32_t src[64];
};

struct WArgs {
  int32_t n;
  int32_t policy;
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* w;
  const int32_t* in_row_ptr;  // in-edges (CSC; the CSR for undirected graphs)
  const int32_t* in_col;
  const int32_t* in_w;
  const int32_t* hc_v;        // in-edge chunks of vertices of in-degree >= heavy
  const int32_t* hc_beg;
  const int32_t* hc_end;
  int64_t num_hc;
  int32_t heavy;
  int32_t* dist;        // [n][L]
  uint64_t* act;        // [2][n]
  int32_t* q;           // [2][n]
  WState* st;
  const WBatch* batch;
};

template <int L>
__global__ void w_fill_kernel(WArgs a) {
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  const int64_t total = (int64_t)a.n * L;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gs) a.dist[i] = kInf;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * (int64_t)a.n; i += gs)
    a.act[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    WState st{};
    st.mode = a.policy == DAWN_POLICY_PULL ? DAWN_MODE_PULL : DAWN_MODE_PUSH;
    *a.st = st;
  }
}

// Reference initialisation, one thread per source lane (each owns its
// column of dist): the listed out-edges write their weight in row order,
// then dist[source] = 0; the out-neighbours other than the source form the
// first frontier.  Isolated sources keep an empty frontier (row INF except
// the source), as MSSP skips them.
template <int L>
__global__ void w_seed_kernel(WArgs a) {
  const int b = threadIdx.x;
  if (blockIdx.x != 0 || b >= a.batch->nlanes) return;
  const int32_t s = a.batch->src[b];
  const int32_t beg = a.row_ptr[s], end = a.row_ptr[s + 1];
  for (int32_t k = beg; k < end; k++) a.dist[(int64_t)a.col[k] * L + b] = a.w[k];
  a.dist[(int64_t)s * L + b] = 0;
  for (int32_t k = beg; k < end; k++) {
    const int32_t v = a.col[k];
    if (v == s) continue;
    const unsigned long long old = atomicOr((unsigned long long*)&a.act[v], 1ull << b);
    if (old == 0) a.q[atomicAdd(&a.st->qlen[0], 1u)] = v;
  }
}

template <int L>
__global__ void __launch_bounds__(kBlock) w_step_kernel(WArgs a) {
  constexpr int G = kWave / L;  // edge groups per wave
  const WState st = *a.st;
  if (st.done || st.mode != DAWN_MODE_PUSH) return;
  const int p = st.step & 1;
  const unsigned qn = st.qlen[p];
  const int32_t* q_in = a.q + (int64_t)p * a.n;
  int32_t* q_out = a.q + (int64_t)(p ^ 1) * a.n;
  uint64_t* act_in = a.act + (int64_t)p * a.n;
  uint64_t* act_out = a.act + (int64_t)(p ^ 1) * a.n;
  unsigned* qlen_out = &a.st->qlen[p ^ 1];
  __shared__ LdsQueue<kWStageCap> s_q;
  if (threadIdx.x == 0) s_q.init();
  __syncthreads();
  const int lane = lane_id();
  const int b = lane % L;
  const int grp = lane / L;
  const uint64_t gmask = L == 64 ? ~0ull : ((1ull << L) - 1);
  // Light vertices (out-degree < kWLight): one lane per queued vertex -- on
  // sparse graphs a vertex is usually active in few lanes, so a wave per
  // vertex would leave most lanes idle.  The lane takes u's edges in turn;
  // per edge it works through the record in chunks of CH = min(L, 16)
  // source lanes that hold an active lane of u: 16-byte loads of u's and
  // the neighbour's distances (u's held in registers when L <= 16), atomics
  // only for the lanes that improve, and one act_next OR per edge.  The
  // wave loops until its busiest lane is done, so every lane of the wave
  // takes part in each (LDS-staged) append.
  {
    constexpr int CH = L < 16 ? L : 16;
    constexpr int NC = L / CH;
    constexpr uint64_t kChunk = CH == 64 ? ~0ull : ((1ull << CH) - 1);
    auto load_chunk = [](const int32_t* p, int32_t (&x)[CH]) {
#pragma unroll
      for (int t = 0; t < CH / 4; t++) {
        const int4 y = reinterpret_cast<const int4*>(p)[t];
        x[4 * t] = y.x;
        x[4 * t + 1] = y.y;
        x[4 * t + 2] = y.z;
        x[4 * t + 3] = y.w;
      }
    };
    for (int64_t base = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < qn;
         base += (int64_t)gridDim.x * kBlock) {  // wave-uniform
      const int64_t i = base + lane;
      int32_t u = 0;
      uint64_t m = 0;
      int beg = 0, deg = 0;
      if (i < qn) {
        u = q_in[i];
        beg = a.row_ptr[u];
        deg = a.row_ptr[u + 1] - beg;
        if (deg < kWLight) m = act_in[u];
      }
      const bool light = i < qn && deg < kWLight;
      const int most = wave_max_i32(m ? deg : 0);
      const int32_t* up = a.dist + (int64_t)u * L;
      int32_t du[CH];
      if (NC == 1 && m) load_chunk(up, du);
      for (int e = 0; e < most; e++) {
        bool append = false;
        int32_t v = -1;
        if (m && e < deg) {
          v = a.col[beg + e];
          const int32_t wt = a.w[beg + e];
          int32_t* vp = a.dist + (int64_t)v * L;
          uint64_t imp = 0;
#pragma unroll 1
          for (int c = 0; c < NC; c++) {
            const uint64_t mc = (m >> (c * CH)) & kChunk;
            if (!mc) continue;
            if (NC > 1) load_chunk(up + c * CH, du);
            int32_t dv[CH];
            load_chunk(vp + c * CH, dv);
#pragma unroll
            for (int bl = 0; bl < CH; bl++) {
              if (((mc >> bl) & 1ull) && du[bl] != kInf) {
                const int32_t nd = minplus(du[bl], wt);
                if (nd < dv[bl] && nd < atomicMin(vp + c * CH + bl, nd)) imp |= 1ull << (c * CH + bl);
              }
            }
          }
          if (imp) append = atomicOr((unsigned long long*)&act_out[v], (unsigned long long)imp) == 0ull;
        }
        s_q.append(append, v, qlen_out, q_out);
      }
      if (light) act_in[u] = 0ull;  // consumed
    }
  }
  // Heavy vertices: one wave each, split into groups of L lanes.
  const int64_t wave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t i = wave; i < qn; i += nwaves) {
    const int32_t u = q_in[i];
    if (a.row_ptr[u + 1] - a.row_ptr[u] < kWLight) continue;
    const uint64_t mask = act_in[u];
    const bool active = (mask >> b) & 1ull;
    const int32_t du = active ? a.dist[(int64_t)u * L + b] : kInf;
    const int beg = a.row_ptr[u], end = a.row_ptr[u + 1];
    for (int e0 = beg; e0 < end; e0 += G) {
      const int e = e0 + grp;
      bool improved = false;
      int32_t v = -1;
      if (e < end) {
        v = a.col[e];
        if (active && du != kInf) {
          const int32_t nd = minplus(du, a.w[e]);
          int32_t* dv = a.dist + (int64_t)v * L + b;
          if (nd < *dv) improved = nd < atomicMin(dv, nd);
        }
      }
      const uint64_t bal = __ballot(improved);
      const uint64_t mine = (bal >> (grp * L)) & gmask;
      // One atomic per group and edge; the first improvement of v in this
      // step appends it to the next queue.
      bool append = false;
      if (b == 0 && mine) append = atomicOr((unsigned long long*)&act_out[v], mine) == 0ull;
      s_q.append(append, v, qlen_out, q_out);
    }
    if (lane == 0) act_in[u] = 0ull;  // consumed (each u is in the queue once)
  }
  s_q.flush(qlen_out, q_out);
}

// Pull step, for large frontiers: every vertex v takes, per source lane b,
// the minimum of dist[u][b] + w over its in-edges (u, w) whose tail
// improved lane b in the last step, and stores the lanes that improve with
// plain stores (v's record is written by v's thread only; a concurrent
// reader of it sees the old or the new value of each lane, both path
// lengths, so the fixed point is unchanged).  Vertices of in-degree >=
// heavy are done by in-edge chunks, a wave each: per lane minima reduced
// over the wave, merged with atomicMin (a vertex's chunks may run in
// different waves).  Light vertices take 16 source lanes at a time from
// 16-byte record loads.
// In-edges per round of a light vertex in the batched pull step.
constexpr int kPullRound = 4;

template <int L>
__global__ void __launch_bounds__(kBlock) w_pull_kernel(WArgs a) {
  constexpr int CH = L < 16 ? L : 16;
  constexpr int NC = L / CH;
  constexpr uint64_t kChunk = CH == 64 ? ~0ull : ((1ull << CH) - 1);
  const WState st = *a.st;
  if (st.done || st.mode != DAWN_MODE_PULL) return;
  const int p = st.step & 1;
  const uint64_t* act_in = a.act + (int64_t)p * a.n;
  uint64_t* act_out = a.act + (int64_t)(p ^ 1) * a.n;
  int32_t* q_out = a.q + (int64_t)(p ^ 1) * a.n;
  unsigned* qlen_out = &a.st->qlen[p ^ 1];
  __shared__ LdsQueue<kWStageCap> s_q;
  if (threadIdx.x == 0) s_q.init();
  __syncthreads();
  const int lane = lane_id();
  auto load_chunk = [](const int32_t* q, int32_t (&x)[CH]) {
#pragma unroll
    for (int t = 0; t < CH / 4; t++) {
      const int4 y = reinterpret_cast<const int4*>(q)[t];
      x[4 * t] = y.x;
      x[4 * t + 1] = y.y;
      x[4 * t + 2] = y.z;
      x[4 * t + 3] = y.w;
    }
  };
  for (int64_t base = (int64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < a.n;
       base += (int64_t)gridDim.x * kBlock) {  // wave-uniform
    const int64_t vv = base + lane;
    const int32_t v = vv < a.n ? (int32_t)vv : 0;
    int beg = 0, deg = 0;
    if (vv < a.n) {
      beg = a.in_row_ptr[v];
      deg = a.in_row_ptr[v + 1] - beg;
    }
    uint64_t imp = 0;
    if (vv < a.n && deg > 0 && deg < a.heavy) {
      int32_t* vp = a.dist + (int64_t)v * L;
#pragma unroll 1
      for (int c = 0; c < NC; c++) {
        int32_t best[CH];
#pragma unroll
        for (int bl = 0; bl < CH; bl++) best[bl] = kInf;
        bool any = false;
        // Rounds of kPullRound in-edges: their columns and weights, then
        // their tails' masks, then the improved tails' records, so that a
        // lane waits on memory once per stage and round, not per edge.
        for (int j0 = 0; j0 < deg; j0 += kPullRound) {
          int32_t u[kPullRound], wt[kPullRound];
#pragma unroll
          for (int t = 0; t < kPullRound; t++) {
            const bool ok = j0 + t < deg;
            u[t] = ok ? a.in_col[beg + j0 + t] : -1;
            wt[t] = ok ? a.in_w[beg + j0 + t] : 0;
          }
          uint64_t m[kPullRound];
#pragma unroll
          for (int t = 0; t < kPullRound; t++) m[t] = u[t] >= 0 ? (act_in[u[t]] >> (c * CH)) & kChunk : 0ull;
#pragma unroll
          for (int t = 0; t < kPullRound; t++) {
            if (!m[t]) continue;
            any = true;
            int32_t du[CH];
            load_chunk(a.dist + (int64_t)u[t] * L + c * CH, du);
#pragma unroll
            for (int bl = 0; bl < CH; bl++)
              if (((m[t] >> bl) & 1ull) && du[bl] != kInf) best[bl] = min(best[bl], minplus(du[bl], wt[t]));
          }
        }
        if (!any) continue;
        int32_t dv[CH];
        load_chunk(vp + c * CH, dv);
        uint64_t ci = 0;
#pragma unroll
        for (int bl = 0; bl < CH; bl++)
          if (best[bl] < dv[bl]) {
            dv[bl] = best[bl];
            ci |= 1ull << bl;
          }
        if (ci) {
#pragma unroll
          for (int t = 0; t < CH / 4; t++)
            reinterpret_cast<int4*>(vp + c * CH)[t] =
                make_int4(dv[4 * t], dv[4 * t + 1], dv[4 * t + 2], dv[4 * t + 3]);
          imp |= ci << (c * CH);
        }
      }
      if (imp) act_out[v] = imp;  // v's own slot (heavy vertices are not here)
    }
    s_q.append(imp != 0, v, qlen_out, q_out);
  }
  // Heavy vertices by in-edge chunks, one wave each.
  const int64_t wave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t hc = wave; hc < a.num_hc; hc += nwaves) {  // wave-uniform
    const int32_t v = a.hc_v[hc];
    const int kb = a.hc_beg[hc], ke = a.hc_end[hc];
    uint64_t vimp = 0;
#pragma unroll 1
    for (int c = 0; c < NC; c++) {
      int32_t best[CH];
#pragma unroll
      for (int bl = 0; bl < CH; bl++) best[bl] = kInf;
      for (int k = kb + lane; k < ke; k += kWave) {
        const int32_t u = a.in_col[k];
        const uint64_t m = (act_in[u] >> (c * CH)) & kChunk;
        if (!m) continue;
        const int32_t wt = a.in_w[k];
        int32_t du[CH];
        load_chunk(a.dist + (int64_t)u * L + c * CH, du);
#pragma unroll
        for (int bl = 0; bl < CH; bl++)
          if (((m >> bl) & 1ull) && du[bl] != kInf) best[bl] = min(best[bl], minplus(du[bl], wt));
      }
      int32_t mine = kInf;
#pragma unroll
      for (int bl = 0; bl < CH; bl++) {
        const int32_t x = wave_min_i32(best[bl]);
        if (lane == bl) mine = x;
      }
      bool improved = false;
      if (lane < CH && mine != kInf) {
        int32_t* dp = a.dist + (int64_t)v * L + c * CH + lane;
        if (mine < *dp) improved = mine < atomicMin(dp, mine);
      }
      vimp |= (__ballot(improved) & kChunk) << (c * CH);
    }
    bool append = false;
    if (lane == 0 && vimp) append = atomicOr((unsigned long long*)&act_out[v], (unsigned long long)vimp) == 0ull;
    s_q.append(append, v, qlen_out, q_out);
  }
  s_q.flush(qlen_out, q_out);
}

// After a pull step: clears the consumed frontier masks (a push step clears
// each as it consumes it; a pull step reads them from many threads).
__global__ void w_clear_kernel(WArgs a) {
  const WState st = *a.st;
  if (st.done || st.mode != DAWN_MODE_PULL) return;
  const int p = st.step & 1;
  const unsigned qn = st.qlen[p];
  const int32_t* q_in = a.q + (int64_t)p * a.n;
  uint64_t* act_in = a.act + (int64_t)p * a.n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < qn;
       i += (int64_t)gridDim.x * blockDim.x)
    act_in[q_in[i]] = 0ull;
}

// Ends a step: the next queue becomes current, and the next step's mode is
// chosen from its size.
__global__ void w_advance_kernel(WState* st, int32_t n, int32_t policy) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->done) return;
  const int p = st->step & 1;
  if (st->mode == DAWN_MODE_PULL) st->pull_steps += 1;
  st->qlen[p] = 0;
  st->step += 1;
  const unsigned nq = st->qlen[p ^ 1];
  if (nq == 0) {
    st->done = 1;
    return;
  }
  int mode = DAWN_MODE_PUSH;
  if (policy == DAWN_POLICY_PULL) mode = DAWN_MODE_PULL;
  else if (policy != DAWN_POLICY_PUSH && (int64_t)nq * kWPullDiv >= (int64_t)n) mode = DAWN_MODE_PULL;
  st->mode = mode;
}

// Distance rows (lane b at out + b * stride) and per-lane sums: a workgroup
// stages 256 vertices x L lanes in LDS (one contiguous read), then each
// wave writes its 64 vertices of every lane as one coalesced row segment.
template <int L>
// acc: [0..63] distance sums, [64..127] reached vertices, [128..191] sum of
// out-degrees of reached vertices, per lane.
__global__ void __launch_bounds__(kBlock) w_final_kernel(WArgs a, int32_t* out, int64_t stride,
                                                         unsigned long long* acc) {
  __shared__ int32_t s_tile[kBlock * L];
  __shared__ unsigned long long s_acc[3 * 64];
  const int nl = a.batch->nlanes;
  if (threadIdx.x < 3 * 64) s_acc[threadIdx.x] = 0;
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < a.n; b0 += (int64_t)gridDim.x * kBlock) {
    const int64_t cnt = std::min<int64_t>(kBlock, a.n - b0);  // block-uniform
    const int32_t* src = a.dist + b0 * L;
    __syncthreads();
    for (int64_t i = threadIdx.x; i < cnt * L; i += kBlock) s_tile[i] = src[i];
    __syncthreads();
    const int64_t v = b0 + wv * kWave + lane;
    const bool ok = v < a.n;
    const unsigned deg = ok ? (unsigned)(a.row_ptr[v + 1] - a.row_ptr[v]) : 0u;
    for (int bl = 0; bl < nl; bl++) {
      const int32_t d = ok ? s_tile[(wv * kWave + lane) * L + bl] : kInf;
      if (out && ok) out[(int64_t)bl * stride + v] = d;
      const bool r = d != kInf;
      const unsigned long long sm = wave_sum_u64(r ? (unsigned long long)d : 0ull);
      const unsigned long long em = wave_sum_u64(r ? (unsigned long long)deg : 0ull);
      const unsigned long long rc = (unsigned long long)__popcll(__ballot(r));
      if (lane == 0) {
        if (sm) atomicAdd(&s_acc[bl], sm);
        if (rc) atomicAdd(&s_acc[64 + bl], rc);
        if (em) atomicAdd(&s_acc[128 + bl], em);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * 64; i += kBlock)
    if ((i & 63) < nl && s_acc[i]) atomicAdd(&acc[i], s_acc[i]);
}

__global__ void w_zero_kernel(unsigned long long* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

}  // namespace

struct WmsWorkspace {
  int32_t n = 0;
  int32_t lanes = 0;
  DevBuf<int32_t> dist, q, rows;
  DevBuf<uint64_t> act;
  DevBuf<WState> st;
  DevBuf<WBatch> batch;
  DevBuf<unsigned long long> sums;
  WState* h_state = nullptr;
  WBatch* h_batch = nullptr;
  std::map<int, hipGraphExec_t> graphs;  // by (L, policy, steps per graph)
  int32_t last_steps = 0;                // steps the last batch took
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

void wms_invalidate_graphs(WmsWorkspace* ws) {
  if (!ws) return;
  for (auto& kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
  ws->graphs.clear();
}

void wms_workspace_free(WmsWorkspace* ws) {
  if (!ws) return;
  for (auto& kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
  if (ws->h_state) (void)hipHostFree(ws->h_state);
  if (ws->h_batch) (void)hipHostFree(ws->h_batch);
  for (hipEvent_t e : {ws->ev0, ws->ev1})
    if (e) (void)hipEventDestroy(e);
  delete ws;
}

namespace {

int ensure_wms(dawn_graph* g, int L) {
  if (!g->wms) g->wms = new WmsWorkspace();
  WmsWorkspace* ws = g->wms;
  if (ws->n == g->n && ws->lanes == L && ws->dist.p) return DAWN_OK;
  for (auto& kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
  ws->graphs.clear();
  ws->dist.release();
  ws->rows.release();
  ws->n = g->n;
  ws->lanes = L;
  DAWN_TRY(ws->dist.alloc((size_t)g->n * L));
  DAWN_TRY(ws->act.alloc((size_t)2 * g->n));
  DAWN_TRY(ws->q.alloc((size_t)2 * g->n));
  DAWN_TRY(ws->st.alloc(1));
  DAWN_TRY(ws->batch.alloc(1));
  DAWN_TRY(ws->sums.alloc(3 * 64));
  if (!ws->h_state) DAWN_HIP(hipHostMalloc((void**)&ws->h_state, sizeof(WState)));
  if (!ws->h_batch) DAWN_HIP(hipHostMalloc((void**)&ws->h_batch, sizeof(WBatch)));
  for (hipEvent_t* e : {&ws->ev0, &ws->ev1})
    if (!*e) DAWN_HIP(hipEventCreate(e));
  return DAWN_OK;
}

template <int L>
int run_w(dawn_graph* g, const int32_t* sources, int32_t S, int32_t* dist, int64_t* sums,
          int32_t policy, dawn_stats* stats) {
  DAWN_TRY(ensure_wms(g, L));
  WmsWorkspace* ws = g->wms;
  hipStream_t s = g->stream;
  const bool dist_dev = dist && is_device_pointer(dist);
  if (dist && !dist_dev) DAWN_TRY(ws->rows.alloc((size_t)L * g->n));
  WArgs a{};
  a.n = g->n;
  a.policy = policy;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.w = g->weight;
  a.in_row_ptr = g->in_row_ptr;
  a.in_col = g->in_col;
  a.in_w = g->in_weight;
  a.hc_v = g->hc_v;
  a.hc_beg = g->hc_beg;
  a.hc_end = g->hc_end;
  a.num_hc = g->num_hc;
  a.heavy = g->knobs.heavy_degree;
  a.dist = ws->dist.p;
  a.act = ws->act.p;
  a.q = ws->q.p;
  a.st = ws->st.p;
  a.batch = ws->batch.p;
  const unsigned fill_grid = grid_for((int64_t)g->n * L);
  const unsigned step_grid = 2048;  // 8 workgroups per CU, waves stride the queue
  // (1024-8192 workgroups measured within noise of each other.)
  const unsigned pull_grid = grid_for(std::max<int32_t>(g->n, 1), kBlock, 4096);
  const unsigned fin_grid = grid_for(std::max<int32_t>(g->n, 1), kBlock, 2048);

  // Steps per captured graph: what the last batch took (+ the one that
  // finds the frontier empty), rounded up to a multiple of 8, so that one
  // graph launch and one read of the done flag usually suffice.
  const int chunk = ws->last_steps > 0 ? std::min(256, (ws->last_steps + 1 + 7) & ~7) : kWChunk;
  const int key = (chunk * 128 + L) * 8 + (policy & 7);
  hipGraphExec_t exec = nullptr;
  auto it = ws->graphs.find(key);
  if (it != ws->graphs.end()) {
    exec = it->second;
  } else {
    hipGraph_t graph;
    DAWN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk; i++) {
      w_step_kernel<L><<<step_grid, kBlock, 0, s>>>(a);
      w_pull_kernel<L><<<pull_grid, kBlock, 0, s>>>(a);
      w_clear_kernel<<<step_grid, kBlock, 0, s>>>(a);
      w_advance_kernel<<<1, 1, 0, s>>>(a.st, a.n, policy);
    }
    DAWN_HIP(hipStreamEndCapture(s, &graph));
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DAWN_HIP(e);
    ws->graphs[key] = exec;
  }

  float total_ms = 0.f;
  std::vector<unsigned long long> h_sums(3 * 64);
  for (int32_t b0 = 0; b0 < S; b0 += L) {
    const int32_t lanes = std::min<int32_t>(L, S - b0);
    WBatch& hb = *ws->h_batch;
    std::memset(&hb, 0, sizeof hb);
    hb.nlanes = lanes;
    for (int i = 0; i < lanes; i++) {
      int64_t v = sources[b0 + i] % (int64_t)g->n;
      if (v < 0) v += g->n;
      hb.src[i] = (int32_t)v;
      if (stats && g->h_row_ptr[v] == g->h_row_ptr[v + 1]) stats->skipped_sources += 1;
    }
    DAWN_HIP(hipMemcpyAsync(ws->batch.p, ws->h_batch, sizeof(WBatch), hipMemcpyHostToDevice, s));
    DAWN_HIP(hipEventRecord(ws->ev0, s));
    w_fill_kernel<L><<<fill_grid, kBlock, 0, s>>>(a);
    w_seed_kernel<L><<<1, 64, 0, s>>>(a);
    int64_t launched = 0;
    for (;;) {
      DAWN_HIP(hipGraphLaunch(exec, s));
      launched += chunk;
      DAWN_HIP(hipMemcpyAsync(ws->h_state, ws->st.p, sizeof(WState), hipMemcpyDeviceToHost, s));
      DAWN_HIP(hipStreamSynchronize(s));
      if (ws->h_state->done) break;
      if (launched > (int64_t)g->n + 2 * chunk) return DAWN_ERR_HIP;
    }
    int32_t* out = dist ? (dist_dev ? dist + (int64_t)b0 * g->n : ws->rows.p) : nullptr;
    w_zero_kernel<<<1, 64, 0, s>>>(ws->sums.p, 3 * 64);
    w_final_kernel<L><<<fin_grid, kBlock, 0, s>>>(a, out, g->n, ws->sums.p);
    DAWN_HIP(hipGetLastError());
    DAWN_HIP(hipEventRecord(ws->ev1, s));
    DAWN_HIP(hipMemcpyAsync(h_sums.data(), ws->sums.p, sizeof(unsigned long long) * 3 * 64,
                            hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    DAWN_HIP(hipEventElapsedTime(&ms, ws->ev0, ws->ev1));
    total_ms += ms;
    if (dist && !dist_dev)
      DAWN_HIP(hipMemcpy(dist + (int64_t)b0 * g->n, ws->rows.p, sizeof(int32_t) * (size_t)lanes * g->n,
                         hipMemcpyDeviceToHost));
    for (int i = 0; i < lanes; i++)
      if (sums) sums[b0 + i] = (int64_t)h_sums[i];
    ws->last_steps = ws->h_state->step;
    if (stats) {
      for (int i = 0; i < lanes; i++) {
        stats->vertices_reached += (int64_t)h_sums[64 + i];
        stats->edges_reached += (int64_t)h_sums[128 + i];
      }
      stats->batches += 1;
      stats->steps += ws->h_state->step;
      stats->mode_steps[DAWN_MODE_PUSH] += ws->h_state->step - ws->h_state->pull_steps;
      stats->mode_steps[DAWN_MODE_PULL] += ws->h_state->pull_steps;
    }
  }
  if (stats) stats->device_ms += total_ms;
  return DAWN_OK;
}

}  // namespace

// Weighted multi-source traversal, batches of up to 64 sources.
int run_wmsbfs(dawn_graph* g, const int32_t* sources, int32_t S, int32_t* dist, int64_t* sums,
               int32_t policy, dawn_stats* stats) {
  if (!g || !g->weighted || S < 0 || (S > 0 && !sources) || g->n <= 0) return DAWN_ERR_INVALID_ARG;
  if (S == 0) return DAWN_OK;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  const int32_t per = std::min<int32_t>(S, 64);
  if (per <= 8) return run_w<8>(g, sources, S, dist, sums, policy, stats);
  if (per <= 16) return run_w<16>(g, sources, S, dist, sums, policy, stats);
  if (per <= 32) return run_w<32>(g, sources, S, dist, sums, policy, stats);
  return run_w<64>(g, sources, S, dist, sums, policy, stats);
}

}  // namespace dawn
