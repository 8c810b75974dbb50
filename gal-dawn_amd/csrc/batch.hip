// Many independent graphs of uneven size in one persistent launch.
//
// Replaces, for a batch of graphs, the reference's host loop of per-source
// launches (DAWN::MSSP_GPU::run, src/algorithm/gpu/mssp.cu:26-43 and
// run_Weighted :53-102): here every (graph, source block) is a task in a
// device work queue, and each workgroup pops tasks with one atomic and runs
// the whole traversal inside the workgroup (barriers are workgroup-local, so
// no grid-wide synchronisation exists and every workgroup exits once the
// task counter passes the last task).
//
// Unweighted task: all sources of one graph (or a share of them on the
// largest graphs) as lanes of one word per vertex, as narrow as the lanes
// allow (<= 64 lanes; 8 sources: a byte), sparse level-synchronous BOVM:
//   Q_l = [(v, lanes that reached v at level l)]
//   expand: for (u, f) in Q_l, w in N(u): nv = f & ~seen[w]; if nv,
//           atomicOr(next[w], nv); the first writer appends w to `touched`
//   commit: for w in touched: nw = next[w] & ~seen[w]; next[w] = 0;
//           seen[w] |= nw; dist[b][w] = l+1 for b in nw; Q_{l+1} += (w, nw)
// Weighted task: one source, int32 min-plus (the fixed point of GOVM,
// src/algorithm/cpu/sssp.cpp:141-199, from the reference's seeding), by
// near-far relaxation: vertices whose distance dropped are relaxed next if
// it is below a threshold (near), else set aside (far); when no near vertex
// is left the threshold moves past the smallest far distance by delta and
// the far pile is split.  Relaxing in distance order re-relaxes far fewer
// edges than plain frontier relaxation (CPU count on the 4.1M-edge ER graph
// of the uneven workload: 5.7 -> 1.2 relaxations per edge); atomicMin
// relaxations in place in the output row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"

namespace dawn {

struct BatchGraph {
  int32_t n;
  int32_t weighted;
  int64_t m;
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* w;
  int64_t dist_off;  // first output row of this graph
};

struct BatchTask {
  int32_t graph;
  int32_t lane0;   // first source index within the graph's block
  int32_t lanes;   // sources in this task (1 for weighted)
  int32_t pad;
};

struct BatchArgs {
  const BatchGraph* graphs;
  const BatchTask* tasks;
  int32_t num_tasks;
  int32_t spg;               // sources per graph
  const int32_t* sources;    // num_graphs * spg (already reduced mod n)
  int32_t* dist;
  unsigned* task_counter;
  unsigned long long* levels_total;
  const unsigned long long* wsum;  // per graph: sum of edge weights (weighted)
  // Per-workgroup scratch, sized for the largest graph (max_n).
  int64_t max_n;
  uint64_t* seen;      // [grid][max_n]
  uint64_t* next;      // [grid][max_n]
  int32_t* qv;         // [grid][2][max_n]
  uint64_t* qw;        // [grid][2][max_n]
  int32_t* touched;    // [grid][max_n]
  uint16_t* stage;     // [grid][max_n][8]: arrival levels of tasks of <= 8 lanes
};

namespace {

// Threads per workgroup of the persistent kernel: a task (one graph's
// traversal) runs inside one workgroup, so the largest graphs' tasks are
// the critical path of a batch, and wider workgroups shorten it (measured
// on 512 graphs of 1k-130k vertices: 1024 threads 30 ms unweighted / 85 ms
// weighted; 256 threads, two or eight per CU: 74 ms / 183-203 ms).  One
// such workgroup is resident per CU (the kernel's scalar registers allow
// seven waves per SIMD); forcing two per CU (eight waves, scalar registers
// spilled) made the batches slower (18.7 -> 24-27 ms unweighted, 46 -> 51 ms
// weighted): twice the concurrent tasks' state no longer fits the caches.
constexpr int kBatchBlock = 1024;

template <int BS>
__device__ __forceinline__ int owner(const int* off, int e) {
  int lo = 0, hi = BS - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Exclusive scan of one int per thread over a BS-thread workgroup.
// scratch: BS/kWave ints of LDS; *total receives the sum.
template <int BS>
__device__ __forceinline__ int scan_bs(int x, int* scratch, int* total) {
  const int l = lane_id();
  const int w = threadIdx.x / kWave;
  int inc = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int y = __shfl_up(inc, o);
    if (l >= o) inc += y;
  }
  if (l == kWave - 1) scratch[w] = inc;
  __syncthreads();
  int base = 0, sum = 0;
#pragma unroll
  for (int i = 0; i < BS / kWave; i++) {
    const int s = scratch[i];
    if (i < w) base += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return base + inc - x;
}

// Appends through an LDS counter (one workgroup owns the list).
__device__ __forceinline__ int lds_append(unsigned* s_count, bool pred) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return -1;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned base = 0;
  if (lane_id() == leader) base = atomicAdd(s_count, (unsigned)__popcll(m));
  base = __shfl(base, leader);
  return pred ? (int)(base + (unsigned)__popcll(m & lanes_below_mask())) : -1;
}

// Stage entry of a level that does not fit 16 bits (written to its row).
constexpr uint16_t kStaged = 0xFFFF;

// atomicOr on a lane word of type LT; sub-word types go through the
// aligned 32-bit word that holds them.  Returns the old value of *p.
template <typename LT>
__device__ __forceinline__ LT atomic_or_lanes(LT* p, LT v) {
  if constexpr (sizeof(LT) == 8) {
    return (LT)atomicOr((unsigned long long*)p, (unsigned long long)v);
  } else if constexpr (sizeof(LT) == 4) {
    return (LT)atomicOr((unsigned*)p, (unsigned)v);
  } else {
    const uintptr_t ad = (uintptr_t)p;
    const int sh = (int)(ad & 3) * 8;
    const unsigned old = atomicOr((unsigned*)(ad & ~(uintptr_t)3), (unsigned)v << sh);
    return (LT)(old >> sh);
  }
}

// LT: the narrowest unsigned type holding the task's lanes (8 sources per
// graph: one byte per vertex).  The per-vertex state (seen, next, queue
// words) is probed and OR-ed at random per scanned edge, so its footprint
// decides how much of it the caches hold: at 8 bytes per vertex the
// workgroups' state spilled out of L2 and the Infinity Cache, and the
// kernel moved about 100 GB for 512 graphs of 14 M vertices.
template <int BS, typename LT>
__device__ void run_unweighted(const BatchArgs& a, const BatchGraph& G, const BatchTask& T,
                               LT* seen, LT* next, int32_t* qv, LT* qw,
                               int32_t* touched, int* s_off, int* s_beg, uint64_t* s_f,
                               int* s_scan, unsigned* s_cnt) {
  const int32_t n = G.n;
  const int64_t mn = a.max_n;
  // Lane b of the task is source lane0 + b of the graph: row drow + b * n.
  int32_t* drow = a.dist + G.dist_off + (int64_t)T.lane0 * n;
  const LT lanemask = (LT)(T.lanes >= 64 ? ~0ull : ((1ull << T.lanes) - 1));
  // Tasks of at most 8 lanes stage each arrival level in stage[v][b]
  // (16 bits; kStaged marks a level that did not fit, written to its row at
  // once) and write the rows in one coalesced pass at the end, instead of
  // 4-byte stores scattered over the rows as vertices are reached.
  constexpr bool kStage = sizeof(LT) == 1;
  uint16_t* stage = a.stage + (int64_t)blockIdx.x * mn * 8;
  for (int32_t v = threadIdx.x; v < n; v += blockDim.x) {
    seen[v] = (LT)~lanemask;
    next[v] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s_cnt[0] = 0;  // |Q_0|
    for (int i = 0; i < T.lanes; i++) {
      const int32_t s = a.sources[(int64_t)T.graph * a.spg + T.lane0 + i];
      const LT bit = (LT)(1ull << i);
      if ((seen[s] & lanemask) == 0) {
        qv[s_cnt[0]] = s;
        qw[s_cnt[0]] = 0;
        s_cnt[0]++;
      }
      seen[s] |= bit;
      if constexpr (kStage) stage[(int64_t)s * 8 + i] = 0;
      else drow[(int64_t)i * n + s] = 0;
    }
    // Words of the level-0 entries.
    for (unsigned j = 0; j < s_cnt[0]; j++) qw[j] = seen[qv[j]] & lanemask;
  }
  __syncthreads();
  unsigned qn = s_cnt[0];
  int par = 0;
  int level = 0;
  while (qn > 0) {
    const int32_t* qv_in = qv + (int64_t)par * mn;
    const LT* qw_in = qw + (int64_t)par * mn;
    int32_t* qv_out = qv + (int64_t)(par ^ 1) * mn;
    LT* qw_out = qw + (int64_t)(par ^ 1) * mn;
    if (threadIdx.x == 0) s_cnt[1] = 0;  // |touched|
    __syncthreads();
    // Expand Q_level with workgroup edge balancing (LDS degree scan).
    for (unsigned c0 = 0; c0 < qn; c0 += BS) {
      const unsigned i = c0 + threadIdx.x;
      int beg = 0, deg = 0;
      uint64_t f = 0;
      if (i < qn) {
        const int32_t u = qv_in[i];
        beg = G.row_ptr[u];
        deg = G.row_ptr[u + 1] - beg;
        f = qw_in[i];
      }
      int total;
      const int off = scan_bs<BS>(deg, s_scan, &total);
      s_off[threadIdx.x] = off;
      s_beg[threadIdx.x] = beg;
      s_f[threadIdx.x] = f;
      __syncthreads();
      for (int e0 = 0; e0 < total; e0 += BS) {
        const int e = e0 + threadIdx.x;
        bool first = false;
        int32_t w = 0;
        if (e < total) {
          const int j = owner<BS>(s_off, e);
          w = G.col[s_beg[j] + (e - s_off[j])];
          const LT nv = (LT)(s_f[j] & ~(uint64_t)seen[w]);
          if (nv) first = atomic_or_lanes<LT>(&next[w], nv) == 0;
        }
        const int slot = lds_append(&s_cnt[1], first);
        if (first) touched[slot] = w;
      }
      __syncthreads();
    }
    // Commit the touched vertices into Q_{level+1} and the distance rows.
    if (threadIdx.x == 0) s_cnt[2] = 0;
    __syncthreads();
    const unsigned tn = s_cnt[1];
    const int32_t lv = level + 1;
    for (unsigned t0 = 0; t0 < tn; t0 += BS) {
      const unsigned t = t0 + threadIdx.x;
      bool isf = false;
      int32_t w = 0;
      LT nw = 0;
      if (t < tn) {
        w = touched[t];
        const LT nx = next[w];
        next[w] = 0;
        nw = (LT)(nx & ~seen[w]);
        if (nw) {
          seen[w] = (LT)(seen[w] | nw);
          isf = true;
          uint64_t bits = (uint64_t)nw;
          while (bits) {
            const int b = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            if constexpr (kStage) {
              if (lv < (int32_t)kStaged) {
                stage[(int64_t)w * 8 + b] = (uint16_t)lv;
                continue;
              }
              stage[(int64_t)w * 8 + b] = kStaged;
            }
            drow[(int64_t)b * n + w] = lv;
          }
        }
      }
      const int slot = lds_append(&s_cnt[2], isf);
      if (isf) {
        qv_out[slot] = w;
        qw_out[slot] = nw;
      }
    }
    __syncthreads();
    qn = s_cnt[2];
    par ^= 1;
    level = lv;
    __syncthreads();
  }
  // Rows: staged levels, INF for lanes that never arrived (one thread per
  // vertex, so each row store of a wave covers 64 consecutive vertices).
  for (int32_t v = threadIdx.x; v < n; v += blockDim.x) {
    const uint64_t got = (uint64_t)(LT)(seen[v] & lanemask);
    if constexpr (kStage) {
      const uint4 q = *reinterpret_cast<const uint4*>(stage + (int64_t)v * 8);
      const uint32_t hw[4] = {q.x, q.y, q.z, q.w};
      for (int b = 0; b < T.lanes; b++) {
        const uint32_t l = (hw[b >> 1] >> (16 * (b & 1))) & 0xFFFFu;
        if (!((got >> b) & 1)) drow[(int64_t)b * n + v] = kInf;
        else if (l != kStaged) drow[(int64_t)b * n + v] = (int32_t)l;
      }
    } else {
      uint64_t miss = (uint64_t)(LT)(~seen[v] & lanemask);
      while (miss) {
        const int b = __ffsll((unsigned long long)miss) - 1;
        miss &= miss - 1;
        drow[(int64_t)b * n + v] = kInf;
      }
    }
  }
  if (threadIdx.x == 0) atomicAdd(a.levels_total, (unsigned long long)level);
}

// One weighted source.  The reference's initialisation
// (src/algorithm/cpu/sssp.cpp:40-46): listed out-edges of the source write
// their weight in row order, the distinct neighbours form the first
// frontier, distance[source] = 0, the source is never relaxed.  Then
// near-far relaxation to the same fixed point.  flags[v], one byte: kNear
// (v is in the near queue, cleared when v is taken out to be relaxed) and
// kFar (v is in the far pile).  A byte per vertex instead of two 32-bit
// words keeps more of the task's randomly probed state in the caches; a
// relaxation clears kNear and then reads dist[u] at the L2, so an
// improvement that found kNear still set is relaxed with its new value.
constexpr unsigned kNear = 1u, kFar = 2u;

__device__ __forceinline__ unsigned flags_or(uint8_t* f, unsigned bits) {
  const uintptr_t ad = (uintptr_t)f;
  const int sh = (int)(ad & 3) * 8;
  return (atomicOr((unsigned*)(ad & ~(uintptr_t)3), bits << sh) >> sh) & 0xFFu;
}
__device__ __forceinline__ void flags_clear(uint8_t* f, unsigned bits) {
  const uintptr_t ad = (uintptr_t)f;
  const int sh = (int)(ad & 3) * 8;
  atomicAnd((unsigned*)(ad & ~(uintptr_t)3), ~(bits << sh));
}

template <int BS>
__device__ void run_weighted(const BatchArgs& a, const BatchGraph& G, const BatchTask& T,
                             uint8_t* flags, int32_t* qv, int32_t* far,
                             int* s_off, int* s_beg, int* s_du, int* s_scan, unsigned* s_cnt,
                             int* s_red) {
  const int32_t n = G.n;
  const int64_t mn = a.max_n;
  const int32_t src = a.sources[(int64_t)T.graph * a.spg + T.lane0];
  int32_t* dist = a.dist + G.dist_off + (int64_t)T.lane0 * n;
  // delta: twice the mean edge weight over the mean degree (on the uneven
  // workload's weighted graphs, scales 1-8 measured within 10%, 2 best).
  const double avg_w = G.m > 0 ? (double)a.wsum[T.graph] / (double)G.m : 1.0;
  const double avg_deg = n > 0 ? (double)G.m / (double)n : 1.0;
  const long long delta = (long long)fmax(1.0, 2.0 * avg_w / fmax(avg_deg, 1.0));
  for (int32_t v = threadIdx.x; v < n; v += blockDim.x) {
    dist[v] = kInf;
    flags[v] = 0;
  }
  __syncthreads();
  // s_cnt: [0] near queue length, [1] next near length, [3] far length,
  // [4] next far length.
  if (threadIdx.x == 0) {
    const int32_t b = G.row_ptr[src], e = G.row_ptr[src + 1];
    for (int32_t k = b; k < e; k++) dist[G.col[k]] = G.w[k];
    dist[src] = 0;
    int32_t dmin = kInf;
    for (int32_t k = b; k < e; k++)
      if (G.col[k] != src) dmin = min(dmin, dist[G.col[k]]);
    const long long thr = (long long)dmin + delta;
    unsigned c = 0, f = 0;
    for (int32_t k = b; k < e; k++) {
      const int32_t v = G.col[k];
      if (v == src || flags[v]) continue;
      if (dist[v] < thr) {
        flags[v] = (uint8_t)kNear;
        qv[c++] = v;
      } else {
        flags[v] = (uint8_t)kFar;
        far[f++] = v;
      }
    }
    s_cnt[0] = c;
    s_cnt[3] = f;
    s_red[1] = (int)min(thr, (long long)kInf);
  }
  __syncthreads();
  unsigned qn = s_cnt[0];
  long long thr = s_red[1];
  int par = 0, fpar = 0;
  uint32_t steps = 0;
  for (;;) {
    while (qn > 0) {
      const int32_t* q_in = qv + (int64_t)par * mn;
      int32_t* q_out = qv + (int64_t)(par ^ 1) * mn;
      int32_t* f_out = far + (int64_t)fpar * mn;
      steps++;
      if (threadIdx.x == 0) s_cnt[1] = 0;
      __syncthreads();
      for (unsigned c0 = 0; c0 < qn; c0 += BS) {
        const unsigned i = c0 + threadIdx.x;
        int beg = 0, deg = 0, du = 0;
        if (i < qn) {
          const int32_t u = q_in[i];
          beg = G.row_ptr[u];
          deg = G.row_ptr[u + 1] - beg;
          flags_clear(flags + u, kNear);
          du = __hip_atomic_load(dist + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        int total;
        const int off = scan_bs<BS>(deg, s_scan, &total);
        s_off[threadIdx.x] = off;
        s_beg[threadIdx.x] = beg;
        s_du[threadIdx.x] = du;
        __syncthreads();
        for (int e0 = 0; e0 < total; e0 += BS) {
          const int e = e0 + threadIdx.x;
          bool add_near = false, add_far = false;
          int32_t v = 0;
          if (e < total) {
            const int j = owner<BS>(s_off, e);
            const int k = s_beg[j] + (e - s_off[j]);
            v = G.col[k];
            const int32_t nd = minplus(s_du[j], G.w[k]);
            if (v != src && nd < dist[v]) {
              const int32_t old = atomicMin(&dist[v], nd);
              if (nd < old) {
                if ((long long)nd < thr) add_near = !(flags_or(flags + v, kNear) & kNear);
                else add_far = !(flags_or(flags + v, kFar) & kFar);
              }
            }
          }
          const int slot = lds_append(&s_cnt[1], add_near);
          if (add_near) q_out[slot] = v;
          const int fslot = lds_append(&s_cnt[3], add_far);
          if (add_far) f_out[fslot] = v;
        }
        __syncthreads();
      }
      qn = s_cnt[1];
      par ^= 1;
      __syncthreads();
    }
    const unsigned fn = s_cnt[3];
    if (fn == 0) break;
    // No near vertex left: the threshold moves past the smallest far
    // distance, and the far pile splits into near and the next far pile.
    // An entry whose distance dropped below the old threshold was relaxed
    // as a near vertex already and leaves the pile.
    const int32_t* f_in = far + (int64_t)fpar * mn;
    int32_t* f_next = far + (int64_t)(fpar ^ 1) * mn;
    int32_t* q_out = qv + (int64_t)par * mn;
    int32_t lmin = kInf;
    for (unsigned i = threadIdx.x; i < fn; i += BS) {
      const int32_t d = dist[f_in[i]];
      if ((long long)d >= thr) lmin = min(lmin, d);
    }
    lmin = min(lmin, __shfl_xor(lmin, 32));
    for (int o = 16; o >= 1; o >>= 1) lmin = min(lmin, __shfl_xor(lmin, o));
    if (threadIdx.x == 0) s_red[0] = kInf;
    __syncthreads();
    if (lane_id() == 0) atomicMin(&s_red[0], lmin);
    __syncthreads();
    const long long old_thr = thr;
    const int32_t fmin = s_red[0];
    thr = fmin == kInf ? old_thr : (long long)fmin + delta;
    if (threadIdx.x == 0) {
      s_cnt[1] = 0;
      s_cnt[4] = 0;
    }
    __syncthreads();
    for (unsigned c0 = 0; c0 < fn; c0 += BS) {
      const unsigned i = c0 + threadIdx.x;
      bool to_near = false, keep = false;
      int32_t v = 0;
      if (i < fn) {
        v = f_in[i];
        const long long d = dist[v];
        if (d < old_thr) {
          flags_clear(flags + v, kFar);  // relaxed as a near vertex already
        } else if (d < thr) {
          flags_clear(flags + v, kFar);
          to_near = !(flags_or(flags + v, kNear) & kNear);
        } else {
          keep = true;
        }
      }
      const int slot = lds_append(&s_cnt[1], to_near);
      if (to_near) q_out[slot] = v;
      const int kslot = lds_append(&s_cnt[4], keep);
      if (keep) f_next[kslot] = v;
    }
    __syncthreads();
    qn = s_cnt[1];
    if (threadIdx.x == 0) s_cnt[3] = s_cnt[4];
    fpar ^= 1;
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(a.levels_total, (unsigned long long)steps);
}

// Sum of the edge weights of each weighted graph (the near-far delta):
// kWsumParts workgroups per graph, one atomic each.
constexpr int kWsumParts = 4;
__global__ void __launch_bounds__(kBlock) wsum_kernel(const BatchGraph* graphs,
                                                      unsigned long long* out) {
  __shared__ unsigned long long s_t;
  const BatchGraph G = graphs[blockIdx.x / kWsumParts];
  const int part = blockIdx.x % kWsumParts;
  if (threadIdx.x == 0) s_t = 0ull;
  __syncthreads();
  unsigned long long t = 0;
  if (G.weighted)
    for (int64_t k = (int64_t)part * kBlock + threadIdx.x; k < G.m; k += (int64_t)kWsumParts * kBlock)
      t += (unsigned)G.w[k];
  t = wave_sum_u64(t);
  if (lane_id() == 0 && t) atomicAdd(&s_t, t);
  __syncthreads();
  if (threadIdx.x == 0 && s_t) atomicAdd(&out[blockIdx.x / kWsumParts], s_t);
}

template <int BS>
__global__ void __launch_bounds__(BS) batch_kernel(BatchArgs a) {
  __shared__ int s_off[BS];
  __shared__ int s_beg[BS];
  __shared__ uint64_t s_f[BS];
  __shared__ int s_scan[BS / kWave];
  __shared__ unsigned s_cnt[5];
  __shared__ int s_red[2];
  __shared__ int s_task;
  const int64_t mn = a.max_n;
  uint64_t* seen = a.seen + (int64_t)blockIdx.x * mn;
  uint64_t* next = a.next + (int64_t)blockIdx.x * mn;
  int32_t* qv = a.qv + (int64_t)blockIdx.x * 2 * mn;
  uint64_t* qw = a.qw + (int64_t)blockIdx.x * 2 * mn;
  int32_t* touched = a.touched + (int64_t)blockIdx.x * mn;
  for (;;) {
    if (threadIdx.x == 0) s_task = (int)atomicAdd(a.task_counter, 1u);
    __syncthreads();
    const int t = s_task;
    __syncthreads();
    if (t >= a.num_tasks) break;  // every workgroup leaves once the queue is empty
    const BatchTask T = a.tasks[t];
    const BatchGraph G = a.graphs[T.graph];
    if (G.weighted)
      run_weighted<BS>(a, G, T, (uint8_t*)next, qv, (int32_t*)qw, s_off, s_beg,
                       (int*)s_f, s_scan, s_cnt, s_red);
    else if (T.lanes <= 8)
      run_unweighted<BS, uint8_t>(a, G, T, (uint8_t*)seen, (uint8_t*)next, qv, (uint8_t*)qw,
                                  touched, s_off, s_beg, s_f, s_scan, s_cnt);
    else if (T.lanes <= 16)
      run_unweighted<BS, uint16_t>(a, G, T, (uint16_t*)seen, (uint16_t*)next, qv, (uint16_t*)qw,
                                   touched, s_off, s_beg, s_f, s_scan, s_cnt);
    else if (T.lanes <= 32)
      run_unweighted<BS, uint32_t>(a, G, T, (uint32_t*)seen, (uint32_t*)next, qv, (uint32_t*)qw,
                                   touched, s_off, s_beg, s_f, s_scan, s_cnt);
    else
      run_unweighted<BS, uint64_t>(a, G, T, seen, next, qv, qw, touched, s_off, s_beg, s_f,
                                   s_scan, s_cnt);
    __syncthreads();
  }
}

}  // namespace
}  // namespace dawn

using namespace dawn;

extern "C" int dawn_batch_mssp(dawn_graph* const* graphs, int32_t num_graphs,
                               const int32_t* sources, int32_t spg, int32_t* dist,
                               const int64_t* dist_offsets, dawn_stats* stats) {
  if (!graphs || num_graphs <= 0 || !sources || spg <= 0 || spg > 64 || !dist_offsets)
    return DAWN_ERR_INVALID_ARG;
  const int device = graphs[0]->device;
  int64_t max_n = 1, total = 0;
  std::vector<BatchGraph> hg(num_graphs);
  std::vector<BatchTask> ht;
  std::vector<int32_t> hs((size_t)num_graphs * spg);
  for (int32_t i = 0; i < num_graphs; i++) {
    dawn_graph* g = graphs[i];
    if (!g || g->device != device || g->n <= 0) return DAWN_ERR_INVALID_ARG;
    max_n = std::max<int64_t>(max_n, g->n);
    hg[i] = BatchGraph{g->n, g->weighted, g->m, g->row_ptr, g->col, g->weight, dist_offsets[i]};
    total = std::max<int64_t>(total, dist_offsets[i] + (int64_t)spg * g->n);
    for (int32_t j = 0; j < spg; j++) {
      int64_t s = sources[(int64_t)i * spg + j] % (int64_t)g->n;
      if (s < 0) s += g->n;
      hs[(size_t)i * spg + j] = (int32_t)s;
      if (stats && g->h_row_ptr[s] == g->h_row_ptr[s + 1]) stats->skipped_sources += 1;
    }
    if (g->weighted) {
      for (int32_t j = 0; j < spg; j++) ht.push_back(BatchTask{i, j, 1, 0});
    } else {
      // The sources of a graph share one task as lanes of a word, except
      // on the largest graphs, whose tasks would be the batch's critical
      // path: from 500k edges two tasks of half the lanes, from 2M edges a
      // task per source (512 graphs of 1k-130k vertices, 8 sources each:
      // 31.5 -> 25.5 ms).
      const int32_t per = g->m >= 2000000 ? 1 : g->m >= 500000 ? std::max(1, spg / 2) : spg;
      for (int32_t j = 0; j < spg; j += per) ht.push_back(BatchTask{i, j, std::min(per, spg - j), 0});
    }
  }
  // Largest graphs first: the long tasks start early.
  std::stable_sort(ht.begin(), ht.end(), [&](const BatchTask& x, const BatchTask& y) {
    return hg[x.graph].n > hg[y.graph].n;
  });
  DAWN_HIP(hipSetDevice(device));
  int cus = 0;
  DAWN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  // One workgroup per CU: as many as are resident at once.
  const int grid = std::max(1, std::min<int>((int)ht.size(), cus));

  DevBuf<BatchGraph> dg;
  DevBuf<BatchTask> dt;
  DevBuf<int32_t> ds, dq, dtouch, dout;
  DevBuf<uint64_t> dseen, dnext, dqw;
  DevBuf<uint16_t> dstage;
  DevBuf<unsigned> dcnt;
  DevBuf<unsigned long long> dlev, dwsum;
  DAWN_TRY(dg.alloc(num_graphs));
  DAWN_TRY(dt.alloc(ht.size()));
  DAWN_TRY(ds.alloc(hs.size()));
  DAWN_TRY(dseen.alloc((size_t)grid * max_n));
  DAWN_TRY(dnext.alloc((size_t)grid * max_n));
  DAWN_TRY(dq.alloc((size_t)grid * 2 * max_n));
  DAWN_TRY(dqw.alloc((size_t)grid * 2 * max_n));
  DAWN_TRY(dtouch.alloc((size_t)grid * max_n));
  DAWN_TRY(dstage.alloc((size_t)grid * max_n * 8));
  DAWN_TRY(dcnt.alloc(1));
  DAWN_TRY(dlev.alloc(1));
  DAWN_TRY(dwsum.alloc(num_graphs));
  const bool dev_out = is_device_pointer(dist);
  int32_t* out = dist;
  if (!dev_out) {
    DAWN_TRY(dout.alloc((size_t)total));
    out = dout.p;
  }
  hipStream_t s = graphs[0]->stream;
  DAWN_HIP(hipMemcpyAsync(dg.p, hg.data(), sizeof(BatchGraph) * hg.size(), hipMemcpyHostToDevice, s));
  DAWN_HIP(hipMemcpyAsync(dt.p, ht.data(), sizeof(BatchTask) * ht.size(), hipMemcpyHostToDevice, s));
  DAWN_HIP(hipMemcpyAsync(ds.p, hs.data(), sizeof(int32_t) * hs.size(), hipMemcpyHostToDevice, s));
  DAWN_HIP(hipMemsetAsync(dcnt.p, 0, sizeof(unsigned), s));
  DAWN_HIP(hipMemsetAsync(dlev.p, 0, sizeof(unsigned long long), s));
  DAWN_HIP(hipMemsetAsync(dwsum.p, 0, sizeof(unsigned long long) * num_graphs, s));
  BatchArgs a{};
  a.graphs = dg.p;
  a.tasks = dt.p;
  a.num_tasks = (int32_t)ht.size();
  a.spg = spg;
  a.sources = ds.p;
  a.dist = out;
  a.task_counter = dcnt.p;
  a.levels_total = dlev.p;
  a.wsum = dwsum.p;
  a.max_n = max_n;
  a.seen = dseen.p;
  a.next = dnext.p;
  a.qv = dq.p;
  a.qw = dqw.p;
  a.touched = dtouch.p;
  a.stage = dstage.p;
  hipEvent_t e0, e1;
  DAWN_HIP(hipEventCreate(&e0));
  DAWN_HIP(hipEventCreate(&e1));
  bool any_weighted = false;
  for (const BatchGraph& x : hg) any_weighted = any_weighted || x.weighted;
  DAWN_HIP(hipEventRecord(e0, s));
  if (any_weighted) wsum_kernel<<<num_graphs * kWsumParts, kBlock, 0, s>>>(dg.p, dwsum.p);
  batch_kernel<kBatchBlock><<<grid, kBatchBlock, 0, s>>>(a);
  DAWN_HIP(hipGetLastError());
  DAWN_HIP(hipEventRecord(e1, s));
  DAWN_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  DAWN_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  unsigned long long levels = 0;
  DAWN_HIP(hipMemcpy(&levels, dlev.p, sizeof levels, hipMemcpyDeviceToHost));
  if (dist && !dev_out)
    DAWN_HIP(hipMemcpy(dist, out, sizeof(int32_t) * (size_t)total, hipMemcpyDeviceToHost));
  if (stats) {
    stats->device_ms += ms;
    stats->steps += (int64_t)levels;
    stats->batches += (int64_t)ht.size();
  }
  return DAWN_OK;
}
