// Bit-packed unweighted DAWN over 64·W sources at once (batched BOVM) on one
// HIP device, W = 1, 2, 4, 8 or 16 lane words per vertex.
//
// Replaces the reference's per-source loops DAWN::MSSP_GPU::run
// (src/algorithm/gpu/mssp.cu:9-51), DAWN::APSP_GPU::run
// (src/algorithm/gpu/apsp.cu:62-111) and DAWN::MSSP_CPU::run
// (src/algorithm/cpu/mssp.cpp:9-32): one step advances 64·W frontiers at
// once; each vertex carries W uint64 words, bit b of word i = source lane
// 64·i + b.  A vertex's W words are contiguous, so every random access of
// a step (one frontier gather, one seen probe, one atomic) moves the lanes
// of all W words in one or two cache lines: the dense steps, bound by
// random accesses, cost far less per source at W = 8 than at W = 1.  At
// W = 16 a record is one 128-byte cache line, split over two lanes (eight
// words each) wherever one lane would hold a whole record.
//
// Layout in HBM (per graph, reused across batches; v-major, W words each):
//   seen   u64[npad][W]        lanes that have reached v (after a pull step
//                              only the live lanes: nothing reads the others)
//   next   u64[npad][W]        scatter target of push / top-down steps and of
//                              heavy pull chunks (zero between steps: every
//                              fold clears what its step wrote)
//   hist   u64[H][npad][W]     level history (H = 32 slots, 16 at W = 16):
//                              slot l%H holds the lanes that
//                              reached v exactly at level l (= frontier of l),
//                              valid only where fbits of that level is set
//                              (entries of other vertices are never read)
//   fbits  u32[H][npad/32]     1 bit per vertex: in the frontier of level l
//                              (any lane), slot l%H
//   done   u32[npad/32]        1 bit per vertex: no live lane is missing (stays
//                              set: seen only grows, live lanes only shrink)
//   q, qb  int32[qcap]         push queue: one entry per 256 out-edges of a
//                              frontier vertex (vertex, first edge)
//   dist   int32[64W][n]       output rows of the batch (caller or workspace)
// Step modes (all produce the same next level; one kernel each, the device
// scheduler's choice runs and the others return at once):
//   PULL    every vertex missing a live lane ORs the previous level's words
//           of its in-neighbours that are in the frontier bitmap (CSC):
//           light vertices one lane (pair) each with early exit once all
//           missing lanes are found, folding their result in place (level
//           record, seen, bitmap words, counters); heavy vertices' in-edge
//           lists in chunks, one wave per chunk, merged by atomicOr into
//           next[] and folded by a separate pass.
//   PUSH    queue entries, workgroup-balanced expansion (LDS degree scan),
//           atomicOr of (frontier & ~seen[v]) into next[v], word by word.
//   BITMAP  dense scan of the frontier bitmap, same scatter as PUSH.
//   PUSH and BITMAP steps end with a finalize pass that folds next into the
//   level slot, seen, the bitmaps and the level counters; vertices marked
//   done are skipped, and level words are written only for new arrivals.
// After the last level one pass turns the history into int32 distances
// (16-byte row stores) and per-lane sums/counts (64x64 bit transposes +
// popcount), so no distance is written twice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "sched.hpp"

namespace dawn {

constexpr int kMaxWords = 16;
typedef int32_t v4i32 __attribute__((ext_vector_type(4)));
constexpr int kMaxLanes = 64 * kMaxWords;
// Bits of a level relative to the window start: 2^slices >= level slots.
constexpr int level_slices(int W) { return W > 8 ? 4 : 5; }

// Level slots of the history ring: 32, or 16 at W = 16 (128-byte records:
// 16 slots are 34 GB at n = 2^24).
constexpr int hist_slots(int W) { return W > 8 ? 16 : 32; }
// Steps per captured chunk.  A chunk writes levels up to `chunk` past the
// newest level kept from the previous chunk, so chunk + 1 <= slots.
constexpr int chunk_steps(int W) { return W > 8 ? 14 : 24; }
static_assert(chunk_steps(16) + 1 <= hist_slots(16) && chunk_steps(8) + 1 <= hist_slots(8), "ring");
// Words of a record one lane handles in the dense steps: W, or 8 at W = 16
// (two lanes per vertex, each holding half of the record).
template <int W>
constexpr int lane_words_of() { return W > 8 ? 8 : W; }

struct MsBatch {
  int32_t nlanes;
  int32_t lo;                      // first level not yet flushed to dist
  uint64_t lanemask[kMaxWords];    // valid lanes per word
  int32_t* dist;                   // output rows: lane b at dist + row[b] * stride
  int64_t stride;
  int32_t src[kMaxLanes];
  int32_t row[kMaxLanes];          // output row of each lane
};

struct MsArgs {
  int32_t n;
  int32_t H;
  int64_t npad;
  int64_t m;
  int64_t zero_in;     // vertices of in-degree 0: done from the start
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* in_row_ptr;
  const int32_t* in_col;
  uint64_t* seen;
  uint64_t* next;
  uint64_t* hist;
  uint32_t* fbits;     // [H][fwords]
  uint32_t* done;      // [fwords]
  int64_t fwords;
  int32_t* q;          // [qcap] frontier entries: vertex
  int32_t* qb;         // [qcap] first out-edge of the entry's chunk
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
  int32_t variant;        // light pull variant (TuneKnobs::pull_variant)
  const int32_t* perm;    // original id -> position (hub-first view), or null
  LevelCounters* ring;
  unsigned long long* part;  // kRing x kMsStripes counter stripes
  SchedState* st;
  MsBatch* batch;
};

// Level counters are added per workgroup into one of kMsStripes stripes of
// the level (stripe = workgroup mod kMsStripes, each on its own 256-byte
// span): the same-address atomics of thousands of workgroups would
// serialise at one line (about 12 ns each).  Stripe layout: fv, fe, scout,
// flanes, uv, ue, alive[16].  The decide kernel sums (ORs) the stripes of
// the level it reads into the level's ring entry.
constexpr int kMsStripes = 64;
constexpr int kMsStripeU64 = 32;
constexpr int kMsPartU64 = kRing * kMsStripes * kMsStripeU64;
static_assert(kMsStripes == kWave, "the decide kernel reduces one stripe per lane");

__device__ __forceinline__ unsigned long long* ms_stripe(const MsArgs& a, int lv) {
  return a.part + ((size_t)(lv % kRing) * kMsStripes + blockIdx.x % kMsStripes) * kMsStripeU64;
}

struct MsOut {
  unsigned long long* sum;  // per lane, indexed [lane]
  unsigned long long* reached;
  unsigned long long* edges;
  int32_t want_edges;
};

struct MsWorkspace {
  int32_t n = 0;
  int32_t words = 0;   // W the buffers are laid out for
  int32_t slots = 0;   // level slots of hist / fbits
  int64_t npad = 0;
  DevBuf<uint64_t> seen, next, hist, seen_bak;
  DevBuf<uint32_t> fbits, done, done_bak;
  DevBuf<int32_t> q, qb, distbuf;
  DevBuf<int32_t> iso;  // (row, source) pairs of isolated sources
  int64_t qcap = 0;
  DevBuf<LevelCounters> ring, ring_bak;
  DevBuf<unsigned long long> part, part_bak;
  DevBuf<SchedState> st;
  DevBuf<MsBatch> batch;
  DevBuf<float> trace_feat;
  DevBuf<int32_t> trace_mode;
  DevBuf<unsigned long long> lane_stats;  // 3 * kMaxLanes
  SchedState* h_state = nullptr;
  MsBatch* h_batch = nullptr;
  // Captured batch graphs keyed by (W, policy, want_edges, scheduler net).
  std::map<std::tuple<int, int, int, const void*>, hipGraphExec_t> graphs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, e_a = nullptr, e_b = nullptr;
};

void ss_invalidate_graphs(SsWorkspace* ws);
void wms_invalidate_graphs(WmsWorkspace* ws);

// Captured step graphs bake grid sizes, knobs and buffer addresses (the
// heavy-chunk lists among them); the autotuner and workspace re-layouts
// drop them.
void invalidate_graphs(dawn_graph* g) {
  for (MsWorkspace* ws : {g->ms, g->ms2, g->ms_tail}) {
    if (!ws) continue;
    for (auto& kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
    ws->graphs.clear();
  }
  ss_invalidate_graphs(g->ss);
  wms_invalidate_graphs(g->wms);
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

// The W words of one vertex: one 8-, 16- or 32-byte load.
template <int W>
__device__ __forceinline__ void load_words(const uint64_t* p, uint64_t (&x)[W]) {
  if constexpr (W == 1) {
    x[0] = p[0];
  } else {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
#pragma unroll
    for (int i = 0; i < W / 2; i++) {
      const ulonglong2 t = q[i];
      x[2 * i] = t.x;
      x[2 * i + 1] = t.y;
    }
  }
}

template <int W>
__device__ __forceinline__ void store_words(uint64_t* p, const uint64_t (&x)[W]) {
  if constexpr (W == 1) {
    p[0] = x[0];
  } else {
    ulonglong2* q = reinterpret_cast<ulonglong2*>(p);
#pragma unroll
    for (int i = 0; i < W / 2; i++) q[i] = make_ulonglong2(x[2 * i], x[2 * i + 1]);
  }
}

template <int W>
__device__ __forceinline__ void or_words(uint64_t* p, const uint64_t (&x)[W]) {
#pragma unroll
  for (int i = 0; i < W; i++)
    if (x[i]) atomicOr((unsigned long long*)&p[i], (unsigned long long)x[i]);
}

// OR over the lanes of the wave that are congruent to this one mod T (a
// power of two): T = 1 is the whole wave.
template <int T, int J = 32>
__device__ __forceinline__ uint64_t wave_or_mod(uint64_t x) {
  if constexpr (J >= T) return wave_or_mod<T, J / 2>(x | xor_lane_u64<J>(x));
  else return x;
}

// Bits 0, 2, 4, ... of a 64-bit mask, packed into 32 bits.
__device__ __forceinline__ uint32_t even_bits(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return (uint32_t)x;
}

// Per-thread counters of a level, reduced over the workgroup and flushed
// with one set of atomics per workgroup into the workgroup's counter stripe
// of the level.  alive holds NA of the W lane words: words 0..W-1
// (NA = W), or words sub*NA .. sub*NA + NA-1 when W/NA threads share a
// record and this thread holds part `sub`.  flush() must be called by every
// thread.
template <int W, int NA>
struct MsAcc {
  unsigned long long fv = 0, fe = 0, scout = 0, lanes = 0, uv = 0, ue = 0;
  uint64_t alive[NA] = {};
  __device__ void flush(unsigned long long* c, int sub) {
    constexpr int R = 6 + W;
    __shared__ unsigned long long s_v[kBlock / kWave][R];
    const int w = threadIdx.x / kWave;
    unsigned long long r[R] = {wave_sum_u64(fv),    wave_sum_u64(fe), wave_sum_u64(scout),
                               wave_sum_u64(lanes), wave_sum_u64(uv), wave_sum_u64(ue)};
#pragma unroll
    for (int i = 0; i < W; i++)
      r[6 + i] = wave_or_u64(NA == W || i / NA == sub ? alive[i % NA] : 0ull);
    if (lane_id() == 0)
      for (int i = 0; i < R; i++) s_v[w][i] = r[i];
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t[R] = {};
      for (int j = 0; j < kBlock / kWave; j++) {
        for (int i = 0; i < 6; i++) t[i] += s_v[j][i];
        for (int i = 6; i < R; i++) t[i] |= s_v[j][i];
      }
      for (int i = 0; i < 6; i++)
        if (t[i]) atomicAdd(&c[i], t[i]);
      for (int i = 0; i < W; i++)
        if (t[6 + i]) atomicOr(&c[6 + i], t[6 + i]);
    }
  }
};

// ---------------------------------------------------------------- kernels ---

template <int W>
__global__ void ms_fill_kernel(MsArgs a, MsOut o) {
  const MsBatch& b = *a.batch;
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  const int64_t words = a.npad * W;
  // next[] is not touched: it is zeroed when allocated and every step
  // leaves it zero (the fold clears what the step wrote).
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < words; x += gs) {
    const int64_t v = x / W;
    a.seen[x] = v < a.n ? ~b.lanemask[x % W] : ~0ull;
  }
  // A vertex of in-degree 0 never receives a lane (a source has it from
  // level 0 on): it starts done, so no step reads its seen words.  The
  // scheduler adds these vertices back to its unvisited count.
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < a.fwords; w += gs) {
    a.fbits[w] = 0u;  // level-0 frontier bitmap
    uint32_t d = 0u;
    if (a.zero_in) {
      const int64_t v0 = w * 32;
      int32_t r = v0 < a.n ? a.in_row_ptr[v0] : 0;
      for (int j = 0; j < 32 && v0 + j < a.n; j++) {
        const int32_t r1 = a.in_row_ptr[v0 + j + 1];
        d |= (r1 == r ? 1u : 0u) << j;
        r = r1;
      }
    }
    a.done[w] = d;
  }
  if (blockIdx.x == 0 && threadIdx.x < kRing) a.ring[threadIdx.x] = LevelCounters{};
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < kMsPartU64; t += gs)
    a.part[t] = 0ull;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 64 * W; t += gs) {
    o.sum[t] = 0;
    o.reached[t] = 0;
    o.edges[t] = 0;
  }
}

// Level 0: the batch's sources (duplicates share one frontier vertex), one
// thread per lane.  The first lane of each source vertex clears its level
// record (the slot may hold words of an earlier batch) and marks it in the
// level-0 bitmap; then every lane sets its bit.
template <int W>
__global__ void __launch_bounds__(kMaxLanes) ms_seed_kernel(MsArgs a) {
  __shared__ int32_t s_src[kMaxLanes];
  __shared__ unsigned long long s_c[4];  // fv, fe, source vertices complete, their in-degree
  __shared__ unsigned long long s_alive[kMaxWords];
  const MsBatch& b = *a.batch;
  const int nl = b.nlanes;
  const int l = threadIdx.x;
  if (l < nl) s_src[l] = b.src[l];
  if (l < 4) s_c[l] = 0;
  if (l < kMaxWords) s_alive[l] = 0;
  __syncthreads();
  int32_t v = -1;
  bool first = false;
  if (l < nl) {
    v = s_src[l];
    first = true;
    for (int k = 0; k < l && first; k++) first = s_src[k] != v;
  }
  if (first) {
    for (int k = 0; k < W; k++) a.hist[(int64_t)v * W + k] = 0ull;
    atomicOr(&a.fbits[v >> 5], 1u << (v & 31));
    atomicAdd(&s_c[0], 1ull);
    atomicAdd(&s_c[1], (unsigned long long)out_deg(a, v));
  }
  __syncthreads();
  if (l < nl) {
    const uint64_t bit = 1ull << (l & 63);
    atomicOr((unsigned long long*)&a.hist[(int64_t)v * W + (l >> 6)], (unsigned long long)bit);
    atomicOr((unsigned long long*)&a.seen[(int64_t)v * W + (l >> 6)], (unsigned long long)bit);
    atomicOr(&s_alive[l >> 6], (unsigned long long)bit);
  }
  __syncthreads();
  // Vertices still missing a lane after level 0, and their in-degree sum:
  // a source vertex leaves the count (once) if every lane has it.
  if (first) {
    bool full = true;
    for (int k = 0; k < W; k++) full = full && a.seen[(int64_t)v * W + k] == ~0ull;
    if (full && a.in_row_ptr[v + 1] > a.in_row_ptr[v]) {  // in-degree 0: in zero_in
      atomicAdd(&s_c[2], 1ull);
      atomicAdd(&s_c[3], (unsigned long long)(a.in_row_ptr[v + 1] - a.in_row_ptr[v]));
    }
  }
  __syncthreads();
  if (l == 0) {
    LevelCounters c{};
    c.fv = s_c[0];
    c.fe = s_c[1];
    c.scout = s_c[1];
    c.flanes = (unsigned long long)nl;
    for (int k = 0; k < W; k++) c.alive[k] = s_alive[k];
    c.uv = (unsigned long long)(a.n - a.zero_in) - s_c[2];
    c.ue = (unsigned long long)a.m - s_c[3];
    a.ring[0] = c;
    // Level 0's counters in its stripe 0 (the other stripes are zero).
    unsigned long long* p0 = a.part;
    p0[0] = c.fv;
    p0[1] = c.fe;
    p0[2] = c.scout;
    p0[3] = c.flanes;
    p0[4] = c.uv;
    p0[5] = c.ue;
    for (int k = 0; k < 16; k++) p0[6 + k] = c.alive[k];
    SchedState st{};
    st.step = 0;
    st.e2c = a.m;
    st.scout = (long long)c.scout;
    st.force_mode = -1;
    st.last_mode = DAWN_MODE_PUSH;
    *a.st = st;
  }
}

// One wave: reduces the counter stripes of the last completed level into
// its ring entry (an assignment, so running it twice is harmless), then
// decides.
__global__ void ms_decide_kernel(MsArgs a, SchedParams p, SchedTrace tr) {
  const SchedState s = *a.st;
  if (!s.done) {
    const int slot = s.step % kRing;
    const unsigned long long* sp = a.part + ((size_t)slot * kMsStripes + lane_id()) * kMsStripeU64;
    unsigned long long v[6];
#pragma unroll
    for (int i = 0; i < 6; i++) v[i] = wave_sum_u64(sp[i]);
    uint64_t al[16];
#pragma unroll
    for (int k = 0; k < 16; k++) al[k] = wave_or_u64(sp[6 + k]);
    if (threadIdx.x == 0) {
      LevelCounters* c = a.ring + slot;
      c->fv = v[0];
      c->fe = v[1];
      c->scout = v[2];
      c->flanes = v[3];
      c->uv = v[4];
      c->ue = v[5];
      for (int k = 0; k < 16; k++) c->alive[k] = al[k];
    }
  }
  sched_decide(a.st, a.ring, p, tr);
}

// Pull steps fold their light vertices in the pull kernel, except the
// edge-balanced variant (1, W <= 4), which leaves every vertex to finalize.
template <int W>
__device__ __forceinline__ bool pull_folds_light(const MsArgs& a) {
  return !(W <= 4 && a.variant == 1);
}

// Light vertices of a pull step (in-degree < a.heavy), one lane each (two
// at W = 16, each with half of the record), with the step's fold done in
// place, so that next[] is not written, read back and cleared for them:
// the lane writes its vertex's level record and seen words, and the wave
// writes the frontier and done bitmap words of its 64 (32) vertices and
// accumulates the level counters.  Heavy vertices (chunked below, merged in
// next[]) are folded by the finalize pass; their bitmap bits are left
// clear here (a done heavy vertex keeps its done bit).
//
// Per lane: U column reads, then U bitmap probes, then the gathers they
// admit, between early-exit checks (all live missing lanes found).  NT:
// column reads as streaming loads.  The halves of a W = 16 record are
// independent: each lane stops once its own half is complete (more
// gathers could only add lanes outside its miss mask).  Vertices marked
// done are skipped before their seen words are read.  seen keeps only the
// live lanes: alive & ~(still missing) -- bits of lanes that are no longer
// alive are never read again (miss, the push masks and the fold only look
// at live lanes; distances come from the level records).
template <int W, int U, bool NT>
__device__ __forceinline__ void pull_light_fused(const MsArgs& a, int lv, const uint64_t* prev,
                                                 const uint32_t* fb,
                                                 const uint64_t (&alive)[lane_words_of<W>()]) {
  constexpr int HW = lane_words_of<W>();
  constexpr int TPV = W / HW;       // lanes per vertex
  constexpr int VPW = kWave / TPV;  // vertices per wave and pass: two bitmap words, or one
  const int lane = lane_id();
  const int h = lane % TPV;
  uint64_t* cur = a.hist + (int64_t)(lv % a.H) * a.npad * W;
  uint32_t* fbo = a.fbits + (int64_t)(lv % a.H) * a.fwords;
  // Level counters: per-lane 32-bit sums (each bounded by m + n < 2^32),
  // vertex counts from ballots, live lanes OR-ed into LDS (checked first,
  // so the atomics stop once the words fill up).
  __shared__ unsigned long long s_alive[W];
  __shared__ unsigned long long s_cnt[kBlock / kWave][6];
  if (threadIdx.x < W) s_alive[threadIdx.x] = 0;
  __syncthreads();
  unsigned c_fe = 0, c_scout = 0, c_ue = 0, c_lanes = 0;
  unsigned long long c_fv = 0, c_uv = 0;  // wave-uniform
  const int64_t vstride = (int64_t)gridDim.x * (kBlock / TPV);
  for (int64_t base = (int64_t)blockIdx.x * (kBlock / TPV) + (threadIdx.x / kWave) * VPW;
       base < a.npad; base += vstride) {  // wave-uniform: the whole wave runs every pass
    const int32_t v = (int32_t)(base + lane / TPV);
    const bool done_before = v >= a.n || ((a.done[v >> 5] >> (v & 31)) & 1u);
    int beg = 0, deg = 0;
    if (!done_before) {
      beg = a.in_row_ptr[v];
      deg = a.in_row_ptr[v + 1] - beg;
    }
    const bool light = !done_before && deg < a.heavy;
    uint64_t nw[HW];
#pragma unroll
    for (int i = 0; i < HW; i++) nw[i] = 0;
    unsigned still = 0;  // this half still misses a live lane after the step
    if (light) {
      uint64_t miss[HW];
      {
        uint64_t sv[HW];
        load_words<HW>(a.seen + (int64_t)v * W + h * HW, sv);
#pragma unroll
        for (int i = 0; i < HW; i++) miss[i] = ~sv[i] & alive[i];
      }
      uint64_t any = 0;
#pragma unroll
      for (int i = 0; i < HW; i++) any |= miss[i];
      bool scan = any != 0;
      if constexpr (TPV > 1) scan = (xor_lane_u32<1>(scan ? 1u : 0u) != 0) || scan;  // pair: both lanes here
      if (scan && deg > 0) {
        const int32_t* nb = a.in_col + beg;
        uint64_t acc_w[HW] = {};
        auto complete = [&]() {
          bool c = true;
#pragma unroll
          for (int i = 0; i < HW; i++) c = c && (acc_w[i] & miss[i]) == miss[i];
          return c;
        };
        auto gather = [&](int32_t u) {
          uint64_t f[HW];
          load_words<HW>(prev + (int64_t)u * W + h * HW, f);
#pragma unroll
          for (int i = 0; i < HW; i++) acc_w[i] |= f[i];
        };
        if constexpr (TPV == 1) {
          int j = 0;
          for (; j + U <= deg; j += U) {
            int32_t u[U];
            uint32_t w[U];
#pragma unroll
            for (int t = 0; t < U; t++) u[t] = NT ? __builtin_nontemporal_load(nb + j + t) : nb[j + t];
#pragma unroll
            for (int t = 0; t < U; t++) w[t] = fb[u[t] >> 5];
#pragma unroll
            for (int t = 0; t < U; t++)
              if ((w[t] >> (u[t] & 31)) & 1u) gather(u[t]);
            if (complete()) break;
          }
          if (!complete())
            for (; j < deg; j++) {
              const int32_t u = NT ? __builtin_nontemporal_load(nb + j) : nb[j];
              if ((fb[u >> 5] >> (u & 31)) & 1u) gather(u);
            }
        } else {
          // Lane pair: each lane probes every other edge (U per lane and
          // round), the probes are exchanged, and both lanes gather the
          // records that are in the frontier -- the even lane's first, so
          // the two halves of a line are loaded by one instruction.  A lane
          // whose half is complete still probes for its partner; the pair
          // stops when both halves are.
          bool mine = complete();
          for (int j = 0; j < deg; j += 2 * U) {
            int32_t r0[U], r1[U];
            bool g0[U], g1[U];
#pragma unroll
            for (int t = 0; t < U; t++) {
              const int e = j + 2 * t + h;
              int32_t u = 0;
              bool hit = false;
              if (e < deg) {
                u = NT ? __builtin_nontemporal_load(nb + e) : nb[e];
                hit = (fb[u >> 5] >> (u & 31)) & 1u;
              }
              const int32_t pu = (int32_t)xor_lane_u32<1>((uint32_t)u);
              const bool ph = xor_lane_u32<1>(hit ? 1u : 0u) != 0;
              r0[t] = h ? pu : u;
              g0[t] = h ? ph : hit;
              r1[t] = h ? u : pu;
              g1[t] = h ? hit : ph;
            }
            if (!mine) {
#pragma unroll
              for (int t = 0; t < U; t++) {
                if (g0[t]) gather(r0[t]);
                if (g1[t]) gather(r1[t]);
              }
              mine = complete();
            }
            const bool both = (xor_lane_u32<1>(mine ? 1u : 0u) != 0) && mine;  // same in both lanes
            if (both) break;
          }
        }
#pragma unroll
        for (int i = 0; i < HW; i++) nw[i] = acc_w[i] & miss[i];
      }
      uint64_t rest = 0, got = 0;
#pragma unroll
      for (int i = 0; i < HW; i++) {
        rest |= miss[i] & ~nw[i];
        got |= nw[i];
      }
      still = rest != 0;
      if (got) {
        uint64_t sn[HW];
#pragma unroll
        for (int i = 0; i < HW; i++) sn[i] = alive[i] & ~(miss[i] & ~nw[i]);
        store_words<HW>(a.seen + (int64_t)v * W + h * HW, sn);
      }
    }
    // Arrivals anywhere in the vertex: its whole level record is written.
    unsigned arr = 0;
#pragma unroll
    for (int i = 0; i < HW; i++) arr |= nw[i] != 0 ? 1u : 0u;
    if constexpr (TPV > 1) {
      arr |= xor_lane_u32<1>(arr);
      still |= xor_lane_u32<1>(still);
    }
    if (light) {
      if (arr) store_words<HW>(cur + (int64_t)v * W + h * HW, nw);
#pragma unroll
      for (int i = 0; i < HW; i++) {
        c_lanes += (unsigned)__popcll(nw[i]);
        const uint64_t x = nw[i] & ~s_alive[h * HW + i];
        if (x) atomicOr(&s_alive[h * HW + i], (unsigned long long)x);
      }
      if (h == 0) {
        if (arr) {
          const int d = out_deg(a, v);
          c_fe += (unsigned)d;
          c_scout += (unsigned)(d > 0 ? d : 1);
        }
        if (still) c_ue += (unsigned)deg;
      }
    }
    // Frontier and done bitmap words of the pass's vertices (heavy ones:
    // frontier bit clear, done bit as before; the finalize pass adds theirs).
    const uint64_t fm = __ballot(light && arr && h == 0);
    const uint64_t dm = __ballot(h == 0 && (done_before || (light && !still)));
    c_fv += (unsigned long long)__popcll(fm);
    c_uv += (unsigned long long)__popcll(__ballot(h == 0 && light && still));
    if (lane == 0) {
      if constexpr (TPV == 1) {
        fbo[base >> 5] = (uint32_t)fm;
        fbo[(base >> 5) + 1] = (uint32_t)(fm >> 32);
        a.done[base >> 5] = (uint32_t)dm;
        a.done[(base >> 5) + 1] = (uint32_t)(dm >> 32);
      } else {
        fbo[base >> 5] = even_bits(fm);
        a.done[base >> 5] = even_bits(dm);
      }
    }
  }
  // Workgroup totals, one set of atomics per workgroup into its stripe.
  const unsigned long long r[6] = {c_fv, wave_sum_u64(c_fe), wave_sum_u64(c_scout),
                                   wave_sum_u64(c_lanes), c_uv, wave_sum_u64(c_ue)};
  if (lane == 0)
    for (int i = 0; i < 6; i++) s_cnt[threadIdx.x / kWave][i] = r[i];
  __syncthreads();
  unsigned long long* c = ms_stripe(a, lv);
  if (threadIdx.x < 6) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / kWave; w++) t += s_cnt[w][threadIdx.x];
    if (t) atomicAdd(&c[threadIdx.x], t);
  }
  if (threadIdx.x < W && s_alive[threadIdx.x])
    atomicOr(&c[6 + threadIdx.x], (unsigned long long)s_alive[threadIdx.x]);
}

// Expansion of one step, one kernel per mode (the device scheduler picks
// the mode; the other two return at once).  Separate kernels keep each
// mode's registers and LDS to itself.  HW words of a record per lane; at
// W = 16 lane pairs share a vertex (TPV = 2), lane l holding words
// h*HW .. h*HW + HW-1, h = l % TPV.
#define MS_STEP_PROLOGUE(MODE)                                              \
  constexpr int HW = lane_words_of<W>();                                    \
  constexpr int TPV = W / HW;                                               \
  const SchedState st = *a.st;                                              \
  if (st.done || st.mode != (MODE)) return;                                 \
  const int lv = st.step;                                                   \
  const uint64_t* prev = a.hist + (int64_t)((lv - 1) % a.H) * a.npad * W;   \
  const int lane = lane_id();                                               \
  const int h = lane % TPV;                                                 \
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ring[(lv + 1) % kRing] = LevelCounters{}; \
  if (blockIdx.x == 0)                                                      \
    for (int i_ = threadIdx.x; i_ < kMsStripes * kMsStripeU64; i_ += blockDim.x) \
      a.part[(size_t)((lv + 1) % kRing) * kMsStripes * kMsStripeU64 + i_] = 0ull; \
  (void)lane;                                                               \
  (void)h

// PUSH: queue entries, workgroup-balanced expansion (LDS degree scan).
template <int W>
__global__ void __launch_bounds__(kBlock) ms_push_kernel(MsArgs a) {
  MS_STEP_PROLOGUE(DAWN_MODE_PUSH);
  __shared__ int s_off[kBlock];
  __shared__ int s_beg[kBlock];
  __shared__ uint64_t s_f[kBlock * W];
  __shared__ unsigned s_nz[kBlock];
  __shared__ int s_scan[kBlock / kWave];
  const int32_t* q_in = a.q;  // built by ms_compact_kernel for this step
  const int32_t* qb_in = a.qb;
  const unsigned qn = a.ring[(lv - 1) % kRing].qlen;
  for (int64_t c0 = (int64_t)blockIdx.x * kBlock; c0 < qn; c0 += (int64_t)gridDim.x * kBlock) {
    const int64_t i = c0 + threadIdx.x;
    int beg = 0, deg = 0;
    unsigned nz = 0;  // nonzero words of the entry's frontier record
    if (i < qn) {
      const int32_t u = q_in[i];
      beg = qb_in[i];
      deg = min(a.row_ptr[u + 1] - beg, kPushChunk);
      uint64_t f[HW];
#pragma unroll
      for (int p = 0; p < TPV; p++) {
        load_words<HW>(prev + (int64_t)u * W + p * HW, f);
#pragma unroll
        for (int k = 0; k < HW; k++) {
          s_f[threadIdx.x * W + p * HW + k] = f[k];
          nz |= (f[k] != 0ull ? 1u : 0u) << (p * HW + k);
        }
      }
    }
    if (!nz) deg = 0;  // an empty record reaches nobody
    s_nz[threadIdx.x] = nz;
    int total;
    const int off = block_exclusive_scan(deg, s_scan, &total);
    s_off[threadIdx.x] = off;
    s_beg[threadIdx.x] = beg;
    __syncthreads();
    // Frontier records of sparse levels hold few nonzero words: an edge
    // touches seen/next only in the words its owner has (8-byte accesses
    // when at most half of them are set, whole-record loads otherwise).
    for (int e = threadIdx.x; e < total; e += kBlock) {
      const int j = owner_of(s_off, e);
      unsigned msk = s_nz[j];
      const int32_t v = a.col[s_beg[j] + (e - s_off[j])];
      const int64_t vw = (int64_t)v * W;
      if (2 * __popc(msk) > W) {
#pragma unroll
        for (int p = 0; p < TPV; p++) {
          uint64_t sv[HW], nv[HW];
          load_words<HW>(a.seen + vw + p * HW, sv);
#pragma unroll
          for (int k = 0; k < HW; k++) nv[k] = s_f[j * W + p * HW + k] & ~sv[k];
          or_words<HW>(a.next + vw + p * HW, nv);
        }
      } else {
        for (; msk; msk &= msk - 1) {
          const int k = __ffs(msk) - 1;
          const uint64_t nv = s_f[j * W + k] & ~a.seen[vw + k];
          if (nv) atomicOr((unsigned long long*)&a.next[vw + k], (unsigned long long)nv);
        }
      }
    }
    __syncthreads();
  }
}

// BITMAP: dense top-down scan of the frontier bitmap.
template <int W>
__global__ void __launch_bounds__(kBlock) ms_topdown_kernel(MsArgs a) {
  MS_STEP_PROLOGUE(DAWN_MODE_BITMAP);
  // Dense top-down: light frontier vertices one lane (pair) each, heavy
  // ones in out-edge chunks of knobs.pull_chunk edges, one wave per chunk.
  const uint32_t* fbp = a.fbits + (int64_t)((lv - 1) % a.H) * a.fwords;
  const int64_t vstride = (int64_t)gridDim.x * (kBlock / TPV);
  for (int64_t base = (int64_t)blockIdx.x * (kBlock / TPV) + (threadIdx.x / kWave) * (kWave / TPV);
       base < a.n; base += vstride) {
    const int32_t u = (int32_t)(base + lane / TPV);
    if (u >= a.n || !((fbp[u >> 5] >> (u & 31)) & 1u)) continue;
    const int beg = a.row_ptr[u];
    const int deg = a.row_ptr[u + 1] - beg;
    if (deg >= a.heavy) continue;
    uint64_t f[HW];
    load_words<HW>(prev + (int64_t)u * W + h * HW, f);
    int nzc = 0;
#pragma unroll
    for (int k = 0; k < HW; k++) nzc += f[k] != 0ull;
    if (!nzc) continue;
    const bool dense = 2 * nzc > HW;  // else: 8-byte accesses of the nonzero words only
    for (int j = 0; j < deg; j++) {
      const int64_t vw = (int64_t)a.col[beg + j] * W + h * HW;
      if (dense) {
        uint64_t sv[HW], nv[HW];
        load_words<HW>(a.seen + vw, sv);
#pragma unroll
        for (int k = 0; k < HW; k++) nv[k] = f[k] & ~sv[k];
        or_words<HW>(a.next + vw, nv);
      } else {
#pragma unroll
        for (int k = 0; k < HW; k++) {
          if (!f[k]) continue;
          const uint64_t nv = f[k] & ~a.seen[vw + k];
          if (nv) atomicOr((unsigned long long*)&a.next[vw + k], (unsigned long long)nv);
        }
      }
    }
  }
  const int64_t wave = (int64_t)(blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t c = wave; c < a.num_oc; c += nwaves) {  // wave-uniform
    const int32_t u = a.oc_v[c];
    if (!((fbp[u >> 5] >> (u & 31)) & 1u)) continue;
    uint64_t f[HW];
    load_words<HW>(prev + (int64_t)u * W + h * HW, f);
    uint64_t any = 0;
#pragma unroll
    for (int k = 0; k < HW; k++) any |= f[k];
    if constexpr (TPV > 1) any |= xor_lane_u64<1>(any);
    if (!any) continue;
    const int e = a.oc_end[c];
    for (int k = a.oc_beg[c] + lane / TPV; k < e; k += kWave / TPV) {
      const int32_t v = a.col[k];
      uint64_t sv[HW], nv[HW];
      load_words<HW>(a.seen + (int64_t)v * W + h * HW, sv);
#pragma unroll
      for (int t = 0; t < HW; t++) nv[t] = f[t] & ~sv[t];
      or_words<HW>(a.next + (int64_t)v * W + h * HW, nv);
    }
  }
}

// PULL: bottom-up over the in-edges (CSC).  PV: the light-vertex variant
// (knobs.pull_variant), a template argument so that each variant gets the
// registers it needs and no more.
template <int W, int PV>
__global__ void __launch_bounds__(kBlock) ms_pull_kernel(MsArgs a) {
  MS_STEP_PROLOGUE(DAWN_MODE_PULL);
  __shared__ int s_hits[TPV > 1 ? kBlock : 1];  // frontier in-neighbours of a heavy chunk round
  // Only lanes alive in the previous level can arrive now: a vertex looks
  // for miss = (lanes it lacks) & (live lanes) and stops once all are found.
  // alive: the live lanes of this lane's HW words.
  uint64_t alive[HW];
#pragma unroll
  for (int k = 0; k < HW; k++) alive[k] = a.ring[(lv - 1) % kRing].alive[h * HW + k];
  // Frontier bitmap of the previous level (2 MB at n = 2^24, L2-resident):
  // frontier words are gathered only for in-neighbours that are in it.
  const uint32_t* fb = a.fbits + (int64_t)((lv - 1) % a.H) * a.fwords;
  if constexpr (W <= 4 && PV == 1) {
    // Light vertices, edge-balanced: the workgroup's 256 vertices' in-edge
    // lists are concatenated (degree scan) and read one edge per thread,
    // coalesced; hits are OR-ed into the owner's LDS words.  Gathers stop
    // for an owner once it has found all its missing lanes.
    // (W >= 8 runs variant 3 instead: these arrays would halve occupancy.)
    __shared__ int s_off[kBlock];
    __shared__ int s_beg[kBlock];
    __shared__ uint64_t s_f[kBlock * W];
    __shared__ int s_scan[kBlock / kWave];
    __shared__ unsigned long long s_acc[kBlock * W];
    const int64_t wstride = (int64_t)gridDim.x * kBlock;
    for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < a.n; b0 += wstride) {
      const int32_t v = (int32_t)(b0 + threadIdx.x);
      uint64_t miss[W] = {};
      uint64_t any = 0;
      if (v < a.n) {
        uint64_t sv[W];
        load_words<W>(a.seen + (int64_t)v * W, sv);
#pragma unroll
        for (int k = 0; k < W; k++) {
          miss[k] = ~sv[k] & alive[k % HW];
          any |= miss[k];
        }
      }
      int beg = 0, deg = 0;
      if (any) {
        beg = a.in_row_ptr[v];
        const int d = a.in_row_ptr[v + 1] - beg;
        deg = d < a.heavy ? d : 0;
      }
      int total;
      const int off = block_exclusive_scan(deg, s_scan, &total);
      s_off[threadIdx.x] = off;
      s_beg[threadIdx.x] = beg;
#pragma unroll
      for (int k = 0; k < W; k++) {
        s_f[threadIdx.x * W + k] = miss[k];
        s_acc[threadIdx.x * W + k] = 0ull;
      }
      __syncthreads();
      for (int e = threadIdx.x; e < total; e += kBlock) {
        const int j = owner_of(s_off, e);
        const int32_t u = a.in_col[s_beg[j] + (e - s_off[j])];
        if ((fb[u >> 5] >> (u & 31)) & 1u) {
          bool done = true;
#pragma unroll
          for (int k = 0; k < W; k++) {
            const uint64_t mj = s_f[j * W + k];
            done = done && (__atomic_load_n(&s_acc[j * W + k], __ATOMIC_RELAXED) & mj) == mj;
          }
          if (!done) {
            uint64_t f[W];
            load_words<W>(prev + (int64_t)u * W, f);
#pragma unroll
            for (int k = 0; k < W; k++) {
              const uint64_t hit = f[k] & s_f[j * W + k];
              if (hit) atomicOr(&s_acc[j * W + k], (unsigned long long)hit);
            }
          }
        }
      }
      __syncthreads();
      if (deg) {
#pragma unroll
        for (int k = 0; k < W; k++)
          if (s_acc[threadIdx.x * W + k]) a.next[(int64_t)v * W + k] = s_acc[threadIdx.x * W + k];
      }
    }
  } else if constexpr (PV == 0) {
    // Gathers in flight per lane: fewer for wide records (registers).
    pull_light_fused<W, (W > 4 ? 2 : 4), false>(a, lv, prev, fb, alive);
  } else if constexpr (PV == 2) {
    pull_light_fused<W, (W > 4 ? 4 : 8), false>(a, lv, prev, fb, alive);
  } else {
    pull_light_fused<W, (W > 4 ? 4 : 8), true>(a, lv, prev, fb, alive);
  }
  // Heavy vertices: their in-edge lists are cut into chunks of at most
  // knobs.pull_chunk edges, one wave per chunk (lane pairs per edge at
  // W = 16), merged with atomicOr.
  const int64_t wave = (int64_t)(blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t c = wave; c < a.num_hc; c += nwaves) {  // wave-uniform
    const int32_t v = a.hc_v[c];
    uint64_t sv[HW], nx[HW], miss[HW];
    load_words<HW>(a.seen + (int64_t)v * W + h * HW, sv);
    load_words<HW>(a.next + (int64_t)v * W + h * HW, nx);
    bool found = true;
    uint64_t any = 0;
#pragma unroll
    for (int k = 0; k < HW; k++) {
      miss[k] = ~sv[k] & alive[k];
      any |= miss[k];
      found = found && (nx[k] & miss[k]) == miss[k];  // other chunks found them all
    }
    if constexpr (TPV > 1) {
      any |= xor_lane_u64<1>(any);
      found = (xor_lane_u32<1>(found ? 1u : 0u) != 0) && found;
    }
    if (!any || found) continue;
    const int he = a.hc_end[c];
    uint64_t part[HW] = {};
    if constexpr (TPV == 1) {
      for (int k = a.hc_beg[c] + lane; k < he; k += kWave) {
        const int32_t u = a.in_col[k];
        if ((fb[u >> 5] >> (u & 31)) & 1u) {
          uint64_t f[HW];
          load_words<HW>(prev + (int64_t)u * W, f);
#pragma unroll
          for (int t = 0; t < HW; t++) part[t] |= f[t];
        }
      }
    } else {
      // Split records: every lane probes one in-edge (64 per round), the
      // in-neighbours found in the frontier are listed in LDS, and lane
      // pairs gather the listed records (two halves of one line each).
      // Sparse frontiers cost one probe per edge, not one per lane pair.
      // The column reads and bitmap probes of R rounds are issued before
      // any is examined, so a wave waits once per R*64 edges, not per round.
      constexpr int R = 8;
      int* hits = s_hits + (threadIdx.x / kWave) * kWave;
      for (int k0 = a.hc_beg[c]; k0 < he; k0 += R * kWave) {  // wave-uniform
        int32_t u[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int k = k0 + r * kWave + lane;
          u[r] = k < he ? a.in_col[k] : -1;
        }
        uint32_t fw[R];
#pragma unroll
        for (int r = 0; r < R; r++) fw[r] = u[r] >= 0 ? fb[u[r] >> 5] : 0u;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const bool hit = u[r] >= 0 && ((fw[r] >> (u[r] & 31)) & 1u);
          const uint64_t m = __ballot(hit);
          if (!m) continue;
          if (hit) hits[__popcll(m & lanes_below_mask())] = u[r];
          wave_lds_sync();
          const int nh = __popcll(m);
          for (int b = 0; b < nh; b += kWave / TPV) {
            const int i = b + lane / TPV;
            if (i < nh) {
              uint64_t f[HW];
              load_words<HW>(prev + (int64_t)hits[i] * W + h * HW, f);
#pragma unroll
              for (int t = 0; t < HW; t++) part[t] |= f[t];
            }
          }
          wave_lds_sync();  // the list is rewritten by the next round
        }
      }
    }
#pragma unroll
    for (int t = 0; t < HW; t++) part[t] = wave_or_mod<TPV>(part[t]) & miss[t];
    if (lane < TPV) or_words<HW>(a.next + (int64_t)v * W + h * HW, part);
  }
}
#undef MS_STEP_PROLOGUE

// Folds next[] into the level slot, seen, the frontier bitmap, the done
// bitmap and the counters.  Each thread takes WPT consecutive words (4, or
// 8 at W = 16) with 16-byte loads per array: 4/W whole vertices (W < 4), one
// vertex (W = 4), or half of one vertex's record (W = 8, 16: two threads
// per vertex, combined with a lane shuffle).  npad is a multiple of 64, so
// a thread's vertices never straddle it.  Vertices already done are
// skipped after one bitmap read.
template <int W>
__global__ void __launch_bounds__(kBlock) ms_finalize_kernel(MsArgs a) {
  constexpr int WPT = W > 8 ? 8 : 4;           // words per thread
  constexpr int TPV = W > WPT ? W / WPT : 1;   // threads per vertex
  constexpr int V = W < WPT ? WPT / W : 1;     // vertices per thread
  constexpr int WPV = W < WPT ? W : WPT;       // this thread's words of one vertex
  constexpr int LPW = 32 * TPV / V;            // lanes per 32-bit bitmap word
  constexpr unsigned kAll = (1u << V) - 1;
  static_assert(LPW <= kWave, "a bitmap word is written by one wave");
  const SchedState st = *a.st;
  if (st.done) return;
  const int lv = st.step;
  uint64_t* cur = a.hist + (int64_t)(lv % a.H) * a.npad * W;
  uint32_t* fb = a.fbits + (int64_t)(lv % a.H) * a.fwords;
  unsigned long long* c_out = ms_stripe(a, lv);
  const int lane = lane_id();
  const int sub = lane % TPV;  // half of the record (W >= 8)
  // Live lanes of the previous level for this thread's words of a vertex
  // (record words sub*WPT + i).
  uint64_t alive_prev[WPV];
  const unsigned long long* al = a.ring[(lv - 1) % kRing].alive;
#pragma unroll
  for (int i = 0; i < WPV; i++) alive_prev[i] = al[sub * WPT + i];
  MsAcc<W, WPV> acc;
  if (st.mode == DAWN_MODE_PULL && pull_folds_light<W>(a)) return;  // ms_fold_heavy_kernel
  constexpr int64_t kTile = (int64_t)kBlock / TPV * V;
  for (int64_t b0 = (int64_t)blockIdx.x * kTile; b0 < a.npad; b0 += (int64_t)gridDim.x * kTile) {
    const int64_t v0 = b0 + (int64_t)(threadIdx.x / TPV) * V;
    unsigned dnib = kAll;
    if (v0 < a.npad) dnib = (a.done[v0 >> 5] >> (v0 & 31)) & kAll;
    const bool work = v0 < a.npad && dnib != kAll;  // same for the threads of a vertex
    const unsigned done_in = dnib;  // (W < 4) siblings done on entry are not counted again
    const int64_t w0 = v0 * W + WPT * sub;
    ulonglong2* np2 = reinterpret_cast<ulonglong2*>(a.next + w0);
    ulonglong2* sp2 = reinterpret_cast<ulonglong2*>(a.seen + w0);
    ulonglong2* cp2 = reinterpret_cast<ulonglong2*>(cur + w0);
    uint64_t sv[WPT], nw[WPT];
    unsigned nzp = 0;  // 16-byte pairs of next[] that hold bits (to clear)
#pragma unroll
    for (int k = 0; k < WPT; k++) sv[k] = nw[k] = 0;
    if (work) {
      ulonglong2 nn[WPT / 2], ss[WPT / 2];
#pragma unroll
      for (int p = 0; p < WPT / 2; p++) nn[p] = np2[p];
#pragma unroll
      for (int p = 0; p < WPT / 2; p++) ss[p] = sp2[p];
#pragma unroll
      for (int p = 0; p < WPT / 2; p++) {
        if (nn[p].x | nn[p].y) nzp |= 1u << p;
        sv[2 * p] = ss[p].x;
        sv[2 * p + 1] = ss[p].y;
        nw[2 * p] = nn[p].x & ~ss[p].x;
        nw[2 * p + 1] = nn[p].y & ~ss[p].y;
      }
    }
    // Per vertex: arrivals, missing live lanes, popcount (this thread's part).
    uint64_t any[V], unv[V];
    unsigned pc[V];
#pragma unroll
    for (int j = 0; j < V; j++) {
      any[j] = 0;
      unv[j] = 0;
      pc[j] = 0;
#pragma unroll
      for (int i = 0; i < WPV; i++) {
        const uint64_t x = nw[j * WPV + i];
        any[j] |= x;
        pc[j] += (unsigned)__popcll(x);
        unv[j] |= ~(sv[j * WPV + i] | x) & alive_prev[i];
        acc.alive[i] |= x;
      }
    }
    if constexpr (TPV > 1) {  // every lane of the wave is here
      any[0] = group_or_u64<TPV>(any[0]);
      unv[0] = group_or_u64<TPV>(unv[0]);
      pc[0] = group_sum_u32<TPV>(pc[0]);
    }
    unsigned nib = 0;
    if (work) {
#pragma unroll
      for (int p = 0; p < WPT / 2; p++)
        if ((nzp >> p) & 1u) np2[p] = make_ulonglong2(0ull, 0ull);
      // Level words only where something arrived (the frontier bitmap marks
      // which entries of the slot are valid), but always a vertex's whole
      // record (W >= 4); for W < 4 a 16-byte pair covers whole vertices.
#pragma unroll
      for (int p = 0; p < WPT / 2; p++) {
        const bool wr = W >= 4 ? any[0] != 0 : (nw[2 * p] | nw[2 * p + 1]) != 0;
        if (wr) {
          cp2[p] = make_ulonglong2(nw[2 * p], nw[2 * p + 1]);
          sp2[p] = make_ulonglong2(sv[2 * p] | nw[2 * p], sv[2 * p + 1] | nw[2 * p + 1]);
        }
      }
      // Padding vertices (v >= n) have seen = ~0 and next = 0: never new,
      // never unvisited.
      if (sub == 0) {
#pragma unroll
        for (int j = 0; j < V; j++) {
          const int32_t v = (int32_t)(v0 + j);
          if (any[j]) {
            nib |= 1u << j;
            const int d = out_deg(a, v);
            acc.fv += 1;
            acc.fe += (unsigned)d;
            acc.scout += (unsigned)(d > 0 ? d : 1);
            acc.lanes += pc[j];
          }
          if (unv[j]) {  // still misses a live lane (or started done: in-degree 0)
            if (!((done_in >> j) & 1u)) {
              acc.uv += 1;
              acc.ue += (unsigned)(a.in_row_ptr[v + 1] - a.in_row_ptr[v]);
            }
          } else {
            dnib |= 1u << j;
          }
        }
      }
    }
    if (sub != 0) dnib = 0;  // the first thread of a vertex reports its bits
    // Frontier and done bitmaps: LPW lanes (32 vertices) per word.
    const int pos = ((lane / TPV) * V) & 31;
    const unsigned x = group_or_u32<LPW>(nib << pos);
    const unsigned y = group_or_u32<LPW>(dnib << pos);
    if ((lane % LPW) == 0 && v0 < a.npad) {
      fb[v0 >> 5] = x;
      a.done[v0 >> 5] = y;
    }
  }
  acc.flush(c_out, sub);
}

// Fold of a pull step's heavy vertices (in-degree >= knobs.heavy_degree),
// whose chunks merged their arrivals in next[] (the pull kernel folded the
// light ones): one group of TPV threads per chunk entry, acting on the
// first chunk of each vertex not yet done.  Same arithmetic as the
// finalize pass; bitmap bits are OR-ed into the words the pull wrote.
template <int W>
__global__ void __launch_bounds__(kBlock) ms_fold_heavy_kernel(MsArgs a) {
  constexpr int WPT = W > 8 ? 8 : 4;
  constexpr int TPV = W > WPT ? W / WPT : 1;
  constexpr int WPV = W < WPT ? W : WPT;
  const SchedState st = *a.st;
  if (st.done || st.mode != DAWN_MODE_PULL || !pull_folds_light<W>(a)) return;
  const int lv = st.step;
  uint64_t* cur = a.hist + (int64_t)(lv % a.H) * a.npad * W;
  uint32_t* fb = a.fbits + (int64_t)(lv % a.H) * a.fwords;
  unsigned long long* c_out = ms_stripe(a, lv);
  const int lane = lane_id();
  const int sub = lane % TPV;
  uint64_t alive_prev[WPV];
  const unsigned long long* al = a.ring[(lv - 1) % kRing].alive;
#pragma unroll
  for (int i = 0; i < WPV; i++) alive_prev[i] = al[sub * WPT + i];
  MsAcc<W, WPV> acc;
  // Pull step: the pull kernel folded the light vertices.  Fold the heavy
  // ones (merged by their chunks in next[]): one group of TPV threads per
  // chunk entry, acting on the first chunk of each vertex not yet done.
  const int64_t gstride = (int64_t)gridDim.x * (kBlock / TPV);
  for (int64_t cb = (int64_t)blockIdx.x * (kBlock / TPV) + (threadIdx.x / kWave) * (kWave / TPV);
       cb < a.num_hc; cb += gstride) {  // wave-uniform
    const int64_t c = cb + lane / TPV;
    int32_t v = 0;
    bool work = false;
    if (c < a.num_hc) {
      v = a.hc_v[c];
      work = (c == 0 || a.hc_v[c - 1] != v) && !((a.done[v >> 5] >> (v & 31)) & 1u);
    }
    uint64_t nx[WPV], sv[WPV], nw[WPV];
#pragma unroll
    for (int i = 0; i < WPV; i++) nx[i] = sv[i] = 0;
    if (work) {
      load_words<WPV>(a.next + (int64_t)v * W + sub * WPV, nx);
      load_words<WPV>(a.seen + (int64_t)v * W + sub * WPV, sv);
    }
    uint64_t any = 0, unv = 0, had = 0;
    unsigned pc = 0;
#pragma unroll
    for (int i = 0; i < WPV; i++) {
      nw[i] = nx[i] & ~sv[i];
      any |= nw[i];
      had |= nx[i];
      unv |= ~(sv[i] | nw[i]) & alive_prev[i];
      pc += (unsigned)__popcll(nw[i]);
    }
    const uint64_t own = any;
    if constexpr (TPV > 1) {  // every lane of the wave is here
      any = group_or_u64<TPV>(any);
      unv = group_or_u64<TPV>(unv);
      pc = group_sum_u32<TPV>(pc);
    }
    if (work) {
      if (any) store_words<WPV>(cur + (int64_t)v * W + sub * WPV, nw);
      if (own) {
        uint64_t sn[WPV];
#pragma unroll
        for (int i = 0; i < WPV; i++) sn[i] = sv[i] | nw[i];
        store_words<WPV>(a.seen + (int64_t)v * W + sub * WPV, sn);
      }
      if (had) {
        uint64_t z[WPV];
#pragma unroll
        for (int i = 0; i < WPV; i++) z[i] = 0;
        store_words<WPV>(a.next + (int64_t)v * W + sub * WPV, z);
      }
#pragma unroll
      for (int i = 0; i < WPV; i++) acc.alive[i] |= nw[i];
      if (sub == 0) {
        const uint32_t bit = 1u << (v & 31);
        if (any) {
          atomicOr(&fb[v >> 5], bit);
          const int d = out_deg(a, v);
          acc.fv += 1;
          acc.fe += (unsigned)d;
          acc.scout += (unsigned)(d > 0 ? d : 1);
          acc.lanes += pc;
        }
        if (unv) {
          acc.uv += 1;
          acc.ue += (unsigned)(a.in_row_ptr[v + 1] - a.in_row_ptr[v]);
        } else {
          atomicOr(&a.done[v >> 5], bit);
        }
      }
    }
  }
  acc.flush(c_out, sub);
}

// Builds the frontier queue of the previous level from its bitmap, only
// for a PUSH step: one entry per kPushChunk out-edges of each frontier
// vertex, so the push step's work per entry is bounded whatever the degree.
__global__ void __launch_bounds__(kBlock) ms_compact_kernel(MsArgs a) {
  // Entries are staged in LDS and the workgroup reserves its queue range
  // with one atomic per kCompactCap entries (the queue length is one
  // address: an atomic per tile from every workgroup would serialise).
  constexpr int kCompactCap = 2048;
  __shared__ int32_t s_v[kCompactCap];
  __shared__ int32_t s_b[kCompactCap];
  __shared__ int s_scan[kBlock / kWave];
  __shared__ unsigned s_n, s_base;
  const SchedState st = *a.st;
  if (st.done || st.mode != DAWN_MODE_PUSH) return;
  const int lv = st.step;
  const uint32_t* fb = a.fbits + (int64_t)((lv - 1) % a.H) * a.fwords;
  int32_t* q = a.q;
  int32_t* qb = a.qb;
  unsigned* qlen = &a.ring[(lv - 1) % kRing].qlen;
  if (threadIdx.x == 0) s_n = 0u;
  __syncthreads();
  auto flush = [&]() {  // block-uniform
    __syncthreads();
    const unsigned cnt = s_n;
    if (threadIdx.x == 0) s_base = cnt ? atomicAdd(qlen, cnt) : 0u;
    __syncthreads();
    for (unsigned i = threadIdx.x; i < cnt; i += kBlock) {
      q[s_base + i] = s_v[i];
      qb[s_base + i] = s_b[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
  };
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < a.npad; b0 += (int64_t)gridDim.x * kBlock) {
    const int32_t v = (int32_t)(b0 + threadIdx.x);
    const bool isf = v < a.n && ((fb[v >> 5] >> (v & 31)) & 1u);
    if (!__syncthreads_or(isf)) continue;  // block-uniform
    const int d = isf ? out_deg(a, v) : 0;
    const int k = (d + kPushChunk - 1) / kPushChunk;
    int total;
    const int off = block_exclusive_scan(k, s_scan, &total);
    if (s_n + (unsigned)total > (unsigned)kCompactCap) flush();  // s_n, total block-uniform
    const int32_t rb = isf ? a.row_ptr[v] : 0;
    if (total > kCompactCap) {
      // More chunks than the stage holds: straight to the queue.
      if (threadIdx.x == 0) s_base = atomicAdd(qlen, (unsigned)total);
      __syncthreads();
      for (int i = 0; i < k; i++) {
        q[s_base + off + i] = v;
        qb[s_base + off + i] = rb + i * kPushChunk;
      }
    } else {
      const unsigned at = s_n;
      for (int i = 0; i < k; i++) {
        s_v[at + off + i] = v;
        s_b[at + off + i] = rb + i * kPushChunk;
      }
      __syncthreads();
      if (threadIdx.x == 0) s_n = at + (unsigned)total;
    }
    __syncthreads();
  }
  flush();
}

// Rows of isolated sources: INF except 0 at the source.  rs holds (row,
// source) pairs; row k goes to dist + rs[2k] * stride.  Each workgroup
// writes whole chunks of kIsoChunk 16-byte vectors of one row in order
// (chunks taken grid-stride), with plain stores: measured on a 16 GB fill,
// chunks of 256 KB written in order reach 6.2 TB/s, as the runtime's own
// fill does, where vectors taken grid-stride over the whole buffer reach
// 5.1-5.8 TB/s and streaming stores 5-8% less than plain ones.
constexpr int64_t kIsoChunk = 16384;
__global__ void __launch_bounds__(kBlock) ms_isolated_rows_kernel(int32_t* dist, int64_t stride,
                                                                  int32_t n, const int32_t* rs,
                                                                  int32_t count) {
  const int64_t per_row = ((int64_t)n + 3) / 4;
  const int64_t cpr = (per_row + kIsoChunk - 1) / kIsoChunk;
  const int64_t chunks = cpr * count;
  const bool vec = (stride % 4 == 0) && (((uintptr_t)dist & 15) == 0);
  for (int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const int64_t k = ch / cpr;
    const int64_t t0 = (ch - k * cpr) * kIsoChunk;
    const int64_t t1 = t0 + kIsoChunk < per_row ? t0 + kIsoChunk : per_row;
    const int32_t row = rs[2 * k], src = rs[2 * k + 1];
    int32_t* rp = dist + (int64_t)row * stride;
    for (int64_t t = t0 + threadIdx.x; t < t1; t += kBlock) {
      const int64_t c = t * 4;
      if (vec && c + 4 <= n) {
        v4i32 val = {kInf, kInf, kInf, kInf};
        if (src >= c && src < c + 4) val[src - c] = 0;
        *reinterpret_cast<v4i32*>(rp + c) = val;
      } else {
        for (int j = 0; j < 4 && c + j < n; j++) rp[c + j] = c + j == src ? 0 : kInf;
      }
    }
  }
}

// out = a + b on N-bit numbers held as bit slices (slice t: bit t of every
// lane's number, one 64-bit mask per slice); out has N + 1 slices.
template <int N>
__device__ __forceinline__ void bitsliced_add(const uint64_t (&a)[N], const uint64_t (&b)[N],
                                              uint64_t (&out)[N + 1]) {
  uint64_t c = 0;
#pragma unroll
  for (int t = 0; t < N; t++) {
    const uint64_t x = a[t] ^ b[t];
    out[t] = x ^ c;
    c = (a[t] & b[t]) | (c & x);
  }
  out[N] = c;
}

// Waves per SIMD the distance pass asks the compiler for: 4 (at most 128
// VGPRs) where that needs no spills, else no constraint.
constexpr int final_waves(int W) { return W == 1 || W == 16 ? 4 : 1; }

// Turns levels [lo, hi) of the history into distances.  from_state: lo and
// hi come from the batch / scheduler state and the pass runs only once the
// traversal is done.  The first window (lo == 0) writes every entry of the
// rows, INF where no level so far holds the lane; later windows (long
// traversals, flushed before the history ring wraps) overwrite the entries
// of the lanes that arrive in them.  So the pass never reads seen[].
// Wave w of a workgroup handles word w % W of a 256-vertex tile (4 vertices
// per lane), so the 64 rows of one word are written with 16-byte stores,
// 1 KB per wave per row, and the W waves of a tile share its cache lines.
// For W > 4 the W/4 workgroups of a tile are placed on one XCD (workgroup
// ids x, x + 8, ... of the round-robin dispatch), so each record line is
// fetched into one L2, not into W/4 of them.
template <int W>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(final_waves(W))))
ms_final_kernel(MsArgs a, MsOut o, int from_state, int lo_arg, int hi_arg) {
  constexpr int KS = level_slices(W);
  static_assert((1 << level_slices(W)) >= hist_slots(W), "window levels fit the slices");
  __shared__ unsigned long long s_sum[64 * W], s_reach[64 * W], s_edges[64 * W];
  int lo = lo_arg, hi = hi_arg;
  if (from_state) {
    const SchedState st = *a.st;
    if (!st.done) return;
    lo = a.batch->lo;
    hi = st.levels;
  }
  const int nl = a.batch->nlanes;
  int32_t* const dist = a.batch->dist;
  const int64_t stride = a.batch->stride;
  __shared__ int32_t s_row[64 * W];
  for (int i = threadIdx.x; i < 64 * W; i += kBlock) {
    s_sum[i] = 0;
    s_reach[i] = 0;
    s_edges[i] = 0;
    s_row[i] = i < nl ? a.batch->row[i] : 0;
  }
  __syncthreads();
  // Levels are stored relative to lo in K bit slices: S_k = OR of the
  // level words whose (l - lo) has bit k set.  hi - lo <= H <= 2^KS.
  int K = 0;
  while ((1 << K) < hi - lo) K++;
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  // Lane word of this wave, and its tile, over the G words that hold the
  // batch's lanes (G = the power of two >= its live words, at most W: a
  // partly filled last batch spreads its few words over every wave instead
  // of leaving the waves of empty words idle).  G <= 4: the workgroup's four
  // waves cover 4/G tiles; G = 8: two workgroups share a tile.  G > 4: the
  // grid is a multiple of kXcds * G/4; workgroup x runs on XCD x % kXcds,
  // and r = x / kXcds counts that XCD's workgroups, G/4 per tile.
  int G = 1;
  while (G < W && 64 * G < nl) G <<= 1;
  const int BPT = G > 4 ? G / 4 : 1;  // workgroups per tile
  const int TPB = G < 4 ? 4 / G : 1;  // vertex tiles per workgroup
  const unsigned xr = blockIdx.x / kXcds;
  const int wi = G > 4 ? (int)(xr % BPT) * 4 + wv : wv % G;
  const int tile = G > 4 ? 0 : wv / G;
  const int64_t tile0 = G > 4 ? (int64_t)(xr / BPT) * kXcds + blockIdx.x % kXcds
                              : (int64_t)blockIdx.x * TPB + tile;
  const int rows = min(64, nl - 64 * wi);  // lanes of this word in the batch
  // The first window writes every (lane, vertex): 16-byte row stores when
  // the rows are aligned.
  const bool first = lo == 0;
  const bool vec_ok = dist && (stride % 4 == 0) && (((uintptr_t)dist & 15) == 0) && first;
  const int64_t wstride = (int64_t)(gridDim.x / BPT) * TPB * 256;
  // A word without lanes in the batch (a partly filled last batch) has no
  // rows and no sums: its wave skips the tiles (no barrier inside the loop).
  const int64_t tiles_end = rows > 0 ? a.n : 0;
  for (int64_t base = tile0 * 256; base < tiles_end; base += wstride) {
    const int64_t v0 = base + 4 * lane;
    uint64_t F[4], S[4][KS];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      F[j] = 0;
#pragma unroll
      for (int k = 0; k < KS; k++) S[j][k] = 0;
    }
    // Rows are in the original order; the records live at q[j], the hub-
    // first position of v0 + j (identity without a view).
    const bool ident = a.perm == nullptr;
    int32_t q[4];  // vertex ids < n < 2^31
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int32_t v = (int32_t)(v0 + j);
      q[j] = ident || v >= a.n ? v : a.perm[v];
    }
    if (v0 + 3 < a.n && ident) {
      for (int l = lo; l < hi; l++) {
        // Level words are valid only for the level's frontier vertices.
        const unsigned fm = (a.fbits[(int64_t)(l % a.H) * a.fwords + (v0 >> 5)] >> (v0 & 31)) & 15u;
        if (!fm) continue;
        const uint64_t* hb = a.hist + (int64_t)(l % a.H) * a.npad * W;
        uint64_t w[4];
        if constexpr (W == 1) {
          const ulonglong2* hp = reinterpret_cast<const ulonglong2*>(hb + v0);
          const ulonglong2 h01 = hp[0], h23 = hp[1];
          w[0] = h01.x; w[1] = h01.y; w[2] = h23.x; w[3] = h23.y;
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (!((fm >> j) & 1u)) w[j] = 0;
        } else {
#pragma unroll
          for (int j = 0; j < 4; j++) w[j] = ((fm >> j) & 1u) ? hb[(v0 + j) * W + wi] : 0ull;
        }
        const int r = l - lo;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          F[j] |= w[j];
#pragma unroll
          for (int k = 0; k < KS; k++)
            if ((r >> k) & 1) S[j][k] |= w[j];
        }
      }
    } else if (v0 + 3 < a.n) {
      // Outside the hub range the four positions are consecutive: their
      // bitmap bits come from one (or two adjacent) words.
      const bool run = q[1] == q[0] + 1 && q[2] == q[0] + 2 && q[3] == q[0] + 3;
      const int32_t qw = q[0] >> 5;
      const int qb = (int)(q[0] & 31);
      for (int l = lo; l < hi; l++) {
        const uint32_t* fbl = a.fbits + (int64_t)(l % a.H) * a.fwords;
        unsigned fm = 0;
        if (run) {
          fm = fbl[qw] >> qb;
          if (qb > 28) fm |= fbl[qw + 1] << (32 - qb);
          fm &= 15u;
        } else {
#pragma unroll
          for (int j = 0; j < 4; j++) fm |= ((fbl[q[j] >> 5] >> (q[j] & 31)) & 1u) << j;
        }
        if (!fm) continue;
        const uint64_t* hb = a.hist + (int64_t)(l % a.H) * a.npad * W;
        const int r = l - lo;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint64_t w = ((fm >> j) & 1u) ? hb[(int64_t)q[j] * W + wi] : 0ull;
          F[j] |= w;
#pragma unroll
          for (int k = 0; k < KS; k++)
            if ((r >> k) & 1) S[j][k] |= w;
        }
      }
    } else {
      for (int j = 0; j < 4; j++) {
        if (v0 + j >= a.n) continue;
        const int64_t v = q[j];
        for (int l = lo; l < hi; l++) {
          if (!((a.fbits[(int64_t)(l % a.H) * a.fwords + (v >> 5)] >> (v & 31)) & 1u)) continue;
          const uint64_t w = a.hist[(int64_t)(l % a.H) * a.npad * W + v * W + wi];
          const int r = l - lo;
          F[j] |= w;
#pragma unroll
          for (int k = 0; k < KS; k++)
            if ((r >> k) & 1) S[j][k] |= w;
        }
      }
    }
    if (dist && v0 < a.n) {
      const bool full = vec_ok && v0 + 3 < a.n;
      // Rows 0-31 of the word from the low halves, 32-63 from the high
      // halves: one bit-field extract per (vertex, slice).  Slices k >= K
      // are zero, so all KS are summed without a test.
#pragma unroll
      for (int h = 0; h < 2; h++) {
        uint32_t Sh[4][KS], Fh[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          Fh[j] = (uint32_t)(F[j] >> (32 * h));
#pragma unroll
          for (int k = 0; k < KS; k++) Sh[j][k] = (uint32_t)(S[j][k] >> (32 * h));
        }
        const int bmax = min(32, rows - 32 * h);
        // Lane rb0 + bb of the batch writes output row s_row[rb0 + bb].
        const int rb0 = 64 * wi + 32 * h;
        // Level of vertex j in row bb relative to lo (slices k >= K are zero).
        auto rel = [&](int j, int bb) {
          uint32_t x = 0;
#pragma unroll
          for (int k = 0; k < KS; k++) x |= __builtin_amdgcn_ubfe(Sh[j][k], bb, 1) << k;
          return (int32_t)x;
        };
        if (full && bmax == 32) {
          // Whole 16-byte row segments, no per-row branch: the level or INF by a
          // bit select on the sign-extended "reached" bit; streaming stores
          // (rows are not read back here; plain stores measured 21% slower
          // in this pass on RMAT-24, while its steps' records share the L2).
#pragma unroll 1
          for (int bb = 0; bb < 32; bb++) {
            v4i32 val;
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const int32_t m = __builtin_amdgcn_sbfe((int32_t)Fh[j], bb, 1);
              val[j] = (m & (lo + rel(j, bb))) | (~m & kInf);
            }
            __builtin_nontemporal_store(
                val, reinterpret_cast<v4i32*>(dist + (int64_t)s_row[rb0 + bb] * stride + v0));
          }
        } else {
          for (int bb = 0; bb < bmax; bb++) {
            int32_t* row = dist + (int64_t)s_row[rb0 + bb] * stride + v0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const bool in = __builtin_amdgcn_ubfe(Fh[j], bb, 1) != 0;
              const int32_t d = in ? lo + rel(j, bb) : kInf;
              if (full) row[j] = d;
              else if (v0 + j < a.n && (in || first)) row[j] = d;
            }
          }
        }
      }
    }
    // Per-lane sums: sum_b = lo * reach_b + sum_k 2^k * |{v : bit b of S_k}|,
    // counted with 64x64 bit transposes (lane b receives lane-b bits).
    unsigned long long sum = 0, reach = 0, edges = 0;
    // edges_b = sum_i [bit i of tf] * deg_i, bit-sliced over the degree:
    // sum_t 2^t * popcount(tf & ballot(bit t of deg)), over the degree bits
    // of this wave's vertices only (hubs are rare).
    int dg[4] = {0, 0, 0, 0};
    int dbits = 0;
    if (o.want_edges) {
#pragma unroll
      for (int j = 0; j < 4; j++) dg[j] = v0 + j < a.n ? out_deg(a, q[j]) : 0;
      const int dmax = wave_max_xor(max(max(dg[0], dg[1]), max(dg[2], dg[3])));
      dbits = dmax > 0 ? 32 - __builtin_clz((unsigned)dmax) : 0;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint64_t tf = wave_transpose64(F[j]);
      const unsigned long long rc = (unsigned long long)__popcll(tf);
      reach += rc;
      sum += rc * (unsigned long long)lo;
      for (int t = 0; t < dbits; t++) {
        const uint64_t dm = __ballot((dg[j] >> t) & 1);
        edges += (unsigned long long)__popcll(tf & dm) << t;
      }
    }
    // The four vertices' relative levels per row, added bit-sliced in the
    // thread (KS+2 bit slices of the row sums), then one transpose per slice.
    {
      uint64_t s01[KS + 1], s23[KS + 1], tot[KS + 2];
      bitsliced_add<KS>(S[0], S[1], s01);
      bitsliced_add<KS>(S[2], S[3], s23);
      bitsliced_add<KS + 1>(s01, s23, tot);
#pragma unroll
      for (int t = 0; t < KS + 2; t++) {
        const uint64_t m = wave_transpose64(tot[t]);
        sum += (unsigned long long)__popcll(m) << t;
      }
    }
    const int gl = 64 * wi + lane;
    if (sum) atomicAdd(&s_sum[gl], sum);
    if (reach) atomicAdd(&s_reach[gl], reach);
    if (edges) atomicAdd(&s_edges[gl], edges);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nl; b += kBlock) {
    if (s_sum[b]) atomicAdd(&o.sum[b], s_sum[b]);
    if (s_reach[b]) atomicAdd(&o.reached[b], s_reach[b]);
    if (s_edges[b]) atomicAdd(&o.edges[b], s_edges[b]);
  }
}

// Lane words per vertex for a traversal of S sources: the knob, but no
// wider than the sources need.
int words_for(const dawn_graph* g, int32_t S) {
  const int knob = std::max(1, std::min(kMaxWords, (int)g->knobs.lane_words));
  int w = 1;
  while (w < knob && 64 * w < S) w <<= 1;
  return w;
}

// The batches of a traversal of S live sources: full ones of 64·W lanes
// (the first S_main sources) and, when the last batch is partly filled and
// its lanes fit fewer words, that batch at Wt words (run_msbfs_w makes the
// same split, and runs the last batch at W if its workspace does not fit).
struct BatchSplit {
  int W = 1, Wt = 1;
  int32_t S_main = 0;
};
BatchSplit split_batches(const dawn_graph* g, int32_t S) {
  BatchSplit b;
  b.W = words_for(g, S);
  const int32_t full = S / (64 * b.W) * (64 * b.W);
  b.Wt = 1;
  while (b.Wt < b.W && 64 * b.Wt < S - full) b.Wt <<= 1;
  b.S_main = full > 0 && b.Wt < b.W ? full : S;
  return b;
}

int ensure_ms(dawn_graph* g, MsWorkspace*& slot, int W) {
  if (!slot) slot = new MsWorkspace();
  MsWorkspace* ws = slot;
  if (ws->n == g->n && ws->words == W && ws->seen.p) return DAWN_OK;
  // New layout: captured graphs hold the old buffers' addresses.
  for (auto& kv : ws->graphs) (void)hipGraphExecDestroy(kv.second);
  ws->graphs.clear();
  ws->seen_bak.release();
  ws->distbuf.release();
  for (DevBuf<uint64_t>* b : {&ws->seen, &ws->next, &ws->hist}) b->release();
  ws->n = g->n;
  ws->words = 0;  // set once every buffer is in place
  ws->slots = hist_slots(W);
  ws->npad = ((int64_t)g->n + 63) / 64 * 64;
  const size_t words = (size_t)ws->npad * W;
  int rc = ws->seen.alloc(words);
  if (!rc) rc = ws->next.alloc(words);
  if (!rc) rc = ws->hist.alloc((size_t)ws->slots * words);
  if (rc) {  // out of memory: drop the partial layout (the next call starts over)
    for (DevBuf<uint64_t>* b : {&ws->seen, &ws->next, &ws->hist}) b->release();
    (void)hipGetLastError();
    return rc;
  }
  DAWN_TRY(ws->fbits.alloc((size_t)ws->slots * (ws->npad / 32)));
  DAWN_TRY(ws->done.alloc((size_t)(ws->npad / 32)));
  // Queue entries per level: one per kPushChunk out-edges of each frontier
  // vertex, at most n + m / kPushChunk.
  ws->qcap = (int64_t)g->n + g->m / kPushChunk + 64;
  DAWN_TRY(ws->q.alloc((size_t)ws->qcap));
  DAWN_TRY(ws->qb.alloc((size_t)ws->qcap));
  DAWN_TRY(ws->ring.alloc(kRing));
  DAWN_TRY(ws->ring_bak.alloc(kRing));
  DAWN_TRY(ws->part.alloc(kMsPartU64));
  DAWN_TRY(ws->part_bak.alloc(kMsPartU64));
  DAWN_TRY(ws->st.alloc(1));
  DAWN_TRY(ws->batch.alloc(1));
  DAWN_TRY(ws->trace_feat.alloc((size_t)kTraceMax * 8));
  DAWN_TRY(ws->trace_mode.alloc(kTraceMax));
  DAWN_TRY(ws->lane_stats.alloc(3 * kMaxLanes));
  DAWN_HIP(hipMemset(ws->next.p, 0, sizeof(uint64_t) * words));
  if (!ws->h_state) DAWN_HIP(hipHostMalloc((void**)&ws->h_state, sizeof(SchedState)));
  if (!ws->h_batch) DAWN_HIP(hipHostMalloc((void**)&ws->h_batch, sizeof(MsBatch)));
  for (hipEvent_t* e : {&ws->ev0, &ws->ev1, &ws->e_a, &ws->e_b})
    if (!*e) DAWN_HIP(hipEventCreate(e));
  ws->words = W;
  return DAWN_OK;
}

MsArgs make_args(dawn_graph* g, MsWorkspace* ws) {
  MsArgs a{};
  a.n = g->n;
  a.H = ws->slots;
  a.npad = ws->npad;
  a.m = g->m;
  a.zero_in = g->zero_in;
  const MsView* vw = g->view;
  a.row_ptr = vw ? vw->row_ptr : g->row_ptr;
  a.col = vw ? vw->col : g->col;
  a.in_row_ptr = vw ? vw->in_row_ptr : g->in_row_ptr;
  a.in_col = vw ? vw->in_col : g->in_col;
  a.perm = vw ? vw->perm : nullptr;
  a.seen = ws->seen.p;
  a.next = ws->next.p;
  a.fbits = ws->fbits.p;
  a.done = ws->done.p;
  a.fwords = ws->npad / 32;
  a.hist = ws->hist.p;
  a.q = ws->q.p;
  a.qb = ws->qb.p;
  a.qcap = ws->qcap;
  a.hc_v = vw ? vw->hc_v : g->hc_v;
  a.hc_beg = vw ? vw->hc_beg : g->hc_beg;
  a.hc_end = vw ? vw->hc_end : g->hc_end;
  a.num_hc = vw ? vw->num_hc : g->num_hc;
  a.oc_v = vw ? vw->oc_v : g->oc_v;
  a.oc_beg = vw ? vw->oc_beg : g->oc_beg;
  a.oc_end = vw ? vw->oc_end : g->oc_end;
  a.num_oc = vw ? vw->num_oc : g->num_oc;
  a.heavy = g->knobs.heavy_degree;
  a.variant = g->knobs.pull_variant;
  a.ring = ws->ring.p;
  a.part = ws->part.p;
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
  p.uv_base = g->zero_in;
  return p;
}

// Dense step kernels: grid capped by knobs.dense_blocks, grid-stride beyond.
unsigned ms_grid(const dawn_graph* g) {
  return grid_for(std::max<int64_t>(g->n, 1), kBlock, (unsigned)g->knobs.dense_blocks);
}
// Finalize pass: 1024/W vertices per workgroup and iteration (128 at
// W = 16); at most 2048 workgroups (8 per CU), since each flushes its
// counters with atomics.
unsigned ms_fin_grid(const dawn_graph* g, int W) {
  return grid_for(std::max<int64_t>(g->n, 1), W > 8 ? kBlock / 2 : kBlock * 4 / W, 2048u);
}
// Distance pass: 1024/W vertices per workgroup (W > 4: W/4 workgroups per
// 256-vertex tile, on one XCD, so the grid is a multiple of kXcds * W/4), at
// most 32768 workgroups (RMAT-24 at W = 16, 4096-source passes alternating
// on one box: 16384 makes the pass about 1.7% slower, 65536 the same).
unsigned ms_final_grid(const dawn_graph* g, int W) {
  unsigned grid = grid_for(std::max<int64_t>(g->n, 1), kBlock * 4 / W, 32768u);
  if (W > 4) {
    const unsigned unit = (unsigned)(kXcds * W / 4);
    grid = std::max(unit, (grid + unit - 1) / unit * unit);
  }
  return grid;
}

// Stages one batch's sources (taken mod n, then to their position in the
// hub-first view if there is one) into the pinned batch block.
void stage_batch(MsWorkspace* ws, const int32_t* sources, int32_t lanes, const MsView* view) {
  MsBatch& b = *ws->h_batch;
  std::memset(&b, 0, sizeof b);
  b.nlanes = lanes;
  b.lo = 0;
  for (int i = 0; i < kMaxWords; i++) {
    const int l = lanes - 64 * i;
    b.lanemask[i] = l >= 64 ? ~0ull : l <= 0 ? 0ull : ((1ull << l) - 1);
  }
  for (int i = 0; i < lanes; i++) {
    int64_t s = sources[i] % (int64_t)ws->n;
    if (s < 0) s += ws->n;
    b.src[i] = view ? view->h_perm[s] : (int32_t)s;
    b.row[i] = i;  // lane i -> row i of the output (callers with a row map override)
  }
}

// Sources without out-edges are not traversed (the reference skips them,
// src/algorithm/cpu/mssp.cpp:16-18): their row is INF except 0 at the
// source.  Only the other sources take lanes, so every lane of a batch
// traverses.  Pairs (output row, source vertex) of the isolated sources.
struct IsoRows {
  const int32_t* pairs = nullptr;
  int32_t count = 0;
};

template <int W>
void launch_pull(const MsArgs& a, unsigned grid, hipStream_t s) {
  switch (a.variant) {
    case 0: ms_pull_kernel<W, 0><<<grid, kBlock, 0, s>>>(a); break;
    case 1: ms_pull_kernel<W, 1><<<grid, kBlock, 0, s>>>(a); break;
    case 2: ms_pull_kernel<W, 2><<<grid, kBlock, 0, s>>>(a); break;
    default: ms_pull_kernel<W, 3><<<grid, kBlock, 0, s>>>(a); break;
  }
}

// Bitmap -> queue compaction: a streaming scan, so few workgroups (each
// reserves its queue range with one atomic on the queue length).
inline unsigned compact_grid(unsigned grid) { return std::min(grid, 1024u); }

template <int W>
void launch_step(const MsArgs& a, unsigned grid, unsigned fin_grid, hipStream_t s) {
  ms_compact_kernel<<<compact_grid(grid), kBlock, 0, s>>>(a);
  ms_push_kernel<W><<<grid, kBlock, 0, s>>>(a);
  ms_topdown_kernel<W><<<grid, kBlock, 0, s>>>(a);
  launch_pull<W>(a, grid, s);
  ms_finalize_kernel<W><<<fin_grid, kBlock, 0, s>>>(a);
  ms_fold_heavy_kernel<W><<<fin_grid, kBlock, 0, s>>>(a);
}

// Runs the batches of an unweighted multi-source traversal, 64·W sources
// per batch.
//   dist: caller buffer (S*n int32, host/device) or null
//   sums/reached/edges: per source (host) or null
//   Processes batches first, first + step, ... on workspace ws and stream s.
template <int W>
int run_msbfs_slot(dawn_graph* g, MsWorkspace* ws, hipStream_t s, const int32_t* sources,
                   const int32_t* rowidx, int32_t S, int32_t first, int32_t step,
                   const IsoRows* iso, int32_t* dist, int64_t* sums, int64_t* reached,
                   int64_t* edges, int32_t policy, dawn_stats* stats, bool rows) {
  constexpr int kB = 64 * W;
  const bool dist_dev = dist && is_device_pointer(dist);
  // Rows of a host output, or (rows) of a call without output, go through
  // the slot's row buffer; without either the distance pass only sums.
  const bool buf_rows = !dist_dev && (dist || rows);
  if (buf_rows) DAWN_TRY(ws->distbuf.alloc((size_t)kB * g->n));

  MsArgs a = make_args(g, ws);
  SchedParams p = make_params(g, policy);
  SchedTrace tr{ws->trace_feat.p, ws->trace_mode.p};
  const unsigned grid = ms_grid(g);
  const unsigned fin_grid = ms_fin_grid(g, W);
  const unsigned fgrid = ms_final_grid(g, W);
  unsigned long long* ls = ws->lane_stats.p;

  // Graph of one batch: fill, seed, kChunk steps, final pass (no-op unless
  // the traversal finished within the chunk).  Sources and the output
  // pointer live in the device batch block, so one graph serves all batches.
  constexpr int kChunk = chunk_steps(W);
  auto get_graph = [&](int want_edges, hipGraphExec_t* exec) -> int {
    const auto key = std::make_tuple(W, (int)policy, want_edges, (const void*)p.mlp);
    auto it = ws->graphs.find(key);
    if (it != ws->graphs.end()) {
      *exec = it->second;
      return DAWN_OK;
    }
    MsOut o{ls, ls + kMaxLanes, ls + 2 * kMaxLanes, want_edges};
    hipGraph_t graph;
    DAWN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    ms_fill_kernel<W><<<grid_for(ws->npad * W), kBlock, 0, s>>>(a, o);
    ms_seed_kernel<W><<<1, kMaxLanes, 0, s>>>(a);
    for (int i = 0; i < kChunk; i++) {
      ms_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
      launch_step<W>(a, grid, fin_grid, s);
    }
    ms_final_kernel<W><<<fgrid, kBlock, 0, s>>>(a, o, 1, 0, 0);
    DAWN_HIP(hipStreamEndCapture(s, &graph));
    hipError_t e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DAWN_HIP(e);
    ws->graphs[key] = *exec;
    return DAWN_OK;
  };

  const int want_edges = (stats || edges) ? 1 : 0;
  std::vector<unsigned long long> h_ls(3 * kMaxLanes);
  std::vector<int32_t> h_modes(kTraceMax);
  float total_ms = 0.f;
  // Output row of entry k of sources.
  auto row_of = [&](int64_t k) -> int64_t { return rowidx ? rowidx[k] : k; };
  for (int32_t b0 = first * kB; b0 < S; b0 += step * kB) {
    const int32_t lanes = std::min<int32_t>(kB, S - b0);
    int32_t* out_base = dist_dev ? dist : buf_rows ? ws->distbuf.p : nullptr;
    hipGraphExec_t exec;
    DAWN_TRY(get_graph(want_edges, &exec));
    stage_batch(ws, sources + b0, lanes, g->view);
    if (dist_dev)
      for (int i = 0; i < lanes; i++) ws->h_batch->row[i] = (int32_t)row_of(b0 + i);
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
      MsOut o{ls, ls + kMaxLanes, ls + 2 * kMaxLanes, want_edges};
      const int stp = ws->h_state->step;
      if (stp > lo) {
        ms_final_kernel<W><<<fgrid, kBlock, 0, s>>>(a, o, 0, lo, stp);
        lo = stp;
      }
      for (int i = 0; i < kChunk; i++) {
        ms_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
        launch_step<W>(a, grid, fin_grid, s);
      }
      guard_steps += kChunk;
      DAWN_HIP(hipMemcpyAsync(ws->h_state, ws->st.p, sizeof(SchedState), hipMemcpyDeviceToHost, s));
      DAWN_HIP(hipStreamSynchronize(s));
      if (ws->h_state->done) {
        ms_final_kernel<W><<<fgrid, kBlock, 0, s>>>(a, o, 0, lo, ws->h_state->levels);
        break;
      }
      if (guard_steps > (int64_t)g->n + 2 * kChunk) return DAWN_ERR_HIP;
    }
    DAWN_HIP(hipGetLastError());
    DAWN_HIP(hipEventRecord(ws->ev1, s));
    DAWN_HIP(hipMemcpyAsync(h_ls.data(), ls, sizeof(unsigned long long) * 3 * kMaxLanes,
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
      // Lane rows of the row buffer to their rows of the host output, runs
      // of consecutive rows in one copy.
      for (int i = 0; i < lanes;) {
        const int64_t r0 = row_of(b0 + i);
        int j = i + 1;
        while (j < lanes && row_of(b0 + j) == r0 + (j - i)) j++;
        DAWN_HIP(hipMemcpyAsync(dist + r0 * g->n, ws->distbuf.p + (int64_t)i * g->n,
                                sizeof(int32_t) * (size_t)(j - i) * g->n, hipMemcpyDeviceToHost,
                                s));
        i = j;
      }
      DAWN_HIP(hipStreamSynchronize(s));
    }
    for (int i = 0; i < lanes; i++) {
      const int64_t r = row_of(b0 + i);
      if (sums) sums[r] = (int64_t)h_ls[i];
      if (reached) reached[r] = (int64_t)h_ls[kMaxLanes + i];
      if (edges) edges[r] = (int64_t)h_ls[2 * kMaxLanes + i];
    }
    if (stats) {
      stats->batches += 1;
      stats->frontier_edges += (int64_t)ws->h_state->work_fe;
      for (int i = 0; i < lanes; i++) {
        stats->vertices_reached += (int64_t)h_ls[kMaxLanes + i];
        stats->edges_reached += (int64_t)h_ls[2 * kMaxLanes + i];
      }
      const int levels = ws->h_state->levels;
      for (int st = 1; st <= levels && st < kTraceMax; st++) {
        stats->steps += 1;
        const int md = h_modes[st];
        if (md >= 0 && md < DAWN_NUM_MODES) stats->mode_steps[md] += 1;
      }
    }
  }
  // Rows of the isolated sources (timed with the batches), after this
  // stream's batches: into the caller's device rows, or the slot's row
  // buffer when rows are computed without an output (row k mod 64·W
  // there).  Written last, the fill overlaps the other stream's last batch,
  // whose steps mostly read (RMAT-24, 4096 sources: 1.2% faster than
  // writing them first, alternating runs on one box).
  if (iso && iso->count > 0 && (dist_dev || buf_rows)) {
    std::vector<int32_t> pairs(iso->pairs, iso->pairs + 2 * (size_t)iso->count);
    if (!dist_dev)
      for (int32_t k = 0; k < iso->count; k++) pairs[2 * k] = k % kB;
    DAWN_TRY(ws->iso.alloc(pairs.size()));
    DAWN_HIP(hipMemcpyAsync(ws->iso.p, pairs.data(), sizeof(int32_t) * pairs.size(),
                            hipMemcpyHostToDevice, s));
    DAWN_HIP(hipEventRecord(ws->ev0, s));
    const int64_t chunks = (((int64_t)g->n + 3) / 4 + kIsoChunk - 1) / kIsoChunk * iso->count;
    ms_isolated_rows_kernel<<<(unsigned)std::min<int64_t>(chunks, 1 << 20), kBlock, 0, s>>>(
        dist_dev ? dist : ws->distbuf.p, g->n, g->n, ws->iso.p, iso->count);
    DAWN_HIP(hipGetLastError());
    DAWN_HIP(hipEventRecord(ws->ev1, s));
    DAWN_HIP(hipEventSynchronize(ws->ev1));
    float ms = 0.f;
    DAWN_HIP(hipEventElapsedTime(&ms, ws->ev0, ws->ev1));
    total_ms += ms;
  }
  if (stats) stats->device_ms += total_ms;
  return DAWN_OK;
}

// Runs an unweighted multi-source traversal in batches of 64·W sources.
// With two or more batches a second workspace and stream (kept with the
// graph) run every other batch from a second host thread, so that one
// batch's memory-bound distance pass overlaps the other's steps; if the
// second workspace cannot be allocated, one stream runs them all.  Then
// stats->device_ms is the wall time of the whole set of batches.
//
// A partly filled last batch that fits fewer lane words runs on records of
// that width, in a workspace of its own (kept with the graph), on the
// stream whose turn it is: a step gathers whole records whatever the
// number of live lanes in them, so a 120-lane batch at 16 words costs about
// as much as a full one.  If that workspace cannot be allocated, the last
// batch runs at W.
template <int W>
int run_msbfs_w(dawn_graph* g, const int32_t* sources, const int32_t* rowidx, int32_t S,
                const IsoRows* iso, int32_t* dist, int64_t* sums, int64_t* reached,
                int64_t* edges, int32_t policy, dawn_stats* stats, bool rows) {
  constexpr int kB = 64 * W;
  DAWN_TRY(ensure_ms(g, g->ms, W));
  DAWN_TRY(ms_view_ensure(g));
  // Row buffers of batches whose rows do not go to a device buffer.
  const bool dist_dev = dist && is_device_pointer(dist);
  const bool buf_rows = !dist_dev && (dist || rows);
  if (buf_rows) DAWN_TRY(g->ms->distbuf.alloc((size_t)kB * g->n));
  bool two = S > kB;
  if (two) {
    if (!g->stream2 && hipStreamCreateWithFlags(&g->stream2, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      g->stream2 = nullptr;
      two = false;
    }
    if (two && (ensure_ms(g, g->ms2, W) != DAWN_OK ||
                (buf_rows && g->ms2->distbuf.alloc((size_t)kB * g->n) != DAWN_OK))) {
      ms_workspace_free(g->ms2);  // not enough memory for a second batch in flight
      g->ms2 = nullptr;
      (void)hipGetLastError();
      two = false;
    }
  }
  const int32_t full = S / kB * kB;
  const int32_t tail = S - full;
  int Wt = 1;
  while (Wt < W && 64 * Wt < tail) Wt <<= 1;
  bool narrow = full > 0 && tail > 0 && Wt < W;
  if (narrow && (ensure_ms(g, g->ms_tail, Wt) != DAWN_OK ||
                 (buf_rows && g->ms_tail->distbuf.alloc((size_t)64 * Wt * g->n) != DAWN_OK))) {
    ms_workspace_free(g->ms_tail);
    g->ms_tail = nullptr;
    (void)hipGetLastError();
    narrow = false;
  }
  std::vector<int32_t> tail_rows;
  const int32_t* trow = nullptr;
  if (narrow) {
    if (rowidx) {
      trow = rowidx + full;
    } else {
      tail_rows.resize(tail);
      for (int32_t i = 0; i < tail; i++) tail_rows[i] = full + i;
      trow = tail_rows.data();
    }
  }
  const int32_t S_main = narrow ? full : S;
  // The narrow last batch, on stream s (nothing to do unless there is one).
  auto run_tail = [&](hipStream_t s, dawn_stats* st) -> int {
    if (!narrow) return DAWN_OK;
    const int32_t* ts = sources + full;
    switch (Wt) {
      case 1:
        return run_msbfs_slot<1>(g, g->ms_tail, s, ts, trow, tail, 0, 1, nullptr, dist, sums,
                                 reached, edges, policy, st, rows);
      case 2:
        return run_msbfs_slot<2>(g, g->ms_tail, s, ts, trow, tail, 0, 1, nullptr, dist, sums,
                                 reached, edges, policy, st, rows);
      case 4:
        return run_msbfs_slot<4>(g, g->ms_tail, s, ts, trow, tail, 0, 1, nullptr, dist, sums,
                                 reached, edges, policy, st, rows);
      default:
        return run_msbfs_slot<8>(g, g->ms_tail, s, ts, trow, tail, 0, 1, nullptr, dist, sums,
                                 reached, edges, policy, st, rows);
    }
  };
  // The narrow batch (batch full/kB) runs on the stream of its turn.
  const int tail_turn = (int)((full / kB) % 2);
  if (!two) {
    DAWN_TRY(run_msbfs_slot<W>(g, g->ms, g->stream, sources, rowidx, S_main, 0, 1, iso, dist,
                               sums, reached, edges, policy, stats, rows));
    return run_tail(g->stream, stats);
  }
  dawn_stats st1{};
  int rc1 = DAWN_OK;
  const auto t0 = std::chrono::steady_clock::now();
  std::thread second([&] {
    if (hipSetDevice(g->device) != hipSuccess) {
      rc1 = DAWN_ERR_HIP;
      return;
    }
    // The second stream runs the fewer batches: it also writes the rows of
    // the isolated sources.
    rc1 = run_msbfs_slot<W>(g, g->ms2, g->stream2, sources, rowidx, S_main, 1, 2, iso, dist, sums,
                            reached, edges, policy, stats ? &st1 : nullptr, rows);
    if (rc1 == DAWN_OK && tail_turn == 1) rc1 = run_tail(g->stream2, stats ? &st1 : nullptr);
  });
  dawn_stats st0{};
  int rc0 = run_msbfs_slot<W>(g, g->ms, g->stream, sources, rowidx, S_main, 0, 2, nullptr, dist,
                              sums, reached, edges, policy, stats ? &st0 : nullptr, rows);
  if (rc0 == DAWN_OK && tail_turn == 0) rc0 = run_tail(g->stream, stats ? &st0 : nullptr);
  second.join();
  const double wall_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc0 != DAWN_OK) return rc0;
  if (rc1 != DAWN_OK) return rc1;
  if (stats) {
    stats->device_ms += wall_ms;
    stats->steps += st0.steps + st1.steps;
    for (int i = 0; i < DAWN_NUM_MODES; i++) stats->mode_steps[i] += st0.mode_steps[i] + st1.mode_steps[i];
    stats->edges_reached += st0.edges_reached + st1.edges_reached;
    stats->vertices_reached += st0.vertices_reached + st1.vertices_reached;
    stats->batches += st0.batches + st1.batches;
    stats->skipped_sources += st0.skipped_sources + st1.skipped_sources;
    stats->frontier_edges += st0.frontier_edges + st1.frontier_edges;
  }
  return DAWN_OK;
}

// Per-step mode timings from identical state (scheduler training data).
// Each step: decide (features), snapshot, then every mode is run `repeats`
// times from the snapshot and timed with HIP events; the minimum is kept.
// The batches of `sources` at W lane words run in workspace `slot`; rows
// are written from index 0 of the given arrays.
template <int W>
int run_ms_record_w(dawn_graph* g, MsWorkspace*& slot, const int32_t* sources, int32_t S,
                    int32_t repeats, float* features, float* times_ms, int32_t* chosen,
                    int32_t max_rows, int32_t* rows_out, int32_t policy) {
  constexpr int kB = 64 * W;
  DAWN_TRY(ensure_ms(g, slot, W));
  DAWN_TRY(ms_view_ensure(g));
  MsWorkspace* ws = slot;
  const size_t words = (size_t)ws->npad * W;
  DAWN_TRY(ws->seen_bak.alloc(words));
  DAWN_TRY(ws->done_bak.alloc((size_t)(ws->npad / 32)));
  const size_t done_bytes = sizeof(uint32_t) * (size_t)(ws->npad / 32);
  hipStream_t s = g->stream;
  MsArgs a = make_args(g, ws);
  SchedParams p = make_params(g, policy);
  SchedTrace tr{ws->trace_feat.p, ws->trace_mode.p};
  const unsigned grid = ms_grid(g);
  const unsigned fin_grid = ms_fin_grid(g, W);
  unsigned long long* ls = ws->lane_stats.p;
  MsOut o{ls, ls + kMaxLanes, ls + 2 * kMaxLanes, 0};
  int32_t rows = 0;
  for (int32_t b0 = 0; b0 < S && rows < max_rows; b0 += kB) {
    const int32_t lanes = std::min<int32_t>(kB, S - b0);
    stage_batch(ws, sources + b0, lanes, g->view);
    DAWN_HIP(hipMemcpyAsync(ws->batch.p, ws->h_batch, sizeof(MsBatch), hipMemcpyHostToDevice, s));
    ms_fill_kernel<W><<<grid_for(ws->npad * W), kBlock, 0, s>>>(a, o);
    ms_seed_kernel<W><<<1, kMaxLanes, 0, s>>>(a);
    for (int step = 1; rows < max_rows && step < a.H; step++) {
      ms_decide_kernel<<<1, 64, 0, s>>>(a, p, tr);
      DAWN_HIP(hipMemcpyAsync(ws->h_state, ws->st.p, sizeof(SchedState), hipMemcpyDeviceToHost, s));
      DAWN_HIP(hipStreamSynchronize(s));
      if (ws->h_state->done) break;
      const SchedState snap = *ws->h_state;
      if (chosen) chosen[rows] = snap.mode;
      DAWN_HIP(hipMemcpyAsync(features + (size_t)rows * 8, ws->trace_feat.p + (size_t)snap.step * 8,
                              sizeof(float) * 8, hipMemcpyDeviceToHost, s));
      DAWN_HIP(hipMemcpyAsync(ws->seen_bak.p, ws->seen.p, sizeof(uint64_t) * words,
                              hipMemcpyDeviceToDevice, s));
      DAWN_HIP(hipMemcpyAsync(ws->ring_bak.p, ws->ring.p, sizeof(LevelCounters) * kRing,
                              hipMemcpyDeviceToDevice, s));
      DAWN_HIP(hipMemcpyAsync(ws->part_bak.p, ws->part.p, sizeof(unsigned long long) * kMsPartU64,
                              hipMemcpyDeviceToDevice, s));
      DAWN_HIP(hipMemcpyAsync(ws->done_bak.p, ws->done.p, done_bytes, hipMemcpyDeviceToDevice, s));
      for (int mode = 0; mode < DAWN_NUM_MODES; mode++) {
        float best = 1e30f;
        for (int r = 0; r < repeats; r++) {
          SchedState st2 = snap;
          st2.mode = mode;
          *ws->h_state = st2;
          DAWN_HIP(hipMemcpyAsync(ws->st.p, ws->h_state, sizeof(SchedState), hipMemcpyHostToDevice, s));
          DAWN_HIP(hipMemcpyAsync(ws->seen.p, ws->seen_bak.p, sizeof(uint64_t) * words,
                                  hipMemcpyDeviceToDevice, s));
          DAWN_HIP(hipMemcpyAsync(ws->ring.p, ws->ring_bak.p, sizeof(LevelCounters) * kRing,
                                  hipMemcpyDeviceToDevice, s));
          DAWN_HIP(hipMemcpyAsync(ws->part.p, ws->part_bak.p,
                                  sizeof(unsigned long long) * kMsPartU64, hipMemcpyDeviceToDevice, s));
          DAWN_HIP(hipMemcpyAsync(ws->done.p, ws->done_bak.p, done_bytes, hipMemcpyDeviceToDevice, s));
          DAWN_HIP(hipEventRecord(ws->e_a, s));
          launch_step<W>(a, grid, fin_grid, s);
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

// Same kernels as run_msbfs, launched one by one with HIP events around each
// launch on the graph's stream, writing distance rows and statistics like a
// timed traversal: the batches of `sources` at W lane words, in workspace
// `slot`.  Adds to kernel_ms / launches, indexed by DAWN_PROF_* (fill, seed,
// decide, expand push, expand pull, expand bitmap, finalize, final);
// levels_out[i]: the number of levels of batch i.
template <int W>
int run_ms_profile_w(dawn_graph* g, MsWorkspace*& slot, const int32_t* sources, int32_t S,
                     int32_t policy, double* kernel_ms, int64_t* launches, int32_t* levels_out) {
  constexpr int kB = 64 * W;
  DAWN_TRY(ensure_ms(g, slot, W));
  DAWN_TRY(ms_view_ensure(g));
  MsWorkspace* ws = slot;
  DAWN_TRY(ws->distbuf.alloc((size_t)kB * g->n));
  hipStream_t s = g->stream;
  MsArgs a = make_args(g, ws);
  SchedParams p = make_params(g, policy);
  SchedTrace tr{ws->trace_feat.p, ws->trace_mode.p};
  const unsigned grid = ms_grid(g);
  const unsigned fin_grid = ms_fin_grid(g, W);
  const unsigned fgrid = ms_final_grid(g, W);
  unsigned long long* ls = ws->lane_stats.p;
  MsOut o{ls, ls + kMaxLanes, ls + 2 * kMaxLanes, 1};
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
  for (int32_t b0 = 0; b0 < S; b0 += kB, batch++) {
    const int32_t lanes = std::min<int32_t>(kB, S - b0);
    stage_batch(ws, sources + b0, lanes, g->view);
    ws->h_batch->dist = ws->distbuf.p;
    ws->h_batch->stride = g->n;
    DAWN_HIP(hipMemcpyAsync(ws->batch.p, ws->h_batch, sizeof(MsBatch), hipMemcpyHostToDevice, s));
    DAWN_TRY(timed(0, [&] { ms_fill_kernel<W><<<grid_for(ws->npad * W), kBlock, 0, s>>>(a, o); }));
    DAWN_TRY(timed(1, [&] { ms_seed_kernel<W><<<1, kMaxLanes, 0, s>>>(a); }));
    int lo = 0;
    for (int64_t guard_steps = 0;; guard_steps++) {
      if (guard_steps > (int64_t)g->n + 2) return DAWN_ERR_HIP;
      DAWN_TRY(timed(2, [&] { ms_decide_kernel<<<1, 64, 0, s>>>(a, p, tr); }));
      DAWN_HIP(hipMemcpy(ws->h_state, ws->st.p, sizeof(SchedState), hipMemcpyDeviceToHost));
      const SchedState st = *ws->h_state;
      if (st.done) {
        DAWN_TRY(timed(7, [&] { ms_final_kernel<W><<<fgrid, kBlock, 0, s>>>(a, o, 0, lo, st.levels); }));
        if (levels_out) levels_out[batch] = st.levels;
        break;
      }
      const int slot = st.mode == DAWN_MODE_PUSH ? 3 : st.mode == DAWN_MODE_PULL ? 4 : 5;
      DAWN_TRY(timed(slot, [&] {
        if (st.mode == DAWN_MODE_PUSH) {
          ms_compact_kernel<<<compact_grid(grid), kBlock, 0, s>>>(a);
          ms_push_kernel<W><<<grid, kBlock, 0, s>>>(a);
        } else if (st.mode == DAWN_MODE_PULL) {
          launch_pull<W>(a, grid, s);
        } else {
          ms_topdown_kernel<W><<<grid, kBlock, 0, s>>>(a);
        }
      }));
      DAWN_TRY(timed(6, [&] {
        ms_finalize_kernel<W><<<fin_grid, kBlock, 0, s>>>(a);
        ms_fold_heavy_kernel<W><<<fin_grid, kBlock, 0, s>>>(a);
      }));
      if (st.step - lo >= a.H - 2) {  // flush before the history ring wraps
        DAWN_TRY(timed(7, [&] { ms_final_kernel<W><<<fgrid, kBlock, 0, s>>>(a, o, 0, lo, st.step); }));
        lo = st.step;
      }
    }
  }
  DAWN_HIP(hipStreamSynchronize(s));
  return DAWN_OK;
}

}  // namespace

// Splits a source list into the sources that traverse (out-degree > 0)
// with their positions in the list, and (position, vertex) pairs of the
// isolated ones.  The traversing sources are ordered by out-degree,
// largest first (stable), so that consecutive batches hold sources of
// similar degree: in a small-world graph a high-degree source reaches most
// vertices a level or more before a leaf does, so a batch of like sources
// has its lanes arrive at a vertex over fewer levels, and the vertex is in
// fewer frontiers (fewer pull gathers).  Lanes are independent: the order
// changes time only; rows and sums go to each source's own position.
void split_live(const dawn_graph* g, const int32_t* sources, int32_t S, std::vector<int32_t>& live,
                std::vector<int32_t>& live_row, std::vector<int32_t>& iso) {
  live.clear();
  live_row.clear();
  iso.clear();
  std::vector<std::pair<int32_t, int32_t>> by_deg;  // (-degree, position)
  for (int32_t i = 0; i < S; i++) {
    int64_t s = sources[i] % (int64_t)g->n;
    if (s < 0) s += g->n;
    const int32_t d = g->h_row_ptr[s + 1] - g->h_row_ptr[s];
    if (d == 0) {
      iso.push_back(i);
      iso.push_back((int32_t)s);
    } else {
      by_deg.push_back({-d, i});
    }
  }
  std::stable_sort(by_deg.begin(), by_deg.end(),
                   [](const std::pair<int32_t, int32_t>& x, const std::pair<int32_t, int32_t>& y) {
                     return x.first < y.first;
                   });
  for (const auto& e : by_deg) {
    live.push_back(sources[e.second]);
    live_row.push_back(e.second);
  }
}

int run_msbfs(dawn_graph* g, const int32_t* sources, int32_t S, int32_t* dist, int64_t* sums,
              int64_t* reached, int64_t* edges, int32_t policy, dawn_stats* stats, bool rows) {
  if (!g || S < 0 || (S > 0 && !sources) || g->n <= 0) return DAWN_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  // Isolated sources take no lane (as the reference skips them,
  // src/algorithm/cpu/mssp.cpp:16-18): sum 0, reached 1 (the source), no
  // edges; their rows are written by the engine when rows are asked for.
  std::vector<int32_t> live, live_row, iso;
  split_live(g, sources, S, live, live_row, iso);
  const int32_t L = (int32_t)live.size();
  const IsoRows isr{iso.data(), (int32_t)(iso.size() / 2)};
  const bool dist_dev = dist && is_device_pointer(dist);
  // Workspaces are sized by the batch width; if the device cannot hold them
  // (they are allocated before any batch runs), narrower batches are tried.
  for (int w = words_for(g, L);; w >>= 1) {
    int rc;
    const int32_t* ls = live.data();
    const int32_t* lr = live_row.data();
    switch (w) {
      case 1: rc = run_msbfs_w<1>(g, ls, lr, L, &isr, dist, sums, reached, edges, policy, stats, rows); break;
      case 2: rc = run_msbfs_w<2>(g, ls, lr, L, &isr, dist, sums, reached, edges, policy, stats, rows); break;
      case 4: rc = run_msbfs_w<4>(g, ls, lr, L, &isr, dist, sums, reached, edges, policy, stats, rows); break;
      case 8: rc = run_msbfs_w<8>(g, ls, lr, L, &isr, dist, sums, reached, edges, policy, stats, rows); break;
      default: rc = run_msbfs_w<16>(g, ls, lr, L, &isr, dist, sums, reached, edges, policy, stats, rows); break;
    }
    if (rc == DAWN_ERR_OUT_OF_MEMORY && w > 1) {
      ms_workspace_free(g->ms);
      g->ms = nullptr;
      ms_workspace_free(g->ms2);
      g->ms2 = nullptr;
      ms_workspace_free(g->ms_tail);
      g->ms_tail = nullptr;
      (void)hipGetLastError();
      continue;
    }
    if (rc != DAWN_OK) return rc;
    break;
  }
  for (int32_t k = 0; k < isr.count; k++) {
    const int32_t i = iso[2 * k], s = iso[2 * k + 1];
    if (sums) sums[i] = 0;
    if (reached) reached[i] = 1;
    if (edges) edges[i] = 0;
    if (dist && !dist_dev) {
      int32_t* row = dist + (int64_t)i * g->n;
      std::fill(row, row + g->n, (int32_t)kInf);
      row[s] = 0;
    }
  }
  if (stats) {
    stats->skipped_sources += isr.count;
    stats->vertices_reached += isr.count;
  }
  return DAWN_OK;
}

int run_ms_record(dawn_graph* g, const int32_t* sources, int32_t S, int32_t repeats,
                  float* features, float* times_ms, int32_t* chosen, int32_t max_rows,
                  int32_t* rows_out, int32_t policy) {
  if (!g || S <= 0 || !sources || repeats < 1) return DAWN_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  // The batches of a traversal: isolated sources take no lane.
  std::vector<int32_t> live, live_row, iso;
  split_live(g, sources, S, live, live_row, iso);
  sources = live.data();
  S = (int32_t)live.size();
  if (S == 0) {
    if (rows_out) *rows_out = 0;
    return DAWN_OK;
  }
  // The batches as the traversal runs them (full ones, then a partly
  // filled last one at its own width); the last batch's rows follow.
  const BatchSplit bs = split_batches(g, S);
  auto run = [&](int w, MsWorkspace*& slot, const int32_t* src, int32_t cnt, int32_t off,
                 int32_t* got) -> int {
    float* f = features + (size_t)off * 8;
    float* t = times_ms + (size_t)off * DAWN_NUM_MODES;
    int32_t* c = chosen ? chosen + off : nullptr;
    const int32_t mr = max_rows - off;
    switch (w) {
      case 1: return run_ms_record_w<1>(g, slot, src, cnt, repeats, f, t, c, mr, got, policy);
      case 2: return run_ms_record_w<2>(g, slot, src, cnt, repeats, f, t, c, mr, got, policy);
      case 4: return run_ms_record_w<4>(g, slot, src, cnt, repeats, f, t, c, mr, got, policy);
      case 8: return run_ms_record_w<8>(g, slot, src, cnt, repeats, f, t, c, mr, got, policy);
      default: return run_ms_record_w<16>(g, slot, src, cnt, repeats, f, t, c, mr, got, policy);
    }
  };
  int32_t r0 = 0, r1 = 0;
  DAWN_TRY(run(bs.W, g->ms, sources, bs.S_main, 0, &r0));
  if (bs.S_main < S && r0 < max_rows)
    DAWN_TRY(run(bs.Wt, g->ms_tail, sources + bs.S_main, S - bs.S_main, r0, &r1));
  if (rows_out) *rows_out = r0 + r1;
  return DAWN_OK;
}

int run_ms_profile(dawn_graph* g, const int32_t* sources, int32_t S, int32_t policy,
                   double* kernel_ms, int64_t* launches, int32_t* levels_out) {
  if (!g || S <= 0 || !sources || !kernel_ms || !launches) return DAWN_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> guard(g->lock);
  DAWN_HIP(hipSetDevice(g->device));
  // The batches of a traversal: isolated sources take no lane (their rows
  // are a plain fill, not profiled here).
  std::vector<int32_t> live, live_row, iso;
  split_live(g, sources, S, live, live_row, iso);
  const int32_t S_all = S;
  sources = live.data();
  S = (int32_t)live.size();
  if (levels_out)
    for (int32_t i = 0; i < (S_all + 63) / 64; i++) levels_out[i] = -1;
  for (int i = 0; i < 8; i++) kernel_ms[i] = 0, launches[i] = 0;
  if (S == 0) return DAWN_OK;
  // The batches as the traversal runs them: full ones at W lane words, a
  // partly filled last one at the fewest words holding its lanes.
  const BatchSplit bs = split_batches(g, S);
  auto run = [&](int w, MsWorkspace*& slot, const int32_t* src, int32_t cnt, int32_t* lv) -> int {
    switch (w) {
      case 1: return run_ms_profile_w<1>(g, slot, src, cnt, policy, kernel_ms, launches, lv);
      case 2: return run_ms_profile_w<2>(g, slot, src, cnt, policy, kernel_ms, launches, lv);
      case 4: return run_ms_profile_w<4>(g, slot, src, cnt, policy, kernel_ms, launches, lv);
      case 8: return run_ms_profile_w<8>(g, slot, src, cnt, policy, kernel_ms, launches, lv);
      default: return run_ms_profile_w<16>(g, slot, src, cnt, policy, kernel_ms, launches, lv);
    }
  };
  DAWN_TRY(run(bs.W, g->ms, sources, bs.S_main, levels_out));
  if (bs.S_main == S) return DAWN_OK;
  return run(bs.Wt, g->ms_tail, sources + bs.S_main, S - bs.S_main,
             levels_out ? levels_out + bs.S_main / (64 * bs.W) : nullptr);
}

}  // namespace dawn
