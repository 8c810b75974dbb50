// Shared declarations of the HIP frontier engine: device graph, status
// handling, wave64 helpers.  gfx950 only (wave64, 160 KB LDS per CU).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/dawn_capi.h"

#if defined(__HIP_DEVICE_COMPILE__) && !defined(__gfx950__)
#error "device code is written for gfx950 (CDNA4, wave64)"
#endif

namespace dawn {

constexpr int kWave = 64;
constexpr int kBlock = 256;          // threads per workgroup for step kernels
constexpr int kHeavyDegree = 32;     // rows >= this are expanded by a whole wave
constexpr int kPullChunk = 256;      // in-edges per heavy pull chunk
constexpr int kPushChunk = 256;      // out-edges per push queue entry
constexpr int kXcds = 8;             // MI355X: workgroups go round-robin to 8 XCDs (own L2 each)
constexpr int32_t kInf = DAWN_INF;

// Records the last HIP error text (per thread) for diagnostics.
void set_hip_error(hipError_t e, const char* what, const char* file, int line);
const char* last_error();
// Records a failure text that is not a HIP error (per thread).
void set_error(const char* text);

#define DAWN_HIP(x)                                                  \
  do {                                                               \
    hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) {                                          \
      ::dawn::set_hip_error(e_, #x, __FILE__, __LINE__);             \
      return e_ == hipErrorOutOfMemory ? DAWN_ERR_OUT_OF_MEMORY      \
                                       : DAWN_ERR_HIP;               \
    }                                                                \
  } while (0)

#define DAWN_TRY(x)           \
  do {                        \
    int rc_ = (x);            \
    if (rc_ != DAWN_OK) return rc_; \
  } while (0)

// Device allocation that frees itself.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int alloc(size_t n) {
    if (n <= count && p) return DAWN_OK;
    release();
    size_t bytes = sizeof(T) * (n ? n : 1);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      set_hip_error(e, "hipMalloc", __FILE__, __LINE__);
      return e == hipErrorOutOfMemory ? DAWN_ERR_OUT_OF_MEMORY : DAWN_ERR_HIP;
    }
    count = n;
    return DAWN_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    count = 0;
  }
};

// Launch knobs chosen by the autotuner (tune.hip) per graph class.
struct TuneKnobs {
  int32_t dense_blocks = 2048;  // grid cap of the batched dense step kernels
  int32_t heavy_degree = 32;    // in/out-degree from which edge lists are chunked
  int32_t pull_chunk = 256;     // edges per heavy chunk (one wave each)
  int32_t ss_chunk = 0;         // single-source steps per captured graph (0: auto)
  int32_t pull_variant = 0;     // light pull: 0 per lane 4 in flight, 1 edge-balanced,
                                // 2 per lane 8 in flight, 3 as 2 with streaming column loads
  int32_t lane_words = 16;      // batched engine: 64-lane words per vertex (1, 2, 4, 8, 16)
  int32_t hub_order = 1;        // batched engine: hub-first vertex order (MsView)
};

// Graphs below this size keep their own order in the batched engine.
constexpr int32_t kViewMinVertices = 1 << 16;

struct MsWorkspace;   // 64-lane batched engine state (msbfs.hip)
struct SsWorkspace;   // single-source engines state (bfs.hip / sssp.hip)
struct WmsWorkspace;  // batched weighted engine state (wmsbfs.hip)
struct MsView;        // hub-first order of the batched engine (view.hip)

}  // namespace dawn

// Device-resident graph.  For undirected graphs the in-edge arrays alias the
// out-edge arrays (the adjacency is symmetric).
struct dawn_graph {
  int32_t n = 0;
  int64_t m = 0;
  int32_t directed = 0;
  int32_t weighted = 0;
  int device = 0;
  int32_t max_out_degree = 0;
  int32_t max_in_degree = 0;
  int64_t zero_in = 0;            // vertices of in-degree 0 (no step ever reaches them)
  int32_t* row_ptr = nullptr;     // [n+1]
  int32_t* col = nullptr;         // [m]
  int32_t* weight = nullptr;      // [m] or null
  int32_t* in_row_ptr = nullptr;  // CSC (directed only; else == row_ptr)
  int32_t* in_col = nullptr;
  int32_t* in_weight = nullptr;
  // In-edge lists of vertices with in-degree >= kHeavyDegree, cut into
  // chunks of at most kPullChunk edges: chunk c covers in_col[beg[c], end[c])
  // of vertex v[c].  One wave per chunk in the pull step.
  int32_t* hc_v = nullptr;
  int32_t* hc_beg = nullptr;
  int32_t* hc_end = nullptr;
  int64_t num_hc = 0;
  // Same for out-edge lists (dense top-down step).  Alias the hc_* arrays
  // for undirected graphs.
  int32_t* oc_v = nullptr;
  int32_t* oc_beg = nullptr;
  int32_t* oc_end = nullptr;
  int64_t num_oc = 0;
  std::vector<int32_t> h_row_ptr; // host copy of offsets (degrees for drivers)
  std::vector<int32_t> h_in_row_ptr;  // host CSC offsets (directed graphs)
  dawn::TuneKnobs knobs;
  hipStream_t stream = nullptr;
  std::mutex lock;                // one traversal per graph at a time
  dawn::MsWorkspace* ms = nullptr;
  dawn::MsWorkspace* ms2 = nullptr;   // second batch in flight (stream2)
  dawn::MsWorkspace* ms_tail = nullptr;  // a partly filled last batch, fewer lane words
  hipStream_t stream2 = nullptr;
  dawn::SsWorkspace* ss = nullptr;
  dawn::WmsWorkspace* wms = nullptr;
  dawn::MsView* view = nullptr;
};

namespace dawn {
// Hub-first renumbering of a graph for the batched unweighted engine: the
// n/16 vertices of largest out-degree first, the others in their order.
// Device arrays are in the new numbering; perm maps old ids to new ones.
struct MsView {
  int32_t k = 0;
  std::vector<int32_t> h_perm, h_row_ptr, h_in_row_ptr;
  int32_t* row_ptr = nullptr;
  int32_t* col = nullptr;
  int32_t* in_row_ptr = nullptr;  // == row_ptr / col for undirected graphs
  int32_t* in_col = nullptr;
  int32_t* perm = nullptr;
  int32_t *hc_v = nullptr, *hc_beg = nullptr, *hc_end = nullptr;
  int64_t num_hc = 0;
  int32_t *oc_v = nullptr, *oc_beg = nullptr, *oc_end = nullptr;
  int64_t num_oc = 0;
};
int ms_view_ensure(dawn_graph* g);
void ms_view_free(dawn_graph* g);
}  // namespace dawn

namespace dawn {

// ---- wave64 device helpers -------------------------------------------------

// Min-plus candidate du + w, summed in 64 bits: kInf when u is unreached or
// the path weight reaches INT32_MAX (such a path is not representable in
// int32 and is ignored, as the CPU restatement skips sums >= INT32_MAX).
__device__ __forceinline__ int32_t minplus(int32_t du, int32_t w) {
  const long long nd = (long long)du + (long long)w;
  return nd >= (long long)kInf ? kInf : (int32_t)nd;
}

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t lanes_below_mask() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Orders the LDS accesses of one wave: what its lanes wrote before is seen
// by its lanes after (LDS executes a wave's operations in order; this keeps
// the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-aggregated append: lanes with pred get consecutive slots of *counter.
// Must be reached by every lane that may append (ballot over active lanes).
__device__ __forceinline__ int wave_append(unsigned int* counter, bool pred) {
  uint64_t m = __ballot(pred);
  if (m == 0) return -1;
  int leader = __ffsll((unsigned long long)m) - 1;
  unsigned int base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (unsigned int)__popcll(m));
  base = __shfl(base, leader);
  return pred ? (int)(base + (unsigned int)__popcll(m & lanes_below_mask())) : -1;
}

// Wave-aggregated append of k >= 0 slots per lane; returns the lane's first
// slot.  Must be called by all 64 lanes.
__device__ __forceinline__ int wave_append_n(unsigned int* counter, int k) {
  const int l = lane_id();
  int inc = k;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(inc, o);
    if (l >= o) inc += y;
  }
  const int total = __shfl(inc, 63);
  unsigned int base = 0;
  if (total > 0) {
    if (l == 0) base = atomicAdd(counter, (unsigned int)total);
    base = __shfl(base, 0);
  }
  return (int)base + inc - k;
}

// Block-aggregated append of k >= 0 slots per thread: one atomic per block.
// Returns the thread's first slot.  Every thread of the kBlock-thread block
// must call it (it synchronises the block).  s_wtot: kBlock/kWave ints of
// LDS, s_base: one unsigned of LDS.
__device__ __forceinline__ int block_append_n(unsigned int* counter, int k, int* s_wtot,
                                              unsigned int* s_base) {
  const int l = lane_id();
  const int w = threadIdx.x / kWave;
  int inc = k;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(inc, o);
    if (l >= o) inc += y;
  }
  if (l == kWave - 1) s_wtot[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int i = 0; i < kBlock / kWave; i++) t += s_wtot[i];
    *s_base = t ? atomicAdd(counter, (unsigned int)t) : 0u;
  }
  __syncthreads();
  int off = 0;
  for (int i = 0; i < w; i++) off += s_wtot[i];
  const int r = (int)*s_base + off + inc - k;
  __syncthreads();  // s_wtot / s_base may be reused after this
  return r;
}

// Queue appends of a workgroup staged in LDS.  A global counter that every
// wave of the grid bumps serialises at one address: thousands of such
// atomics per step cost more than the step's own memory traffic.  Waves
// reserve LDS slots with LDS atomics instead, and the workgroup reserves
// its global queue range once, in flush().  Appends beyond the LDS
// capacity go straight to the global queue (wave-aggregated), so nothing
// is lost on overflow.
template <int Cap>
struct LdsQueue {
  int32_t v[Cap];
  unsigned n;     // slots reserved (may pass Cap)
  unsigned base;  // global position of v[0], set in flush()

  // One thread; the caller synchronises the workgroup before any append.
  __device__ void init() { n = 0u; }

  // Called by all 64 lanes of a wave.
  __device__ void append(bool pred, int32_t x, unsigned* qlen, int32_t* q_out) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    const unsigned cnt = (unsigned)__popcll(m);
    unsigned b = 0;
    if (lane_id() == leader) b = atomicAdd(&n, cnt);
    b = __shfl(b, leader);
    const unsigned idx = b + (unsigned)__popcll(m & lanes_below_mask());
    if (b + cnt <= (unsigned)Cap) {
      if (pred) v[idx] = x;
      return;
    }
    // LDS full: slots [b, Cap) are still filled here (flush copies them),
    // the rest go to the global queue.
    const unsigned first_global = b > (unsigned)Cap ? b : (unsigned)Cap;
    const unsigned ng = b + cnt - first_global;
    unsigned g = 0;
    if (lane_id() == leader) g = atomicAdd(qlen, ng);
    g = __shfl(g, leader);
    if (pred) {
      if (idx < (unsigned)Cap) v[idx] = x;
      else q_out[g + (idx - first_global)] = x;
    }
  }

  // Every thread of the workgroup, once, after its last append.
  __device__ void flush(unsigned* qlen, int32_t* q_out) {
    __syncthreads();
    const unsigned cnt = n < (unsigned)Cap ? n : (unsigned)Cap;
    if (threadIdx.x == 0) base = cnt ? atomicAdd(qlen, cnt) : 0u;
    __syncthreads();
    for (unsigned i = threadIdx.x; i < cnt; i += blockDim.x) q_out[base + i] = v[i];
  }
};

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
  for (int o = 32; o >= 1; o >>= 1) {
    unsigned lo = __shfl_xor((unsigned)(x & 0xffffffffu), o);
    unsigned hi = __shfl_xor((unsigned)(x >> 32), o);
    x += ((unsigned long long)hi << 32) | lo;
  }
  return x;
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t x) {
  for (int o = 32; o >= 1; o >>= 1) {
    unsigned lo = __shfl_xor((unsigned)(x & 0xffffffffu), o);
    unsigned hi = __shfl_xor((unsigned)(x >> 32), o);
    x |= ((uint64_t)hi << 32) | lo;
  }
  return x;
}

__device__ __forceinline__ int wave_max_i32(int x) {
  for (int o = 32; o >= 1; o >>= 1) x = max(x, __shfl_xor(x, o));
  return x;
}

__device__ __forceinline__ int wave_min_i32(int x) {
  for (int o = 32; o >= 1; o >>= 1) x = min(x, __shfl_xor(x, o));
  return x;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int o) {
  unsigned lo = __shfl_xor((unsigned)(x & 0xffffffffu), o);
  unsigned hi = __shfl_xor((unsigned)(x >> 32), o);
  return ((uint64_t)hi << 32) | lo;
}

// Value of lane (this lane ^ J), J a power of two below 64: DPP quad
// permutes for J = 1, 2, ds_swizzle (bit mode, within 32 lanes) for
// J = 4..16, ds_bpermute across the halves for J = 32.  All lanes active.
template <int J>
__device__ __forceinline__ uint32_t xor_lane_u32(uint32_t x) {
  static_assert(J >= 1 && J <= 32 && (J & (J - 1)) == 0, "power of two");
  if constexpr (J == 1) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
  } else if constexpr (J == 2) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
  } else if constexpr (J < 32) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (J << 10));
  } else {
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((__lane_id() ^ 32u) << 2), (int)x);
  }
}

template <int J>
__device__ __forceinline__ uint64_t xor_lane_u64(uint64_t x) {
  const uint32_t lo = xor_lane_u32<J>((uint32_t)x);
  const uint32_t hi = xor_lane_u32<J>((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// OR (or sum) over the lanes l ^ m, m < N (N a power of two, at most 64):
// the lanes of an aligned group of N.  All lanes active.
template <int N>
__device__ __forceinline__ uint32_t group_or_u32(uint32_t x) {
  if constexpr (N >= 2) return group_or_u32<N / 2>(x | xor_lane_u32<N / 2>(x));
  else return x;
}
template <int N>
__device__ __forceinline__ uint64_t group_or_u64(uint64_t x) {
  if constexpr (N >= 2) return group_or_u64<N / 2>(x | xor_lane_u64<N / 2>(x));
  else return x;
}
template <int N>
__device__ __forceinline__ uint32_t group_sum_u32(uint32_t x) {
  if constexpr (N >= 2) return group_sum_u32<N / 2>(x + xor_lane_u32<N / 2>(x));
  else return x;
}

// Maximum over the wave through xor_lane exchanges (all lanes active).
template <int J = 32>
__device__ __forceinline__ int wave_max_xor(int x) {
  if constexpr (J >= 1) return wave_max_xor<J / 2>(max(x, (int)xor_lane_u32<J>((uint32_t)x)));
  else return x;
}

// One round of the 64x64 bit transpose: swaps bit J of the row index (the
// lane) with bit J of the column index.  M: the low J bits of every 2J-bit
// block.  Branch-free: the lane's own half is kept, the other half comes
// from lane ^ J shifted into place.
template <int J, uint64_t M>
__device__ __forceinline__ uint64_t transpose_round(uint64_t x, int lane) {
  const uint64_t o = xor_lane_u64<J>(x);
  const bool upper = (lane & J) != 0;
  const uint64_t y = upper ? (o >> J) : (o << J);
  const uint64_t keep = upper ? ~M : M;
  return (x & keep) | (y & ~keep);
}

// Transpose the 64x64 bit matrix held one row per lane: afterwards lane j
// holds bit i = bit j of lane i's input.  Six butterfly rounds.
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t x) {
  const int l = lane_id();
  x = transpose_round<32, 0x00000000FFFFFFFFull>(x, l);
  x = transpose_round<16, 0x0000FFFF0000FFFFull>(x, l);
  x = transpose_round<8, 0x00FF00FF00FF00FFull>(x, l);
  x = transpose_round<4, 0x0F0F0F0F0F0F0F0Full>(x, l);
  x = transpose_round<2, 0x3333333333333333ull>(x, l);
  x = transpose_round<1, 0x5555555555555555ull>(x, l);
  return x;
}

// In-register 32x32 bit transpose: afterwards bit r of A[c] is bit c of the
// input A[r] (recursive block swaps, 5 rounds of 16 masked exchanges).
__host__ __device__ __forceinline__ void transpose32(uint32_t (&A)[32]) {
  uint32_t m = 0x0000FFFFu;
#pragma unroll
  for (int j = 16; j != 0; j >>= 1, m ^= (m << j)) {
#pragma unroll
    for (int k = 0; k < 32; k = ((k | j) + 1) & ~j) {
      const uint32_t t = ((A[k] >> j) ^ A[k | j]) & m;
      A[k] ^= t << j;
      A[k | j] ^= t;
    }
  }
}

// Exclusive scan of one int per thread over a kBlock-thread workgroup.
// scratch: kBlock/kWave ints of LDS.  Returns the exclusive prefix; *total
// receives the block sum.
__device__ __forceinline__ int block_exclusive_scan(int x, int* scratch, int* total) {
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
  for (int i = 0; i < kBlock / kWave; i++) {
    int s = scratch[i];
    if (i < w) base += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return base + inc - x;
}

// Grid for grid-stride kernels: enough workgroups to fill 256 CUs several
// times over, capped by the work available.
inline unsigned grid_for(int64_t items, int per_block = kBlock, unsigned cap = 8192) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > (int64_t)cap) b = cap;
  return (unsigned)b;
}

// Pointer kind of a caller buffer.
bool is_device_pointer(const void* p);

// (Re)builds the heavy chunk lists of g from g->knobs (graph_dev.hip).
int build_chunks(dawn_graph* g);
// Drops captured step graphs (their launch shapes depend on the knobs).
void invalidate_graphs(dawn_graph* g);
// Knobs cached for the graph's class, if any (tune.hip).
bool tuned_knobs_for(const dawn_graph* g, TuneKnobs* out);

// Workspace release hooks (defined by the engines).
void ms_workspace_free(MsWorkspace* ws);
void ss_workspace_free(SsWorkspace* ws);
void wms_workspace_free(WmsWorkspace* ws);

}  // namespace dawn
