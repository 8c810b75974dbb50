// Device-side graph construction: edge list -> CSR (+ CSC for directed
// graphs) and the seeded generators, on the GPU, with the same result as the
// host build (host_graph.cpp) bit for bit: self-loops dropped, optional
// symmetrization, rows sorted by (column, weight), optional dedup keeping the
// minimum weight.
//
// Reference: DAWN::Matrix::coo2Csr / coo2Csr_Weighted (src/matrix.cpp:28-82:
// row counts, prefix sum, copy of the row-sorted COO), DAWN::Matrix::
// transpose / transpose_Weighted (matrix.cpp:3-26: a sort of the COO by
// (row, column[, value])), DAWN::IO::readGraph* (src/io.cpp:2-192: 0-based
// entries, self-loops dropped, both directions for undirected graphs),
// DAWN::Graph::createGraph (src/graph.cpp:9-77).
//
// Pipeline, on the graph's stream:
//   emit     one 64-bit key per directed entry, u << vb | v (vb = bits of
//            n), and its weight; a self-loop becomes the key ~0, which sorts
//            after every real key, and is counted.
//   sort     stable LSD radix sort, 8-bit digits over the 2*vb key bits
//            (own kernels: per-tile digit counts, a device-wide exclusive
//            scan of the counts in digit-major order, and a scatter that
//            ranks each tile in LDS by wave ballots and writes it out in
//            digit runs).  Weighted graphs first sort the entry indices by
//            weight, so that equal keys come out weight-ascending.
//   compact  the first entry of each run of equal keys (dedup) or every
//            entry: column and weight in place, row offsets from the row
//            boundaries; for directed graphs the transposed keys v << vb | u
//            of the kept entries, sorted the same way, give the in-edge CSC.
// Heavy-vertex chunk lists come from the device offsets (count, scan,
// write), as the host build does from its offsets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace dawn {

int build_chunks(dawn_graph* g);
int validate_csr(const dawn_csr* csr);

namespace {

constexpr int kRsPer = 16;                 // keys per thread per tile
constexpr int kRsTile = kBlock * kRsPer;   // 4096 keys per workgroup
constexpr int kScanPer = 16;
constexpr int kScanTile = kBlock * kScanPer;
constexpr uint64_t kSentinel = ~0ull;

// ---- counter-based random numbers (the host generators' draws) -----------

__device__ __forceinline__ uint64_t d_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t d_rng(uint64_t seed, uint64_t stream, uint64_t index) {
  return d_splitmix64(d_splitmix64(seed * 0xD1B54A32D192ED03ull + stream) ^
                      (index * 0x9E3779B97F4A7C15ull));
}

__device__ __forceinline__ uint32_t d_bounded32(uint64_t r, uint32_t range) {
  return (uint32_t)(((r >> 32) * (uint64_t)range) >> 32);
}

// Streams of the host generators (host_graph.cpp).
enum : uint64_t { kStPairU = 1, kStPairV = 2, kStWeight = 3, kStRmat = 4, kStExtra = 5 };

// ---- workgroup scans ------------------------------------------------------

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (l >= o) x += y;
  }
  return x;
}

// Exclusive scan of one value per thread over the workgroup (kBlock
// threads); *total receives the sum.  Every thread must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* total) {
  __shared__ uint32_t s_w[kBlock / kWave];
  const int w = threadIdx.x / kWave;
  const uint32_t inc = wave_incl_scan_u32(x);
  if (lane_id() == kWave - 1) s_w[w] = inc;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kBlock / kWave; i++) {
    if (i < w) base += s_w[i];
    tot += s_w[i];
  }
  __syncthreads();  // s_w is reused by the next call
  *total = tot;
  return base + inc - x;
}

// Sum of each tile of kScanTile values.
__global__ void __launch_bounds__(kBlock) scan_sums_kernel(const uint32_t* in, int64_t n,
                                                           uint32_t* sums) {
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
  uint32_t x = 0;
#pragma unroll
  for (int r = 0; r < kScanPer; r++) {
    const int64_t i = t0 + r * kBlock + threadIdx.x;
    if (i < n) x += in[i];
  }
  uint32_t tot;
  (void)block_excl_scan(x, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan of each tile in place, plus the tile's offset.
__global__ void __launch_bounds__(kBlock) scan_tiles_kernel(uint32_t* data, int64_t n,
                                                            const uint32_t* tile_off) {
  __shared__ uint32_t s[kScanTile];
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
  for (int r = 0; r < kScanPer; r++) {
    const int64_t i = t0 + r * kBlock + threadIdx.x;
    s[r * kBlock + threadIdx.x] = i < n ? data[i] : 0u;
  }
  __syncthreads();
  uint32_t loc[kScanPer], x = 0;
#pragma unroll
  for (int j = 0; j < kScanPer; j++) {
    loc[j] = x;
    x += s[threadIdx.x * kScanPer + j];
  }
  uint32_t tot;
  const uint32_t base = block_excl_scan(x, &tot) + (tile_off ? tile_off[blockIdx.x] : 0u);
#pragma unroll
  for (int j = 0; j < kScanPer; j++) s[threadIdx.x * kScanPer + j] = base + loc[j];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kScanPer; r++) {
    const int64_t i = t0 + r * kBlock + threadIdx.x;
    if (i < n) data[i] = s[r * kBlock + threadIdx.x];
  }
}

// Device-wide exclusive scan of n uint32 values in place (sum < 2^32).
struct ScanTemp {
  std::vector<DevBuf<uint32_t>*> levels;
  ~ScanTemp() {
    for (auto* b : levels) delete b;
  }
};

int excl_scan(uint32_t* data, int64_t n, hipStream_t s, ScanTemp& t, size_t level = 0) {
  if (n <= 0) return DAWN_OK;
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    scan_tiles_kernel<<<1, kBlock, 0, s>>>(data, n, nullptr);
    DAWN_HIP(hipGetLastError());
    return DAWN_OK;
  }
  if (t.levels.size() <= level) t.levels.push_back(new DevBuf<uint32_t>());
  DevBuf<uint32_t>& sums = *t.levels[level];
  DAWN_TRY(sums.alloc((size_t)tiles));
  scan_sums_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(data, n, sums.p);
  DAWN_HIP(hipGetLastError());
  DAWN_TRY(excl_scan(sums.p, tiles, s, t, level + 1));
  scan_tiles_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(data, n, sums.p);
  DAWN_HIP(hipGetLastError());
  return DAWN_OK;
}

// ---- LSD radix sort ---------------------------------------------------------

// Digit counts of each tile, digit-major: hist[d * tiles + tile].
__global__ void __launch_bounds__(kBlock) rs_hist_kernel(const uint64_t* keys, int64_t N,
                                                         int shift, int64_t tiles,
                                                         uint32_t* hist) {
  __shared__ uint32_t s_cnt[256];
  s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t t0 = (int64_t)blockIdx.x * kRsTile;
#pragma unroll
  for (int r = 0; r < kRsPer; r++) {
    const int64_t i = t0 + r * kBlock + threadIdx.x;
    if (i < N) atomicAdd(&s_cnt[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * tiles + blockIdx.x] = s_cnt[threadIdx.x];
}

// One stable scatter pass: the tile is ranked in LDS (key order within a
// digit = index order), then written out digit run by digit run.
// offs = scanned counts: offs[d * tiles + tile] = output position of the
// tile's first key of digit d.
template <bool VALS>
__global__ void __launch_bounds__(kBlock) rs_scatter_kernel(const uint64_t* kin,
                                                            const int32_t* vin, uint64_t* kout,
                                                            int32_t* vout, int64_t N, int shift,
                                                            int64_t tiles, const uint32_t* offs) {
  __shared__ uint64_t s_key[kRsTile];
  __shared__ int32_t s_val[VALS ? kRsTile : 1];
  __shared__ uint32_t s_start[256], s_run[256], s_gbase[256];
  __shared__ uint32_t s_wcnt[kBlock / kWave][256];
  const int tid = threadIdx.x;
  const int wv = tid / kWave;
  const int64_t t0 = (int64_t)blockIdx.x * kRsTile;
  const int nt = (int)min<int64_t>(kRsTile, N - t0);
  // Digit counts of the tile -> local digit starts.
  s_run[tid] = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; w++) s_wcnt[w][tid] = 0;
  __syncthreads();
  uint64_t k[kRsPer];
  int32_t v[kRsPer];
  uint32_t d[kRsPer];
#pragma unroll
  for (int r = 0; r < kRsPer; r++) {
    const int li = r * kBlock + tid;
    k[r] = li < nt ? kin[t0 + li] : 0ull;
    v[r] = VALS && li < nt ? vin[t0 + li] : 0;
    d[r] = (uint32_t)(k[r] >> shift) & 255u;
    if (li < nt) atomicAdd(&s_run[d[r]], 1u);
  }
  __syncthreads();
  uint32_t tot;
  const uint32_t st = block_excl_scan(s_run[tid], &tot);
  __syncthreads();
  s_start[tid] = st;
  s_run[tid] = st;
  s_gbase[tid] = offs[(int64_t)tid * tiles + blockIdx.x];
  __syncthreads();
  const uint64_t below = lanes_below_mask();
#pragma unroll 1
  for (int r = 0; r < kRsPer; r++) {
    const bool valid = r * kBlock + tid < nt;
    uint64_t eq = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d[r] >> b) & 1u;
      const uint64_t m = __ballot(bit);
      eq &= bit ? m : ~m;
    }
    const uint32_t wrank = (uint32_t)__popcll(eq & below);
    if (valid && wrank == 0) s_wcnt[wv][d[r]] = (uint32_t)__popcll(eq);
    __syncthreads();
    if (valid) {
      uint32_t pos = s_run[d[r]] + wrank;
      for (int w = 0; w < wv; w++) pos += s_wcnt[w][d[r]];
      s_key[pos] = k[r];
      if (VALS) s_val[pos] = v[r];
    }
    __syncthreads();
    {
      uint32_t add = 0;
#pragma unroll
      for (int w = 0; w < kBlock / kWave; w++) {
        add += s_wcnt[w][tid];
        s_wcnt[w][tid] = 0;
      }
      s_run[tid] += add;
    }
    __syncthreads();
  }
  for (int i = tid; i < nt; i += kBlock) {
    const uint64_t key = s_key[i];
    const uint32_t dg = (uint32_t)(key >> shift) & 255u;
    const int64_t pos = (int64_t)s_gbase[dg] + (i - s_start[dg]);
    kout[pos] = key;
    if (VALS) vout[pos] = s_val[i];
  }
}

struct SortTemp {
  DevBuf<uint32_t> hist;
  ScanTemp scan;
};

// Stable LSD radix sort of N keys on their low `bits` bits, with optional
// int32 payload; ping-pongs between (k0, v0) and (k1, v1).  *in1 tells
// which buffer holds the result.
int radix_sort(uint64_t* k0, uint64_t* k1, int32_t* v0, int32_t* v1, int64_t N, int bits,
               hipStream_t s, SortTemp& t, bool* in1) {
  *in1 = false;
  if (N <= 1 || bits <= 0) return DAWN_OK;
  const int64_t tiles = (N + kRsTile - 1) / kRsTile;
  DAWN_TRY(t.hist.alloc((size_t)tiles * 256));
  bool cur1 = false;
  for (int shift = 0; shift < bits; shift += 8) {
    const uint64_t* ki = cur1 ? k1 : k0;
    uint64_t* ko = cur1 ? k0 : k1;
    const int32_t* vi = cur1 ? v1 : v0;
    int32_t* vo = cur1 ? v0 : v1;
    rs_hist_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(ki, N, shift, tiles, t.hist.p);
    DAWN_HIP(hipGetLastError());
    DAWN_TRY(excl_scan(t.hist.p, tiles * 256, s, t.scan));
    if (v0)
      rs_scatter_kernel<true><<<(unsigned)tiles, kBlock, 0, s>>>(ki, vi, ko, vo, N, shift, tiles,
                                                                 t.hist.p);
    else
      rs_scatter_kernel<false><<<(unsigned)tiles, kBlock, 0, s>>>(ki, nullptr, ko, nullptr, N,
                                                                  shift, tiles, t.hist.p);
    DAWN_HIP(hipGetLastError());
    cur1 = !cur1;
  }
  *in1 = cur1;
  return DAWN_OK;
}

// ---- emission of directed entries -----------------------------------------

// Generator / edge-list parameters of one emit launch.
struct EmitArgs {
  int kind;             // 0 edge list, 1 ER, 2 RMAT, 3 grid
  int32_t n;
  int64_t count;        // pairs
  int symmetrize;
  int weighted;
  int vb;               // key bits of a vertex id
  uint64_t seed;
  int32_t wmin, wmax;
  // ER / grid
  int32_t side;
  int64_t horiz, grid_edges;
  // RMAT
  int32_t scale;
  uint64_t ta, tb, tc;
  // edge list
  const int32_t* u;
  const int32_t* v;
  const int32_t* w;
  // out
  uint64_t* keys;
  int32_t* wts;
  unsigned long long* selfloops;
  unsigned int* bad;    // edge list: an id out of range or a negative weight
};

__device__ __forceinline__ int32_t d_weight(const EmitArgs& a, int64_t i) {
  const uint32_t range = (uint32_t)((int64_t)a.wmax - a.wmin + 1);
  return a.wmin + (int32_t)d_bounded32(d_rng(a.seed, kStWeight, (uint64_t)i), range);
}

// Pair i of the generator or edge list (the host generators' formulas).
__device__ __forceinline__ void d_pair(const EmitArgs& a, int64_t i, int32_t& u, int32_t& v,
                                       int32_t& w) {
  w = 1;
  if (a.kind == 0) {
    u = a.u[i];
    v = a.v[i];
    if (a.w) w = a.w[i];
    return;
  }
  if (a.kind == 1) {
    u = (int32_t)d_bounded32(d_rng(a.seed, kStPairU, (uint64_t)i), (uint32_t)a.n);
    v = (int32_t)d_bounded32(d_rng(a.seed, kStPairV, (uint64_t)i), (uint32_t)a.n);
  } else if (a.kind == 2) {
    uint32_t uu = 0, vv = 0;
    uint64_t r = 0;
    for (int32_t l = 0; l < a.scale; l++) {
      if ((l & 1) == 0) r = d_rng(a.seed, kStRmat, (uint64_t)i * 16 + (uint64_t)(l >> 1));
      const uint64_t x = (l & 1) ? (r >> 32) : (r & 0xffffffffull);
      uint32_t ub, vbit;
      if (x < a.ta) { ub = 0; vbit = 0; }
      else if (x < a.tb) { ub = 0; vbit = 1; }
      else if (x < a.tc) { ub = 1; vbit = 0; }
      else { ub = 1; vbit = 1; }
      uu = (uu << 1) | ub;
      vv = (vv << 1) | vbit;
    }
    u = (int32_t)uu;
    v = (int32_t)vv;
  } else {
    if (i < a.horiz) {
      const int64_t r = i / (a.side - 1), c = i % (a.side - 1);
      u = (int32_t)(r * a.side + c);
      v = u + 1;
    } else if (i < a.grid_edges) {
      const int64_t j = i - a.horiz;
      const int64_t r = j / a.side, c = j % a.side;
      u = (int32_t)(r * a.side + c);
      v = u + a.side;
    } else {
      const int64_t j = i - a.grid_edges;
      u = (int32_t)d_bounded32(d_rng(a.seed, kStExtra, (uint64_t)(2 * j)), (uint32_t)a.n);
      v = (int32_t)d_bounded32(d_rng(a.seed, kStExtra, (uint64_t)(2 * j + 1)), (uint32_t)a.n);
    }
  }
  if (a.weighted) w = d_weight(a, i);
}

// Entries of pair i at 2i, 2i+1 (symmetrized) or i.
__global__ void __launch_bounds__(kBlock) emit_kernel(EmitArgs a) {
  unsigned long long loops = 0;
  unsigned bad = 0;
  const int per = a.symmetrize ? 2 : 1;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.count;
       i += (int64_t)gridDim.x * kBlock) {
    int32_t u, v, w;
    d_pair(a, i, u, v, w);
    if (a.kind == 0 && (u < 0 || u >= a.n || v < 0 || v >= a.n || w < 0)) {
      bad = 1;
      u = v = 0;  // recorded as a self-loop; the call fails
    }
    uint64_t k0 = kSentinel, k1 = kSentinel;
    if (u != v) {
      k0 = ((uint64_t)u << a.vb) | (uint64_t)v;
      k1 = ((uint64_t)v << a.vb) | (uint64_t)u;
    } else {
      loops += 1;
    }
    a.keys[per * i] = k0;
    if (a.weighted) a.wts[per * i] = w;
    if (per == 2) {
      a.keys[2 * i + 1] = k1;
      if (a.weighted) a.wts[2 * i + 1] = w;
    }
  }
  loops = wave_sum_u64(loops);
  if (lane_id() == 0 && loops) atomicAdd(a.selfloops, loops);
  if (bad) atomicOr(a.bad, 1u);
}

// Weight-first order: keys sorted by weight carry their index; gather the
// entries in that order.
__global__ void iota_kernel(int32_t* p, int64_t N) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

__global__ void weight_keys_kernel(const int32_t* w, uint64_t* k, int64_t N) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    k[i] = (uint64_t)(uint32_t)w[i];
}

__global__ void gather_kernel(const int32_t* perm, const uint64_t* kin, const int32_t* win,
                              uint64_t* kout, int32_t* wout, int64_t N) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t p = perm[i];
    kout[i] = kin[p];
    wout[i] = win[p];
  }
}

// ---- compaction -------------------------------------------------------------

__device__ __forceinline__ bool is_head(const uint64_t* keys, int64_t i, int dedup) {
  return !dedup || i == 0 || keys[i] != keys[i - 1];
}

// Kept entries per tile of kScanTile.
__global__ void __launch_bounds__(kBlock) cmp_count_kernel(const uint64_t* keys, int64_t E,
                                                           int dedup, uint32_t* cnt) {
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
  uint32_t x = 0;
#pragma unroll
  for (int r = 0; r < kScanPer; r++) {
    const int64_t i = t0 + r * kBlock + threadIdx.x;
    if (i < E && is_head(keys, i, dedup)) x++;
  }
  uint32_t tot;
  (void)block_excl_scan(x, &tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// Writes the kept entries (column = low vb bits, weight) at their ranks and
// the row offsets at the row boundaries; with tkeys, also the transposed
// keys (column << vb | row) and weights of the kept entries (in-edge CSC).
__global__ void __launch_bounds__(kBlock) cmp_write_kernel(
    const uint64_t* keys, const int32_t* w, int64_t E, int dedup, int vb, const uint32_t* tile_off,
    int32_t n, int64_t m, int32_t* col, int32_t* w_out, int32_t* row_ptr, uint64_t* tkeys,
    int32_t* tw) {
  __shared__ uint32_t s_flag[kScanTile];
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
  const uint64_t vmask = (1ull << vb) - 1;
#pragma unroll
  for (int r = 0; r < kScanPer; r++) {
    const int64_t i = t0 + r * kBlock + threadIdx.x;
    s_flag[r * kBlock + threadIdx.x] = (i < E && is_head(keys, i, dedup)) ? 1u : 0u;
  }
  __syncthreads();
  uint32_t x = 0;
#pragma unroll
  for (int j = 0; j < kScanPer; j++) x += s_flag[threadIdx.x * kScanPer + j];
  uint32_t tot;
  uint32_t pos = block_excl_scan(x, &tot) + tile_off[blockIdx.x];
  for (int j = 0; j < kScanPer; j++) {
    const int64_t i = t0 + threadIdx.x * kScanPer + j;
    if (i >= E) break;
    const uint64_t key = keys[i];
    const int32_t u = (int32_t)(key >> vb);
    const int32_t v = (int32_t)(key & vmask);
    const int32_t uprev = i == 0 ? -1 : (int32_t)(keys[i - 1] >> vb);
    if (s_flag[threadIdx.x * kScanPer + j]) {
      col[pos] = v;
      if (w_out) w_out[pos] = w[i];
      if (tkeys) {
        tkeys[pos] = ((uint64_t)v << vb) | (uint64_t)u;
        if (tw) tw[pos] = w[i];
      }
      if (u != uprev)
        for (int32_t r = uprev + 1; r <= u; r++) row_ptr[r] = (int32_t)pos;
      pos++;
    }
    if (i == E - 1)
      for (int32_t r = u + 1; r <= n; r++) row_ptr[r] = (int32_t)m;
  }
}

__global__ void fill_i32_kernel(int32_t* p, int64_t N, int32_t value) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

// Sorted keys [0, E) -> CSR arrays (allocated here) on the device.
int compact(const uint64_t* keys, const int32_t* w, int64_t E, int dedup, int vb, int32_t n,
            hipStream_t s, SortTemp& t, int32_t** row_ptr, int32_t** col, int32_t** w_out,
            int64_t* m_out, uint64_t* tkeys, int32_t* tw) {
  auto dalloc = [](int32_t** p, size_t count) -> int {
    DAWN_HIP(hipMalloc((void**)p, sizeof(int32_t) * std::max<size_t>(count, 1)));
    return DAWN_OK;
  };
  DAWN_TRY(dalloc(row_ptr, (size_t)n + 1));
  int64_t m = 0;
  const int64_t tiles = (E + kScanTile - 1) / kScanTile;
  DevBuf<uint32_t> cnt;
  if (E > 0) {
    DAWN_TRY(cnt.alloc((size_t)tiles));
    cmp_count_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(keys, E, dedup, cnt.p);
    DAWN_HIP(hipGetLastError());
    uint32_t last = 0;
    DAWN_HIP(hipMemcpyAsync(&last, cnt.p + tiles - 1, sizeof last, hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    DAWN_TRY(excl_scan(cnt.p, tiles, s, t.scan));
    uint32_t off = 0;
    DAWN_HIP(hipMemcpyAsync(&off, cnt.p + tiles - 1, sizeof off, hipMemcpyDeviceToHost, s));
    DAWN_HIP(hipStreamSynchronize(s));
    m = (int64_t)off + last;
  }
  if (m > (int64_t)INT32_MAX) return DAWN_ERR_OVERFLOW;
  DAWN_TRY(dalloc(col, (size_t)m));
  if (w) DAWN_TRY(dalloc(w_out, (size_t)m));
  if (E > 0) {
    cmp_write_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(keys, w, E, dedup, vb, cnt.p, n, m, *col,
                                                        w ? *w_out : nullptr, *row_ptr, tkeys, tw);
  } else {
    fill_i32_kernel<<<grid_for((int64_t)n + 1), kBlock, 0, s>>>(*row_ptr, (int64_t)n + 1, 0);
  }
  DAWN_HIP(hipGetLastError());
  DAWN_HIP(hipStreamSynchronize(s));
  *m_out = m;
  return DAWN_OK;
}

int key_bits(int32_t n) {
  int b = 1;
  while (b < 31 && (1ll << b) < (int64_t)n) b++;
  return b;
}

int weight_bits(int32_t wmax) {
  int b = 1;
  while (b < 31 && (1ll << b) <= (int64_t)wmax) b++;
  return b;
}

// Builds the device graph from one emit launch.
int build_device(EmitArgs ea, int dedup, int32_t wmax_bound, dawn_graph** out) {
  if (!out) return DAWN_ERR_INVALID_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return DAWN_ERR_NO_DEVICE;
  }
  const int32_t n = ea.n;
  const int per = ea.symmetrize ? 2 : 1;
  const int64_t N = ea.count * per;
  if (N > (int64_t)INT32_MAX) return DAWN_ERR_OVERFLOW;
  dawn_graph* g = new dawn_graph();
  auto fail = [&](int rc) {
    dawn_graph_destroy(g);
    return rc;
  };
  g->n = n;
  g->directed = ea.symmetrize ? 0 : 1;
  g->weighted = ea.weighted ? 1 : 0;
  if (hipGetDevice(&g->device) != hipSuccess) return fail(DAWN_ERR_HIP);
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(DAWN_ERR_HIP);
  hipStream_t s = g->stream;
  int rc = DAWN_OK;
  {
    SortTemp t;
    DevBuf<uint64_t> ka, kb;
    DevBuf<int32_t> wa, wb, perm;
    DevBuf<unsigned long long> loops;
    DevBuf<unsigned int> bad;
    if ((rc = ka.alloc((size_t)N)) || (rc = kb.alloc((size_t)N))) return fail(rc);
    if (ea.weighted && ((rc = wa.alloc((size_t)N)) || (rc = wb.alloc((size_t)N)))) return fail(rc);
    if ((rc = loops.alloc(1)) || (rc = bad.alloc(1))) return fail(rc);
    if (hipMemsetAsync(loops.p, 0, sizeof(unsigned long long), s) != hipSuccess ||
        hipMemsetAsync(bad.p, 0, sizeof(unsigned int), s) != hipSuccess)
      return fail(DAWN_ERR_HIP);
    ea.vb = key_bits(n);
    ea.keys = ka.p;
    ea.wts = ea.weighted ? wa.p : nullptr;
    ea.selfloops = loops.p;
    ea.bad = bad.p;
    if (ea.count > 0) {
      emit_kernel<<<grid_for(ea.count, kBlock, 65536u), kBlock, 0, s>>>(ea);
      if (hipGetLastError() != hipSuccess) return fail(DAWN_ERR_HIP);
    }
    unsigned long long h_loops = 0;
    unsigned int h_bad = 0;
    if (hipMemcpyAsync(&h_loops, loops.p, sizeof h_loops, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&h_bad, bad.p, sizeof h_bad, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(DAWN_ERR_HIP);
    if (h_bad) return fail(DAWN_ERR_INVALID_ARG);
    const int64_t E = N - (int64_t)h_loops * per;  // entries before the sentinels
    uint64_t* keys = ka.p;
    int32_t* wts = ea.weighted ? wa.p : nullptr;
    bool in1 = false;
    if (ea.weighted && N > 1) {
      // Weight-first: entry indices sorted by weight, then the entries
      // gathered in that order, so the stable key sort leaves equal keys
      // weight-ascending (dedup keeps the minimum, as the host build).
      if ((rc = perm.alloc((size_t)N))) return fail(rc);
      DevBuf<int32_t> perm2;
      if ((rc = perm2.alloc((size_t)N))) return fail(rc);
      weight_keys_kernel<<<grid_for(N, kBlock, 65536u), kBlock, 0, s>>>(wa.p, kb.p, N);
      iota_kernel<<<grid_for(N, kBlock, 65536u), kBlock, 0, s>>>(perm.p, N);
      // kb holds the weight keys; ka keeps the entry keys for the gather.
      DevBuf<uint64_t> kc;
      if ((rc = kc.alloc((size_t)N))) return fail(rc);
      if ((rc = radix_sort(kb.p, kc.p, perm.p, perm2.p, N, weight_bits(wmax_bound), s, t, &in1)))
        return fail(rc);
      const int32_t* order = in1 ? perm2.p : perm.p;
      gather_kernel<<<grid_for(N, kBlock, 65536u), kBlock, 0, s>>>(order, ka.p, wa.p, kb.p, wb.p, N);
      if (hipGetLastError() != hipSuccess) return fail(DAWN_ERR_HIP);
      // Sort (kb, wb) by key, ping-ponging with (ka, wa).
      if ((rc = radix_sort(kb.p, ka.p, wb.p, wa.p, N, 2 * ea.vb, s, t, &in1))) return fail(rc);
      keys = in1 ? ka.p : kb.p;
      wts = in1 ? wa.p : wb.p;
    } else {
      if ((rc = radix_sort(ka.p, kb.p, nullptr, nullptr, N, 2 * ea.vb, s, t, &in1))) return fail(rc);
      keys = in1 ? kb.p : ka.p;
    }
    // Directed graphs: the kept entries' transposed keys go to the free
    // key buffer, for the in-edge CSC.
    uint64_t* tk = nullptr;
    int32_t* tw = nullptr;
    if (g->directed) {
      tk = keys == ka.p ? kb.p : ka.p;
      if (ea.weighted) tw = wts == wa.p ? wb.p : wa.p;
    }
    int64_t m = 0;
    if ((rc = compact(keys, wts, E, dedup, ea.vb, n, s, t, &g->row_ptr, &g->col,
                      &g->weight, &m, tk, tw)))
      return fail(rc);
    g->m = m;
    if (g->directed) {
      // tk/tw hold m transposed entries in (row, column, weight) order;
      // sort them by (column, row) into the free buffers.
      uint64_t* tk2 = tk == ka.p ? kb.p : ka.p;
      int32_t* tw2 = ea.weighted ? (tw == wa.p ? wb.p : wa.p) : nullptr;
      if ((rc = radix_sort(tk, tk2, tw, tw2, m, 2 * ea.vb, s, t, &in1))) return fail(rc);
      const uint64_t* ck = in1 ? tk2 : tk;
      const int32_t* cw = ea.weighted ? (in1 ? tw2 : tw) : nullptr;
      int64_t m2 = 0;
      if ((rc = compact(ck, cw, m, 0, ea.vb, n, s, t, &g->in_row_ptr, &g->in_col,
                        &g->in_weight, &m2, nullptr, nullptr)))
        return fail(rc);
    } else {
      g->in_row_ptr = g->row_ptr;
      g->in_col = g->col;
      g->in_weight = g->weight;
    }
  }
  // Host copies of the offsets (degrees for the drivers) and degree stats.
  g->h_row_ptr.resize((size_t)n + 1);
  if (hipMemcpy(g->h_row_ptr.data(), g->row_ptr, sizeof(int32_t) * ((size_t)n + 1),
                hipMemcpyDeviceToHost) != hipSuccess)
    return fail(DAWN_ERR_HIP);
  if (g->directed) {
    g->h_in_row_ptr.resize((size_t)n + 1);
    if (hipMemcpy(g->h_in_row_ptr.data(), g->in_row_ptr, sizeof(int32_t) * ((size_t)n + 1),
                  hipMemcpyDeviceToHost) != hipSuccess)
      return fail(DAWN_ERR_HIP);
  }
  int32_t maxo = 0, maxi = 0;
  int64_t z = 0;
  const std::vector<int32_t>& ip = g->directed ? g->h_in_row_ptr : g->h_row_ptr;
  for (int32_t v = 0; v < n; v++) {
    maxo = std::max(maxo, g->h_row_ptr[v + 1] - g->h_row_ptr[v]);
    maxi = std::max(maxi, ip[v + 1] - ip[v]);
    z += ip[v + 1] == ip[v];
  }
  g->max_out_degree = maxo;
  g->max_in_degree = maxi;
  g->zero_in = z;
  (void)tuned_knobs_for(g, &g->knobs);
  if ((rc = build_chunks(g))) return fail(rc);
  *out = g;
  return DAWN_OK;
}

EmitArgs base_args(int kind, int32_t n, int64_t count, int symmetrize, int weighted, uint64_t seed,
                   int32_t wmin, int32_t wmax) {
  EmitArgs a{};
  a.kind = kind;
  a.n = n;
  a.count = count;
  a.symmetrize = symmetrize;
  a.weighted = weighted;
  a.seed = seed;
  a.wmin = wmin;
  a.wmax = wmax;
  return a;
}

bool weights_ok(int32_t weighted, int32_t wmin, int32_t wmax) {
  return !weighted || (wmin >= 0 && wmax >= wmin);
}

}  // namespace
}  // namespace dawn

using namespace dawn;

extern "C" {

int dawn_graph_from_edges(int32_t n, int64_t count, const int32_t* u, const int32_t* v,
                          const int32_t* w, int32_t symmetrize, int32_t dedup, dawn_graph** out) {
  if (!out || n <= 0 || count < 0 || (count > 0 && (!u || !v))) return DAWN_ERR_INVALID_ARG;
  // Host arrays are staged on the device; device arrays are read in place.
  DevBuf<int32_t> du, dv, dw;
  const int32_t *pu = u, *pv = v, *pw = w;
  int32_t wmax = 0;
  if (count > 0 && !is_device_pointer(u)) {
    DAWN_TRY(du.alloc((size_t)count));
    DAWN_TRY(dv.alloc((size_t)count));
    DAWN_HIP(hipMemcpy(du.p, u, sizeof(int32_t) * count, hipMemcpyHostToDevice));
    DAWN_HIP(hipMemcpy(dv.p, v, sizeof(int32_t) * count, hipMemcpyHostToDevice));
    pu = du.p;
    pv = dv.p;
  }
  if (w && count > 0) {
    std::vector<int32_t> hw;
    const int32_t* wh = w;
    if (is_device_pointer(w)) {
      hw.resize((size_t)count);
      DAWN_HIP(hipMemcpy(hw.data(), w, sizeof(int32_t) * count, hipMemcpyDeviceToHost));
      wh = hw.data();
    } else {
      DAWN_TRY(dw.alloc((size_t)count));
      DAWN_HIP(hipMemcpy(dw.p, w, sizeof(int32_t) * count, hipMemcpyHostToDevice));
      pw = dw.p;
    }
    for (int64_t i = 0; i < count; i++) wmax = std::max(wmax, wh[i]);
  }
  EmitArgs a = base_args(0, n, count, symmetrize ? 1 : 0, w ? 1 : 0, 0, 0, 0);
  a.u = pu;
  a.v = pv;
  a.w = pw;
  return build_device(a, dedup ? 1 : 0, std::max(wmax, 1), out);
}

int dawn_graph_gen_er(int32_t n, double avg_degree, uint64_t seed, int32_t weighted, int32_t wmin,
                      int32_t wmax, dawn_graph** out) {
  if (!out || n <= 1 || avg_degree < 0 || !weights_ok(weighted, wmin, wmax))
    return DAWN_ERR_INVALID_ARG;
  const int64_t M = (int64_t)std::llround((double)n * avg_degree / 2.0);
  EmitArgs a = base_args(1, n, M, 1, weighted ? 1 : 0, seed, wmin, wmax);
  return build_device(a, 1, std::max(wmax, 1), out);
}

int dawn_graph_gen_rmat(int32_t scale, int32_t edge_factor, double ra, double rb, double rc,
                        uint64_t seed, int32_t weighted, int32_t wmin, int32_t wmax,
                        dawn_graph** out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || ra < 0 || rb < 0 || rc < 0 ||
      ra + rb + rc > 1.0 || !weights_ok(weighted, wmin, wmax))
    return DAWN_ERR_INVALID_ARG;
  const int32_t n = (int32_t)(1u << scale);
  const int64_t M = (int64_t)edge_factor << scale;
  EmitArgs a = base_args(2, n, M, 1, weighted ? 1 : 0, seed, wmin, wmax);
  a.scale = scale;
  a.ta = (uint64_t)std::llround(ra * 4294967296.0);
  a.tb = (uint64_t)std::llround((ra + rb) * 4294967296.0);
  a.tc = (uint64_t)std::llround((ra + rb + rc) * 4294967296.0);
  return build_device(a, 1, std::max(wmax, 1), out);
}

int dawn_graph_gen_grid(int32_t side, double extra_fraction, uint64_t seed, int32_t weighted,
                        int32_t wmin, int32_t wmax, dawn_graph** out) {
  if (!out || side < 2 || side > 46340 || extra_fraction < 0 || !weights_ok(weighted, wmin, wmax))
    return DAWN_ERR_INVALID_ARG;
  const int32_t n = side * side;
  const int64_t horiz = (int64_t)side * (side - 1);
  const int64_t grid_edges = 2 * horiz;
  const int64_t extra = (int64_t)std::llround(extra_fraction * (double)grid_edges);
  EmitArgs a = base_args(3, n, grid_edges + extra, 1, weighted ? 1 : 0, seed, wmin, wmax);
  a.side = side;
  a.horiz = horiz;
  a.grid_edges = grid_edges;
  return build_device(a, 1, std::max(wmax, 1), out);
}

int dawn_graph_download(const dawn_graph* g, dawn_csr* out) {
  if (!g || !out) return DAWN_ERR_INVALID_ARG;
  DAWN_HIP(hipSetDevice(g->device));
  const size_t n = (size_t)g->n, m = (size_t)g->m;
  int32_t* rp = (int32_t*)std::malloc(sizeof(int32_t) * (n + 1));
  int32_t* cc = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(m, 1));
  int32_t* ww = g->weighted ? (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(m, 1)) : nullptr;
  if (!rp || !cc || (g->weighted && !ww)) {
    std::free(rp);
    std::free(cc);
    std::free(ww);
    return DAWN_ERR_OUT_OF_MEMORY;
  }
  hipError_t e = hipMemcpy(rp, g->row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost);
  if (e == hipSuccess && m) e = hipMemcpy(cc, g->col, sizeof(int32_t) * m, hipMemcpyDeviceToHost);
  if (e == hipSuccess && m && ww)
    e = hipMemcpy(ww, g->weight, sizeof(int32_t) * m, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    std::free(rp);
    std::free(cc);
    std::free(ww);
    set_hip_error(e, "hipMemcpy", __FILE__, __LINE__);
    return DAWN_ERR_HIP;
  }
  out->n = g->n;
  out->m = g->m;
  out->directed = g->directed;
  out->weighted = g->weighted;
  out->row_ptr = rp;
  out->col = cc;
  out->weight = ww;
  return DAWN_OK;
}

}  // extern "C"
