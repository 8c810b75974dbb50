// Host-side graph construction: seeded generators, edge-list -> CSR build,
// MatrixMarket reader/writer, source lists and distance output files.
//
// Reference semantics followed:
//   MTX header / dispatch      src/graph.cpp:9-77
//   edge readers               src/io.cpp:2-192 (1-based -> 0-based, self-loops
//                              dropped, undirected => both directions,
//                              rows sorted, duplicates kept)
//   COO -> CSR                 src/matrix.cpp:28-82
//   readList / outfile         src/io.cpp:194-231
//
// The generators are new (the reference loads files only).  Every random
// draw is a pure function of (seed, stream, index), so output is identical
// for any OpenMP thread count.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dawn_capi.h"
#include "host_util.hpp"

namespace {

using dawn::rng_u64;

// 32-bit uniform integer in [0, range) from a 64-bit random (multiply-high).
inline uint32_t bounded32(uint64_t r, uint32_t range) {
  return (uint32_t)(((r >> 32) * (uint64_t)range) >> 32);
}

enum : uint64_t {
  kStreamPairU = 1,
  kStreamPairV = 2,
  kStreamWeight = 3,
  kStreamRmat = 4,
  kStreamExtra = 5,
  kStreamSources = 6
};

// Generic CSR builder.  pair(i, u, v, w) yields edge i; returns false to skip.
// Steps: count out-degrees (atomic), exclusive scan, scatter via atomic
// cursors, per-row sort (+ optional dedup keeping min weight), compaction.
// The final CSR does not depend on scatter order, so it is deterministic.
template <class PairFn>
int build_csr(int32_t n, int64_t count, bool symmetrize, bool dedup,
              bool weighted, PairFn pair, dawn_csr* out) {
  if (n < 0 || count < 0) return DAWN_ERR_INVALID_ARG;
  std::vector<std::atomic<int64_t>> deg(n);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; i++) deg[i].store(0, std::memory_order_relaxed);

#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; i++) {
    int32_t u, v, w;
    if (!pair(i, u, v, w)) continue;
    if (u == v) continue;
    deg[u].fetch_add(1, std::memory_order_relaxed);
    if (symmetrize) deg[v].fetch_add(1, std::memory_order_relaxed);
  }
  std::vector<int64_t> off(n + 1);
  off[0] = 0;
  for (int32_t i = 0; i < n; i++)
    off[i + 1] = off[i] + deg[i].load(std::memory_order_relaxed);
  const int64_t raw = off[n];

  std::vector<int32_t> col(raw);
  std::vector<int32_t> wt(weighted ? raw : 0);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; i++) deg[i].store(off[i], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; i++) {
    int32_t u, v, w;
    if (!pair(i, u, v, w)) continue;
    if (u == v) continue;
    int64_t p = deg[u].fetch_add(1, std::memory_order_relaxed);
    col[p] = v;
    if (weighted) wt[p] = w;
    if (symmetrize) {
      int64_t q = deg[v].fetch_add(1, std::memory_order_relaxed);
      col[q] = u;
      if (weighted) wt[q] = w;
    }
  }

  // Sort each row by (col, weight); dedup keeps the first = minimum weight.
  std::vector<int64_t> newdeg(n);
#pragma omp parallel
  {
    std::vector<std::pair<int32_t, int32_t>> tmp;
#pragma omp for schedule(dynamic, 1024)
    for (int32_t r = 0; r < n; r++) {
      const int64_t b = off[r], e = off[r + 1];
      int64_t len = e - b;
      if (weighted) {
        tmp.resize(len);
        for (int64_t k = 0; k < len; k++) tmp[k] = {col[b + k], wt[b + k]};
        std::sort(tmp.begin(), tmp.end());
        int64_t o = 0;
        for (int64_t k = 0; k < len; k++) {
          if (dedup && o > 0 && tmp[o - 1].first == tmp[k].first) continue;
          tmp[o++] = tmp[k];
        }
        for (int64_t k = 0; k < o; k++) {
          col[b + k] = tmp[k].first;
          wt[b + k] = tmp[k].second;
        }
        len = o;
      } else {
        std::sort(col.begin() + b, col.begin() + e);
        if (dedup) len = std::unique(col.begin() + b, col.begin() + e) - (col.begin() + b);
      }
      newdeg[r] = len;
    }
  }
  int64_t m = 0;
  for (int32_t r = 0; r < n; r++) m += newdeg[r];
  if (m > (int64_t)INT32_MAX) return DAWN_ERR_OVERFLOW;

  int32_t* rp = (int32_t*)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
  int32_t* cc = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
  int32_t* ww = weighted ? (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1)) : nullptr;
  if (!rp || !cc || (weighted && !ww)) {
    std::free(rp); std::free(cc); std::free(ww);
    return DAWN_ERR_OUT_OF_MEMORY;
  }
  rp[0] = 0;
  for (int32_t r = 0; r < n; r++) rp[r + 1] = rp[r] + (int32_t)newdeg[r];
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t r = 0; r < n; r++) {
    std::memcpy(cc + rp[r], col.data() + off[r], sizeof(int32_t) * newdeg[r]);
    if (weighted) std::memcpy(ww + rp[r], wt.data() + off[r], sizeof(int32_t) * newdeg[r]);
  }
  out->n = n;
  out->m = m;
  out->directed = symmetrize ? 0 : 1;
  out->weighted = weighted ? 1 : 0;
  out->row_ptr = rp;
  out->col = cc;
  out->weight = ww;
  return DAWN_OK;
}

inline int32_t draw_weight(uint64_t seed, int64_t i, int32_t wmin, int32_t wmax) {
  uint32_t range = (uint32_t)((int64_t)wmax - wmin + 1);
  return wmin + (int32_t)bounded32(rng_u64(seed, kStreamWeight, (uint64_t)i), range);
}

bool valid_weights(int32_t weighted, int32_t wmin, int32_t wmax) {
  return !weighted || (wmin >= 0 && wmax >= wmin);
}

// ---- MatrixMarket ---------------------------------------------------------

struct MtxEdge {
  int32_t r, c;
  double w;
};

char* next_line(char*& p, char* end) {
  if (p >= end) return nullptr;
  char* s = p;
  char* nl = (char*)std::memchr(p, '\n', end - p);
  if (nl) {
    *nl = '\0';
    p = nl + 1;
  } else {
    p = end;
  }
  return s;
}

}  // namespace

namespace dawn {

// Structure checks before upload: monotone offsets, columns in range,
// non-negative weights (the min-plus engine assumes w >= 0).
int validate_csr(const dawn_csr* csr) {
  const int32_t n = csr->n;
  const int64_t m = csr->m;
  if (csr->row_ptr[0] != 0 || (int64_t)csr->row_ptr[n] != m) return DAWN_ERR_INVALID_ARG;
  for (int32_t i = 0; i < n; i++)
    if (csr->row_ptr[i + 1] < csr->row_ptr[i]) return DAWN_ERR_INVALID_ARG;
  int bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
  for (int64_t k = 0; k < m; k++) {
    if (csr->col[k] < 0 || csr->col[k] >= n) bad |= 1;
    if (csr->weight && csr->weight[k] < 0) bad |= 1;
  }
  return bad ? DAWN_ERR_INVALID_ARG : DAWN_OK;
}

// In-edge CSC (column-major copy of the adjacency): ip[n+1], ic[m] = source
// rows ascending within each column, iw[m] their weights.
int build_csc(const dawn_csr* csr, std::vector<int32_t>& ip, std::vector<int32_t>& ic,
              std::vector<int32_t>& iw, int32_t* max_in_degree) {
  const int32_t n = csr->n;
  const int64_t m = csr->m;
  ip.assign((size_t)n + 1, 0);
  ic.resize((size_t)m);
  iw.resize(csr->weight ? (size_t)m : 0);
  for (int64_t k = 0; k < m; k++) ip[csr->col[k] + 1]++;
  int32_t maxi = 0;
  for (int32_t i = 0; i < n; i++) {
    maxi = std::max(maxi, ip[i + 1]);
    ip[i + 1] += ip[i];
  }
  std::vector<int32_t> cur(ip.begin(), ip.end() - 1);
  for (int32_t r = 0; r < n; r++)
    for (int32_t k = csr->row_ptr[r]; k < csr->row_ptr[r + 1]; k++) {
      const int32_t p = cur[csr->col[k]]++;
      ic[p] = r;
      if (csr->weight) iw[p] = csr->weight[k];
    }
  *max_in_degree = maxi;
  return DAWN_OK;
}

}  // namespace dawn

extern "C" {

const char* dawn_status_string(int status) {
  switch (status) {
    case DAWN_OK: return "ok";
    case DAWN_ERR_INVALID_ARG: return "invalid argument";
    case DAWN_ERR_ISOLATED_SOURCE: return "source is an isolated node";
    case DAWN_ERR_NO_DEVICE: return "no HIP device";
    case DAWN_ERR_HIP: return "HIP runtime error";
    case DAWN_ERR_OUT_OF_MEMORY: return "out of memory";
    case DAWN_ERR_IO: return "cannot open file";
    case DAWN_ERR_FORMAT: return "invalid file format";
    case DAWN_ERR_OVERFLOW: return "size exceeds int32 CSR offsets";
    case DAWN_ERR_UNSUPPORTED: return "unsupported";
    case DAWN_ERR_NOT_WEIGHTED: return "graph has no weights";
    case DAWN_ERR_SCRATCH_LIMIT: return "graph exceeds the fixed scratch region";
    default: return "unknown status";
  }
}

void dawn_free(void* p) { std::free(p); }

void dawn_csr_free(dawn_csr* csr) {
  if (!csr) return;
  std::free(csr->row_ptr);
  std::free(csr->col);
  std::free(csr->weight);
  csr->row_ptr = nullptr;
  csr->col = nullptr;
  csr->weight = nullptr;
  csr->n = 0;
  csr->m = 0;
}

int dawn_gen_er(int32_t n, double avg_degree, uint64_t seed, int32_t weighted,
                int32_t wmin, int32_t wmax, dawn_csr* out) {
  if (!out || n <= 1 || avg_degree < 0 || !valid_weights(weighted, wmin, wmax))
    return DAWN_ERR_INVALID_ARG;
  // G(n, M): M = round(n * d / 2) uniform undirected pairs, then symmetrize,
  // dedup and drop self-loops (=> about n*d directed edges).
  const int64_t M = (int64_t)std::llround((double)n * avg_degree / 2.0);
  auto pair = [&](int64_t i, int32_t& u, int32_t& v, int32_t& w) {
    u = (int32_t)bounded32(rng_u64(seed, kStreamPairU, (uint64_t)i), (uint32_t)n);
    v = (int32_t)bounded32(rng_u64(seed, kStreamPairV, (uint64_t)i), (uint32_t)n);
    w = weighted ? draw_weight(seed, i, wmin, wmax) : 1;
    return true;
  };
  return build_csr(n, M, true, true, weighted != 0, pair, out);
}

int dawn_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                  double c, uint64_t seed, int32_t weighted, int32_t wmin,
                  int32_t wmax, dawn_csr* out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 ||
      c < 0 || a + b + c > 1.0 || !valid_weights(weighted, wmin, wmax))
    return DAWN_ERR_INVALID_ARG;
  const int32_t n = (int32_t)(1u << scale);
  const int64_t M = (int64_t)edge_factor << scale;
  // Quadrant thresholds on a 32-bit uniform: [0,ta) -> (0,0), [ta,tb) -> (0,1),
  // [tb,tc) -> (1,0), [tc, 2^32) -> (1,1).  No noise, no vertex relabeling.
  const uint64_t ta = (uint64_t)std::llround(a * 4294967296.0);
  const uint64_t tb = (uint64_t)std::llround((a + b) * 4294967296.0);
  const uint64_t tc = (uint64_t)std::llround((a + b + c) * 4294967296.0);
  auto pair = [&](int64_t i, int32_t& u, int32_t& v, int32_t& w) {
    uint32_t uu = 0, vv = 0;
    uint64_t r = 0;
    for (int32_t l = 0; l < scale; l++) {
      if ((l & 1) == 0) r = rng_u64(seed, kStreamRmat, (uint64_t)i * 16 + (uint64_t)(l >> 1));
      uint64_t x = (l & 1) ? (r >> 32) : (r & 0xffffffffull);
      uint32_t ub, vb;
      if (x < ta) { ub = 0; vb = 0; }
      else if (x < tb) { ub = 0; vb = 1; }
      else if (x < tc) { ub = 1; vb = 0; }
      else { ub = 1; vb = 1; }
      uu = (uu << 1) | ub;
      vv = (vv << 1) | vb;
    }
    u = (int32_t)uu;
    v = (int32_t)vv;
    w = weighted ? draw_weight(seed, i, wmin, wmax) : 1;
    return true;
  };
  return build_csr(n, M, true, true, weighted != 0, pair, out);
}

int dawn_gen_grid(int32_t side, double extra_fraction, uint64_t seed,
                  int32_t weighted, int32_t wmin, int32_t wmax, dawn_csr* out) {
  if (!out || side < 2 || side > 46340 || extra_fraction < 0 ||
      !valid_weights(weighted, wmin, wmax))
    return DAWN_ERR_INVALID_ARG;
  const int32_t n = side * side;
  const int64_t horiz = (int64_t)side * (side - 1);
  const int64_t grid_edges = 2 * horiz;  // horizontal + vertical
  const int64_t extra = (int64_t)std::llround(extra_fraction * (double)grid_edges);
  auto pair = [&](int64_t i, int32_t& u, int32_t& v, int32_t& w) {
    if (i < horiz) {  // (r, c) - (r, c+1)
      int64_t r = i / (side - 1), c = i % (side - 1);
      u = (int32_t)(r * side + c);
      v = u + 1;
    } else if (i < grid_edges) {  // (r, c) - (r+1, c)
      int64_t j = i - horiz;
      int64_t r = j / side, c = j % side;
      u = (int32_t)(r * side + c);
      v = u + side;
    } else {  // uniform random long-range pair
      int64_t j = i - grid_edges;
      u = (int32_t)bounded32(rng_u64(seed, kStreamExtra, (uint64_t)(2 * j)), (uint32_t)n);
      v = (int32_t)bounded32(rng_u64(seed, kStreamExtra, (uint64_t)(2 * j + 1)), (uint32_t)n);
    }
    w = weighted ? draw_weight(seed, i, wmin, wmax) : 1;
    return true;
  };
  return build_csr(n, grid_edges + extra, true, true, weighted != 0, pair, out);
}

int dawn_gen_sources(int32_t n, int32_t count, uint64_t seed, int32_t* out) {
  if (n <= 0 || count < 0 || (count > 0 && !out)) return DAWN_ERR_INVALID_ARG;
  for (int32_t i = 0; i < count; i++)
    out[i] = (int32_t)bounded32(rng_u64(seed, kStreamSources, (uint64_t)i), (uint32_t)n);
  return DAWN_OK;
}

int dawn_csr_from_edges(int32_t n, int64_t count, const int32_t* u,
                        const int32_t* v, const int32_t* w, int32_t symmetrize,
                        int32_t dedup, dawn_csr* out) {
  if (!out || n < 0 || count < 0 || (count > 0 && (!u || !v)))
    return DAWN_ERR_INVALID_ARG;
  for (int64_t i = 0; i < count; i++) {
    if (u[i] < 0 || u[i] >= n || v[i] < 0 || v[i] >= n) return DAWN_ERR_INVALID_ARG;
    if (w && w[i] < 0) return DAWN_ERR_INVALID_ARG;
  }
  auto pair = [&](int64_t i, int32_t& a, int32_t& b, int32_t& c) {
    a = u[i];
    b = v[i];
    c = w ? w[i] : 1;
    return true;
  };
  return build_csr(n, count, symmetrize != 0, dedup != 0, w != nullptr, pair, out);
}

int dawn_read_mtx(const char* path, int32_t weighted, dawn_csr* out) {
  return dawn_read_mtx_as(path, weighted, -1, out);
}

int dawn_read_mtx_as(const char* path, int32_t weighted, int32_t as_directed, dawn_csr* out) {
  if (!path || !out || as_directed < -1 || as_directed > 1) return DAWN_ERR_INVALID_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) return DAWN_ERR_IO;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<char> buf((size_t)std::max<long>(sz, 0) + 1);
  size_t got = sz > 0 ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
  std::fclose(f);
  buf[got] = '\0';
  char* p = buf.data();
  char* end = p + got;

  // Header: %%MatrixMarket matrix coordinate <field> <symmetry>
  char* line = next_line(p, end);
  if (!line) return DAWN_ERR_FORMAT;
  char fmt[64] = {0}, obj[64] = {0}, mtype[64] = {0}, dtype[64] = {0}, sym[64] = {0};
  int nf = std::sscanf(line, "%63s %63s %63s %63s %63s", fmt, obj, mtype, dtype, sym);
  if (nf < 3 || std::strcmp(fmt, "%%MatrixMarket") != 0 || std::strcmp(mtype, "coordinate") != 0)
    return DAWN_ERR_FORMAT;  // reference prints "invalid file" (src/graph.cpp:45-48)
  const bool pattern = std::strcmp(dtype, "pattern") == 0;
  // src/graph.cpp:40-44, unless the caller names the reader (src/io.cpp:2-192).
  const bool directed = as_directed >= 0 ? as_directed == 1 : std::strcmp(sym, "symmetric") != 0;
  // A pattern file read as weighted uses weight 1 (the reference reads an
  // unset value there, src/io.cpp:107).
  const bool use_w = weighted != 0;

  long long rows = -1, cols = -1, nnz = -1;
  while ((line = next_line(p, end))) {
    if (line[0] == '%' || line[0] == '\0' || line[0] == '\r') continue;
    if (std::sscanf(line, "%lld %lld %lld", &rows, &cols, &nnz) < 2) return DAWN_ERR_FORMAT;
    break;
  }
  if (rows <= 0 || rows > INT32_MAX) return DAWN_ERR_FORMAT;
  const int32_t n = (int32_t)rows;

  std::vector<MtxEdge> edges;
  if (nnz > 0) edges.reserve((size_t)nnz * (directed ? 1 : 2));
  while ((line = next_line(p, end))) {
    if (line[0] == '%' || line[0] == '\0' || line[0] == '\r') continue;
    char* q = line;
    char* e1;
    long long r = std::strtoll(q, &e1, 10);
    if (e1 == q) continue;
    char* e2;
    long long c = std::strtoll(e1, &e2, 10);
    if (e2 == e1) return DAWN_ERR_FORMAT;
    double w = 1.0;
    if (!pattern) {
      char* e3;
      double x = std::strtod(e2, &e3);
      if (e3 != e2) w = x;
    }
    r -= 1;
    c -= 1;
    if (r < 0 || r >= n || c < 0 || c >= n) return DAWN_ERR_FORMAT;
    if (r == c) continue;  // self-loops dropped (src/io.cpp:19,71,111,172)
    if (use_w && (w < 0 || w > (double)INT32_MAX || std::floor(w) != w)) return DAWN_ERR_FORMAT;
    edges.push_back({(int32_t)r, (int32_t)c, w});
    if (!directed) edges.push_back({(int32_t)c, (int32_t)r, w});
  }
  if ((int64_t)edges.size() > (int64_t)INT32_MAX) return DAWN_ERR_OVERFLOW;
  // Ascending (row, col, weight) with duplicates kept: the order the
  // reference's priority_queue drain / std::sort produces (src/io.cpp:31-38,
  // src/matrix.cpp:3-26).
  std::sort(edges.begin(), edges.end(), [](const MtxEdge& x, const MtxEdge& y) {
    if (x.r != y.r) return x.r < y.r;
    if (x.c != y.c) return x.c < y.c;
    return x.w < y.w;
  });
  const int64_t m = (int64_t)edges.size();
  int32_t* rp = (int32_t*)std::calloc((size_t)n + 1, sizeof(int32_t));
  int32_t* cc = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
  int32_t* ww = use_w ? (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1)) : nullptr;
  if (!rp || !cc || (use_w && !ww)) {
    std::free(rp); std::free(cc); std::free(ww);
    return DAWN_ERR_OUT_OF_MEMORY;
  }
  for (int64_t i = 0; i < m; i++) rp[edges[i].r + 1]++;
  for (int32_t i = 0; i < n; i++) rp[i + 1] += rp[i];
  for (int64_t i = 0; i < m; i++) {
    cc[i] = edges[i].c;
    if (use_w) ww[i] = (int32_t)edges[i].w;
  }
  out->n = n;
  out->m = m;
  out->directed = directed ? 1 : 0;
  out->weighted = use_w ? 1 : 0;
  out->row_ptr = rp;
  out->col = cc;
  out->weight = ww;
  return DAWN_OK;
}

int dawn_write_mtx(const char* path, const dawn_csr* csr) {
  if (!path || !csr || !csr->row_ptr) return DAWN_ERR_INVALID_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return DAWN_ERR_IO;
  const bool w = csr->weight != nullptr;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate %s %s\n",
               w ? "integer" : "pattern", csr->directed ? "general" : "symmetric");
  // Symmetric files store each undirected edge once (lower triangle).
  int64_t cnt = 0;
  for (int32_t r = 0; r < csr->n; r++)
    for (int32_t k = csr->row_ptr[r]; k < csr->row_ptr[r + 1]; k++)
      if (csr->directed || csr->col[k] < r) cnt++;
  std::fprintf(f, "%d %d %lld\n", csr->n, csr->n, (long long)cnt);
  for (int32_t r = 0; r < csr->n; r++)
    for (int32_t k = csr->row_ptr[r]; k < csr->row_ptr[r + 1]; k++) {
      if (!csr->directed && csr->col[k] >= r) continue;
      if (w)
        std::fprintf(f, "%d %d %d\n", r + 1, csr->col[k] + 1, csr->weight[k]);
      else
        std::fprintf(f, "%d %d\n", r + 1, csr->col[k] + 1);
    }
  std::fclose(f);
  return DAWN_OK;
}

int dawn_read_list(const char* path, int32_t** out, int64_t* count) {
  if (!path || !out || !count) return DAWN_ERR_INVALID_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return DAWN_ERR_IO;
  std::vector<int32_t> v;
  char line[256];
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '%') continue;
    char* e;
    errno = 0;
    long long x = std::strtoll(line, &e, 10);
    if (e == line) continue;
    // std::stringstream >> int saturates on overflow; keep that value.
    if (x > INT32_MAX) x = INT32_MAX;
    if (x < INT32_MIN) x = INT32_MIN;
    v.push_back((int32_t)x);
  }
  std::fclose(f);
  int32_t* buf = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(v.size(), 1));
  if (!buf) return DAWN_ERR_OUT_OF_MEMORY;
  if (!v.empty()) std::memcpy(buf, v.data(), sizeof(int32_t) * v.size());
  *out = buf;
  *count = (int64_t)v.size();
  return DAWN_OK;
}

int dawn_write_dist(const char* path, int32_t n, const int32_t* dist, int32_t source) {
  if (!path || !dist || n < 0) return DAWN_ERR_INVALID_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return DAWN_ERR_IO;
  for (int32_t j = 0; j < n; j++)
    if (j != source && dist[j] > 0 && dist[j] != DAWN_INF)
      std::fprintf(f, "%d %d %d\n", source, j, dist[j]);
  std::fclose(f);
  return DAWN_OK;
}

}  // extern "C"
