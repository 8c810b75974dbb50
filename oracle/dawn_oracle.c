/*
 * dawn_oracle.c — CPU oracle for the DAWN frontier engine.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker and the CPU baseline; never
 * linked into or called by the product library.
 *
 * Each function restates the reference's CPU algorithm (file:line cited at
 * the function) in plain C, sequentially and deterministically (the
 * reference's OpenMP steps have benign or lossy races; see the comments).
 */
#include "dawn_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* SOVMP, src/algorithm/cpu/bfs.cpp:159-185: for every frontier vertex j,
 * every out-neighbour with distance 0 ("unvisited") gets distance = step and
 * joins the next frontier; j leaves the frontier.  Returns 1 if nothing was
 * discovered (converged).  The reference runs the j loop under OpenMP with
 * same-value writes; the sequential order gives the same result. */
static int sovmp(int32_t n, const int32_t* row_ptr, const int32_t* col,
                 unsigned char* alpha, unsigned char* beta, int32_t* dist,
                 int32_t step) {
  int converged = 1;
  for (int32_t j = 0; j < n; j++) {
    if (!alpha[j]) continue;
    for (int32_t k = row_ptr[j]; k < row_ptr[j + 1]; k++) {
      int32_t v = col[k];
      if (!dist[v]) {
        dist[v] = step;
        beta[v] = 1;
        converged = 0;
      }
    }
    alpha[j] = 0;
  }
  return converged;
}

/* BFS_CPU::BFS, src/algorithm/cpu/bfs.cpp:23-77.  The source's neighbours
 * start at distance 1 and in the frontier (:40-44); steps 2,3,... alternate
 * the two bool arrays (:47-58); the source itself reads as unvisited (0) and
 * is re-discovered at step 2, which is harmless; distance[source] is reset
 * to 0 at the end (:62). */
int oracle_bfs_dense(int32_t n, const int32_t* row_ptr, const int32_t* col,
                     int32_t source, int32_t* dist, int32_t* steps) {
  if (n <= 0 || source < 0 || source >= n) return -1;
  if (row_ptr[source] == row_ptr[source + 1]) return -2; /* bfs.cpp:12-15 */
  unsigned char* alpha = (unsigned char*)calloc((size_t)n, 1);
  unsigned char* beta = (unsigned char*)calloc((size_t)n, 1);
  if (!alpha || !beta) { free(alpha); free(beta); return -5; }
  memset(dist, 0, sizeof(int32_t) * (size_t)n);
  for (int32_t i = row_ptr[source]; i < row_ptr[source + 1]; i++) {
    alpha[col[i]] = 1;
    dist[col[i]] = 1;
  }
  int32_t step = 1;
  while (step < n) {
    step++;
    int conv = (step % 2 == 0) ? sovmp(n, row_ptr, col, alpha, beta, dist, step)
                               : sovmp(n, row_ptr, col, beta, alpha, dist, step);
    if (conv) break;
  }
  dist[source] = 0;
  if (steps) *steps = step;
  free(alpha);
  free(beta);
  return 0;
}

/* SOVM, src/algorithm/cpu/bfs.cpp:134-157: queue frontier. */
static int32_t sovm(const int32_t* row_ptr, const int32_t* col,
                    const int32_t* alpha, int32_t* beta, int32_t* dist,
                    int32_t step, int32_t entry) {
  int32_t out = 0;
  for (int32_t j = 0; j < entry; j++) {
    int32_t u = alpha[j];
    for (int32_t k = row_ptr[u]; k < row_ptr[u + 1]; k++) {
      int32_t v = col[k];
      if (!dist[v]) {
        dist[v] = step;
        beta[out++] = v;
      }
    }
  }
  return out;
}

/* BFS_CPU::BFS_kernel, src/algorithm/cpu/bfs.cpp:79-132. */
static int bfs_sparse_ws(int32_t n, const int32_t* row_ptr, const int32_t* col,
                         int32_t source, int32_t* dist, int32_t* steps,
                         int32_t* qa, int32_t* qb) {
  memset(dist, 0, sizeof(int32_t) * (size_t)n);
  int32_t entry = 0;
  /* The reference appends every listed neighbour, duplicates included
   * (bfs.cpp:96-99); expanding a vertex twice changes nothing, so the queue
   * keeps one copy. */
  for (int32_t i = row_ptr[source]; i < row_ptr[source + 1]; i++) {
    if (dist[col[i]] == 0) qa[entry++] = col[i];
    dist[col[i]] = 1;
  }
  int32_t step = 1;
  while (step < n) {
    step++;
    if (step % 2 == 0)
      entry = sovm(row_ptr, col, qa, qb, dist, step, entry);
    else
      entry = sovm(row_ptr, col, qb, qa, dist, step, entry);
    if (!entry) break;
  }
  dist[source] = 0; /* bfs.cpp:117 */
  if (steps) *steps = step;
  return 0;
}

int oracle_bfs_sparse(int32_t n, const int32_t* row_ptr, const int32_t* col,
                      int32_t source, int32_t* dist, int32_t* steps) {
  if (n <= 0 || source < 0 || source >= n) return -1;
  if (row_ptr[source] == row_ptr[source + 1]) return -2;
  int32_t* qa = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* qb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!qa || !qb) { free(qa); free(qb); return -5; }
  int rc = bfs_sparse_ws(n, row_ptr, col, source, dist, steps, qa, qb);
  free(qa);
  free(qb);
  return rc;
}

void oracle_ref_to_inf(int32_t n, int32_t* dist, int32_t source) {
  for (int32_t v = 0; v < n; v++)
    if (v != source && dist[v] == 0) dist[v] = ORACLE_INF;
}

/* GOVMP, src/algorithm/cpu/sssp.cpp:169-199, int32 weights: for every
 * frontier vertex j and out-edge k, tmp = distance[j] + w[k]; if it improves
 * distance[col[k]] (and the target is not the source) the target takes it
 * and joins the next frontier.  Reads distances updated earlier in the same
 * step (Gauss-Seidel order), like the reference.  Sequential: the
 * reference's unsynchronised OpenMP min can lose updates (sssp.cpp:186-190). */
static int govmp(int32_t n, const int32_t* row_ptr, const int32_t* col,
                 const int32_t* w, unsigned char* alpha, unsigned char* beta,
                 int32_t* dist, int32_t source, int* overflow) {
  int converged = 1;
  for (int32_t j = 0; j < n; j++) {
    if (!alpha[j]) continue;
    for (int32_t k = row_ptr[j]; k < row_ptr[j + 1]; k++) {
      int32_t v = col[k];
      int64_t tmp = (int64_t)dist[j] + (int64_t)w[k];
      if (tmp >= ORACLE_INF) { *overflow = 1; continue; }
      if ((int64_t)dist[v] > tmp && v != source) {
        dist[v] = (int32_t)tmp;
        beta[v] = 1;
        converged = 0;
      }
    }
    alpha[j] = 0;
  }
  return converged;
}

/* SSSP_CPU::SSSP, src/algorithm/cpu/sssp.cpp:23-80: distances start at INF,
 * the source's neighbours take the listed edge weight in row order (a later
 * duplicate overwrites an earlier one, :41-44), distance[source] = 0 (:46),
 * then GOVMP alternates the frontier arrays until converged. */
int oracle_sssp_dense(int32_t n, const int32_t* row_ptr, const int32_t* col,
                      const int32_t* w, int32_t source, int32_t* dist,
                      int32_t* steps) {
  if (n <= 0 || source < 0 || source >= n || !w) return -1;
  if (row_ptr[source] == row_ptr[source + 1]) return -2; /* sssp.cpp:12-15 */
  unsigned char* alpha = (unsigned char*)calloc((size_t)n, 1);
  unsigned char* beta = (unsigned char*)calloc((size_t)n, 1);
  if (!alpha || !beta) { free(alpha); free(beta); return -5; }
  for (int32_t v = 0; v < n; v++) dist[v] = ORACLE_INF;
  for (int32_t i = row_ptr[source]; i < row_ptr[source + 1]; i++) {
    dist[col[i]] = w[i];
    alpha[col[i]] = 1;
  }
  dist[source] = 0;
  int overflow = 0;
  int32_t step = 1;
  while (step < n) {
    step++;
    int conv = (step % 2 == 0)
                   ? govmp(n, row_ptr, col, w, alpha, beta, dist, source, &overflow)
                   : govmp(n, row_ptr, col, w, beta, alpha, dist, source, &overflow);
    if (conv) break;
  }
  if (steps) *steps = step;
  free(alpha);
  free(beta);
  return overflow ? -8 : 0;
}

/* GOVM + SSSP_kernel, src/algorithm/cpu/sssp.cpp:82-167, with the frontier
 * vertex's distance distance[alpha[j]] (the reference reads distance[j]) and
 * one queue entry per vertex per step. */
static int sssp_sparse_ws(int32_t n, const int32_t* row_ptr, const int32_t* col,
                          const int32_t* w, int32_t source, int32_t* dist,
                          int32_t* steps, int32_t* qa, int32_t* qb,
                          int32_t* mark) {
  for (int32_t v = 0; v < n; v++) { dist[v] = ORACLE_INF; mark[v] = 0; }
  int32_t entry = 0;
  for (int32_t i = row_ptr[source]; i < row_ptr[source + 1]; i++) {
    int32_t v = col[i];
    dist[v] = w[i];
    if (mark[v] != 1) { mark[v] = 1; qa[entry++] = v; }
  }
  dist[source] = 0;
  int overflow = 0;
  int32_t step = 1;
  int32_t* in = qa;
  int32_t* out = qb;
  while (step < n) {
    step++;
    int32_t cnt = 0;
    for (int32_t j = 0; j < entry; j++) {
      int32_t u = in[j];
      for (int32_t k = row_ptr[u]; k < row_ptr[u + 1]; k++) {
        int32_t v = col[k];
        int64_t tmp = (int64_t)dist[u] + (int64_t)w[k];
        if (tmp >= ORACLE_INF) { overflow = 1; continue; }
        if ((int64_t)dist[v] > tmp && v != source) {
          dist[v] = (int32_t)tmp;
          if (mark[v] != step) { mark[v] = step; out[cnt++] = v; }
        }
      }
    }
    entry = cnt;
    int32_t* t = in; in = out; out = t;
    if (!entry) break;
  }
  if (steps) *steps = step;
  return overflow ? -8 : 0;
}

int oracle_sssp_sparse(int32_t n, const int32_t* row_ptr, const int32_t* col,
                       const int32_t* w, int32_t source, int32_t* dist,
                       int32_t* steps) {
  if (n <= 0 || source < 0 || source >= n || !w) return -1;
  if (row_ptr[source] == row_ptr[source + 1]) return -2;
  int32_t* qa = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* qb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* mk = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!qa || !qb || !mk) { free(qa); free(qb); free(mk); return -5; }
  int rc = sssp_sparse_ws(n, row_ptr, col, w, source, dist, steps, qa, qb, mk);
  free(qa); free(qb); free(mk);
  return rc;
}

/* ---- textbook Dijkstra (binary heap of (dist, vertex)) ------------------ */
typedef struct { int64_t d; int32_t v; } hitem;

static void heap_push(hitem* h, int64_t* sz, hitem x) {
  int64_t i = (*sz)++;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h[p].d <= x.d) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = x;
}

static hitem heap_pop(hitem* h, int64_t* sz) {
  hitem top = h[0];
  hitem x = h[--(*sz)];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1;
    if (l >= *sz) break;
    int64_t c = (l + 1 < *sz && h[l + 1].d < h[l].d) ? l + 1 : l;
    if (h[c].d >= x.d) break;
    h[i] = h[c];
    i = c;
  }
  if (*sz > 0) h[i] = x;
  return top;
}

int oracle_dijkstra(int32_t n, const int32_t* row_ptr, const int32_t* col,
                    const int32_t* w, int32_t source, int32_t* dist) {
  if (n <= 0 || source < 0 || source >= n) return -1;
  int64_t m = row_ptr[n];
  hitem* h = (hitem*)malloc(sizeof(hitem) * (size_t)(m + 2));
  int64_t* d = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  if (!h || !d) { free(h); free(d); return -5; }
  for (int32_t v = 0; v < n; v++) d[v] = INT64_MAX;
  int64_t sz = 0;
  d[source] = 0;
  hitem s0 = {0, source};
  heap_push(h, &sz, s0);
  while (sz > 0) {
    hitem it = heap_pop(h, &sz);
    if (it.d != d[it.v]) continue;
    for (int32_t k = row_ptr[it.v]; k < row_ptr[it.v + 1]; k++) {
      int64_t nd = it.d + (w ? w[k] : 1);
      int32_t v = col[k];
      if (nd < d[v]) {
        d[v] = nd;
        hitem x = {nd, v};
        heap_push(h, &sz, x);
      }
    }
  }
  int rc = 0;
  for (int32_t v = 0; v < n; v++) {
    if (d[v] == INT64_MAX) dist[v] = ORACLE_INF;
    else if (d[v] >= ORACLE_INF) { dist[v] = ORACLE_INF; rc = -8; }
    else dist[v] = (int32_t)d[v];
  }
  free(h);
  free(d);
  return rc;
}

/* MSSP_CPU::run, src/algorithm/cpu/mssp.cpp:9-32: OpenMP over the source
 * list, source = msource[i] % rows (:15), isolated sources skipped
 * (:16-18), each thread a sequential sparse traversal (BFS_kernel or
 * SSSP_kernel).  Distances are kept here (the reference discards them). */
int oracle_mssp(int32_t n, const int32_t* row_ptr, const int32_t* col,
                const int32_t* w, const int32_t* sources, int32_t S,
                int32_t* dist, int64_t* sums, int32_t* skipped,
                int32_t threads, double* seconds) {
  if (n <= 0 || S < 0 || (S > 0 && !sources)) return -1;
  int nth = threads > 0 ? threads : omp_get_max_threads();
  int rc_all = 0;
  double t0 = now_s();
#pragma omp parallel num_threads(nth)
  {
    int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* qa = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* qb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* mk = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
#pragma omp for schedule(dynamic, 1)
    for (int32_t i = 0; i < S; i++) {
      int32_t s = (int32_t)(((int64_t)sources[i] % n + n) % n);
      int32_t* row = dist ? dist + (size_t)i * (size_t)n : d;
      int rc = 0;
      if (row_ptr[s] == row_ptr[s + 1]) {
        for (int32_t v = 0; v < n; v++) row[v] = ORACLE_INF;
        row[s] = 0;
        if (skipped) skipped[i] = 1;
        if (sums) sums[i] = 0;
        continue;
      }
      if (skipped) skipped[i] = 0;
      if (w) {
        rc = sssp_sparse_ws(n, row_ptr, col, w, s, row, NULL, qa, qb, mk);
      } else {
        rc = bfs_sparse_ws(n, row_ptr, col, s, row, NULL, qa, qb);
        oracle_ref_to_inf(n, row, s);
      }
      if (rc) {
#pragma omp atomic write
        rc_all = rc;
      }
      if (sums) {
        int64_t acc = 0;
        for (int32_t v = 0; v < n; v++)
          if (row[v] != ORACLE_INF) acc += row[v];
        sums[i] = acc;
      }
    }
    free(d); free(qa); free(qb); free(mk);
  }
  if (seconds) *seconds = now_s() - t0;
  return rc_all;
}

double oracle_time_bfs_dense(int32_t n, const int32_t* row_ptr,
                             const int32_t* col, int32_t source,
                             int32_t threads, int32_t* dist) {
  /* Same loop as oracle_bfs_dense with SOVMP's j loop under OpenMP, as the
   * reference runs it (bfs.cpp:167); distance writes are same-value. */
  if (row_ptr[source] == row_ptr[source + 1]) return -1.0;
  int nth = threads > 0 ? threads : omp_get_max_threads();
  unsigned char* a = (unsigned char*)calloc((size_t)n, 1);
  unsigned char* b = (unsigned char*)calloc((size_t)n, 1);
  memset(dist, 0, sizeof(int32_t) * (size_t)n);
  for (int32_t i = row_ptr[source]; i < row_ptr[source + 1]; i++) {
    a[col[i]] = 1;
    dist[col[i]] = 1;
  }
  double t0 = now_s();
  int32_t step = 1;
  while (step < n) {
    step++;
    unsigned char* in = (step % 2 == 0) ? a : b;
    unsigned char* out = (step % 2 == 0) ? b : a;
    int converged = 1;
#pragma omp parallel for num_threads(nth) schedule(static) reduction(&& : converged)
    for (int32_t j = 0; j < n; j++) {
      if (!in[j]) continue;
      for (int32_t k = row_ptr[j]; k < row_ptr[j + 1]; k++) {
        int32_t v = col[k];
        int32_t cur;
#pragma omp atomic read
        cur = dist[v];
        if (!cur) {
#pragma omp atomic write
          dist[v] = step;
          out[v] = 1;
          converged = 0;
        }
      }
      in[j] = 0;
    }
    if (converged) break;
  }
  double t = now_s() - t0;
  dist[source] = 0;
  free(a);
  free(b);
  return t;
}

/* CC_CPU::kernel, src/algorithm/cpu/cc.cpp:76-78: closeness from the
 * distance row.  The reference divides (row-1) by an int32 accumulate with
 * integer division; the engine's value is (n-1)/sum with sum exact in int64
 * and one double division. */
double oracle_closeness(int32_t n, const int32_t* dist, int64_t* sum,
                        float* ref_int) {
  int64_t acc = 0;
  uint32_t acc32 = 0; /* int32 accumulate with wrap-around */
  for (int32_t v = 0; v < n; v++) {
    if (dist[v] == ORACLE_INF) continue; /* reference encodes these as 0 */
    acc += dist[v];
    acc32 += (uint32_t)dist[v];
  }
  if (sum) *sum = acc;
  if (ref_int) {
    int32_t a = (int32_t)acc32;
    *ref_int = a != 0 ? (float)((n - 1) / a) : 0.0f;
  }
  return acc > 0 ? (double)(n - 1) / (double)acc : 0.0;
}

/* GAPBS DOBFS, test/gabps/bfs.cc:122-182.  Only the push/pull decisions are
 * restated; both steps discover exactly the next BFS level, so the schedule
 * depends on level sizes and degree sums alone.  scout_count adds
 * max(out_degree, 1) per discovered vertex (InitParent stores -1 for
 * degree-0 vertices, bfs.cc:114-120, and TDStep adds -parent, :79-82). */
int oracle_dobfs_modes(int32_t n, const int32_t* row_ptr, const int32_t* col,
                       int32_t source, int32_t alpha, int32_t beta,
                       int32_t* modes, int32_t max) {
  int32_t* level = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!level || !q) { free(level); free(q); return -5; }
  for (int32_t v = 0; v < n; v++) level[v] = -1;
  /* Level-synchronous BFS; lstart[l] = start of level l in q. */
  int32_t* lstart = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 2));
  int32_t head = 0, tail = 0, nl = 0;
  level[source] = 0;
  q[tail++] = source;
  lstart[0] = 0;
  while (head < tail) {
    int32_t lend = tail;
    lstart[++nl] = lend;
    for (; head < lend; head++) {
      int32_t u = q[head];
      for (int32_t k = row_ptr[u]; k < row_ptr[u + 1]; k++) {
        int32_t v = col[k];
        if (level[v] < 0) { level[v] = level[u] + 1; q[tail++] = v; }
      }
    }
  }
  /* levels 0..nl-1 non-empty; level nl is empty. */
#define LSIZE(l) ((l) < nl ? (int64_t)(lstart[(l) + 1] - lstart[(l)]) : (int64_t)0)
  int64_t edges_to_check = row_ptr[n];
  int64_t scout = row_ptr[source + 1] - row_ptr[source];
  int32_t cur = 0; /* current frontier = level cur */
  int32_t out = 0;
  while (cur < nl) { /* queue not empty */
    if (scout > edges_to_check / alpha) {
      int64_t awake = LSIZE(cur), old;
      do {
        old = awake;
        if (out < max) modes[out] = 1;
        out++;
        cur++;
        awake = LSIZE(cur);
      } while (awake >= old || awake > (int64_t)n / beta);
      scout = 1;
    } else {
      edges_to_check -= scout;
      if (out < max) modes[out] = 0;
      out++;
      cur++;
      int64_t sc = 0;
      if (cur < nl) {
        for (int32_t i = lstart[cur]; i < lstart[cur + 1]; i++) {
          int32_t v = q[i];
          int64_t d = row_ptr[v + 1] - row_ptr[v];
          sc += d != 0 ? d : 1;
        }
      }
      scout = sc;
    }
  }
#undef LSIZE
  free(level);
  free(q);
  free(lstart);
  return out < max ? out : max;
}

/* ---- MLP 8 -> 32 -> 32 -> 3 (ReLU), double precision -------------------- */
#define IN 8
#define H 32
#define OUT 3
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + H * IN)
#define OFF_W2 (OFF_B1 + H)
#define OFF_B2 (OFF_W2 + H * H)
#define OFF_W3 (OFF_B2 + H)
#define OFF_B3 (OFF_W3 + OUT * H)

static void mlp_row(const double* p, const double* x, double* h1, double* h2,
                    double* o) {
  for (int i = 0; i < H; i++) {
    double a = p[OFF_B1 + i];
    for (int j = 0; j < IN; j++) a += p[OFF_W1 + i * IN + j] * x[j];
    h1[i] = a > 0 ? a : 0;
  }
  for (int i = 0; i < H; i++) {
    double a = p[OFF_B2 + i];
    for (int j = 0; j < H; j++) a += p[OFF_W2 + i * H + j] * h1[j];
    h2[i] = a > 0 ? a : 0;
  }
  for (int i = 0; i < OUT; i++) {
    double a = p[OFF_B3 + i];
    for (int j = 0; j < H; j++) a += p[OFF_W3 + i * H + j] * h2[j];
    o[i] = a;
  }
}

void oracle_mlp_forward(const double* params, const double* x, int32_t N,
                        double* out) {
  double h1[H], h2[H];
  for (int32_t r = 0; r < N; r++) mlp_row(params, x + r * IN, h1, h2, out + r * OUT);
}

double oracle_mlp_loss_grad(const double* p, const double* x, const double* y,
                            int32_t N, double* g) {
  memset(g, 0, sizeof(double) * (OFF_B3 + OUT));
  double loss = 0;
  double h1[H], h2[H], o[OUT], d3[OUT], d2[H], d1[H];
  const double scale = 1.0 / ((double)N * OUT);
  for (int32_t r = 0; r < N; r++) {
    const double* xr = x + r * IN;
    mlp_row(p, xr, h1, h2, o);
    for (int i = 0; i < OUT; i++) {
      double e = o[i] - y[r * OUT + i];
      loss += e * e * scale;
      d3[i] = 2.0 * e * scale;
    }
    for (int i = 0; i < OUT; i++) {
      g[OFF_B3 + i] += d3[i];
      for (int j = 0; j < H; j++) g[OFF_W3 + i * H + j] += d3[i] * h2[j];
    }
    for (int j = 0; j < H; j++) {
      double a = 0;
      for (int i = 0; i < OUT; i++) a += p[OFF_W3 + i * H + j] * d3[i];
      d2[j] = h2[j] > 0 ? a : 0;
    }
    for (int i = 0; i < H; i++) {
      g[OFF_B2 + i] += d2[i];
      for (int j = 0; j < H; j++) g[OFF_W2 + i * H + j] += d2[i] * h1[j];
    }
    for (int j = 0; j < H; j++) {
      double a = 0;
      for (int i = 0; i < H; i++) a += p[OFF_W2 + i * H + j] * d2[i];
      d1[j] = h1[j] > 0 ? a : 0;
    }
    for (int i = 0; i < H; i++) {
      g[OFF_B1 + i] += d1[i];
      for (int j = 0; j < IN; j++) g[OFF_W1 + i * IN + j] += d1[i] * xr[j];
    }
  }
  return loss;
}

/* ---- shortest-path certificate (no reference counterpart) --------------- */

/* The textbook optimality conditions of single-source shortest paths with
 * positive weights (w = 1 when w is NULL).  For reached u (d[u] < INF) and
 * every edge u->v: d[v] <= d[u] + w (a sum at or above INF bounds nothing,
 * as the engine and the restatement above ignore such paths); every reached
 * v != source has an in-edge u->v with d[u] + w == d[v]; d[source] == 0.
 * Together they fix d: the parent edges give a path of weight d[v] (d drops
 * by >= 1 per edge, so the chain ends at the source), and the edge
 * inequality bounds d[v] by every path's weight; an unreached vertex with a
 * reachable in-neighbour breaks the inequality.  O(m), OpenMP over rows. */
int oracle_certify(int32_t n, const int32_t* row_ptr, const int32_t* col,
                   const int32_t* w, int32_t source, const int32_t* d,
                   int32_t threads) {
  if (n <= 0 || source < 0 || source >= n) return -1;
  if (d[source] != 0) return 0;
  unsigned char* parent = (unsigned char*)calloc((size_t)n, 1);
  if (!parent) return -5;
  int nth = threads > 0 ? threads : omp_get_max_threads();
  int ok = 1, bad_weight = 0;
#pragma omp parallel for num_threads(nth) schedule(dynamic, 4096) reduction(&& : ok) reduction(|| : bad_weight)
  for (int32_t u = 0; u < n; u++) {
    if (d[u] < 0) { ok = 0; continue; }
    if (d[u] == ORACLE_INF) continue;
    for (int32_t k = row_ptr[u]; k < row_ptr[u + 1]; k++) {
      const int32_t wk = w ? w[k] : 1;
      if (wk <= 0) { bad_weight = 1; continue; }
      const int32_t v = col[k];
      const int64_t through = (int64_t)d[u] + (int64_t)wk;
      if (through >= ORACLE_INF) continue;
      if ((int64_t)d[v] > through) ok = 0;
      if ((int64_t)d[v] == through) {
#pragma omp atomic write
        parent[v] = 1;
      }
    }
  }
  if (bad_weight) { free(parent); return -1; }
  for (int32_t v = 0; v < n && ok; v++)
    if (v != source && d[v] != ORACLE_INF && !parent[v]) ok = 0;
  free(parent);
  return ok;
}
