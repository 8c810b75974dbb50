/*
 * dawn_oracle.h — CPU oracle for the DAWN frontier engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's CPU DAWN (lxrzlyr/GAL-DAWN, the .cpp files under src/algorithm/cpu) and of the
 * GAPBS direction-optimizing rule (test/gabps/bfs.cc).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker / CPU baseline — never as a product code path.
 *
 * Conventions: int32 CSR (row_ptr[n+1], col[m], optional int32 weight[m]).
 * "ref" distance arrays use the reference's encoding (0 = unvisited for BFS);
 * "inf" arrays use the engine's encoding (INT32_MAX = unreachable).
 */
#ifndef DAWN_ORACLE_H
#define DAWN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_INF 2147483647

/* BFS_CPU::BFS + SOVMP (src/algorithm/cpu/bfs.cpp:23-77, 159-185).
 * Dense bool frontier; output in reference encoding (0 for source and for
 * unreached).  *steps = value of `step` when the loop ended. */
int oracle_bfs_dense(int32_t n, const int32_t* row_ptr, const int32_t* col,
                     int32_t source, int32_t* dist_ref, int32_t* steps);

/* BFS_CPU::BFS_kernel + SOVM (src/algorithm/cpu/bfs.cpp:79-157).
 * Sparse queue frontier; reference encoding. */
int oracle_bfs_sparse(int32_t n, const int32_t* row_ptr, const int32_t* col,
                      int32_t source, int32_t* dist_ref, int32_t* steps);

/* Reference encoding -> engine encoding: 0 -> INF for every v != source. */
void oracle_ref_to_inf(int32_t n, int32_t* dist, int32_t source);

/* SSSP_CPU::SSSP + GOVMP (src/algorithm/cpu/sssp.cpp:23-80, 169-199) with
 * int32 weights and sums (the reference uses float), run sequentially
 * (the reference's OpenMP min is unsynchronised).  INF = INT32_MAX.
 * Returns -8 if a path sum overflows int32. */
int oracle_sssp_dense(int32_t n, const int32_t* row_ptr, const int32_t* col,
                      const int32_t* w, int32_t source, int32_t* dist,
                      int32_t* steps);

/* SSSP_CPU::SSSP_kernel + GOVM (src/algorithm/cpu/sssp.cpp:82-167) with the
 * frontier vertex's own distance (the reference reads distance[j], a queue
 * index, at sssp.cpp:157) and a de-duplicated queue. */
int oracle_sssp_sparse(int32_t n, const int32_t* row_ptr, const int32_t* col,
                       const int32_t* w, int32_t source, int32_t* dist,
                       int32_t* steps);

/* Textbook binary-heap Dijkstra (independent cross-check; not a reference
 * restatement).  INF = INT32_MAX. */
int oracle_dijkstra(int32_t n, const int32_t* row_ptr, const int32_t* col,
                    const int32_t* w, int32_t source, int32_t* dist);

/* MSSP_CPU::run (src/algorithm/cpu/mssp.cpp:9-32): OpenMP over sources,
 * source = sources[i] mod n, isolated sources skipped (skipped[i] = 1, row
 * INF except dist[src] = 0, sum 0).  dist: S*n (engine encoding) or NULL.
 * sums: S int64 sums over reached vertices or NULL.  threads <= 0: default.
 * Returns wall seconds of the traversal in *seconds if non-NULL. */
int oracle_mssp(int32_t n, const int32_t* row_ptr, const int32_t* col,
                const int32_t* w, const int32_t* sources, int32_t S,
                int32_t* dist, int64_t* sums, int32_t* skipped,
                int32_t threads, double* seconds);

/* Dense single-source BFS timed the way BFS_CPU::BFS times it (wall clock
 * around the step loop), with OpenMP inside SOVMP.  Returns seconds. */
double oracle_time_bfs_dense(int32_t n, const int32_t* row_ptr,
                             const int32_t* col, int32_t source,
                             int32_t threads, int32_t* dist_ref);

/* Closeness from an engine-encoded distance row (src/algorithm/cpu/cc.cpp:
 * 76-78): *sum = exact int64 sum over reached vertices; returns
 * (n-1)/sum in double (0 if sum == 0).  *ref_int = the reference's value:
 * (row-1) / int32 accumulate with integer division, as float. */
double oracle_closeness(int32_t n, const int32_t* dist, int64_t* sum,
                        float* ref_int);

/* GAPBS DOBFS push/pull schedule (test/gabps/bfs.cc:122-182, alpha=15,
 * beta=18) for a single-source BFS: modes[i] (0 = push/top-down,
 * 1 = pull/bottom-up) for the step that discovers level i+1.  Returns the
 * number of steps written (<= max). */
int oracle_dobfs_modes(int32_t n, const int32_t* row_ptr, const int32_t* col,
                       int32_t source, int32_t alpha, int32_t beta,
                       int32_t* modes, int32_t max);

/* Shortest-path certificate (the optimality conditions; no reference
 * counterpart): returns 1 iff d (engine encoding) is the distance array from
 * source over the out-edges with positive weights w (NULL: unit weights),
 * 0 if not, -1 for invalid arguments or a weight <= 0.  Paths of weight >=
 * INT32_MAX count as absent.  O(m), OpenMP (threads <= 0: default). */
int oracle_certify(int32_t n, const int32_t* row_ptr, const int32_t* col,
                   const int32_t* w, int32_t source, const int32_t* d,
                   int32_t threads);

/* 8->32->32->3 ReLU MLP in double precision.  params layout:
 * W1[32][8] b1[32] W2[32][32] b2[32] W3[3][32] b3[3] (1443 values).
 * loss = mean over N*3 of (out - y)^2. */
void oracle_mlp_forward(const double* params, const double* x, int32_t N,
                        double* out);
double oracle_mlp_loss_grad(const double* params, const double* x,
                            const double* y, int32_t N, double* grad);

#ifdef __cplusplus
}
#endif
#endif
