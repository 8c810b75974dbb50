/*
 * dawn_capi.h — C ABI of the MI355X-native DAWN frontier engine.
 *
 * Plain C types only (pointers + sizes); no HIP or torch types cross this
 * boundary.  Every compute entry point returns a status code (DAWN_OK == 0)
 * instead of calling exit(); the C++ drop-in layer (dawn_amd.hxx) maps the
 * isolated-source status back to the reference's behaviour.
 *
 * Distance conventions (all entry points):
 *   dist[source] = 0, dist[v] = hop count (BFS) or int32 path weight (SSSP),
 *   unreachable vertices = DAWN_INF (INT32_MAX).
 * The reference writes 0 for unreachable vertices in its unweighted path
 * (src/algorithm/cpu/bfs.cpp:38,62) and FLT_MAX in its weighted path
 * (src/algorithm/cpu/sssp.cpp:40); dawn_write_dist() reproduces its output
 * file format from these arrays.
 */
#ifndef DAWN_CAPI_H
#define DAWN_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DAWN_INF 2147483647

/* Status codes. */
enum {
  DAWN_OK = 0,
  DAWN_ERR_INVALID_ARG = -1,
  DAWN_ERR_ISOLATED_SOURCE = -2, /* source has no out-edges (reference: exit(0)) */
  DAWN_ERR_NO_DEVICE = -3,
  DAWN_ERR_HIP = -4,
  DAWN_ERR_OUT_OF_MEMORY = -5,
  DAWN_ERR_IO = -6,          /* file cannot be opened (reference: cerr + return) */
  DAWN_ERR_FORMAT = -7,      /* not a MatrixMarket coordinate file / bad weight */
  DAWN_ERR_OVERFLOW = -8,    /* sizes beyond int32 CSR offsets */
  DAWN_ERR_UNSUPPORTED = -9,
  DAWN_ERR_NOT_WEIGHTED = -10,
  DAWN_ERR_SCRATCH_LIMIT = -11  /* graph larger than the fixed scratch region */
};

/* Step modes of the frontier engine.  Every mode produces identical
 * distances; the choice changes only time. */
enum {
  DAWN_MODE_PUSH = 0,   /* sparse frontier queue, top-down scatter          */
  DAWN_MODE_PULL = 1,   /* bottom-up gather over unvisited vertices (CSC)   */
  DAWN_MODE_BITMAP = 2  /* dense top-down scan of the frontier bitmap/words */
};
#define DAWN_NUM_MODES 3

/* Scheduling policies. */
enum {
  DAWN_POLICY_FIXED = 0,   /* GAPBS alpha/beta rule (test/gabps/bfs.cc:146,161) */
  DAWN_POLICY_LEARNED = 1, /* FP32 MLP 8->32->32->3 (falls back to FIXED if unset) */
  DAWN_POLICY_PUSH = 2,    /* force one mode for every step */
  DAWN_POLICY_PULL = 3,
  DAWN_POLICY_BITMAP = 4
};

/* Host-side CSR, owned by the library when returned by a generator or
 * reader (free with dawn_csr_free), borrowed when passed in. */
typedef struct dawn_csr {
  int32_t n;          /* vertices (reference Graph_t::rows)               */
  int64_t m;          /* directed edges (reference Graph_t::nnz)          */
  int32_t directed;   /* 0: symmetric adjacency                           */
  int32_t weighted;   /* weight != NULL                                   */
  int32_t* row_ptr;   /* [n+1]                                            */
  int32_t* col;       /* [m]                                              */
  int32_t* weight;    /* [m] or NULL                                      */
} dawn_csr;

/* Per-call statistics. */
typedef struct dawn_stats {
  double device_ms;          /* HIP-event time of the traversal            */
  int64_t steps;             /* frontier steps executed (summed over batches) */
  int64_t mode_steps[DAWN_NUM_MODES];
  int64_t edges_reached;     /* sum over sources of out-degree of reached vertices */
  int64_t vertices_reached;  /* sum over sources of reached vertices (incl. source) */
  int64_t batches;
  int32_t skipped_sources;   /* isolated sources (reference skips them, mssp.cpp:16-18) */
  int32_t reserved;
  int64_t frontier_edges;    /* sum over steps of the out-degrees of the frontier
                                vertices a step expands: the edge relaxations of a
                                push schedule (unweighted batches: per batch, not
                                per lane; 0 where an engine does not count it) */
} dawn_stats;

/* ---------------------------------------------------------------- host --- */

/* Library / device info.  Returns DAWN_OK and the number of visible GPUs. */
int dawn_device_count(int32_t* count);
const char* dawn_status_string(int status);
/* Text of the last HIP failure seen by this thread ("" if none). */
const char* dawn_last_error(void);

/* Seeded synthetic generators (deterministic for any thread count).
 * All produce symmetrized, deduplicated CSR without self-loops; when
 * weighted, weights are uniform integers in [wmin, wmax], equal in both
 * directions, duplicates keep the minimum. */
int dawn_gen_er(int32_t n, double avg_degree, uint64_t seed, int32_t weighted,
                int32_t wmin, int32_t wmax, dawn_csr* out);
int dawn_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                  double c, uint64_t seed, int32_t weighted, int32_t wmin,
                  int32_t wmax, dawn_csr* out);
int dawn_gen_grid(int32_t side, double extra_fraction, uint64_t seed,
                  int32_t weighted, int32_t wmin, int32_t wmax, dawn_csr* out);
/* Uniform sources in [0, n) from a seeded stream. */
int dawn_gen_sources(int32_t n, int32_t count, uint64_t seed, int32_t* out);

/* Build a CSR from an edge list (u[i] -> v[i], optional w[i]).
 * symmetrize: add v->u; dedup: drop repeated (u,v) keeping min weight.
 * Self-loops are always dropped (reference src/io.cpp:19,71,111,172). */
int dawn_csr_from_edges(int32_t n, int64_t count, const int32_t* u,
                        const int32_t* v, const int32_t* w, int32_t symmetrize,
                        int32_t dedup, dawn_csr* out);

/* MatrixMarket reader with the reference's semantics: pattern => unweighted,
 * symmetric => undirected (both directions pushed), 1-based -> 0-based,
 * self-loops dropped, duplicates KEPT, rows sorted by (col, weight).
 * Reference: DAWN::Graph::createGraph (src/graph.cpp:9-77) and
 * DAWN::IO::readGraph* (src/io.cpp:2-192).  Weights must be integral. */
int dawn_read_mtx(const char* path, int32_t weighted, dawn_csr* out);
/* The same reader with the edge semantics chosen by the caller, as the
 * reference's four readers are called directly: as_directed 0 pushes both
 * directions of every listed pair (DAWN::IO::readGraph / readGraph_Weighted,
 * src/io.cpp:2-45, 91-143), 1 keeps them as listed (readGraph_Directed /
 * readGraph_Directed_Weighted, io.cpp:47-89, 145-192), -1 follows the
 * header's symmetry field (dawn_read_mtx). */
int dawn_read_mtx_as(const char* path, int32_t weighted, int32_t as_directed,
                     dawn_csr* out);
/* Write a CSR as a MatrixMarket coordinate file (general/symmetric). */
int dawn_write_mtx(const char* path, const dawn_csr* csr);
void dawn_csr_free(dawn_csr* csr);

/* Source list: one int per line, '%' lines skipped (src/io.cpp:194-213).
 * *out is malloc'd; free with dawn_free. */
int dawn_read_list(const char* path, int32_t** out, int64_t* count);
void dawn_free(void* p);
/* "src dst dist" lines for reached non-source vertices, the format of
 * DAWN::IO::outfile (src/io.cpp:215-231). */
int dawn_write_dist(const char* path, int32_t n, const int32_t* dist,
                    int32_t source);

/* -------------------------------------------------------------- device --- */

typedef struct dawn_graph dawn_graph;

/* Upload a CSR to the current HIP device; builds the in-edge CSC for
 * directed graphs and the degree bins used by the step kernels. */
int dawn_graph_create(const dawn_csr* csr, dawn_graph** out);
int dawn_graph_destroy(dawn_graph* g);
int dawn_graph_info(const dawn_graph* g, int32_t* n, int64_t* m,
                    int32_t* directed, int32_t* weighted);

/* Graph built on the device: the edge list is turned into the CSR (and the
 * in-edge CSC of a directed graph) by GPU kernels -- emission of directed
 * entries, a stable radix sort by (row, column, weight), compaction -- with
 * the same arrays as dawn_csr_from_edges + dawn_graph_create, bit for bit:
 * self-loops dropped, symmetrize adds v->u, dedup keeps the minimum weight,
 * rows sorted by (column, weight).  u, v, w: host or device pointers (w may
 * be NULL: unweighted); an id outside [0, n) or a negative weight gives
 * DAWN_ERR_INVALID_ARG.  Reference: DAWN::Matrix::coo2Csr / coo2Csr_Weighted
 * (src/matrix.cpp:28-82) after DAWN::IO::readGraph* (src/io.cpp:2-192). */
int dawn_graph_from_edges(int32_t n, int64_t count, const int32_t* u,
                          const int32_t* v, const int32_t* w,
                          int32_t symmetrize, int32_t dedup, dawn_graph** out);
/* The seeded generators above, run on the device (same graphs as
 * dawn_gen_* + dawn_graph_create, bit for bit). */
int dawn_graph_gen_er(int32_t n, double avg_degree, uint64_t seed,
                      int32_t weighted, int32_t wmin, int32_t wmax,
                      dawn_graph** out);
int dawn_graph_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                        double c, uint64_t seed, int32_t weighted,
                        int32_t wmin, int32_t wmax, dawn_graph** out);
int dawn_graph_gen_grid(int32_t side, double extra_fraction, uint64_t seed,
                        int32_t weighted, int32_t wmin, int32_t wmax,
                        dawn_graph** out);
/* Copies a device graph's CSR to a host dawn_csr (free with dawn_csr_free). */
int dawn_graph_download(const dawn_graph* g, dawn_csr* out);

/* Single-source unweighted DAWN (BOVM).  dist: n int32, host or device
 * pointer.  Reference: DAWN::BFS_GPU::run (src/algorithm/gpu/bfs.cu:107-129),
 * DAWN::BFS_CPU::run (src/algorithm/cpu/bfs.cpp:9-20). */
int dawn_bfs(const dawn_graph* g, int32_t source, int32_t* dist,
             int32_t policy, dawn_stats* stats);

/* Single-source int32 min-plus DAWN (SOVM).  Reference: DAWN::SSSP_GPU::run
 * (src/algorithm/gpu/sssp.cu:131-157), DAWN::SSSP_CPU::run
 * (src/algorithm/cpu/sssp.cpp:9-21). */
int dawn_sssp(const dawn_graph* g, int32_t source, int32_t* dist,
              int32_t policy, dawn_stats* stats);

/* dawn_bfs without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and statuses as
 * dawn_bfs, but the traversal state lives in a fixed scratch region that
 * the library's code object carries (66 MB per GPU, held until the process
 * ends) instead of the graph's workspace, so graphs of at most 4,194,304
 * vertices: larger ones return DAWN_ERR_SCRATCH_LIMIT.  The call allocates
 * nothing (no hipMalloc / hipFree, no captured graph), on the graph's
 * device and stream, and returns when the traversal is done.  The learned
 * policy uses the single-source scheduler network (dawn_sched_set_mlp_single)
 * only if it is already on the device (dawn_bfs has used it there);
 * otherwise the fixed rule.
 * stats->device_ms is host wall time of the traversal.  All calls on one
 * GPU share its region: the calls of a process are serialised per device
 * (calls that overlapped on one GPU would race). */
int dawn_bfs_noalloc(const dawn_graph* g, int32_t source, int32_t* dist,
                     int32_t policy, dawn_stats* stats);

/* Multi-source shortest paths.  Unweighted graphs run 64*W sources per
 * bit-packed batch (W = the lane-words knob, up to 16: 1024 sources);
 * weighted graphs run the batched min-plus engine, up to 64 sources per
 * batch.
 * sources are taken modulo n (reference: msource[i] % rows, mssp.cpp:15);
 * isolated sources are skipped (dist row left INF except dist[src]=0, sum 0).
 * dist: S*n int32 row-major (host or device), or NULL to keep only sums
 * (the rows are still computed, each batch's into a device scratch buffer,
 * as the reference computes every source's distance array and discards it).
 * sums: S int64 sums of distances over reached vertices, or NULL.
 * Reference: DAWN::MSSP_GPU::run / run_Weighted (src/algorithm/gpu/mssp.cu:9-102),
 * DAWN::MSSP_CPU::run (src/algorithm/cpu/mssp.cpp:9-32). */
int dawn_mssp(const dawn_graph* g, const int32_t* sources, int32_t count,
              int32_t* dist, int64_t* sums, int32_t policy, dawn_stats* stats);

/* All-pairs: sources [first, first+count) streamed in batches; dist is
 * count*n (host/device) or NULL.  Reference: DAWN::APSP_GPU::run
 * (src/algorithm/gpu/apsp.cu:62-111). */
int dawn_apsp_block(const dawn_graph* g, int32_t first, int32_t count,
                    int32_t* dist, int64_t* sums, int32_t policy,
                    dawn_stats* stats);

/* Closeness centrality (n-1)/sum(dist), sum exact in int64, one IEEE double
 * division.  closeness/sums: count entries (host).  ref_int: if non-NULL,
 * receives the reference's unweighted value (row-1)/int32-sum with integer
 * division (src/algorithm/cpu/cc.cpp:77-78) as float.
 * Unweighted graphs: the batched engine computes only the sums (no
 * distance rows are written).
 * Reference: DAWN::CC_GPU::run (src/algorithm/gpu/cc.cu:9-103). */
int dawn_closeness(const dawn_graph* g, const int32_t* sources, int32_t count,
                   double* closeness, int64_t* sums, float* ref_int,
                   int32_t policy, dawn_stats* stats);

/* Uneven batch: many independent graphs in one persistent launch.
 * graphs[i] are device graphs; sources are count_per_graph per graph,
 * concatenated.  dist_offsets[i] = offset of graph i's first row in dist
 * (rows of n_i); dist may be host or device, or NULL (rows computed into a
 * scratch buffer and dropped, e.g. for timing).  Weighted graphs run
 * min-plus. */
int dawn_batch_mssp(dawn_graph* const* graphs, int32_t num_graphs,
                    const int32_t* sources, int32_t count_per_graph,
                    int32_t* dist, const int64_t* dist_offsets,
                    dawn_stats* stats);

/* Profiling pass of dawn_mssp (unweighted): the same kernels launched one
 * at a time with HIP events around each launch on the graph's stream,
 * writing distance rows and statistics like a timed traversal.
 * kernel_ms[8] / launches[8] are indexed by DAWN_PROF_*: EXPAND_* is the
 * step kernel of the chosen mode (a pull step includes its in-place fold of
 * the light vertices), FINALIZE the fold that follows it (the full fold
 * after push / top-down steps, the heavy-vertex fold after pull steps);
 * levels (ceil(count/64) entries) receives the level count of each batch of
 * 64·W sources (W = lane words in use), then -1 for the unused entries. */
enum {
  DAWN_PROF_FILL = 0, DAWN_PROF_SEED = 1, DAWN_PROF_DECIDE = 2,
  DAWN_PROF_EXPAND_PUSH = 3, DAWN_PROF_EXPAND_PULL = 4,
  DAWN_PROF_EXPAND_BITMAP = 5, DAWN_PROF_FINALIZE = 6, DAWN_PROF_FINAL = 7
};
int dawn_mssp_profile(const dawn_graph* g, const int32_t* sources,
                      int32_t count, int32_t policy, double* kernel_ms,
                      int64_t* launches, int32_t* levels);

/* Plain device-to-device copy of `bytes` (read + write), `iters` timed
 * launches: achieved GB/s.  No counterpart in the reference; used to
 * calibrate rocprof traffic counters and as the box's achievable HBM
 * bandwidth next to the 8 TB/s peak. */
int dawn_probe_copy(int64_t bytes, int32_t iters, double* gbps);

/* Device memory for callers without a GPU framework of their own (the
 * reference keeps its device arrays in thrust::device_vector,
 * include/dawn/algorithm/gpu/bfs.cuh:17-18): allocation on the current
 * device, and synchronous copies / fills (hipMemcpyDefault: any direction).
 * Pointers from another HIP runtime loaded in the same process (a
 * framework that bundles its own) are not valid here.
 * dawn_graph_synchronize waits for the work queued on g's stream. */
int dawn_device_malloc(int64_t bytes, void** out);
int dawn_device_free(void* p);
int dawn_memcpy(void* dst, const void* src, int64_t bytes);
int dawn_memset(void* p, int32_t byte_value, int64_t bytes);
int dawn_graph_synchronize(const dawn_graph* g);

/* ----------------------------------------------------------- scheduler --- */

/* Number of per-step features read by the scheduler MLP. */
#define DAWN_SCHED_FEATURES 8
#define DAWN_SCHED_HIDDEN 32
/* Parameter count of the 8->32->32->3 MLP: 8*32+32 + 32*32+32 + 32*3+3. */
#define DAWN_SCHED_PARAMS 1443

/* Install / clear the process-wide scheduler network (1443 floats, layout:
 * W1[32][8], b1[32], W2[32][32], b2[32], W3[3][32], b3[3]) and the feature
 * normalisation (mean[8], inv_std[8]).  params == NULL clears it. */
int dawn_sched_set_mlp(const float* params, const float* feat_mean,
                       const float* feat_inv_std);

/* Record per-step mode timings: for each batch step every mode is run from
 * the same state and timed with HIP events (minimum of `repeats`); then the
 * engine advances.  features: rows*8 floats, times_ms: rows*3 floats,
 * fixed_modes (nullable): rows ints, the mode the fixed alpha/beta rule
 * chose at that step.  *rows receives the number of rows written
 * (<= max_rows).  Unweighted graphs only. */
int dawn_sched_record(const dawn_graph* g, const int32_t* sources,
                      int32_t count, int32_t repeats, float* features,
                      float* times_ms, int32_t* fixed_modes, int32_t max_rows,
                      int32_t* rows);

/* Per-step trace of a policy run: modes[i] and features (rows*8). */
int dawn_sched_trace(const dawn_graph* g, const int32_t* sources,
                     int32_t count, int32_t policy, int32_t* modes,
                     float* features, float* times_ms, int32_t max_rows,
                     int32_t* rows);

/* The same for the single-source BOVM engine (dawn_bfs).  Its learned
 * policy has its own network (trained on single-source steps), installed
 * with dawn_sched_set_mlp_single; without one it takes the fixed rule.
 * Recording: every mode discovers the same levels, so each source is
 * traversed once per mode with that mode forced, steps launched one at a
 * time between HIP events (minimum of `repeats`), and once under the fixed
 * rule (dawn_sched_trace_single: under `policy`) for the features and the
 * rule's choices; one row per (source, step), isolated sources skipped. */
int dawn_sched_set_mlp_single(const float* params, const float* feat_mean,
                              const float* feat_inv_std);
int dawn_sched_record_single(const dawn_graph* g, const int32_t* sources,
                             int32_t count, int32_t repeats, float* features,
                             float* times_ms, int32_t* fixed_modes,
                             int32_t max_rows, int32_t* rows);
int dawn_sched_trace_single(const dawn_graph* g, const int32_t* sources,
                            int32_t count, int32_t policy, int32_t* modes,
                            float* features, float* times_ms,
                            int32_t max_rows, int32_t* rows);

/* MLP on the GPU (hand-written HIP kernels).  x: N*8, y: N*3 targets
 * (log step time per mode; the network regresses them), params: 1443.
 * dawn_mlp_forward: out N*3.  dawn_mlp_loss_grad: mean-squared loss and its
 * gradient w.r.t. params (1443).  dawn_mlp_train: Adam, in-place params. */
int dawn_mlp_forward(const float* params, const float* x, int32_t N,
                     float* out);
int dawn_mlp_loss_grad(const float* params, const float* x, const float* y,
                       int32_t N, float* loss, float* grad);
int dawn_mlp_train(float* params, const float* x, const float* y, int32_t N,
                   int32_t epochs, int32_t batch, float lr, uint64_t seed,
                   float* final_loss);

/* ---------------------------------------------------------- step level --- */

/* One frontier step on caller-owned device state: the step-level
 * equivalent of the reference's SOVMP (BOVM dense step,
 * src/algorithm/cpu/bfs.cpp:159-185) and GOVMP (min-plus dense step,
 * src/algorithm/cpu/sssp.cpp:169-199), and of the CUDA-typed GPU kernel()
 * overloads (include/dawn/algorithm/gpu/bfs.cuh:14-19, sssp.cuh:12-18),
 * with an opaque stream instead of CUDA types.  All pointers are device
 * memory of g's device.
 *   alpha    frontier bitmap, bit v of word v/32: read, then cleared (as
 *            SOVMP clears alpha[j]);
 *   beta     next-frontier bitmap: bits of updated vertices are OR-ed in
 *            (the caller clears it before the step);
 *   dist     int32[n], INT32_MAX = unreached.  Unweighted: every unreached
 *            out-neighbour of a frontier vertex gets `level`.  Weighted:
 *            dist[v] = min(dist[v], dist[u] + w(u,v)) over frontier u;
 *   changed  int64 counter, set to the number of vertices whose distance
 *            changed in this step (0 = converged).
 * mode: DAWN_MODE_PUSH (scan the frontier's out-edges) or DAWN_MODE_PULL
 * (unweighted only: unreached vertices scan their in-edges for a frontier
 * vertex); both give the same result.  stream: a hipStream_t, or NULL for
 * the graph's own stream.  The call is asynchronous on that stream. */
typedef struct dawn_frontier {
  uint32_t* alpha;
  uint32_t* beta;
  int32_t* dist;
  int64_t* changed;
  int32_t level;
  int32_t reserved;
} dawn_frontier;
int dawn_step(const dawn_graph* g, const dawn_frontier* f, int32_t mode, void* stream);

/* ------------------------------------------------------------ autotune --- */

/* Autotuned launch knobs per graph class (no counterpart in the reference,
 * whose launch shape is fixed).  dawn_tune runs a sweep on g (unweighted:
 * whole traversals of 4096 uniform pseudo-random sources under the learned
 * schedule, or the fixed rule when no network is installed; weighted: two
 * single-source traversals), applies the winners to g (a candidate must be
 * 0.5% faster than the best so far) and stores them in the in-process
 * cache, from which graphs
 * created later in the same class start; load/save persist the cache as
 * text, one line per class: "class dense_blocks heavy_degree pull_chunk
 * ss_chunk pull_variant lane_words hub_order". */
int dawn_tune(const dawn_graph* g, int32_t verbose);
int dawn_tune_load(const char* path);
int dawn_tune_save(const char* path);

/* The knobs of one graph, in the order of DAWN_KNOB_*: grid cap of the
 * dense batched kernels, degree from which edge lists are cut into wave
 * chunks, chunk length, single-source steps per captured graph (0 = auto),
 * light-pull variant (0..3), 64-lane words per vertex of the batched engine
 * (1, 2, 4, 8 or 16: 64 to 1024 sources per batch), hub-first vertex
 * order of the batched engine (0 or 1; graphs of 65,536 vertices or more).
 * Knobs change time only, never results.  dawn_graph_set_knobs rejects
 * values out of range. */
enum {
  DAWN_KNOB_DENSE_BLOCKS = 0,
  DAWN_KNOB_HEAVY_DEGREE = 1,
  DAWN_KNOB_PULL_CHUNK = 2,
  DAWN_KNOB_SS_CHUNK = 3,
  DAWN_KNOB_PULL_VARIANT = 4,
  DAWN_KNOB_LANE_WORDS = 5,
  DAWN_KNOB_HUB_ORDER = 6,
  DAWN_NUM_KNOBS = 7
};
int dawn_graph_get_knobs(const dawn_graph* g, int32_t* knobs);
int dawn_graph_set_knobs(dawn_graph* g, const int32_t* knobs);

#ifdef __cplusplus
}
#endif

#endif /* DAWN_CAPI_H */
