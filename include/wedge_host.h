/*
 * wedge_host.h -- C host library of the MI355X pull engine.
 *
 * Graph ingress (the reference's binary edge-list format), the seeded
 * synthetic generators, and the execution drivers (the per-application
 * iteration loops with the plain threshold heuristic or the learned
 * schedule) that call the device phases declared in wedge_gpu.h.
 *
 * Reference counterparts:
 *   - binary edge-list reader:    source/graphdata.c:946-1133 (header :589-642,
 *                                 exact size check :963-979, weight word :196-202)
 *   - execution drivers:          source/execution_{bfs,cc,pr,sssp}.c
 *                                 (execution_impl_* and execution_initialize_*)
 *   - statistics block:           source/main.c:153-182
 *
 * All functions return 0 on success and a negative WG_E* code on failure; they
 * never exit the process.
 */
#ifndef WEDGE_HOST_H
#define WEDGE_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "wedge_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- in-memory graph: CSC grouped by destination, 32-bit indices ------- */

typedef struct wg_graph {
    uint64_t num_vertices;
    uint64_t num_edges;
    uint32_t* offsets;  /* num_vertices + 1 entries; in-edges of v are [offsets[v], offsets[v+1]) */
    uint32_t* sources;  /* num_edges entries, ascending within each destination */
    void* weights;      /* num_edges entries of double (WG_WEIGHT_F64) or uint32_t (WG_WEIGHT_U32) */
    int weight_kind;    /* WG_WEIGHT_NONE / WG_WEIGHT_F64 / WG_WEIGHT_U32 */
} wg_graph;

/* Number of host threads used by the generators and the loader
 * (WEDGE_THREADS if set, else min(online CPUs, 16)). */
int wg_host_threads(void);

/* R-MAT generator: 2^scale vertices, edge_factor * 2^scale draws, quadrant
 * probabilities a, b, c (d = 1 - a - b - c).  Self-loops and duplicate edges are
 * removed.  symmetrize != 0 adds the reverse of every draw before dedup.
 * weight_kind WG_WEIGHT_F64 attaches wg_weight_f64(src, dst, seed). */
int wg_graph_rmat(wg_graph* g, int scale, int edge_factor, double a, double b, double c,
                  uint64_t seed, int symmetrize, int weight_kind);

/* Road-like grid: rows x cols lattice, each undirected lattice edge kept with
 * probability p_keep, each down-right diagonal with p_diag, both directions
 * stored, integer weights uniform in [1, wmax] per undirected edge. */
int wg_graph_grid(wg_graph* g, uint32_t rows, uint32_t cols, double p_keep, double p_diag,
                  uint32_t wmax, uint64_t seed);

/* Reads the reference's binary edge list: u64 V, u64 E, then E records of
 * (u64 src, u64 dst) or, when weight_kind != WG_WEIGHT_NONE, (src, dst, weight)
 * with the weight word read as a double (WG_WEIGHT_F64) or an unsigned integer
 * (WG_WEIGHT_U32, must be < 2^32).  Records must be grouped by ascending
 * destination.  Duplicates and self-loops are kept, as the reference keeps them. */
int wg_graph_read(wg_graph* g, const char* path, int weight_kind);

/* Writes g in the same binary format (weights as the third word if present). */
int wg_graph_write(const wg_graph* g, const char* path);

void wg_graph_free(wg_graph* g);

/* Disjoint union of n member graphs (all of one weight kind): member g's
 * vertex v becomes first[g] + v.  first receives n + 1 entries.  The union is
 * what wg_set_members expects (include/wedge_gpu.h).  WG_E_RANGE if the union
 * needs more than 32-bit vertex or edge indices. */
int wg_graph_union(const wg_graph* members, uint32_t n, wg_graph* out, uint64_t* first);

/* Out-degree of every vertex (out must hold num_vertices entries). */
int wg_graph_outdegrees(const wg_graph* g, uint32_t* out);

/* Vertex with the largest out-degree (lowest id on ties). */
uint64_t wg_graph_max_outdegree_vertex(const wg_graph* g);

/* Per-edge FP64 weight in (0, 1] derived from (src, dst, seed). */
double wg_weight_f64(uint64_t src, uint64_t dst, uint64_t seed);

/* Counter-based 64-bit random value: output number `counter` of the
 * splitmix64 stream seeded with `key` (Steele, Lea, Flood 2014; Vigna). */
uint64_t wg_rand64(uint64_t key, uint64_t counter);

/* ---- execution drivers ---------------------------------------------------- */

enum wg_schedule_kind {
    WG_SCHED_PLAIN = 0,   /* frontier pull iff activity <= E * t / 100 (reference heuristic) */
    WG_SCHED_LEARNED = 1, /* FP32 MLP picks mode and block granularity */
    WG_SCHED_DENSE = 2,   /* always dense pull (testing) */
    WG_SCHED_FRONTIER = 3 /* always frontier pull (testing) */
};

typedef struct wg_run_opts {
    int schedule;            /* wg_schedule_kind */
    uint32_t threshold_pct;  /* plain heuristic threshold, reference default 20 */
    int block_log2;          /* frontier granularity in edges per bit (log2); 0 = default */
    uint32_t iterations;     /* PageRank iterations */
    double damping;          /* PageRank damping factor */
    int contrib_fp32;        /* PageRank: store contributions in FP32 */
    uint64_t root;           /* BFS / SSSP root */
    int record_trace;        /* record per-iteration trace rows (wg_trace_row) */
    uint32_t max_iterations; /* safety cap for converging apps, 0 = none */
} wg_run_opts;

void wg_run_opts_default(wg_run_opts* o);

/* Per-run statistics (the reference's statistics block, main.c:153-182). */
typedef struct wg_run_stats {
    double running_ms;        /* pull + frontier generation (+ device loop overhead) */
    double pull_ms;
    double frontier_pull_ms;
    double dense_pull_ms;
    double wedge_ms;          /* frontier (Wedge) generation */
    uint64_t iterations;
    uint64_t frontier_iterations;
    uint64_t edges_reached;   /* BFS/SSSP: edges whose source was reached */
    uint64_t vertices_reached;
    double special;           /* PR: total rank (sum); other apps: 0 */
} wg_run_stats;

/* One row per iteration of a converging app (trace recorder). */
typedef struct wg_trace_row {
    uint32_t iteration;
    uint32_t mode;              /* WG_MODE_* used */
    int32_t block_log2;
    uint32_t pad;
    uint64_t frontier_vertices; /* vertices active at iteration start */
    uint64_t frontier_edges;    /* their out-edges (reference num_edges_active) */
    uint64_t unvisited;         /* BFS / SSSP: vertices not reached yet; CC: 0 */
    double pull_ms;             /* host wall time of the pull call */
    double wedge_ms;            /* host wall time of frontier build + advance */
    double gpu_pull_ms;         /* device time of the pull kernels (HIP events) */
    double gpu_build_ms;        /* device time of the frontier build for the next pull */
} wg_trace_row;

/* Runs an application end to end on a context that already holds the graph.
 * out_values receives V entries: BFS uint64 packed (depth << 32 | parent),
 * CC uint32 labels, PR double ranks, SSSP double or uint32 distances. */
int wg_run_pr(wg_ctx* ctx, const wg_run_opts* o, double* ranks, wg_run_stats* st);
int wg_run_bfs(wg_ctx* ctx, const wg_run_opts* o, uint64_t* packed, wg_run_stats* st);
int wg_run_cc(wg_ctx* ctx, const wg_run_opts* o, uint32_t* labels, wg_run_stats* st);
int wg_run_sssp(wg_ctx* ctx, const wg_run_opts* o, void* dist, wg_run_stats* st);

/* Trace rows recorded by the last run with record_trace set. */
uint64_t wg_trace_count(void);
int wg_trace_copy(wg_trace_row* rows, uint64_t max_rows);

/* ---- learned schedule ------------------------------------------------------ */

/* The schedule model (FP32 MLP, layout in wedge_gpu.h) used by
 * WG_SCHED_LEARNED: load / save a model file, set / get the parameters
 * (WG_MLP_PARAMS floats).  Without a model the learned schedule falls back to
 * the plain rule.  The device loop evaluates the same model on the GPU. */
int wg_mlp_load(const char* path);
int wg_mlp_save(const char* path);
int wg_mlp_set(const float* params, uint64_t count);
int wg_mlp_get(float* params, uint64_t count);
void wg_mlp_unload(void);
int wg_mlp_loaded(void);

/* Feature vector (WG_MLP_IN floats) of a loop state, as the schedule sees it:
 * app, iterations completed, graph size, vertices and out-edges activated by
 * the last pull, vertices not reached yet (BFS, SSSP; 0 for CC). */
int wg_mlp_feature_vector(int app, uint64_t iteration, uint64_t num_vertices, uint64_t num_edges,
                          uint64_t frontier_vertices, uint64_t frontier_edges, uint64_t unvisited, float* out);

/* log2 work estimate of every action (WG_MLP_ACTIONS floats) for a loop
 * state: the part of the prediction that is not learned. */
int wg_mlp_baseline(uint64_t num_vertices, uint64_t num_edges, uint64_t frontier_edges, float* out);

/* Host evaluation of the loaded model: predicted log2 times of every action
 * (network output + base; base may be NULL) into out (WG_MLP_ACTIONS floats,
 * may be NULL); returns the action chosen for app (>= 0), settled against the
 * plain rule's action plain_action (-1: none), or an error code. */
int wg_mlp_predict(int app, const float* features, const float* base, int plain_action, float* out);

/* Loads a graph into a context (wraps wg_load_csc). */
int wg_ctx_load_graph(wg_ctx* ctx, const wg_graph* g);

#ifdef __cplusplus
}
#endif

#endif /* WEDGE_HOST_H */
