/*
 * wedge_cpu.h -- ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of Wedge's pull engine (stanford-mast/Wedge), written in
 * plain C with pthreads.  It is the correctness checker for the HIP path and
 * the CPU baseline that bench.py times; nothing in the product links it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.
 *
 * What it restates (reference file:line):
 *   edge-vector layout        source/graphdata.c:212-243 (4 lanes: bit 63 valid,
 *                             bits 47:0 source, bits 62:48 destination pieces)
 *   vertex index              source/graphdata.c:402-427
 *   out-degrees (valid lanes) source/graphdata.c:430-490
 *   Wedge CSRI                source/graphdata.c:493-585
 *   frontier sizing           source/graphdata.c:804-892
 *   work units / scheduling   32 dynamically scheduled units per thread,
 *                             e.g. source/phase_edge_pull_pr.c:36-53
 *   PR edge phase             source/phase_edge_pull_pr.c:28-116
 *   PR merge                  source/phases.c:69-114
 *   PR vertex phase           source/phase_vertex_pr.c:26-77
 *   BFS phases                source/phase_edge_pull_bfs.c:109-332
 *   CC phases                 source/phase_edge_pull_cc.c:53-297
 *   SSSP phases               source/phase_edge_pull_sssp.c:57-306
 *   idempotent merge          source/phases.c:54-65
 *   Wedge frontier            source/wedge.c:37-155
 *   drivers + heuristic       source/execution_{pr,bfs,cc,sssp}.c
 *
 * Deliberate, documented differences from the reference:
 *   - damping is a parameter (the reference hardcodes 0.875,
 *     include/constants.h:39);
 *   - integer SSSP takes a full 64-bit minimum; the reference's minimum works
 *     on 32-bit halves (source/phase_edge_pull_sssp.c:140-142), which agrees
 *     with a 64-bit minimum whenever distances are below 2^32;
 *   - FP64-weight SSSP is added (dist = min over fl(dist[u] + w));
 *   - BFS additionally records the depth of every vertex (the iteration of its
 *     first discovery, root = 0); the parent array follows the reference,
 *     including its root quirk (the root is rediscovered and given a child as
 *     parent, source/execution_bfs.c:106-117 with phase_edge_pull_bfs.c:163);
 *   - work units that receive no edge vectors are skipped instead of reading
 *     past the edge list.
 */
#ifndef WEDGE_CPU_H
#define WEDGE_CPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wcpu wcpu;

#define WCPU_WEIGHT_NONE 0
#define WCPU_WEIGHT_F64 1
#define WCPU_WEIGHT_U32 2

typedef struct wcpu_stats {
    double running_ms;       /* pull + Wedge generation, as source/main.c:153-160 */
    double pull_ms;
    double frontier_pull_ms;
    double dense_pull_ms;
    double wedge_ms;
    uint64_t iterations;
    uint64_t frontier_iterations;
    double special;          /* PR: sum of ranks */
} wcpu_stats;

/* Builds the reference's in-memory layout from a destination-grouped CSC.
 * nthreads = 0 uses all online CPUs.  Returns NULL on failure. */
wcpu* wcpu_create(uint64_t num_vertices, uint64_t num_edges, const uint32_t* offsets,
                  const uint32_t* sources, const void* weights, int weight_kind, int nthreads);
void wcpu_destroy(wcpu* w);
int wcpu_threads(const wcpu* w);
uint64_t wcpu_num_vectors(const wcpu* w);
/* Out-degree as counted from valid lanes (source/graphdata.c:430-490). */
int wcpu_outdegrees(const wcpu* w, uint64_t* out);

/* PageRank: ranks[v] = contrib[v] * od'[v] after `iterations` iterations. */
int wcpu_pr(wcpu* w, uint32_t iterations, double damping, double* ranks, wcpu_stats* st);
/* BFS: parent (UINT64_MAX unreached) and depth (UINT32_MAX unreached). */
int wcpu_bfs(wcpu* w, uint64_t root, uint32_t threshold_pct, uint64_t* parent, uint32_t* depth,
             wcpu_stats* st);
/* CC: FP64 labels (integers) on the given graph. */
int wcpu_cc(wcpu* w, uint32_t threshold_pct, double* labels, wcpu_stats* st);
/* SSSP: integer graphs -> uint64 distances (UINT64_MAX unreached);
 *       FP64 graphs    -> double distances (+inf unreached). */
int wcpu_sssp(wcpu* w, uint64_t root, uint32_t threshold_pct, void* dist, wcpu_stats* st);

/* Personalised PageRank (no counterpart in the reference; its PageRank
 * teleports uniformly).  For each seed s, in FP64, starting from r = e_s:
 *   r(v) = d * (sum over in-edges u->v of c(u) + [v = s] * D) + [v = s] * (1 - d)
 *   c(u) = r(u) / od(u) (0 when od(u) = 0), D = sum of r over vertices with od = 0
 * ranks receives nseeds rows of V values after `iterations` iterations.  In-edge
 * sums follow the edge-vector order (ascending source within a destination). */
int wcpu_ppr(wcpu* w, const uint64_t* seeds, uint32_t nseeds, uint32_t iterations, double damping, double* ranks);

#ifdef __cplusplus
}
#endif

#endif /* WEDGE_CPU_H */
