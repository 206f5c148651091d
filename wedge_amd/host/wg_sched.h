/*
 * wg_sched.h -- per-iteration schedule choice: pull mode and Wedge block
 * granularity.  The plain rule is the reference heuristic
 * (source/execution_bfs.c:132-141); the learned rule is the FP32 MLP.
 */
#ifndef WG_SCHED_H
#define WG_SCHED_H

#include <stdint.h>

#include "wedge_host.h"

typedef struct wg_sched_features {
    int app;
    uint32_t iteration;
    uint64_t num_vertices;
    uint64_t num_edges;
    uint64_t frontier_vertices; /* vertices activated by the last pull */
    uint64_t frontier_edges;    /* their out-edges: the reference's num_edges_active */
    uint64_t unvisited;         /* vertices not reached yet (BFS, SSSP) */
    int binned_dense;           /* dense pulls run on the binned layout (wg_dense_pull_binned) */
} wg_sched_features;

/* Fills *mode and *block_log2 for the next pull. */
void wg_sched_decide(const wg_run_opts* o, uint64_t threshold, int plain_block_log2, const wg_sched_features* f,
                     int* mode, int* block_log2);

#endif /* WG_SCHED_H */
