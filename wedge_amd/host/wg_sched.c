/*
 * wg_sched.c -- schedule choice for the converging applications.
 */
#include "wg_sched.h"

#include "wg_mlp.h"
#include "wg_mlp_math.h"

void wg_sched_decide(const wg_run_opts* o, uint64_t threshold, int plain_block_log2, const wg_sched_features* f,
                     int* mode, int* block_log2)
{
    *block_log2 = plain_block_log2;
    switch (o->schedule) {
    case WG_SCHED_DENSE:
        *mode = WG_MODE_DENSE;
        return;
    case WG_SCHED_FRONTIER:
        *mode = WG_MODE_FRONTIER;
        return;
    case WG_SCHED_LEARNED:
        if (wg_sched_binned_dense(f->binned_dense, f->frontier_edges, f->num_edges)) {
            *mode = WG_MODE_DENSE;
            return;
        }
        /* below the gate: the plain rule's frontier pull (wg_mlp_gated) */
        if (!wg_mlp_gated(f->frontier_edges, f->num_edges) &&
            wg_mlp_decide(f, f->frontier_edges > threshold, mode, block_log2) == WG_OK)
            return;
        /* no model loaded: the plain heuristic is the fallback */
        /* fall through */
    default:
        /* source/execution_bfs.c:141 */
        *mode = (f->frontier_edges > threshold) ? WG_MODE_DENSE : WG_MODE_FRONTIER;
        return;
    }
}
