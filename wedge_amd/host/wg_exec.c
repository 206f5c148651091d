/*
 * wg_exec.c -- execution drivers: the per-application iteration loops.
 *
 * Each driver mirrors the reference's execution_impl_* loop
 * (source/execution_bfs.c:122-365, source/execution_cc.c:83-325,
 * source/execution_pr.c:89-179, source/execution_sssp.c:93-334):
 *   decide the pull mode  ->  edge pull  ->  activity count  ->  stop at 0
 *   ->  build the Wedge frontier if the next pull is frontier-filtered
 *   ->  frontier swap / clear.
 * The plain schedule is the reference heuristic (dense pull iff
 * num_edges_active > E * t / 100, CC's first iteration always dense).  The
 * learned schedule asks the MLP (wg_sched.c) for the mode and the block
 * granularity; it falls back to the plain heuristic when no model is loaded.
 * Running time covers the loop only, as the reference's statistics do
 * (source/main.c:153-160).
 */
#define _GNU_SOURCE
#include "wedge_host.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "wg_mlp.h"
#include "wg_sched.h"

static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC_RAW, &t);
    return (double)t.tv_sec * 1e3 + (double)t.tv_nsec * 1e-6;
}

/* ---- trace recorder --------------------------------------------------------- */

static wg_trace_row* g_trace = NULL;
static uint64_t g_trace_n = 0, g_trace_cap = 0;

static void trace_reset(void)
{
    g_trace_n = 0;
}

static void trace_push(const wg_trace_row* r)
{
    if (g_trace_n == g_trace_cap) {
        const uint64_t cap = g_trace_cap ? g_trace_cap * 2 : 1024;
        wg_trace_row* p = (wg_trace_row*)realloc(g_trace, cap * sizeof(wg_trace_row));
        if (!p)
            return;
        g_trace = p;
        g_trace_cap = cap;
    }
    g_trace[g_trace_n++] = *r;
}

uint64_t wg_trace_count(void)
{
    return g_trace_n;
}

int wg_trace_copy(wg_trace_row* rows, uint64_t max_rows)
{
    if (!rows)
        return WG_E_INVAL;
    const uint64_t n = g_trace_n < max_rows ? g_trace_n : max_rows;
    memcpy(rows, g_trace, n * sizeof(wg_trace_row));
    return (int)n;
}

/* ---- options ------------------------------------------------------------------- */

void wg_run_opts_default(wg_run_opts* o)
{
    memset(o, 0, sizeof(*o));
    o->schedule = WG_SCHED_PLAIN;
    o->threshold_pct = 20;                 /* include/cmdline.h:30 */
    o->block_log2 = WG_DEFAULT_BLOCK_LOG2; /* include/configuration.h:19-24 */
    o->iterations = 1;                     /* include/cmdline.h:28 */
    o->damping = 0.875;                    /* include/constants.h:39 */
    o->contrib_fp32 = 0;
    o->root = 0;
    o->record_trace = 0;
    o->max_iterations = 0;
}

int wg_ctx_load_graph(wg_ctx* ctx, const wg_graph* g)
{
    if (!ctx || !g)
        return WG_E_INVAL;
    return wg_load_csc(ctx, g->num_vertices, g->num_edges, g->offsets, g->sources, g->weights, g->weight_kind);
}

static int block_log2_of(const wg_run_opts* o)
{
    int b = o->block_log2 ? o->block_log2 : WG_DEFAULT_BLOCK_LOG2;
    if (b < WG_MIN_BLOCK_LOG2)
        b = WG_MIN_BLOCK_LOG2;
    if (b > WG_MAX_BLOCK_LOG2)
        b = WG_MAX_BLOCK_LOG2;
    return b;
}

/* ---- PageRank --------------------------------------------------------------------- */

int wg_run_pr(wg_ctx* ctx, const wg_run_opts* o, double* ranks, wg_run_stats* st)
{
    if (!ctx || !o || o->iterations == 0)
        return WG_E_INVAL;
    wg_run_stats s;
    memset(&s, 0, sizeof(s));
    int rc = wg_init_pr(ctx, o->damping, o->contrib_fp32);
    if (rc)
        return rc;
    const double t0 = now_ms();
    for (uint32_t it = 0; it < o->iterations; ++it) {
        rc = wg_pr_iteration(ctx, NULL);
        if (rc)
            return rc;
    }
    rc = wg_sync(ctx);
    if (rc)
        return rc;
    s.dense_pull_ms = now_ms() - t0;
    s.pull_ms = s.dense_pull_ms;
    s.running_ms = s.pull_ms;
    s.iterations = o->iterations;
    if (ranks) {
        rc = wg_read_ranks(ctx, ranks);
        if (rc)
            return rc;
        uint64_t V = 0;
        wg_graph_info(ctx, &V, NULL);
        double sum = 0.0;
        for (uint64_t v = 0; v < V; ++v)
            sum += ranks[v];
        s.special = sum;
    }
    if (st)
        *st = s;
    return WG_OK;
}

/* ---- converging applications -------------------------------------------------------- */

typedef int (*pull_fn)(wg_ctx*, int, uint64_t*);

/* Device-resident loop: the schedule is decided on the GPU, no per-iteration
 * host round trip (used unless per-iteration trace rows are requested). */
static int run_device_loop(wg_ctx* ctx, const wg_run_opts* o, wg_run_stats* st)
{
    wg_loop_params p;
    p.schedule = o->schedule;
    p.threshold_pct = o->threshold_pct;
    p.block_log2 = block_log2_of(o);
    p.max_iterations = o->max_iterations;
    wg_run_stats s;
    memset(&s, 0, sizeof(s));
    uint64_t it = 0, fit = 0;
    int rc = wg_set_schedule_model(ctx, o->schedule == WG_SCHED_LEARNED ? wg_mlp_params() : NULL, WG_MLP_PARAMS);
    if (rc)
        return rc;
    const double t0 = now_ms();
    rc = wg_converge(ctx, &p, &it, &fit);
    if (rc)
        return rc;
    s.running_ms = now_ms() - t0;
    s.pull_ms = s.running_ms;
    s.iterations = it;
    s.frontier_iterations = fit;
    uint64_t rv = 0, re = 0;
    rc = wg_reached(ctx, &rv, &re);
    if (rc)
        return rc;
    s.vertices_reached = rv;
    s.edges_reached = re;
    if (st)
        *st = s;
    return WG_OK;
}

/* Device time of the kernels in ids (terminated by -1) since the last profile reset. */
static const int pull_kernels[] = { WG_K_MIN_TILE, WG_K_MIN_BLOCKS, WG_K_BFS_DENSE, WG_K_BFS_HEAVY, WG_K_BFS_BLOCKS, -1 };
static const int build_kernels[] = { WG_K_WEDGE_BUILD, -1 };

static int gpu_ms(wg_ctx* ctx, const int* ids, double* ms)
{
    *ms = 0.0;
    for (; *ids >= 0; ++ids) {
        uint64_t n = 0;
        double t = 0.0;
        int rc = wg_profile_get(ctx, *ids, &n, &t);
        if (rc)
            return rc;
        *ms += t;
    }
    return WG_OK;
}

static int run_converging(wg_ctx* ctx, const wg_run_opts* o, int app, pull_fn pull, wg_run_stats* st)
{
    if (!o->record_trace && (o->schedule != WG_SCHED_LEARNED || wg_mlp_loaded()))
        return run_device_loop(ctx, o, st);
    uint64_t V = 0, E = 0;
    wg_graph_info(ctx, &V, &E);
    const uint64_t threshold = (E * (uint64_t)o->threshold_pct) / 100ull;
    const int plain_blog2 = block_log2_of(o);
    wg_run_stats s;
    memset(&s, 0, sizeof(s));
    if (o->record_trace) {
        trace_reset();
        int rc = wg_profile_enable(ctx, 1);
        if (rc)
            return rc;
    }

    uint64_t active_edges = E;          /* CC: source/execution_cc.c:95 */
    uint64_t frontier_vertices = V;
    if (app != WG_APP_CC) {             /* BFS / SSSP: od[root] */
        uint32_t od_root = 0;
        int rc = wg_vertex_outdegree(ctx, o->root, &od_root);
        if (rc)
            return rc;
        active_edges = od_root;
        frontier_vertices = 1;
    }
    uint64_t unvisited = (app == WG_APP_CC) ? 0 : V - 1; /* not reached yet (BFS, SSSP) */

    wg_sched_features f;
    memset(&f, 0, sizeof(f));
    f.app = app;
    f.num_vertices = V;
    f.num_edges = E;
    int rcb = wg_dense_pull_binned(ctx, &f.binned_dense);
    if (rcb)
        return rcb;
    int mode, blog2;
    f.iteration = 0;
    f.frontier_vertices = frontier_vertices;
    f.frontier_edges = active_edges;
    f.unvisited = unvisited;
    wg_sched_decide(o, threshold, plain_blog2, &f, &mode, &blog2);
    if (app == WG_APP_CC)
        mode = WG_MODE_DENSE; /* source/execution_cc.c:103 */

    int rc;
    const double t_start = now_ms();
    if (mode == WG_MODE_FRONTIER) {
        /* source/execution_bfs.c:116-117, source/execution_sssp.c:87-88 */
        rc = wg_wedge_activate_single_vertex(ctx, o->root, blog2);
        if (rc)
            return rc;
    }
    uint32_t it = 0;
    for (;;) {
        if (o->record_trace && (rc = wg_profile_reset(ctx)))
            return rc;
        const double t0 = now_ms();
        uint64_t act = 0;
        rc = pull(ctx, mode, &act);
        if (rc)
            return rc;
        const double t1 = now_ms();
        const double dt_pull = t1 - t0;
        if (mode == WG_MODE_DENSE) {
            s.dense_pull_ms += dt_pull;
        } else {
            s.frontier_pull_ms += dt_pull;
            s.frontier_iterations += 1;
        }
        it += 1;
        s.iterations = it;
        wg_trace_row row;
        memset(&row, 0, sizeof(row));
        row.iteration = it - 1;
        row.mode = (uint32_t)mode;
        row.block_log2 = blog2;
        row.frontier_vertices = frontier_vertices;
        row.frontier_edges = active_edges;
        row.unvisited = unvisited;
        row.pull_ms = dt_pull;
        if (o->record_trace && (rc = gpu_ms(ctx, pull_kernels, &row.gpu_pull_ms)))
            return rc;
        if (it > 2 * V + 64)
            return WG_E_STATE; /* cannot happen for a correct pull: each iteration activates a vertex */
        if (act == 0 || (o->max_iterations && it >= o->max_iterations)) {
            if (o->record_trace)
                trace_push(&row);
            break;
        }
        active_edges = act;
        uint64_t nact = 0, fresh = 0;
        wg_last_pull_counts(ctx, NULL, &nact, &fresh);
        frontier_vertices = nact;
        unvisited = (unvisited >= fresh) ? unvisited - fresh : 0;
        /* next iteration's schedule */
        int next_mode, next_blog2;
        f.iteration = it;
        f.frontier_edges = active_edges;
        f.frontier_vertices = frontier_vertices;
        f.unvisited = unvisited;
        const double t2 = now_ms();
        wg_sched_decide(o, threshold, plain_blog2, &f, &next_mode, &next_blog2);
        if (next_mode == WG_MODE_FRONTIER) {
            rc = wg_wedge_generate_frontier(ctx, next_blog2);
            if (rc)
                return rc;
        }
        rc = wg_advance_frontier(ctx, NULL);
        if (rc)
            return rc;
        if (o->record_trace && (rc = wg_sync(ctx)))
            return rc; /* charge the frontier work to this row, not the next pull */
        const double t3 = now_ms();
        s.wedge_ms += t3 - t2;
        row.wedge_ms = t3 - t2;
        if (o->record_trace && (rc = gpu_ms(ctx, build_kernels, &row.gpu_build_ms)))
            return rc;
        if (o->record_trace)
            trace_push(&row);
        mode = next_mode;
        blog2 = next_blog2;
    }
    s.running_ms = now_ms() - t_start;
    s.pull_ms = s.dense_pull_ms + s.frontier_pull_ms;
    if (o->record_trace && (rc = wg_profile_enable(ctx, 0)))
        return rc;
    uint64_t rv = 0, re = 0;
    rc = wg_reached(ctx, &rv, &re);
    if (rc)
        return rc;
    s.vertices_reached = rv;
    s.edges_reached = re;
    if (st)
        *st = s;
    return WG_OK;
}

int wg_run_bfs(wg_ctx* ctx, const wg_run_opts* o, uint64_t* packed, wg_run_stats* st)
{
    if (!ctx || !o)
        return WG_E_INVAL;
    int rc = wg_init_bfs(ctx, o->root);
    if (rc)
        return rc;
    rc = run_converging(ctx, o, WG_APP_BFS, wg_edge_pull_bfs, st);
    if (rc)
        return rc;
    return packed ? wg_read_bfs(ctx, packed) : WG_OK;
}

int wg_run_cc(wg_ctx* ctx, const wg_run_opts* o, uint32_t* labels, wg_run_stats* st)
{
    if (!ctx || !o)
        return WG_E_INVAL;
    int rc = wg_init_cc(ctx);
    if (rc)
        return rc;
    rc = run_converging(ctx, o, WG_APP_CC, wg_edge_pull_cc, st);
    if (rc)
        return rc;
    return labels ? wg_read_cc(ctx, labels) : WG_OK;
}

int wg_run_sssp(wg_ctx* ctx, const wg_run_opts* o, void* dist, wg_run_stats* st)
{
    if (!ctx || !o)
        return WG_E_INVAL;
    int rc = wg_init_sssp(ctx, o->root);
    if (rc)
        return rc;
    rc = run_converging(ctx, o, WG_APP_SSSP, wg_edge_pull_sssp, st);
    if (rc)
        return rc;
    return dist ? wg_read_sssp(ctx, dist) : WG_OK;
}
