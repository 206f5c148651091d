/*
 * wedge_gpu.h -- C-ABI of the MI355X (gfx950) pull-engine phases.
 *
 * This is the drop-in boundary for the reference's phase operators
 * (include/phases.h:27-88, include/wedge.h:23-27).  The reference passes data
 * through process globals (include/graphdata.h:31-110) and returns nothing;
 * here every device buffer belongs to a context, host pointers are borrowed
 * only for the duration of a call, and every entry point returns an int status
 * (0 = WG_OK, negative = WG_E_*).  Nothing in this library exits or aborts.
 *
 * Calls on one context are serialised on the context's HIP stream; a context
 * must be used by one host thread at a time.  Scalars that the reference
 * returns through graph_reduce_buffer (source/phases.c:26-50) come back through
 * optional out-parameters: passing a non-NULL pointer makes the call wait for
 * the device and copy the value; passing NULL leaves the value on the device
 * and returns without waiting.
 */
#ifndef WEDGE_GPU_H
#define WEDGE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wg_ctx wg_ctx;

/* ---- status codes --------------------------------------------------------- */
#define WG_OK           0
#define WG_E_INVAL     -1  /* bad argument */
#define WG_E_NOMEM     -2  /* device or host allocation failed */
#define WG_E_HIP       -3  /* a HIP runtime call failed (see wg_last_hip_error) */
#define WG_E_STATE     -4  /* call not valid in the context's current state */
#define WG_E_RANGE     -5  /* vertex id or size out of range */
#define WG_E_IO        -6  /* file could not be opened / read / written */
#define WG_E_FORMAT    -7  /* input file malformed (size check, ordering) */
#define WG_E_OVERFLOW  -8  /* integer distance exceeded 32 bits */
#define WG_E_NODEV     -9  /* no HIP device */

/* ---- enums ---------------------------------------------------------------- */
#define WG_WEIGHT_NONE 0
#define WG_WEIGHT_F64  1
#define WG_WEIGHT_U32  2

#define WG_APP_NONE 0
#define WG_APP_BFS  1
#define WG_APP_CC   2
#define WG_APP_PR   3
#define WG_APP_SSSP 4
#define WG_APP_MSBFS 5   /* batched sources (wg_msbfs) */
#define WG_APP_MSSSSP 6  /* batched sources (wg_mssssp) */
#define WG_APP_PPR 7     /* batched personalised PageRank (wg_ppr) */

#define WG_MODE_DENSE    0  /* reference: perform_edge_pull_phase_X */
#define WG_MODE_FRONTIER 1  /* reference: perform_edge_pull_phase_X_with_wedge_frontier */

/* Default frontier granularity: 2^4 = 16 edges per Wedge bit, the reference's
 * 4 edge vectors of 4 lanes per bit (include/configuration.h:19-24). */
#define WG_DEFAULT_BLOCK_LOG2 4
#define WG_MIN_BLOCK_LOG2 3
#define WG_MAX_BLOCK_LOG2 10

/* Kernel ids for wg_profile_get. */
#define WG_K_PR_TILE        0  /* fused PageRank edge pull + vertex phase, edge tiles */
#define WG_K_PR_SPAN        1  /* PageRank destinations spanning several tiles */
#define WG_K_PR_VERTEX      2  /* stand-alone vertex phase */
#define WG_K_MIN_TILE       3  /* CC / SSSP dense pull */
#define WG_K_MIN_BLOCKS     4  /* CC / SSSP frontier pull */
#define WG_K_BFS_DENSE      5
#define WG_K_BFS_HEAVY      6
#define WG_K_BFS_BLOCKS     7
#define WG_K_WEDGE_BUILD    8
#define WG_K_ADVANCE        9
#define WG_K_PR_ACC        10  /* stand-alone PageRank edge pull (accumulators) */
#define WG_K_LOOP          11  /* whole device-resident iteration loop (wg_converge) */
#define WG_K_BATCH_PR      12  /* whole batched PageRank run (wg_batch_pr) */
#define WG_K_PB_PUSH       13  /* binned PageRank: window partial sums */
#define WG_K_PB_PULL       14  /* binned PageRank: bin accumulation + vertex phase */
#define WG_K_COUNT         15

/* ---- context -------------------------------------------------------------- */

/* Opens a context on HIP device `device` with its own stream. */
int wg_open(wg_ctx** out, int device);
void wg_close(wg_ctx* ctx);
const char* wg_error_string(int code);
/* Last hipError_t seen by the context (0 if none). */
int wg_last_hip_error(const wg_ctx* ctx);
/* Waits for all work queued on the context's stream. */
int wg_sync(wg_ctx* ctx);

/* Context options (set before wg_load_csc). */
#define WG_OPT_RELABEL 1  /* 1: store vertices in descending total-degree order on
                             the device; 0: keep the caller's order; 2 (default):
                             relabel only skewed graphs (maximum in-degree above
                             64 and above 16x the mean degree).  Every id and
                             value crossing this API stays in the caller's
                             original numbering either way */
#define WG_RELABEL_OFF 0
#define WG_RELABEL_ON 1
#define WG_RELABEL_AUTO 2
#define WG_OPT_PR_BIN_SLICE 2  /* binned PageRank pull (WG_PR_BINNED): window-destination
                                  sums one pull workgroup adds, 1..2^24 (default
                                  262144); a bin with more is split over several
                                  workgroups whose accumulators are then added.
                                  Takes effect at the next layout build */
#define WG_OPT_MS_BINNED 3     /* dense levels of the 64-source BFS (wg_msbfs) on the binned
                                  layout of WG_PR_BINNED (OR of frontier masks per
                                  window-destination pair, then per bin): 0 off (edge
                                  tiles), 1 on, 2 (default) on for graphs of 2^24 edges
                                  or more */
#define WG_OPT_COUNT_WORK 4    /* 1: the BFS / CC / SSSP pulls count the work that
                                  wg_work_counts reports (kernels built with the
                                  counts, a little slower); 0 (default): they do not.
                                  May be changed between runs */
#define WG_OPT_MIN_BINNED 5    /* dense CC / SSSP pulls (wg_edge_pull_cc / _sssp in WG_MODE_DENSE,
                                  and the dense iterations of wg_converge) on the binned
                                  layout of WG_PR_BINNED: per window-destination pair the
                                  minimum of the window's offers (label, or distance plus
                                  weight), then per bin the minimum against the current
                                  value.  A dense pull then reads the values of the
                                  previous iteration (the tile pull may also see values
                                  lowered earlier in the same pull), so iteration counts
                                  can differ; the fixpoint is the same.  0 off, 1 on,
                                  2 (default) on for graphs of 2^24 edges or more and at
                                  least 8 edges per vertex */
int wg_set_option(wg_ctx* ctx, int option, int64_t value);

/* Uploads a destination-grouped CSC and builds every derived device structure:
 * out-degrees, per-edge destination column, the source-to-edge index (the
 * reference's CSRI), the PageRank edge tiles.  offsets has V+1 entries with
 * offsets[V] == E < 2^32; sources has E entries, ascending within each
 * destination.  weights (E entries) is read according to weight_kind.
 * Replaces graph_data_read_from_file + graph_data_initialize_data_structures
 * (source/graphdata.c:946-1133, :648-942) for the device. */
int wg_load_csc(wg_ctx* ctx, uint64_t num_vertices, uint64_t num_edges,
                const uint32_t* offsets, const uint32_t* sources,
                const void* weights, int weight_kind);
/* Builds the graph on the device from an edge list in any order (no reference
 * counterpart; the reference reads edges already grouped by destination,
 * source/graphdata.c:247-585): src[e] -> dst[e] with optional weights[e]
 * (layout of weight_kind).  The edges are sorted by (destination, source) with
 * a stable radix sort; flags drop self-loops (WG_EDGES_DROP_SELF_LOOPS) and
 * repeated (src, dst) pairs (WG_EDGES_DEDUP, the first occurrence's weight
 * stays).  The result is loaded as by wg_load_csc.  kept_edges (optional)
 * receives the number of edges kept.  WG_E_RANGE if an endpoint is >= V. */
#define WG_EDGES_DROP_SELF_LOOPS 1
#define WG_EDGES_DEDUP 2
int wg_load_edges(wg_ctx* ctx, uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst, const void* weights,
                  int weight_kind, int flags, uint64_t* kept_edges);

/* Number of vertices / edges loaded. */
int wg_graph_info(const wg_ctx* ctx, uint64_t* num_vertices, uint64_t* num_edges);
/* Out-degree array (V entries) as computed on the device. */
int wg_read_outdegrees(wg_ctx* ctx, uint32_t* host_out);
/* Out-degree of one vertex (the reference reads graph_vertex_outdegrees[root]
 * to seed num_edges_active, source/execution_bfs.c:133). */
int wg_vertex_outdegree(wg_ctx* ctx, uint64_t vertex, uint32_t* out);

/* ---- per-application initialisation (execution_initialize_*) -------------- */

/* source/execution_pr.c:73-85: contrib = (1/N)/od', accumulators = 0.
 * contrib_fp32 != 0 stores contributions in FP32 (accumulation stays FP64). */
int wg_init_pr(wg_ctx* ctx, double damping, int contrib_fp32);
/* source/execution_bfs.c:84-118: root visited at depth 0, root in frontier. */
int wg_init_bfs(wg_ctx* ctx, uint64_t root);
/* source/execution_cc.c:69-79: label[v] = v. */
int wg_init_cc(wg_ctx* ctx);
/* source/execution_sssp.c:70-89: dist = +inf, dist[root] = 0, root active.
 * Distances are FP64 for WG_WEIGHT_F64 graphs, uint32 for WG_WEIGHT_U32. */
int wg_init_sssp(wg_ctx* ctx, uint64_t root);
/* The same with several roots at depth / distance 0 (repeats allowed): on a
 * graph made of disjoint members, one root per member runs every member's
 * BFS / SSSP in one traversal (wg_converge runs it). */
int wg_init_bfs_roots(wg_ctx* ctx, const uint64_t* roots, uint32_t n);
int wg_init_sssp_roots(wg_ctx* ctx, const uint64_t* roots, uint32_t n);

/* ---- PageRank phases ------------------------------------------------------ */

/* perform_edge_pull_phase_pr (source/phase_edge_pull_pr.c:28-116) followed by
 * edge_pull_op_merge_with_merge_buffer (source/phases.c:69-114) and the
 * reduce-buffer combine (source/phases.c:40-50): accumulators[v] = sum of
 * contributions over in-edges; total_rank = sum of all accumulators. */
int wg_edge_pull_pr(wg_ctx* ctx, double* total_rank);
/* perform_vertex_phase_pr (source/phase_vertex_pr.c:26-77), using the
 * total_rank of the preceding wg_edge_pull_pr. */
int wg_vertex_pr(wg_ctx* ctx);
/* One whole PageRank iteration (edge pull + vertex phase) as one fused tile
 * kernel plus a tile-spanning fix-up; the total rank the vertex phase needs is
 * carried from the previous iteration.  total_rank (optional) is the sum of
 * the accumulators of this iteration. */
int wg_pr_iteration(wg_ctx* ctx, double* total_rank);
/* Final ranks: rank[v] = contrib[v] * od'[v] (source/execution_pr.c:39-42). */
int wg_read_ranks(wg_ctx* ctx, double* host_ranks);

/* ---- traversal phases ----------------------------------------------------- */

/* perform_edge_pull_phase_bfs (source/phase_edge_pull_bfs.c:109-210) for
 * WG_MODE_DENSE and perform_edge_pull_phase_bfs_with_wedge_frontier (:212-332)
 * for WG_MODE_FRONTIER.  edges_activated = sum of out-degrees of the vertices
 * discovered (the reference's num_edges_changed summed over threads). */
int wg_edge_pull_bfs(wg_ctx* ctx, int mode, uint64_t* edges_activated);
/* perform_edge_pull_phase_cc[_with_wedge_frontier]
 * (source/phase_edge_pull_cc.c:53-164, :166-297) plus the idempotent merge
 * (source/phases.c:54-65). */
int wg_edge_pull_cc(wg_ctx* ctx, int mode, uint64_t* edges_activated);
/* perform_edge_pull_phase_sssp[_with_wedge_frontier]
 * (source/phase_edge_pull_sssp.c:57-169, :171-306) plus the idempotent merge.
 * Returns WG_E_OVERFLOW if an integer distance reached 2^32 - 1. */
int wg_edge_pull_sssp(wg_ctx* ctx, int mode, uint64_t* edges_activated);

/* Edges and vertices activated by the last pull whose edges_activated
 * out-parameter was non-NULL, and how many of those vertices were reached for
 * the first time (BFS: all; SSSP: distance was infinite; CC: 0).  Values
 * already on the host; no device wait. */
int wg_last_pull_counts(wg_ctx* ctx, uint64_t* edges, uint64_t* vertices, uint64_t* newly_reached);

/* wedge_generate_frontier (source/wedge.c:63-155): marks every 2^block_log2
 * edge block that holds an out-edge of a vertex activated by the last pull,
 * and queues those blocks for the next WG_MODE_FRONTIER pull. */
int wg_wedge_generate_frontier(wg_ctx* ctx, int block_log2);
/* wedge_activate_single_vertex (source/wedge.c:52-59). */
int wg_wedge_activate_single_vertex(wg_ctx* ctx, uint64_t vertex, int block_log2);
/* End of an iteration: the vertices activated by the last pull become the
 * current frontier and the next-frontier bitmask is cleared (the reference's
 * active/snapshot swap and memset, source/execution_bfs.c:326-338 and
 * source/execution_cc.c:243-277).  frontier_vertices (optional) = number of
 * vertices in the new frontier. */
int wg_advance_frontier(wg_ctx* ctx, uint64_t* frontier_vertices);

/* Runs the initialised BFS / CC / SSSP to convergence with the per-iteration
 * schedule decided on the device (the reference's execution_impl_* loop,
 * source/execution_bfs.c:122-365 and its cc/sssp twins): a one-thread decide
 * kernel reads the activity count of each pull, stops at 0, and picks the
 * next pull's mode (and Wedge granularity); every pull, frontier-build and
 * advance kernel reads that choice and exits at once when not selected.  The
 * host enqueues iterations in growing batches and never waits per iteration.
 * schedule: 0 plain threshold rule, 1 learned (the model of
 * wg_set_schedule_model evaluated on the device; the plain rule if no model is
 * set; either way a CC / SSSP pull is dense whenever its dense pulls run on the
 * binned layout, wg_dense_pull_binned, and the activity exceeds E / 64),
 * 2 always dense, 3 always frontier. */
typedef struct wg_loop_params {
    int schedule;
    uint32_t threshold_pct;  /* plain rule: dense iff activity > E * pct / 100 */
    int block_log2;          /* Wedge granularity of the plain rule */
    uint32_t max_iterations; /* 0 = until convergence */
} wg_loop_params;
int wg_converge(wg_ctx* ctx, const wg_loop_params* params, uint64_t* iterations, uint64_t* frontier_iterations);
/* *on = 1 if the initialised CC / SSSP's dense pulls run on the binned layout
 * (WG_OPT_MIN_BINNED and the graph's size), else 0.  No reference counterpart:
 * the learned schedule of the host-driven loop (wg_run_cc / wg_run_sssp with a
 * trace recorded) reads it to take the dense pulls as wg_converge does. */
int wg_dense_pull_binned(const wg_ctx* ctx, int* on);
/* Fused small-frontier levels.  No reference counterpart: an execution
 * option of wg_converge.  While the chosen pull is frontier-filtered and a
 * level is small (at most 1024 activated vertices and 8192 out-edges per
 * workgroup of the run), the loop runs its levels back to back in one launch
 * of `workgroups` resident workgroups (one grid barrier per level) instead of
 * four launches per level, each level touching only the out-edges of its
 * active vertices.  Sets the workgroup count for every application whose bit
 * (1 << WG_APP_BFS, WG_APP_CC, WG_APP_SSSP) is in apps; 0 turns the option
 * off for them; at most 256 and at most the device's CU count.  BFS results,
 * iteration counts and schedule decisions are unchanged; CC labels and SSSP
 * distances are unchanged, their iteration counts may differ (a
 * frontier-filtered pull reads whole edge blocks, which can settle values in
 * fewer iterations).  Default: BFS with 32 workgroups, CC and SSSP off. */
int wg_set_fused_levels(wg_ctx* ctx, uint32_t workgroups, uint32_t apps);
/* Iterations of the last wg_converge that ran fused (included in its
 * iteration counts). */
int wg_fused_level_count(wg_ctx* ctx, uint64_t* levels);

/* ---- batched sources -------------------------------------------------------
 * No reference counterpart (the reference's BFS and SSSP take one root,
 * source/execution_bfs.c:84-118, source/execution_sssp.c:70-89): up to
 * WG_MS_MAX_SOURCES sources traverse the graph together, lane s of a
 * per-vertex 64-bit mask (BFS) or of a per-vertex distance row (SSSP)
 * belonging to source s.  Each source's result equals that of a single-source
 * run from it.  params: schedule 0 = plain rule per level (on the summed
 * out-degree of the vertices any source newly reached), 2 = dense, 3 =
 * frontier (1 behaves as 0); threshold_pct, block_log2, max_iterations as in
 * wg_converge.  Roots may repeat. */
#define WG_MS_MAX_SOURCES 64
/* levels: BFS levels run; edges_reached: sum over sources of the out-degrees
 * of the vertices each reached (both optional).  WG_E_OVERFLOW if a depth
 * would exceed 65534.  Its device buffers (128V bytes of 16-bit depth rows,
 * 64 per vertex, three 8V-byte masks and 128V bytes keeping the new-source
 * masks of the first 16 levels, from which the depth rows are written once
 * per vertex) stay allocated until another application is initialised or the
 * graph changes, so that further calls on the same graph do not allocate
 * them again.  A threshold_pct of 1-2 suits batched levels (the Python
 * binding's default is 2). */
int wg_msbfs(wg_ctx* ctx, const uint64_t* roots, uint32_t nroots, const wg_loop_params* params, uint64_t* levels,
             uint64_t* edges_reached);
/* Depths (V entries) from root number `source`, UINT32_MAX if unreached. */
int wg_read_msbfs(wg_ctx* ctx, uint32_t source, uint32_t* host_depths);
/* SSSP from up to 64 roots (weighted graphs): per vertex a row of 64
 * distances.  iterations: relaxation rounds run (optional).  WG_E_OVERFLOW if
 * an integer distance reached 2^32 - 1. */
int wg_mssssp(wg_ctx* ctx, const uint64_t* roots, uint32_t nroots, const wg_loop_params* params,
              uint64_t* iterations);
/* Distances (V entries) from root number `source`: double (+inf unreached) on
 * FP64-weighted graphs, uint32 (0xFFFFFFFF unreached) on integer ones. */
int wg_read_mssssp(wg_ctx* ctx, uint32_t source, void* host_dist);

/* Personalised PageRank for up to WG_PPR_MAX_SEEDS seeds in one pass over
 * the edges per iteration (no reference counterpart; the reference's PageRank
 * teleports uniformly, source/phase_vertex_pr.c:26-77).  Seed s restarts at s
 * and its dangling mass returns to s; starting from all mass on s:
 *   r(v) = d * (sum over in-edges u->v of r(u)/od(u) + [v = s] * D) + [v = s] * (1 - d)
 * with D the mass on vertices without out-edges.  Contributions are stored in
 * FP32 (a 128-byte row per vertex), sums and ranks are FP64, in a fixed order.
 * The device buffers (2 x 128V bytes of contributions, 256V bytes of ranks,
 * partials) stay allocated until another application is initialised or the
 * graph changes, so that further calls on the same graph reuse them. */
#define WG_PPR_MAX_SEEDS 32
int wg_ppr(wg_ctx* ctx, const uint64_t* seeds, uint32_t nseeds, uint32_t iterations, double damping);
/* Ranks (V entries, FP64) for seed number `seed`. */
int wg_read_ppr(wg_ctx* ctx, uint32_t seed, double* host_ranks);

/* ---- validation (no reference counterpart) -------------------------------
 * Graph500-style certificates checked on the device against the loaded graph,
 * one pass over the edges and one over the vertices:
 *  BFS   roots: depth 0 and parent = itself; every edge u->v with u reached has
 *        v reached and depth(v) <= depth(u) + 1; every other reached v has its
 *        parent as an in-neighbour one level up.
 *  SSSP  roots: distance 0; every edge u->v with u reached has v reached and
 *        dist(v) <= dist(u) + w; every other reached v has an in-edge with
 *        dist(u) + w == dist(v) (with positive weights: the shortest paths).
 *  CC    every edge u->v has label(v) <= label(u); label(v) <= v; label(v) == v
 *        or an in-neighbour carries label(v) (the minimum reaching label).
 * violations = failing edges + failing vertices (0: the result is correct). */
/* The context's current BFS / CC / SSSP result. */
int wg_validate(wg_ctx* ctx, uint64_t* violations);
/* A result computed elsewhere (app WG_APP_BFS / CC / SSSP), in the caller's
 * numbering and the layouts of wg_read_bfs / wg_read_cc / wg_read_sssp, with
 * its roots (BFS / SSSP). */
int wg_validate_values(wg_ctx* ctx, int app, const uint64_t* roots, uint32_t nroots, const void* values,
                       uint64_t* violations);
/* wg_validate without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Its bitmaps and
 * counter live in a fixed device array of 2 * WG_NOALLOC_MAX_VERTICES / 8 +
 * 256 bytes (32 MiB) that each GPU the library runs on holds until the
 * process ends.  Graphs with more than WG_NOALLOC_MAX_VERTICES vertices are
 * rejected with WG_E_RANGE.  Every call on one GPU shares that GPU's array,
 * and each call clears what it uses first: calls that overlap on one GPU (from
 * several threads or contexts) race and must be serialised by the caller. */
#define WG_NOALLOC_MAX_VERTICES (1ull << 27)
int wg_validate_noalloc(wg_ctx* ctx, uint64_t* violations);

/* ---- batches of independent graphs (no reference counterpart) ------------
 * A batch of small graphs is loaded as one graph, their disjoint union:
 * member g owns the vertices [first[g], first[g+1]) and every edge stays
 * inside its member.  Load the union with WG_OPT_RELABEL = WG_RELABEL_OFF so
 * that the members keep their vertex ranges (a single member, n = 1, may be
 * relabelled).  Then
 *  - BFS / SSSP from one root per member (wg_init_bfs_roots /
 *    wg_init_sssp_roots + wg_converge) and CC (wg_init_cc + wg_converge) run
 *    every member in one traversal, with results equal to per-member runs
 *    shifted by first[g];
 *  - wg_batch_pr runs the reference's PageRank (source/phase_edge_pull_pr.c,
 *    source/phase_vertex_pr.c:26-77) on every member at once, each with its
 *    own N = member size and S = sum of the member's accumulators. */
int wg_set_members(wg_ctx* ctx, uint32_t n, const uint64_t* first /* n + 1 entries */);

/* A launch shape of the batched PageRank: edges per lane of a tile (a tile
 * holds up to 64 * edges_per_lane edges; 4, 8 or 16, and for the single-graph
 * PageRank also 2), workgroups per compute
 * unit of the tile kernel (1..16; for the single-graph PageRank, 1 means one
 * workgroup of 16 wavefronts per CU with a 64 KB LDS contribution cache,
 * the others workgroups of 4 wavefronts), and xcd_split: 1 = each of the 8 XCDs
 * walks its own contiguous eighth of the edges (so a member's contributions
 * stay in one XCD's L2), 0 = all workgroups stride over all tiles. */
typedef struct wg_batch_shape {
    uint32_t edges_per_lane;
    uint32_t blocks_per_cu;
    uint32_t xcd_split;
} wg_batch_shape;
#define WG_BATCH_NSHAPES 24
/* The autotuner's candidate number i (0 <= i < WG_BATCH_NSHAPES). */
int wg_batch_shape_at(uint32_t i, wg_batch_shape* out);

/* `iterations` PageRank iterations on every member from the uniform start.
 * shape: NULL = the tuned shape (wg_batch_autotune) or {8, 8, 0}.  use_graph: 1 =
 * all iterations are captured once into a HIP graph and replayed as one
 * launch (the graph is kept and reused while iterations, damping, precision
 * and shape stay the same); 0 = plain launches.  Both give bit-identical
 * ranks.  Returns when the work is queued: wg_sync or wg_read_batch_ranks
 * waits for it. */
int wg_batch_pr(wg_ctx* ctx, uint32_t iterations, double damping, int contrib_fp32, const wg_batch_shape* shape,
                int use_graph);
/* Times every candidate shape (plain launches of `iterations` iterations,
 * best of three after a warm-up), keeps the fastest for shape == NULL and
 * returns it; ms_per_shape (WG_BATCH_NSHAPES entries, optional) receives each
 * candidate's time.  The next wg_batch_pr with use_graph = 1 captures the
 * chosen shape. */
int wg_batch_autotune(wg_ctx* ctx, uint32_t iterations, double damping, int contrib_fp32, wg_batch_shape* best,
                      float* ms_per_shape);
/* Ranks of the last wg_batch_pr (V entries, FP64; each member sums to ~1). */
int wg_read_batch_ranks(wg_ctx* ctx, double* host_ranks);

/* The tile shape of the single-graph PageRank (wg_pr_iteration,
 * wg_edge_pull_pr): the same candidates as the batch.  NULL restores the
 * default {8, 8, 0}.  The shape stays until the next graph load; results do
 * not depend on it beyond FP64 summation order (every tile shape sums in a
 * fixed order).
 * xcd_split == WG_PR_BINNED selects the binned pull instead of edge tiles
 * (only with edges_per_lane 8): sources in windows whose contributions fill
 * one workgroup's LDS, destinations in bins whose FP64 accumulators do, one
 * streaming pass that sums each window's edges per destination and one that
 * adds those sums per bin and runs the vertex phase.  blocks_per_cu 1: 128 KiB
 * windows and bins (one 1024-thread workgroup per CU); 2: 64 KiB.  Its layout
 * is built on the first iteration after a change of graph, shape or
 * precision (about 2 bytes per edge plus 10 per window-destination pair, and
 * one FP64 sum per pair while iterating).  Its per-bin sums use LDS atomics,
 * so results are reproducible to rounding, not bit for bit. */
#define WG_PR_BINNED 2
int wg_set_pr_shape(wg_ctx* ctx, const wg_batch_shape* shape);
int wg_get_pr_shape(const wg_ctx* ctx, wg_batch_shape* shape);
/* The single-graph PageRank autotuner's candidate number i
 * (0 <= i < WG_PR_NSHAPES): the batch's WG_BATCH_NSHAPES shapes, then
 * {4|8|16, 1, 0} (one 16-wavefront workgroup per CU), then 128-edge tiles
 * {2, 1, 0}, {2, 4, 0}, {2, 16, 0} (2 edges per lane), then the binned pull
 * {8, 1, WG_PR_BINNED}, {8, 2, WG_PR_BINNED}. */
#define WG_PR_NSHAPES (WG_BATCH_NSHAPES + 8)
int wg_pr_shape_at(uint32_t i, wg_batch_shape* out);
/* Size of the binned pull's layout for the current graph, precision and
 * shape (built if needed; WG_E_STATE unless PageRank is initialised with a
 * binned shape): window-destination pairs, code entries (edges plus padding),
 * push workgroups, pull workgroups. */
int wg_pr_bins_info(wg_ctx* ctx, uint64_t* pairs, uint64_t* code_entries, uint32_t* push_groups,
                    uint32_t* pull_groups);
/* After wg_init_pr: times `iterations` fused iterations per candidate (best of
 * two after a warm-up), keeps the fastest shape, and initialises PageRank
 * again with the same damping and precision.  ms_per_shape (optional,
 * WG_PR_NSHAPES entries) receives the time per iteration of each candidate. */
int wg_pr_autotune(wg_ctx* ctx, uint32_t iterations, wg_batch_shape* best, float* ms_per_shape);

/* BFS result: packed[v] = (depth << 32) | parent, UINT64_MAX if unreached;
 * the root holds (0 << 32) | root. */
int wg_read_bfs(wg_ctx* ctx, uint64_t* host_packed);
/* CC result: uint32 minimum label per vertex. */
int wg_read_cc(wg_ctx* ctx, uint32_t* host_labels);
/* SSSP result: double (F64 graphs, +inf unreached) or uint32 (U32 graphs,
 * 0xFFFFFFFF unreached). */
int wg_read_sssp(wg_ctx* ctx, void* host_dist);

/* Activity counters of the current state: vertices reached (BFS: visited;
 * SSSP: finite distance) and the sum of their out-degrees. */
int wg_reached(wg_ctx* ctx, uint64_t* vertices, uint64_t* edges);
/* Work of the BFS / CC / SSSP pulls since the application's init (no
 * reference counterpart; the measure of a traversal's minimum traffic):
 * edges_read = edges whose source a pull read (an in-list entry, a marked
 * block's edge, a pushed out-edge), dests_examined = destination states a
 * pull examined (a destination group of a dense pull, a run of one
 * destination's edges in a marked block, the destination of a pushed
 * out-edge).  WG_E_STATE unless WG_OPT_COUNT_WORK is 1. */
int wg_work_counts(wg_ctx* ctx, uint64_t* edges_read, uint64_t* dests_examined);

/* ---- learned schedule --------------------------------------------------------
 * No reference counterpart: replaces the fixed threshold rule of
 * source/execution_bfs.c:141 (and its cc/sssp twins).  An FP32 MLP maps
 * WG_MLP_IN features of the state before a pull to the predicted log2 time of
 * each of WG_MLP_ACTIONS pull choices, as a learned correction to an analytic
 * log2 work estimate of each choice; the cheapest is taken.  It only picks
 * schedules, so results are the same whatever it predicts.
 *
 * Features (wg_mlp_features in wedge_amd/csrc/wg_mlp_math.h) are scale-free,
 * with L(x) = log2(1 + x), fv / fe the vertices / out-edges activated by the
 * last pull and u the vertices not reached yet (BFS, SSSP): L(fv) - L(V),
 * L(fe) - L(E), L(u) - L(V), L(E) - L(V), L(fe) - L(fv), L(iteration) - L(V)/2,
 * app is BFS, app is CC, app is SSSP, L(fe) - L(V), L(fv) - L(E), L(u) - L(E)
 * (the u terms are 0 for CC).
 * The plain rule's choice stands unless the model predicts the cheapest
 * action to be faster than it by more than a factor 2^margin[app] (margins
 * for BFS, CC, SSSP are part of the model).
 * Actions: 0 dense; 1..5 frontier at block_log2 3, 4, 5, 6, 8.  BFS may take
 * any action; CC and SSSP only 0 and 2 (dense, or frontier at the default
 * granularity): their iteration count depends on the granularity, which a
 * per-pull cost does not see, while a BFS level is the same in every mode.
 * Parameter layout (floats): lo[IN], hi[IN] (inputs are clamped to the
 * training range), mean[IN], inv_std[IN], W1[IN][H], b1[H], W2[H][H], b2[H],
 * W3[H][A], b3[A], margin[3], lv_max; W[i][j] is input i to unit j.  CC and
 * SSSP keep the plain rule on graphs with log2(1 + V) > lv_max (larger than
 * the training graphs). */
#define WG_MLP_IN 12
#define WG_MLP_HIDDEN 64
#define WG_MLP_ACTIONS 6
#define WG_MLP_PARAMS                                                                                    \
    (4 * WG_MLP_IN + WG_MLP_IN * WG_MLP_HIDDEN + WG_MLP_HIDDEN + WG_MLP_HIDDEN * WG_MLP_HIDDEN + \
     WG_MLP_HIDDEN + WG_MLP_HIDDEN * WG_MLP_ACTIONS + WG_MLP_ACTIONS + 4)

/* Mode and block granularity of action a (WG_E_INVAL if out of range). */
int wg_mlp_action(int action, int* mode, int* block_log2);

/* Sets (params != NULL, count == WG_MLP_PARAMS) or clears (params == NULL)
 * the model the device loop uses for schedule 1. */
int wg_set_schedule_model(wg_ctx* ctx, const float* params, uint64_t count);

/* Batch inference with the context's model and the device decision code:
 * features[n][WG_MLP_IN], work estimates base[n][WG_MLP_ACTIONS] (may be
 * NULL: 0) and the plain rule's actions plain[n] (may be NULL: none) ->
 * predicted log2 times out[n][WG_MLP_ACTIONS] (may be NULL) and the chosen
 * actions (may be NULL). */
int wg_mlp_infer(wg_ctx* ctx, int app, const float* features, const float* base, const int32_t* plain, uint64_t n,
                 float* out, int32_t* actions);

/* Training on the device (hand-written kernels; minibatch Adam on the squared
 * error of the taken action's output, fixed-order gradient sums, so a run is
 * reproducible).  samples: features[n][IN], the action taken and the target
 * of its network output (measured log2 time minus the work estimate).
 * params: in = initial parameters including lo/hi/mean/inv_std, the
 * margins and lv_max (none of which is trained), out = trained.  final_loss: mean
 * squared error over all samples after training. */
typedef struct wg_mlp_train_opts {
    uint32_t epochs;
    uint32_t batch;     /* samples per step */
    float lr;           /* Adam step size */
    float beta1, beta2, eps;
    uint64_t seed;      /* minibatch order */
} wg_mlp_train_opts;
int wg_mlp_train(wg_ctx* ctx, const float* features, const int32_t* actions, const float* targets, uint64_t n,
                 const wg_mlp_train_opts* opts, float* params, float* final_loss);

/* ---- measurement ---------------------------------------------------------- */

/* Enables per-launch HIP-event timing of every kernel on the context stream. */
int wg_profile_enable(wg_ctx* ctx, int on);
/* Launch count and summed device time of kernel `kernel_id` since enabling
 * or the last reset (waits for the stream). */
int wg_profile_get(wg_ctx* ctx, int kernel_id, uint64_t* launches, double* total_ms);
int wg_profile_reset(wg_ctx* ctx);
/* Device-to-device streaming copy of `bytes` (read + write), best of five
 * after a warm-up: the bandwidth an HBM-bound kernel reaches on this device,
 * reported next to the nominal 8 TB/s. */
int wg_measure_copy(wg_ctx* ctx, uint64_t bytes, double* gb_per_s);

#ifdef __cplusplus
}
#endif

#endif /* WEDGE_GPU_H */
