// wg_internal.h -- context layout and helpers shared by the HIP translation units.
//
// Device layout of a loaded graph (all 32-bit indices, one HBM pool):
//   off[V+1]        CSC offsets by destination
//   src[E]          source of each edge, ascending within a destination
//   dst[E]          destination of each edge (read only by frontier-mode pulls,
//                   which touch only the marked edge blocks)
//   w[E]            weights: double (SSSP FP64) or uint32 (SSSP integer)
//   od[V]           out-degree
//   csri_off[V+1],  source -> positions of its out-edges in src[]
//   csri_pos[E]     (the reference's CSRI, source/graphdata.c:493-585, but holding
//                   edge positions so any block granularity can be derived)
//   tile_v[T+1],    edge tiles for the dense pulls, one wavefront per tile: a
//   tile_e[T+1]     light tile owns whole destinations [tile_v[t], tile_v[t+1])
//                   with <= TILE_EDGES edges and <= TILE_VERTS vertices; a
//                   destination with more than TILE_EDGES in-edges is cut into
//                   heavy tiles of TILE_EDGES edges whose tile_v word is that
//                   vertex with WG_HEAVY_BIT set
//   heavy_v[H], heavy_t[H]  heavy destinations and their first tile
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <vector>

#include "wedge_gpu.h"

#define WG_WAVE 64
#define WG_BLOCK 256
#define WG_WPB (WG_BLOCK / WG_WAVE)           // wavefronts per block
#define WG_IPT 8                              // edges per lane in a tile
#define WG_TILE_EDGES (WG_WAVE * WG_IPT)      // 512 edges per wavefront tile
#define WG_TILE_VERTS 256                     // destinations per light tile
#define WG_HEAVY_BIT 0x80000000u              // tile_v flag: heavy tile
#define WG_BFS_SEQ 16                         // edges a lane scans alone in dense BFS
#define WG_BFS_HEAVY 2048                     // in-degree above which dense BFS splits a vertex
#define WG_PUSH_ROOTS_MAX 4096u               // roots up to which wg_converge pushes the first BFS level

// Counts that many workgroups add to are spread over WG_CSLOTS slots, one
// 256-byte line each: atomics from many workgroups on one address serialise
// in one L2 channel (about 13 ns each on gfx950), while different lines are
// served in parallel.  A fold (one wavefront) adds the slots into the scalar
// fields below and zeroes them.
#define WG_CSLOTS 32
enum { WG_CNT_ACTIVITY = 0, WG_CNT_ACTIVATED, WG_CNT_FRESH, WG_CNT_FRONTIER, WG_CNT_REACHED_V, WG_CNT_REACHED_E,
       WG_CNT_EDGES_READ, WG_CNT_DESTS, WG_CNT_N };
struct __attribute__((aligned(256))) wg_count_slot {
    unsigned long long v[WG_CNT_N];
    unsigned long long pad[32 - WG_CNT_N];
};

// Device-side counters, one struct per context.  The host mirror copies the
// fields before 'slots' only.
struct wg_dev_counters {
    unsigned long long activity;        // sum of out-degrees of vertices activated this pull
    unsigned long long activated;       // number of vertices activated this pull
    unsigned long long fresh;           // of which reached for the first time (BFS, SSSP)
    unsigned long long frontier_verts;  // popcount of the frontier made current by advance
    unsigned int run_err;               // a fused-level run gave up waiting at its grid barrier
    unsigned int heavy_len;             // high out-degree vertices deferred by the frontier build
    unsigned int overflow;              // integer SSSP distance reached 2^32-1
    unsigned int ticket;                // last-block-done ticket (PR span kernel)
    double S[2];                        // PR: sum of ranks of vertices with out-degree > 0
    double total_rank;                  // PR: sum of accumulators of the last iteration
    unsigned long long reached_verts;
    unsigned long long reached_edges;
    // work of the BFS / CC / SSSP pulls since their init (wg_work_counts):
    unsigned long long edges_read;      // edges whose source the pulls read
    unsigned long long dests_examined;  // destination states the pulls examined
    // control block of the device-resident iteration loop (wg_converge)
    unsigned int mode;                  // pull mode of the next iteration
    int blog2;                          // its Wedge granularity
    unsigned int done;                  // converged (or iteration cap reached)
    unsigned int iters;                 // iterations completed
    unsigned int fiters;                // of which frontier-mode
    unsigned int level;                 // BFS: depth of the current frontier
    unsigned int sched;                 // schedule kind (wg_loop_params.schedule)
    int plain_blog2;                    // granularity the plain rule uses
    unsigned int max_iters;             // 0 = no cap
    unsigned int run_levels;            // iterations run by fused small-frontier runs
    unsigned long long threshold;       // plain rule: dense iff activity > threshold
    unsigned long long last_activity;   // activity of the last completed pull
    unsigned long long last_activated;
    unsigned long long unvisited;       // vertices not reached yet (BFS, SSSP): learned-rule feature
    unsigned int app;                   // features of the learned rule: application,
    unsigned int min_binned;            // dense CC/SSSP pulls run on the binned layout (learned rule)
    unsigned long long num_vertices;    // graph size
    unsigned long long num_edges;
    wg_count_slot slots[WG_CSLOTS];
};
#define WG_CTR_HOST_BYTES offsetof(wg_dev_counters, slots)

// Gate of a kernel launched by the device-resident loop: it runs only when the
// control block selects it.  WG_GATE_ALWAYS: launched by the phase API, runs.
#define WG_GATE_ALWAYS (-1)
#define WG_GATE_DENSE 0     // runs if the selected pull mode is dense
#define WG_GATE_FRONTIER 1  // runs if the selected pull mode is frontier
#define WG_GATE_LIVE 2      // runs until the loop has converged

__device__ __forceinline__ bool wg_gated_off(const wg_dev_counters* c, int gate)
{
    if (gate < 0)
        return false;
    if (c->done)
        return true;
    if (gate == WG_GATE_LIVE)
        return false;
    return c->mode != (unsigned int)gate;
}

// Fused small-frontier levels (k_bfs_run, k_min_run): a queue entry is up to
// WG_RUN_CH out-edges of one vertex.  A run of G workgroups takes levels that
// activate at most WG_RUN_NACT_PER_WG * G vertices with at most
// WG_RUN_ACT_PER_WG * G out-edges: at most 2048 entries (8 rounds of its
// threads) per workgroup, so a queue of WG_RUN_QCAP entries holds any level a
// run of WG_RUN_MAX_WGS workgroups continues with.
#define WG_RUN_CH 8u
#define WG_RUN_NACT_PER_WG 1024ull
#define WG_RUN_ACT_PER_WG 8192ull
#define WG_RUN_MAX_WGS 256u
#define WG_RUN_QCAP (1u << 19)
#define WG_RUN_DEFAULT_WGS 32u

struct __attribute__((aligned(256))) wg_run_line {
    unsigned long long v;
    unsigned long long pad[31];
};
// Control block of a run: its grid barrier and, by level mod 3, the queue
// tails and the level's counts (one 256-byte line each).
struct wg_run_ctl {
    wg_run_line count;   // barrier arrivals
    wg_run_line gen;     // barrier generation
    wg_run_line tail[3];
    wg_run_line act[3];
    wg_run_line nact[3];
    wg_run_line fresh[3];
};

struct wg_prof {
    int enabled;
    uint64_t launches[WG_K_COUNT];
    double total_ms[WG_K_COUNT];
    // pending event pairs
    int npending;
    int cap;
    hipEvent_t* ev;   // 2 * cap
    int* kid;         // cap
};

// Edge tiles of one tile shape (see build_tiles in wg_ctx.hip).
struct wg_tiling {
    uint32_t ntiles, nheavy;
    uint32_t *tile_v, *tile_e, *heavy_v, *heavy_t;
    uint32_t split[9];        // edge-balanced tile ranges of the 8 XCDs
};

struct wg_ctx {
    int device;
    hipStream_t stream;
    int last_hip_error;
    int num_cu;

    // graph
    uint64_t V, E;
    int wkind;
    uint32_t* off;
    uint32_t* src;
    uint32_t* dst;
    void* w;
    uint32_t* od;
    uint32_t* csri_off;
    uint32_t* csri_pos;
    uint32_t ntiles;
    uint32_t* tile_v;
    uint32_t* tile_e;
    uint32_t nheavy;
    uint32_t* heavy_v;
    uint32_t* heavy_t;
    uint32_t nbfs_heavy;      // dense BFS chunks of vertices with in-degree > WG_BFS_HEAVY
    uint32_t* first_src;      // dense BFS: smallest in-neighbour of each vertex, UINT32_MAX if it has
                              // none or more than WG_BFS_HEAVY (built by the first BFS init)
    uint32_t* csr_dst;        // BFS: destination of out-edge position p, dst[csri_pos[p]] (E entries,
                              // built by the first BFS init)
    uint32_t* bfs_heavy_v;
    uint32_t* bfs_heavy_e;    // chunk start edge (chunk end = min(start + WG_BFS_HEAVY, off[v+1]))
    uint32_t n_out_pos;       // vertices with out-degree > 0
    int relabel;              // option: store vertices in descending-degree order
    uint32_t* perm;           // device id -> original id (NULL when not relabelled)
    uint32_t* inv;            // original id -> device id (NULL when not relabelled)

    // application state
    int app;
    void* prop[2];            // PR contributions (double buffer) / BFS packed / CC labels / SSSP dist
    void* ms[4];              // batched sources: per-vertex masks or rows (freed with the app state)
    uint32_t ms_n;            // batched sources: number of roots / seeds
    bool ms_nl_clean;         // 64-source BFS: nl is all zero (the last run completed)
    size_t ms_aux_bytes;      // personalised PageRank: size of ms[2] (its reduction partials)
    void* idmap;              // wg_to_device_ids: original ids in, device ids out (grown as needed)
    size_t idmap_bytes;
    int pr_cur;               // PR: index of the current contribution buffer
    int pr_parity;            // PR: index of S[] holding the current iteration's sum
    int pr_fp32;
    double damping;
    double* acc;              // PR accumulators (stand-alone edge phase)
    double* tile_part;        // PR: per-tile partial sums of heavy destinations
    double* blk_part;         // PR: per-block partial sums for S
    uint32_t pr_grid;         // blocks of the tile kernel
    uint32_t span_grid;       // blocks of the span kernel
    // PageRank tile shape (wg_set_pr_shape / wg_pr_autotune): edges per lane
    // 4, 8 (the tiles above) or 16, workgroups per CU, XCD split
    int pr_ipt, pr_bpc, pr_xcd;
    wg_tiling pr_alt[3];      // tilings with 4, 16 and 2 edges per lane (built on first use)
    int pr_alt_built;
    uint32_t pr_split8[9];    // XCD split of the default tiles
    struct wg_pb* pb;         // layout of the binned PageRank pull (wg_pb.hip), built on first use
    uint32_t pb_slice;        // option: pairs per pull workgroup of that layout (0 = default)
    int ms_binned;            // option WG_OPT_MS_BINNED + 1 (0 = unset: the default)
    int min_binned;           // option WG_OPT_MIN_BINNED + 1 (0 = unset: the default)
    int count_work;           // option WG_OPT_COUNT_WORK
    uint64_t bm_words;        // 64-bit words per vertex bitmask
    uint64_t* bm_cur;         // BFS: current frontier
    uint64_t* bm_next;        // activated by the running pull
    uint64_t* bm_visited;     // BFS: visited before this level
    uint32_t* wedge;          // Wedge frontier, one bit per edge block
    uint64_t wedge_words;     // 32-bit words allocated (finest granularity)
    int wedge_log2;           // granularity of the marked blocks, -1 if none are marked
    uint32_t* heavy_list;     // frontier build: deferred out-edge chunks (vertex, first position)
    uint32_t heavy_cap;       // chunks heavy_list holds
    float* mlp;               // schedule model parameters (WG_MLP_PARAMS floats)
    int mlp_set;              // a model is set for the learned schedule
    wg_dev_counters* ctr;     // device
    wg_dev_counters* hctr;    // pinned host mirror
    wg_dev_counters* hctr_dev; // the mirror's device address (null if it has none)
    uint32_t root_dev;        // BFS / SSSP root in device numbering (the first root)
    uint32_t* roots;          // BFS / SSSP roots in device numbering, distinct
    uint32_t* root_off;       // BFS, at most WG_PUSH_ROOTS_MAX roots: exclusive prefix of the roots'
                              // out-degrees (nroots + 1 entries), for the first level's push
    uint32_t nroots;
    unsigned long long root_activity; // summed out-degree of the roots
    int roots_frontier;       // the BFS frontier is still the roots (no pull or advance since the init)
    uint32_t push_roots;      // roots the running wg_converge pushes on its first level (0: none)
    struct wg_members* members;      // batch of member graphs (wg_set_members), NULL if none
    // fused small-frontier levels of the device loop (k_bfs_run)
    uint32_t run_wgs[8];      // workgroups of a run, by WG_APP_*, 0 = off
    struct wg_run_ctl* run_ctl;
    uint4* run_q;             // two queues of WG_RUN_QCAP entries
    uint64_t* run_spare;      // CC / SSSP runs: second activation mask (bm_words), zero between runs
    int iter;                 // iterations run since init

    wg_prof prof;
};

// ---- error plumbing --------------------------------------------------------

#define WG_HIP(ctx, call)                                   \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) {                             \
            if (ctx)                                        \
                (ctx)->last_hip_error = (int)_e;            \
            return WG_E_HIP;                                \
        }                                                   \
    } while (0)

// Launch bracket for per-kernel timing.
int wg_prof_begin(wg_ctx* ctx, int kid);
int wg_prof_end(wg_ctx* ctx, int kid, int slot);

static inline int wg_check_launch(wg_ctx* ctx)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->last_hip_error = (int)e;
        return WG_E_HIP;
    }
    return WG_OK;
}

static inline uint32_t wg_div_up(uint64_t a, uint64_t b)
{
    return (uint32_t)((a + b - 1) / b);
}

// Number of blocks for a grid-stride kernel: enough to fill the chip several
// times over (256 CUs in 8 XCDs) but no more than the work needs.
static inline uint32_t wg_grid(const wg_ctx* ctx, uint64_t work_items, uint32_t items_per_block, uint32_t per_cu)
{
    uint64_t need = (work_items + items_per_block - 1) / items_per_block;
    uint64_t cap = (uint64_t)ctx->num_cu * per_cu;
    if (need > cap)
        need = cap;
    if (need < 1)
        need = 1;
    return (uint32_t)need;
}

// ---- device helpers ---------------------------------------------------------

__device__ __forceinline__ int wg_lane() { return (int)(threadIdx.x & (WG_WAVE - 1)); }

__device__ __forceinline__ unsigned long long wg_ballot(int pred)
{
    return __ballot(pred);
}

template <typename T>
__device__ __forceinline__ T wg_wave_sum(T v)
{
    for (int o = WG_WAVE / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, WG_WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wg_wave_min(T v)
{
    for (int o = WG_WAVE / 2; o > 0; o >>= 1) {
        const T x = __shfl_xor(v, o, WG_WAVE);
        v = x < v ? x : v;
    }
    return v;
}

// Orders this wavefront's LDS writes before its following LDS reads by other
// lanes (no block-wide barrier: wavefront tiles are independent).
__device__ __forceinline__ void wg_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// internal entry points shared across translation units
int wg_reset_counters(wg_ctx* ctx, unsigned int flags); // RST_* flags (wg_frontier.h)
int wg_read_counters(wg_ctx* ctx);                      // fold the slots, copy to ctx->hctr, wait
// fold the slots, publish activity / activated / fresh to ctx->hctr and clear
// them on the device for the next level (one launch), wait
int wg_take_activity(wg_ctx* ctx);
int wg_launch_wedge_build(wg_ctx* ctx, const uint64_t* active, int block_log2);
uint32_t wg_frontier_grid(const wg_ctx* ctx);
uint32_t wg_wedge_build_grid(const wg_ctx* ctx);
uint32_t wg_advance_grid(const wg_ctx* ctx);
int wg_pr_reset(wg_ctx* ctx);
int wg_free_app_state(wg_ctx* ctx);
int wg_alloc_bitmasks(wg_ctx* ctx);
int wg_clear_u64(wg_ctx* ctx, uint64_t* p, uint64_t words);
int wg_fill_u32(wg_ctx* ctx, uint32_t* p, uint64_t n, uint32_t value);
int wg_fill_u64(wg_ctx* ctx, uint64_t* p, uint64_t n, uint64_t value);
// original vertex id -> device vertex id (identity without relabelling)
int wg_to_device_id(wg_ctx* ctx, uint64_t orig, uint32_t* out);
// the same for n ids with one round trip to the device (WG_E_RANGE if any is out of range)
int wg_to_device_ids(wg_ctx* ctx, const uint64_t* orig, uint32_t n, uint32_t* out);
// copies V values stored in device order to host memory in original order
int wg_read_vertex_array(wg_ctx* ctx, const void* dev, void* host, size_t elem_size);
// edge tiles of at most tile_edges edges (and WG_TILE_VERTS destinations)
void wg_build_tiles(const uint32_t* off, uint64_t V, uint32_t tile_edges, std::vector<uint32_t>& tv,
                    std::vector<uint32_t>& te, std::vector<uint32_t>& hv, std::vector<uint32_t>& ht);
// releases the member batch (with its HIP graph); called when the graph goes
void wg_free_members(wg_ctx* ctx);
// releases the loaded graph and everything built on it
void wg_drop_graph(wg_ctx* ctx);
// XCD split points of a tiling (first tile at or after x/8 of the edges)
void wg_tile_split_points(const std::vector<uint32_t>& te, uint64_t E, uint32_t start[9]);
// builds the 4- and 16-edges-per-lane PageRank tilings once per graph
int wg_build_pr_tilings(wg_ctx* ctx);
// binned PageRank pull (wg_pb.hip): one iteration (mode 0 fused, 1 edge phase
// only), building the layout for the current precision and shape first if needed
int wg_pb_iteration(wg_ctx* ctx, int mode, const void* cin, void* cout, double* S_in, double* S_out);
int wg_pb_prepare(wg_ctx* ctx);
int wg_pb_stats(const wg_ctx* ctx, uint64_t* pairs, uint64_t* code_entries, uint32_t* tasks, uint32_t* slices);
// a dense level of the 64-source BFS on the binned layout, committed: with
// n[v] = (OR of the in-neighbours' front_in & live) & ~seen[v], seen |= n,
// front_out = n (front_out may be front_in), the depth entries of the new
// sources = level when rows is set (else the caller writes them from
// front_out), the frontier vertex bitmask (bm_next) and the level's counts in
// ctx->ctr
// dense CC / SSSP pulls on the binned layout (WG_OPT_MIN_BINNED): whether the
// context's graph and application use them, the layout (with the weights in
// code order) built ahead of a loop, and one dense pull over ctx->prop[0]
// (push and pull kernels gated by `gate`, as the tile pull they replace)
bool wg_pb_min_on(const wg_ctx* ctx);
int wg_pb_min_prepare(wg_ctx* ctx);
int wg_pb_min_pull(wg_ctx* ctx, int gate);
int wg_pb_msbfs_level(wg_ctx* ctx, const unsigned long long* front_in, unsigned long long* front_out,
                      unsigned long long* seen, uint16_t* depth, unsigned long long live, uint32_t level, bool rows);
// releases the binned layout (graph change)
void wg_free_pb(wg_ctx* ctx);
