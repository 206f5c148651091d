// wg_batch.hip -- many BFS sources in one traversal (no counterpart in the
// reference, whose BFS takes one root: source/execution_bfs.c:84-118).
//
// Up to 64 sources share every pull: lane s of a vertex's 64-bit masks
// belongs to source s.
//   seen[v]   sources that have reached v
//   front[v]  sources that reached v at the last level (the frontier)
//   nl[v]     sources reaching v at the running level (filled by the pull)
//   save[k][v]  the new sources of level k (k < 16): each level's frontier
//                 masks are kept, and the depth rows written from them once
//                 per vertex after the run (k_ms_rows); levels past the 16th
//                 use front and write their rows as they commit
//   depth[v][64]  uint16 level at which source s reached v, valid where
//                 seen[v] has bit s (vertex-major: a vertex's row is written
//                 by one store; reads mask entries with seen, so the matrix
//                 needs no clearing between runs)
// A dense level is the bottom-up pull of source/phase_edge_pull_bfs.c:109-210
// for all sources at once: a vertex ORs the frontier masks of its in-edges,
// restricted to the sources that have not reached it.  A frontier level walks the Wedge frontier
// built from the vertices with any new source (source/wedge.c:63-155).  The
// schedule per level is the plain rule on the summed out-degree of those
// vertices.  Each source's depths equal those of a single-source BFS.
#include "wg_frontier.h"

#include <algorithm>
#include <vector>

#define MS_LEVEL_MAX 0xFFFEu
#define MS_SAVE 16u // levels per run whose new-source masks are kept for k_ms_rows (slot k: level k)
#define MS_ROW_PAD 9u // uint4 per staged depth row in LDS (8 used, one of padding against bank conflicts)

__device__ __forceinline__ unsigned long long wave_or64(unsigned long long x)
{
    for (int o = WG_WAVE / 2; o > 0; o >>= 1)
        x |= (unsigned long long)__shfl_xor((long long)x, o, WG_WAVE);
    return x;
}

// Roots: source i starts at roots[i]; the activity of the root level is the
// summed out-degree of the distinct roots.
__global__ void k_ms_roots(const uint32_t* __restrict__ roots, uint32_t n, unsigned long long* __restrict__ seen,
                           unsigned long long* __restrict__ front, uint16_t* __restrict__ depth,
                           unsigned long long* __restrict__ bm, const uint32_t* __restrict__ od,
                           wg_dev_counters* __restrict__ ctr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t r = roots[i];
    atomicOr(&seen[r], 1ull << i);
    atomicOr(&front[r], 1ull << i);
    const unsigned long long bit = 1ull << (r & 63);
    if (!(atomicOr(&bm[r >> 6], bit) & bit))
        atomicAdd(&ctr->slots[0].v[WG_CNT_ACTIVITY], (unsigned long long)od[r]);
    depth[(uint64_t)r * WG_MS_MAX_SOURCES + i] = 0;
}

// Dense level: one lane per destination; in-edge lists longer than WG_BFS_SEQ
// are finished by the whole wavefront, 64 edges per step.
// Frontier mask of u.  While the frontier is small (TEST), the vertex bitmask
// (V/8 bytes, cache-resident) is tested first and the mask (8V bytes) read only
// for frontier vertices; once most vertices are in it, the extra random access
// costs more than it saves.
template <bool TEST>
__device__ __forceinline__ unsigned long long front_of(const uint64_t* __restrict__ fbm,
                                                       const unsigned long long* __restrict__ front, uint32_t u)
{
    if (TEST)
        return ((fbm[u >> 6] >> (u & 63)) & 1ull) ? front[u] : 0ull;
    return front[u];
}

// Dense level over the edge tiles (the PageRank tiles: 512 consecutive edges
// per wavefront, whole destinations per light tile).  The sources a
// destination still needs are staged in LDS first; a tile none of whose
// destinations needs anything is skipped before its edges are read.  Each
// lane ORs the frontier masks of its 8 edges per destination run into the
// destination's LDS word; the light tile owns its destinations and stores
// their new sources, a heavy tile ORs its part into nl with one atomic.
template <bool TEST>
__global__ __launch_bounds__(WG_BLOCK) void k_ms_tile(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                       const uint64_t* __restrict__ fbm,
                                                       const unsigned long long* __restrict__ front,
                                                       const unsigned long long* __restrict__ seen,
                                                       unsigned long long* __restrict__ nl,
                                                       const uint32_t* __restrict__ tile_v,
                                                       const uint32_t* __restrict__ tile_e, uint32_t ntiles,
                                                       unsigned long long live)
{
    constexpr int NOFF = (WG_TILE_VERTS + 1 + WG_WAVE - 1) / WG_WAVE;
    constexpr int NVTX = WG_TILE_VERTS / WG_WAVE;
    __shared__ uint32_t s_off_all[WG_WPB][WG_TILE_VERTS + 1];
    __shared__ unsigned long long s_need_all[WG_WPB][WG_TILE_VERTS];
    __shared__ unsigned long long s_acc_all[WG_WPB][WG_TILE_VERTS];
    const int wave = threadIdx.x / WG_WAVE;
    const int lane = wg_lane();
    uint32_t* s_off = s_off_all[wave];
    unsigned long long* s_need = s_need_all[wave];
    unsigned long long* s_acc = s_acc_all[wave];
    const uint32_t stride = gridDim.x * WG_WPB;
    for (uint32_t t = blockIdx.x * WG_WPB + wave; t < ntiles; t += stride) {
        const uint32_t tv = tile_v[t];
        const uint32_t e0 = tile_e[t], e1 = tile_e[t + 1];
        const uint32_t v0 = tv & ~WG_HEAVY_BIT;
        const uint32_t a = e0 + (uint32_t)lane * WG_IPT;
        if (tv & WG_HEAVY_BIT) {
            const unsigned long long need = live & ~seen[v0];
            if (!need)
                continue;
            unsigned long long acc = 0ull;
#pragma unroll
            for (int j = 0; j < WG_IPT; ++j)
                if (a + j < e1)
                    acc |= front_of<TEST>(fbm, front, src[a + j]) & need;
            acc = wave_or64(acc);
            if (lane == 0 && acc)
                atomicOr(&nl[v0], acc);
            continue;
        }
        const uint32_t nv = (tile_v[t + 1] & ~WG_HEAVY_BIT) - v0;
        bool any = false;
#pragma unroll
        for (int k = 0; k < NVTX; ++k) {
            const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
            if (i < nv) {
                const unsigned long long nd = live & ~seen[v0 + i];
                s_need[i] = nd;
                s_acc[i] = 0ull;
                any = any || nd != 0ull;
            }
        }
        if (wg_ballot(any) == 0ull)
            continue;
#pragma unroll
        for (int k = 0; k < NOFF; ++k) {
            const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
            if (i <= nv)
                s_off[i] = off[v0 + i];
        }
        wg_wave_sync();
        if (a < e1) {
            uint32_t lo = 0, hi = nv - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= a)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            // destinations of the lane's edges (LDS walk), then every needed
            // gather issued at once (independent loads), then the runs
            uint32_t u[WG_IPT];
            uint16_t dd[WG_IPT];
            unsigned long long f[WG_IPT];
            {
                uint32_t d = lo, dend = s_off[d + 1];
#pragma unroll
                for (int j = 0; j < WG_IPT; ++j) {
                    const uint32_t e = a + j;
                    u[j] = e < e1 ? src[e] : 0u;
                    if (e < e1 && e == dend) {
                        ++d;
                        while (d < nv && s_off[d + 1] == s_off[d])
                            ++d;
                        dend = (d < nv) ? s_off[d + 1] : 0xffffffffu;
                    }
                    dd[j] = (uint16_t)d;
                }
            }
#pragma unroll
            for (int j = 0; j < WG_IPT; ++j) {
                const unsigned long long nd = (a + j < e1) ? s_need[dd[j]] : 0ull;
                f[j] = nd ? (front_of<TEST>(fbm, front, u[j]) & nd) : 0ull;
            }
            unsigned long long run = 0ull;
#pragma unroll
            for (int j = 0; j < WG_IPT; ++j) {
                if (a + j < e1) {
                    run |= f[j];
                    const bool last = (j + 1 == WG_IPT) || (a + j + 1 >= e1) || dd[j + 1 < WG_IPT ? j + 1 : j] != dd[j];
                    if (last) {
                        if (run)
                            atomicOr(&s_acc[dd[j]], run);
                        run = 0ull;
                    }
                }
            }
        }
        wg_wave_sync();
#pragma unroll
        for (int k = 0; k < NVTX; ++k) {
            const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
            if (i < nv) {
                const unsigned long long acc = s_acc[i];
                if (acc)
                    nl[v0 + i] = acc;
            }
        }
        wg_wave_sync();
    }
}

// Frontier level over the marked edge blocks.
__global__ __launch_bounds__(WG_BLOCK) void k_ms_blocks(int blog2, uint32_t E, const uint32_t* __restrict__ src,
                                                         const uint32_t* __restrict__ dst, const uint64_t* __restrict__ fbm,
                                                         const unsigned long long* __restrict__ front,
                                                         const unsigned long long* __restrict__ seen,
                                                         unsigned long long* __restrict__ nl, uint32_t* __restrict__ wedge,
                                                         unsigned long long live)
{
    walk_marked_blocks(wedge, blog2, E, [&](uint32_t p, bool ok, bool) {
        if (ok) {
            const unsigned long long f = front_of<true>(fbm, front, src[p]) & live;
            if (f) {
                const uint32_t v = dst[p];
                const unsigned long long add = f & ~seen[v];
                if (add)
                    atomicOr(&nl[v], add);
            }
        }
    });
}

// End of a level: the new sources of each vertex become its frontier, join
// seen, and write the vertex's depth row; the frontier vertex bitmask is
// rewritten word by word for the next frontier build; counts: out-degree and
// number of vertices with any new source, and (vertex, source) pairs reached.
template <bool ROWS>
__global__ __launch_bounds__(WG_BLOCK) void k_ms_commit(unsigned long long* __restrict__ nl,
                                                         unsigned long long* __restrict__ seen,
                                                         unsigned long long* __restrict__ front,
                                                         uint16_t* __restrict__ depth, uint64_t* __restrict__ bm,
                                                         const uint32_t* __restrict__ od, uint32_t V, uint32_t level,
                                                         wg_dev_counters* __restrict__ ctr)
{
    const int lane = wg_lane();
    const uint64_t nwords = ((uint64_t)V + 63) / 64;
    const uint64_t nwaves = (uint64_t)gridDim.x * WG_WPB;
    act_tally t = { 0ull, 0ull, 0ull };
    for (uint64_t w = (uint64_t)blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE; w < nwords; w += nwaves) {
        const uint32_t v = (uint32_t)(w * 64 + lane);
        unsigned long long n = 0ull;
        if (v < V) {
            n = nl[v];
            if (n) {
                nl[v] = 0ull;
                seen[v] |= n;
                t.edges += od[v];
                t.verts += 1ull;
                t.fresh += (unsigned long long)__popcll(n);
            }
            if (front[v] != n)
                front[v] = n;
        }
        unsigned long long has = wg_ballot(n != 0ull);
        if (lane == 0)
            bm[w] = has;
        if (!ROWS) // front is a kept slot: k_ms_rows writes the rows
            continue;
        while (has) { // one row store per vertex with new sources
            const int l = __ffsll((long long)has) - 1;
            has &= has - 1ull;
            const unsigned long long bits = (unsigned long long)__shfl((long long)n, l);
            if ((bits >> lane) & 1ull)
                depth[(w * 64 + (uint64_t)l) * WG_MS_MAX_SOURCES + (uint64_t)lane] = (uint16_t)level;
        }
    }
    add_tally(t, ctr);
}

// Depth rows written once per vertex: levels 0 .. nsave - 1 keep their
// new-source masks (slot k of `save`, V words each, holds level k) instead of
// storing a partial row per vertex per level.  A source reaches a vertex at
// one level only, so its entry is the index of the one slot holding its bit:
// bit j of that index is bit s of plane j, the OR of the slots whose index
// has bit j.  Entries of sources that have not reached the vertex (in no
// slot) are 0xFFFF, which reads never see (they mask with seen) and which a
// later level overwrites when its source arrives.  Each lane builds its
// vertex's 128-byte row in LDS; the wavefront then stores its 64 rows (8 KB,
// contiguous in depth) 1 KB per instruction, skipping rows of vertices that
// no kept level reached.
__global__ __launch_bounds__(WG_BLOCK) void k_ms_rows(const unsigned long long* __restrict__ save, uint32_t nsave,
                                                       uint32_t V, uint16_t* __restrict__ depth)
{
    __shared__ uint4 s_rows_all[WG_WPB][WG_WAVE * MS_ROW_PAD];
    const int lane = wg_lane();
    uint4* s_rows = s_rows_all[threadIdx.x / WG_WAVE];
    const uint64_t nwords = ((uint64_t)V + 63) / 64;
    const uint64_t nwaves = (uint64_t)gridDim.x * WG_WPB;
    for (uint64_t w = (uint64_t)blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE; w < nwords; w += nwaves) {
        const uint32_t v = (uint32_t)(w * 64 + lane);
        unsigned long long pl[4] = { 0ull, 0ull, 0ull, 0ull }, any = 0ull;
#pragma unroll
        for (uint32_t k = 0; k < MS_SAVE; ++k) {
            const unsigned long long m = (k < nsave && v < V) ? save[(uint64_t)k * V + v] : 0ull;
            any |= m;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((k >> j) & 1u)
                    pl[j] |= m;
        }
        const uint32_t h[5][2] = { { (uint32_t)pl[0], (uint32_t)(pl[0] >> 32) },
                                   { (uint32_t)pl[1], (uint32_t)(pl[1] >> 32) },
                                   { (uint32_t)pl[2], (uint32_t)(pl[2] >> 32) },
                                   { (uint32_t)pl[3], (uint32_t)(pl[3] >> 32) },
                                   { (uint32_t)any, (uint32_t)(any >> 32) } };
#pragma unroll
        for (int q = 0; q < 8; ++q) { // sources 8q .. 8q + 7
            uint32_t wd[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                uint32_t pair = 0u;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int src = 8 * q + 2 * t + u, hw = src >> 5, b = src & 31;
                    const uint32_t e = ((h[0][hw] >> b) & 1u) | (((h[1][hw] >> b) & 1u) << 1) |
                                       (((h[2][hw] >> b) & 1u) << 2) | (((h[3][hw] >> b) & 1u) << 3);
                    pair |= (((h[4][hw] >> b) & 1u) ? e : 0xffffu) << (16 * u);
                }
                wd[t] = pair;
            }
            s_rows[lane * MS_ROW_PAD + q] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
        }
        const unsigned long long has = wg_ballot(any != 0ull);
        wg_wave_sync();
        uint4* out = reinterpret_cast<uint4*>(depth + w * 64 * WG_MS_MAX_SOURCES);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t idx = (uint32_t)i * WG_WAVE + (uint32_t)lane; // uint4 of the wavefront's 8 KB
            const uint32_t row = idx >> 3;
            if ((has >> row) & 1ull)
                out[idx] = s_rows[row * MS_ROW_PAD + (idx & 7u)];
        }
        wg_wave_sync();
    }
}

// Edges reached over all sources: sum of od[v] * |seen[v]|.
__global__ void k_ms_reached(const unsigned long long* __restrict__ seen, const uint32_t* __restrict__ od, uint32_t V,
                             wg_dev_counters* __restrict__ c)
{
    unsigned long long ne = 0ull, nv = 0ull;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = (unsigned long long)__popcll(seen[v]);
        nv += k;
        ne += k * od[v];
    }
    add_count(nv, c, WG_CNT_REACHED_V);
    add_count(ne, c, WG_CNT_REACHED_E);
}

__global__ void k_ms_column(const uint16_t* __restrict__ depth, const unsigned long long* __restrict__ seen, uint32_t s,
                            uint32_t V, uint32_t* __restrict__ out)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = ((seen[v] >> s) & 1ull) ? (uint32_t)depth[v * WG_MS_MAX_SOURCES + s] : 0xffffffffu;
}

// ---- host side ----------------------------------------------------------------------

static int ms_alloc(wg_ctx* ctx, size_t bytes, void** p)
{
    WG_HIP(ctx, hipMalloc(p, bytes ? bytes : 8));
    return WG_OK;
}

extern "C" int wg_msbfs(wg_ctx* ctx, const uint64_t* roots, uint32_t nroots, const wg_loop_params* params,
                        uint64_t* levels, uint64_t* edges_reached)
{
    if (!ctx || !roots || !params)
        return WG_E_INVAL;
    if (!ctx->off)
        return WG_E_STATE;
    if (nroots == 0 || nroots > WG_MS_MAX_SOURCES)
        return WG_E_INVAL;
    const int blog2 = params->block_log2 ? params->block_log2 : WG_DEFAULT_BLOCK_LOG2;
    if (blog2 < WG_MIN_BLOCK_LOG2 || blog2 > WG_MAX_BLOCK_LOG2 || params->schedule < 0 || params->schedule > 3)
        return WG_E_INVAL;
    std::vector<uint32_t> hroots(nroots);
    int rc;
    if ((rc = wg_to_device_ids(ctx, roots, nroots, hroots.data())))
        return rc;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t V = (uint32_t)ctx->V, E = (uint32_t)ctx->E;
    void *seen = nullptr, *front = nullptr, *nl = nullptr, *depth = nullptr, *droots = nullptr;
    if (ctx->app == WG_APP_MSBFS) { // the last run's buffers (sized by V, which a new graph frees)
        depth = ctx->prop[0];
        seen = ctx->ms[0];
        front = ctx->ms[1];
        nl = ctx->ms[2];
    } else {
        wg_free_app_state(ctx);
        if ((rc = ms_alloc(ctx, (size_t)V * 8, &seen)) || (rc = ms_alloc(ctx, (size_t)V * 8, &front)) ||
            (rc = ms_alloc(ctx, (size_t)V * 8 + WG_MS_MAX_SOURCES * 4, &nl)) || // nl, then the device roots
            (rc = ms_alloc(ctx, (size_t)V * WG_MS_MAX_SOURCES * 2, &depth))) {
            (void)hipFree(seen);
            (void)hipFree(front);
            (void)hipFree(nl);
            (void)hipFree(depth);
            return rc;
        }
        ctx->prop[0] = depth; // freed with the application state
        ctx->ms[0] = seen;
        ctx->ms[1] = front;
        ctx->ms[2] = nl;
        ctx->app = WG_APP_MSBFS;
        ctx->ms_nl_clean = false;
    }
    droots = (char*)nl + (size_t)V * 8;
    if (!ctx->ms[3] && (rc = ms_alloc(ctx, (size_t)V * 8 * MS_SAVE, &ctx->ms[3])))
        return rc;
    ctx->ms_n = nroots;
    {
        unsigned long long *dseen = (unsigned long long*)seen, *dfront = (unsigned long long*)front,
                           *dnl = (unsigned long long*)nl;
        unsigned long long* const save = (unsigned long long*)ctx->ms[3];
        const unsigned long long live = nroots == 64 ? ~0ull : ((1ull << nroots) - 1ull);
        // Each level's new sources go to the next kept slot (the roots to slot
        // 0) and k_ms_rows writes the depth rows after the run, whole.  When
        // the slots run out, k_ms_rows writes the rows of the kept levels and
        // the later levels write theirs as they commit, into front.
        uint32_t nsave = 1, level = 0;
        bool keeping = true;
        unsigned long long* cur = save; // the last level's new sources (the frontier masks)
        WG_HIP(ctx, hipMemcpyAsync(droots, hroots.data(), (size_t)nroots * 4, hipMemcpyHostToDevice, ctx->stream));
        // nl is zero after a completed run (each commit clears what it takes)
        if ((rc = wg_clear_u64(ctx, (uint64_t*)seen, V)) || (rc = wg_clear_u64(ctx, (uint64_t*)save, V)) ||
            (!ctx->ms_nl_clean && (rc = wg_clear_u64(ctx, (uint64_t*)nl, V))) ||
            (rc = wg_clear_u64(ctx, ctx->bm_next, ctx->bm_words)) ||
            (rc = wg_fill_u32(ctx, ctx->wedge, ctx->wedge_words, 0u)))
            goto out;
        ctx->ms_nl_clean = false; // until this run completes
        if ((rc = wg_reset_counters(ctx, RST_ACTIVITY)))
            goto out;
        hipLaunchKernelGGL(k_ms_roots, dim3((nroots + 63) / 64), dim3(64), 0, ctx->stream, (const uint32_t*)droots,
                           nroots, dseen, cur, (uint16_t*)depth, (unsigned long long*)ctx->bm_next, ctx->od, ctx->ctr);
        if ((rc = wg_check_launch(ctx)) || (rc = wg_take_activity(ctx)))
            goto out;
        unsigned long long act = ctx->hctr->activity; // summed out-degree of the distinct roots
        uint64_t fverts = nroots;                     // vertices in the frontier (at most)
        const unsigned long long threshold = (ctx->E * (unsigned long long)params->threshold_pct) / 100ull;
        const uint32_t cgrid = wg_grid(ctx, ctx->bm_words, WG_WPB, 16);
        const int opt = ctx->ms_binned ? ctx->ms_binned - 1 : 2;
        const bool binned = opt == 1 || (opt == 2 && ctx->E >= (1ull << 24));
        for (;;) {
            const bool dense = params->schedule == 2 || (params->schedule != 3 && act > threshold);
            if (level + 1 > MS_LEVEL_MAX) {
                rc = WG_E_OVERFLOW; // depths are stored in 16 bits
                goto out;
            }
            if (keeping && nsave == MS_SAVE) { // no slot left: the kept levels' rows now
                hipLaunchKernelGGL(k_ms_rows, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream,
                                   (const unsigned long long*)save, nsave, V, (uint16_t*)depth);
                if ((rc = wg_check_launch(ctx)))
                    goto out;
                keeping = false;
            }
            unsigned long long* out = keeping ? save + (size_t)nsave * V : dfront;
            if (dense && binned) { // the binned pull commits the level itself
                if ((rc = wg_pb_msbfs_level(ctx, cur, out, dseen, (uint16_t*)depth, live, level + 1, !keeping)))
                    goto out;
            } else {
                if (dense && fverts < (uint64_t)V / 8)
                    hipLaunchKernelGGL(k_ms_tile<true>, dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off,
                                       ctx->src, ctx->bm_next, cur, dseen, dnl, ctx->tile_v, ctx->tile_e, ctx->ntiles,
                                       live);
                else if (dense)
                    hipLaunchKernelGGL(k_ms_tile<false>, dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off,
                                       ctx->src, ctx->bm_next, cur, dseen, dnl, ctx->tile_v, ctx->tile_e, ctx->ntiles,
                                       live);
                else {
                    if ((rc = wg_launch_wedge_build(ctx, ctx->bm_next, blog2)))
                        goto out;
                    hipLaunchKernelGGL(k_ms_blocks, dim3(wg_frontier_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream, blog2,
                                       E, ctx->src, ctx->dst, ctx->bm_next, cur, dseen, dnl, ctx->wedge, live);
                }
                if ((rc = wg_check_launch(ctx)))
                    goto out;
                // out may be cur (front, once the slots are gone): the pull above has read it
                if (keeping)
                    hipLaunchKernelGGL(k_ms_commit<false>, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream, dnl, dseen,
                                       out, (uint16_t*)depth, ctx->bm_next, ctx->od, V, level + 1, ctx->ctr);
                else
                    hipLaunchKernelGGL(k_ms_commit<true>, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream, dnl, dseen,
                                       out, (uint16_t*)depth, ctx->bm_next, ctx->od, V, level + 1, ctx->ctr);
            }
            if ((rc = wg_check_launch(ctx)) || (rc = wg_take_activity(ctx))) // the counters start the next level at 0
                goto out;
            if (keeping)
                ++nsave; // slot nsave - 1 holds level nsave - 1
            cur = out;
            ++level;
            if (ctx->hctr->activated == 0)
                break;
            act = ctx->hctr->activity;
            fverts = ctx->hctr->activated;
            if (params->max_iterations && level >= params->max_iterations)
                break;
        }
        if (keeping) {
            hipLaunchKernelGGL(k_ms_rows, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream,
                               (const unsigned long long*)save, nsave, V, (uint16_t*)depth);
            if ((rc = wg_check_launch(ctx)))
                goto out;
        }
        if (levels)
            *levels = level;
        if (edges_reached) {
            WG_HIP(ctx, hipMemsetAsync(&ctx->ctr->reached_verts, 0, 2 * sizeof(unsigned long long), ctx->stream));
            hipLaunchKernelGGL(k_ms_reached, dim3(wg_grid(ctx, V, 1024, 8)), dim3(256), 0, ctx->stream, dseen, ctx->od, V,
                               ctx->ctr);
            if ((rc = wg_check_launch(ctx)) || (rc = wg_read_counters(ctx)))
                goto out;
            *edges_reached = ctx->hctr->reached_edges;
        }
    }
out:
    if (rc == WG_OK) {
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->ms_nl_clean = true;
    }
    return rc;
}

extern "C" int wg_read_msbfs(wg_ctx* ctx, uint32_t source, uint32_t* host_depths)
{
    if (!ctx || !host_depths)
        return WG_E_INVAL;
    if (ctx->app != WG_APP_MSBFS)
        return WG_E_STATE;
    if (source >= ctx->ms_n)
        return WG_E_RANGE;
    uint32_t* tmp = nullptr;
    WG_HIP(ctx, hipMalloc((void**)&tmp, ctx->V * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_ms_column, dim3(wg_grid(ctx, ctx->V, 1024, 8)), dim3(256), 0, ctx->stream,
                       (const uint16_t*)ctx->prop[0], (const unsigned long long*)ctx->ms[0], source, (uint32_t)ctx->V,
                       tmp);
    int rc = wg_check_launch(ctx);
    if (!rc)
        rc = wg_read_vertex_array(ctx, tmp, host_depths, sizeof(uint32_t));
    (void)hipFree(tmp);
    return rc;
}

// ---- SSSP from up to 64 sources ---------------------------------------------------------
//
//   dist[v][64]  distance from source s (lane s): FP64 bit patterns (non-negative
//                doubles order like their bits) or uint32, INF if unreached
//   chg[v]       sources whose distance to v improved in the last iteration
// An iteration relaxes destinations, one wavefront per destination, lane s
// taking min over in-edges u->v of dist[u][s] + w; the row of u is read only
// when some source improved u.  Dense iterations relax every destination;
// sparse ones first mark the out-neighbours of the improved vertices and
// relax only those.  Minimum is order-independent, so the result is each
// source's single-source SSSP fixpoint (source/phase_edge_pull_sssp.c:57-169).

struct MsF64 {
    typedef unsigned long long T;
    static constexpr unsigned long long INF = 0x7FF0000000000000ull;
    __device__ static T add(T d, const void* __restrict__ w, uint32_t p, unsigned int* __restrict__)
    {
        return (T)__double_as_longlong(__longlong_as_double((long long)d) + ((const double*)w)[p]);
    }
};

struct MsU32 {
    typedef uint32_t T;
    static constexpr uint32_t INF = 0xffffffffu;
    __device__ static T add(T d, const void* __restrict__ w, uint32_t p, unsigned int* __restrict__ ovf)
    {
        const uint32_t s = d + ((const uint32_t*)w)[p];
        if (s < d || s == INF) {
            *ovf = 1u;
            return INF;
        }
        return s;
    }
};

template <class P>
__global__ void k_mss_init(typename P::T* __restrict__ dist, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dist[i] = P::INF;
}

template <class P>
__global__ void k_mss_roots(const uint32_t* __restrict__ roots, uint32_t n, typename P::T* __restrict__ dist,
                            unsigned long long* __restrict__ chg, unsigned long long* __restrict__ bm,
                            const uint32_t* __restrict__ od, wg_dev_counters* __restrict__ ctr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t r = roots[i];
    dist[(uint64_t)r * WG_MS_MAX_SOURCES + i] = 0;
    atomicOr(&chg[r], 1ull << i);
    const unsigned long long bit = 1ull << (r & 63);
    if (!(atomicOr(&bm[r >> 6], bit) & bit))
        atomicAdd(&ctr->slots[0].v[WG_CNT_ACTIVITY], (unsigned long long)od[r]);
}

// Marks the out-neighbours of the vertices set in `active` in `touched`.
// Lanes take vertices with few out-edges; longer lists are walked by the
// whole wavefront.
__global__ __launch_bounds__(WG_BLOCK) void k_mss_touch(const uint64_t* __restrict__ active, uint64_t nwords,
                                                         const uint32_t* __restrict__ csri_off,
                                                         const uint32_t* __restrict__ csri_pos,
                                                         const uint32_t* __restrict__ dst,
                                                         unsigned long long* __restrict__ touched)
{
    const int lane = wg_lane();
    const uint64_t nwaves = (uint64_t)gridDim.x * WG_WPB;
    for (uint64_t w = (uint64_t)blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE; w < nwords; w += nwaves) {
        const uint64_t bits = active[w];
        if (!bits)
            continue;
        const uint32_t u = (uint32_t)(w * 64 + (uint64_t)lane);
        uint32_t s = 0, e = 0;
        if ((bits >> lane) & 1ull) {
            s = csri_off[u];
            e = csri_off[u + 1];
        }
        const bool big = e - s > WG_WEDGE_LANE_MAX;
        if (!big)
            for (uint32_t p = s; p < e; ++p) {
                const uint32_t v = dst[csri_pos[p]];
                atomicOr(&touched[v >> 6], 1ull << (v & 63));
            }
        unsigned long long many = wg_ballot(big);
        while (many) {
            const int l = __ffsll((long long)many) - 1;
            many &= many - 1ull;
            const uint32_t ls = __shfl(s, l), le = __shfl(e, l);
            for (uint32_t p = ls + (uint32_t)lane; p < le; p += WG_WAVE) {
                const uint32_t v = dst[csri_pos[p]];
                atomicOr(&touched[v >> 6], 1ull << (v & 63));
            }
        }
    }
}

// Minimum over in-edges [b, e) of dist[u][lane] + w for the sources improved
// at u (lane = source), four edges in flight.
template <class P>
__device__ __forceinline__ typename P::T mss_min(uint32_t b, uint32_t e, const uint32_t* __restrict__ src,
                                                 const void* __restrict__ w, const typename P::T* __restrict__ dist,
                                                 const unsigned long long* __restrict__ chg, int lane,
                                                 unsigned int* __restrict__ ovf)
{
    typedef typename P::T T;
    T m = P::INF;
    for (uint32_t p = b; p < e; p += 4) {
        uint32_t u[4];
        unsigned long long c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            u[k] = p + k < e ? src[p + k] : 0u;
            c[k] = p + k < e ? chg[u[k]] : 0ull;
        }
        T x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            x[k] = ((c[k] >> lane) & 1ull) ? dist[(uint64_t)u[k] * WG_MS_MAX_SOURCES + lane] : P::INF;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (x[k] != P::INF) {
                const T cand = P::add(x[k], w, p + k, ovf);
                m = cand < m ? cand : m;
            }
    }
    return m;
}

// Relaxation over the edge tiles, one wavefront per tile, lane = source.  A
// light tile relaxes its destinations one after the other and owns them; a
// heavy tile (part of one destination's in-edges) folds its minimum in with
// per-lane atomics.  With `sel`, only the destinations set in it are relaxed.
// Improved sources go to chg_next[v], improved vertices to the bitmask bm.
template <class P>
__global__ __launch_bounds__(WG_BLOCK) void k_mss_relax(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                         const void* __restrict__ w, const uint32_t* __restrict__ od,
                                                         const uint32_t* __restrict__ tile_v,
                                                         const uint32_t* __restrict__ tile_e, uint32_t ntiles,
                                                         typename P::T* __restrict__ dist,
                                                         const unsigned long long* __restrict__ chg,
                                                         unsigned long long* __restrict__ chg_next,
                                                         const uint64_t* __restrict__ sel, unsigned long long* __restrict__ bm,
                                                         unsigned int* __restrict__ ovf, wg_dev_counters* __restrict__ ctr)
{
    typedef typename P::T T;
    const int lane = wg_lane();
    const uint32_t stride = gridDim.x * WG_WPB;
    act_tally t = { 0ull, 0ull, 0ull };
    for (uint32_t ti = blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE; ti < ntiles; ti += stride) {
        const uint32_t tv = tile_v[ti];
        const uint32_t v0 = tv & ~WG_HEAVY_BIT;
        if (tv & WG_HEAVY_BIT) {
            if (sel && !((sel[v0 >> 6] >> (v0 & 63)) & 1ull))
                continue;
            const T m = mss_min<P>(tile_e[ti], tile_e[ti + 1], src, w, dist, chg, lane, ovf);
            const uint64_t slot = (uint64_t)v0 * WG_MS_MAX_SOURCES + lane;
            bool better = false;
            if (m < dist[slot]) {
                const T old = atomic_min_t<T>(&dist[slot], m);
                better = m < old;
                if (better && old == P::INF)
                    t.fresh += 1ull; // counted once: only the atomic that replaced INF sees it
            }
            const unsigned long long nm = wg_ballot(better);
            if (lane == 0 && nm) {
                const unsigned long long was = atomicOr(&chg_next[v0], nm);
                if (!was) { // first improvement of v0 this iteration
                    t.edges += od[v0];
                    t.verts += 1ull;
                    atomicOr(&bm[v0 >> 6], 1ull << (v0 & 63));
                }
            }
            continue;
        }
        const uint32_t v1 = tile_v[ti + 1] & ~WG_HEAVY_BIT;
        for (uint32_t v = v0; v < v1; ++v) {
            if (sel && !((sel[v >> 6] >> (v & 63)) & 1ull))
                continue;
            const T m = mss_min<P>(off[v], off[v + 1], src, w, dist, chg, lane, ovf);
            const uint64_t slot = (uint64_t)v * WG_MS_MAX_SOURCES + lane;
            const T old = dist[slot];
            const bool better = m < old;
            if (better)
                dist[slot] = m;
            const unsigned long long nm = wg_ballot(better);
            if (nm) {
                const unsigned long long fresh = wg_ballot(better && old == P::INF);
                if (lane == 0) {
                    chg_next[v] = nm;
                    t.edges += od[v];
                    t.verts += 1ull;
                    t.fresh += (unsigned long long)__popcll(fresh);
                    atomicOr(&bm[v >> 6], 1ull << (v & 63));
                }
            }
        }
    }
    add_tally(t, ctr); // counts sit in lane 0; the wave sum adds the zeros of the others
}

template <class P>
__global__ void k_mss_column(const typename P::T* __restrict__ dist, uint32_t s, uint64_t V, typename P::T* __restrict__ out)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = dist[v * WG_MS_MAX_SOURCES + s];
}

template <class P>
static int mssssp_run(wg_ctx* ctx, const std::vector<uint32_t>& roots, const wg_loop_params* params, uint64_t* iters)
{
    typedef typename P::T T;
    const uint32_t V = (uint32_t)ctx->V, n = (uint32_t)roots.size();
    const uint64_t nw = ctx->bm_words;
    void *dist = nullptr, *c0 = nullptr, *c1 = nullptr, *aux = nullptr;
    int rc;
    if ((rc = ms_alloc(ctx, (size_t)V * WG_MS_MAX_SOURCES * sizeof(T), &dist)) ||
        (rc = ms_alloc(ctx, (size_t)V * 8, &c0)) || (rc = ms_alloc(ctx, (size_t)V * 8, &c1)) ||
        (rc = ms_alloc(ctx, nw * 8 + (size_t)n * 4, &aux))) {
        (void)hipFree(dist);
        (void)hipFree(c0);
        (void)hipFree(c1);
        (void)hipFree(aux);
        return rc;
    }
    ctx->prop[0] = dist;
    ctx->ms[0] = c0;
    ctx->ms[1] = c1;
    ctx->ms[2] = aux;
    ctx->ms_n = n;
    ctx->app = WG_APP_MSSSSP;
    unsigned long long* touched = (unsigned long long*)aux;
    uint32_t* droots = (uint32_t*)((char*)aux + nw * 8);
    unsigned long long *chg = (unsigned long long*)c0, *chg_next = (unsigned long long*)c1;
    unsigned int* ovf = &ctx->ctr->overflow;
    WG_HIP(ctx, hipMemcpyAsync(droots, roots.data(), (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_mss_init<P>, dim3(wg_grid(ctx, (uint64_t)V * WG_MS_MAX_SOURCES, 1024, 8)), dim3(256), 0,
                       ctx->stream, (T*)dist, (uint64_t)V * WG_MS_MAX_SOURCES);
    if ((rc = wg_check_launch(ctx)) || (rc = wg_clear_u64(ctx, (uint64_t*)chg, V)) ||
        (rc = wg_clear_u64(ctx, (uint64_t*)chg_next, V)) || (rc = wg_clear_u64(ctx, (uint64_t*)touched, nw)) ||
        (rc = wg_clear_u64(ctx, ctx->bm_next, nw)) || (rc = wg_reset_counters(ctx, RST_ACTIVITY | RST_OVERFLOW)))
        return rc;
    hipLaunchKernelGGL(k_mss_roots<P>, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, droots, n, (T*)dist, chg,
                       (unsigned long long*)ctx->bm_next, ctx->od, ctx->ctr);
    if ((rc = wg_check_launch(ctx)) || (rc = wg_read_counters(ctx)))
        return rc;
    unsigned long long act = ctx->hctr->activity;
    const unsigned long long threshold = (ctx->E * (unsigned long long)params->threshold_pct) / 100ull;
    const uint64_t cap = 64ull * V + 64ull;
    uint64_t it = 0;
    for (;;) {
        const bool dense = params->schedule == 2 || (params->schedule != 3 && act > threshold);
        if ((rc = wg_reset_counters(ctx, RST_ACTIVITY)))
            return rc;
        if (!dense) {
            hipLaunchKernelGGL(k_mss_touch, dim3(wg_wedge_build_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream,
                               ctx->bm_next, nw, ctx->csri_off, ctx->csri_pos, ctx->dst, touched);
            if ((rc = wg_check_launch(ctx)))
                return rc;
        }
        // the improved-vertex bitmask is rebuilt by the relaxation
        if ((rc = wg_clear_u64(ctx, ctx->bm_next, nw)))
            return rc;
        hipLaunchKernelGGL(k_mss_relax<P>, dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->src, ctx->w,
                           ctx->od, ctx->tile_v, ctx->tile_e, ctx->ntiles, (T*)dist, chg, chg_next,
                           dense ? (const uint64_t*)nullptr : (const uint64_t*)touched,
                           (unsigned long long*)ctx->bm_next, ovf, ctx->ctr);
        if ((rc = wg_check_launch(ctx)) || (!dense && (rc = wg_clear_u64(ctx, (uint64_t*)touched, nw))) ||
            (rc = wg_read_counters(ctx)))
            return rc;
        ++it;
        if (ctx->hctr->activated == 0)
            break;
        if (it > cap || (params->max_iterations && it >= params->max_iterations))
            break;
        act = ctx->hctr->activity;
        // the improved masks become the next iteration's input; the old ones are cleared for reuse
        unsigned long long* tmp = chg;
        chg = chg_next;
        chg_next = tmp;
        if ((rc = wg_clear_u64(ctx, (uint64_t*)chg_next, V)))
            return rc;
    }
    if (iters)
        *iters = it;
    return ctx->hctr->overflow ? WG_E_OVERFLOW : WG_OK;
}

extern "C" int wg_mssssp(wg_ctx* ctx, const uint64_t* roots, uint32_t nroots, const wg_loop_params* params,
                         uint64_t* iterations)
{
    if (!ctx || !roots || !params)
        return WG_E_INVAL;
    if (!ctx->off || ctx->wkind == WG_WEIGHT_NONE)
        return WG_E_STATE;
    if (nroots == 0 || nroots > WG_MS_MAX_SOURCES || params->schedule < 0 || params->schedule > 3)
        return WG_E_INVAL;
    std::vector<uint32_t> hroots(nroots);
    int rc;
    if ((rc = wg_to_device_ids(ctx, roots, nroots, hroots.data())))
        return rc;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    wg_free_app_state(ctx);
    rc = ctx->wkind == WG_WEIGHT_F64 ? mssssp_run<MsF64>(ctx, hroots, params, iterations)
                                     : mssssp_run<MsU32>(ctx, hroots, params, iterations);
    if (rc == WG_OK || rc == WG_E_OVERFLOW)
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return rc;
}

extern "C" int wg_read_mssssp(wg_ctx* ctx, uint32_t source, void* host_dist)
{
    if (!ctx || !host_dist)
        return WG_E_INVAL;
    if (ctx->app != WG_APP_MSSSSP)
        return WG_E_STATE;
    if (source >= ctx->ms_n)
        return WG_E_RANGE;
    const size_t esz = ctx->wkind == WG_WEIGHT_F64 ? 8 : 4;
    void* tmp = nullptr;
    WG_HIP(ctx, hipMalloc(&tmp, ctx->V * esz));
    const uint32_t grid = wg_grid(ctx, ctx->V, 1024, 8);
    if (esz == 8)
        hipLaunchKernelGGL(k_mss_column<MsF64>, dim3(grid), dim3(256), 0, ctx->stream,
                           (const unsigned long long*)ctx->prop[0], source, ctx->V, (unsigned long long*)tmp);
    else
        hipLaunchKernelGGL(k_mss_column<MsU32>, dim3(grid), dim3(256), 0, ctx->stream, (const uint32_t*)ctx->prop[0],
                           source, ctx->V, (uint32_t*)tmp);
    int rc = wg_check_launch(ctx);
    if (!rc)
        rc = wg_read_vertex_array(ctx, tmp, host_dist, esz);
    (void)hipFree(tmp);
    return rc;
}
