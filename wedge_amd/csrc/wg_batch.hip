// wg_batch.hip -- many BFS sources in one traversal (no counterpart in the
// reference, whose BFS takes one root: source/execution_bfs.c:84-118).
//
// Up to 64 sources share every pull: lane s of a vertex's 64-bit masks
// belongs to source s.
//   seen[v]   sources that have reached v
//   front[v]  sources that reached v at the last level (the frontier)
//   nl[v]     sources reaching v at the running level (filled by the pull)
//   depth[v][64]  uint16 level at which source s reached v, 0xFFFF if never
//                 (vertex-major: a vertex's row is written by one store)
// A dense level is the bottom-up pull of source/phase_edge_pull_bfs.c:109-210
// for all sources at once: a vertex ORs the frontier masks of its in-edges,
// restricted to the sources that have not reached it.  A frontier level walks the Wedge frontier
// built from the vertices with any new source (source/wedge.c:63-155).  The
// schedule per level is the plain rule on the summed out-degree of those
// vertices.  Each source's depths equal those of a single-source BFS.
#include "wg_frontier.h"

#include <algorithm>
#include <vector>

#define MS_DEPTH_NONE 0xFFFFu
#define MS_LEVEL_MAX 0xFFFEu

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

__global__ void k_ms_column(const uint16_t* __restrict__ depth, uint32_t s, uint32_t V, uint32_t* __restrict__ out)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint16_t d = depth[v * WG_MS_MAX_SOURCES + s];
        out[v] = d == MS_DEPTH_NONE ? 0xffffffffu : (uint32_t)d;
    }
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
    for (uint32_t i = 0; i < nroots; ++i)
        if ((rc = wg_to_device_id(ctx, roots[i], &hroots[i])))
            return rc;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    wg_free_app_state(ctx);
    const uint32_t V = (uint32_t)ctx->V, E = (uint32_t)ctx->E;
    void *seen = nullptr, *front = nullptr, *nl = nullptr, *depth = nullptr, *droots = nullptr;
    if ((rc = ms_alloc(ctx, (size_t)V * 8, &seen)) || (rc = ms_alloc(ctx, (size_t)V * 8, &front)) ||
        (rc = ms_alloc(ctx, (size_t)V * 8, &nl)) || (rc = ms_alloc(ctx, (size_t)V * WG_MS_MAX_SOURCES * 2, &depth)) ||
        (rc = ms_alloc(ctx, (size_t)nroots * 4, &droots))) {
        (void)hipFree(seen);
        (void)hipFree(front);
        (void)hipFree(nl);
        (void)hipFree(depth);
        (void)hipFree(droots);
        return rc;
    }
    ctx->prop[0] = depth;  // freed with the application state
    ctx->ms[0] = seen;
    ctx->ms[1] = front;
    ctx->ms[2] = nl;
    ctx->ms_n = nroots;
    ctx->app = WG_APP_MSBFS;
    unsigned long long *dseen = (unsigned long long*)seen, *dfront = (unsigned long long*)front,
                       *dnl = (unsigned long long*)nl;
    const unsigned long long live = nroots == 64 ? ~0ull : ((1ull << nroots) - 1ull);
    WG_HIP(ctx, hipMemcpyAsync(droots, hroots.data(), (size_t)nroots * 4, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = wg_clear_u64(ctx, (uint64_t*)seen, V)) || (rc = wg_clear_u64(ctx, (uint64_t*)front, V)) ||
        (rc = wg_clear_u64(ctx, (uint64_t*)nl, V)) || (rc = wg_clear_u64(ctx, ctx->bm_next, ctx->bm_words)) ||
        (rc = wg_fill_u32(ctx, (uint32_t*)depth, (uint64_t)V * (WG_MS_MAX_SOURCES / 2), 0xffffffffu)) ||
        (rc = wg_fill_u32(ctx, ctx->wedge, ctx->wedge_words, 0u)))
        goto out;
    if ((rc = wg_reset_counters(ctx, RST_ACTIVITY)))
        goto out;
    hipLaunchKernelGGL(k_ms_roots, dim3((nroots + 63) / 64), dim3(64), 0, ctx->stream, (const uint32_t*)droots, nroots,
                       dseen, dfront, (uint16_t*)depth, (unsigned long long*)ctx->bm_next, ctx->od, ctx->ctr);
    if ((rc = wg_check_launch(ctx)) || (rc = wg_read_counters(ctx)))
        goto out;
    {
        unsigned long long act = ctx->hctr->activity; // summed out-degree of the distinct roots
        uint64_t fverts = nroots;                     // vertices in the frontier (at most)
        const unsigned long long threshold = (ctx->E * (unsigned long long)params->threshold_pct) / 100ull;
        const uint32_t cgrid = wg_grid(ctx, ctx->bm_words, WG_WPB, 16);
        uint32_t level = 0;
        for (;;) {
            const bool dense = params->schedule == 2 || (params->schedule != 3 && act > threshold);
            if ((rc = wg_reset_counters(ctx, RST_ACTIVITY)))
                goto out;
            if (dense) {
                if (fverts < (uint64_t)V / 8)
                    hipLaunchKernelGGL(k_ms_tile<true>, dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off,
                                       ctx->src, ctx->bm_next, dfront, dseen, dnl, ctx->tile_v, ctx->tile_e, ctx->ntiles,
                                       live);
                else
                    hipLaunchKernelGGL(k_ms_tile<false>, dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off,
                                       ctx->src, ctx->bm_next, dfront, dseen, dnl, ctx->tile_v, ctx->tile_e, ctx->ntiles,
                                       live);
            } else {
                if ((rc = wg_launch_wedge_build(ctx, ctx->bm_next, blog2)))
                    goto out;
                hipLaunchKernelGGL(k_ms_blocks, dim3(wg_frontier_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream, blog2, E,
                                   ctx->src, ctx->dst, ctx->bm_next, dfront, dseen, dnl, ctx->wedge, live);
            }
            if ((rc = wg_check_launch(ctx)))
                goto out;
            if (level + 1 > MS_LEVEL_MAX) {
                rc = WG_E_OVERFLOW; // depths are stored in 16 bits
                goto out;
            }
            hipLaunchKernelGGL(k_ms_commit, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream, dnl, dseen, dfront,
                               (uint16_t*)depth, ctx->bm_next, ctx->od, V, level + 1, ctx->ctr);
            if ((rc = wg_check_launch(ctx)) || (rc = wg_read_counters(ctx)))
                goto out;
            ++level;
            if (ctx->hctr->activated == 0)
                break;
            act = ctx->hctr->activity;
            fverts = ctx->hctr->activated;
            if (params->max_iterations && level >= params->max_iterations)
                break;
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
    (void)hipFree(droots);
    if (rc == WG_OK)
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
                       (const uint16_t*)ctx->prop[0], source, (uint32_t)ctx->V, tmp);
    int rc = wg_check_launch(ctx);
    if (!rc)
        rc = wg_read_vertex_array(ctx, tmp, host_depths, sizeof(uint32_t));
    (void)hipFree(tmp);
    return rc;
}
