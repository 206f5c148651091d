// wg_pr.hip -- PageRank edge pull and vertex phase on gfx950.
//
// Reference: perform_edge_pull_phase_pr (source/phase_edge_pull_pr.c:28-116),
// the merge of destinations split across work units (source/phases.c:69-114),
// the reduce-buffer combine (source/phases.c:40-50) and perform_vertex_phase_pr
// (source/phase_vertex_pr.c:26-77) with its formula
//     contrib'[v] = ((acc[v] + (1 - S) / N) * d + (1 - d) / N) / od'[v],
//     od'[v] = N if od[v] == 0, S = sum of all accumulators.
//
// MI355X design.  One wavefront processes one edge tile at a time (grid-stride
// over tiles).  A light tile holds whole destinations (<= 512 edges, <= 256
// vertices): each lane sums 8 consecutive edges, a destination that lies
// inside one lane is written to LDS by that lane, a destination that spans
// lanes is summed in lane order from per-lane first/last pieces, and the
// wavefront then runs the vertex phase for the tile's vertices with coalesced
// reads of od[] and writes of contrib[].  All sums are taken in a fixed order,
// so results are bit-reproducible run to run.  Destinations with more than 512
// in-edges are cut into heavy tiles whose partial sums are added in tile
// order by a second small kernel, which also reduces the per-block partial
// sums of S.
//
// The fused iteration needs S = sum of accumulators before any accumulator of
// the iteration exists.  It uses the identity
//     sum_v acc_t[v] = sum_u od[u] * contrib_{t-1}[u] = sum_{u: od[u] > 0} rank_{t-1}[u],
// so the previous iteration produces S for the next one; one pass over the
// edges per iteration remains (4E + 24V algorithmic bytes with FP64
// contributions, 4E + 16V with FP32 contributions).
#include "wg_pr_tile.h"

// Stand-alone vertex phase (source/phase_vertex_pr.c:26-77) with S taken from
// the preceding edge phase; also produces the sum for a following fused
// iteration.
template <typename T>
__global__ __launch_bounds__(WG_BLOCK) void k_pr_vertex(const double* __restrict__ acc, const uint32_t* __restrict__ od,
                                                         T* __restrict__ cout, uint64_t V, const double* __restrict__ total,
                                                         double damping, double N, double* __restrict__ blk_part,
                                                         double* __restrict__ S_out, unsigned int* __restrict__ ticket)
{
    __shared__ double s_red[WG_WPB];
    __shared__ bool s_last;
    const double corr = (1.0 - *total) / N;
    const double omd_n = (1.0 - damping) / N;
    double s_part = 0.0;
    for (uint64_t v = blockIdx.x * (uint64_t)WG_BLOCK + threadIdx.x; v < V; v += (uint64_t)gridDim.x * WG_BLOCK) {
        const uint32_t odv = od[v];
        const double t1 = acc[v] + corr;
        const double t2 = t1 * damping;
        const double rank = t2 + omd_n;
        cout[v] = (T)(rank / (odv ? (double)odv : N));
        if (odv)
            s_part += rank;
    }
    s_part = block_sum_det(s_part, s_red);
    if (threadIdx.x == 0) {
        blk_part[blockIdx.x] = s_part;
        __threadfence();
        const unsigned int prev = atomicAdd(ticket, 1u);
        s_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        const double tot = reduce_partials(blk_part, gridDim.x, s_red);
        if (threadIdx.x == 0) {
            *S_out = tot;
            *ticket = 0u;
        }
    }
}

template <typename T>
__global__ void k_pr_init(const uint32_t* __restrict__ od, T* __restrict__ c, uint64_t V, double N)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = od[v];
        c[v] = (T)((1.0 / N) / (o ? (double)o : N));
    }
}

template <typename T>
__global__ void k_pr_ranks(const uint32_t* __restrict__ od, const T* __restrict__ c, double* __restrict__ r, uint64_t V,
                           double N)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = od[v];
        r[v] = (double)c[v] * (o ? (double)o : N);
    }
}

// ---- host side ------------------------------------------------------------------

extern "C" int wg_init_pr(wg_ctx* ctx, double damping, int contrib_fp32)
{
    if (!ctx || !ctx->off)
        return WG_E_STATE;
    if (!(damping >= 0.0 && damping <= 1.0))
        return WG_E_INVAL;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    wg_free_app_state(ctx);
    const size_t esz = contrib_fp32 ? sizeof(float) : sizeof(double);
    WG_HIP(ctx, hipMalloc(&ctx->prop[0], ctx->V * esz));
    WG_HIP(ctx, hipMalloc(&ctx->prop[1], ctx->V * esz));
    ctx->app = WG_APP_PR;
    ctx->pr_fp32 = contrib_fp32 ? 1 : 0;
    ctx->damping = damping;
    ctx->pr_cur = 0;
    ctx->pr_parity = 0;
    ctx->iter = 0;
    const double N = (double)ctx->V;
    const uint32_t grid = wg_grid(ctx, ctx->V, 1024, 8);
    if (ctx->pr_fp32)
        hipLaunchKernelGGL(k_pr_init<float>, dim3(grid), dim3(256), 0, ctx->stream, ctx->od, (float*)ctx->prop[0], ctx->V, N);
    else
        hipLaunchKernelGGL(k_pr_init<double>, dim3(grid), dim3(256), 0, ctx->stream, ctx->od, (double*)ctx->prop[0], ctx->V, N);
    int rc = wg_check_launch(ctx);
    if (rc)
        return rc;
    // sum over vertices with out-degree > 0 of the initial rank 1/N
    const double s1 = (double)ctx->n_out_pos * (1.0 / N);
    WG_HIP(ctx, hipMemcpyAsync(&ctx->ctr->S[0], &s1, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    WG_HIP(ctx, hipMemsetAsync(&ctx->ctr->ticket, 0, sizeof(unsigned int), ctx->stream));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return WG_OK;
}

// The tiling and launch shape of the PageRank edge pull.
struct pr_shape {
    const uint32_t *tile_v, *tile_e, *heavy_v, *heavy_t;
    uint32_t ntiles, nheavy, grid, sgrid;
    wg_tile_split split;
};

static pr_shape current_shape(const wg_ctx* ctx)
{
    pr_shape s;
    const uint32_t* split = ctx->pr_split8;
    if (ctx->pr_ipt == WG_IPT) {
        s.tile_v = ctx->tile_v;
        s.tile_e = ctx->tile_e;
        s.heavy_v = ctx->heavy_v;
        s.heavy_t = ctx->heavy_t;
        s.ntiles = ctx->ntiles;
        s.nheavy = ctx->nheavy;
    } else {
        const wg_tiling& t = ctx->pr_alt[ctx->pr_ipt == 4 ? 0 : ctx->pr_ipt == 16 ? 1 : 2];
        s.tile_v = t.tile_v;
        s.tile_e = t.tile_e;
        s.heavy_v = t.heavy_v;
        s.heavy_t = t.heavy_t;
        s.ntiles = t.ntiles;
        s.nheavy = t.nheavy;
        split = t.split;
    }
    // blocks_per_cu 1: one workgroup of WG_PR_BIG_WPB wavefronts per CU
    s.grid = ctx->pr_bpc == 1 ? wg_grid(ctx, s.ntiles, WG_PR_BIG_WPB, 1) : wg_grid(ctx, s.ntiles, WG_WPB, (uint32_t)ctx->pr_bpc);
    s.sgrid = wg_grid(ctx, s.nheavy, 4, 4);
    s.split.on = ctx->pr_xcd ? 1u : 0u;
    for (int x = 0; x <= WG_NXCD; ++x)
        s.split.start[x] = split[x];
    if (s.split.on) // whole rows of 8 workgroups, one per XCD
        s.grid = (s.grid + WG_NXCD - 1) / WG_NXCD * WG_NXCD;
    return s;
}

template <typename T, int MODE, int IPT>
static void pr_launch_shape(wg_ctx* ctx, const pr_shape& s, const T* cin, T* cout, double* S_in, double* S_out, double N)
{
    // the hottest contributions in LDS, if the vertices are in degree order
    const bool big = ctx->pr_bpc == 1;
    const uint64_t hub_max = big ? wg_pr_hub_max<T, WG_PR_BIG_WPB>() : wg_pr_hub_max<T, WG_WPB>();
    const uint32_t nhub = ctx->perm ? (uint32_t)(ctx->V < hub_max ? ctx->V : hub_max) : 0u;
    if (big)
        hipLaunchKernelGGL((k_pr_tile<T, MODE, IPT, true, WG_PR_BIG_WPB>), dim3(s.grid), dim3(WG_PR_BIG_WPB * WG_WAVE), 0,
                           ctx->stream, ctx->off, ctx->src, cin, cout, ctx->acc, ctx->od, s.tile_v, s.tile_e, s.ntiles,
                           S_in, ctx->damping, N, ctx->tile_part, ctx->blk_part, s.split, nhub);
    else
        hipLaunchKernelGGL((k_pr_tile<T, MODE, IPT, true>), dim3(s.grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off,
                           ctx->src, cin, cout, ctx->acc, ctx->od, s.tile_v, s.tile_e, s.ntiles, S_in, ctx->damping, N,
                           ctx->tile_part, ctx->blk_part, s.split, nhub);
    hipLaunchKernelGGL((k_pr_span<T, MODE, IPT>), dim3(s.sgrid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, s.heavy_v,
                       s.heavy_t, s.nheavy, ctx->tile_part, cout, ctx->acc, ctx->od, S_in, ctx->damping, N, ctx->blk_part,
                       s.grid, S_out, &ctx->ctr->ticket);
}

template <typename T>
static int pr_launch(wg_ctx* ctx, int mode)
{
    const double N = (double)ctx->V;
    const T* cin = (const T*)ctx->prop[ctx->pr_cur];
    T* cout = (T*)ctx->prop[ctx->pr_cur ^ 1];
    double* S_in = &ctx->ctr->S[ctx->pr_parity];
    double* S_out = &ctx->ctr->S[ctx->pr_parity ^ 1];
    if (mode == 1) {
        S_out = &ctx->ctr->total_rank;
        cout = nullptr;
    }
    if (ctx->pr_xcd == WG_PR_BINNED)
        return wg_pb_iteration(ctx, mode, cin, cout, S_in, S_out);
    const pr_shape s = current_shape(ctx);
    // one bracket around tile + span: the iteration's device time
    const int slot = wg_prof_begin(ctx, mode == 0 ? WG_K_PR_TILE : WG_K_PR_ACC);
    switch (ctx->pr_ipt * 2 + mode) {
    case 4: pr_launch_shape<T, 0, 2>(ctx, s, cin, cout, S_in, S_out, N); break;
    case 5: pr_launch_shape<T, 1, 2>(ctx, s, cin, cout, S_in, S_out, N); break;
    case 8: pr_launch_shape<T, 0, 4>(ctx, s, cin, cout, S_in, S_out, N); break;
    case 9: pr_launch_shape<T, 1, 4>(ctx, s, cin, cout, S_in, S_out, N); break;
    case 16: pr_launch_shape<T, 0, 8>(ctx, s, cin, cout, S_in, S_out, N); break;
    case 17: pr_launch_shape<T, 1, 8>(ctx, s, cin, cout, S_in, S_out, N); break;
    case 32: pr_launch_shape<T, 0, 16>(ctx, s, cin, cout, S_in, S_out, N); break;
    case 33: pr_launch_shape<T, 1, 16>(ctx, s, cin, cout, S_in, S_out, N); break;
    default: return WG_E_STATE;
    }
    const int rc = wg_check_launch(ctx);
    wg_prof_end(ctx, mode == 0 ? WG_K_PR_TILE : WG_K_PR_ACC, slot);
    return rc;
}

extern "C" int wg_set_pr_shape(wg_ctx* ctx, const wg_batch_shape* shape)
{
    if (!ctx)
        return WG_E_INVAL;
    if (!ctx->off)
        return WG_E_STATE;
    if (!shape) {
        ctx->pr_ipt = WG_IPT;
        ctx->pr_bpc = 8;
        ctx->pr_xcd = 0;
        return WG_OK;
    }
    const uint32_t ipt = shape->edges_per_lane;
    if (shape->xcd_split == WG_PR_BINNED) {
        if (ipt != 8 || shape->blocks_per_cu < 1 || shape->blocks_per_cu > 2)
            return WG_E_INVAL;
        ctx->pr_ipt = (int)ipt;
        ctx->pr_bpc = (int)shape->blocks_per_cu;
        ctx->pr_xcd = WG_PR_BINNED;
        return WG_OK;
    }
    if ((ipt != 2 && ipt != 4 && ipt != 8 && ipt != 16) || shape->blocks_per_cu < 1 || shape->blocks_per_cu > 16 ||
        shape->xcd_split > 1)
        return WG_E_INVAL;
    if (ipt != WG_IPT) {
        WG_HIP(ctx, hipSetDevice(ctx->device));
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        int rc = wg_build_pr_tilings(ctx);
        if (rc)
            return rc;
    }
    ctx->pr_ipt = (int)ipt;
    ctx->pr_bpc = (int)shape->blocks_per_cu;
    ctx->pr_xcd = (int)shape->xcd_split;
    return WG_OK;
}

extern "C" int wg_get_pr_shape(const wg_ctx* ctx, wg_batch_shape* shape)
{
    if (!ctx || !shape)
        return WG_E_INVAL;
    shape->edges_per_lane = (uint32_t)(ctx->pr_ipt ? ctx->pr_ipt : WG_IPT);
    shape->blocks_per_cu = (uint32_t)(ctx->pr_bpc ? ctx->pr_bpc : 8);
    shape->xcd_split = (uint32_t)ctx->pr_xcd;
    return WG_OK;
}

extern "C" int wg_pr_shape_at(uint32_t i, wg_batch_shape* out)
{
    // one 16-wave workgroup per CU for each tile size, then 128-edge tiles
    static const wg_batch_shape extra[WG_PR_NSHAPES - WG_BATCH_NSHAPES] = {
        {4, 1, 0}, {8, 1, 0}, {16, 1, 0}, {2, 1, 0}, {2, 4, 0}, {2, 16, 0}, {8, 1, WG_PR_BINNED}, {8, 2, WG_PR_BINNED}};
    if (!out || i >= WG_PR_NSHAPES)
        return WG_E_INVAL;
    if (i < WG_BATCH_NSHAPES)
        return wg_batch_shape_at(i, out);
    *out = extra[i - WG_BATCH_NSHAPES];
    return WG_OK;
}

extern "C" int wg_pr_autotune(wg_ctx* ctx, uint32_t iterations, wg_batch_shape* best, float* ms_per_shape)
{
    if (!ctx || iterations == 0)
        return WG_E_INVAL;
    if (ctx->app != WG_APP_PR)
        return WG_E_STATE;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const double damping = ctx->damping;
    const int fp32 = ctx->pr_fp32;
    hipEvent_t ev[2];
    WG_HIP(ctx, hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return WG_E_HIP;
    }
    int rc = WG_OK;
    float best_ms = 0.0f;
    wg_batch_shape best_s = {WG_IPT, 8, 0};
    for (uint32_t i = 0; i < WG_PR_NSHAPES && rc == WG_OK; ++i) {
        wg_batch_shape s;
        wg_pr_shape_at(i, &s);
        if ((rc = wg_set_pr_shape(ctx, &s)) || (rc = wg_pr_iteration(ctx, nullptr))) // warm-up
            break;
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && rc == WG_OK; ++rep) {
            if (hipEventRecord(ev[0], ctx->stream) != hipSuccess) {
                rc = WG_E_HIP;
                break;
            }
            for (uint32_t it = 0; it < iterations && rc == WG_OK; ++it)
                rc = wg_pr_iteration(ctx, nullptr);
            float t = 0.0f;
            if (rc || hipEventRecord(ev[1], ctx->stream) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess) {
                rc = rc ? rc : WG_E_HIP;
                break;
            }
            t /= (float)iterations;
            ms = (rep == 0 || t < ms) ? t : ms;
        }
        if (rc)
            break;
        if (ms_per_shape)
            ms_per_shape[i] = ms;
        if (i == 0 || ms < best_ms) {
            best_ms = ms;
            best_s = s;
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc)
        return rc;
    // the timed iterations advanced the state: start again from the uniform ranks
    if ((rc = wg_set_pr_shape(ctx, &best_s)) || (rc = wg_init_pr(ctx, damping, fp32)))
        return rc;
    if (best)
        *best = best_s;
    return WG_OK;
}

extern "C" int wg_pr_bins_info(wg_ctx* ctx, uint64_t* pairs, uint64_t* code_entries, uint32_t* push_groups,
                               uint32_t* pull_groups)
{
    if (!ctx || !pairs || !code_entries || !push_groups || !pull_groups)
        return WG_E_INVAL;
    if (ctx->app != WG_APP_PR || ctx->pr_xcd != WG_PR_BINNED)
        return WG_E_STATE;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const int rc = wg_pb_prepare(ctx);
    if (rc)
        return rc;
    return wg_pb_stats(ctx, pairs, code_entries, push_groups, pull_groups);
}

extern "C" int wg_pr_iteration(wg_ctx* ctx, double* total_rank)
{
    if (!ctx || ctx->app != WG_APP_PR)
        return WG_E_STATE;
    const int parity = ctx->pr_parity;
    int rc = ctx->pr_fp32 ? pr_launch<float>(ctx, 0) : pr_launch<double>(ctx, 0);
    if (rc)
        return rc;
    ctx->pr_cur ^= 1;
    ctx->pr_parity ^= 1;
    ctx->iter += 1;
    if (total_rank) {
        // the sum of this iteration's accumulators is the S it consumed
        WG_HIP(ctx, hipMemcpyAsync(total_rank, &ctx->ctr->S[parity], sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return WG_OK;
}

extern "C" int wg_edge_pull_pr(wg_ctx* ctx, double* total_rank)
{
    if (!ctx || ctx->app != WG_APP_PR)
        return WG_E_STATE;
    if (!ctx->acc)
        WG_HIP(ctx, hipMalloc((void**)&ctx->acc, ctx->V * sizeof(double)));
    int rc = ctx->pr_fp32 ? pr_launch<float>(ctx, 1) : pr_launch<double>(ctx, 1);
    if (rc)
        return rc;
    if (total_rank) {
        WG_HIP(ctx, hipMemcpyAsync(total_rank, &ctx->ctr->total_rank, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return WG_OK;
}

extern "C" int wg_vertex_pr(wg_ctx* ctx)
{
    if (!ctx || ctx->app != WG_APP_PR || !ctx->acc)
        return WG_E_STATE;
    const double N = (double)ctx->V;
    const uint32_t grid = wg_grid(ctx, ctx->V, WG_BLOCK * 4, 8);
    int slot = wg_prof_begin(ctx, WG_K_PR_VERTEX);
    if (ctx->pr_fp32)
        hipLaunchKernelGGL(k_pr_vertex<float>, dim3(grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->acc, ctx->od,
                           (float*)ctx->prop[ctx->pr_cur], ctx->V, &ctx->ctr->total_rank, ctx->damping, N, ctx->blk_part,
                           &ctx->ctr->S[ctx->pr_parity], &ctx->ctr->ticket);
    else
        hipLaunchKernelGGL(k_pr_vertex<double>, dim3(grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->acc, ctx->od,
                           (double*)ctx->prop[ctx->pr_cur], ctx->V, &ctx->ctr->total_rank, ctx->damping, N,
                           ctx->blk_part, &ctx->ctr->S[ctx->pr_parity], &ctx->ctr->ticket);
    int rc = wg_check_launch(ctx);
    if (rc)
        return rc;
    wg_prof_end(ctx, WG_K_PR_VERTEX, slot);
    ctx->iter += 1;
    return WG_OK;
}

extern "C" int wg_read_ranks(wg_ctx* ctx, double* host_ranks)
{
    if (!ctx || ctx->app != WG_APP_PR || !host_ranks)
        return WG_E_STATE;
    double* tmp = nullptr;
    WG_HIP(ctx, hipMalloc((void**)&tmp, ctx->V * sizeof(double)));
    const double N = (double)ctx->V;
    const uint32_t grid = wg_grid(ctx, ctx->V, 1024, 8);
    if (ctx->pr_fp32)
        hipLaunchKernelGGL(k_pr_ranks<float>, dim3(grid), dim3(256), 0, ctx->stream, ctx->od,
                           (const float*)ctx->prop[ctx->pr_cur], tmp, ctx->V, N);
    else
        hipLaunchKernelGGL(k_pr_ranks<double>, dim3(grid), dim3(256), 0, ctx->stream, ctx->od,
                           (const double*)ctx->prop[ctx->pr_cur], tmp, ctx->V, N);
    int rc = wg_check_launch(ctx);
    if (rc == WG_OK)
        rc = wg_read_vertex_array(ctx, tmp, host_ranks, sizeof(double));
    (void)hipFree(tmp);
    return rc;
}
