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
#include "wg_internal.h"

// Deterministic block-wide sum (xor butterfly inside each wave, then the
// four wave sums in wave order).  All threads receive the result.
__device__ __forceinline__ double block_sum_det(double v, double* s_red)
{
    v = wg_wave_sum(v);
    const int wave = threadIdx.x / WG_WAVE;
    __syncthreads();
    if ((threadIdx.x & (WG_WAVE - 1)) == 0)
        s_red[wave] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < WG_WPB; ++i)
        r += s_red[i];
    return r;
}

template <typename T>
__device__ __forceinline__ double load_contrib(const T* __restrict__ p, uint32_t i)
{
    return (double)p[i];
}

// MODE 0: fused iteration (contrib_out + partial S of ranks)
// MODE 1: edge phase only (acc_out + partial sum of accumulators)
//
// One wavefront per tile, tiles taken grid-stride.  All global loads of a
// tile (its sources, offset window and out-degrees) are issued before the
// first dependent use, the source stream is read non-temporally (it is read
// once per iteration and should not displace contributions from L2), and
// wavefronts never wait for each other.
template <typename T, int MODE>
__global__ __launch_bounds__(WG_BLOCK) void k_pr_tile(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                       const T* __restrict__ cin, T* __restrict__ cout,
                                                       double* __restrict__ acc_out, const uint32_t* __restrict__ od,
                                                       const uint32_t* __restrict__ tile_v,
                                                       const uint32_t* __restrict__ tile_e, uint32_t ntiles,
                                                       const double* __restrict__ S_in, double damping, double N,
                                                       double* __restrict__ tile_part, double* __restrict__ blk_part)
{
    constexpr int NOFF = (WG_TILE_VERTS + 1 + WG_WAVE - 1) / WG_WAVE;
    constexpr int NVTX = WG_TILE_VERTS / WG_WAVE;
    __shared__ uint32_t s_off_all[WG_WPB][WG_TILE_VERTS + 1];
    __shared__ double s_acc_all[WG_WPB][WG_TILE_VERTS];
    __shared__ double s_fp_all[WG_WPB][WG_WAVE];
    __shared__ double s_lp_all[WG_WPB][WG_WAVE];
    __shared__ double s_red[WG_WPB];

    const int wave = threadIdx.x / WG_WAVE;
    const int lane = wg_lane();
    uint32_t* s_off = s_off_all[wave];
    double* s_acc = s_acc_all[wave];
    double* s_fp = s_fp_all[wave];
    double* s_lp = s_lp_all[wave];
    double corr = 0.0, omd_n = 0.0;
    if (MODE == 0) {
        corr = (1.0 - *S_in) / N;
        omd_n = (1.0 - damping) / N;
    }
    double s_part = 0.0;

    const uint32_t stride = gridDim.x * WG_WPB;
    for (uint32_t t = blockIdx.x * WG_WPB + wave; t < ntiles; t += stride) {
        const uint32_t tv = tile_v[t];
        const uint32_t e0 = tile_e[t], e1 = tile_e[t + 1];
        const uint32_t v0 = tv & ~WG_HEAVY_BIT;
        const uint32_t a = e0 + (uint32_t)lane * WG_IPT;
        uint32_t s_idx[WG_IPT];
#pragma unroll
        for (int j = 0; j < WG_IPT; ++j)
            s_idx[j] = (a + j < e1) ? __builtin_nontemporal_load(&src[a + j]) : 0u;
        if (tv & WG_HEAVY_BIT) {
            // heavy tile: one destination, partial sum in a fixed order
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < WG_IPT; ++j)
                if (a + j < e1)
                    s += load_contrib(cin, s_idx[j]);
            s = wg_wave_sum(s);
            if (lane == 0)
                tile_part[t] = s;
            continue;
        }
        const uint32_t nv = (tile_v[t + 1] & ~WG_HEAVY_BIT) - v0;
        uint32_t offw[NOFF];
#pragma unroll
        for (int k = 0; k < NOFF; ++k) {
            const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
            offw[k] = (i <= nv) ? off[v0 + i] : 0u;
        }
        uint32_t odw[NVTX];
        if (MODE == 0) {
#pragma unroll
            for (int k = 0; k < NVTX; ++k) {
                const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
                odw[k] = (i < nv) ? od[v0 + i] : 0u;
            }
        }
        double g[WG_IPT];
#pragma unroll
        for (int j = 0; j < WG_IPT; ++j)
            g[j] = (a + j < e1) ? load_contrib(cin, s_idx[j]) : 0.0;
#pragma unroll
        for (int k = 0; k < NOFF; ++k) {
            const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
            if (i <= nv)
                s_off[i] = offw[k];
        }
        wg_wave_sync();

        double fp = 0.0, lp = 0.0;
        if (a < e1) {
            // owner of edge a: the largest window index with s_off[i] <= a
            uint32_t lo = 0, hi = nv - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= a)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            uint32_t d = lo;
            uint32_t dend = s_off[d + 1];
            bool cont = s_off[d] < a; // this piece continues a destination from an earlier lane
            bool first = true;
            double run = 0.0;
#pragma unroll
            for (int j = 0; j < WG_IPT; ++j) {
                const uint32_t e = a + j;
                if (e < e1) {
                    run += g[j];
                    if (e + 1 == dend) {
                        if (first && cont)
                            fp = run;
                        else
                            s_acc[d] = run;
                        run = 0.0;
                        first = false;
                        cont = false;
                        ++d;
                        while (d < nv && s_off[d + 1] == s_off[d])
                            ++d;
                        dend = (d < nv) ? s_off[d + 1] : 0xffffffffu;
                    }
                }
            }
            const uint32_t b = (a + WG_IPT < e1) ? a + WG_IPT : e1;
            if (b < dend && d < nv) {
                // an open piece remains at the end of the run
                if (first && cont)
                    fp = run;
                else
                    lp = run;
            }
        }
        s_fp[lane] = fp;
        s_lp[lane] = lp;
        wg_wave_sync();

#pragma unroll
        for (int k = 0; k < NVTX; ++k) {
            const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
            if (i < nv) {
                const uint32_t vs = s_off[i], ve = s_off[i + 1];
                double accv = 0.0;
                if (vs != ve) {
                    const uint32_t k0 = (vs - e0) / WG_IPT, k1 = (ve - 1 - e0) / WG_IPT;
                    if (k0 == k1) {
                        accv = s_acc[i];
                    } else {
                        accv = s_lp[k0];
                        for (uint32_t q = k0 + 1; q <= k1; ++q)
                            accv += s_fp[q];
                    }
                }
                const uint32_t v = v0 + i;
                if (MODE == 0) {
                    const uint32_t odv = odw[k];
                    const double t1 = accv + corr;
                    const double t2 = t1 * damping;
                    const double rank = t2 + omd_n;
                    cout[v] = (T)(rank / (odv ? (double)odv : N));
                    if (odv)
                        s_part += rank;
                } else {
                    acc_out[v] = accv;
                    s_part += accv;
                }
            }
        }
        wg_wave_sync(); // the next tile reuses this wavefront's LDS
    }
    s_part = block_sum_det(s_part, s_red);
    if (threadIdx.x == 0)
        blk_part[blockIdx.x] = s_part;
}

// Deterministic reduction of n partials by the calling block.
__device__ double reduce_partials(const double* p, uint32_t n, double* s_red)
{
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += WG_BLOCK)
        s += __builtin_nontemporal_load(&p[i]);
    return block_sum_det(s, s_red);
}

// Heavy destinations (sum of their tile partials in tile order) and the final
// reduction of all block partials by the last block to finish.
template <typename T, int MODE>
__global__ __launch_bounds__(WG_BLOCK) void k_pr_span(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ heavy_v,
                                                       const uint32_t* __restrict__ heavy_t, uint32_t nheavy,
                                                       const double* __restrict__ tile_part, T* __restrict__ cout,
                                                       double* __restrict__ acc_out, const uint32_t* __restrict__ od,
                                                       const double* __restrict__ S_in, double damping, double N,
                                                       double* __restrict__ blk_part, uint32_t main_blocks,
                                                       double* __restrict__ S_out, unsigned int* __restrict__ ticket)
{
    __shared__ double s_red[WG_WPB];
    __shared__ bool s_last;
    const int lane = wg_lane();
    double corr = 0.0, omd_n = 0.0;
    if (MODE == 0) {
        corr = (1.0 - *S_in) / N;
        omd_n = (1.0 - damping) / N;
    }
    double s_part = 0.0;
    const uint32_t waves = gridDim.x * (WG_BLOCK / WG_WAVE);
    for (uint32_t h = blockIdx.x * (WG_BLOCK / WG_WAVE) + threadIdx.x / WG_WAVE; h < nheavy; h += waves) {
        const uint32_t v = heavy_v[h], t0 = heavy_t[h];
        const uint32_t nch = (off[v + 1] - off[v] + WG_TILE_EDGES - 1) / WG_TILE_EDGES;
        double s = 0.0;
        for (uint32_t c = lane; c < nch; c += WG_WAVE)
            s += tile_part[t0 + c];
        s = wg_wave_sum(s);
        if (lane == 0) {
            if (MODE == 0) {
                const uint32_t odv = od[v];
                const double t1 = s + corr;
                const double t2 = t1 * damping;
                const double rank = t2 + omd_n;
                cout[v] = (T)(rank / (odv ? (double)odv : N));
                if (odv)
                    s_part += rank;
            } else {
                acc_out[v] = s;
                s_part += s;
            }
        }
    }
    s_part = block_sum_det(s_part, s_red);
    if (threadIdx.x == 0) {
        blk_part[main_blocks + blockIdx.x] = s_part;
        __threadfence();
        const unsigned int prev = atomicAdd(ticket, 1u);
        s_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        const double total = reduce_partials(blk_part, main_blocks + gridDim.x, s_red);
        if (threadIdx.x == 0) {
            *S_out = total;
            *ticket = 0u;
        }
    }
}

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
    int slot = wg_prof_begin(ctx, mode == 0 ? WG_K_PR_TILE : WG_K_PR_ACC);
    if (mode == 0)
        hipLaunchKernelGGL((k_pr_tile<T, 0>), dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->src, cin,
                           cout, ctx->acc, ctx->od, ctx->tile_v, ctx->tile_e, ctx->ntiles, S_in, ctx->damping, N,
                           ctx->tile_part, ctx->blk_part);
    else
        hipLaunchKernelGGL((k_pr_tile<T, 1>), dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->src, cin,
                           cout, ctx->acc, ctx->od, ctx->tile_v, ctx->tile_e, ctx->ntiles, S_in, ctx->damping, N,
                           ctx->tile_part, ctx->blk_part);
    int rc = wg_check_launch(ctx);
    if (rc)
        return rc;
    wg_prof_end(ctx, WG_K_PR_TILE, slot);
    slot = wg_prof_begin(ctx, WG_K_PR_SPAN);
    if (mode == 0)
        hipLaunchKernelGGL((k_pr_span<T, 0>), dim3(ctx->span_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->heavy_v,
                           ctx->heavy_t, ctx->nheavy, ctx->tile_part, cout, ctx->acc, ctx->od, S_in, ctx->damping, N,
                           ctx->blk_part, ctx->pr_grid, S_out, &ctx->ctr->ticket);
    else
        hipLaunchKernelGGL((k_pr_span<T, 1>), dim3(ctx->span_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->heavy_v,
                           ctx->heavy_t, ctx->nheavy, ctx->tile_part, cout, ctx->acc, ctx->od, S_in, ctx->damping, N,
                           ctx->blk_part, ctx->pr_grid, S_out, &ctx->ctr->ticket);
    rc = wg_check_launch(ctx);
    if (rc)
        return rc;
    wg_prof_end(ctx, WG_K_PR_SPAN, slot);
    return WG_OK;
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
