// wg_ppr.hip -- personalised PageRank for up to 32 seeds in one pass over the
// edges (no counterpart in the reference, whose PageRank has a uniform
// teleport: source/phase_vertex_pr.c:26-77).
//
// For seed s (iteration t -> t+1, d = damping):
//   r_s(v) = d * (sum over in-edges u->v of c_s(u) + [v = s] * D_s) + [v = s] * (1 - d)
//   c_s(u) = r_s(u) / od(u) for od(u) > 0, 0 otherwise
//   D_s    = sum of r_s over vertices without out-edges (their mass returns to the seed)
// starting from r_s = [v = s].  Every iteration keeps sum_v r_s(v) = 1.
//
// Layout: contributions C[v][32] in FP32 (a 128-byte row per vertex, so an
// in-edge gathers one coalesced row, lane s = seed s), ranks R[v][32] in FP64
// written by the last iteration.  Sums run in FP64 in a fixed order (each
// destination's in-edges in CSC order; heavy destinations through per-tile
// partials summed in tile order; the dangling mass through per-wave partials
// summed in wave order), so results are reproducible run to run.
#include "wg_frontier.h"

#include <vector>

#define PPR_LANES 32

// the rank of v for seed lane s from its accumulated in-edge sum (fixed order
// of operations: d * (acc + [seed] * D) + [seed] * (1 - d))
__device__ __forceinline__ double ppr_rank(double acc, bool is_seed, double dang, double d)
{
    const double a = is_seed ? acc + dang : acc;
    const double r = a * d;
    return is_seed ? r + (1.0 - d) : r;
}

__global__ void k_ppr_init(const uint32_t* __restrict__ seeds, uint32_t nseeds, const uint32_t* __restrict__ od,
                           float* __restrict__ C, double* __restrict__ D, uint64_t V)
{
    // zero every row, then the seeds' rows
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < V * PPR_LANES;
         i += (uint64_t)gridDim.x * blockDim.x)
        C[i] = 0.0f;
}

__global__ void k_ppr_seeds(const uint32_t* __restrict__ seeds, uint32_t nseeds, const uint32_t* __restrict__ od,
                            float* __restrict__ C, double* __restrict__ D)
{
    const uint32_t s = threadIdx.x;
    if (s >= PPR_LANES)
        return;
    if (s >= nseeds) {
        D[s] = 0.0;
        return;
    }
    const uint32_t v = seeds[s], o = od[v];
    C[(uint64_t)v * PPR_LANES + s] = o ? (float)(1.0 / (double)o) : 0.0f;
    D[s] = o ? 0.0 : 1.0;
}

// One iteration over the edge tiles, one wavefront per tile.  Light tile: the
// two half-waves take alternate destinations, lane s of a half summing seed
// s.  Heavy tile (part of one destination): the halves take alternate edges,
// the tile's sum goes to part[t][s].
__global__ __launch_bounds__(WG_BLOCK) void k_ppr_tile(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                        const uint32_t* __restrict__ od, const uint32_t* __restrict__ tile_v,
                                                        const uint32_t* __restrict__ tile_e, uint32_t ntiles,
                                                        const uint32_t* __restrict__ seeds, uint32_t nseeds,
                                                        const float* __restrict__ C, float* __restrict__ Cn,
                                                        double* __restrict__ R, double* __restrict__ part,
                                                        const double* __restrict__ D, double* __restrict__ wpart,
                                                        double damping)
{
    const int lane = wg_lane();
    const uint32_t half = (uint32_t)lane >> 5, s = (uint32_t)lane & 31u;
    const uint32_t seed = s < nseeds ? seeds[s] : 0xffffffffu;
    const double dang = D[s];
    const uint32_t wave = blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE;
    const uint32_t stride = gridDim.x * WG_WPB;
    double dsum = 0.0; // mass of this lane's destinations without out-edges
    for (uint32_t t = wave; t < ntiles; t += stride) {
        const uint32_t tv = tile_v[t];
        const uint32_t e0 = tile_e[t], e1 = tile_e[t + 1];
        const uint32_t v0 = tv & ~WG_HEAVY_BIT;
        if (tv & WG_HEAVY_BIT) {
            double a = 0.0;
            uint32_t p = e0 + half;
            for (; p + 6 < e1; p += 8) { // four edges in flight per half
                const uint32_t u0 = src[p], u1 = src[p + 2], u2 = src[p + 4], u3 = src[p + 6];
                const float x0 = C[(uint64_t)u0 * PPR_LANES + s], x1 = C[(uint64_t)u1 * PPR_LANES + s];
                const float x2 = C[(uint64_t)u2 * PPR_LANES + s], x3 = C[(uint64_t)u3 * PPR_LANES + s];
                a += (double)x0;
                a += (double)x1;
                a += (double)x2;
                a += (double)x3;
            }
            for (; p < e1; p += 2)
                a += (double)C[(uint64_t)src[p] * PPR_LANES + s];
            const double lo = __shfl(a, (int)s), hi = __shfl(a, (int)s + 32);
            if (lane < 32)
                part[(uint64_t)t * PPR_LANES + s] = lo + hi;
            continue;
        }
        const uint32_t nv = (tile_v[t + 1] & ~WG_HEAVY_BIT) - v0;
        for (uint32_t i = half; i < nv; i += 2) {
            const uint32_t v = v0 + i;
            const uint32_t b = off[v], e = off[v + 1];
            double a = 0.0;
            uint32_t p = b;
            for (; p + 3 < e; p += 4) {
                const uint32_t u0 = src[p], u1 = src[p + 1], u2 = src[p + 2], u3 = src[p + 3];
                const float x0 = C[(uint64_t)u0 * PPR_LANES + s], x1 = C[(uint64_t)u1 * PPR_LANES + s];
                const float x2 = C[(uint64_t)u2 * PPR_LANES + s], x3 = C[(uint64_t)u3 * PPR_LANES + s];
                a += (double)x0;
                a += (double)x1;
                a += (double)x2;
                a += (double)x3;
            }
            for (; p < e; ++p)
                a += (double)C[(uint64_t)src[p] * PPR_LANES + s];
            const double r = ppr_rank(a, v == seed, dang, damping);
            const uint32_t o = od[v];
            Cn[(uint64_t)v * PPR_LANES + s] = o ? (float)(r / (double)o) : 0.0f;
            if (R)
                R[(uint64_t)v * PPR_LANES + s] = r;
            if (!o)
                dsum += r;
        }
    }
    const double lo = __shfl(dsum, (int)s), hi = __shfl(dsum, (int)s + 32);
    if (lane < 32)
        wpart[(uint64_t)wave * PPR_LANES + s] = lo + hi;
}

// Heavy destinations: the tile partials in tile order, then the rank.
__global__ __launch_bounds__(WG_BLOCK) void k_ppr_span(const uint32_t* __restrict__ off, const uint32_t* __restrict__ heavy_v,
                                                        const uint32_t* __restrict__ heavy_t, uint32_t nheavy,
                                                        const uint32_t* __restrict__ od, const uint32_t* __restrict__ seeds,
                                                        uint32_t nseeds, const double* __restrict__ part,
                                                        float* __restrict__ Cn, double* __restrict__ R,
                                                        const double* __restrict__ D, double* __restrict__ spart,
                                                        double damping)
{
    const int lane = wg_lane();
    const uint32_t s = (uint32_t)lane & 31u;
    const uint32_t seed = s < nseeds ? seeds[s] : 0xffffffffu;
    const double dang = D[s];
    const uint32_t wave = blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE;
    const uint32_t stride = gridDim.x * WG_WPB;
    double dsum = 0.0;
    for (uint32_t h = wave; h < nheavy; h += stride) {
        const uint32_t v = heavy_v[h], t0 = heavy_t[h];
        const uint32_t nch = (off[v + 1] - off[v] + WG_TILE_EDGES - 1) / WG_TILE_EDGES;
        double a = 0.0;
        for (uint32_t c = 0; c < nch; ++c)
            a += part[(uint64_t)(t0 + c) * PPR_LANES + s];
        const double r = ppr_rank(a, v == seed, dang, damping);
        const uint32_t o = od[v];
        if (lane < 32) {
            Cn[(uint64_t)v * PPR_LANES + s] = o ? (float)(r / (double)o) : 0.0f;
            if (R)
                R[(uint64_t)v * PPR_LANES + s] = r;
        }
        if (!o)
            dsum += r;
    }
    if (lane < 32)
        spart[(uint64_t)wave * PPR_LANES + s] = dsum;
}

// D_next[s] = the per-wave dangling partials in wave order.
// Dangling mass per seed: the n partial rows p[n][32] (wave partials, then
// span partials) summed per lane in a fixed order.  Block b takes the rows
// [b n / B, (b + 1) n / B); its thread t sums lane t % 32 over the rows of
// group t / 32 (every 8th row, four running sums), the 8 group sums are added
// in group order, and the last block to finish adds the B block sums in block
// order.
#define PPR_RED_BLOCKS 64
__global__ __launch_bounds__(256) void k_ppr_reduce(const double* __restrict__ p, uint32_t n, double* __restrict__ bpart,
                                                    unsigned int* __restrict__ ticket, double* __restrict__ Dn)
{
    __shared__ double s_grp[8][PPR_LANES];
    __shared__ bool s_last;
    const uint32_t lane = threadIdx.x & (PPR_LANES - 1), grp = threadIdx.x / PPR_LANES;
    const uint32_t r0 = (uint32_t)((uint64_t)n * blockIdx.x / gridDim.x);
    const uint32_t r1 = (uint32_t)((uint64_t)n * (blockIdx.x + 1) / gridDim.x);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    uint32_t r = r0 + grp;
    for (; r + 24 < r1; r += 32) {
        a0 += __builtin_nontemporal_load(&p[(uint64_t)r * PPR_LANES + lane]);
        a1 += __builtin_nontemporal_load(&p[(uint64_t)(r + 8) * PPR_LANES + lane]);
        a2 += __builtin_nontemporal_load(&p[(uint64_t)(r + 16) * PPR_LANES + lane]);
        a3 += __builtin_nontemporal_load(&p[(uint64_t)(r + 24) * PPR_LANES + lane]);
    }
    for (; r < r1; r += 8)
        a0 += __builtin_nontemporal_load(&p[(uint64_t)r * PPR_LANES + lane]);
    s_grp[grp][lane] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if (threadIdx.x < PPR_LANES) {
        double b = 0.0;
        for (int g = 0; g < 8; ++g)
            b += s_grp[g][lane];
        bpart[(uint64_t)blockIdx.x * PPR_LANES + lane] = b;
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0)
        s_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (s_last) {
        __threadfence();
        if (threadIdx.x < PPR_LANES) {
            double d = 0.0;
            for (uint32_t b = 0; b < gridDim.x; ++b)
                d += __builtin_nontemporal_load(&bpart[(uint64_t)b * PPR_LANES + lane]);
            Dn[lane] = d;
        }
        if (threadIdx.x == 0)
            *ticket = 0u;
    }
}

__global__ void k_ppr_column(const double* __restrict__ R, uint32_t s, uint64_t V, double* __restrict__ out)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = R[v * PPR_LANES + s];
}

extern "C" int wg_ppr(wg_ctx* ctx, const uint64_t* seeds, uint32_t nseeds, uint32_t iterations, double damping)
{
    if (!ctx || !seeds)
        return WG_E_INVAL;
    if (!ctx->off)
        return WG_E_STATE;
    if (nseeds == 0 || nseeds > WG_PPR_MAX_SEEDS || iterations == 0 || !(damping >= 0.0 && damping <= 1.0))
        return WG_E_INVAL;
    std::vector<uint32_t> hs(nseeds);
    int rc;
    if ((rc = wg_to_device_ids(ctx, seeds, nseeds, hs.data())))
        return rc;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint64_t V = ctx->V;
    const uint32_t grid = ctx->pr_grid, sgrid = ctx->span_grid;
    const uint32_t nw = grid * WG_WPB, ns = sgrid * WG_WPB;
    void *C0 = nullptr, *C1 = nullptr, *R = nullptr, *aux = nullptr;
    // aux: seeds (32 u32) | D[2][32] | tile partials [ntiles][32] | wave partials | span partials |
    //      reduction block partials [PPR_RED_BLOCKS][32] | reduction ticket (u32)
    const size_t aux_doubles = 2 * PPR_LANES + ((size_t)ctx->ntiles + nw + ns + PPR_RED_BLOCKS) * PPR_LANES;
    const size_t aux_bytes = PPR_LANES * sizeof(uint32_t) + aux_doubles * sizeof(double) + sizeof(unsigned int);
    if (ctx->app == WG_APP_PPR && ctx->ms_aux_bytes == aux_bytes) {
        // the last call's buffers (sized by V, which a new graph frees): the
        // init below rewrites C0, every iteration the other matrix, the last
        // one R, and the partials are rewritten before they are read
        C0 = ctx->ms[0];
        C1 = ctx->ms[1];
        aux = ctx->ms[2];
        R = ctx->prop[0];
    } else {
        wg_free_app_state(ctx);
        if (hipMalloc(&C0, V * PPR_LANES * sizeof(float)) != hipSuccess ||
            hipMalloc(&C1, V * PPR_LANES * sizeof(float)) != hipSuccess ||
            hipMalloc(&R, V * PPR_LANES * sizeof(double)) != hipSuccess || hipMalloc(&aux, aux_bytes) != hipSuccess) {
            (void)hipFree(C0);
            (void)hipFree(C1);
            (void)hipFree(R);
            (void)hipFree(aux);
            return WG_E_NOMEM;
        }
        ctx->ms[0] = C0;
        ctx->ms[1] = C1;
        ctx->ms[2] = aux;
        ctx->prop[0] = R;
        ctx->ms_aux_bytes = aux_bytes;
        ctx->app = WG_APP_PPR;
    }
    ctx->ms_n = nseeds;
    uint32_t* dseeds = (uint32_t*)aux;
    double* D = (double*)((char*)aux + PPR_LANES * sizeof(uint32_t));
    double* part = D + 2 * PPR_LANES;
    double* wpart = part + (size_t)ctx->ntiles * PPR_LANES;
    double* spart = wpart + (size_t)nw * PPR_LANES;
    double* rpart = spart + (size_t)ns * PPR_LANES;
    unsigned int* ticket = (unsigned int*)(rpart + (size_t)PPR_RED_BLOCKS * PPR_LANES);
    WG_HIP(ctx, hipMemsetAsync(ticket, 0, sizeof(unsigned int), ctx->stream));
    WG_HIP(ctx, hipMemcpyAsync(dseeds, hs.data(), nseeds * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_ppr_init, dim3(wg_grid(ctx, V * PPR_LANES, 1024, 8)), dim3(256), 0, ctx->stream, dseeds, nseeds,
                       ctx->od, (float*)C0, D, V);
    hipLaunchKernelGGL(k_ppr_seeds, dim3(1), dim3(64), 0, ctx->stream, dseeds, nseeds, ctx->od, (float*)C0, D);
    if ((rc = wg_check_launch(ctx)))
        return rc;
    float* cur = (float*)C0;
    float* nxt = (float*)C1;
    for (uint32_t it = 0; it < iterations; ++it) {
        double* Dcur = D + (it & 1u) * PPR_LANES;
        double* Dnxt = D + ((it + 1u) & 1u) * PPR_LANES;
        double* Rout = (it + 1 == iterations) ? (double*)R : nullptr;
        hipLaunchKernelGGL(k_ppr_tile, dim3(grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->src, ctx->od, ctx->tile_v,
                           ctx->tile_e, ctx->ntiles, dseeds, nseeds, cur, nxt, Rout, part, Dcur, wpart, damping);
        hipLaunchKernelGGL(k_ppr_span, dim3(sgrid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->heavy_v, ctx->heavy_t,
                           ctx->nheavy, ctx->od, dseeds, nseeds, part, nxt, Rout, Dcur, spart, damping);
        hipLaunchKernelGGL(k_ppr_reduce, dim3(PPR_RED_BLOCKS), dim3(256), 0, ctx->stream, wpart, nw + ns, rpart, ticket,
                           Dnxt);
        if ((rc = wg_check_launch(ctx)))
            return rc;
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return WG_OK;
}

extern "C" int wg_read_ppr(wg_ctx* ctx, uint32_t seed, double* host_ranks)
{
    if (!ctx || !host_ranks)
        return WG_E_INVAL;
    if (ctx->app != WG_APP_PPR)
        return WG_E_STATE;
    if (seed >= ctx->ms_n)
        return WG_E_RANGE;
    double* tmp = nullptr;
    WG_HIP(ctx, hipMalloc((void**)&tmp, ctx->V * sizeof(double)));
    hipLaunchKernelGGL(k_ppr_column, dim3(wg_grid(ctx, ctx->V, 1024, 8)), dim3(256), 0, ctx->stream,
                       (const double*)ctx->prop[0], seed, ctx->V, tmp);
    int rc = wg_check_launch(ctx);
    if (!rc)
        rc = wg_read_vertex_array(ctx, tmp, host_ranks, sizeof(double));
    (void)hipFree(tmp);
    return rc;
}
