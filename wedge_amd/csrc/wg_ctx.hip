// wg_ctx.hip -- context, graph upload and derived device structures, readback,
// per-kernel timing.
//
// Replaces the reference's ingress + data-structure initialisation
// (source/graphdata.c:648-942, :946-1133) for the device: instead of 4-lane
// edge vectors with a per-destination vector index and a CSRI of vector
// groups, the device holds a 32-bit CSC plus the derived arrays listed in
// wg_internal.h.
#include <hipcub/hipcub.hpp>

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "wg_internal.h"

// ---- small kernels ------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t value)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t value)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Destination column: a block walks one tile; light tiles binary-search the
// tile's vertex window, heavy tiles belong to one vertex.
__global__ __launch_bounds__(WG_BLOCK) void k_build_dst(const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ tile_v,
                                                         const uint32_t* __restrict__ tile_e, uint32_t ntiles,
                                                         uint32_t* __restrict__ dst)
{
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const bool heavy = (tile_v[t] & WG_HEAVY_BIT) != 0u;
        const uint32_t v0 = tile_v[t] & ~WG_HEAVY_BIT, v1 = tile_v[t + 1] & ~WG_HEAVY_BIT;
        const uint32_t e0 = tile_e[t], e1 = tile_e[t + 1];
        for (uint32_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
            uint32_t d = v0;
            if (!heavy) {
                uint32_t lo = v0, hi = v1 - 1; // largest v with off[v] <= e
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (off[mid] <= e)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                d = lo;
            }
            dst[e] = d;
        }
    }
}

// Input check on the caller's CSC: sources in range and ascending within each
// destination.
__global__ void k_validate(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src, uint64_t V, uint64_t E,
                           unsigned int* __restrict__ bad)
{
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < E; p += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = V - 1;
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) >> 1;
            if (off[mid] <= p)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint32_t s = src[p];
        bool ok = s < V;
        if (p + 1 < off[lo + 1])
            ok = ok && s <= src[p + 1];
        if (!ok)
            atomicOr(bad, 1u);
    }
}

__global__ void k_outdeg(const uint32_t* __restrict__ src, uint64_t E, uint32_t* __restrict__ od)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&od[src[i]], 1u);
}

__global__ void k_csri_scatter(const uint32_t* __restrict__ src, uint64_t E, uint32_t* __restrict__ cursor,
                               uint32_t* __restrict__ pos)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t at = atomicAdd(&cursor[src[i]], 1u);
        pos[at] = (uint32_t)i;
    }
}

// ---- degree relabelling ------------------------------------------------------------

__global__ void k_degree_key(const uint32_t* __restrict__ off, const uint32_t* __restrict__ od, uint64_t V,
                             uint32_t* __restrict__ key, uint32_t* __restrict__ ids)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t d = (uint64_t)(off[v + 1] - off[v]) + od[v];
        key[v] = d > 0xffffffffull ? 0xffffffffu : (uint32_t)d;
        ids[v] = (uint32_t)v;
    }
}

// inv[perm[i]] = i; nin[i] = in-degree of the vertex now at position i
__global__ void k_invert(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ off, uint64_t V,
                         uint32_t* __restrict__ inv, uint32_t* __restrict__ nin)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < V; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = perm[i];
        inv[o] = (uint32_t)i;
        nin[i] = off[o + 1] - off[o];
    }
}

// New CSC: position p of the relabelled layout takes the edge at the same rank
// of the same (original) destination; sources are renamed, their original
// order inside each destination is kept.
template <typename W>
__global__ void k_relabel_edges(const uint32_t* __restrict__ noff, const uint32_t* __restrict__ perm,
                                const uint32_t* __restrict__ ooff, const uint32_t* __restrict__ osrc,
                                const W* __restrict__ ow, const uint32_t* __restrict__ inv, uint64_t V, uint64_t E,
                                uint32_t* __restrict__ nsrc, W* __restrict__ nw)
{
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < E; p += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = V - 1; // largest i with noff[i] <= p
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) >> 1;
            if (noff[mid] <= p)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint64_t op = ooff[perm[lo]] + (p - noff[lo]);
        nsrc[p] = inv[osrc[op]];
        if (nw)
            nw[p] = ow[op];
    }
}

template <typename T>
__global__ void k_to_original(const T* __restrict__ in, const uint32_t* __restrict__ inv, uint64_t V, T* __restrict__ out)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = in[inv[v]];
}

// ---- helpers shared with other translation units -------------------------------

int wg_fill_u32(wg_ctx* ctx, uint32_t* p, uint64_t n, uint32_t value)
{
    if (n == 0)
        return WG_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(wg_grid(ctx, n, 1024, 8)), dim3(256), 0, ctx->stream, p, n, value);
    return wg_check_launch(ctx);
}

int wg_fill_u64(wg_ctx* ctx, uint64_t* p, uint64_t n, uint64_t value)
{
    if (n == 0)
        return WG_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(wg_grid(ctx, n, 1024, 8)), dim3(256), 0, ctx->stream, p, n, value);
    return wg_check_launch(ctx);
}

int wg_clear_u64(wg_ctx* ctx, uint64_t* p, uint64_t words)
{
    return wg_fill_u64(ctx, p, words, 0);
}

// ---- profiling --------------------------------------------------------------------

static int prof_flush(wg_ctx* ctx)
{
    wg_prof* p = &ctx->prof;
    if (p->npending == 0)
        return WG_OK;
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < p->npending; ++i) {
        float ms = 0.f;
        WG_HIP(ctx, hipEventElapsedTime(&ms, p->ev[2 * i], p->ev[2 * i + 1]));
        p->total_ms[p->kid[i]] += (double)ms;
        p->launches[p->kid[i]] += 1;
    }
    p->npending = 0;
    return WG_OK;
}

int wg_prof_begin(wg_ctx* ctx, int kid)
{
    wg_prof* p = &ctx->prof;
    if (!p->enabled)
        return -1;
    if (p->npending == p->cap) {
        if (prof_flush(ctx) != WG_OK)
            return -1;
    }
    const int slot = p->npending++;
    p->kid[slot] = kid;
    if (hipEventRecord(p->ev[2 * slot], ctx->stream) != hipSuccess)
        return -1;
    return slot;
}

int wg_prof_end(wg_ctx* ctx, int kid, int slot)
{
    (void)kid;
    if (slot < 0 || !ctx->prof.enabled)
        return WG_OK;
    WG_HIP(ctx, hipEventRecord(ctx->prof.ev[2 * slot + 1], ctx->stream));
    return WG_OK;
}

extern "C" int wg_profile_enable(wg_ctx* ctx, int on)
{
    if (!ctx)
        return WG_E_INVAL;
    wg_prof* p = &ctx->prof;
    if (on && !p->ev) {
        p->cap = 2048;
        p->ev = (hipEvent_t*)calloc(2 * (size_t)p->cap, sizeof(hipEvent_t));
        p->kid = (int*)calloc((size_t)p->cap, sizeof(int));
        if (!p->ev || !p->kid)
            return WG_E_NOMEM;
        for (int i = 0; i < 2 * p->cap; ++i)
            WG_HIP(ctx, hipEventCreate(&p->ev[i]));
    }
    if (!on) {
        int rc = prof_flush(ctx);
        if (rc)
            return rc;
    }
    p->enabled = on ? 1 : 0;
    return WG_OK;
}

extern "C" int wg_profile_get(wg_ctx* ctx, int kid, uint64_t* launches, double* total_ms)
{
    if (!ctx || kid < 0 || kid >= WG_K_COUNT)
        return WG_E_INVAL;
    int rc = prof_flush(ctx);
    if (rc)
        return rc;
    if (launches)
        *launches = ctx->prof.launches[kid];
    if (total_ms)
        *total_ms = ctx->prof.total_ms[kid];
    return WG_OK;
}

extern "C" int wg_profile_reset(wg_ctx* ctx)
{
    if (!ctx)
        return WG_E_INVAL;
    int rc = prof_flush(ctx);
    if (rc)
        return rc;
    memset(ctx->prof.launches, 0, sizeof(ctx->prof.launches));
    memset(ctx->prof.total_ms, 0, sizeof(ctx->prof.total_ms));
    return WG_OK;
}


extern "C" int wg_set_option(wg_ctx* ctx, int option, int64_t value)
{
    if (!ctx)
        return WG_E_INVAL;
    switch (option) {
    case WG_OPT_RELABEL:
        if (value < WG_RELABEL_OFF || value > WG_RELABEL_AUTO)
            return WG_E_INVAL;
        ctx->relabel = (int)value;
        return WG_OK;
    case WG_OPT_MS_BINNED:
        if (value < 0 || value > 2)
            return WG_E_INVAL;
        ctx->ms_binned = (int)value + 1; // 0 = unset: the default
        return WG_OK;
    case WG_OPT_MIN_BINNED:
        if (value < 0 || value > 2)
            return WG_E_INVAL;
        ctx->min_binned = (int)value + 1; // 0 = unset: the default
        return WG_OK;
    case WG_OPT_COUNT_WORK:
        if (value < 0 || value > 1)
            return WG_E_INVAL;
        ctx->count_work = (int)value;
        return WG_OK;
    case WG_OPT_PR_BIN_SLICE:
        if (value < 1 || value > (1 << 24))
            return WG_E_INVAL;
        ctx->pb_slice = (uint32_t)value;
        if (ctx->pb) { // rebuilt with the new size on the next binned iteration
            WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
            wg_free_pb(ctx);
        }
        return WG_OK;
    default:
        return WG_E_INVAL;
    }
}

int wg_to_device_id(wg_ctx* ctx, uint64_t orig, uint32_t* out)
{
    if (orig >= ctx->V)
        return WG_E_RANGE;
    if (!ctx->inv) {
        *out = (uint32_t)orig;
        return WG_OK;
    }
    uint32_t v = 0;
    WG_HIP(ctx, hipMemcpyAsync(&v, ctx->inv + orig, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = v;
    return WG_OK;
}

__global__ void k_map_ids(const uint64_t* __restrict__ in, uint32_t n, const uint32_t* __restrict__ inv,
                          uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = inv[in[i]];
}

int wg_to_device_ids(wg_ctx* ctx, const uint64_t* orig, uint32_t n, uint32_t* out)
{
    for (uint32_t i = 0; i < n; ++i)
        if (orig[i] >= ctx->V)
            return WG_E_RANGE;
    if (!ctx->inv || n == 0) {
        for (uint32_t i = 0; i < n; ++i)
            out[i] = (uint32_t)orig[i];
        return WG_OK;
    }
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const size_t need = (size_t)n * (sizeof(uint64_t) + sizeof(uint32_t));
    if (ctx->idmap_bytes < need) {
        if (ctx->idmap)
            (void)hipFree(ctx->idmap);
        ctx->idmap = nullptr;
        ctx->idmap_bytes = 0;
        WG_HIP(ctx, hipMalloc(&ctx->idmap, need));
        ctx->idmap_bytes = need;
    }
    uint64_t* din = (uint64_t*)ctx->idmap;
    uint32_t* dout = (uint32_t*)(din + n);
    WG_HIP(ctx, hipMemcpyAsync(din, orig, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_map_ids, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, (const uint64_t*)din, n,
                       (const uint32_t*)ctx->inv, dout);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipMemcpyAsync(out, dout, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return WG_OK;
}

int wg_read_vertex_array(wg_ctx* ctx, const void* dev, void* host, size_t esz)
{
    if (!ctx->inv) {
        WG_HIP(ctx, hipMemcpyAsync(host, dev, ctx->V * esz, hipMemcpyDeviceToHost, ctx->stream));
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return WG_OK;
    }
    void* tmp = nullptr;
    WG_HIP(ctx, hipMalloc(&tmp, ctx->V * esz));
    const uint32_t grid = wg_grid(ctx, ctx->V, 1024, 8);
    if (esz == 8)
        hipLaunchKernelGGL(k_to_original<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream, (const uint64_t*)dev, ctx->inv,
                           ctx->V, (uint64_t*)tmp);
    else
        hipLaunchKernelGGL(k_to_original<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream, (const uint32_t*)dev, ctx->inv,
                           ctx->V, (uint32_t*)tmp);
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipMemcpyAsync(host, tmp, ctx->V * esz, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e3 = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        ctx->last_hip_error = (int)(e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3);
        return WG_E_HIP;
    }
    return WG_OK;
}

// ---- context ----------------------------------------------------------------------

extern "C" const char* wg_error_string(int code)
{
    switch (code) {
    case WG_OK: return "ok";
    case WG_E_INVAL: return "invalid argument";
    case WG_E_NOMEM: return "out of memory";
    case WG_E_HIP: return "HIP runtime error";
    case WG_E_STATE: return "call not valid in the current state";
    case WG_E_RANGE: return "value out of range";
    case WG_E_IO: return "I/O error";
    case WG_E_FORMAT: return "malformed input";
    case WG_E_OVERFLOW: return "integer distance overflow";
    case WG_E_NODEV: return "no HIP device";
    default: return "unknown error";
    }
}

extern "C" int wg_last_hip_error(const wg_ctx* ctx)
{
    return ctx ? ctx->last_hip_error : 0;
}

extern "C" int wg_open(wg_ctx** out, int device)
{
    if (!out)
        return WG_E_INVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return WG_E_NODEV;
    if (device < 0 || device >= n)
        return WG_E_INVAL;
    wg_ctx* ctx = (wg_ctx*)calloc(1, sizeof(wg_ctx));
    if (!ctx)
        return WG_E_NOMEM;
    ctx->device = device;
    ctx->relabel = WG_RELABEL_AUTO;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        free(ctx);
        return WG_E_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cu = prop.multiProcessorCount;
    else
        ctx->num_cu = 256;
    if (hipMalloc((void**)&ctx->ctr, sizeof(wg_dev_counters)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->hctr, sizeof(wg_dev_counters), hipHostMallocDefault) != hipSuccess ||
        hipMemset(ctx->ctr, 0, sizeof(wg_dev_counters)) != hipSuccess) {
        if (ctx->ctr)
            (void)hipFree(ctx->ctr);
        (void)hipStreamDestroy(ctx->stream);
        free(ctx);
        return WG_E_NOMEM;
    }
    memset(ctx->hctr, 0, sizeof(wg_dev_counters));
    if (hipHostGetDevicePointer((void**)&ctx->hctr_dev, ctx->hctr, 0) != hipSuccess) {
        ctx->hctr_dev = nullptr; // wg_take_activity copies instead
        (void)hipGetLastError();
    }
    ctx->run_wgs[WG_APP_BFS] = WG_RUN_DEFAULT_WGS;
    *out = ctx;
    return WG_OK;
}

static void free_dev(void** p)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
}

int wg_free_app_state(wg_ctx* ctx)
{
    free_dev(&ctx->prop[0]);
    free_dev(&ctx->prop[1]);
    free_dev(&ctx->ms[0]);
    free_dev(&ctx->ms[1]);
    free_dev(&ctx->ms[2]);
    free_dev(&ctx->ms[3]);
    ctx->ms_n = 0;
    free_dev((void**)&ctx->acc);
    free_dev((void**)&ctx->roots);
    free_dev((void**)&ctx->root_off);
    ctx->nroots = 0;
    ctx->app = WG_APP_NONE;
    return WG_OK;
}

static void free_graph(wg_ctx* ctx)
{
    wg_free_members(ctx);
    wg_free_pb(ctx);
    wg_free_app_state(ctx);
    free_dev((void**)&ctx->off);
    free_dev((void**)&ctx->src);
    free_dev((void**)&ctx->dst);
    free_dev(&ctx->w);
    free_dev((void**)&ctx->od);
    free_dev((void**)&ctx->csri_off);
    free_dev((void**)&ctx->csri_pos);
    free_dev((void**)&ctx->tile_v);
    free_dev((void**)&ctx->tile_e);
    free_dev((void**)&ctx->heavy_v);
    free_dev((void**)&ctx->heavy_t);
    free_dev((void**)&ctx->bfs_heavy_v);
    free_dev((void**)&ctx->bfs_heavy_e);
    free_dev((void**)&ctx->tile_part);
    free_dev((void**)&ctx->blk_part);
    for (auto& t : ctx->pr_alt) {
        free_dev((void**)&t.tile_v);
        free_dev((void**)&t.tile_e);
        free_dev((void**)&t.heavy_v);
        free_dev((void**)&t.heavy_t);
        t.ntiles = t.nheavy = 0;
    }
    ctx->pr_alt_built = 0;
    ctx->pr_ipt = WG_IPT;
    ctx->pr_bpc = 8;
    ctx->pr_xcd = 0;
    free_dev((void**)&ctx->bm_cur);
    free_dev((void**)&ctx->bm_next);
    free_dev((void**)&ctx->bm_visited);
    free_dev((void**)&ctx->wedge);
    free_dev((void**)&ctx->heavy_list);
    free_dev((void**)&ctx->run_spare);
    free_dev((void**)&ctx->first_src);
    free_dev((void**)&ctx->csr_dst);
    free_dev((void**)&ctx->perm);
    free_dev((void**)&ctx->inv);
    ctx->V = ctx->E = 0;
    ctx->ntiles = ctx->nheavy = ctx->nbfs_heavy = 0;
}

extern "C" void wg_close(wg_ctx* ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_graph(ctx);
    if (ctx->prof.ev) {
        for (int i = 0; i < 2 * ctx->prof.cap; ++i)
            (void)hipEventDestroy(ctx->prof.ev[i]);
        free(ctx->prof.ev);
        free(ctx->prof.kid);
    }
    (void)hipFree(ctx->mlp);
    (void)hipFree(ctx->idmap);
    (void)hipFree(ctx->run_ctl);
    (void)hipFree(ctx->run_q);
    (void)hipFree(ctx->ctr);
    (void)hipHostFree(ctx->hctr);
    (void)hipStreamDestroy(ctx->stream);
    free(ctx);
}

// Streaming copy, 16 bytes per lane per step: the bandwidth an HBM-bound
// kernel reaches on this device.
typedef unsigned int wg_u32x4 __attribute__((ext_vector_type(4)));
__global__ void k_stream_copy(const wg_u32x4* __restrict__ in, wg_u32x4* __restrict__ out, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

extern "C" int wg_measure_copy(wg_ctx* ctx, uint64_t bytes, double* gb_per_s)
{
    if (!ctx || !gb_per_s || bytes < 16)
        return WG_E_INVAL;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const uint64_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = WG_OK;
    float best = 0.0f;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return WG_E_NOMEM;
    }
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess ||
        hipMemsetAsync(a, 0, n * 16, ctx->stream) != hipSuccess) {
        rc = WG_E_HIP;
    } else {
        const uint32_t grid = ctx->num_cu * 4; // the fastest of 2..32 per CU, with or without non-temporal hints
        for (int rep = 0; rep < 6 && rc == WG_OK; ++rep) {
            float t = 0.0f;
            const hipError_t e0 = hipEventRecord(ev[0], ctx->stream);
            hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, (const wg_u32x4*)a,
                               (wg_u32x4*)b, n);
            const hipError_t e1 = hipEventRecord(ev[1], ctx->stream);
            if (e0 != hipSuccess || e1 != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess)
                rc = WG_E_HIP;
            else if (rep > 0 && (best == 0.0f || t < best)) // the first run warms up
                best = t;
        }
    }
    if (ev[0])
        (void)hipEventDestroy(ev[0]);
    if (ev[1])
        (void)hipEventDestroy(ev[1]);
    (void)hipFree(a);
    (void)hipFree(b);
    if (rc == WG_OK)
        *gb_per_s = 2.0 * (double)(n * 16) / ((double)best * 1e-3) / 1e9;
    return rc;
}

extern "C" int wg_sync(wg_ctx* ctx)
{
    if (!ctx)
        return WG_E_INVAL;
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return WG_OK;
}

// Host-side tiling of the destinations (see wg_internal.h) for tiles of at
// most tile_edges edges and WG_TILE_VERTS destinations.
void wg_build_tiles(const uint32_t* off, uint64_t V, uint32_t tile_edges, std::vector<uint32_t>& tv,
                    std::vector<uint32_t>& te, std::vector<uint32_t>& hv, std::vector<uint32_t>& ht)
{
    uint64_t v = 0;
    while (v < V) {
        const uint32_t d = off[v + 1] - off[v];
        if (d > tile_edges) {
            hv.push_back((uint32_t)v);
            ht.push_back((uint32_t)tv.size());
            for (uint64_t e = off[v]; e < off[v + 1]; e += tile_edges) {
                tv.push_back((uint32_t)v | WG_HEAVY_BIT);
                te.push_back((uint32_t)e);
            }
            ++v;
            continue;
        }
        tv.push_back((uint32_t)v);
        te.push_back(off[v]);
        uint32_t ecount = 0, vcount = 0;
        while (v < V && vcount < WG_TILE_VERTS) {
            const uint32_t dd = off[v + 1] - off[v];
            if (dd > tile_edges || ecount + dd > tile_edges)
                break;
            ecount += dd;
            ++vcount;
            ++v;
        }
    }
    tv.push_back((uint32_t)V);
    te.push_back(off[V]);
}

static void build_tiles(const uint32_t* off, uint64_t V, std::vector<uint32_t>& tv, std::vector<uint32_t>& te,
                        std::vector<uint32_t>& hv, std::vector<uint32_t>& ht, std::vector<uint32_t>& bv,
                        std::vector<uint32_t>& be)
{
    wg_build_tiles(off, V, WG_TILE_EDGES, tv, te, hv, ht);
    for (uint64_t u = 0; u < V; ++u) {
        const uint32_t d = off[u + 1] - off[u];
        if (d > WG_BFS_HEAVY)
            for (uint64_t e = off[u]; e < off[u + 1]; e += WG_BFS_HEAVY) {
                bv.push_back((uint32_t)u);
                be.push_back((uint32_t)e);
            }
    }
}

template <typename T>
static int upload(wg_ctx* ctx, T** dptr, const T* host, uint64_t n)
{
    WG_HIP(ctx, hipMalloc((void**)dptr, (n ? n : 1) * sizeof(T)));
    if (n && host)
        WG_HIP(ctx, hipMemcpy(*dptr, host, n * sizeof(T), hipMemcpyHostToDevice));
    return WG_OK;
}

// First tile starting at or after x/8 of the edges, x = 0..8.
void wg_tile_split_points(const std::vector<uint32_t>& te, uint64_t E, uint32_t start[9])
{
    const uint32_t ntiles = (uint32_t)(te.size() - 1);
    for (int x = 0; x <= 8; ++x)
        start[x] = (x == 8) ? ntiles
                            : (uint32_t)(std::lower_bound(te.begin(), te.end() - 1, (uint32_t)(E * (uint64_t)x / 8)) -
                                         te.begin());
}

// Tilings with 4, 16 and 2 edges per lane for the PageRank shapes, from the
// device offsets (in device order).
int wg_build_pr_tilings(wg_ctx* ctx)
{
    if (ctx->pr_alt_built)
        return WG_OK;
    std::vector<uint32_t> hoff(ctx->V + 1);
    WG_HIP(ctx, hipMemcpy(hoff.data(), ctx->off, (ctx->V + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
    uint32_t max_tiles = ctx->ntiles;
    const uint32_t ipts[3] = {4, 16, 2};
    for (int i = 0; i < 3; ++i) {
        std::vector<uint32_t> tv, te, hv, ht;
        wg_build_tiles(hoff.data(), ctx->V, WG_WAVE * ipts[i], tv, te, hv, ht);
        wg_tiling& t = ctx->pr_alt[i];
        t.ntiles = (uint32_t)(tv.size() - 1);
        t.nheavy = (uint32_t)hv.size();
        wg_tile_split_points(te, ctx->E, t.split);
        max_tiles = std::max(max_tiles, t.ntiles);
        int rc;
        if ((rc = upload(ctx, &t.tile_v, tv.data(), tv.size())) || (rc = upload(ctx, &t.tile_e, te.data(), te.size())) ||
            (rc = upload(ctx, &t.heavy_v, hv.data(), hv.size())) || (rc = upload(ctx, &t.heavy_t, ht.data(), ht.size())))
            return rc;
    }
    if (max_tiles > ctx->ntiles) { // heavy-tile partials for the finest tiling
        free_dev((void**)&ctx->tile_part);
        int rc = upload<double>(ctx, &ctx->tile_part, nullptr, (uint64_t)max_tiles + 1);
        if (rc)
            return rc;
    }
    ctx->pr_alt_built = 1;
    return WG_OK;
}


// Relabelling pays off when a few vertices hold many edges (their values are
// gathered over and over and end up in one place); on near-regular graphs such
// as road networks the caller's order usually keeps neighbours closer.
static bool skewed_degrees(const uint32_t* off, uint64_t V, uint64_t E)
{
    uint32_t maxin = 0;
    for (uint64_t v = 0; v < V; ++v) {
        const uint32_t d = off[v + 1] - off[v];
        maxin = d > maxin ? d : maxin;
    }
    const double mean = (double)E / (double)V;
    return maxin > 64u && (double)maxin > 16.0 * mean;
}

// Replaces the device CSC (off/src/w/od, original numbering) by the same graph
// with vertices in descending total-degree order (ties by original id), and
// returns the new offsets on the host for tiling.
static int relabel_graph(wg_ctx* ctx, std::vector<uint32_t>& hoff)
{
    const uint64_t V = ctx->V, E = ctx->E;
    const size_t wsz = ctx->wkind == WG_WEIGHT_F64 ? 8 : ctx->wkind == WG_WEIGHT_U32 ? 4 : 0;
    uint32_t *key = nullptr, *ids = nullptr, *key2 = nullptr, *perm = nullptr, *inv = nullptr, *nin = nullptr;
    uint32_t *noff = nullptr, *nsrc = nullptr, *nod = nullptr;
    void *nw = nullptr, *tmp = nullptr;
    int rc = WG_OK;
    const uint32_t gv = wg_grid(ctx, V, 1024, 8), ge = wg_grid(ctx, E, 1024, 8);
    size_t b1 = 0, b2 = 0;
    const uint32_t e32 = (uint32_t)E; // read by the stream before the synchronise below
    hipError_t e = hipSuccess;
#define RL(call)                          \
    do {                                  \
        if ((e = (call)) != hipSuccess) { \
            rc = WG_E_HIP;                \
            goto done;                    \
        }                                 \
    } while (0)
    RL(hipMalloc((void**)&key, V * 4));
    RL(hipMalloc((void**)&ids, V * 4));
    RL(hipMalloc((void**)&key2, V * 4));
    RL(hipMalloc((void**)&perm, V * 4));
    RL(hipMalloc((void**)&inv, V * 4));
    RL(hipMalloc((void**)&nin, V * 4));
    RL(hipMalloc((void**)&noff, (V + 1) * 4));
    RL(hipMalloc((void**)&nsrc, (E ? E : 1) * 4));
    RL(hipMalloc((void**)&nod, V * 4));
    if (wsz)
        RL(hipMalloc(&nw, (E ? E : 1) * wsz));
    hipLaunchKernelGGL(k_degree_key, dim3(gv), dim3(256), 0, ctx->stream, ctx->off, ctx->od, V, key, ids);
    RL(hipGetLastError());
    RL(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b1, key, key2, ids, perm, (int)V, 0, 32, ctx->stream));
    RL(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, nin, noff, (int)V, ctx->stream));
    RL(hipMalloc(&tmp, (b1 > b2 ? b1 : b2) + 16));
    RL(hipcub::DeviceRadixSort::SortPairsDescending(tmp, b1, key, key2, ids, perm, (int)V, 0, 32, ctx->stream));
    hipLaunchKernelGGL(k_invert, dim3(gv), dim3(256), 0, ctx->stream, perm, ctx->off, V, inv, nin);
    RL(hipGetLastError());
    RL(hipcub::DeviceScan::ExclusiveSum(tmp, b2, nin, noff, (int)V, ctx->stream));
    RL(hipMemcpyAsync(noff + V, &e32, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    if (E) {
        if (wsz == 8)
            hipLaunchKernelGGL(k_relabel_edges<uint64_t>, dim3(ge), dim3(256), 0, ctx->stream, noff, perm, ctx->off,
                               ctx->src, (const uint64_t*)ctx->w, inv, V, E, nsrc, (uint64_t*)nw);
        else if (wsz == 4)
            hipLaunchKernelGGL(k_relabel_edges<uint32_t>, dim3(ge), dim3(256), 0, ctx->stream, noff, perm, ctx->off,
                               ctx->src, (const uint32_t*)ctx->w, inv, V, E, nsrc, (uint32_t*)nw);
        else
            hipLaunchKernelGGL(k_relabel_edges<uint32_t>, dim3(ge), dim3(256), 0, ctx->stream, noff, perm, ctx->off,
                               ctx->src, (const uint32_t*)nullptr, inv, V, E, nsrc, (uint32_t*)nullptr);
        RL(hipGetLastError());
    }
    hipLaunchKernelGGL(k_to_original<uint32_t>, dim3(gv), dim3(256), 0, ctx->stream, ctx->od, perm, V, nod);
    RL(hipGetLastError());
    hoff.resize(V + 1);
    RL(hipMemcpyAsync(hoff.data(), noff, (V + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    RL(hipStreamSynchronize(ctx->stream));
    // swap in the relabelled arrays
    (void)hipFree(ctx->off);
    (void)hipFree(ctx->src);
    (void)hipFree(ctx->od);
    if (ctx->w)
        (void)hipFree(ctx->w);
    ctx->off = noff;
    ctx->src = nsrc;
    ctx->od = nod;
    ctx->w = nw;
    ctx->perm = perm;
    ctx->inv = inv;
    noff = nsrc = nod = perm = inv = nullptr;
    nw = nullptr;
done:
#undef RL
    if (rc)
        ctx->last_hip_error = (int)e;
    (void)hipFree(key);
    (void)hipFree(ids);
    (void)hipFree(key2);
    (void)hipFree(nin);
    (void)hipFree(tmp);
    if (noff)
        (void)hipFree(noff);
    if (nsrc)
        (void)hipFree(nsrc);
    if (nod)
        (void)hipFree(nod);
    if (nw)
        (void)hipFree(nw);
    if (perm)
        (void)hipFree(perm);
    if (inv)
        (void)hipFree(inv);
    return rc;
}

void wg_drop_graph(wg_ctx* ctx)
{
    free_graph(ctx);
}

extern "C" int wg_load_csc(wg_ctx* ctx, uint64_t V, uint64_t E, const uint32_t* off, const uint32_t* src,
                           const void* weights, int weight_kind)
{
    if (!ctx || !off || (E && !src) || V == 0)
        return WG_E_INVAL;
    if (V >= (uint64_t)WG_HEAVY_BIT || E >= 0xffffffffull)
        return WG_E_RANGE;
    if (weight_kind != WG_WEIGHT_NONE && weight_kind != WG_WEIGHT_F64 && weight_kind != WG_WEIGHT_U32)
        return WG_E_INVAL;
    if (weight_kind != WG_WEIGHT_NONE && E && !weights)
        return WG_E_INVAL;
    if (off[0] != 0 || off[V] != E)
        return WG_E_FORMAT;
    for (uint64_t v = 0; v < V; ++v)
        if (off[v + 1] < off[v])
            return WG_E_FORMAT;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_graph(ctx);

    int rc;
    std::vector<uint32_t> tv, te, hv, ht, bv, be;
    std::vector<uint32_t> hoff;
    const uint32_t* toff = off;
    ctx->V = V;
    ctx->E = E;
    ctx->wkind = weight_kind;
    if ((rc = upload(ctx, &ctx->off, off, V + 1)) || (rc = upload(ctx, &ctx->src, src, E)) ||
        (rc = upload<uint32_t>(ctx, &ctx->od, nullptr, V)))
        goto fail;
    if (weight_kind == WG_WEIGHT_F64) {
        if ((rc = upload(ctx, (double**)&ctx->w, (const double*)weights, E)))
            goto fail;
    } else if (weight_kind == WG_WEIGHT_U32) {
        if ((rc = upload(ctx, (uint32_t**)&ctx->w, (const uint32_t*)weights, E)))
            goto fail;
    }
    {
        // validate the caller's CSC, then out-degrees in the original numbering
        unsigned int* bad = &ctx->ctr->overflow;
        WG_HIP(ctx, hipMemsetAsync(bad, 0, sizeof(unsigned int), ctx->stream));
        if (E) {
            hipLaunchKernelGGL(k_validate, dim3(wg_grid(ctx, E, 1024, 8)), dim3(256), 0, ctx->stream, ctx->off, ctx->src,
                               V, E, bad);
            if ((rc = wg_check_launch(ctx)))
                goto fail;
        }
        unsigned int hbad = 0;
        WG_HIP(ctx, hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, ctx->stream));
        WG_HIP(ctx, hipMemsetAsync(ctx->od, 0, V * sizeof(uint32_t), ctx->stream));
        if (E) {
            hipLaunchKernelGGL(k_outdeg, dim3(wg_grid(ctx, E, 1024, 8)), dim3(256), 0, ctx->stream, ctx->src, E, ctx->od);
            if ((rc = wg_check_launch(ctx)))
                goto fail;
        }
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (hbad) {
            rc = WG_E_FORMAT;
            goto fail;
        }
    }
    if (V > 1 && (ctx->relabel == WG_RELABEL_ON || (ctx->relabel == WG_RELABEL_AUTO && skewed_degrees(off, V, E)))) {
        if ((rc = relabel_graph(ctx, hoff)))
            goto fail;
        toff = hoff.data();
    }
    tv.reserve(E / WG_TILE_EDGES + V / WG_TILE_VERTS + 16);
    te.reserve(tv.capacity());
    build_tiles(toff, V, tv, te, hv, ht, bv, be);
    if ((rc = upload<uint32_t>(ctx, &ctx->dst, nullptr, E)) ||
        (rc = upload<uint32_t>(ctx, &ctx->csri_off, nullptr, V + 1)) ||
        (rc = upload<uint32_t>(ctx, &ctx->csri_pos, nullptr, E)))
        goto fail;
    ctx->ntiles = (uint32_t)(tv.size() - 1);
    ctx->nheavy = (uint32_t)hv.size();
    ctx->nbfs_heavy = (uint32_t)bv.size();
    if ((rc = upload(ctx, &ctx->tile_v, tv.data(), tv.size())) || (rc = upload(ctx, &ctx->tile_e, te.data(), te.size())) ||
        (rc = upload(ctx, &ctx->heavy_v, hv.data(), hv.size())) || (rc = upload(ctx, &ctx->heavy_t, ht.data(), ht.size())) ||
        (rc = upload(ctx, &ctx->bfs_heavy_v, bv.data(), bv.size())) ||
        (rc = upload(ctx, &ctx->bfs_heavy_e, be.data(), be.size())))
        goto fail;

    {
        if (E) {
            hipLaunchKernelGGL(k_build_dst, dim3(wg_grid(ctx, ctx->ntiles, 1, 16)), dim3(WG_BLOCK), 0, ctx->stream,
                               ctx->off, ctx->tile_v, ctx->tile_e, ctx->ntiles, ctx->dst);
            if ((rc = wg_check_launch(ctx)))
                goto fail;
        }
        // the source -> edge-position index (the reference's CSRI)
        size_t tmp_bytes = 0;
        WG_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, ctx->od, ctx->csri_off, (int)V, ctx->stream));
        void* tmp = nullptr;
        uint32_t* cursor = nullptr;
        WG_HIP(ctx, hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1));
        if (hipMalloc((void**)&cursor, V * sizeof(uint32_t)) != hipSuccess) {
            (void)hipFree(tmp);
            rc = WG_E_NOMEM;
            goto fail;
        }
        hipError_t e1 = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, ctx->od, ctx->csri_off, (int)V, ctx->stream);
        const uint32_t e32 = (uint32_t)E; // read by the stream before the synchronise below
        hipError_t e2 = hipMemcpyAsync(ctx->csri_off + V, &e32, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream);
        hipError_t e3 = hipMemcpyAsync(cursor, ctx->csri_off, V * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream);
        if (E)
            hipLaunchKernelGGL(k_csri_scatter, dim3(wg_grid(ctx, E, 1024, 8)), dim3(256), 0, ctx->stream, ctx->src, E,
                               cursor, ctx->csri_pos);
        hipError_t e4 = hipGetLastError();
        hipError_t e5 = hipStreamSynchronize(ctx->stream);
        (void)hipFree(tmp);
        (void)hipFree(cursor);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess || e5 != hipSuccess) {
            ctx->last_hip_error = (int)(e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3 != hipSuccess ? e3 : e4 != hipSuccess ? e4 : e5);
            rc = WG_E_HIP;
            goto fail;
        }
    }
    {
        // number of vertices with out-degree > 0 (the first PageRank sum)
        std::vector<uint32_t> hod(V);
        WG_HIP(ctx, hipMemcpy(hod.data(), ctx->od, V * sizeof(uint32_t), hipMemcpyDeviceToHost));
        uint32_t n_out = 0;
        for (uint64_t v = 0; v < V; ++v)
            n_out += hod[v] ? 1u : 0u;
        ctx->n_out_pos = n_out;
    }
    // scratch shared by the apps
    ctx->bm_words = (V + 63) / 64;
    ctx->wedge_words = (((E + (1ull << WG_MIN_BLOCK_LOG2) - 1) >> WG_MIN_BLOCK_LOG2) + 31) / 32 + 1;
    ctx->pr_grid = wg_grid(ctx, ctx->ntiles, WG_WPB, 8);
    wg_tile_split_points(te, E, ctx->pr_split8);
    ctx->heavy_cap = (uint32_t)(E / 32 + E / 4096 + 16); // bound on deferred out-edge chunks per build
    ctx->span_grid = wg_grid(ctx, ctx->nheavy, 4, 4);
    if ((rc = upload<uint64_t>(ctx, &ctx->bm_cur, nullptr, ctx->bm_words)) ||
        (rc = upload<uint64_t>(ctx, &ctx->bm_next, nullptr, ctx->bm_words)) ||
        (rc = upload<uint64_t>(ctx, &ctx->bm_visited, nullptr, ctx->bm_words)) ||
        (rc = upload<uint32_t>(ctx, &ctx->wedge, nullptr, ctx->wedge_words)) ||
        (rc = upload<uint32_t>(ctx, &ctx->heavy_list, nullptr, 2 * (uint64_t)ctx->heavy_cap)) ||
        (rc = upload<double>(ctx, &ctx->tile_part, nullptr, ctx->ntiles + 1)) ||
        (rc = upload<double>(ctx, &ctx->blk_part, nullptr, (uint64_t)ctx->num_cu * 20 + 32)))
        goto fail;
    if ((rc = wg_clear_u64(ctx, ctx->bm_cur, ctx->bm_words)) || (rc = wg_clear_u64(ctx, ctx->bm_next, ctx->bm_words)) ||
        (rc = wg_clear_u64(ctx, ctx->bm_visited, ctx->bm_words)) || (rc = wg_fill_u32(ctx, ctx->wedge, ctx->wedge_words, 0)))
        goto fail;
    WG_HIP(ctx, hipMemsetAsync(ctx->ctr, 0, sizeof(wg_dev_counters), ctx->stream));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return WG_OK;
fail:
    free_graph(ctx);
    return rc;
}

extern "C" int wg_graph_info(const wg_ctx* ctx, uint64_t* V, uint64_t* E)
{
    if (!ctx)
        return WG_E_INVAL;
    if (V)
        *V = ctx->V;
    if (E)
        *E = ctx->E;
    return WG_OK;
}

extern "C" int wg_read_outdegrees(wg_ctx* ctx, uint32_t* host_out)
{
    if (!ctx || !host_out || !ctx->od)
        return WG_E_INVAL;
    return wg_read_vertex_array(ctx, ctx->od, host_out, sizeof(uint32_t));
}

extern "C" int wg_vertex_outdegree(wg_ctx* ctx, uint64_t vertex, uint32_t* out)
{
    if (!ctx || !out || !ctx->od)
        return WG_E_INVAL;
    uint32_t dv = 0;
    int rc = wg_to_device_id(ctx, vertex, &dv);
    if (rc)
        return rc;
    uint32_t v = 0;
    WG_HIP(ctx, hipMemcpyAsync(&v, ctx->od + dv, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = v;
    return WG_OK;
}
