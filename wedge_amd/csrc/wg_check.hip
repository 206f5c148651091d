// wg_check.hip -- on-device validation of BFS, CC and SSSP results against
// the loaded graph (no counterpart in the reference, which prints results
// unchecked).  Graph500-style certificates, checked with one pass over the
// edges and one over the vertices:
//   BFS   roots have depth 0 and themselves as parent; every edge u -> v with
//         u reached has v reached and depth(v) <= depth(u) + 1; every other
//         reached v has its parent p as an in-neighbour with
//         depth(p) = depth(v) - 1.
//   SSSP  roots have distance 0; every edge u -> v with u reached has v
//         reached and dist(v) <= dist(u) + w (one IEEE add for FP64); every
//         other reached v has an in-edge with dist(u) + w == dist(v).  With
//         positive weights this pins the shortest-path distances.
//   CC    every edge u -> v has label(v) <= label(u); label(v) <= v; and
//         label(v) == v or some in-neighbour u has label(u) == label(v).
//         Together: label(v) = min of {v} and the vertices reaching v.
// A violation is one failing edge or vertex; the count is returned.
#include "wg_frontier.h"

#include <mutex>
#include <vector>

namespace {

__device__ __forceinline__ void count_bad(unsigned long long n, unsigned long long* bad)
{
    n = wg_wave_sum(n);
    if (wg_lane() == 0 && n)
        atomicAdd(bad, n);
}

__device__ __forceinline__ uint32_t orig_id(const uint32_t* perm, uint32_t v)
{
    return perm ? perm[v] : v;
}

__device__ __forceinline__ bool bit_of(const uint64_t* bm, uint32_t v)
{
    return (bm[v >> 6] >> (v & 63)) & 1ull;
}

__device__ __forceinline__ void set_bit(uint64_t* bm, uint32_t v)
{
    atomicOr((unsigned long long*)&bm[v >> 6], 1ull << (v & 63));
}

} // namespace

__global__ void k_chk_roots(const uint32_t* __restrict__ roots, uint32_t n, uint64_t* __restrict__ rootbm)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        set_bit(rootbm, roots[i]);
}

// Values given in original numbering -> device numbering.
template <typename T>
__global__ void k_chk_to_device(const T* __restrict__ in, const uint32_t* __restrict__ perm, uint64_t V,
                                T* __restrict__ out)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = in[perm[v]];
}

// BFS packed words: depth << 32 | parent (original id), UINT64_MAX unreached.
__global__ void k_chk_bfs_edges(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E,
                                const uint64_t* __restrict__ packed, const uint32_t* __restrict__ perm,
                                uint64_t* __restrict__ ok, unsigned long long* __restrict__ bad)
{
    unsigned long long nbad = 0;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = src[e], v = dst[e];
        const uint64_t pu = packed[u], pv = packed[v];
        if (pu == ~0ull)
            continue;
        const uint64_t du = pu >> 32;
        if (pv == ~0ull || (pv >> 32) > du + 1) {
            ++nbad;
            continue;
        }
        if ((pv >> 32) == du + 1 && (uint32_t)pv == orig_id(perm, u))
            set_bit(ok, v);
    }
    count_bad(nbad, bad);
}

__global__ void k_chk_bfs_vertices(const uint64_t* __restrict__ packed, uint64_t V, const uint64_t* __restrict__ rootbm,
                                   const uint32_t* __restrict__ perm, const uint64_t* __restrict__ ok,
                                   unsigned long long* __restrict__ bad)
{
    unsigned long long nbad = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t pv = packed[v];
        if (bit_of(rootbm, (uint32_t)v))
            nbad += (pv != (uint64_t)orig_id(perm, (uint32_t)v)) ? 1 : 0; // depth 0, parent itself
        else if (pv != ~0ull && ((pv >> 32) == 0 || !bit_of(ok, (uint32_t)v)))
            ++nbad;
    }
    count_bad(nbad, bad);
}

// SSSP: FP64 bit patterns (INF = +inf) or u32 (INF = 0xFFFFFFFF).
template <bool F64>
__global__ void k_chk_sssp_edges(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E,
                                 const void* __restrict__ w, const void* __restrict__ dist, uint64_t* __restrict__ ok,
                                 unsigned long long* __restrict__ bad)
{
    unsigned long long nbad = 0;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = src[e], v = dst[e];
        if (F64) {
            const double du = ((const double*)dist)[u], dv = ((const double*)dist)[v];
            if (du == __builtin_inf())
                continue;
            const double t = du + ((const double*)w)[e]; // one IEEE add (no contraction)
            if (!(dv <= t))
                ++nbad;
            else if (dv == t)
                set_bit(ok, v);
        } else {
            const uint32_t du = ((const uint32_t*)dist)[u], dv = ((const uint32_t*)dist)[v];
            if (du == 0xffffffffu)
                continue;
            const uint64_t t = (uint64_t)du + ((const uint32_t*)w)[e];
            if ((uint64_t)dv > t || dv == 0xffffffffu)
                ++nbad;
            else if ((uint64_t)dv == t)
                set_bit(ok, v);
        }
    }
    count_bad(nbad, bad);
}

template <bool F64>
__global__ void k_chk_sssp_vertices(const void* __restrict__ dist, uint64_t V, const uint64_t* __restrict__ rootbm,
                                    const uint64_t* __restrict__ ok, unsigned long long* __restrict__ bad)
{
    unsigned long long nbad = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        bool reached, zero;
        if (F64) {
            const double d = ((const double*)dist)[v];
            reached = d != __builtin_inf();
            zero = d == 0.0;
        } else {
            const uint32_t d = ((const uint32_t*)dist)[v];
            reached = d != 0xffffffffu;
            zero = d == 0u;
        }
        if (bit_of(rootbm, (uint32_t)v))
            nbad += zero ? 0 : 1;
        else if (reached && !bit_of(ok, (uint32_t)v))
            ++nbad;
    }
    count_bad(nbad, bad);
}

// Edges are grouped by destination, so a wavefront's 64 edges hold a few runs
// of one destination each: the first lane of a run sets the destination's ok
// bit once if any edge of the run justifies its label (on a graph with few
// components nearly every edge does, and one atomic per edge on one word
// serialised the kernel).  Every lane of a wavefront takes the same steps.
__global__ void k_chk_cc_edges(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E,
                               const uint32_t* __restrict__ label, uint64_t* __restrict__ ok,
                               unsigned long long* __restrict__ bad)
{
    unsigned long long nbad = 0;
    const int lane = wg_lane();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e0 = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~(WG_WAVE - 1u)); e0 < E; e0 += stride) {
        const uint64_t e = e0 + (uint64_t)lane;
        bool eq = false;
        uint32_t d = 0xffffffffu;
        if (e < E) {
            d = dst[e];
            const uint32_t lu = label[src[e]], lv = label[d];
            if (lv > lu)
                ++nbad;
            else
                eq = lv == lu;
        }
        const uint32_t dprev = (uint32_t)__shfl_up((int)d, 1, WG_WAVE);
        const unsigned long long heads = wg_ballot(lane == 0 || dprev != d);
        const unsigned long long eqm = wg_ballot(eq);
        if (((heads >> lane) & 1ull) && e < E) {
            const unsigned long long above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
            const unsigned long long upto = above ? (above & (0ull - above)) - 1ull : ~0ull; // lanes below the next run
            if (eqm & upto & ~((1ull << lane) - 1ull))
                set_bit(ok, d);
        }
    }
    count_bad(nbad, bad);
}

__global__ void k_chk_cc_vertices(const uint32_t* __restrict__ label, uint64_t V, const uint32_t* __restrict__ perm,
                                  const uint64_t* __restrict__ ok, unsigned long long* __restrict__ bad)
{
    unsigned long long nbad = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = orig_id(perm, (uint32_t)v), l = label[v];
        if (l > o || (l != o && !bit_of(ok, (uint32_t)v)))
            ++nbad;
    }
    count_bad(nbad, bad);
}

// ---- host side ------------------------------------------------------------------

// Runs the checks on device-order values (vals) with device-id roots, in the
// buffers ok and rootbm (bm_words words each) and the counter bad, which it
// clears first.
static int run_checks_in(wg_ctx* ctx, int app, const void* vals, const uint32_t* droots, uint32_t nroots,
                         uint64_t* ok, uint64_t* rootbm, unsigned long long* bad, uint64_t* violations)
{
    const uint64_t V = ctx->V, E = ctx->E, words = ctx->bm_words;
    int rc = WG_OK;
    unsigned long long hbad = 0;
    if ((rc = wg_clear_u64(ctx, ok, words)) || (rc = wg_clear_u64(ctx, rootbm, words)) ||
        (rc = wg_clear_u64(ctx, (uint64_t*)bad, 1)))
        return rc;
    if (nroots)
        hipLaunchKernelGGL(k_chk_roots, dim3((nroots + 255) / 256), dim3(256), 0, ctx->stream, droots, nroots, rootbm);
    {
        const uint32_t eg = wg_grid(ctx, E, 1024, 8), vg = wg_grid(ctx, V, 1024, 8);
        if (app == WG_APP_BFS) {
            if (E)
                hipLaunchKernelGGL(k_chk_bfs_edges, dim3(eg), dim3(256), 0, ctx->stream, ctx->src, ctx->dst, E,
                                   (const uint64_t*)vals, ctx->perm, ok, bad);
            hipLaunchKernelGGL(k_chk_bfs_vertices, dim3(vg), dim3(256), 0, ctx->stream, (const uint64_t*)vals, V, rootbm,
                               ctx->perm, ok, bad);
        } else if (app == WG_APP_SSSP) {
            const bool f64 = ctx->wkind == WG_WEIGHT_F64;
            if (E) {
                if (f64)
                    hipLaunchKernelGGL(k_chk_sssp_edges<true>, dim3(eg), dim3(256), 0, ctx->stream, ctx->src, ctx->dst,
                                       E, ctx->w, vals, ok, bad);
                else
                    hipLaunchKernelGGL(k_chk_sssp_edges<false>, dim3(eg), dim3(256), 0, ctx->stream, ctx->src, ctx->dst,
                                       E, ctx->w, vals, ok, bad);
            }
            if (f64)
                hipLaunchKernelGGL(k_chk_sssp_vertices<true>, dim3(vg), dim3(256), 0, ctx->stream, vals, V, rootbm, ok,
                                   bad);
            else
                hipLaunchKernelGGL(k_chk_sssp_vertices<false>, dim3(vg), dim3(256), 0, ctx->stream, vals, V, rootbm, ok,
                                   bad);
        } else {
            if (E)
                hipLaunchKernelGGL(k_chk_cc_edges, dim3(eg), dim3(256), 0, ctx->stream, ctx->src, ctx->dst, E,
                                   (const uint32_t*)vals, ok, bad);
            hipLaunchKernelGGL(k_chk_cc_vertices, dim3(vg), dim3(256), 0, ctx->stream, (const uint32_t*)vals, V, ctx->perm,
                               ok, bad);
        }
    }
    if ((rc = wg_check_launch(ctx)))
        return rc;
    if (hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        ctx->last_hip_error = (int)hipGetLastError();
        return WG_E_HIP;
    }
    *violations = hbad;
    return WG_OK;
}

// run_checks_in with buffers allocated for the call.
static int run_checks(wg_ctx* ctx, int app, const void* vals, const uint32_t* droots, uint32_t nroots,
                      uint64_t* violations)
{
    const uint64_t words = ctx->bm_words;
    uint64_t* ok = nullptr;
    uint64_t* rootbm = nullptr;
    unsigned long long* bad = nullptr;
    if (hipMalloc((void**)&ok, words * 8) != hipSuccess || hipMalloc((void**)&rootbm, words * 8) != hipSuccess ||
        hipMalloc((void**)&bad, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(ok);
        (void)hipFree(rootbm);
        return WG_E_NOMEM;
    }
    const int rc = run_checks_in(ctx, app, vals, droots, nroots, ok, rootbm, bad, violations);
    (void)hipStreamSynchronize(ctx->stream); // the kernels are done with the buffers before they are freed
    (void)hipFree(ok);
    (void)hipFree(rootbm);
    (void)hipFree(bad);
    return rc;
}

extern "C" int wg_validate(wg_ctx* ctx, uint64_t* violations)
{
    if (!ctx || !violations)
        return WG_E_INVAL;
    if (ctx->app != WG_APP_BFS && ctx->app != WG_APP_CC && ctx->app != WG_APP_SSSP)
        return WG_E_STATE;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    return run_checks(ctx, ctx->app, ctx->prop[0], ctx->roots, ctx->app == WG_APP_CC ? 0u : ctx->nroots, violations);
}

// ---- allocation-free validation -----------------------------------------------------
//
// wg_validate_noalloc keeps its two bitmaps and its counter in `scratch`, a
// device array of fixed size that every GPU holds from the library's load until
// the process ends.  Host code never passes the array's name to a kernel: the
// name is a host-side stand-in, and each device has its own copy at its own
// address, which hipGetSymbolAddress returns with that device current.
#define WG_NOALLOC_WORDS (WG_NOALLOC_MAX_VERTICES / 64)
#define WG_NOALLOC_SCRATCH_WORDS (2 * WG_NOALLOC_WORDS + 32) // ok, roots, counter line
static __device__ __align__(256) uint64_t scratch[WG_NOALLOC_SCRATCH_WORDS];

#define WG_NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static uint64_t* g_scratch_ptr[WG_NOALLOC_MAX_DEVICES];

// Device address of this device's scratch, looked up once per device (the
// device must be current).
static int scratch_on_device(wg_ctx* ctx, uint64_t** out)
{
    if (ctx->device < 0 || ctx->device >= WG_NOALLOC_MAX_DEVICES)
        return WG_E_RANGE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ctx->last_hip_error = (int)e;
            return WG_E_HIP;
        }
        g_scratch_ptr[ctx->device] = (uint64_t*)p;
    }
    *out = g_scratch_ptr[ctx->device];
    return WG_OK;
}

extern "C" int wg_validate_noalloc(wg_ctx* ctx, uint64_t* violations)
{
    if (!ctx || !violations)
        return WG_E_INVAL;
    if (ctx->app != WG_APP_BFS && ctx->app != WG_APP_CC && ctx->app != WG_APP_SSSP)
        return WG_E_STATE;
    if (ctx->V > WG_NOALLOC_MAX_VERTICES || ctx->bm_words > WG_NOALLOC_WORDS)
        return WG_E_RANGE;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    uint64_t* base = nullptr;
    int rc = scratch_on_device(ctx, &base);
    if (rc)
        return rc;
    uint64_t* ok = base;
    uint64_t* rootbm = base + WG_NOALLOC_WORDS;
    unsigned long long* bad = (unsigned long long*)(base + 2 * WG_NOALLOC_WORDS);
    return run_checks_in(ctx, ctx->app, ctx->prop[0], ctx->roots, ctx->app == WG_APP_CC ? 0u : ctx->nroots, ok, rootbm,
                         bad, violations);
}

extern "C" int wg_validate_values(wg_ctx* ctx, int app, const uint64_t* roots, uint32_t nroots, const void* values,
                                  uint64_t* violations)
{
    if (!ctx || !values || !violations || (nroots && !roots))
        return WG_E_INVAL;
    if (app != WG_APP_BFS && app != WG_APP_CC && app != WG_APP_SSSP)
        return WG_E_INVAL;
    if (!ctx->off || (app == WG_APP_SSSP && ctx->wkind == WG_WEIGHT_NONE))
        return WG_E_STATE;
    std::vector<uint32_t> dr(nroots);
    int rc;
    if ((rc = wg_to_device_ids(ctx, roots, (uint32_t)nroots, dr.data())))
        return rc;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const size_t esz = (app == WG_APP_BFS || (app == WG_APP_SSSP && ctx->wkind == WG_WEIGHT_F64)) ? 8 : 4;
    const uint64_t V = ctx->V;
    void *raw = nullptr, *dev = nullptr;
    uint32_t* droots = nullptr;
    if (hipMalloc(&raw, V * esz) != hipSuccess || hipMalloc(&dev, V * esz) != hipSuccess ||
        hipMalloc((void**)&droots, (nroots ? nroots : 1) * sizeof(uint32_t)) != hipSuccess) {
        (void)hipFree(raw);
        (void)hipFree(dev);
        return WG_E_NOMEM;
    }
    rc = WG_OK;
    if (hipMemcpyAsync(raw, values, V * esz, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        (nroots && hipMemcpyAsync(droots, dr.data(), nroots * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) !=
                       hipSuccess)) {
        ctx->last_hip_error = (int)hipGetLastError();
        rc = WG_E_HIP;
    } else {
        const void* vals = raw;
        if (ctx->perm) { // the caller's numbering -> device numbering
            const uint32_t g = wg_grid(ctx, V, 1024, 8);
            if (esz == 8)
                hipLaunchKernelGGL(k_chk_to_device<uint64_t>, dim3(g), dim3(256), 0, ctx->stream, (const uint64_t*)raw,
                                   ctx->perm, V, (uint64_t*)dev);
            else
                hipLaunchKernelGGL(k_chk_to_device<uint32_t>, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t*)raw,
                                   ctx->perm, V, (uint32_t*)dev);
            vals = dev;
            rc = wg_check_launch(ctx);
        }
        if (rc == WG_OK)
            rc = run_checks(ctx, app, vals, droots, app == WG_APP_CC ? 0u : nroots, violations);
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(raw);
    (void)hipFree(dev);
    (void)hipFree(droots);
    return rc;
}
