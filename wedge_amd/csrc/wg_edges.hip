// wg_edges.hip -- graph construction on the device from an edge list in any
// order (no counterpart in the reference, whose ingress reads edges already
// grouped by destination, source/graphdata.c:247-585,946-1133).
//
//   1. key = dst << b | src (b = bits of V - 1), value = the edge's input
//      position; one hipcub radix sort over the 2b key bits groups the edges
//      by destination with ascending sources (a stable sort keeps the input
//      order of equal pairs);
//   2. keep flags: self-loops and repeats of the previous key dropped on
//      request; an exclusive scan gives each kept edge its position;
//   3. scatter: source, destination and (through the input position) weight
//      of every kept edge;
//   4. in-degrees by a wave-aggregated histogram over the sorted destinations
//      (one atomic per run of equal destinations per wavefront), offsets by an
//      exclusive scan.
// The CSC then goes through wg_load_csc (validation, out-degrees, optional
// relabelling, tiles, CSRI), so both entry points produce the same layout.
#include "wg_internal.h"

#include <hipcub/hipcub.hpp>

#include <vector>

__global__ void k_edge_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E, uint64_t V,
                            int b, uint64_t* __restrict__ key, uint32_t* __restrict__ idx,
                            unsigned int* __restrict__ bad)
{
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = src[e], d = dst[e];
        if (s >= V || d >= V)
            atomicOr(bad, 1u);
        key[e] = ((uint64_t)d << b) | (uint64_t)s;
        idx[e] = (uint32_t)e;
    }
}

__global__ void k_edge_flags(const uint64_t* __restrict__ key, uint64_t E, int b, int drop_self, int dedup,
                             uint32_t* __restrict__ keep)
{
    const uint64_t mask = (1ull << b) - 1ull;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = key[e];
        bool ok = true;
        if (drop_self && (k >> b) == (k & mask))
            ok = false;
        if (dedup && e > 0 && key[e - 1] == k)
            ok = false;
        keep[e] = ok ? 1u : 0u;
    }
}

template <typename W>
__global__ void k_edge_scatter(const uint64_t* __restrict__ key, const uint32_t* __restrict__ idx,
                               const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos, uint64_t E, int b,
                               const W* __restrict__ w_in, uint32_t* __restrict__ src_out, uint32_t* __restrict__ dst_out,
                               W* __restrict__ w_out)
{
    const uint64_t mask = (1ull << b) - 1ull;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
        if (!keep[e])
            continue;
        const uint64_t k = key[e];
        const uint32_t p = pos[e];
        src_out[p] = (uint32_t)(k & mask);
        dst_out[p] = (uint32_t)(k >> b);
        if (w_in)
            w_out[p] = w_in[idx[e]];
    }
}

// cnt[d] += length of each run of equal destinations inside a wavefront.
__global__ void k_edge_indeg(const uint32_t* __restrict__ dst, uint64_t n, uint32_t* __restrict__ cnt)
{
    const int lane = wg_lane();
    const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / WG_WAVE);
    for (uint64_t w0 = ((uint64_t)blockIdx.x * (blockDim.x / WG_WAVE) + threadIdx.x / WG_WAVE) * WG_WAVE; w0 < n;
         w0 += nwaves * WG_WAVE) {
        const uint64_t i = w0 + (uint64_t)lane;
        const bool live = i < n;
        const uint32_t d = live ? dst[i] : 0xffffffffu;
        const uint32_t dprev = __shfl_up(d, 1);
        const bool start = live && (lane == 0 || dprev != d);
        const unsigned long long starts = wg_ballot(start);
        const unsigned long long lives = wg_ballot(live);
        if (start) {
            const unsigned long long later = (lane == WG_WAVE - 1) ? 0ull : (starts >> (lane + 1)) << (lane + 1);
            const int end = later ? __ffsll((long long)later) - 1 : WG_WAVE - __clzll((long long)lives);
            atomicAdd(&cnt[d], (uint32_t)(end - lane));
        }
    }
}

namespace {

struct dev_buf {
    void* p = nullptr;
    ~dev_buf()
    {
        if (p)
            (void)hipFree(p);
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
    }
};

} // namespace

#define WG_TRY(call)                                 \
    do {                                             \
        hipError_t _e = (call);                      \
        if (_e != hipSuccess) {                      \
            ctx->last_hip_error = (int)_e;           \
            return _e == hipErrorOutOfMemory ? WG_E_NOMEM : WG_E_HIP; \
        }                                            \
    } while (0)

static int build_csc(wg_ctx* ctx, uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     const void* weights, size_t wsz, int flags, std::vector<uint32_t>& hoff,
                     std::vector<uint32_t>& hsrc, std::vector<char>& hw)
{
    int b = 1;
    while ((1ull << b) < V)
        ++b;
    hipStream_t st = ctx->stream;
    dev_buf dsrc, ddst, key0, key1, idx0, idx1, flag, pos, tmp, bad;
    WG_TRY(hipMalloc(&dsrc.p, E * 4));
    WG_TRY(hipMalloc(&ddst.p, E * 4));
    WG_TRY(hipMalloc(&key0.p, E * 8));
    WG_TRY(hipMalloc(&idx0.p, E * 4));
    WG_TRY(hipMalloc(&bad.p, sizeof(unsigned int)));
    WG_TRY(hipMemcpyAsync(dsrc.p, src, E * 4, hipMemcpyHostToDevice, st));
    WG_TRY(hipMemcpyAsync(ddst.p, dst, E * 4, hipMemcpyHostToDevice, st));
    WG_TRY(hipMemsetAsync(bad.p, 0, sizeof(unsigned int), st));
    hipLaunchKernelGGL(k_edge_keys, dim3(wg_grid(ctx, E, 1024, 8)), dim3(256), 0, st, (const uint32_t*)dsrc.p,
                       (const uint32_t*)ddst.p, E, V, b, (uint64_t*)key0.p, (uint32_t*)idx0.p, (unsigned int*)bad.p);
    WG_TRY(hipGetLastError());
    unsigned int hbad = 0;
    WG_TRY(hipMemcpyAsync(&hbad, bad.p, sizeof(hbad), hipMemcpyDeviceToHost, st));
    WG_TRY(hipStreamSynchronize(st));
    if (hbad)
        return WG_E_RANGE;
    dsrc.release();
    ddst.release();

    // 1. sort by (dst, src)
    WG_TRY(hipMalloc(&key1.p, E * 8));
    WG_TRY(hipMalloc(&idx1.p, E * 4));
    size_t tbytes = 0;
    WG_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, (const uint64_t*)key0.p, (uint64_t*)key1.p,
                                              (const uint32_t*)idx0.p, (uint32_t*)idx1.p, E, 0, 2 * b, st));
    WG_TRY(hipMalloc(&tmp.p, tbytes ? tbytes : 1));
    WG_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tbytes, (const uint64_t*)key0.p, (uint64_t*)key1.p,
                                              (const uint32_t*)idx0.p, (uint32_t*)idx1.p, E, 0, 2 * b, st));
    WG_TRY(hipStreamSynchronize(st));
    tmp.release();
    key0.release();
    idx0.release();

    // 2. keep flags and positions
    WG_TRY(hipMalloc(&flag.p, E * 4));
    WG_TRY(hipMalloc(&pos.p, E * 4));
    hipLaunchKernelGGL(k_edge_flags, dim3(wg_grid(ctx, E, 1024, 8)), dim3(256), 0, st, (const uint64_t*)key1.p, E, b,
                       (flags & WG_EDGES_DROP_SELF_LOOPS) ? 1 : 0, (flags & WG_EDGES_DEDUP) ? 1 : 0, (uint32_t*)flag.p);
    WG_TRY(hipGetLastError());
    tbytes = 0;
    WG_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, (const uint32_t*)flag.p, (uint32_t*)pos.p, E, st));
    WG_TRY(hipMalloc(&tmp.p, tbytes ? tbytes : 1));
    WG_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tbytes, (const uint32_t*)flag.p, (uint32_t*)pos.p, E, st));
    uint32_t last[2] = {0, 0};
    WG_TRY(hipMemcpyAsync(&last[0], (uint32_t*)pos.p + (E - 1), 4, hipMemcpyDeviceToHost, st));
    WG_TRY(hipMemcpyAsync(&last[1], (uint32_t*)flag.p + (E - 1), 4, hipMemcpyDeviceToHost, st));
    WG_TRY(hipStreamSynchronize(st));
    tmp.release();
    const uint64_t K = (uint64_t)last[0] + last[1];

    // 3. scatter the kept edges
    dev_buf osrc, odst, ow, win, cnt, off;
    WG_TRY(hipMalloc(&osrc.p, (K ? K : 1) * 4));
    WG_TRY(hipMalloc(&odst.p, (K ? K : 1) * 4));
    if (wsz) {
        WG_TRY(hipMalloc(&win.p, E * wsz));
        WG_TRY(hipMalloc(&ow.p, (K ? K : 1) * wsz));
        WG_TRY(hipMemcpyAsync(win.p, weights, E * wsz, hipMemcpyHostToDevice, st));
    }
    const uint32_t eg = wg_grid(ctx, E, 1024, 8);
    if (wsz == 8)
        hipLaunchKernelGGL(k_edge_scatter<uint64_t>, dim3(eg), dim3(256), 0, st, (const uint64_t*)key1.p,
                           (const uint32_t*)idx1.p, (const uint32_t*)flag.p, (const uint32_t*)pos.p, E, b,
                           (const uint64_t*)win.p, (uint32_t*)osrc.p, (uint32_t*)odst.p, (uint64_t*)ow.p);
    else
        hipLaunchKernelGGL(k_edge_scatter<uint32_t>, dim3(eg), dim3(256), 0, st, (const uint64_t*)key1.p,
                           (const uint32_t*)idx1.p, (const uint32_t*)flag.p, (const uint32_t*)pos.p, E, b,
                           (const uint32_t*)(wsz ? win.p : nullptr), (uint32_t*)osrc.p, (uint32_t*)odst.p,
                           (uint32_t*)ow.p);
    WG_TRY(hipGetLastError());
    WG_TRY(hipStreamSynchronize(st));
    key1.release();
    idx1.release();
    flag.release();
    pos.release();
    win.release();

    // 4. offsets
    WG_TRY(hipMalloc(&cnt.p, (V + 1) * 4));
    WG_TRY(hipMalloc(&off.p, (V + 1) * 4));
    WG_TRY(hipMemsetAsync(cnt.p, 0, (V + 1) * 4, st));
    if (K)
        hipLaunchKernelGGL(k_edge_indeg, dim3(wg_grid(ctx, K, 1024, 8)), dim3(256), 0, st, (const uint32_t*)odst.p, K,
                           (uint32_t*)cnt.p);
    WG_TRY(hipGetLastError());
    tbytes = 0;
    WG_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, (const uint32_t*)cnt.p, (uint32_t*)off.p, V + 1, st));
    WG_TRY(hipMalloc(&tmp.p, tbytes ? tbytes : 1));
    WG_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tbytes, (const uint32_t*)cnt.p, (uint32_t*)off.p, V + 1, st));

    hoff.resize(V + 1);
    hsrc.resize(K);
    hw.resize(K * wsz);
    WG_TRY(hipMemcpyAsync(hoff.data(), off.p, (V + 1) * 4, hipMemcpyDeviceToHost, st));
    if (K)
        WG_TRY(hipMemcpyAsync(hsrc.data(), osrc.p, K * 4, hipMemcpyDeviceToHost, st));
    if (K && wsz)
        WG_TRY(hipMemcpyAsync(hw.data(), ow.p, K * wsz, hipMemcpyDeviceToHost, st));
    WG_TRY(hipStreamSynchronize(st));
    return WG_OK;
}

extern "C" int wg_load_edges(wg_ctx* ctx, uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                             const void* weights, int weight_kind, int flags, uint64_t* kept_edges)
{
    if (!ctx || V == 0 || (E && (!src || !dst)))
        return WG_E_INVAL;
    if (V >= (uint64_t)WG_HEAVY_BIT || E >= 0xffffffffull)
        return WG_E_RANGE;
    if (weight_kind != WG_WEIGHT_NONE && weight_kind != WG_WEIGHT_F64 && weight_kind != WG_WEIGHT_U32)
        return WG_E_INVAL;
    if (weight_kind != WG_WEIGHT_NONE && E && !weights)
        return WG_E_INVAL;
    if (flags & ~(WG_EDGES_DROP_SELF_LOOPS | WG_EDGES_DEDUP))
        return WG_E_INVAL;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    wg_drop_graph(ctx); // the sort needs the memory
    const size_t wsz = weight_kind == WG_WEIGHT_F64 ? 8 : weight_kind == WG_WEIGHT_U32 ? 4 : 0;
    std::vector<uint32_t> hoff, hsrc;
    std::vector<char> hw;
    int rc;
    if (E) {
        if ((rc = build_csc(ctx, V, E, src, dst, weights, wsz, flags, hoff, hsrc, hw)))
            return rc;
    } else {
        hoff.assign(V + 1, 0u);
    }
    if (kept_edges)
        *kept_edges = hsrc.size();
    return wg_load_csc(ctx, V, hsrc.size(), hoff.data(), hsrc.data(), wsz ? (const void*)hw.data() : nullptr,
                       weight_kind);
}
