// wg_pb.hip -- the binned PageRank edge pull: two streaming passes per
// iteration over a static reordering of the edges, for graphs whose
// contributions are far larger than the caches (rmat-s27: 1.07 GB of FP64
// contributions against 4 MB of L2 per XCD and 256 MB of Infinity Cache).
//
// Reference: perform_edge_pull_phase_pr (source/phase_edge_pull_pr.c:28-116),
// the merge of partial sums of one destination (source/phases.c:69-114) and
// perform_vertex_phase_pr (source/phase_vertex_pr.c:26-77, formula and
// operation order).  The sums are the same as the tile kernel's; only the
// order of the FP64 additions differs.
//
// The tile kernel (wg_pr_tile.h) gathers one contribution per edge from
// anywhere in the array, and a gather that misses the caches costs a whole
// line (86 bytes per edge below L2 at s27).  Here every global access is a
// stream:
//   windows  sources cut into windows of Ws = 2^ws_log vertices, whose
//            contributions fill one workgroup's LDS (128 or 64 KiB)
//   bins     destinations cut into bins of Wd = 2^wd_log vertices, whose FP64
//            accumulators fill one workgroup's LDS
//   push     per window: a workgroup copies the window's contributions to LDS
//            and streams the window's edges, ordered by (destination, source),
//            2 bytes each; each run of edges with one destination (a "pair":
//            window x destination) becomes one FP64 partial sum, stored in
//            window order (each wavefront writes one contiguous range)
//   pull     per bin: a workgroup gathers the partials of every window (one
//            run of consecutive positions per window) into LDS accumulators,
//            then runs the vertex phase on the bin's vertices
// Bytes per iteration: 2 per edge, 18 per pair (the FP64 partial written and
// read back, and its 2-byte destination; 10 with the FP32-stored
// contributions, whose partials are stored in FP32), the windows' LDS copies,
// and the vertex phase's 16 (FP64) or 12 (FP32) per vertex plus offsets-free
// out-degrees.
// After the degree relabelling most edges of a window share their destination
// with other edges of that window (rmat-s24: 0.27 pairs per edge with 16K-vertex
// windows, rmat-s27: 0.40), so an iteration moves about 7-11 bytes per edge.
//
// FP64 order: a partial adds its edges in source order within each lane, then
// the lanes' open runs by a segmented scan (a fixed order); the pull adds
// partials with LDS atomics (ds_add_f64) in no fixed order, so results agree
// with the tile kernel to rounding, not bit for bit, and may differ in the last
// bits from run to run.
//
// Layout (built once per graph and window size by pb_build, device side):
//   code[q]        the edges by (window, destination, source), each window's run
//                  padded to a multiple of WG_PB_TILE entries: bits 13..0 =
//                  source - window start, bit 15 = last edge of its pair (also
//                  forced on a tile's last entry, so no pair crosses tiles)
//   tile_pbase[t]  pairs before tile t (WG_PB_TILE entries per tile): pair p's
//                  partial is part[p]
//   task[k]        push work: {window, first tile, end tile, end of the real
//                  entries (code index)}, at most WG_PB_TASK_TILES tiles
//   blocks         runs of pairs of one (window, bin), numbered in bin order
//                  (window order within a bin); bin order position o of block
//                  kb holds pair o - boff[kb] (32-bit wrap-around)
//   pdst[o]        destination - bin start of the pair at bin-order position o,
//                  bit 15 set where a block starts
//   chunk_block[c] the block holding position 64c (c <= positions / 64: one
//                  past the end holds the block count), followed by
//                  WG_PB_CHUNK_OFFS block offsets per chunk: boff of the
//                  chunk's first blocks, so that the pull of a chunk with at
//                  most that many blocks needs no dependent load for them
//   slice[j]       pull work: {bin, first position, end position, accumulator
//                  slot or WG_PB_NOSLOT when the slice is its bin's only one},
//                  at most WG_PB_SLICE pairs, a contiguous range of positions
//   bin_slot[b]    {first slot, slices} of a bin cut into several slices: each
//                  slice writes its accumulators to its slot, and the last one
//                  to finish (a per-bin ticket) adds the slots in slice order
//   wq[q]          on a weighted graph, code entry q's edge weight (FP64 or
//                  u32; 0 for padding), for the binned SSSP pull
#include "wg_frontier.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#define WG_PB_THREADS 1024                      // 16 wavefronts, one workgroup per CU
#define WG_PB_WAVES (WG_PB_THREADS / WG_WAVE)
#define WG_PB_TILE 512u                         // code entries per wavefront tile
#define WG_PB_TASK_TILES 1024u                  // tiles per push workgroup (512K edges)
#define WG_PB_STAGE 256u                        // partials a push wavefront stages in LDS
#define WG_PB_SLICE (256u * 1024u)              // pairs per pull workgroup (default)
#define WG_PB_NOSLOT 0xffffffffu
#define WG_PB_END 0x8000u                       // code: last edge of a pair
#define WG_PB_SRC 0x3fffu                       // code: source - window start
#define WG_PB_MAX_WS_LOG 14
#define WG_PB_BFLAG 0x8000u                     // pdst: a block starts here
#define WG_PB_DMASK 0x7fffu                     // pdst: destination - bin start
#define WG_PB_CHUNK_OFFS 8u                     // block offsets kept with each 64-position chunk

struct wg_pb {
    int ws_log, wd_log;
    uint32_t S, B;
    uint64_t nq, npairs;
    uint32_t ntiles, ntasks, nblocks, nslices, nslots;
    uint16_t* code;
    uint32_t* tile_pbase;
    uint4* task;
    uint32_t* boff;
    uint32_t* chunk_block;
    uint16_t* pdst;
    double* part;
    uint4* slice;
    uint2* bin_slot;
    double* slot_acc;
    unsigned int* bin_ticket;
    double* bin_part;
    void* wq;
};

namespace {

template <typename P>
void pb_free_dev(P*& p)
{
    if (p)
        (void)hipFree((void*)p);
    p = nullptr;
}

void pb_release(wg_pb* pb)
{
    pb_free_dev(pb->code);
    pb_free_dev(pb->tile_pbase);
    pb_free_dev(pb->task);
    pb_free_dev(pb->boff);
    pb_free_dev(pb->chunk_block);
    pb_free_dev(pb->pdst);
    pb_free_dev(pb->part);
    pb_free_dev(pb->slice);
    pb_free_dev(pb->bin_slot);
    pb_free_dev(pb->slot_acc);
    pb_free_dev(pb->bin_ticket);
    pb_free_dev(pb->bin_part);
    pb_free_dev(pb->wq);
}

struct pb_tmp {
    void* p = nullptr;
    ~pb_tmp()
    {
        if (p)
            (void)hipFree(p);
    }
};

} // namespace

void wg_free_pb(wg_ctx* ctx)
{
    if (!ctx->pb)
        return;
    pb_release(ctx->pb);
    delete ctx->pb;
    ctx->pb = nullptr;
}

// ---- layout build ---------------------------------------------------------------

__global__ void k_pb_keys(const uint32_t* __restrict__ src, uint64_t E, int ws_log, uint16_t* __restrict__ key,
                          uint32_t* __restrict__ idx)
{
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
        key[e] = (uint16_t)(src[e] >> ws_log);
        idx[e] = (uint32_t)e;
    }
}

// ws[s] = first position of window s in the sorted keys (ws[S] = E)
__global__ void k_pb_window_starts(const uint16_t* __restrict__ key, uint64_t E, uint32_t S, uint64_t* __restrict__ ws)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > S)
        return;
    uint64_t lo = 0, hi = E;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((uint32_t)key[mid] < s)
            lo = mid + 1;
        else
            hi = mid;
    }
    ws[s] = lo;
}

// Edge (sorted position) of code entry q, or UINT64_MAX for padding.
__device__ __forceinline__ uint64_t pb_entry(uint64_t q, const uint32_t* __restrict__ tile_win,
                                             const uint64_t* __restrict__ ws, const uint64_t* __restrict__ pq,
                                             uint32_t* s_out, uint64_t* r_out, uint64_t* cnt_out)
{
    const uint32_t s = tile_win[q / WG_PB_TILE];
    const uint64_t r = q - pq[s], cnt = ws[s + 1] - ws[s];
    *s_out = s;
    *r_out = r;
    *cnt_out = cnt;
    return r < cnt ? ws[s] + r : ~0ull;
}

__global__ void k_pb_codes(uint64_t nq, const uint32_t* __restrict__ tile_win, const uint64_t* __restrict__ ws,
                           const uint64_t* __restrict__ pq, const uint32_t* __restrict__ perm,
                           const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint32_t wmask,
                           uint16_t* __restrict__ code)
{
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nq; q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t s;
        uint64_t r, cnt;
        const uint64_t p = pb_entry(q, tile_win, ws, pq, &s, &r, &cnt);
        uint16_t c = 0;
        if (p != ~0ull) {
            const uint32_t e = perm[p];
            const bool end = (r + 1 == cnt) || ((q % WG_PB_TILE) == WG_PB_TILE - 1) || dst[perm[p + 1]] != dst[e];
            c = (uint16_t)((src[e] & wmask) | (end ? WG_PB_END : 0u));
        }
        code[q] = c;
    }
}

// wq[q] = the weight of code entry q's edge (0 for padding).
template <typename W>
__global__ void k_pb_weights(uint64_t nq, const uint32_t* __restrict__ tile_win, const uint64_t* __restrict__ ws,
                             const uint64_t* __restrict__ pq, const uint32_t* __restrict__ perm,
                             const W* __restrict__ w, W* __restrict__ wq)
{
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nq; q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t s;
        uint64_t r, cnt;
        const uint64_t p = pb_entry(q, tile_win, ws, pq, &s, &r, &cnt);
        wq[q] = p != ~0ull ? w[perm[p]] : W(0);
    }
}

struct pb_end_flag {
    __host__ __device__ uint32_t operator()(uint16_t c) const { return (c & WG_PB_END) ? 1u : 0u; }
};

__global__ void k_pb_pairs(uint64_t nq, const uint16_t* __restrict__ code, const uint32_t* __restrict__ pidx,
                           const uint32_t* __restrict__ tile_win, const uint64_t* __restrict__ ws,
                           const uint64_t* __restrict__ pq, const uint32_t* __restrict__ perm,
                           const uint32_t* __restrict__ dst, int wd_log, uint16_t* __restrict__ pdst,
                           uint32_t* __restrict__ pkey)
{
    const uint32_t dmask = (1u << wd_log) - 1u;
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nq; q += (uint64_t)gridDim.x * blockDim.x) {
        if (!(code[q] & WG_PB_END))
            continue;
        uint32_t s;
        uint64_t r, cnt;
        const uint64_t p = pb_entry(q, tile_win, ws, pq, &s, &r, &cnt);
        const uint32_t d = dst[perm[p]];
        const uint32_t i = pidx[q];
        pdst[i] = (uint16_t)(d & dmask);
        pkey[i] = (s << 16) | (d >> wd_log);
    }
}

__global__ void k_pb_tile_base(const uint32_t* __restrict__ pidx, uint32_t ntiles, uint32_t npairs,
                               uint32_t* __restrict__ tile_pbase)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntiles)
        tile_pbase[t] = pidx[(uint64_t)t * WG_PB_TILE];
    else if (t == ntiles)
        tile_pbase[t] = npairs;
}

__global__ void k_pb_block_flags(const uint32_t* __restrict__ pkey, uint64_t n, uint8_t* __restrict__ flag)
{
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x)
        flag[p] = (p == 0 || pkey[p] != pkey[p - 1]) ? 1u : 0u;
}

__global__ void k_pb_block_keys(const uint32_t* __restrict__ start, uint32_t nb, const uint32_t* __restrict__ pkey,
                                uint32_t* __restrict__ key)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb)
        key[i] = pkey[start[i]];
}

// Number of entries of the ascending array a[0..n) below x.
__device__ __forceinline__ uint32_t pb_lower_bound(const uint32_t* __restrict__ a, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Pair destinations into bin order: pair p of window-order block i goes to
// position p + woff[i]; the first pair of each block carries WG_PB_BFLAG.
__global__ void k_pb_place(const uint16_t* __restrict__ pdst_w, uint32_t npairs, const uint32_t* __restrict__ bstart,
                           uint32_t nb, const uint32_t* __restrict__ woff, uint16_t* __restrict__ pdst_b)
{
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += gridDim.x * blockDim.x) {
        const uint32_t i = pb_lower_bound(bstart, nb, p + 1u) - 1u; // the last block starting at or before p
        pdst_b[p + woff[i]] = (uint16_t)(pdst_w[p] | (bstart[i] == p ? WG_PB_BFLAG : 0u));
    }
}

#define PB_TRY(call)                                                   \
    do {                                                               \
        hipError_t _e = (call);                                        \
        if (_e != hipSuccess) {                                        \
            ctx->last_hip_error = (int)_e;                             \
            return _e == hipErrorOutOfMemory ? WG_E_NOMEM : WG_E_HIP;  \
        }                                                              \
    } while (0)

template <typename T>
static int pb_upload(wg_ctx* ctx, T** d, const std::vector<T>& h)
{
    PB_TRY(hipMalloc((void**)d, (h.empty() ? 1 : h.size()) * sizeof(T)));
    if (!h.empty())
        PB_TRY(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return WG_OK;
}

static int pb_build(wg_ctx* ctx, wg_pb* pb, int ws_log, int wd_log)
{
    const uint64_t V = ctx->V, E = ctx->E;
    const uint32_t S = (uint32_t)((V + (1ull << ws_log) - 1) >> ws_log);
    const uint32_t B = (uint32_t)((V + (1ull << wd_log) - 1) >> wd_log);
    if (S > 65536u || B > 65536u || ws_log > WG_PB_MAX_WS_LOG || wd_log > 15)
        return WG_E_RANGE;
    hipStream_t st = ctx->stream;
    pb->ws_log = ws_log;
    pb->wd_log = wd_log;
    pb->S = S;
    pb->B = B;
    int bits = 1;
    while ((1u << bits) < S)
        ++bits;

    // 1. the edges in (window, destination, source) order: a stable sort of the
    //    CSC order, which is (destination, source), by window
    pb_tmp key0, key1, idx0, perm, tmp, wsd;
    std::vector<uint64_t> ws(S + 1, 0);
    const uint32_t g = wg_grid(ctx, E ? E : 1, 1024, 8);
    if (E) {
        PB_TRY(hipMalloc(&key0.p, E * 2));
        PB_TRY(hipMalloc(&key1.p, E * 2));
        PB_TRY(hipMalloc(&idx0.p, E * 4));
        PB_TRY(hipMalloc(&perm.p, E * 4));
        hipLaunchKernelGGL(k_pb_keys, dim3(g), dim3(256), 0, st, ctx->src, E, ws_log, (uint16_t*)key0.p,
                           (uint32_t*)idx0.p);
        size_t tb = 0;
        PB_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const uint16_t*)key0.p, (uint16_t*)key1.p,
                                                  (const uint32_t*)idx0.p, (uint32_t*)perm.p, (uint64_t)E, 0, bits, st));
        PB_TRY(hipMalloc(&tmp.p, tb ? tb : 1));
        PB_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, (const uint16_t*)key0.p, (uint16_t*)key1.p,
                                                  (const uint32_t*)idx0.p, (uint32_t*)perm.p, (uint64_t)E, 0, bits, st));
        (void)hipFree(tmp.p);
        tmp.p = nullptr;
        (void)hipFree(idx0.p);
        idx0.p = nullptr;
        (void)hipFree(key0.p);
        key0.p = nullptr;
        PB_TRY(hipMalloc(&wsd.p, (S + 1) * sizeof(uint64_t)));
        hipLaunchKernelGGL(k_pb_window_starts, dim3((S + 256) / 256), dim3(256), 0, st, (const uint16_t*)key1.p, E, S,
                           (uint64_t*)wsd.p);
        PB_TRY(hipGetLastError());
        PB_TRY(hipMemcpyAsync(ws.data(), wsd.p, (S + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        PB_TRY(hipStreamSynchronize(st));
        (void)hipFree(key1.p);
        key1.p = nullptr;
    }

    // 2. padded code ranges, tiles and push tasks (host)
    std::vector<uint64_t> pq(S + 1, 0);
    std::vector<uint32_t> tile_win;
    std::vector<uint4> task;
    for (uint32_t s = 0; s < S; ++s) {
        const uint64_t cnt = ws[s + 1] - ws[s];
        const uint64_t tiles = (cnt + WG_PB_TILE - 1) / WG_PB_TILE;
        const uint32_t t0 = (uint32_t)(pq[s] / WG_PB_TILE);
        for (uint64_t k = 0; k < tiles; k += WG_PB_TASK_TILES) {
            const uint64_t k1 = std::min<uint64_t>(tiles, k + WG_PB_TASK_TILES);
            task.push_back(make_uint4(s, t0 + (uint32_t)k, t0 + (uint32_t)k1, 0u));
        }
        for (uint64_t k = 0; k < tiles; ++k)
            tile_win.push_back(s);
        pq[s + 1] = pq[s] + tiles * WG_PB_TILE;
    }
    const uint64_t nq = pq[S];
    if (nq >= 0xffffffffull) // code indices are 32-bit
        return WG_E_RANGE;
    for (auto& tk : task) // end of the window's real entries, as a code index
        tk.w = (uint32_t)(pq[tk.x] + (ws[tk.x + 1] - ws[tk.x]));
    pb->nq = nq;
    pb->ntiles = (uint32_t)(nq / WG_PB_TILE);
    pb->ntasks = (uint32_t)task.size();

    // 3. codes, pair numbers (window order), pair destinations and (window, bin) keys
    pb_tmp twin, pqd, pidx, pkey, pdst_w, bflag, bst, nsel, bk;
    uint32_t npairs = 0;
    int rc;
    if (nq) {
        if ((rc = pb_upload(ctx, (uint32_t**)&twin.p, tile_win)) || (rc = pb_upload(ctx, (uint64_t**)&pqd.p, pq)))
            return rc;
        PB_TRY(hipMalloc((void**)&pb->code, nq * sizeof(uint16_t)));
        const uint32_t gq = wg_grid(ctx, nq, 1024, 8);
        hipLaunchKernelGGL(k_pb_codes, dim3(gq), dim3(256), 0, st, nq, (const uint32_t*)twin.p, (const uint64_t*)wsd.p,
                           (const uint64_t*)pqd.p, (const uint32_t*)perm.p, ctx->src, ctx->dst,
                           (1u << ws_log) - 1u, pb->code);
        PB_TRY(hipGetLastError());
        if (ctx->w && ctx->wkind != WG_WEIGHT_NONE) { // weights in code order, for the binned SSSP pull
            if (ctx->wkind == WG_WEIGHT_F64) {
                PB_TRY(hipMalloc(&pb->wq, nq * sizeof(double)));
                hipLaunchKernelGGL(k_pb_weights<double>, dim3(gq), dim3(256), 0, st, nq, (const uint32_t*)twin.p,
                                   (const uint64_t*)wsd.p, (const uint64_t*)pqd.p, (const uint32_t*)perm.p,
                                   (const double*)ctx->w, (double*)pb->wq);
            } else {
                PB_TRY(hipMalloc(&pb->wq, nq * sizeof(uint32_t)));
                hipLaunchKernelGGL(k_pb_weights<uint32_t>, dim3(gq), dim3(256), 0, st, nq, (const uint32_t*)twin.p,
                                   (const uint64_t*)wsd.p, (const uint64_t*)pqd.p, (const uint32_t*)perm.p,
                                   (const uint32_t*)ctx->w, (uint32_t*)pb->wq);
            }
            PB_TRY(hipGetLastError());
        }
        PB_TRY(hipMalloc(&pidx.p, nq * sizeof(uint32_t)));
        hipcub::TransformInputIterator<uint32_t, pb_end_flag, const uint16_t*> flags(pb->code, pb_end_flag());
        size_t tb = 0;
        PB_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flags, (uint32_t*)pidx.p, (uint64_t)nq, st));
        PB_TRY(hipMalloc(&tmp.p, tb ? tb : 1));
        PB_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flags, (uint32_t*)pidx.p, (uint64_t)nq, st));
        uint32_t last_idx = 0;
        uint16_t last_code = 0;
        PB_TRY(hipMemcpyAsync(&last_idx, (uint32_t*)pidx.p + nq - 1, 4, hipMemcpyDeviceToHost, st));
        PB_TRY(hipMemcpyAsync(&last_code, pb->code + nq - 1, 2, hipMemcpyDeviceToHost, st));
        PB_TRY(hipStreamSynchronize(st));
        (void)hipFree(tmp.p);
        tmp.p = nullptr;
        npairs = last_idx + ((last_code & WG_PB_END) ? 1u : 0u);
    }
    pb->npairs = npairs;
    PB_TRY(hipMalloc((void**)&pb->tile_pbase, ((uint64_t)pb->ntiles + 1) * sizeof(uint32_t)));
    PB_TRY(hipMalloc((void**)&pb->pdst, (npairs ? npairs : 1) * sizeof(uint16_t)));
    PB_TRY(hipMalloc((void**)&pb->part, (npairs ? npairs : 1) * sizeof(double)));
    std::vector<uint32_t> bstart, bkey;
    if (nq) {
        hipLaunchKernelGGL(k_pb_tile_base, dim3(pb->ntiles / 256 + 1), dim3(256), 0, st, (const uint32_t*)pidx.p,
                           pb->ntiles, npairs, pb->tile_pbase);
        PB_TRY(hipMalloc(&pkey.p, (npairs ? npairs : 1) * sizeof(uint32_t)));
        PB_TRY(hipMalloc(&pdst_w.p, (npairs ? npairs : 1) * sizeof(uint16_t)));
        const uint32_t gq = wg_grid(ctx, nq, 1024, 8);
        hipLaunchKernelGGL(k_pb_pairs, dim3(gq), dim3(256), 0, st, nq, pb->code, (const uint32_t*)pidx.p,
                           (const uint32_t*)twin.p, (const uint64_t*)wsd.p, (const uint64_t*)pqd.p,
                           (const uint32_t*)perm.p, ctx->dst, wd_log, (uint16_t*)pdst_w.p, (uint32_t*)pkey.p);
        PB_TRY(hipGetLastError());
        (void)hipFree(perm.p);
        perm.p = nullptr;

        // 4. blocks: runs of equal (window, bin) among the pairs
        PB_TRY(hipMalloc(&bflag.p, npairs));
        PB_TRY(hipMalloc(&bst.p, (uint64_t)npairs * 4));
        PB_TRY(hipMalloc(&nsel.p, sizeof(uint64_t)));
        hipLaunchKernelGGL(k_pb_block_flags, dim3(wg_grid(ctx, npairs, 1024, 8)), dim3(256), 0, st,
                           (const uint32_t*)pkey.p, (uint64_t)npairs, (uint8_t*)bflag.p);
        PB_TRY(hipGetLastError());
        (void)hipFree(pidx.p);
        pidx.p = nullptr;
        hipcub::CountingInputIterator<uint32_t> pos(0u);
        size_t tb = 0;
        PB_TRY(hipcub::DeviceSelect::Flagged(nullptr, tb, pos, (const uint8_t*)bflag.p, (uint32_t*)bst.p,
                                             (uint64_t*)nsel.p, (int64_t)npairs, st));
        PB_TRY(hipMalloc(&tmp.p, tb ? tb : 1));
        PB_TRY(hipcub::DeviceSelect::Flagged(tmp.p, tb, pos, (const uint8_t*)bflag.p, (uint32_t*)bst.p,
                                             (uint64_t*)nsel.p, (int64_t)npairs, st));
        uint64_t nb = 0;
        PB_TRY(hipMemcpyAsync(&nb, nsel.p, sizeof(nb), hipMemcpyDeviceToHost, st));
        PB_TRY(hipStreamSynchronize(st));
        (void)hipFree(tmp.p);
        tmp.p = nullptr;
        PB_TRY(hipMalloc(&bk.p, (nb ? nb : 1) * 4));
        hipLaunchKernelGGL(k_pb_block_keys, dim3((uint32_t)(nb / 256 + 1)), dim3(256), 0, st, (const uint32_t*)bst.p,
                           (uint32_t)nb, (const uint32_t*)pkey.p, (uint32_t*)bk.p);
        PB_TRY(hipGetLastError());
        bstart.resize(nb);
        bkey.resize(nb);
        if (nb) {
            PB_TRY(hipMemcpyAsync(bstart.data(), bst.p, nb * 4, hipMemcpyDeviceToHost, st));
            PB_TRY(hipMemcpyAsync(bkey.data(), bk.p, nb * 4, hipMemcpyDeviceToHost, st));
        }
        PB_TRY(hipStreamSynchronize(st));
    }

    // 5. bin order of the blocks (window order kept within a bin), their
    //    offsets, the block of every 64th position, slices and accumulator
    //    slots (host)
    const uint64_t nb = bstart.size();
    pb->nblocks = (uint32_t)nb;
    std::vector<uint64_t> bin_first(B + 1, 0);
    for (uint64_t i = 0; i < nb; ++i)
        bin_first[(bkey[i] & 0xffffu) + 1]++;
    for (uint32_t b = 0; b < B; ++b)
        bin_first[b + 1] += bin_first[b];
    std::vector<uint32_t> by_bin(nb);
    {
        std::vector<uint64_t> fill(bin_first.begin(), bin_first.end() - 1);
        for (uint64_t i = 0; i < nb; ++i) // blocks are in (window, bin) order: stable by bin
            by_bin[fill[bkey[i] & 0xffffu]++] = (uint32_t)i;
    }
    std::vector<uint32_t> woff(nb ? nb : 1, 0u); // by window-order block
    std::vector<uint32_t> boff(nb ? nb : 1, 0u); // by bin-order block
    const uint64_t nchunks = ((uint64_t)npairs + 63u) / 64u;
    std::vector<uint32_t> chunk_block(nchunks + 1 + nchunks * WG_PB_CHUNK_OFFS, (uint32_t)nb);
    std::vector<uint32_t> bin_out(B + 1, 0u); // first position of each bin in part[]
    const uint32_t slice_max = ctx->pb_slice ? ctx->pb_slice : WG_PB_SLICE;
    std::vector<uint4> slice;
    std::vector<uint2> bin_slot(B, make_uint2(0u, 0u));
    uint32_t nslots = 0;
    {
        uint32_t o = 0;
        uint64_t c = 0; // next chunk whose first position has no block yet
        for (uint32_t b = 0; b < B; ++b) {
            bin_out[b] = o;
            for (uint64_t k = bin_first[b]; k < bin_first[b + 1]; ++k) {
                const uint32_t i = by_bin[k];
                const uint32_t p0 = bstart[i];
                const uint32_t p1 = (i + 1 < nb) ? bstart[i + 1] : npairs;
                woff[i] = o - p0; // wraps around: p + woff[i] = o + (p - p0)
                boff[k] = o - p0;
                o += p1 - p0;
                for (; c < nchunks && c * 64u < (uint64_t)o; ++c) // positions [o - (p1 - p0), o) are block k's
                    chunk_block[c] = (uint32_t)k;
            }
            const uint32_t end = o;
            const size_t first_slice = slice.size();
            uint32_t q = bin_out[b];
            do { // every bin has a slice, even without pairs: its vertex phase
                const uint32_t q1 = (end - q > slice_max) ? q + slice_max : end;
                slice.push_back(make_uint4(b, q, q1, WG_PB_NOSLOT));
                q = q1;
            } while (q < end);
            const uint32_t nsl = (uint32_t)(slice.size() - first_slice);
            if (nsl > 1) {
                bin_slot[b] = make_uint2(nslots, nsl);
                for (uint32_t j = 0; j < nsl; ++j)
                    slice[first_slice + j].w = nslots + j;
                nslots += nsl;
            }
        }
        bin_out[B] = o;
    }
    for (uint64_t c = 0; c < nchunks; ++c) { // the chunk's first block offsets, after the block numbers
        const uint64_t k0 = chunk_block[c];
        for (uint32_t j = 0; j < WG_PB_CHUNK_OFFS; ++j)
            chunk_block[nchunks + 1 + c * WG_PB_CHUNK_OFFS + j] = k0 + j < nb ? boff[k0 + j] : 0u;
    }
    pb->nslices = (uint32_t)slice.size();
    pb->nslots = nslots;
    pb_tmp woffd;
    if ((rc = pb_upload(ctx, &pb->task, task)) || (rc = pb_upload(ctx, &pb->slice, slice)) ||
        (rc = pb_upload(ctx, &pb->bin_slot, bin_slot)) || (rc = pb_upload(ctx, &pb->boff, boff)) ||
        (rc = pb_upload(ctx, &pb->chunk_block, chunk_block)) || (rc = pb_upload(ctx, (uint32_t**)&woffd.p, woff)))
        return rc;
    if (npairs) // destinations in bin order
        hipLaunchKernelGGL(k_pb_place, dim3(wg_grid(ctx, npairs, 1024, 8)), dim3(256), 0, st,
                           (const uint16_t*)pdst_w.p, npairs, (const uint32_t*)bst.p, (uint32_t)nb,
                           (const uint32_t*)woffd.p, pb->pdst);
    PB_TRY(hipGetLastError());
    PB_TRY(hipMalloc((void**)&pb->slot_acc, ((uint64_t)(nslots ? nslots : 1) << wd_log) * sizeof(double)));
    PB_TRY(hipMalloc((void**)&pb->bin_ticket, B * sizeof(unsigned int)));
    PB_TRY(hipMemsetAsync(pb->bin_ticket, 0, B * sizeof(unsigned int), st));
    PB_TRY(hipMalloc((void**)&pb->bin_part, B * sizeof(double)));
    PB_TRY(hipStreamSynchronize(st));
    return WG_OK;
}

// Window and bin sizes of a binned shape: 16K-vertex windows (at most what the
// 14-bit source field holds) and 16K-vertex bins, 128 KiB of FP64 accumulators,
// for one 1024-thread workgroup per CU (blocks_per_cu 1); 8K-vertex windows
// and bins, 64 KiB each, for two (blocks_per_cu 2).  FP32 windows of 16K
// vertices take 64 KiB.
static void pb_sizes(int bpc, int* ws_log, int* wd_log)
{
    *ws_log = bpc == 1 ? 14 : 13;
    *wd_log = bpc == 1 ? 14 : 13;
}

static int pb_prepare_sizes(wg_ctx* ctx, int ws_log, int wd_log)
{
    if (ctx->pb && ctx->pb->ws_log == ws_log && ctx->pb->wd_log == wd_log)
        return WG_OK;
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    wg_free_pb(ctx);
    ctx->pb = new wg_pb();
    const int rc = pb_build(ctx, ctx->pb, ws_log, wd_log);
    if (rc)
        wg_free_pb(ctx);
    return rc;
}

int wg_pb_prepare(wg_ctx* ctx)
{
    int ws_log, wd_log;
    pb_sizes(ctx->pr_bpc, &ws_log, &wd_log);
    return pb_prepare_sizes(ctx, ws_log, wd_log);
}

// ---- iteration -------------------------------------------------------------------

// Deterministic sum over the workgroup (xor butterfly in each wavefront, then
// the wavefront sums in wavefront order); every thread receives it.
static __device__ __forceinline__ double block_sum_pb(double v, double* s_red)
{
    v = wg_wave_sum(v);
    __syncthreads();
    if (wg_lane() == 0)
        s_red[threadIdx.x / WG_WAVE] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < WG_PB_WAVES; ++i)
        r += s_red[i];
    return r;
}

// s[i] = f(src[i]) for i < n <= N: every load issued before the first LDS
// store (one wait per window instead of one per step).
template <int N, typename S, typename D, class F>
__device__ __forceinline__ void pb_fill_lds(D* s, const S* __restrict__ src, uint32_t n, F f)
{
    static_assert(N % WG_PB_THREADS == 0, "whole rounds");
    constexpr int R = N / WG_PB_THREADS;
    S x[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const uint32_t i = threadIdx.x + (uint32_t)u * WG_PB_THREADS;
        x[u] = i < n ? src[i] : S(0);
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const uint32_t i = threadIdx.x + (uint32_t)u * WG_PB_THREADS;
        if (i < n)
            s[i] = f(x[u]);
    }
}

// Codes of the lane's WG_PB_IPT entries of a tile (one 16-byte load).
#define WG_PB_IPT 8
__device__ __forceinline__ uint4 pb_load_codes(const uint16_t* __restrict__ code, uint32_t t, int lane)
{
    return *reinterpret_cast<const uint4*>(code + (uint64_t)t * WG_PB_TILE + (uint32_t)lane * WG_PB_IPT);
}

// What the push combines: PageRank sums FP64 contributions, the 64-source BFS
// ORs 64-bit frontier masks.
struct pb_sum {
    typedef double V;
    static constexpr bool ZERO_BITS_0 = true; // the identity's bit pattern is all zeros
    static __device__ __forceinline__ V zero() { return 0.0; }
    static __device__ __forceinline__ V op(V a, V b) { return a + b; }
    static __device__ __forceinline__ unsigned long long bits(V x) { return (unsigned long long)__double_as_longlong(x); }
    static __device__ __forceinline__ V from_bits(unsigned long long x) { return __longlong_as_double((long long)x); }
};
struct pb_or {
    typedef unsigned long long V;
    static constexpr bool ZERO_BITS_0 = true;
    static __device__ __forceinline__ V zero() { return 0ull; }
    static __device__ __forceinline__ V op(V a, V b) { return a | b; }
    static __device__ __forceinline__ unsigned long long bits(V x) { return x; }
    static __device__ __forceinline__ V from_bits(unsigned long long x) { return x; }
};
// Minimum of CC labels or SSSP distances (P::T: u32, or FP64 bit patterns,
// which order as the distances do since they are never negative).
template <class P>
struct pb_min {
    typedef typename P::T V;
    static constexpr bool ZERO_BITS_0 = false;
    static __device__ __forceinline__ V zero() { return P::INF; }
    static __device__ __forceinline__ V op(V a, V b) { return b < a ? b : a; }
    static __device__ __forceinline__ unsigned long long bits(V x) { return (unsigned long long)x; }
    static __device__ __forceinline__ V from_bits(unsigned long long x) { return (V)x; }
};

// Cross-lane moves for the push's segmented scan, as DPP controls (GFX9
// encodings): row_shr:n shifts within rows of 16 lanes, row_bcast:15 / :31
// hand the last lane of a row (of two rows) to the next row(s), wave_shr:1
// shifts the whole wavefront by one lane.  A lane with no source (or a row
// left out by the row mask) receives the combine's identity (OP::zero()).
#define WG_DPP_ROW_SHR(n) (0x110 + (n))
#define WG_DPP_ROW_BCAST15 0x142
#define WG_DPP_ROW_BCAST31 0x143
#define WG_DPP_WAVE_SHR1 0x138

// With every row enabled and an all-zero identity, bound_ctrl writes the zero
// into lanes without a source, so no old value has to be set up first.
template <class OP, int CTRL, int ROWS>
__device__ __forceinline__ typename OP::V pb_dpp_v(typename OP::V v)
{
    constexpr bool BC = OP::ZERO_BITS_0 && ROWS == 0xf;
    const unsigned long long x = OP::bits(v), z = BC ? 0ull : OP::bits(OP::zero());
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)z, (int)(unsigned)x, CTRL, ROWS, 0xf, BC);
    if constexpr (sizeof(typename OP::V) == 4)
        return OP::from_bits((unsigned)lo);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(z >> 32), (int)(unsigned)(x >> 32), CTRL, ROWS, 0xf, BC);
    return OP::from_bits(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// One step of the segmented inclusive scan: (v, f) of this lane combined
// with the partial result CTRL brings from a lower lane (a set flag starts a
// new segment, so a lane with f keeps its own value).
template <class OP, int CTRL, int ROWS>
__device__ __forceinline__ void pb_seg_dpp(typename OP::V& v, int& f)
{
    const typename OP::V o = pb_dpp_v<OP, CTRL, ROWS>(v);
    const int of = __builtin_amdgcn_update_dpp(0, f, CTRL, ROWS, 0xf, ROWS == 0xf);
    if (!f)
        v = OP::op(o, v);
    f |= of;
}

// A tile's inputs: the lane's codes and the tile's first pair number (a
// scalar load: t is uniform).
struct pb_tile_in {
    uint4 codes;
    uint32_t pbase;
};

__device__ __forceinline__ pb_tile_in pb_tile_load(const uint16_t* __restrict__ code,
                                                   const uint32_t* __restrict__ tile_pbase, uint32_t t, int lane)
{
    pb_tile_in in;
    t = __builtin_amdgcn_readfirstlane(t);
    in.codes = pb_load_codes(code, t, lane);
    in.pbase = tile_pbase[t];
    return in;
}

// One lane's share of a push tile in flight: its entries' values and end
// flags, then its open run and the run's segmented scan.
template <class OP>
struct pb_lane_tile {
    typename OP::V v[WG_PB_IPT];
    bool e[WG_PB_IPT];
    typename OP::V sc;
    int f;
};

// Tile phase 1 (one wavefront): lane l takes the tile's entries [8l, 8l + 8)
// and combines them in order; an entry flagged as a pair's last closes a pair,
// and the value still open at the lane's end is its run.  Every entry reads
// the window (padding entries read slot 0), so the pass has no branches.  It
// runs in two steps: pb_tile_fetch issues the entries' window reads (and the
// loader's per-entry loads), pb_tile_combine sums them once they are in, so
// the reads of two tiles can all be in flight before the first sum.  An
// invalid tile (past the task's end) has no entries.
template <class OP, bool WHOLE, typename W, class LOAD>
__device__ __forceinline__ void pb_tile_fetch(const pb_tile_in& in, uint32_t t, bool valid, const W* s_win,
                                              uint32_t wmask, pb_lane_tile<OP>& w, LOAD load)
{
    const uint4 raw = in.codes;
    const uint32_t c[WG_PB_IPT] = {raw.x & 0xffffu, raw.x >> 16, raw.y & 0xffffu, raw.y >> 16,
                                   raw.z & 0xffffu, raw.z >> 16, raw.w & 0xffffu, raw.w >> 16};
    // code index of each entry, for loaders that read per-entry data (edge
    // weights): an invalid tile's t is past the task (its codes are another
    // tile's), so its entries use index 0, which every layout has
    const uint32_t q0 = valid ? t * WG_PB_TILE + (uint32_t)wg_lane() * WG_PB_IPT : 0u;
#pragma unroll
    for (int j = 0; j < WG_PB_IPT; ++j)
        w.v[j] = load(s_win[c[j] & wmask], WHOLE ? q0 + (uint32_t)j : (valid ? q0 + (uint32_t)j : 0u));
}

template <class OP, bool WHOLE>
__device__ __forceinline__ void pb_tile_combine(const pb_tile_in& in, uint32_t t, bool valid, uint32_t wend,
                                                pb_lane_tile<OP>& w)
{
    typedef typename OP::V V;
    const uint4 raw = in.codes;
    const uint32_t c[WG_PB_IPT] = {raw.x & 0xffffu, raw.x >> 16, raw.y & 0xffffu, raw.y >> 16,
                                   raw.z & 0xffffu, raw.z >> 16, raw.w & 0xffffu, raw.w >> 16};
    const uint32_t q0 = valid ? t * WG_PB_TILE + (uint32_t)wg_lane() * WG_PB_IPT : 0u;
    V run = OP::zero();
    bool has_end = false;
#pragma unroll
    for (int j = 0; j < WG_PB_IPT; ++j) {
        // a WHOLE tile is valid and ends at or before the window's last real entry
        const bool real = WHOLE || (valid && q0 + (uint32_t)j < wend);
        if constexpr (!WHOLE)
            w.v[j] = real ? w.v[j] : OP::zero();
        w.e[j] = real && (c[j] & WG_PB_END);
        run = OP::op(run, w.v[j]);
        run = w.e[j] ? OP::zero() : run;
        has_end = has_end || w.e[j];
    }
    w.sc = run;
    w.f = has_end ? 1 : 0;
}

// Tile phase 2 for two tiles at once: the segmented inclusive scan of the
// lanes' runs (a lane with an end starts a new segment): shifts of 1, 2, 4, 8
// within each row of 16, then the row totals into rows 1 and 3, then the
// first two rows' total into rows 2 and 3.  A lane's f is set once its value
// reaches back to its segment's start (lane 0's segment starts with the tile),
// and the later steps leave such a lane as it is: the scan stops when every
// lane of both tiles has its f.  Most lanes hold a pair's end (a pair averages
// 2.5 entries), so most tile pairs stop after the first shift.  Each tile's
// sc becomes its carry, the scan one lane down (the value flowing into the
// lane's first entry).
template <class OP>
__device__ __forceinline__ bool pb_scan_open(const pb_lane_tile<OP>& a, const pb_lane_tile<OP>& b)
{
    return wg_ballot(!(a.f && b.f)) != 0ull;
}

template <class OP>
__device__ __forceinline__ void pb_tile_scan2(pb_lane_tile<OP>& a, pb_lane_tile<OP>& b)
{
    const int first = wg_lane() == 0 ? 1 : 0;
    a.f |= first;
    b.f |= first;
    pb_seg_dpp<OP, WG_DPP_ROW_SHR(1), 0xf>(a.sc, a.f);
    pb_seg_dpp<OP, WG_DPP_ROW_SHR(1), 0xf>(b.sc, b.f);
    if (pb_scan_open(a, b)) {
        pb_seg_dpp<OP, WG_DPP_ROW_SHR(2), 0xf>(a.sc, a.f);
        pb_seg_dpp<OP, WG_DPP_ROW_SHR(2), 0xf>(b.sc, b.f);
        if (pb_scan_open(a, b)) {
            pb_seg_dpp<OP, WG_DPP_ROW_SHR(4), 0xf>(a.sc, a.f);
            pb_seg_dpp<OP, WG_DPP_ROW_SHR(4), 0xf>(b.sc, b.f);
            pb_seg_dpp<OP, WG_DPP_ROW_SHR(8), 0xf>(a.sc, a.f);
            pb_seg_dpp<OP, WG_DPP_ROW_SHR(8), 0xf>(b.sc, b.f);
            if (pb_scan_open(a, b)) {
                pb_seg_dpp<OP, WG_DPP_ROW_BCAST15, 0xa>(a.sc, a.f);
                pb_seg_dpp<OP, WG_DPP_ROW_BCAST15, 0xa>(b.sc, b.f);
                pb_seg_dpp<OP, WG_DPP_ROW_BCAST31, 0xc>(a.sc, a.f);
                pb_seg_dpp<OP, WG_DPP_ROW_BCAST31, 0xc>(b.sc, b.f);
            }
        }
    }
    a.sc = pb_dpp_v<OP, WG_DPP_WAVE_SHR1, 0xf>(a.sc);
    b.sc = pb_dpp_v<OP, WG_DPP_WAVE_SHR1, 0xf>(b.sc);
}

// Tile phase 3: each entry flagged as a pair's last gives that pair's value,
// for part[p].  A tile's pairs are consecutive, so its stores fill one
// contiguous range of part[]: they pass through s_stage, this wavefront's LDS
// row, so that each store instruction writes 64 consecutive partials (one
// lane's pairs are consecutive, 64 lanes' are not: stored directly, each
// 8-byte store was its own L2 request).
template <class OP, typename PT>
__device__ __forceinline__ void pb_tile_store(const pb_tile_in& in, const pb_lane_tile<OP>& w, PT* __restrict__ part,
                                              typename OP::V* s_stage)
{
    typedef typename OP::V V;
    const int lane = wg_lane();
    // pre: pairs closed by lower lanes (mbcnt counts the mask's bits below
    // this lane and adds its second operand: two instructions per entry)
    uint32_t pre = 0, all = 0;
#pragma unroll
    for (int j = 0; j < WG_PB_IPT; ++j) {
        const unsigned long long m = wg_ballot(w.e[j]);
        pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, pre));
        all += (uint32_t)__popcll(m);
    }
    // values to the LDS row in pair order, WG_PB_STAGE pairs at a time (the
    // whole wavefront takes the same passes), then coalesced stores; a tile
    // with at most WG_PB_STAGE pairs (most) needs no range check
    const bool one = all <= WG_PB_STAGE; // uniform
    for (uint32_t h = 0; h < all; h += WG_PB_STAGE) {
        V r = w.sc;
        uint32_t i = pre;
#pragma unroll
        for (int j = 0; j < WG_PB_IPT; ++j) {
            r = OP::op(r, w.v[j]);
            if (w.e[j]) {
                if (one || i - h < WG_PB_STAGE) // i >= h: unsigned wrap puts i < h out of range
                    s_stage[i - h] = r;
                ++i;
                r = OP::zero();
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t n = all - h < WG_PB_STAGE ? all - h : WG_PB_STAGE;
        for (uint32_t k = (uint32_t)lane; k < n; k += WG_WAVE)
            part[in.pbase + h + k] = (PT)s_stage[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); // reads done before the row is written again
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Two tiles of one wavefront, their phases interleaved (independent
// instructions for the scheduler: a wavefront's tile is a chain of dependent
// steps, and a CU holds only the 16 wavefronts of one workgroup).
template <class OP, typename W, typename PT, class LOAD>
__device__ __forceinline__ void pb_push_pair(const pb_tile_in& x, uint32_t tx, const pb_tile_in& y, uint32_t ty,
                                             bool y_valid, const W* s_win, uint32_t wmask, uint32_t wend,
                                             PT* __restrict__ part, typename OP::V* s_stage, LOAD load)
{
    pb_lane_tile<OP> a, b;
    // both tiles free of padding and of the task's end (all but a window's
    // last tile): no per-entry validity selects
    const bool whole = (tx + 1u) * WG_PB_TILE <= wend && y_valid && (ty + 1u) * WG_PB_TILE <= wend;
    // all 16 window reads of the two tiles issued before the first sum (the
    // scheduler left four in flight, waiting for each group)
    if (whole) {
        pb_tile_fetch<OP, true>(x, tx, true, s_win, wmask, a, load);
        pb_tile_fetch<OP, true>(y, ty, true, s_win, wmask, b, load);
        __builtin_amdgcn_sched_barrier(0);
        pb_tile_combine<OP, true>(x, tx, true, wend, a);
        pb_tile_combine<OP, true>(y, ty, true, wend, b);
    } else {
        pb_tile_fetch<OP, false>(x, tx, true, s_win, wmask, a, load);
        pb_tile_fetch<OP, false>(y, ty, y_valid, s_win, wmask, b, load);
        __builtin_amdgcn_sched_barrier(0);
        pb_tile_combine<OP, false>(x, tx, true, wend, a);
        pb_tile_combine<OP, false>(y, ty, y_valid, wend, b);
    }
    pb_tile_scan2<OP>(a, b);
    pb_tile_store<OP>(x, a, part, s_stage);
    pb_tile_store<OP>(y, b, part, s_stage);
}

// Push body: one workgroup per task (a window's run of tiles), one wavefront
// per tile pair.  Four tiles' inputs rotate through four sets of registers,
// each pair reloaded (index clamped, never skipped) as soon as it is done, so
// one pair's loads are in flight while the other pair runs.  s_win holds the
// window (filled by the caller, converted by LOAD on read; wmask = its size
// - 1).
template <class OP, typename W, typename PT, class LOAD>
__device__ __forceinline__ void pb_push_tiles(const uint16_t* __restrict__ code, const W* s_win, uint32_t wmask,
                                              const uint4 tk, const uint32_t* __restrict__ tile_pbase,
                                              PT* __restrict__ part, typename OP::V* s_stage, LOAD load)
{
    constexpr uint32_t S = WG_PB_WAVES; // tiles between one wavefront's consecutive tiles
    const int lane = wg_lane();
    uint32_t t = tk.y + (uint32_t)(threadIdx.x / WG_WAVE);
    if (t >= tk.z)
        return;
    const uint32_t tlast = tk.z - 1u;
    auto ld = [&](uint32_t u) { return pb_tile_load(code, tile_pbase, u < tk.z ? u : tlast, lane); };
    pb_tile_in a = ld(t), b = ld(t + S), c = ld(t + 2 * S), d = ld(t + 3 * S);
    for (;;) {
        pb_push_pair<OP>(a, t, b, t + S, t + S < tk.z, s_win, wmask, tk.w, part, s_stage, load);
        a = ld(t + 4 * S);
        b = ld(t + 5 * S);
        if (t + 2 * S >= tk.z)
            break;
        pb_push_pair<OP>(c, t + 2 * S, d, t + 3 * S, t + 3 * S < tk.z, s_win, wmask, tk.w, part, s_stage, load);
        c = ld(t + 6 * S);
        d = ld(t + 7 * S);
        t += 4 * S;
        if (t >= tk.z)
            break;
    }
}

// PageRank push: the window holds contributions (FP64, or FP32 stored).  A
// pair's partial is summed in FP64 and stored as T: FP64, or FP32 with the
// FP32-stored contributions (one more rounding per pair, in the error the
// FP32 leg reports).
template <typename T, int WS>
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_push(const uint16_t* __restrict__ code, const T* __restrict__ cin,
                                                           const uint4* __restrict__ task, uint32_t ntasks,
                                                           const uint32_t* __restrict__ tile_pbase,
                                                           T* __restrict__ part, uint32_t V)
{
    static_assert(WG_WAVE * WG_PB_IPT == WG_PB_TILE, "a tile is one wavefront's entries");
    static_assert(WS <= WG_PB_SRC + 1, "window larger than the source field");
    __shared__ T s_win[WS];
    __shared__ double s_stage[WG_PB_WAVES][WG_PB_STAGE];
    for (uint32_t k = blockIdx.x; k < ntasks; k += gridDim.x) {
        const uint4 tk = task[k];
        const uint32_t v0 = tk.x * (uint32_t)WS;
        const uint32_t nv = (V - v0 < (uint32_t)WS) ? V - v0 : (uint32_t)WS;
        __syncthreads(); // the previous task's wavefronts are done with s_win
        pb_fill_lds<WS>(s_win, cin + v0, nv, [](T x) { return x; });
        __syncthreads();
        pb_push_tiles<pb_sum>(code, s_win, (uint32_t)WS - 1u, tk, tile_pbase, part, s_stage[threadIdx.x / WG_WAVE],
                              [](T x, uint32_t) { return (double)x; });
    }
}

// The slice a pull workgroup takes.  Slices of neighbouring bins share the
// lines where a window's blocks for the two bins meet, and the hardware hands
// workgroups to the 8 XCDs in turn (workgroup i to XCD i mod 8), each with its
// own L2.  Within each group of 128 workgroups, workgroup 8a + b takes slice
// 16b + a, so each XCD runs 16 consecutive slices side by side and reads their
// shared lines from one L2 (runs of 4, 8, 16 and 32 measured: 16 is best).  A
// bijection on [0, n): the last partial group keeps its order.
#define WG_PB_XCD_RUN 16u
__device__ __forceinline__ uint32_t pb_slice_of(uint32_t i, uint32_t n)
{
    constexpr uint32_t G = 8u * WG_PB_XCD_RUN;
    const uint32_t g = i - i % G;
    if (g + G > n)
        return i;
    const uint32_t r = i % G;
    return g + (r % 8u) * WG_PB_XCD_RUN + r / 8u;
}

// The pull's pass over one slice, bin-order positions [b0, b1):
// ADD(destination, partial) for every pair.  Position o of block kb holds pair
// o - boff[kb] of part[] (window order).  Positions go 64 to a wavefront step
// (chunk c = [64c, 64c + 64)): the block-start flags in pdst count the blocks
// after the one holding 64c, and the offsets of the chunk's first
// WG_PB_CHUNK_OFFS blocks, kept with the chunk, come from the lanes that
// loaded them, so the destinations, the offsets and then the partials are
// two rounds of loads.  A chunk with more blocks (rare) loads its block
// number, chunk_block[c], then the blocks' offsets, boff[].  R chunks per wavefront are
// in flight at a time: the pull waits on this chain of loads, so more is
// better as far as registers allow (the PageRank pull at s27: R = 6, 8, 10,
// 12 took 3.66, 3.60, 3.57, 3.53 ms; 12 for FP64 PageRank partials, 8 for
// FP32 ones (12 took their pull from 3.27 to 4.45 ms) and for the minimum,
// 4 for the 64-source BFS, whose epilogue keeps more registers).
template <int R, typename X, class ADD>
__device__ __forceinline__ void pb_pull_range(const X* __restrict__ part, const uint16_t* __restrict__ pdst,
                                              const uint32_t* __restrict__ chunk_block,
                                              const uint32_t* __restrict__ boff, uint32_t nblocks, uint32_t npairs,
                                              uint32_t b0, uint32_t b1, ADD add)
{
    if (b0 >= b1)
        return;
    const int lane = wg_lane();
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull; // lanes 0..lane
    const uint32_t wave = threadIdx.x / WG_WAVE;
    const uint32_t c1 = ((b1 - 1u) >> 6) + 1u;
    const uint32_t* __restrict__ cbo = chunk_block + ((npairs + 63u) >> 6) + 1u;
    for (uint32_t c = (b0 >> 6) + wave; c < c1; c += WG_PB_WAVES * R) {
        uint32_t o[R], pd[R], bo[R];
        X x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t cc = c + (uint32_t)r * WG_PB_WAVES; // uniform
            o[r] = cc * 64u + (uint32_t)lane;
            pd[r] = 0u;
            bo[r] = 0u;
            if (cc < c1) {
                if (o[r] < npairs)
                    pd[r] = pdst[o[r]];
                if ((uint32_t)lane < WG_PB_CHUNK_OFFS)
                    bo[r] = cbo[(uint64_t)cc * WG_PB_CHUNK_OFFS + (uint32_t)lane];
            }
        }
        unsigned slow = 0u; // uniform: chunks with more blocks than kept offsets, done below
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // blocks started after the chunk's first position: up to this one, and in all
            const unsigned long long starts = wg_ballot(lane > 0 && (pd[r] & WG_PB_BFLAG));
            const uint32_t k = (uint32_t)__popcll(starts & upto);
            if ((uint32_t)__popcll(starts) >= WG_PB_CHUNK_OFFS)
                slow |= 1u << r;
            const uint32_t off = (uint32_t)__shfl((int)bo[r], (int)k, WG_WAVE);
            x[r] = X(0);
            if (!((slow >> r) & 1u) && o[r] >= b0 && o[r] < b1)
                x[r] = __builtin_nontemporal_load(&part[o[r] - off]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (!((slow >> r) & 1u) && o[r] >= b0 && o[r] < b1)
                add(pd[r] & WG_PB_DMASK, x[r]);
        while (slow) { // the block number, then every block offset the chunk needs
            const uint32_t r = (uint32_t)__builtin_ctz(slow);
            slow &= slow - 1u;
            const uint32_t cc = c + r * WG_PB_WAVES;
            const uint32_t oo = cc * 64u + (uint32_t)lane;
            const uint32_t pdv = oo < npairs ? pdst[oo] : 0u;
            const uint32_t k0 = chunk_block[cc];
            const uint32_t bov = k0 + (uint32_t)lane < nblocks ? boff[k0 + (uint32_t)lane] : 0u;
            const uint32_t k = (uint32_t)__popcll(wg_ballot(lane > 0 && (pdv & WG_PB_BFLAG)) & upto);
            const uint32_t off = (uint32_t)__shfl((int)bov, (int)k, WG_WAVE);
            if (oo >= b0 && oo < b1)
                add(pdv & WG_PB_DMASK, __builtin_nontemporal_load(&part[oo - off]));
        }
    }
}

// A bin cut into several slices: each slice stores its LDS words in its slot,
// and the last slice of the bin to finish (a per-bin ticket) combines the
// slots in slice order into its LDS words.  Returns false in the other slices.
// One thread releases the workgroup's stores and acquires the other slices'
// (device-scope fences after the workgroup barrier).
template <typename X, class COMBINE>
__device__ __forceinline__ bool pb_join_slices(X* s_acc, uint32_t nv, uint32_t b, uint32_t slot,
                                               const uint2* __restrict__ bin_slot, X* __restrict__ slot_acc,
                                               unsigned int* __restrict__ bin_ticket, uint32_t wd, int* s_flag,
                                               COMBINE combine)
{
    const uint2 bs = bin_slot[b];
    X* mine = slot_acc + (uint64_t)slot * wd;
    for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS)
        mine[i] = s_acc[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const bool last = atomicAdd(&bin_ticket[b], 1u) == bs.y - 1u;
        if (last) {
            __threadfence();
            bin_ticket[b] = 0u; // for the next iteration (the kernel boundary orders it)
        }
        *s_flag = last ? 1 : 0;
    }
    __syncthreads();
    if (!*s_flag)
        return false;
    for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS) {
        X a = slot_acc[(uint64_t)bs.x * wd + i];
        for (uint32_t j = 1; j < bs.y; ++j)
            a = combine(a, slot_acc[(uint64_t)(bs.x + j) * wd + i]);
        s_acc[i] = a;
    }
    return true;
}

// Pull: one workgroup per slice, a contiguous range of part[] within one bin.
// Adds the slice's partials into the bin's accumulators in LDS (ds_add_f64);
// a bin cut into several slices joins them (pb_join_slices).  The workgroup
// holding a bin's complete accumulators then runs the vertex phase (MODE 0:
// contributions and the sum of ranks of vertices with out-degree > 0; MODE 1:
// accumulators and their sum) and stores the bin's sum in bin_part[bin];
// k_pb_total adds those.
template <typename T, int MODE, int WD>
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_pull(
    const T* __restrict__ part, const uint16_t* __restrict__ pdst, const uint32_t* __restrict__ chunk_block,
    const uint32_t* __restrict__ boff, uint32_t nblocks, uint32_t npairs, const uint4* __restrict__ slice,
    uint32_t nslices, const uint2* __restrict__ bin_slot, double* __restrict__ slot_acc,
    unsigned int* __restrict__ bin_ticket, double* __restrict__ bin_part, T* __restrict__ cout,
    double* __restrict__ acc_out, const uint32_t* __restrict__ od, uint32_t V, const double* __restrict__ S_in,
    double damping, double N)
{
    __shared__ double s_acc[WD];
    __shared__ double s_red[WG_PB_WAVES];
    __shared__ int s_flag;
    double corr = 0.0, omd_n = 0.0;
    if (MODE == 0) {
        corr = (1.0 - *S_in) / N;
        omd_n = (1.0 - damping) / N;
    }
    for (uint32_t k = blockIdx.x; k < nslices; k += gridDim.x) {
        const uint4 sl = slice[pb_slice_of(k, nslices)];
        const uint32_t b = sl.x, v0 = b * (uint32_t)WD;
        const uint32_t nv = (V - v0 < (uint32_t)WD) ? V - v0 : (uint32_t)WD;
        // the bin's out-degrees for the vertex phase, loaded while the slice streams
        constexpr int R = WD / WG_PB_THREADS;
        uint32_t odr[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const uint32_t i = threadIdx.x + (uint32_t)u * WG_PB_THREADS;
            odr[u] = (MODE == 0 && i < nv) ? od[v0 + i] : 0u;
        }
        __syncthreads(); // the previous slice is done with s_acc
        for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS)
            s_acc[i] = 0.0;
        __syncthreads();
        pb_pull_range<sizeof(T) == 8 ? 12 : 8>(part, pdst, chunk_block, boff, nblocks, npairs, sl.y, sl.z,
                                               [&](uint32_t d, T x) { atomicAdd(&s_acc[d], (double)x); });
        __syncthreads();
        if (sl.w != WG_PB_NOSLOT &&
            !pb_join_slices(s_acc, nv, b, sl.w, bin_slot, slot_acc, bin_ticket, (uint32_t)WD, &s_flag,
                            [](double a, double x) { return a + x; }))
            continue;
        double s_part = 0.0;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const uint32_t i = threadIdx.x + (uint32_t)u * WG_PB_THREADS;
            if (i >= nv) // and every later u
                continue;
            const uint32_t v = v0 + i;
            const double accv = s_acc[i];
            if (MODE == 0) {
                const uint32_t odv = odr[u];
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
        s_part = block_sum_pb(s_part, s_red);
        if (threadIdx.x == 0)
            bin_part[b] = s_part;
    }
}

// S_out = the bins' sums, added in a fixed order (one workgroup).
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_total(const double* __restrict__ bin_part, uint32_t nbins,
                                                            double* __restrict__ S_out)
{
    __shared__ double s_red[WG_PB_WAVES];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < nbins; i += WG_PB_THREADS)
        s += bin_part[i];
    s = block_sum_pb(s, s_red);
    if (threadIdx.x == 0)
        *S_out = s;
}

// ---- host side --------------------------------------------------------------------

template <typename T, int WS, int WD>
static int pb_launch(wg_ctx* ctx, int mode, const T* cin, T* cout, double* S_in, double* S_out)
{
    const wg_pb* pb = ctx->pb;
    const double N = (double)ctx->V;
    const uint32_t V = (uint32_t)ctx->V;
    if (pb->ntasks) {
        const int slot = wg_prof_begin(ctx, WG_K_PB_PUSH);
        hipLaunchKernelGGL((k_pb_push<T, WS>), dim3(pb->ntasks), dim3(WG_PB_THREADS), 0, ctx->stream, pb->code, cin,
                           pb->task, pb->ntasks, pb->tile_pbase, reinterpret_cast<T*>(pb->part), V);
        int rc = wg_check_launch(ctx);
        if (rc)
            return rc;
        wg_prof_end(ctx, WG_K_PB_PUSH, slot);
    }
    const int slot = wg_prof_begin(ctx, WG_K_PB_PULL);
    if (mode == 0)
        hipLaunchKernelGGL((k_pb_pull<T, 0, WD>), dim3(pb->nslices), dim3(WG_PB_THREADS), 0, ctx->stream,
                           (const T*)reinterpret_cast<T*>(pb->part),
                           pb->pdst, pb->chunk_block, pb->boff, pb->nblocks, (uint32_t)pb->npairs, pb->slice, pb->nslices, pb->bin_slot, pb->slot_acc, pb->bin_ticket, pb->bin_part,
                           cout, nullptr, ctx->od, V, S_in, ctx->damping, N);
    else
        hipLaunchKernelGGL((k_pb_pull<T, 1, WD>), dim3(pb->nslices), dim3(WG_PB_THREADS), 0, ctx->stream,
                           (const T*)reinterpret_cast<T*>(pb->part),
                           pb->pdst, pb->chunk_block, pb->boff, pb->nblocks, (uint32_t)pb->npairs, pb->slice, pb->nslices, pb->bin_slot, pb->slot_acc, pb->bin_ticket, pb->bin_part,
                           nullptr, ctx->acc, ctx->od, V, S_in, ctx->damping, N);
    int rc = wg_check_launch(ctx);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_pb_total, dim3(1), dim3(WG_PB_THREADS), 0, ctx->stream, pb->bin_part, pb->B, S_out);
    if ((rc = wg_check_launch(ctx)))
        return rc;
    wg_prof_end(ctx, WG_K_PB_PULL, slot);
    return WG_OK;
}

int wg_pb_iteration(wg_ctx* ctx, int mode, const void* cin, void* cout, double* S_in, double* S_out)
{
    int rc = wg_pb_prepare(ctx);
    if (rc)
        return rc;
    const bool big = ctx->pr_bpc == 1;
    if (ctx->pr_fp32)
        return big ? pb_launch<float, 16384, 16384>(ctx, mode, (const float*)cin, (float*)cout, S_in, S_out)
                   : pb_launch<float, 8192, 8192>(ctx, mode, (const float*)cin, (float*)cout, S_in, S_out);
    return big ? pb_launch<double, 16384, 16384>(ctx, mode, (const double*)cin, (double*)cout, S_in, S_out)
               : pb_launch<double, 8192, 8192>(ctx, mode, (const double*)cin, (double*)cout, S_in, S_out);
}

int wg_pb_stats(const wg_ctx* ctx, uint64_t* pairs, uint64_t* code_entries, uint32_t* tasks, uint32_t* slices)
{
    if (!ctx->pb)
        return WG_E_STATE;
    *pairs = ctx->pb->npairs;
    *code_entries = ctx->pb->nq;
    *tasks = ctx->pb->ntasks;
    *slices = ctx->pb->nslices;
    return WG_OK;
}

// ---- 64-source BFS: a dense level on the binned layout ------------------------------
//
// The dense level of the batched BFS (wg_batch.hip) ORs, for every vertex, the
// frontier masks of its in-neighbours (64 sources, one bit each).  OR is
// associative, commutative and idempotent, so the same two passes apply, with
// results independent of the order: the push ORs each (window, destination)
// pair's frontier masks, the pull ORs the pairs into the bin's LDS words
// (ds_or_b64) and keeps the sources each vertex still needs.

// Push: as k_pb_push, with 64-bit masks (front & live) in place of
// contributions and OR in place of the sum.
template <int WS>
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_push_or(const uint16_t* __restrict__ code,
                                                              const unsigned long long* __restrict__ front,
                                                              unsigned long long live, const uint4* __restrict__ task,
                                                              uint32_t ntasks, const uint32_t* __restrict__ tile_pbase,
                                                              unsigned long long* __restrict__ part, uint32_t V)
{
    __shared__ unsigned long long s_win[WS];
    __shared__ unsigned long long s_stage[WG_PB_WAVES][WG_PB_STAGE];
    for (uint32_t k = blockIdx.x; k < ntasks; k += gridDim.x) {
        const uint4 tk = task[k];
        const uint32_t v0 = tk.x * (uint32_t)WS;
        const uint32_t nv = (V - v0 < (uint32_t)WS) ? V - v0 : (uint32_t)WS;
        __syncthreads();
        pb_fill_lds<WS>(s_win, front + v0, nv, [live](unsigned long long x) { return x & live; });
        __syncthreads();
        pb_push_tiles<pb_or>(code, s_win, (uint32_t)WS - 1u, tk, tile_pbase, part, s_stage[threadIdx.x / WG_WAVE],
                             [](unsigned long long x, uint32_t) { return x; });
    }
}

// Pull: ORs a slice's masks into the bin's LDS words; the workgroup holding a
// bin's complete words then commits the level for the bin's vertices (what
// k_ms_commit does after a frontier level): n = the live sources not seen yet,
// seen |= n, front = n, the depth row entries of the new sources, the frontier
// vertex bitmask, and the counts (out-degree and number of vertices with new
// sources, new (vertex, source) pairs).
template <int WD, bool ROWS>
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_pull_or(
    const unsigned long long* __restrict__ part, const uint16_t* __restrict__ pdst,
    const uint32_t* __restrict__ chunk_block, const uint32_t* __restrict__ boff, uint32_t nblocks, uint32_t npairs,
    const uint4* __restrict__ slice, uint32_t nslices, const uint2* __restrict__ bin_slot,
    unsigned long long* __restrict__ slot_acc, unsigned int* __restrict__ bin_ticket,
    unsigned long long* __restrict__ seen, unsigned long long* __restrict__ front, uint16_t* __restrict__ depth,
    uint64_t* __restrict__ bm, const uint32_t* __restrict__ od, unsigned long long live, uint32_t level,
    wg_dev_counters* __restrict__ ctr, uint32_t V)
{
    __shared__ unsigned long long s_acc[WD];
    __shared__ int s_flag;
    const int lane = wg_lane();
    act_tally t = { 0ull, 0ull, 0ull };
    for (uint32_t k = blockIdx.x; k < nslices; k += gridDim.x) {
        const uint4 sl = slice[pb_slice_of(k, nslices)];
        const uint32_t b = sl.x, v0 = b * (uint32_t)WD;
        const uint32_t nv = (V - v0 < (uint32_t)WD) ? V - v0 : (uint32_t)WD;
        // the bin's seen masks for the commit, loaded while the slice streams
        constexpr int R = WD / WG_PB_THREADS;
        unsigned long long sr[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const uint32_t i = threadIdx.x + (uint32_t)u * WG_PB_THREADS;
            sr[u] = i < nv ? seen[v0 + i] : ~0ull;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS)
            s_acc[i] = 0ull;
        __syncthreads();
        pb_pull_range<4>(part, pdst, chunk_block, boff, nblocks, npairs, sl.y, sl.z, [&](uint32_t d, unsigned long long x) {
            if (x)
                atomicOr(&s_acc[d], x);
        });
        __syncthreads();
        if (sl.w != WG_PB_NOSLOT &&
            !pb_join_slices(s_acc, nv, b, sl.w, bin_slot, slot_acc, bin_ticket, (uint32_t)WD, &s_flag,
                            [](unsigned long long a, unsigned long long x) { return a | x; }))
            continue;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const uint32_t i = threadIdx.x + (uint32_t)u * WG_PB_THREADS; // a wavefront: 64 consecutive vertices
            const uint32_t v = v0 + i;
            const unsigned long long a = i < nv ? s_acc[i] : 0ull;
            const unsigned long long n = a & live & ~sr[u];
            if (i < nv) {
                if (n) {
                    seen[v] = sr[u] | n;
                    t.edges += od[v];
                    t.verts += 1ull;
                    t.fresh += (unsigned long long)__popcll(n);
                }
                front[v] = n;
            }
            unsigned long long has = wg_ballot(n != 0ull);
            const uint32_t w0 = v - (uint32_t)lane; // the wavefront's first vertex (a multiple of 64)
            if (lane == 0 && i < nv)
                bm[w0 >> 6] = has;
            if (!ROWS) // the caller keeps front and writes the rows later (k_ms_rows)
                continue;
            while (has) { // one row store per vertex with new sources
                const int l = __ffsll((long long)has) - 1;
                has &= has - 1ull;
                const unsigned long long bits = (unsigned long long)__shfl((long long)n, l);
                if ((bits >> lane) & 1ull)
                    depth[((uint64_t)w0 + (uint64_t)l) * WG_MS_MAX_SOURCES + (uint64_t)lane] = (uint16_t)level;
            }
        }
    }
    add_tally(t, ctr);
}

int wg_pb_msbfs_level(wg_ctx* ctx, const unsigned long long* front_in, unsigned long long* front_out,
                      unsigned long long* seen, uint16_t* depth, unsigned long long live, uint32_t level, bool rows)
{
    int rc = pb_prepare_sizes(ctx, 14, 14); // 16K 8-byte masks per window, 16K words per bin
    if (rc)
        return rc;
    const wg_pb* pb = ctx->pb;
    const uint32_t V = (uint32_t)ctx->V;
    unsigned long long* part = reinterpret_cast<unsigned long long*>(pb->part);
    if (pb->ntasks)
        hipLaunchKernelGGL(k_pb_push_or<16384>, dim3(pb->ntasks), dim3(WG_PB_THREADS), 0, ctx->stream, pb->code,
                           front_in, live, pb->task, pb->ntasks, pb->tile_pbase, part, V);
    if ((rc = wg_check_launch(ctx)))
        return rc;
    if (rows)
        hipLaunchKernelGGL((k_pb_pull_or<16384, true>), dim3(pb->nslices), dim3(WG_PB_THREADS), 0, ctx->stream,
                           (const unsigned long long*)part, pb->pdst, pb->chunk_block, pb->boff, pb->nblocks,
                           (uint32_t)pb->npairs, pb->slice, pb->nslices, pb->bin_slot,
                           reinterpret_cast<unsigned long long*>(pb->slot_acc), pb->bin_ticket, seen, front_out,
                           depth, ctx->bm_next, ctx->od, live, level, ctx->ctr, V);
    else
        hipLaunchKernelGGL((k_pb_pull_or<16384, false>), dim3(pb->nslices), dim3(WG_PB_THREADS), 0, ctx->stream,
                           (const unsigned long long*)part, pb->pdst, pb->chunk_block, pb->boff, pb->nblocks,
                           (uint32_t)pb->npairs, pb->slice, pb->nslices, pb->bin_slot,
                           reinterpret_cast<unsigned long long*>(pb->slot_acc), pb->bin_ticket, seen, front_out,
                           depth, ctx->bm_next, ctx->od, live, level, ctx->ctr, V);
    return wg_check_launch(ctx);
}

// ---- dense CC / SSSP pulls (minimum) ---------------------------------------------
//
// A dense pull of perform_edge_pull_phase_cc / _sssp (source/phase_edge_pull_cc.c,
// source/phase_edge_pull_sssp.c): every destination takes the minimum of what
// its in-edges offer (the source's label, or its distance plus the edge
// weight) when that is below its value, and is activated.  On the binned
// layout the push takes, per (window, destination) pair, the minimum of the
// window's offers (the window's values in LDS, the weights streamed in code
// order) and the pull the minimum of a bin's pairs with LDS atomics, then
// applies it: a lower value is stored, the vertex is set in `next` and
// counted (out-degree, and whether it was reached for the first time) as the
// tile pull counts it.  The push reads the values the previous pull left, so
// the pull is a Jacobi step (the tile pull may see values lowered earlier in
// the same pull); the fixpoint is the same, since minimum is idempotent.

// Push: the window holds values (labels, or distances as bit patterns).
template <bool CW, class P, int WS>
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_push_min(const uint16_t* __restrict__ code,
                                                               const typename P::T* __restrict__ prop,
                                                               const void* __restrict__ wq, const uint4* __restrict__ task,
                                                               uint32_t ntasks, const uint32_t* __restrict__ tile_pbase,
                                                               typename P::T* __restrict__ part, uint32_t V,
                                                               wg_dev_counters* __restrict__ ctr,
                                                               unsigned int* __restrict__ ovf, int gate)
{
    typedef typename P::T T;
    if (wg_gated_off(ctr, gate))
        return;
    __shared__ T s_win[WS];
    __shared__ T s_stage[WG_PB_WAVES][WG_PB_STAGE];
    work_tally<CW> work;
    for (uint32_t k = blockIdx.x; k < ntasks; k += gridDim.x) {
        const uint4 tk = task[k];
        const uint32_t v0 = tk.x * (uint32_t)WS;
        const uint32_t nv = (V - v0 < (uint32_t)WS) ? V - v0 : (uint32_t)WS;
        __syncthreads(); // the previous task's wavefronts are done with s_win
        pb_fill_lds<WS>(s_win, prop + v0, nv, [](T x) { return x; });
        __syncthreads();
        pb_push_tiles<pb_min<P>>(code, s_win, (uint32_t)WS - 1u, tk, tile_pbase, part, s_stage[threadIdx.x / WG_WAVE],
                                 [&](T du, uint32_t q) { return P::val_from(du, wq, q, ovf); });
        if (threadIdx.x == 0) { // the task's edges: its real entries
            const uint32_t end = tk.z * WG_PB_TILE < tk.w ? tk.z * WG_PB_TILE : tk.w;
            work.edges += end - tk.y * WG_PB_TILE;
        }
    }
    add_work(work, ctr);
}

// Pull: the minimum of a slice's pairs per destination in LDS; the workgroup
// holding a bin's complete minima applies them to the bin's vertices.
template <bool CW, class P, int WD>
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_pull_min(
    const typename P::T* __restrict__ part, const uint16_t* __restrict__ pdst, const uint32_t* __restrict__ chunk_block,
    const uint32_t* __restrict__ boff, uint32_t nblocks, uint32_t npairs, const uint4* __restrict__ slice,
    uint32_t nslices, const uint2* __restrict__ bin_slot, typename P::T* __restrict__ slot_acc,
    unsigned int* __restrict__ bin_ticket, typename P::T* __restrict__ prop, const uint32_t* __restrict__ od,
    uint64_t* __restrict__ next, uint32_t V, wg_dev_counters* __restrict__ ctr, int gate)
{
    typedef typename P::T T;
    if (wg_gated_off(ctr, gate))
        return;
    __shared__ T s_acc[WD];
    __shared__ int s_flag;
    act_tally t = { 0ull, 0ull, 0ull };
    work_tally<CW> work;
    for (uint32_t k = blockIdx.x; k < nslices; k += gridDim.x) {
        const uint4 sl = slice[pb_slice_of(k, nslices)];
        const uint32_t b = sl.x, v0 = b * (uint32_t)WD;
        const uint32_t nv = (V - v0 < (uint32_t)WD) ? V - v0 : (uint32_t)WD;
        __syncthreads(); // the previous slice is done with s_acc
        for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS)
            s_acc[i] = P::INF;
        __syncthreads();
        pb_pull_range<8>(part, pdst, chunk_block, boff, nblocks, npairs, sl.y, sl.z, [&](uint32_t d, T x) {
            if (x < s_acc[d])
                atomicMin(&s_acc[d], x);
        });
        __syncthreads();
        if (sl.w != WG_PB_NOSLOT &&
            !pb_join_slices(s_acc, nv, b, sl.w, bin_slot, slot_acc, bin_ticket, (uint32_t)WD, &s_flag,
                            [](T a, T x) { return x < a ? x : a; }))
            continue;
        if (threadIdx.x == 0)
            work.dests += nv;
        constexpr int R = WD / WG_PB_THREADS;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const uint32_t i = threadIdx.x + (uint32_t)u * WG_PB_THREADS; // a wavefront: 64 consecutive vertices
            bool changed = false;
            if (i < nv) {
                const T m = s_acc[i];
                if (m != P::INF) {
                    const T old = prop[v0 + i];
                    if (m < old) { // the bin's vertices are this workgroup's alone
                        prop[v0 + i] = m;
                        changed = true;
                        if (old == P::INF)
                            t.fresh += 1ull;
                    }
                }
            }
            activate_contiguous(next, od, v0 + i - (uint32_t)wg_lane(), changed, t);
        }
    }
    add_tally(t, ctr);
    add_work(work, ctr);
}

bool wg_pb_min_on(const wg_ctx* ctx)
{
    if (ctx->app != WG_APP_CC && ctx->app != WG_APP_SSSP)
        return false;
    if (ctx->app == WG_APP_SSSP && ctx->wkind == WG_WEIGHT_NONE)
        return false;
    const int opt = ctx->min_binned ? ctx->min_binned - 1 : 2;
    return opt == 1 || (opt == 2 && ctx->E >= (1ull << 24) && ctx->E >= 8ull * ctx->V);
}

extern "C" int wg_dense_pull_binned(const wg_ctx* ctx, int* on)
{
    if (!ctx || !on)
        return WG_E_INVAL;
    *on = wg_pb_min_on(ctx) ? 1 : 0;
    return WG_OK;
}

int wg_pb_min_prepare(wg_ctx* ctx)
{
    int rc = pb_prepare_sizes(ctx, 14, 14); // 16K values per window, 16K minima per bin
    if (rc)
        return rc;
    if (ctx->app == WG_APP_SSSP && ctx->pb->nq && !ctx->pb->wq) // built before the weights were loaded
        return WG_E_STATE;
    return WG_OK;
}

template <class P>
static int pb_min_launch(wg_ctx* ctx, int gate)
{
    typedef typename P::T T;
    const wg_pb* pb = ctx->pb;
    const uint32_t V = (uint32_t)ctx->V;
    T* part = reinterpret_cast<T*>(pb->part);
    T* prop = (T*)ctx->prop[0];
    // gated-off launches stay cheap: at most 4 workgroups per CU, each looping over tasks / slices
    const uint32_t cap = (uint32_t)ctx->num_cu * 4u;
    const uint32_t pg = pb->ntasks < cap ? pb->ntasks : cap, lg = pb->nslices < cap ? pb->nslices : cap;
    if (pb->ntasks) {
        if (ctx->count_work)
            hipLaunchKernelGGL((k_pb_push_min<true, P, 16384>), dim3(pg), dim3(WG_PB_THREADS), 0, ctx->stream, pb->code,
                               (const T*)prop, pb->wq, pb->task, pb->ntasks, pb->tile_pbase, part, V, ctx->ctr,
                               &ctx->ctr->overflow, gate);
        else
            hipLaunchKernelGGL((k_pb_push_min<false, P, 16384>), dim3(pg), dim3(WG_PB_THREADS), 0, ctx->stream, pb->code,
                               (const T*)prop, pb->wq, pb->task, pb->ntasks, pb->tile_pbase, part, V, ctx->ctr,
                               &ctx->ctr->overflow, gate);
    }
    int rc = wg_check_launch(ctx);
    if (rc)
        return rc;
    if (ctx->count_work)
        hipLaunchKernelGGL((k_pb_pull_min<true, P, 16384>), dim3(lg), dim3(WG_PB_THREADS), 0, ctx->stream,
                           (const T*)part, pb->pdst, pb->chunk_block, pb->boff, pb->nblocks, (uint32_t)pb->npairs,
                           pb->slice, pb->nslices, pb->bin_slot, reinterpret_cast<T*>(pb->slot_acc), pb->bin_ticket,
                           prop, ctx->od, ctx->bm_next, V, ctx->ctr, gate);
    else
        hipLaunchKernelGGL((k_pb_pull_min<false, P, 16384>), dim3(lg), dim3(WG_PB_THREADS), 0, ctx->stream,
                           (const T*)part, pb->pdst, pb->chunk_block, pb->boff, pb->nblocks, (uint32_t)pb->npairs,
                           pb->slice, pb->nslices, pb->bin_slot, reinterpret_cast<T*>(pb->slot_acc), pb->bin_ticket,
                           prop, ctx->od, ctx->bm_next, V, ctx->ctr, gate);
    return wg_check_launch(ctx);
}

int wg_pb_min_pull(wg_ctx* ctx, int gate)
{
    if (!ctx->pb || ctx->pb->ws_log != 14 || ctx->pb->wd_log != 14)
        return WG_E_STATE; // wg_pb_min_prepare first
    if (ctx->app == WG_APP_CC)
        return pb_min_launch<PolCC>(ctx, gate);
    if (ctx->wkind == WG_WEIGHT_F64)
        return pb_min_launch<PolSF>(ctx, gate);
    return pb_min_launch<PolSU>(ctx, gate);
}
