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
//            window x destination) becomes one FP64 partial sum
//   pull     per bin: a workgroup adds the partials of every window into LDS
//            accumulators, then runs the vertex phase on the bin's vertices
// Bytes per iteration: 2 per edge, 18 per pair (the partial written and read
// back, and its 2-byte destination), the windows' LDS copies, and the vertex
// phase's 16 (FP64) or 12 (FP32) per vertex plus offsets-free out-degrees.
// After the degree relabelling most edges of a window share their destination
// with other edges of that window (rmat-s24: 0.27 pairs per edge with 16K-vertex
// windows, rmat-s25: 0.31), so an iteration moves about 7-11 bytes per edge.
//
// FP64 order: a partial adds its edges in source order within each lane, then
// the lanes' open runs by a segmented scan (a fixed order); the pull adds
// partials with LDS atomics (ds_add_f64) in no fixed order, so results agree
// with the tile kernel to rounding, not bit for bit, and may differ in the last
// bits from run to run.
//
// Layout (built once per graph and window size by pb_build, device side):
//   code[q]        the edges by (window, destination, source), each window's run
//                  padded to a multiple of WG_PB_TILE entries: bits 14..0 =
//                  source - window start, bit 15 = last edge of its pair (also
//                  forced on a tile's last entry, so no pair crosses tiles)
//   tile_pbase[t]  number of pairs before tile t (WG_PB_TILE entries per tile)
//   task[k]        push work: {window, first tile, end tile, end of the real
//                  entries (code index)}, at most WG_PB_TASK_TILES tiles
//   pdst[p]        pair p's destination - bin start
//   part[p]        pair p's partial sum (written by the push every iteration)
//   piece[i]       {first pair, count}: runs of pairs of one (window, bin), at
//                  most WG_PB_PIECE pairs, grouped by bin in window order
//   slice[j]       pull work: {bin, first piece, end piece, accumulator slot or
//                  WG_PB_NOSLOT when the slice is its bin's only one}; a slice
//                  holds at most WG_PB_SLICE pairs (or one piece)
//   bin_slot[b]    {first slot, slices} of a bin cut into several slices: each
//                  slice writes its accumulators to its slot, and the last one
//                  to finish (a per-bin ticket) adds the slots in slice order
#include "wg_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#define WG_PB_THREADS 1024                      // 16 wavefronts, one workgroup per CU
#define WG_PB_WAVES (WG_PB_THREADS / WG_WAVE)
#define WG_PB_TILE 512u                         // code entries per wavefront tile
#define WG_PB_TASK_TILES 256u                   // tiles per push workgroup (128K edges)
#define WG_PB_PIECE 4096u                       // pairs per piece (at most)
#define WG_PB_SLICE (256u * 1024u)              // pairs per pull workgroup (default)
#define WG_PB_NOSLOT 0xffffffffu
#define WG_PB_END 0x8000u

struct wg_pb {
    int ws_log, wd_log;
    uint32_t S, B;
    uint64_t nq, npairs;
    uint32_t ntiles, ntasks, npieces, nslices, nslots;
    uint16_t* code;
    uint32_t* tile_pbase;
    uint4* task;
    uint16_t* pdst;
    double* part;
    uint2* piece;
    uint4* slice;
    uint2* bin_slot;
    double* slot_acc;
    unsigned int* bin_ticket;
    double* bin_part;
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
    pb_free_dev(pb->pdst);
    pb_free_dev(pb->part);
    pb_free_dev(pb->piece);
    pb_free_dev(pb->slice);
    pb_free_dev(pb->bin_slot);
    pb_free_dev(pb->slot_acc);
    pb_free_dev(pb->bin_ticket);
    pb_free_dev(pb->bin_part);
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
    if (S > 65536u || B > 65536u || ws_log > 15 || wd_log > 16)
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

    // 3. codes, pair numbers, pair destinations and (window, bin) keys
    pb_tmp twin, pqd, pidx, pkey;
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
        const uint32_t gq = wg_grid(ctx, nq, 1024, 8);
        hipLaunchKernelGGL(k_pb_pairs, dim3(gq), dim3(256), 0, st, nq, pb->code, (const uint32_t*)pidx.p,
                           (const uint32_t*)twin.p, (const uint64_t*)wsd.p, (const uint64_t*)pqd.p,
                           (const uint32_t*)perm.p, ctx->dst, wd_log, pb->pdst, (uint32_t*)pkey.p);
        PB_TRY(hipGetLastError());
        (void)hipFree(pidx.p);
        pidx.p = nullptr;
        (void)hipFree(perm.p);
        perm.p = nullptr;

        // 4. blocks: runs of equal (window, bin) among the pairs
        pb_tmp bflag, bst, nsel, bk;
        PB_TRY(hipMalloc(&bflag.p, npairs));
        PB_TRY(hipMalloc(&bst.p, (uint64_t)npairs * 4));
        PB_TRY(hipMalloc(&nsel.p, sizeof(uint64_t)));
        hipLaunchKernelGGL(k_pb_block_flags, dim3(wg_grid(ctx, npairs, 1024, 8)), dim3(256), 0, st,
                           (const uint32_t*)pkey.p, (uint64_t)npairs, (uint8_t*)bflag.p);
        PB_TRY(hipGetLastError());
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

    // 5. pieces grouped by bin (window order kept), slices and slots (host)
    const uint64_t nb = bstart.size();
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
    const uint32_t slice_max = ctx->pb_slice ? ctx->pb_slice : WG_PB_SLICE;
    const uint32_t piece_max = std::min<uint32_t>(WG_PB_PIECE, slice_max);
    std::vector<uint2> piece;
    std::vector<uint4> slice;
    std::vector<uint2> bin_slot(B, make_uint2(0u, 0u));
    uint32_t nslots = 0;
    for (uint32_t b = 0; b < B; ++b) {
        const size_t first_slice = slice.size();
        uint4 cur = make_uint4(b, (uint32_t)piece.size(), (uint32_t)piece.size(), WG_PB_NOSLOT);
        uint64_t in_slice = 0;
        for (uint64_t k = bin_first[b]; k < bin_first[b + 1]; ++k) {
            const uint32_t i = by_bin[k];
            const uint32_t p0 = bstart[i];
            const uint32_t p1 = (i + 1 < nb) ? bstart[i + 1] : npairs;
            for (uint32_t p = p0; p < p1; p += piece_max) {
                const uint32_t n = std::min<uint32_t>(piece_max, p1 - p);
                if (in_slice && in_slice + n > slice_max) {
                    slice.push_back(cur);
                    cur.y = cur.z;
                    in_slice = 0;
                }
                piece.push_back(make_uint2(p, n));
                cur.z = (uint32_t)piece.size();
                in_slice += n;
            }
        }
        slice.push_back(cur); // every bin has a slice, even without pairs: its vertex phase
        const uint32_t nsl = (uint32_t)(slice.size() - first_slice);
        if (nsl > 1) {
            bin_slot[b] = make_uint2(nslots, nsl);
            for (uint32_t j = 0; j < nsl; ++j)
                slice[first_slice + j].w = nslots + j;
            nslots += nsl;
        }
    }
    pb->npieces = (uint32_t)piece.size();
    pb->nslices = (uint32_t)slice.size();
    pb->nslots = nslots;
    if ((rc = pb_upload(ctx, &pb->task, task)) || (rc = pb_upload(ctx, &pb->piece, piece)) ||
        (rc = pb_upload(ctx, &pb->slice, slice)) || (rc = pb_upload(ctx, &pb->bin_slot, bin_slot)))
        return rc;
    PB_TRY(hipMalloc((void**)&pb->slot_acc, ((uint64_t)(nslots ? nslots : 1) << wd_log) * sizeof(double)));
    PB_TRY(hipMalloc((void**)&pb->bin_ticket, B * sizeof(unsigned int)));
    PB_TRY(hipMemsetAsync(pb->bin_ticket, 0, B * sizeof(unsigned int), st));
    PB_TRY(hipMalloc((void**)&pb->bin_part, B * sizeof(double)));
    PB_TRY(hipStreamSynchronize(st));
    return WG_OK;
}

// Window and bin sizes of a binned shape: one 1024-thread workgroup per CU
// holds 128 KiB of window and of accumulators (blocks_per_cu 1), or two hold
// 64 KiB each (blocks_per_cu 2).
static void pb_sizes(int bpc, size_t esz, int* ws_log, int* wd_log)
{
    const int kib_log = bpc == 1 ? 17 : 16;
    *ws_log = kib_log - (esz == 8 ? 3 : 2);
    *wd_log = kib_log - 3;
}

int wg_pb_prepare(wg_ctx* ctx)
{
    int ws_log, wd_log;
    pb_sizes(ctx->pr_bpc, ctx->pr_fp32 ? 4 : 8, &ws_log, &wd_log);
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

template <typename T, int IPT>
struct pb_codes;

template <typename T>
struct pb_codes<T, 8> {
    static __device__ __forceinline__ void load(const uint16_t* p, uint32_t* c)
    {
        const uint4 r = *reinterpret_cast<const uint4*>(p);
        c[0] = r.x & 0xffffu, c[1] = r.x >> 16, c[2] = r.y & 0xffffu, c[3] = r.y >> 16;
        c[4] = r.z & 0xffffu, c[5] = r.z >> 16, c[6] = r.w & 0xffffu, c[7] = r.w >> 16;
    }
};

template <typename T>
struct pb_codes<T, 4> {
    static __device__ __forceinline__ void load(const uint16_t* p, uint32_t* c)
    {
        const uint2 r = *reinterpret_cast<const uint2*>(p);
        c[0] = r.x & 0xffffu, c[1] = r.x >> 16, c[2] = r.y & 0xffffu, c[3] = r.y >> 16;
    }
};

// Push: one workgroup per task (a window's run of tiles), one wavefront per
// tile.  Lane l takes the tile's entries [l * IPT, (l + 1) * IPT): it sums its
// entries in order, and each entry flagged as a pair's last emits that pair's
// partial; the sum still open at the lane's end flows into the next lanes by
// a segmented inclusive scan (a lane with an end starts a new segment).
template <typename T, int IPT, int WS>
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_push(const uint16_t* __restrict__ code, const T* __restrict__ cin,
                                                           const uint4* __restrict__ task, uint32_t ntasks,
                                                           const uint32_t* __restrict__ tile_pbase,
                                                           double* __restrict__ part, uint32_t V)
{
    static_assert(WG_WAVE * IPT == WG_PB_TILE, "a tile is one wavefront's entries");
    __shared__ T s_win[WS];
    const int lane = wg_lane();
    const int wave = threadIdx.x / WG_WAVE;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t k = blockIdx.x; k < ntasks; k += gridDim.x) {
        const uint4 tk = task[k];
        const uint32_t v0 = tk.x * (uint32_t)WS;
        const uint32_t nv = (V - v0 < (uint32_t)WS) ? V - v0 : (uint32_t)WS;
        __syncthreads(); // the previous task's wavefronts are done with s_win
        for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS)
            s_win[i] = cin[v0 + i];
        __syncthreads();
        for (uint32_t t = tk.y + (uint32_t)wave; t < tk.z; t += WG_PB_WAVES) {
            const uint32_t q0 = t * WG_PB_TILE + (uint32_t)lane * IPT;
            uint32_t c[IPT];
            pb_codes<T, IPT>::load(code + q0, c);
            double v[IPT];
            bool e[IPT];
            double run = 0.0;
            bool has_end = false;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const bool real = q0 + (uint32_t)j < tk.w;
                v[j] = real ? (double)s_win[c[j] & 0x7fffu] : 0.0;
                e[j] = real && (c[j] & WG_PB_END);
                run += v[j];
                if (e[j]) {
                    run = 0.0;
                    has_end = true;
                }
            }
            // segmented inclusive scan of the open sums over the lanes
            double sc = run;
            int f = has_end ? 1 : 0;
#pragma unroll
            for (int d = 1; d < WG_WAVE; d <<= 1) {
                const double o = __shfl_up(sc, d, WG_WAVE);
                const int of = __shfl_up(f, d, WG_WAVE);
                if (lane >= d) {
                    if (!f)
                        sc = o + sc;
                    f |= of;
                }
            }
            double carry = __shfl_up(sc, 1, WG_WAVE);
            if (lane == 0)
                carry = 0.0;
            uint32_t pre = 0;
#pragma unroll
            for (int j = 0; j < IPT; ++j)
                pre += (uint32_t)__popcll(wg_ballot(e[j]) & below);
            double* out = part + tile_pbase[t] + pre;
            double r = carry;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                r += v[j];
                if (e[j]) {
                    *out++ = r;
                    r = 0.0;
                }
            }
        }
    }
}

// Pull: one workgroup per slice.  Adds the slice's partials into the bin's
// accumulators in LDS; a bin cut into several slices has each slice write its
// accumulators to its slot, and the last slice to finish adds the slots in
// slice order.  The workgroup holding a bin's complete accumulators then runs
// the vertex phase (MODE 0: contributions and the sum of ranks of vertices
// with out-degree > 0; MODE 1: accumulators and their sum), and the last bin
// to finish reduces the bins' sums into S_out.
template <typename T, int MODE, int WD>
__global__ __launch_bounds__(WG_PB_THREADS) void k_pb_pull(
    const double* __restrict__ part, const uint16_t* __restrict__ pdst, const uint2* __restrict__ piece,
    const uint4* __restrict__ slice, uint32_t nslices, const uint2* __restrict__ bin_slot, double* __restrict__ slot_acc,
    unsigned int* __restrict__ bin_ticket, double* __restrict__ bin_part, uint32_t nbins, T* __restrict__ cout,
    double* __restrict__ acc_out, const uint32_t* __restrict__ od, uint32_t V, const double* __restrict__ S_in,
    double damping, double N, double* __restrict__ S_out, unsigned int* __restrict__ ticket)
{
    __shared__ double s_acc[WD];
    __shared__ double s_red[WG_PB_WAVES];
    __shared__ int s_flag;
    const int lane = wg_lane();
    const int wave = threadIdx.x / WG_WAVE;
    double corr = 0.0, omd_n = 0.0;
    if (MODE == 0) {
        corr = (1.0 - *S_in) / N;
        omd_n = (1.0 - damping) / N;
    }
    for (uint32_t k = blockIdx.x; k < nslices; k += gridDim.x) {
        const uint4 sl = slice[k];
        const uint32_t b = sl.x, v0 = b * (uint32_t)WD;
        const uint32_t nv = (V - v0 < (uint32_t)WD) ? V - v0 : (uint32_t)WD;
        __syncthreads(); // the previous slice is done with s_acc
        for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS)
            s_acc[i] = 0.0;
        __syncthreads();
        for (uint32_t pc = sl.y + (uint32_t)wave; pc < sl.z; pc += WG_PB_WAVES) {
            const uint2 pz = piece[pc];
            for (uint32_t i = (uint32_t)lane; i < pz.y; i += WG_WAVE)
                atomicAdd(&s_acc[pdst[pz.x + i]], __builtin_nontemporal_load(&part[pz.x + i]));
        }
        __syncthreads();
        if (sl.w != WG_PB_NOSLOT) {
            const uint2 bs = bin_slot[b];
            double* mine = slot_acc + ((uint64_t)sl.w * (uint32_t)WD);
            for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS)
                mine[i] = s_acc[i];
            __threadfence();
            __syncthreads();
            if (threadIdx.x == 0)
                s_flag = (atomicAdd(&bin_ticket[b], 1u) == bs.y - 1u) ? 1 : 0;
            __syncthreads();
            if (!s_flag)
                continue;
            __threadfence();
            for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS) {
                double a = 0.0;
                for (uint32_t j = 0; j < bs.y; ++j)
                    a += slot_acc[((uint64_t)(bs.x + j) * (uint32_t)WD) + i];
                s_acc[i] = a;
            }
            if (threadIdx.x == 0)
                bin_ticket[b] = 0u;
        }
        double s_part = 0.0;
        for (uint32_t i = threadIdx.x; i < nv; i += WG_PB_THREADS) {
            const uint32_t v = v0 + i;
            const double accv = s_acc[i];
            if (MODE == 0) {
                const uint32_t odv = od[v];
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
        if (threadIdx.x == 0) {
            bin_part[b] = s_part;
            __threadfence();
            s_flag = (atomicAdd(ticket, 1u) == nbins - 1u) ? 1 : 0;
        }
        __syncthreads();
        if (s_flag) {
            __threadfence();
            double s = 0.0;
            for (uint32_t i = threadIdx.x; i < nbins; i += WG_PB_THREADS)
                s += __builtin_nontemporal_load(&bin_part[i]);
            s = block_sum_pb(s, s_red);
            if (threadIdx.x == 0) {
                *S_out = s;
                *ticket = 0u;
            }
        }
    }
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
        hipLaunchKernelGGL((k_pb_push<T, 8, WS>), dim3(pb->ntasks), dim3(WG_PB_THREADS), 0, ctx->stream, pb->code, cin,
                           pb->task, pb->ntasks, pb->tile_pbase, pb->part, V);
        int rc = wg_check_launch(ctx);
        if (rc)
            return rc;
        wg_prof_end(ctx, WG_K_PB_PUSH, slot);
    }
    const int slot = wg_prof_begin(ctx, WG_K_PB_PULL);
    if (mode == 0)
        hipLaunchKernelGGL((k_pb_pull<T, 0, WD>), dim3(pb->nslices), dim3(WG_PB_THREADS), 0, ctx->stream, pb->part,
                           pb->pdst, pb->piece, pb->slice, pb->nslices, pb->bin_slot, pb->slot_acc, pb->bin_ticket,
                           pb->bin_part, pb->B, cout, nullptr, ctx->od, V, S_in, ctx->damping, N, S_out,
                           &ctx->ctr->ticket);
    else
        hipLaunchKernelGGL((k_pb_pull<T, 1, WD>), dim3(pb->nslices), dim3(WG_PB_THREADS), 0, ctx->stream, pb->part,
                           pb->pdst, pb->piece, pb->slice, pb->nslices, pb->bin_slot, pb->slot_acc, pb->bin_ticket,
                           pb->bin_part, pb->B, nullptr, ctx->acc, ctx->od, V, S_in, ctx->damping, N, S_out,
                           &ctx->ctr->ticket);
    const int rc = wg_check_launch(ctx);
    if (rc)
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
        return big ? pb_launch<float, 32768, 16384>(ctx, mode, (const float*)cin, (float*)cout, S_in, S_out)
                   : pb_launch<float, 16384, 8192>(ctx, mode, (const float*)cin, (float*)cout, S_in, S_out);
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
