// wg_frontier.h -- device helpers shared by the traversal kernels: vertex
// activation with per-wavefront counting, the slotted counters and their fold,
// the walk over marked Wedge-frontier blocks and block marking.
#pragma once

#include "wg_internal.h"

__device__ __forceinline__ bool test_bit64(const uint64_t* __restrict__ bm, uint32_t i)
{
    return (bm[i >> 6] >> (i & 63)) & 1ull;
}

// Per-thread activation tally: activated vertices and their out-edges.
struct act_tally {
    unsigned long long edges;
    unsigned long long verts;
    unsigned long long fresh; // activated vertices reached for the first time
};

// Marks vertex v activated (the reference's cond_set_bit + out-degree count,
// source/phase_edge_pull_cc.c:119-123); counts it if this call set the bit.
__device__ __forceinline__ void activate(uint64_t* __restrict__ next, const uint32_t* __restrict__ od, uint32_t v,
                                         act_tally& t)
{
    const unsigned long long m = 1ull << (v & 63);
    const unsigned long long old = atomicOr((unsigned long long*)&next[v >> 6], m);
    if (!(old & m)) {
        t.edges += od[v];
        t.verts += 1ull;
    }
}

// Lanes hold the consecutive vertices base + lane; the changed ones are set in
// 'next' with at most two 64-bit atomics per wavefront (the range straddles at
// most two bitmask words) and counted when newly set.
__device__ __forceinline__ void activate_contiguous(uint64_t* __restrict__ next, const uint32_t* __restrict__ od,
                                                    uint32_t base, bool changed, act_tally& t)
{
    const unsigned long long cm = wg_ballot(changed);
    if (cm == 0ull)
        return;
    const int lane = wg_lane();
    const uint32_t w0 = base >> 6, sh = base & 63u;
    const unsigned long long lo = cm << sh;                  // bits landing in word w0
    const unsigned long long hi = sh ? (cm >> (64u - sh)) : 0ull; // bits landing in word w0 + 1
    unsigned long long old_lo = 0ull, old_hi = 0ull;
    if (lane == 0 && lo)
        old_lo = atomicOr((unsigned long long*)&next[w0], lo);
    if (lane == 1 && hi)
        old_hi = atomicOr((unsigned long long*)&next[w0 + 1], hi);
    old_lo = __shfl(old_lo, 0);
    old_hi = __shfl(old_hi, 1);
    if (changed) {
        const uint32_t pos = sh + (uint32_t)lane;
        const bool was = (pos < 64u) ? ((old_lo >> pos) & 1ull) : ((old_hi >> (pos - 64u)) & 1ull);
        if (!was) {
            t.edges += od[base + (uint32_t)lane];
            t.verts += 1ull;
        }
    }
}

// Adds a per-lane count to field f of this workgroup's counter slot.
__device__ __forceinline__ void add_count(unsigned long long a, wg_dev_counters* __restrict__ c, int f)
{
    a = wg_wave_sum(a);
    if (wg_lane() == 0 && a)
        atomicAdd(&c->slots[blockIdx.x & (WG_CSLOTS - 1)].v[f], a);
}

// Work of a pull (edges whose source it read, destination states it
// examined), counted per thread in the kernels instantiated with CW = true
// (WG_OPT_COUNT_WORK); with CW = false the counts compile away.
template <bool CW>
struct work_tally {
    unsigned long long edges = 0ull;
    unsigned long long dests = 0ull;
};

struct work_none {
    __device__ __forceinline__ work_none& operator+=(unsigned long long) { return *this; }
};

template <>
struct work_tally<false> {
    work_none edges, dests;
};

template <bool CW>
__device__ __forceinline__ void add_work(const work_tally<CW>& t, wg_dev_counters* __restrict__ c)
{
    if constexpr (CW) {
        add_count(t.edges, c, WG_CNT_EDGES_READ);
        add_count(t.dests, c, WG_CNT_DESTS);
    }
}

__device__ __forceinline__ void add_tally(const act_tally& t, wg_dev_counters* __restrict__ c)
{
    add_count(t.edges, c, WG_CNT_ACTIVITY);
    add_count(t.verts, c, WG_CNT_ACTIVATED);
    add_count(t.fresh, c, WG_CNT_FRESH);
}

__device__ __forceinline__ unsigned long long* count_field(wg_dev_counters* c, int f)
{
    switch (f) {
    case WG_CNT_ACTIVITY:
        return &c->activity;
    case WG_CNT_ACTIVATED:
        return &c->activated;
    case WG_CNT_FRESH:
        return &c->fresh;
    case WG_CNT_FRONTIER:
        return &c->frontier_verts;
    case WG_CNT_REACHED_V:
        return &c->reached_verts;
    case WG_CNT_REACHED_E:
        return &c->reached_edges;
    case WG_CNT_EDGES_READ:
        return &c->edges_read;
    default:
        return &c->dests_examined;
    }
}

// Adds the slots into the scalar counts and zeroes them; every lane gets the
// slot totals in tot (if given).  Called by the first wavefront of a launch
// only; nothing else may be adding to the slots.
__device__ inline void fold_counts(wg_dev_counters* c, unsigned long long* tot = nullptr)
{
    const int lane = wg_lane();
    for (int f = 0; f < WG_CNT_N; ++f) {
        unsigned long long x = 0ull;
        if (lane < WG_CSLOTS) {
            x = c->slots[lane].v[f];
            c->slots[lane].v[f] = 0ull;
        }
        x = wg_wave_sum(x);
        if (tot)
            tot[f] = x;
        if (lane == 0)
            *count_field(c, f) += x;
    }
}



#define RST_ACTIVITY 1u
#define RST_FRONTIER 4u
#define RST_OVERFLOW 8u
#define RST_HEAVY 16u
#define RST_WORK 32u

// Lane -> (marked block, edge) mapping shared by the frontier pulls.
struct block_walk {
    uint32_t per_wave;   // blocks per wave step
    uint32_t lane_bits;  // log2 of lanes per block
    uint32_t steps;      // 64-edge steps per block
};

__device__ __forceinline__ block_walk make_walk(int blog2)
{
    block_walk b;
    b.lane_bits = blog2 < 6 ? (uint32_t)blog2 : 6u;
    b.per_wave = 64u >> b.lane_bits;
    b.steps = blog2 > 6 ? (1u << (blog2 - 6)) : 1u;
    return b;
}

// Words of the Wedge frontier at granularity 2^blog2 for E edges.
__device__ __forceinline__ uint32_t wedge_nwords(uint32_t E, int blog2)
{
    const uint64_t nblk = ((uint64_t)E + (1ull << blog2) - 1ull) >> blog2;
    return (uint32_t)((nblk + 31ull) >> 5);
}

// Item k of a wavefront whose lane l holds cnt_l items (inc: the inclusive
// prefix sum of cnt over the lanes, total: its last value, k < total): the
// owning lane and the rank of k among that lane's items.  Every lane calls
// it (a uniform loop of shuffles); lanes may ask for different k.
__device__ __forceinline__ void wave_item_owner(uint32_t inc, uint32_t cnt, uint32_t k, int* owner, uint32_t* rank)
{
    int lo = 0, hi = WG_WAVE - 1;
#pragma unroll
    for (int it = 0; it < 6; ++it) {
        const int mid = (lo + hi) >> 1;
        if ((uint32_t)__shfl((int)inc, mid) > k)
            hi = mid;
        else
            lo = mid + 1;
    }
    *owner = lo;
    *rank = k - (uint32_t)__shfl((int)(inc - cnt), lo);
}

// Position of the rank-th set bit of w (w has more than rank set bits).
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t w, uint32_t rank)
{
    for (uint32_t i = 0; i < rank; ++i)
        w &= w - 1ull;
    return (uint32_t)(__ffsll((long long)w) - 1);
}

// Inclusive prefix sum of x over the lanes of the wavefront.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    const int lane = wg_lane();
#pragma unroll
    for (int d = 1; d < WG_WAVE; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d);
        if (lane >= d)
            x += y;
    }
    return x;
}

// Walks the marked edge blocks (source/phase_edge_pull_bfs.c:245-268): each
// wavefront takes 64 consecutive frontier words, one per lane, zeroes the
// non-zero ones, and hands their marked blocks, in order, to lane groups of
// 2^min(b,6) lanes, 64 >> min(b,6) blocks per step, whichever words they
// come from (a prefix sum of the words' bit counts), so sparse words do not
// cost a step each.  f(p, ok, first) is called by every lane of the
// wavefront for each 64-edge step: p is the edge position, ok whether the
// lane holds an edge, first whether the lane is the first of its block's
// group.
template <class F>
__device__ __forceinline__ void walk_marked_blocks(uint32_t* __restrict__ wedge, int blog2, uint32_t E, F&& f)
{
    const block_walk bw = make_walk(blog2);
    const uint32_t nwords = wedge_nwords(E, blog2);
    const int lane = wg_lane();
    const uint32_t sub = (uint32_t)lane & ((1u << bw.lane_bits) - 1u);
    const uint32_t slot = (uint32_t)lane >> bw.lane_bits;
    const uint32_t nwaves = gridDim.x * WG_WPB;
    for (uint32_t w0 = (blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE) * WG_WAVE; w0 < nwords; w0 += nwaves * WG_WAVE) {
        const uint32_t wi = w0 + (uint32_t)lane;
        uint32_t word = 0u;
        if (wi < nwords) {
            word = wedge[wi];
            if (word)
                wedge[wi] = 0u; // consumed, source/phase_edge_pull_bfs.c:252-256
        }
        const uint32_t cnt = (uint32_t)__popc(word);
        const uint32_t inc = wave_incl_scan(cnt);
        const uint32_t total = (uint32_t)__shfl((int)inc, WG_WAVE - 1);
        for (uint32_t k0 = 0; k0 < total; k0 += bw.per_wave) {
            const uint32_t k = k0 + slot;
            const bool have = k < total;
            int owner;
            uint32_t rank;
            wave_item_owner(inc, cnt, have ? k : 0u, &owner, &rank);
            const uint32_t ow = (uint32_t)__shfl((int)word, owner);
            const uint64_t blk =
                (uint64_t)(w0 + (uint32_t)owner) * 32u + (have ? (uint64_t)nth_set_bit(ow, rank) : 0ull);
            for (uint32_t st = 0; st < bw.steps; ++st) {
                const uint64_t p64 = (blk << blog2) + (uint64_t)st * 64u + sub;
                f((uint32_t)p64, have && p64 < E, sub == 0u);
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ T atomic_min_t(T* p, T v);
template <>
__device__ __forceinline__ uint32_t atomic_min_t<uint32_t>(uint32_t* p, uint32_t v)
{
    return atomicMin(p, v);
}
template <>
__device__ __forceinline__ unsigned long long atomic_min_t<unsigned long long>(unsigned long long* p, unsigned long long v)
{
    return atomicMin(p, v);
}

// Marks edge block blk (source/wedge.c:22-31).  The result of the atomic is
// not needed, so the lane does not wait for it.
__device__ __forceinline__ void mark_block(uint32_t blk, uint32_t* __restrict__ wedge)
{
    const uint32_t m = 1u << (blk & 31);
    uint32_t* wp = &wedge[blk >> 5];
    if (!(*wp & m))
        atomicOr(wp, m);
}

#define WG_WEDGE_LANE_MAX 32u
#define WG_WEDGE_CHUNK 4096u

// ---- CC / SSSP value policies (the dense, frontier and binned min pulls) ----

// val_from(du, w, p): the value edge p offers its destination when its
// source holds du; val reads du from prop.
struct PolCC {
    typedef uint32_t T;
    static constexpr uint32_t INF = 0xffffffffu;
    static constexpr bool HUB = true; // dense pulls read the first labels from LDS
    static constexpr int DENSE_IPT = WG_IPT; // 512-edge tiles (11.2 ms against 11.6 with 256 on rmat-s24 CC)
    __device__ static T val_from(T du, const void* __restrict__ w, uint32_t p, unsigned int* __restrict__ ovf)
    {
        (void)w;
        (void)p;
        (void)ovf;
        return du;
    }
    __device__ static T val(const T* __restrict__ prop, const void* __restrict__ w, uint32_t u, uint32_t p,
                            unsigned int* __restrict__ ovf)
    {
        return val_from(prop[u], w, p, ovf);
    }
};

struct PolSF { // FP64 distances as bit patterns
    typedef unsigned long long T;
    static constexpr unsigned long long INF = 0x7FF0000000000000ull;
    static constexpr bool HUB = false;
    static constexpr int DENSE_IPT = 4; // 256-edge tiles (rmat-s24 SSSP 37.5 -> 32.1 ms)
    __device__ static T val_from(T du, const void* __restrict__ w, uint32_t p, unsigned int* __restrict__ ovf)
    {
        (void)ovf;
        if (du >= INF)
            return INF;
        const double s = __longlong_as_double((long long)du) + ((const double*)w)[p];
        return (T)__double_as_longlong(s);
    }
    __device__ static T val(const T* __restrict__ prop, const void* __restrict__ w, uint32_t u, uint32_t p,
                            unsigned int* __restrict__ ovf)
    {
        return val_from(prop[u], w, p, ovf);
    }
};

struct PolSU { // 32-bit integer distances
    typedef uint32_t T;
    static constexpr uint32_t INF = 0xffffffffu;
    static constexpr bool HUB = false;
    static constexpr int DENSE_IPT = 4;
    __device__ static T val_from(T du, const void* __restrict__ w, uint32_t p, unsigned int* __restrict__ ovf)
    {
        if (du == INF)
            return INF;
        const uint32_t s = du + ((const uint32_t*)w)[p];
        if (s < du || s == INF) {
            *ovf = 1u;
            return INF;
        }
        return s;
    }
    __device__ static T val(const T* __restrict__ prop, const void* __restrict__ w, uint32_t u, uint32_t p,
                            unsigned int* __restrict__ ovf)
    {
        return val_from(prop[u], w, p, ovf);
    }
};

