// wg_traverse.hip -- BFS, CC and SSSP edge pulls, the Wedge frontier, and the
// end-of-iteration frontier handling on gfx950.
//
// Reference phases: source/phase_edge_pull_bfs.c:109-332,
// source/phase_edge_pull_cc.c:53-297, source/phase_edge_pull_sssp.c:57-306,
// source/wedge.c:37-155, the idempotent merge source/phases.c:54-65, and the
// frontier swap/clear in source/execution_{bfs,cc,sssp}.c.
//
// Vertex state on the device:
//   BFS   packed[v] = (depth << 32) | parent, UINT64_MAX while unvisited.  The
//         parent is the smallest in-neighbour on the previous level in both
//         pull modes (dense: first hit in ascending source order; frontier:
//         atomicMin of the packed word), so the result does not depend on the
//         mode sequence.  The reference keeps only a parent (its largest
//         active source within the first matching edge vector).
//   CC    label[v] uint32, min-label propagation in place.
//   SSSP  FP64 distances as IEEE bit patterns in uint64 (non-negative doubles
//         order like their bit patterns, so atomicMin on the bits is the FP64
//         minimum) or uint32 integer distances.
// Bitmasks (one bit per vertex, 64-bit words): cur (BFS frontier), next
// (activated by the running pull: the reference's active bitmask with its
// cond_set_bit semantics), visited (BFS, visited before the running level).
// The Wedge frontier is one bit per edge block of 2^b consecutive CSC edges
// (32-bit words).  A frontier pull scans the words, consumes (zeroes) the
// non-zero ones and spreads their marked blocks over the lanes of a wavefront,
// so the edges it reads follow the marked blocks; the word scan itself is
// E / 2^b / 8 bytes.
#include "wg_frontier.h"

#include <algorithm>
#include <vector>
#include "wg_mlp_wave.h"

// ---- counters -------------------------------------------------------------------

__global__ void k_fold(wg_dev_counters* c)
{
    if (threadIdx.x < WG_WAVE)
        fold_counts(c);
}


__global__ void k_reset(wg_dev_counters* c, unsigned int flags)
{
    const unsigned int t = threadIdx.x;
    if (blockIdx.x == 0 && t < WG_CSLOTS) {
        if (flags & RST_ACTIVITY) {
            c->slots[t].v[WG_CNT_ACTIVITY] = 0ull;
            c->slots[t].v[WG_CNT_ACTIVATED] = 0ull;
            c->slots[t].v[WG_CNT_FRESH] = 0ull;
        }
        if (flags & RST_FRONTIER)
            c->slots[t].v[WG_CNT_FRONTIER] = 0ull;
        if (flags & RST_WORK) {
            c->slots[t].v[WG_CNT_EDGES_READ] = 0ull;
            c->slots[t].v[WG_CNT_DESTS] = 0ull;
        }
    }
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (flags & RST_ACTIVITY) {
            c->activity = 0ull;
            c->activated = 0ull;
            c->fresh = 0ull;
        }
        if (flags & RST_FRONTIER)
            c->frontier_verts = 0ull;
        if (flags & RST_OVERFLOW)
            c->overflow = 0u;
        if (flags & RST_HEAVY)
            c->heavy_len = 0u;
        if (flags & RST_WORK) {
            c->edges_read = 0ull;
            c->dests_examined = 0ull;
        }
    }
}

int wg_reset_counters(wg_ctx* ctx, unsigned int flags)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->ctr, flags);
    return wg_check_launch(ctx);
}

// ---- BFS -------------------------------------------------------------------------

// The first WG_BFS_CUR_LDS words of the frontier bitmask, copied to LDS by a
// dense level: in-lists ascend in source number, and after the degree
// relabelling the first vertices are the hubs, so most in-list scans test
// their first sources here (rmat-s24: the first in-neighbour of a vertex of
// in-degree 8 is among the first 32K vertices 97% of the time).
#define WG_BFS_CUR_LDS 512u
__device__ __forceinline__ bool test_cur(const uint64_t* __restrict__ cur, const uint64_t* s_cur, uint32_t u)
{
    if (u < WG_BFS_CUR_LDS * 64u)
        return (s_cur[u >> 6] >> (u & 63)) & 1ull;
    return test_bit64(cur, u);
}

__device__ __forceinline__ void stage_cur(const uint64_t* __restrict__ cur, uint32_t V, uint64_t* s_cur)
{
    const uint32_t nw = (V + 63u) / 64u;
    for (uint32_t i = threadIdx.x; i < WG_BFS_CUR_LDS; i += blockDim.x)
        s_cur[i] = i < nw ? cur[i] : 0ull;
    __syncthreads();
}

// first[v]: the smallest in-neighbour of v (the first of its ascending
// in-list), UINT32_MAX if v has no in-edges or more than WG_BFS_HEAVY.
__global__ void k_first_src(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src, uint64_t V,
                            uint32_t* __restrict__ first)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = off[v], d = off[v + 1] - s;
        first[v] = (d == 0u || d > WG_BFS_HEAVY) ? 0xffffffffu : src[s];
    }
}

// csr_dst[p] = dst[csri_pos[p]]: the out-edges' destinations in out-edge
// order, so that walking a vertex's out-edges takes one load per edge.
__global__ void k_csr_dst(const uint32_t* __restrict__ csri_pos, const uint32_t* __restrict__ dst, uint64_t E,
                          uint32_t* __restrict__ out)
{
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < E; p += (uint64_t)gridDim.x * blockDim.x)
        out[p] = dst[csri_pos[p]];
}

// Dense bottom-up level: one wave per 64 consecutive vertices.  A lane first
// tests its vertex's smallest in-neighbour (first[], read coalesced): after
// the degree relabelling that is usually a hub, and most vertices of a dense
// level stop there without reading their in-list.  Otherwise the lane scans
// the rest of up to WG_BFS_SEQ in-edges; vertices with more edges are
// finished by the whole wave, 64 edges per step, stopping at the first hit.
// Vertices with more than WG_BFS_HEAVY in-edges are left to k_bfs_heavy.
// s_cur: the frontier's first words in LDS (stage_cur).
template <bool CW>
__device__ void bfs_dense_body(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                               const uint32_t* __restrict__ first, const uint32_t* __restrict__ od,
                               const uint64_t* __restrict__ cur, const uint64_t* __restrict__ visited,
                               uint64_t* __restrict__ next, uint64_t* __restrict__ packed, uint32_t V,
                               const uint32_t* __restrict__ perm, wg_dev_counters* __restrict__ ctr,
                               const uint64_t* s_cur)
{
    const uint32_t newdepth = ctr->level + 1u;
    const int lane = wg_lane();
    const uint64_t nwords = ((uint64_t)V + 63) / 64;
    const uint64_t nwaves = (uint64_t)gridDim.x * (WG_BLOCK / WG_WAVE);
    act_tally act = { 0ull, 0ull, 0ull };
    work_tally<CW> work;
    for (uint64_t w = (uint64_t)blockIdx.x * (WG_BLOCK / WG_WAVE) + threadIdx.x / WG_WAVE; w < nwords; w += nwaves) {
        const uint64_t vis = visited[w];
        if (vis == ~0ull)
            continue;
        const uint32_t v = (uint32_t)(w * 64 + lane);
        bool todo = (v < V) && !((vis >> lane) & 1ull);
        uint32_t f = 0xffffffffu;
        if (todo)
            f = first[v];
        todo = todo && f != 0xffffffffu; // has 1..WG_BFS_HEAVY in-edges
        if (wg_ballot(todo) == 0ull)
            continue;
        int64_t parent = -1;
        if (todo) {
            work.dests += 1ull;
            work.edges += 1ull; // the smallest in-neighbour, from first[]
            if (test_cur(cur, s_cur, f))
                parent = f;
        }
        // the rest of the in-list, if the smallest in-neighbour is not on the frontier
        const bool more = todo && parent < 0;
        uint32_t s = 0, e = 0;
        if (more) {
            s = off[v];
            e = off[v + 1];
            const uint32_t lim = (e - s < WG_BFS_SEQ) ? e : s + WG_BFS_SEQ;
            for (uint32_t p = s + 1; p < lim; ++p) {
                const uint32_t u = src[p];
                work.edges += 1ull;
                if (test_cur(cur, s_cur, u)) {
                    parent = u;
                    break;
                }
            }
        }
        const uint32_t deg = e - s;
        unsigned long long need = wg_ballot(more && parent < 0 && deg > WG_BFS_SEQ);
        while (need) {
            const int l = __ffsll((long long)need) - 1;
            need &= need - 1ull;
            const uint32_t ls = __shfl(s, l) + WG_BFS_SEQ;
            const uint32_t le = __shfl(e, l);
            int64_t found = -1;
            for (uint32_t p = ls; p < le && found < 0; p += WG_WAVE) {
                const uint32_t q = p + lane;
                uint32_t u = 0;
                bool hit = false;
                if (q < le) {
                    u = src[q];
                    work.edges += 1ull;
                    hit = test_cur(cur, s_cur, u);
                }
                const unsigned long long hm = wg_ballot(hit);
                if (hm) {
                    const int fl = __ffsll((long long)hm) - 1;
                    found = (int64_t)__shfl(u, fl);
                }
            }
            if (lane == l && found >= 0)
                parent = found;
        }
        const bool disc = parent >= 0;
        const unsigned long long dm = wg_ballot(disc);
        if (disc) {
            const uint32_t pid = perm ? perm[(uint32_t)parent] : (uint32_t)parent;
            packed[v] = ((uint64_t)newdepth << 32) | (uint64_t)pid;
            act.edges += od[v];
            act.verts += 1ull;
        }
        if (lane == 0 && dm)
            atomicOr((unsigned long long*)&next[w], dm);
    }
    act.fresh = act.verts; // a BFS activation is always a first visit
    add_tally(act, ctr);
    add_work(work, ctr);
}

template <bool CW>
__global__ __launch_bounds__(WG_BLOCK) void k_bfs_dense(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                         const uint32_t* __restrict__ first,
                                                         const uint32_t* __restrict__ od, const uint64_t* __restrict__ cur,
                                                         const uint64_t* __restrict__ visited, uint64_t* __restrict__ next,
                                                         uint64_t* __restrict__ packed, uint32_t V,
                                                         const uint32_t* __restrict__ perm, wg_dev_counters* __restrict__ ctr,
                                                         int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    __shared__ uint64_t s_cur[WG_BFS_CUR_LDS];
    stage_cur(cur, V, s_cur);
    bfs_dense_body<CW>(off, src, first, od, cur, visited, next, packed, V, perm, ctr, s_cur);
}

// Dense level for chunks of WG_BFS_HEAVY in-edges of high in-degree vertices,
// one wave per chunk.
template <bool CW>
__device__ void bfs_heavy_body(const uint32_t* __restrict__ hv, const uint32_t* __restrict__ he, uint32_t n,
                               const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                               const uint32_t* __restrict__ od, const uint64_t* __restrict__ cur,
                               const uint64_t* __restrict__ visited, uint64_t* __restrict__ next,
                               uint64_t* __restrict__ packed, const uint32_t* __restrict__ perm,
                               wg_dev_counters* __restrict__ ctr)
{
    const uint32_t newdepth = ctr->level + 1u;
    const int lane = wg_lane();
    const uint32_t nwaves = gridDim.x * (WG_BLOCK / WG_WAVE);
    act_tally act = { 0ull, 0ull, 0ull };
    work_tally<CW> work;
    for (uint32_t c = blockIdx.x * (WG_BLOCK / WG_WAVE) + threadIdx.x / WG_WAVE; c < n; c += nwaves) {
        const uint32_t v = hv[c];
        if (test_bit64(visited, v))
            continue;
        if (lane == 0)
            work.dests += 1ull;
        const uint32_t cs = he[c];
        const uint32_t vend = off[v + 1];
        const uint32_t ce = (cs + WG_BFS_HEAVY < vend) ? cs + WG_BFS_HEAVY : vend;
        // a lower chunk already found a parent smaller than anything here
        // (parents are original ids; in-lists ascend in original id)
        const uint64_t pk = packed[v];
        const uint32_t first_id = perm ? perm[src[cs]] : src[cs];
        if ((uint32_t)(pk >> 32) == newdepth && (uint32_t)pk < first_id)
            continue;
        int64_t found = -1;
        for (uint32_t p = cs; p < ce && found < 0; p += WG_WAVE) {
            const uint32_t q = p + lane;
            uint32_t u = 0;
            bool hit = false;
            if (q < ce) {
                u = src[q];
                work.edges += 1ull;
                hit = test_bit64(cur, u);
            }
            const unsigned long long hm = wg_ballot(hit);
            if (hm) {
                const int fl = __ffsll((long long)hm) - 1;
                found = (int64_t)__shfl(u, fl);
            }
        }
        if (found >= 0 && lane == 0) {
            const uint32_t pid = perm ? perm[(uint32_t)found] : (uint32_t)found;
            atomicMin((unsigned long long*)&packed[v], ((uint64_t)newdepth << 32) | (uint64_t)pid);
            activate(next, od, v, act);
        }
    }
    act.fresh = act.verts; // a BFS activation is always a first visit
    add_tally(act, ctr);
    add_work(work, ctr);
}

template <bool CW>
__global__ __launch_bounds__(WG_BLOCK) void k_bfs_heavy(const uint32_t* __restrict__ hv, const uint32_t* __restrict__ he,
                                                         uint32_t n, const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ src, const uint32_t* __restrict__ od,
                                                         const uint64_t* __restrict__ cur, const uint64_t* __restrict__ visited,
                                                         uint64_t* __restrict__ next, uint64_t* __restrict__ packed,
                                                         const uint32_t* __restrict__ perm,
                                                         wg_dev_counters* __restrict__ ctr, int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    bfs_heavy_body<CW>(hv, he, n, off, src, od, cur, visited, next, packed, perm, ctr);
}

// Frontier-mode BFS level over the marked edge blocks.
template <bool CW>
__device__ void bfs_blocks_body(wg_dev_counters* __restrict__ ctr, int blog2, uint32_t E,
                                const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                const uint32_t* __restrict__ od, const uint64_t* __restrict__ cur,
                                const uint64_t* __restrict__ visited, uint64_t* __restrict__ next,
                                uint64_t* __restrict__ packed, uint32_t* __restrict__ wedge,
                                const uint32_t* __restrict__ perm)
{
    const uint32_t newdepth = ctr->level + 1u;
    const int lane = wg_lane();
    act_tally act = { 0ull, 0ull, 0ull };
    work_tally<CW> work;
    walk_marked_blocks(wedge, blog2, E, [&](uint32_t p, bool ok, bool first) {
        bool cand = false;
        uint32_t u = 0, v = 0;
        if (ok) {
            u = src[p];
            v = dst[p];
            cand = test_bit64(cur, u) && !test_bit64(visited, v);
        }
        // only the first candidate of a run of equal destinations within one
        // block needs the atomic: sources ascend within a destination (lanes
        // of different blocks are not compared: their blocks need not be
        // adjacent)
        const uint32_t pv = __shfl_up(v, 1);
        const unsigned long long cm = wg_ballot(cand);
        const bool prev_same = !first && ((cm >> (lane - 1)) & 1ull) && pv == v;
        const unsigned long long okm = wg_ballot(ok);
        if (ok) { // a destination's state is examined once per run of its edges in a block
            work.edges += 1ull;
            if (first || !((okm >> (lane - 1)) & 1ull) || pv != v)
                work.dests += 1ull;
        }
        if (cand && !prev_same) {
            const uint32_t pid = perm ? perm[u] : u;
            atomicMin((unsigned long long*)&packed[v], ((uint64_t)newdepth << 32) | (uint64_t)pid);
            activate(next, od, v, act);
        }
    });
    act.fresh = act.verts; // a BFS activation is always a first visit
    add_tally(act, ctr);
    add_work(work, ctr);
}

// The first frontier-filtered level of a traversal whose frontier is still its
// roots.  The pull over the blocks marked by the roots' out-edges
// (source/phase_edge_pull_bfs.c over source/wedge.c:63-155) tests exactly the
// pairs (root r, out-neighbour v) with v unvisited, and gives v depth 1 and the
// smallest such r as parent.  Going along the roots' out-edges (their
// destinations in out-edge order, csr_dst) reaches the same pairs and keeps the
// same minimum, without marking blocks or reading the other edges of each
// block.  The out-edges of all roots are numbered one after the other
// (root_off: the prefix of the roots' out-degrees) and spread over the grid;
// each thread finds the root of its edge by binary search.
template <bool CW>
__device__ void bfs_push_roots_body(const uint32_t* __restrict__ roots, uint32_t nroots,
                                    const uint32_t* __restrict__ root_off,
                                    const uint32_t* __restrict__ csri_off, const uint32_t* __restrict__ csr_dst,
                                    const uint32_t* __restrict__ od,
                                    const uint64_t* __restrict__ visited, uint64_t* __restrict__ next,
                                    uint64_t* __restrict__ packed, const uint32_t* __restrict__ perm,
                                    wg_dev_counters* __restrict__ ctr)
{
    const uint64_t newdepth = ctr->level + 1u;
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = root_off[nroots];
    act_tally act = { 0ull, 0ull, 0ull };
    work_tally<CW> work;
    for (uint64_t g = tid; g < total; g += nthreads) {
        work.edges += 1ull; // one out-edge of a root and its destination's state
        work.dests += 1ull;
        // the last root whose out-edges start at or before g (roots without
        // out-edges share their successor's start and are passed over)
        uint32_t lo = 0, hi = nroots - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (root_off[mid] <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint32_t r = roots[lo];
        const uint64_t val = (newdepth << 32) | (uint64_t)(perm ? perm[r] : r);
        const uint32_t v = csr_dst[csri_off[r] + (uint32_t)(g - root_off[lo])];
        if (!test_bit64(visited, v)) {
            // packed[v] of an unvisited vertex holds ~0 until the first
            // write of this level: that writer alone sets and counts v
            const unsigned long long old = atomicMin((unsigned long long*)&packed[v], (unsigned long long)val);
            if (old == ~0ull) {
                atomicOr((unsigned long long*)&next[v >> 6], 1ull << (v & 63));
                act.edges += od[v];
                act.verts += 1ull;
            }
        }
    }
    act.fresh = act.verts; // a BFS activation is always a first visit
    add_tally(act, ctr);
    add_work(work, ctr);
}

// root_off[i] = out-degree sum of roots[0..i), root_off[n] = the total; one
// workgroup of WG_BLOCK threads, each summing a contiguous run of roots.
__global__ __launch_bounds__(WG_BLOCK) void k_root_offsets(const uint32_t* __restrict__ roots, uint32_t n,
                                                            const uint32_t* __restrict__ csri_off,
                                                            uint32_t* __restrict__ root_off)
{
    __shared__ uint32_t s_sum[WG_BLOCK];
    const uint32_t per = (n + WG_BLOCK - 1) / WG_BLOCK;
    const uint32_t b = threadIdx.x * per < n ? threadIdx.x * per : n;
    const uint32_t e = b + per < n ? b + per : n;
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i)
        sum += csri_off[roots[i] + 1] - csri_off[roots[i]];
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int t = 0; t < WG_BLOCK; ++t) {
            const uint32_t x = s_sum[t];
            s_sum[t] = acc;
            acc += x;
        }
        root_off[n] = acc;
    }
    __syncthreads();
    uint32_t acc = s_sum[threadIdx.x];
    for (uint32_t i = b; i < e; ++i) {
        root_off[i] = acc;
        acc += csri_off[roots[i] + 1] - csri_off[roots[i]];
    }
}

template <bool CW>
__global__ __launch_bounds__(WG_BLOCK) void k_bfs_blocks(wg_dev_counters* __restrict__ ctr, int blog2_arg, uint32_t E,
                                                          const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                                          const uint32_t* __restrict__ od, const uint64_t* __restrict__ cur,
                                                          const uint64_t* __restrict__ visited, uint64_t* __restrict__ next,
                                                          uint64_t* __restrict__ packed, uint32_t* __restrict__ wedge,
                                                          const uint32_t* __restrict__ perm, int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    bfs_blocks_body<CW>(ctr, gate < 0 ? blog2_arg : ctr->blog2, E, src, dst, od, cur, visited, next, packed, wedge, perm);
}

// End of a BFS level: visited |= next, the new frontier replaces the old one
// (copied, so kernels need no pointer swap), next is cleared, the new
// frontier is counted and the level advances.
__device__ void bfs_advance_body(uint64_t* __restrict__ visited, uint64_t* __restrict__ next, uint64_t* __restrict__ cur,
                                 uint64_t nwords, wg_dev_counters* __restrict__ ctr)
{
    unsigned long long cnt = 0;
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwords; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t n = next[w];
        const uint64_t c = cur[w];
        if (n) {
            visited[w] |= n;
            cnt += (unsigned long long)__popcll(n);
            next[w] = 0ull;
        }
        if (c != n)
            cur[w] = n;
    }
    add_count(cnt, ctr, WG_CNT_FRONTIER);
    // the level counter moves once all blocks have read it: the pulls of the
    // next level run after this kernel in stream order
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(&ctr->level, 1u);
}

__global__ void k_bfs_advance(uint64_t* __restrict__ visited, uint64_t* __restrict__ next, uint64_t* __restrict__ cur,
                              uint64_t nwords, wg_dev_counters* __restrict__ ctr, int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    bfs_advance_body(visited, next, cur, nwords, ctr);
}

__device__ void count_clear_body(uint64_t* __restrict__ bm, uint64_t nwords, wg_dev_counters* __restrict__ ctr)
{
    unsigned long long cnt = 0;
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwords; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = bm[w];
        if (b) {
            cnt += (unsigned long long)__popcll(b);
            bm[w] = 0ull;
        }
    }
    add_count(cnt, ctr, WG_CNT_FRONTIER);
}

__global__ void k_count_clear(uint64_t* __restrict__ bm, uint64_t nwords, wg_dev_counters* __restrict__ ctr, int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    count_clear_body(bm, nwords, ctr);
}

// ---- CC / SSSP (min propagation) -------------------------------------------------

// Values of the first vertices in LDS for the dense pulls (policies with HUB):
// after the degree relabelling they are the sources of a large share of the
// edges.  The copy is taken when the pull starts; a value that drops during the
// pull activates its vertex, whose out-edges the next pull reads again, so a
// copy that is older than prop changes when values settle, not what they
// settle to.  CC labels: 5 iterations either way on the symmetrised
// rmat-s24, 12.0 -> 11.2 ms.  SSSP distances on rmat-s24 took 23 iterations
// instead of 20 with the copy (37.4 -> 44.5 ms), and still 23 with only the
// copy's workgroup barrier left in (how far a pull propagates in place depends
// on how its waves interleave), so SSSP keeps the plain path, without either.
#define WG_MIN_HUB_BYTES 16384u
template <class P>
__device__ __forceinline__ void stage_hub(const typename P::T* __restrict__ prop, uint32_t nhub, typename P::T* s_hub)
{
    if constexpr (P::HUB) {
        for (uint32_t i = threadIdx.x; i < nhub; i += blockDim.x)
            s_hub[i] = prop[i];
        __syncthreads();
    }
}

template <class P>
__device__ __forceinline__ void min_apply(typename P::T* __restrict__ prop, uint64_t* __restrict__ next,
                                          const uint32_t* __restrict__ od, uint32_t v, typename P::T m, act_tally& t)
{
    if (m >= prop[v])
        return;
    const typename P::T old = atomic_min_t<typename P::T>(&prop[v], m);
    if (old > m) {
        activate(next, od, v, t);
        if (old == P::INF)
            t.fresh += 1ull;
    }
}

// Dense pull over the edge tiles, one wavefront per tile: per-destination
// minimum in LDS, then one global update per changed destination.  Minimum is
// order-independent, so pieces of a destination spread over lanes or tiles
// meet through atomics (the reference's idempotent merge,
// source/phases.c:54-65).
template <bool CW, class P, int IPT = WG_IPT>
__device__ void min_tile_body(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                              const void* __restrict__ w, typename P::T* __restrict__ prop,
                              const uint32_t* __restrict__ od, const uint32_t* __restrict__ tile_v,
                              const uint32_t* __restrict__ tile_e, uint32_t ntiles, uint64_t* __restrict__ next,
                              wg_dev_counters* __restrict__ ctr, unsigned int* __restrict__ ovf,
                              const typename P::T* s_hub, uint32_t nhub)
{
    typedef typename P::T T;
    constexpr int NOFF = (WG_TILE_VERTS + 1 + WG_WAVE - 1) / WG_WAVE;
    constexpr int NVTX = WG_TILE_VERTS / WG_WAVE;
    __shared__ uint32_t s_off_all[WG_WPB][WG_TILE_VERTS + 1];
    __shared__ T s_min_all[WG_WPB][WG_TILE_VERTS];
    const int wave = threadIdx.x / WG_WAVE;
    const int lane = wg_lane();
    uint32_t* s_off = s_off_all[wave];
    T* s_min = s_min_all[wave];
    act_tally act = { 0ull, 0ull, 0ull };
    work_tally<CW> work;
    const uint32_t stride = gridDim.x * WG_WPB;
    for (uint32_t t = blockIdx.x * WG_WPB + wave; t < ntiles; t += stride) {
        const uint32_t tv = tile_v[t];
        const uint32_t e0 = tile_e[t], e1 = tile_e[t + 1];
        const uint32_t v0 = tv & ~WG_HEAVY_BIT;
        const uint32_t a = e0 + (uint32_t)lane * IPT;
        work.edges += (a < e1) ? ((e1 - a < (uint32_t)IPT) ? e1 - a : (uint32_t)IPT) : 0u;
        T g[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t e = a + j;
            if constexpr (P::HUB) {
                g[j] = P::INF;
                if (e < e1) {
                    const uint32_t u = __builtin_nontemporal_load(&src[e]);
                    g[j] = P::val_from(u < nhub ? s_hub[u] : prop[u], w, e, ovf);
                }
            } else {
                g[j] = (e < e1) ? P::val(prop, w, __builtin_nontemporal_load(&src[e]), e, ovf) : P::INF;
            }
        }
        if (tv & WG_HEAVY_BIT) {
            T m = P::INF;
#pragma unroll
            for (int j = 0; j < IPT; ++j)
                m = g[j] < m ? g[j] : m;
            m = wg_wave_min(m);
            if (lane == 0) {
                work.dests += 1ull;
                if (m != P::INF)
                    min_apply<P>(prop, next, od, v0, m, act);
            }
            continue;
        }
        const uint32_t nv = (tile_v[t + 1] & ~WG_HEAVY_BIT) - v0;
        if (lane == 0)
            work.dests += nv;
#pragma unroll
        for (int k = 0; k < NOFF; ++k) {
            const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
            if (i <= nv)
                s_off[i] = off[v0 + i];
            if (i < nv)
                s_min[i] = P::INF;
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
            uint32_t d = lo;
            uint32_t dend = s_off[d + 1];
            T run = P::INF;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t e = a + j;
                if (e < e1) {
                    run = g[j] < run ? g[j] : run;
                    if (e + 1 == dend) {
                        if (run != P::INF)
                            atomic_min_t<T>(&s_min[d], run);
                        run = P::INF;
                        ++d;
                        while (d < nv && s_off[d + 1] == s_off[d])
                            ++d;
                        dend = (d < nv) ? s_off[d + 1] : 0xffffffffu;
                    }
                }
            }
            if (run != P::INF && d < nv)
                atomic_min_t<T>(&s_min[d], run);
        }
        wg_wave_sync();
#pragma unroll
        for (int k = 0; k < NVTX; ++k) {
            const uint32_t i = (uint32_t)lane + (uint32_t)k * WG_WAVE;
            bool changed = false;
            const uint32_t v = v0 + i;
            if (i < nv) {
                const T m = s_min[i];
                if (m != P::INF && m < prop[v]) {
                    const T old = atomic_min_t<T>(&prop[v], m);
                    changed = old > m;
                    if (changed && old == P::INF)
                        act.fresh += 1ull;
                }
            }
            activate_contiguous(next, od, v0 + (uint32_t)k * WG_WAVE, changed, act);
        }
        wg_wave_sync();
    }
    add_tally(act, ctr);
    add_work(work, ctr);
}

template <bool CW, class P, int IPT = WG_IPT>
__global__ __launch_bounds__(WG_BLOCK) void k_min_tile(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                        const void* __restrict__ w, typename P::T* __restrict__ prop,
                                                        const uint32_t* __restrict__ od, const uint32_t* __restrict__ tile_v,
                                                        const uint32_t* __restrict__ tile_e, uint32_t ntiles,
                                                        uint64_t* __restrict__ next, wg_dev_counters* __restrict__ ctr,
                                                        unsigned int* __restrict__ ovf, int gate, uint32_t nhub)
{
    if (wg_gated_off(ctr, gate))
        return;
    __shared__ typename P::T s_hub[P::HUB ? WG_MIN_HUB_BYTES / sizeof(typename P::T) : 1];
    stage_hub<P>(prop, nhub, s_hub);
    min_tile_body<CW, P, IPT>(off, src, w, prop, od, tile_v, tile_e, ntiles, next, ctr, ovf, s_hub, nhub);
}

// Frontier-mode pull over the marked edge blocks.
template <bool CW, class P>
__device__ void min_blocks_body(wg_dev_counters* __restrict__ ctr, int blog2, uint32_t E,
                                const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                const void* __restrict__ w, typename P::T* __restrict__ prop,
                                const uint32_t* __restrict__ od, uint64_t* __restrict__ next,
                                uint32_t* __restrict__ wedge, unsigned int* __restrict__ ovf)
{
    typedef typename P::T T;
    act_tally act = { 0ull, 0ull, 0ull };
    work_tally<CW> work;
    // one step behind the walk: a step's source and destination loads are
    // issued before the previous step's chain (its source's value, the
    // destination's value, the atomics) runs, so the two overlap; the
    // wavefront still relaxes its steps in walk order, each from the source
    // values the steps before it left (reading them a step earlier as well
    // re-activated more vertices: road-grid SSSP read 1.7 times the edges)
    bool pok = false;
    uint32_t pp = 0, pu = 0, pv = 0;
    auto relax = [&](bool ok, uint32_t p, uint32_t u, uint32_t v) {
        if (ok) {
            work.edges += 1ull;
            const T m = P::val(prop, w, u, p, ovf);
            if (m != P::INF) {
                work.dests += 1ull;
                min_apply<P>(prop, next, od, v, m, act);
            }
        }
    };
    walk_marked_blocks(wedge, blog2, E, [&](uint32_t p, bool ok, bool) {
        const uint32_t u = ok ? src[p] : 0u, v = ok ? dst[p] : 0u;
        relax(pok, pp, pu, pv);
        pok = ok;
        pp = p;
        pu = u;
        pv = v;
    });
    relax(pok, pp, pu, pv);
    add_tally(act, ctr);
    add_work(work, ctr);
}

template <bool CW, class P>
__global__ __launch_bounds__(WG_BLOCK) void k_min_blocks(wg_dev_counters* __restrict__ ctr, int blog2_arg, uint32_t E,
                                                          const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                                          const void* __restrict__ w, typename P::T* __restrict__ prop,
                                                          const uint32_t* __restrict__ od, uint64_t* __restrict__ next,
                                                          uint32_t* __restrict__ wedge, unsigned int* __restrict__ ovf,
                                                          int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    min_blocks_body<CW, P>(ctr, gate < 0 ? blog2_arg : ctr->blog2, E, src, dst, w, prop, od, next, wedge, ovf);
}

// ---- Wedge frontier ---------------------------------------------------------------


// source/wedge.c:63-155: every set vertex marks the blocks of its out-edges.
// Each lane loads one bitmask word; the set vertices of the wavefront's 64
// words are then handed out 64 at a time, one per lane, whichever words they
// come from (a prefix sum of the words' bit counts), so a sparse bitmask does
// not cost a step per non-zero word.  Vertices with more than
// WG_WEDGE_LANE_MAX out-edges are deferred to k_wedge_heavy as chunks of
// WG_WEDGE_CHUNK out-edges (pairs vertex, first position), so that a hub's
// list is spread over many wavefronts.  At most E/32 + E/WG_WEDGE_CHUNK + 1
// chunks are ever listed; heavy_cap bounds the writes regardless.
__global__ __launch_bounds__(WG_BLOCK) void k_wedge_build(const uint64_t* __restrict__ active, uint64_t nwords,
                                                           const uint32_t* __restrict__ csri_off,
                                                           const uint32_t* __restrict__ csri_pos, int blog2_arg,
                                                           uint32_t* __restrict__ wedge, uint32_t* __restrict__ heavy,
                                                           uint32_t heavy_cap, wg_dev_counters* __restrict__ ctr, int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    const int blog2 = gate < 0 ? blog2_arg : ctr->blog2;
    const int lane = wg_lane();
    const uint64_t nwaves = (uint64_t)gridDim.x * WG_WPB;
    for (uint64_t w0 = ((uint64_t)blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE) * WG_WAVE; w0 < nwords;
         w0 += nwaves * WG_WAVE) {
        const uint64_t wi = w0 + (uint64_t)lane;
        const uint64_t word = (wi < nwords) ? active[wi] : 0ull;
        const uint32_t cnt = (uint32_t)__popcll(word);
        const uint32_t inc = wave_incl_scan(cnt);
        const uint32_t total = (uint32_t)__shfl((int)inc, WG_WAVE - 1);
        for (uint32_t k0 = 0; k0 < total; k0 += WG_WAVE) {
            const uint32_t k = k0 + (uint32_t)lane;
            int owner;
            uint32_t rank;
            wave_item_owner(inc, cnt, k < total ? k : 0u, &owner, &rank);
            const uint64_t ow = __shfl(word, owner);
            if (k >= total)
                continue;
            const uint32_t u = (uint32_t)((w0 + (uint64_t)owner) * 64u + (uint64_t)nth_set_bit(ow, rank));
            const uint32_t s0 = csri_off[u], e0 = csri_off[u + 1];
            if (e0 - s0 > WG_WEDGE_LANE_MAX) {
                const uint32_t nch = (e0 - s0 + WG_WEDGE_CHUNK - 1) / WG_WEDGE_CHUNK;
                const uint32_t base = atomicAdd(&ctr->heavy_len, nch);
                for (uint32_t c = 0; c < nch && base + c < heavy_cap; ++c) {
                    heavy[2 * (base + c)] = u;
                    heavy[2 * (base + c) + 1] = s0 + c * WG_WEDGE_CHUNK;
                }
                continue;
            }
            for (uint32_t p = s0; p < e0; ++p)
                mark_block(csri_pos[p] >> blog2, wedge);
        }
    }
}

// Out-edge chunks deferred by k_wedge_build: one wave per chunk.
__device__ void wedge_heavy_body(const uint32_t* __restrict__ heavy, uint32_t heavy_cap,
                                 const uint32_t* __restrict__ csri_off, const uint32_t* __restrict__ csri_pos, int blog2,
                                 uint32_t* __restrict__ wedge, wg_dev_counters* __restrict__ ctr)
{
    const uint32_t n = ctr->heavy_len < heavy_cap ? ctr->heavy_len : heavy_cap;
    const uint32_t nwaves = gridDim.x * WG_WPB;
    for (uint32_t i = blockIdx.x * WG_WPB + threadIdx.x / WG_WAVE; i < n; i += nwaves) {
        const uint32_t u = heavy[2 * i], s = heavy[2 * i + 1];
        const uint32_t end = csri_off[u + 1];
        const uint32_t e = (s + WG_WEDGE_CHUNK < end) ? s + WG_WEDGE_CHUNK : end;
        for (uint32_t p = s + (uint32_t)wg_lane(); p < e; p += WG_WAVE)
            mark_block(csri_pos[p] >> blog2, wedge);
    }
}

__global__ __launch_bounds__(WG_BLOCK) void k_wedge_heavy(const uint32_t* __restrict__ heavy, uint32_t heavy_cap,
                                                           const uint32_t* __restrict__ csri_off,
                                                           const uint32_t* __restrict__ csri_pos, int blog2_arg,
                                                           uint32_t* __restrict__ wedge, wg_dev_counters* __restrict__ ctr,
                                                           int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    wedge_heavy_body(heavy, heavy_cap, csri_off, csri_pos, gate < 0 ? blog2_arg : ctr->blog2, wedge, ctr);
}

// source/wedge.c:52-59
__global__ void k_wedge_single(uint32_t u, const uint32_t* __restrict__ csri_off, const uint32_t* __restrict__ csri_pos,
                               int blog2_arg, uint32_t* __restrict__ wedge, wg_dev_counters* __restrict__ ctr, int gate)
{
    if (wg_gated_off(ctr, gate))
        return;
    const int blog2 = gate < 0 ? blog2_arg : ctr->blog2;
    const uint32_t s = csri_off[u], e = csri_off[u + 1];
    for (uint32_t p = s + blockIdx.x * blockDim.x + threadIdx.x; p < e; p += gridDim.x * blockDim.x)
        mark_block(csri_pos[p] >> blog2, wedge);
}

__global__ void k_reached_bfs(const uint64_t* __restrict__ packed, const uint32_t* __restrict__ od, uint64_t V,
                              wg_dev_counters* __restrict__ c)
{
    unsigned long long nv = 0, ne = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        if (packed[v] != ~0ull) {
            nv += 1;
            ne += od[v];
        }
    add_count(nv, c, WG_CNT_REACHED_V);
    add_count(ne, c, WG_CNT_REACHED_E);
}

template <typename T>
__global__ void k_reached_min(const T* __restrict__ prop, T inf, const uint32_t* __restrict__ od, uint64_t V,
                              wg_dev_counters* __restrict__ c)
{
    unsigned long long nv = 0, ne = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        if (prop[v] < inf) {
            nv += 1;
            ne += od[v];
        }
    add_count(nv, c, WG_CNT_REACHED_V);
    add_count(ne, c, WG_CNT_REACHED_E);
}

__global__ void k_set_u64(uint64_t* p, uint64_t i, uint64_t v)
{
    p[i] = v;
}

__global__ void k_or_bit(uint64_t* bm, uint32_t v)
{
    bm[v >> 6] |= 1ull << (v & 63);
}

// label[v] = original id of v
__global__ void k_iota_u32(uint32_t* p, uint64_t n, const uint32_t* __restrict__ perm)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = perm ? perm[i] : (uint32_t)i;
}

// ---- device-resident iteration loop ------------------------------------------------

#define LOOP_SCHED_PLAIN 0u
#define LOOP_SCHED_LEARNED 1u
#define LOOP_SCHED_DENSE 2u
#define LOOP_SCHED_FRONTIER 3u

// Schedule inputs of the device loop.
struct loop_cfg {
    unsigned int sched;
    unsigned int app;
    int plain_blog2;
    unsigned int binned; // dense pulls on the binned layout
    unsigned long long threshold;
    unsigned long long nv, ne;
};

__device__ __forceinline__ loop_cfg read_cfg(const wg_dev_counters* c)
{
    loop_cfg g;
    g.sched = c->sched;
    g.app = c->app;
    g.plain_blog2 = c->plain_blog2;
    g.binned = c->min_binned;
    g.threshold = c->threshold;
    g.nv = c->num_vertices;
    g.ne = c->num_edges;
    return g;
}

// Schedule of the next pull from the state after the last one.  Called by all
// lanes of one wavefront (the learned rule evaluates the model across it).
__device__ __forceinline__ void loop_choose_wave(const loop_cfg& g, const float* __restrict__ mlp, unsigned long long act,
                                 unsigned long long nact, unsigned int iteration, unsigned long long unvisited, int* mode,
                                 int* blog2)
{
    *blog2 = g.plain_blog2;
    switch (g.sched) {
    case LOOP_SCHED_DENSE:
        *mode = WG_MODE_DENSE;
        return;
    case LOOP_SCHED_FRONTIER:
        *mode = WG_MODE_FRONTIER;
        return;
    case LOOP_SCHED_LEARNED:
        if (wg_sched_binned_dense((int)g.binned, act, g.ne)) { // dense pulls on the binned layout
            *mode = WG_MODE_DENSE;
            return;
        }
        // below the gate: the plain rule's frontier pull (wg_mlp_gated)
        if (mlp && !wg_mlp_gated(act, g.ne) && !wg_mlp_out_of_range(mlp, (int)g.app, g.nv)) {
            float x[WG_MLP_IN];
            float base[WG_MLP_ACTIONS];
            wg_mlp_features((int)g.app, iteration, g.nv, g.ne, nact, act, unvisited, x);
            wg_mlp_base(g.nv, g.ne, act, base);
            const int plain = wg_mlp_plain_action(act > g.threshold, g.plain_blog2);
            const int a = wg_mlp_choose_wave(mlp, (int)g.app, x, base, plain, nullptr);
            *mode = a == 0 ? WG_MODE_DENSE : WG_MODE_FRONTIER;
            if (a)
                *blog2 = WG_MLP_ACTION_BLOG2(a);
            return;
        }
        // no model: the plain rule
        [[fallthrough]];
    default:
        // source/execution_bfs.c:141
        *mode = act > g.threshold ? WG_MODE_DENSE : WG_MODE_FRONTIER;
        return;
    }
}

// One wavefront.  initial_vertices / initial_activity describe the state
// before the first pull (the root, or every vertex for CC).
__global__ void k_loop_init(wg_dev_counters* c, const float* __restrict__ mlp, unsigned int sched,
                            unsigned long long threshold, int plain_blog2, unsigned int max_iters,
                            unsigned long long initial_activity, unsigned long long initial_vertices, int dense_first,
                            unsigned long long unvisited, unsigned int app, unsigned long long nv, unsigned long long ne,
                            unsigned int binned)
{
    if (blockIdx.x != 0 || threadIdx.x >= WG_WAVE)
        return;
    loop_cfg g;
    g.sched = sched;
    g.app = app;
    g.plain_blog2 = plain_blog2;
    g.binned = binned;
    g.threshold = threshold;
    g.nv = nv;
    g.ne = ne;
    int mode = 0, blog2 = plain_blog2;
    loop_choose_wave(g, mlp, initial_activity, initial_vertices, 0u, unvisited, &mode, &blog2);
    if (dense_first)
        mode = WG_MODE_DENSE; // CC: source/execution_cc.c:103
    if (threadIdx.x != 0)
        return;
    c->sched = sched;
    c->app = app;
    c->min_binned = binned;
    c->num_vertices = nv;
    c->num_edges = ne;
    c->threshold = threshold;
    c->plain_blog2 = plain_blog2;
    c->max_iters = max_iters;
    c->done = 0u;
    c->run_err = 0u;
    c->run_levels = 0u;
    c->iters = 0u;
    c->fiters = 0u;
    c->activity = 0ull;
    c->activated = 0ull;
    c->fresh = 0ull;
    c->heavy_len = 0u;
    c->last_activity = initial_activity;
    c->last_activated = initial_vertices;
    c->unvisited = unvisited;
    c->mode = (unsigned int)mode;
    c->blog2 = blog2;
}

// End of a pull (one wavefront): fold the counts, account the iteration,
// stop at zero activity, choose the next schedule.
__device__ void decide_wave(wg_dev_counters* c, const float* __restrict__ mlp)
{
    if (c->done)
        return;
    const unsigned long long act0 = c->activity, nact0 = c->activated, fresh0 = c->fresh;
    const unsigned int iters = c->iters + 1u, fiters = c->fiters + (c->mode == WG_MODE_FRONTIER ? 1u : 0u);
    const unsigned int max_iters = c->max_iters;
    const unsigned long long unv0 = c->unvisited;
    const loop_cfg g = read_cfg(c);
    unsigned long long tot[WG_CNT_N];
    fold_counts(c, tot);
    const unsigned long long act = act0 + tot[WG_CNT_ACTIVITY], nact = nact0 + tot[WG_CNT_ACTIVATED];
    const unsigned long long fresh = fresh0 + tot[WG_CNT_FRESH];
    const unsigned long long unvisited = unv0 >= fresh ? unv0 - fresh : 0ull;
    const bool stop = act == 0ull || (max_iters && iters >= max_iters);
    int mode = 0, blog2 = g.plain_blog2;
    if (!stop)
        loop_choose_wave(g, mlp, act, nact, iters, unvisited, &mode, &blog2);
    if (threadIdx.x != 0)
        return;
    c->iters = iters;
    c->fiters = fiters;
    c->activity = 0ull;
    c->activated = 0ull;
    c->fresh = 0ull;
    c->heavy_len = 0u;
    c->frontier_verts = 0ull;
    c->last_activity = act;
    c->last_activated = nact;
    c->unvisited = unvisited;
    if (stop) {
        c->done = 1u;
        return;
    }
    c->mode = (unsigned int)mode;
    c->blog2 = blog2;
}

// Also clears the counts of the fused-level run that may follow (rc, if any).
__global__ void k_decide(wg_dev_counters* c, const float* __restrict__ mlp, wg_run_ctl* __restrict__ rc)
{
    if (blockIdx.x != 0)
        return;
    if (threadIdx.x < WG_WAVE)
        decide_wave(c, mlp);
    if (rc && threadIdx.x < 3) {
        rc->tail[threadIdx.x].v = 0ull;
        rc->act[threadIdx.x].v = 0ull;
        rc->nact[threadIdx.x].v = 0ull;
        rc->fresh[threadIdx.x].v = 0ull;
    }
}

// One level of the device-resident loop in four launches: the pull the
// control block selected (dense or frontier, one kernel), the decision for
// the next level, the frontier build for it, and the advance (with the hub
// out-lists the build deferred).  Every launch costs a dispatch, gated off or
// not (back-to-back launches start about 3 us apart on this device), and each
// kernel its own chain of memory round trips, which dominates levels with
// small frontiers.  The decision stays a kernel of its
// own: taking it in the last workgroup of the pull needs device-scope fences
// in every wave, which cost more than the launch.
template <bool CW>
__global__ __launch_bounds__(WG_BLOCK) void k_bfs_level(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                         const uint32_t* __restrict__ dst, const uint32_t* __restrict__ od,
                                                         const uint64_t* __restrict__ cur, const uint64_t* __restrict__ visited,
                                                         uint64_t* __restrict__ next, uint64_t* __restrict__ packed,
                                                         uint32_t V, uint32_t E, uint32_t* __restrict__ wedge,
                                                         const uint32_t* __restrict__ perm, wg_dev_counters* __restrict__ ctr,
                                                         const uint32_t* __restrict__ hv, const uint32_t* __restrict__ he,
                                                         uint32_t nheavy, const uint32_t* __restrict__ first,
                                                         const uint32_t* __restrict__ roots, uint32_t npush,
                                                         const uint32_t* __restrict__ root_off,
                                                         const uint32_t* __restrict__ csri_off,
                                                         const uint32_t* __restrict__ csr_dst)
{
    if (wg_gated_off(ctr, WG_GATE_LIVE))
        return;
    if (ctr->mode == WG_MODE_DENSE) {
        __shared__ uint64_t s_cur[WG_BFS_CUR_LDS];
        stage_cur(cur, V, s_cur);
        bfs_dense_body<CW>(off, src, first, od, cur, visited, next, packed, V, perm, ctr, s_cur);
        // and the high in-degree destinations the dense body leaves out
        bfs_heavy_body<CW>(hv, he, nheavy, off, src, od, cur, visited, next, packed, perm, ctr);
    } else if (npush && ctr->iters == 0u) // the frontier is the npush roots; no blocks were marked
        bfs_push_roots_body<CW>(roots, npush, root_off, csri_off, csr_dst, od, visited, next, packed, perm, ctr);
    else
        bfs_blocks_body<CW>(ctr, ctr->blog2, E, src, dst, od, cur, visited, next, packed, wedge, perm);
}

template <bool CW, class P, int IPT = WG_IPT>
__global__ __launch_bounds__(WG_BLOCK) void k_min_level(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                         const uint32_t* __restrict__ dst, const void* __restrict__ w,
                                                         typename P::T* __restrict__ prop, const uint32_t* __restrict__ od,
                                                         const uint32_t* __restrict__ tile_v,
                                                         const uint32_t* __restrict__ tile_e, uint32_t ntiles, uint32_t E,
                                                         uint64_t* __restrict__ next, uint32_t* __restrict__ wedge,
                                                         wg_dev_counters* __restrict__ ctr, unsigned int* __restrict__ ovf,
                                                         uint32_t nhub, int binned)
{
    if (wg_gated_off(ctr, WG_GATE_LIVE))
        return;
    if (ctr->mode == WG_MODE_DENSE) {
        if (binned) // the binned push and pull launched ahead of this kernel did the dense pull
            return;
        __shared__ typename P::T s_hub[P::HUB ? WG_MIN_HUB_BYTES / sizeof(typename P::T) : 1];
        stage_hub<P>(prop, nhub, s_hub);
        min_tile_body<CW, P, IPT>(off, src, w, prop, od, tile_v, tile_e, ntiles, next, ctr, ovf, s_hub, nhub);
    } else
        min_blocks_body<CW, P>(ctr, ctr->blog2, E, src, dst, w, prop, od, next, wedge, ovf);
}

// Advance (BFS) or count-and-clear (CC, SSSP) of the activated vertices, and
// first the hub out-lists the frontier build deferred (those marks are read
// only by the next pull).
__global__ __launch_bounds__(WG_BLOCK) void k_level_advance(uint64_t* __restrict__ visited, uint64_t* __restrict__ next,
                                                             uint64_t* __restrict__ cur, uint64_t nwords, int bfs,
                                                             const uint32_t* __restrict__ heavy, uint32_t heavy_cap,
                                                             const uint32_t* __restrict__ csri_off,
                                                             const uint32_t* __restrict__ csri_pos,
                                                             uint32_t* __restrict__ wedge, wg_dev_counters* __restrict__ ctr)
{
    if (wg_gated_off(ctr, WG_GATE_LIVE))
        return;
    if (ctr->mode == WG_MODE_FRONTIER)
        wedge_heavy_body(heavy, heavy_cap, csri_off, csri_pos, ctr->blog2, wedge, ctr);
    if (bfs)
        bfs_advance_body(visited, next, cur, nwords, ctr);
    else
        count_clear_body(next, nwords, ctr);
}

// ---- fused small-frontier levels ---------------------------------------------------
//
// Levels whose frontier is small cost the loop four dispatches each, while
// their edges keep a handful of workgroups busy for a few microseconds.  A run
// executes such levels back to back in one launch of a few resident
// workgroups, with one grid barrier per level:
//   - the frontier is a queue of entries (vertex, out-edge range of at most
//     WG_RUN_CH CSRI positions, first-chunk flag), so a level touches only
//     the out-edges of its active vertices;
//   - BFS: each out-edge u -> v does the frontier pull's atomicMin of the
//     packed (depth, parent) word, so the smallest parent on the previous
//     level wins as in every other pull; the one thread whose atomic found v
//     unvisited appends v's out-edges to the next queue and sets v in next;
//   - CC / SSSP: each out-edge offers v the value of u (plus the edge
//     weight) with the pulls' atomicMin; a vertex whose value drops is
//     activated once per level (a bit in one of two alternating bitmasks) and
//     appended to the next queue;
//   - every workgroup then takes the decision k_decide takes (the same counts
//     and schedule code, so all take the same one) and the run goes on while
//     the schedule stays frontier-filtered and the level stays small
//     (WG_RUN_ACT_MAX, WG_RUN_NACT_MAX).
// A run starts after k_decide chose a frontier-filtered pull for a small
// activated set (the bits of next) and ends in the state a pull followed by
// k_decide leaves: next holds the last level's activated vertices, the
// control block the counts and the next schedule; the frontier build and the
// advance follow as after any pull.  BFS levels, parents and iteration counts
// are those of the level-per-launch loop.  CC labels and SSSP distances are
// too (the fixpoint is unique), but a run activates exactly the vertices whose
// value dropped while a frontier-filtered pull reads whole edge blocks, so
// the two may take different numbers of iterations to get there.  bm_cur is
// not kept up to date during a BFS run (nothing reads it before the advance
// rewrites it); visited gains each vertex when its first entry is expanded.

// Grid barrier of a run.  A run has at most WG_RUN_MAX_WGS workgroups (no
// more than the device has CUs), launched after the previous kernel of the
// stream has finished, so all are resident together and an arrival counter
// with a generation word suffices.  The wait is bounded: a workgroup that has
// polled WG_RUN_SPIN times sets run_err and done in the control block and
// returns, so the kernel ends whatever happens and the host reports the
// error.
#define WG_RUN_SPIN (1u << 19)

__device__ bool run_barrier(wg_run_ctl* __restrict__ rc, unsigned int G, wg_dev_counters* __restrict__ ctr)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        unsigned int* cnt = (unsigned int*)&rc->count.v;
        unsigned int* gen = (unsigned int*)&rc->gen.v;
        const unsigned int g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        if (atomicAdd(cnt, 1u) == G - 1u) {
            __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(gen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned int n = 0;
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (++n > WG_RUN_SPIN) {
                    ok = 0;
                    atomicOr(&ctr->run_err, 1u);
                    atomicOr(&ctr->done, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// Exclusive prefix sum over the workgroup; *total receives the sum.
__device__ uint32_t block_excl_scan(uint32_t x, uint32_t* total)
{
    __shared__ uint32_t s_w[WG_WPB];
    const int lane = wg_lane(), wv = (int)(threadIdx.x / WG_WAVE);
    uint32_t inc = x;
    for (int d = 1; d < WG_WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d);
        if (lane >= d)
            inc += y;
    }
    if (lane == WG_WAVE - 1)
        s_w[wv] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int i = 0; i < WG_WPB; ++i) {
        const uint32_t t = s_w[i];
        before += (i < wv) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return before + inc - x;
}

// Workgroup-wide sum (every thread gets it).
__device__ unsigned long long block_sum_u64(unsigned long long x)
{
    __shared__ unsigned long long s_w[WG_WPB];
    x = wg_wave_sum(x);
    if (wg_lane() == 0)
        s_w[threadIdx.x / WG_WAVE] = x;
    __syncthreads();
    unsigned long long t = 0ull;
    for (int i = 0; i < WG_WPB; ++i)
        t += s_w[i];
    __syncthreads();
    return t;
}

__device__ __forceinline__ uint32_t run_entries(uint32_t len)
{
    return len ? (len + WG_RUN_CH - 1u) / WG_RUN_CH : 1u;
}

// Writes the entries of vertex v (out-edges [s, e) of the CSRI) from idx on.
// Entries past WG_RUN_QCAP are dropped: the level that produced them has more
// activity than a run takes, so the run ends and never reads the queue.
__device__ __forceinline__ uint32_t run_put(uint4* __restrict__ q, uint32_t idx, uint32_t v, uint32_t s, uint32_t e)
{
    uint32_t c = s;
    do {
        if (idx >= WG_RUN_QCAP)
            return idx;
        const uint32_t ce = (e - c > WG_RUN_CH) ? c + WG_RUN_CH : e;
        q[idx] = make_uint4(v, c, ce, c == s ? 1u : 0u);
        ++idx;
        c = ce;
    } while (c < e);
    return idx;
}

// Appends the entries of the k vertices a thread activated (vertex vv[i]
// with out-edges [so[i], eo[i]) if nw[i]) to the queue whose tail is *tail:
// one atomic per workgroup.  A vertex with more than WG_RUN_SERIAL entries
// is written by its whole wavefront, one entry per lane.  A workgroup whose
// entries would pass WG_RUN_QCAP writes none: the level has more activity
// than a run takes, so the run ends without reading them.
#define WG_RUN_SERIAL 8u
template <int N>
__device__ __forceinline__ void run_append(uint4* __restrict__ q, unsigned int* __restrict__ tail, const uint32_t* vv,
                                           const uint32_t* so, const uint32_t* eo, const bool* nw, uint32_t k)
{
    __shared__ uint32_t s_base;
    uint32_t tot;
    const uint32_t pre = block_excl_scan(k, &tot);
    if (threadIdx.x == 0)
        s_base = tot ? atomicAdd(tail, tot) : 0u;
    __syncthreads();
    const uint32_t base = s_base;
    if (tot && (uint64_t)base + tot <= WG_RUN_QCAP) { // the same in the whole workgroup
        const int lane = wg_lane();
        uint32_t idx = base + pre;
        uint32_t bidx[N];
        bool big[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            bidx[j] = idx;
            big[j] = false;
            if (nw[j]) {
                const uint32_t n = run_entries(eo[j] - so[j]);
                if (n <= WG_RUN_SERIAL)
                    run_put(q, idx, vv[j], so[j], eo[j]);
                else
                    big[j] = true;
                idx += n;
            }
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            unsigned long long m = wg_ballot(big[j]);
            while (m) {
                const int l = __ffsll((long long)m) - 1;
                m &= m - 1ull;
                const uint32_t v = __shfl(vv[j], l), s = __shfl(so[j], l), e = __shfl(eo[j], l);
                const uint32_t b = __shfl(bidx[j], l);
                const uint32_t n = run_entries(e - s);
                for (uint32_t c = (uint32_t)lane; c < n; c += WG_WAVE) {
                    const uint32_t cs = s + c * WG_RUN_CH;
                    const uint32_t ce = (e - cs > WG_RUN_CH) ? cs + WG_RUN_CH : e;
                    q[b + c] = make_uint4(v, cs, ce, c == 0 ? 1u : 0u);
                }
            }
        }
    }
    __syncthreads();
}

// The run's copy of the loop state (the same in every thread).
struct run_state {
    loop_cfg g;
    unsigned int max_iters, level0, iters, fiters;
    unsigned long long unvisited;
};

// Gate of a run, the same in every workgroup: the control block changes only
// after the last barrier of a run.
__device__ __forceinline__ bool run_gate(const wg_dev_counters* ctr, run_state& st)
{
    if (ctr->done || ctr->mode != WG_MODE_FRONTIER || ctr->last_activity > WG_RUN_ACT_PER_WG * gridDim.x ||
        ctr->last_activated > WG_RUN_NACT_PER_WG * gridDim.x)
        return false;
    st.g = read_cfg(ctr);
    st.max_iters = ctr->max_iters;
    st.level0 = ctr->level;
    st.iters = ctr->iters;
    st.fiters = ctr->fiters;
    st.unvisited = ctr->unvisited;
    return true;
}

// Copies the schedule model to LDS for the run's per-level decisions, if the
// learned schedule is on and has a model; returns whether it is.
__device__ __forceinline__ bool run_stage_model(const float* __restrict__ mlp, const run_state& st, float* s_mlp)
{
    if (!mlp || st.g.sched != LOOP_SCHED_LEARNED)
        return false;
    for (int i = threadIdx.x; i < WG_MLP_OFF_END; i += WG_BLOCK)
        s_mlp[i] = mlp[i];
    __syncthreads();
    return true;
}

// Entry: the vertices the last pull activated (bits of act) form queue q.
// Each thread reads 16 words at a time, strided by the grid, so that the
// loads are coalesced and many are in flight.
#define WG_RUN_SCAN_BATCH 16
__device__ void run_entry(const uint64_t* __restrict__ act, uint64_t nwords, const uint32_t* __restrict__ csri_off,
                          uint4* __restrict__ q, wg_run_ctl* __restrict__ rc)
{
    const uint64_t stride = (uint64_t)gridDim.x * WG_BLOCK;
    const uint64_t t = (uint64_t)blockIdx.x * WG_BLOCK + threadIdx.x;
    uint32_t k = 0;
    for (uint64_t w0 = t; w0 < nwords; w0 += WG_RUN_SCAN_BATCH * stride) {
        uint64_t x[WG_RUN_SCAN_BATCH];
#pragma unroll
        for (int j = 0; j < WG_RUN_SCAN_BATCH; ++j) {
            const uint64_t w = w0 + (uint64_t)j * stride;
            x[j] = (w < nwords) ? act[w] : 0ull;
        }
#pragma unroll
        for (int j = 0; j < WG_RUN_SCAN_BATCH; ++j)
            for (uint64_t b = x[j]; b; b &= b - 1ull) {
                const uint32_t v = (uint32_t)((w0 + (uint64_t)j * stride) * 64u + (uint64_t)(__ffsll((long long)b) - 1));
                k += run_entries(csri_off[v + 1] - csri_off[v]);
            }
    }
    __shared__ uint32_t s_base;
    uint32_t tot;
    const uint32_t pre = block_excl_scan(k, &tot);
    if (threadIdx.x == 0)
        s_base = tot ? atomicAdd((unsigned int*)&rc->tail[0].v, tot) : 0u;
    __syncthreads();
    uint32_t idx = s_base + pre;
    if (k)
        for (uint64_t w0 = t; w0 < nwords; w0 += WG_RUN_SCAN_BATCH * stride) {
            uint64_t x[WG_RUN_SCAN_BATCH];
#pragma unroll
            for (int j = 0; j < WG_RUN_SCAN_BATCH; ++j) {
                const uint64_t w = w0 + (uint64_t)j * stride;
                x[j] = (w < nwords) ? act[w] : 0ull;
            }
#pragma unroll
            for (int j = 0; j < WG_RUN_SCAN_BATCH; ++j)
                for (uint64_t b = x[j]; b; b &= b - 1ull) {
                    const uint32_t v =
                        (uint32_t)((w0 + (uint64_t)j * stride) * 64u + (uint64_t)(__ffsll((long long)b) - 1));
                    idx = run_put(q, idx, v, csri_off[v], csri_off[v + 1]);
                }
        }
}

// End of level r: the workgroup's counts into the level's totals, the
// barrier, then the decision k_decide would take after this pull (with the
// model in LDS, s_mlp, if learned).  Returns whether the run goes on; when it
// does not, the control block holds the state after this level.  *ok turns
// false if the barrier gave up.
__device__ __forceinline__ bool run_level_end(wg_run_ctl* __restrict__ rc, wg_dev_counters* __restrict__ ctr,
                                              unsigned int r, unsigned long long act, unsigned long long nact,
                                              unsigned long long fresh, run_state& st, const float* s_mlp, bool learned,
                                              bool* ok)
{
    const unsigned int rn = (r + 1u) % 3u, rz = (r + 2u) % 3u;
    __shared__ int s_dec[2];
    act = block_sum_u64(act);
    nact = block_sum_u64(nact);
    fresh = block_sum_u64(fresh);
    if (threadIdx.x == 0) {
        if (act)
            atomicAdd(&rc->act[rn].v, act);
        if (nact)
            atomicAdd(&rc->nact[rn].v, nact);
        if (fresh)
            atomicAdd(&rc->fresh[rn].v, fresh);
        if (blockIdx.x == 0) { // every workgroup read these before the last barrier
            __hip_atomic_store(&rc->tail[rz].v, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&rc->act[rz].v, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&rc->nact[rz].v, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&rc->fresh[rz].v, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (!run_barrier(rc, gridDim.x, ctr)) {
        *ok = false;
        return false;
    }
    const unsigned long long A = __hip_atomic_load(&rc->act[rn].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long N = __hip_atomic_load(&rc->nact[rn].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long F = __hip_atomic_load(&rc->fresh[rn].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long T = __hip_atomic_load(&rc->tail[rn].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st.iters += 1u;
    st.fiters += 1u;
    st.unvisited = st.unvisited >= F ? st.unvisited - F : 0ull;
    const bool stop = A == 0ull || (st.max_iters && st.iters >= st.max_iters);
    if (!stop && threadIdx.x < WG_WAVE) {
        int mode = 0, blog2 = st.g.plain_blog2;
        if (learned) // s_mlp is the caller's __shared__ array: LDS reads
            loop_choose_wave(st.g, s_mlp, A, N, st.iters, st.unvisited, &mode, &blog2);
        else
            loop_choose_wave(st.g, nullptr, A, N, st.iters, st.unvisited, &mode, &blog2);
        if (threadIdx.x == 0) {
            s_dec[0] = mode;
            s_dec[1] = blog2;
        }
    }
    __syncthreads();
    const int mode = stop ? 0 : s_dec[0];
    const int blog2 = stop ? st.g.plain_blog2 : s_dec[1];
    __syncthreads();
    const bool cont =
        !stop && mode == WG_MODE_FRONTIER && A <= WG_RUN_ACT_PER_WG * gridDim.x &&
        N <= WG_RUN_NACT_PER_WG * gridDim.x && T <= WG_RUN_QCAP;
    if (!cont && blockIdx.x == 0 && threadIdx.x == 0) {
        ctr->iters = st.iters;
        ctr->fiters = st.fiters;
        ctr->last_activity = A;
        ctr->last_activated = N;
        ctr->unvisited = st.unvisited;
        ctr->level = st.level0 + r + 1u;
        ctr->run_levels += r + 1u;
        if (stop) {
            ctr->done = 1u;
        } else {
            ctr->mode = (unsigned int)mode;
            ctr->blog2 = blog2;
        }
    }
    return cont;
}

__device__ __forceinline__ uint32_t run_queue_len(const wg_run_ctl* rc, unsigned int r)
{
    const uint32_t n = __hip_atomic_load((const unsigned int*)&rc->tail[r % 3u].v, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return n < WG_RUN_QCAP ? n : WG_RUN_QCAP;
}

template <bool CW>
__global__ __launch_bounds__(WG_BLOCK) void k_bfs_run(const uint32_t* __restrict__ csri_off,
                                                       const uint32_t* __restrict__ csr_dst, uint64_t* __restrict__ next,
                                                       uint64_t* __restrict__ visited, uint64_t* __restrict__ packed,
                                                       uint64_t nwords, const uint32_t* __restrict__ perm,
                                                       uint4* __restrict__ qa, uint4* __restrict__ qb,
                                                       wg_run_ctl* __restrict__ rc, wg_dev_counters* __restrict__ ctr,
                                                       const float* __restrict__ mlp)
{
    run_state st;
    if (!run_gate(ctr, st))
        return;
    __shared__ float s_mlp[WG_MLP_OFF_END];
    const bool learned = run_stage_model(mlp, st, s_mlp);
    run_entry(next, nwords, csri_off, qa, rc);
    if (!run_barrier(rc, gridDim.x, ctr))
        return;
    uint4* qc = qa;
    uint4* qn = qb;
    work_tally<CW> work;
    for (unsigned int r = 0;; ++r) {
        const uint32_t nc = run_queue_len(rc, r);
        const unsigned long long newdepth = (unsigned long long)(st.level0 + 2u + r);
        unsigned int* tn = (unsigned int*)&rc->tail[(r + 1u) % 3u].v;
        unsigned long long act = 0ull, nact = 0ull;
        for (uint32_t i0 = blockIdx.x * WG_BLOCK; i0 < nc; i0 += gridDim.x * WG_BLOCK) {
            const uint32_t i = i0 + threadIdx.x;
            uint32_t vv[WG_RUN_CH], so[WG_RUN_CH], eo[WG_RUN_CH];
            bool nw[WG_RUN_CH];
#pragma unroll
            for (int j = 0; j < (int)WG_RUN_CH; ++j) {
                vv[j] = so[j] = eo[j] = 0u;
                nw[j] = false;
            }
            uint32_t k = 0;
            if (i < nc) {
                const uint4 en = qc[i];
                const uint32_t u = en.x;
                if (en.w) { // u leaves next for visited, as at an advance
                    const unsigned long long m = 1ull << (u & 63u);
                    atomicOr((unsigned long long*)&visited[u >> 6], m);
                    atomicAnd((unsigned long long*)&next[u >> 6], ~m);
                }
                const uint32_t pid = perm ? perm[u] : u;
                const unsigned long long val = (newdepth << 32) | (unsigned long long)pid;
                work.edges += en.z - en.y; // each out-edge and its destination's state
                work.dests += en.z - en.y;
#pragma unroll
                for (int j = 0; j < (int)WG_RUN_CH; ++j)
                    if (en.y + (uint32_t)j < en.z)
                        vv[j] = csr_dst[en.y + (uint32_t)j];
#pragma unroll
                for (int j = 0; j < (int)WG_RUN_CH; ++j)
                    if (en.y + (uint32_t)j < en.z)
                        nw[j] = atomicMin((unsigned long long*)&packed[vv[j]], val) == ~0ull;
#pragma unroll
                for (int j = 0; j < (int)WG_RUN_CH; ++j)
                    if (nw[j]) {
                        so[j] = csri_off[vv[j]];
                        eo[j] = csri_off[vv[j] + 1];
                        atomicOr((unsigned long long*)&next[vv[j] >> 6], 1ull << (vv[j] & 63u));
                    }
#pragma unroll
                for (int j = 0; j < (int)WG_RUN_CH; ++j)
                    if (nw[j]) {
                        k += run_entries(eo[j] - so[j]);
                        act += eo[j] - so[j];
                        nact += 1ull;
                    }
            }
            run_append<WG_RUN_CH>(qn, tn, vv, so, eo, nw, k);
        }
        bool ok = true;
        add_work(work, ctr);
        work = work_tally<CW>();
        // a BFS activation is always a first visit
        if (!run_level_end(rc, ctr, r, act, nact, nact, st, s_mlp, learned, &ok))
            return;
        uint4* t = qc;
        qc = qn;
        qn = t;
    }
}

// CC / SSSP.  Level r expands queue r, whose vertices are set in mask
// D[r & 1], and activates into D[(r + 1) & 1] (D[0] = next, D[1] = spare):
// a vertex leaves D[r & 1] when its first entry is expanded, and nothing
// sets D[r & 1] during level r, so a vertex whose value drops during level r
// (even after some of its entries were expanded with the old value) is
// appended again.  A run that ends after an even level moves its activated
// set from spare to next.
template <bool CW, class P>
__global__ __launch_bounds__(WG_BLOCK) void k_min_run(const uint32_t* __restrict__ csri_off,
                                                       const uint32_t* __restrict__ csri_pos,
                                                       const uint32_t* __restrict__ dst, const void* __restrict__ w,
                                                       typename P::T* __restrict__ prop, uint64_t* __restrict__ next,
                                                       uint64_t* __restrict__ spare, uint64_t nwords,
                                                       uint4* __restrict__ qa, uint4* __restrict__ qb,
                                                       wg_run_ctl* __restrict__ rc, wg_dev_counters* __restrict__ ctr,
                                                       unsigned int* __restrict__ ovf, const float* __restrict__ mlp)
{
    typedef typename P::T T;
    run_state st;
    if (!run_gate(ctr, st))
        return;
    __shared__ float s_mlp[WG_MLP_OFF_END];
    const bool learned = run_stage_model(mlp, st, s_mlp);
    run_entry(next, nwords, csri_off, qa, rc);
    if (!run_barrier(rc, gridDim.x, ctr))
        return;
    uint4* qc = qa;
    uint4* qn = qb;
    work_tally<CW> work;
    for (unsigned int r = 0;; ++r) {
        uint64_t* dc = (r & 1u) ? spare : next;
        uint64_t* dn = (r & 1u) ? next : spare;
        const uint32_t nc = run_queue_len(rc, r);
        unsigned int* tn = (unsigned int*)&rc->tail[(r + 1u) % 3u].v;
        unsigned long long act = 0ull, nact = 0ull, fresh = 0ull;
        for (uint32_t i0 = blockIdx.x * WG_BLOCK; i0 < nc; i0 += gridDim.x * WG_BLOCK) {
            const uint32_t i = i0 + threadIdx.x;
            uint32_t vv[WG_RUN_CH], so[WG_RUN_CH], eo[WG_RUN_CH], pos[WG_RUN_CH];
            bool nw[WG_RUN_CH];
#pragma unroll
            for (int j = 0; j < (int)WG_RUN_CH; ++j) {
                vv[j] = so[j] = eo[j] = pos[j] = 0u;
                nw[j] = false;
            }
            uint32_t k = 0;
            if (i < nc) {
                const uint4 en = qc[i];
                const uint32_t u = en.x;
                if (en.w)
                    atomicAnd((unsigned long long*)&dc[u >> 6], ~(1ull << (u & 63u)));
                work.edges += en.z - en.y; // each out-edge and its destination's state
                work.dests += en.z - en.y;
#pragma unroll
                for (int j = 0; j < (int)WG_RUN_CH; ++j)
                    if (en.y + (uint32_t)j < en.z) {
                        pos[j] = csri_pos[en.y + (uint32_t)j];
                        vv[j] = dst[pos[j]];
                    }
#pragma unroll
                for (int j = 0; j < (int)WG_RUN_CH; ++j)
                    if (en.y + (uint32_t)j < en.z) {
                        const T m = P::val(prop, w, u, pos[j], ovf);
                        if (m < prop[vv[j]]) {
                            const T old = atomic_min_t<T>(&prop[vv[j]], m);
                            if (old > m) {
                                const unsigned long long b = 1ull << (vv[j] & 63u);
                                nw[j] = !(atomicOr((unsigned long long*)&dn[vv[j] >> 6], b) & b);
                                fresh += (old == P::INF) ? 1ull : 0ull;
                            }
                        }
                    }
#pragma unroll
                for (int j = 0; j < (int)WG_RUN_CH; ++j)
                    if (nw[j]) {
                        so[j] = csri_off[vv[j]];
                        eo[j] = csri_off[vv[j] + 1];
                    }
#pragma unroll
                for (int j = 0; j < (int)WG_RUN_CH; ++j)
                    if (nw[j]) {
                        k += run_entries(eo[j] - so[j]);
                        act += eo[j] - so[j];
                        nact += 1ull;
                    }
            }
            run_append<WG_RUN_CH>(qn, tn, vv, so, eo, nw, k);
        }
        bool ok = true;
        add_work(work, ctr);
        work = work_tally<CW>();
        if (!run_level_end(rc, ctr, r, act, nact, fresh, st, s_mlp, learned, &ok)) {
            if (ok && (r & 1u) == 0u) // the last activations are in spare
                for (uint64_t wd = (uint64_t)blockIdx.x * WG_BLOCK + threadIdx.x; wd < nwords;
                     wd += (uint64_t)gridDim.x * WG_BLOCK) {
                    const uint64_t x = spare[wd];
                    if (x) {
                        next[wd] = x;
                        spare[wd] = 0ull;
                    }
                }
            return;
        }
        uint4* t = qc;
        qc = qn;
        qn = t;
    }
}

// ---- host side ----------------------------------------------------------------------

// The instance of a pull kernel with (CW = true) or without the work counts.
#define WG_CW(ctx, K, ...) ((ctx)->count_work ? K<true, ##__VA_ARGS__> : K<false, ##__VA_ARGS__>)

// Folds the counter slots and copies the counters to the host mirror.
int wg_read_counters(wg_ctx* ctx)
{
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(WG_WAVE), 0, ctx->stream, ctx->ctr);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipMemcpyAsync(ctx->hctr, ctx->ctr, WG_CTR_HOST_BYTES, hipMemcpyDeviceToHost, ctx->stream));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return WG_OK;
}

// Folds the slots and hands the level's activity counts to the host mirror
// (mapped pinned memory), then clears them for the next level.
__global__ void k_take_activity(wg_dev_counters* c, wg_dev_counters* host)
{
    if (threadIdx.x < WG_WAVE)
        fold_counts(c);
    if (threadIdx.x == 0) { // the lane that folded the totals
        host->activity = c->activity;
        host->activated = c->activated;
        host->fresh = c->fresh;
        c->activity = 0ull;
        c->activated = 0ull;
        c->fresh = 0ull;
        __threadfence_system();
    }
}

int wg_take_activity(wg_ctx* ctx)
{
    if (!ctx->hctr_dev) {
        int rc = wg_read_counters(ctx);
        return rc ? rc : wg_reset_counters(ctx, RST_ACTIVITY);
    }
    hipLaunchKernelGGL(k_take_activity, dim3(1), dim3(WG_WAVE), 0, ctx->stream, ctx->ctr, ctx->hctr_dev);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return WG_OK;
}

static int clear_wedge_state(wg_ctx* ctx)
{
    int rc = wg_fill_u32(ctx, ctx->wedge, ctx->wedge_words, 0u);
    if (rc)
        return rc;
    rc = wg_reset_counters(ctx, RST_HEAVY);
    ctx->wedge_log2 = -1;
    return rc;
}

static int init_common(wg_ctx* ctx, int app)
{
    if (!ctx || !ctx->off)
        return WG_E_STATE;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    wg_free_app_state(ctx);
    int rc;
    if ((rc = wg_clear_u64(ctx, ctx->bm_cur, ctx->bm_words)) || (rc = wg_clear_u64(ctx, ctx->bm_next, ctx->bm_words)) ||
        (rc = wg_clear_u64(ctx, ctx->bm_visited, ctx->bm_words)) || (rc = clear_wedge_state(ctx)) ||
        (rc = wg_reset_counters(ctx, RST_ACTIVITY | RST_FRONTIER | RST_OVERFLOW | RST_WORK)))
        return rc;
    ctx->app = app;
    ctx->iter = 0;
    ctx->roots_frontier = 0;
    // the 256-edge tiles of the dense SSSP pulls (built once per graph)
    if (app == WG_APP_SSSP && (rc = wg_build_pr_tilings(ctx)))
        return rc;
    // the smallest in-neighbours the dense BFS levels test first (once per graph)
    if (app == WG_APP_BFS && !ctx->first_src) {
        WG_HIP(ctx, hipMalloc((void**)&ctx->first_src, (ctx->V ? ctx->V : 1) * sizeof(uint32_t)));
        hipLaunchKernelGGL(k_first_src, dim3(wg_grid(ctx, ctx->V, 256, 8)), dim3(256), 0, ctx->stream, ctx->off, ctx->src,
                           ctx->V, ctx->first_src);
        if ((rc = wg_check_launch(ctx)))
            return rc;
    }
    // the out-edges' destinations in out-edge order, for the first level's
    // push and the fused levels (once per graph)
    if (app == WG_APP_BFS && !ctx->csr_dst) {
        WG_HIP(ctx, hipMalloc((void**)&ctx->csr_dst, (ctx->E ? ctx->E : 1) * sizeof(uint32_t)));
        if (ctx->E)
            hipLaunchKernelGGL(k_csr_dst, dim3(wg_grid(ctx, ctx->E, 1024, 8)), dim3(256), 0, ctx->stream, ctx->csri_pos,
                               ctx->dst, ctx->E, ctx->csr_dst);
        if ((rc = wg_check_launch(ctx)))
            return rc;
    }
    WG_HIP(ctx, hipMemsetAsync(&ctx->ctr->level, 0, sizeof(unsigned int), ctx->stream));
    WG_HIP(ctx, hipMemsetAsync(&ctx->ctr->done, 0, sizeof(unsigned int), ctx->stream));
    return WG_OK;
}

// Roots (device ids, distinct): BFS state depth 0 / parent = the root itself
// (original id), SSSP distance 0; every root is set in bm_cur (BFS frontier;
// for SSSP the set the first frontier build starts from) and, for BFS, in
// bm_visited; the activity of the root level (summed out-degree) is counted.
__global__ void k_init_roots(const uint32_t* __restrict__ roots, uint32_t n, int app, int wkind, void* __restrict__ prop,
                             const uint32_t* __restrict__ perm, uint64_t* __restrict__ cur, uint64_t* __restrict__ visited,
                             const uint32_t* __restrict__ od, wg_dev_counters* __restrict__ ctr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t r = roots[i];
    const unsigned long long bit = 1ull << (r & 63);
    atomicOr((unsigned long long*)&cur[r >> 6], bit);
    if (app == WG_APP_BFS) {
        atomicOr((unsigned long long*)&visited[r >> 6], bit);
        ((uint64_t*)prop)[r] = perm ? perm[r] : r;
    } else if (wkind == WG_WEIGHT_F64) {
        ((uint64_t*)prop)[r] = 0ull;
    } else {
        ((uint32_t*)prop)[r] = 0u;
    }
    atomicAdd(&ctr->slots[0].v[WG_CNT_ACTIVITY], (unsigned long long)od[r]);
}

static int init_roots(wg_ctx* ctx, int app, const uint64_t* roots, uint32_t n)
{
    if (!ctx || !roots || n == 0)
        return WG_E_INVAL;
    if (!ctx->off || (app == WG_APP_SSSP && ctx->wkind == WG_WEIGHT_NONE))
        return WG_E_STATE;
    std::vector<uint32_t> dev(n);
    int rc;
    if ((rc = wg_to_device_ids(ctx, roots, n, dev.data())))
        return rc;
    std::sort(dev.begin(), dev.end());
    dev.erase(std::unique(dev.begin(), dev.end()), dev.end());
    if ((rc = init_common(ctx, app)))
        return rc;
    const size_t esz = (app == WG_APP_BFS || ctx->wkind == WG_WEIGHT_F64) ? 8 : 4;
    WG_HIP(ctx, hipMalloc(&ctx->prop[0], ctx->V * esz));
    WG_HIP(ctx, hipMalloc((void**)&ctx->roots, dev.size() * sizeof(uint32_t)));
    ctx->nroots = (uint32_t)dev.size();
    ctx->root_dev = dev[0];
    WG_HIP(ctx, hipMemcpyAsync(ctx->roots, dev.data(), dev.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    if (esz == 8)
        rc = wg_fill_u64(ctx, (uint64_t*)ctx->prop[0], ctx->V, app == WG_APP_BFS ? ~0ull : (uint64_t)PolSF::INF);
    else
        rc = wg_fill_u32(ctx, (uint32_t*)ctx->prop[0], ctx->V, PolSU::INF);
    if (rc || (rc = wg_reset_counters(ctx, RST_ACTIVITY)))
        return rc;
    hipLaunchKernelGGL(k_init_roots, dim3((ctx->nroots + 255) / 256), dim3(256), 0, ctx->stream, ctx->roots, ctx->nroots,
                       app, ctx->wkind, ctx->prop[0], ctx->perm, ctx->bm_cur, ctx->bm_visited, ctx->od, ctx->ctr);
    if ((rc = wg_check_launch(ctx)) || (rc = wg_read_counters(ctx)))
        return rc;
    ctx->root_activity = ctx->hctr->activity;
    ctx->roots_frontier = app == WG_APP_BFS;
    if (app == WG_APP_BFS && ctx->nroots <= WG_PUSH_ROOTS_MAX) { // for the first level's push
        WG_HIP(ctx, hipMalloc((void**)&ctx->root_off, (ctx->nroots + 1) * sizeof(uint32_t)));
        hipLaunchKernelGGL(k_root_offsets, dim3(1), dim3(WG_BLOCK), 0, ctx->stream, ctx->roots, ctx->nroots,
                           ctx->csri_off, ctx->root_off);
        if ((rc = wg_check_launch(ctx)))
            return rc;
    }
    return WG_OK;
}

extern "C" int wg_init_bfs_roots(wg_ctx* ctx, const uint64_t* roots, uint32_t n)
{
    return init_roots(ctx, WG_APP_BFS, roots, n);
}

extern "C" int wg_init_bfs(wg_ctx* ctx, uint64_t root)
{
    return init_roots(ctx, WG_APP_BFS, &root, 1);
}

extern "C" int wg_init_cc(wg_ctx* ctx)
{
    int rc = init_common(ctx, WG_APP_CC);
    if (rc)
        return rc;
    WG_HIP(ctx, hipMalloc(&ctx->prop[0], ctx->V * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_iota_u32, dim3(wg_grid(ctx, ctx->V, 1024, 8)), dim3(256), 0, ctx->stream, (uint32_t*)ctx->prop[0],
                       ctx->V, ctx->perm);
    if ((rc = wg_check_launch(ctx)))
        return rc;
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return WG_OK;
}

extern "C" int wg_init_sssp_roots(wg_ctx* ctx, const uint64_t* roots, uint32_t n)
{
    return init_roots(ctx, WG_APP_SSSP, roots, n);
}

extern "C" int wg_init_sssp(wg_ctx* ctx, uint64_t root)
{
    return init_roots(ctx, WG_APP_SSSP, &root, 1);
}

// Launch shapes.  Frontier pulls: one wavefront per 64 frontier words at the
// finest granularity.  Frontier build: one wavefront per 64 bitmask words.
// End-of-iteration bitmask kernels: one thread per bitmask word.
uint32_t wg_frontier_grid(const wg_ctx* ctx)
{
    return wg_grid(ctx, ctx->wedge_words, WG_WPB * WG_WAVE, 8);
}

uint32_t wg_wedge_build_grid(const wg_ctx* ctx)
{
    return wg_grid(ctx, ctx->bm_words, WG_WPB * WG_WAVE, 8);
}

uint32_t wg_advance_grid(const wg_ctx* ctx)
{
    return wg_grid(ctx, ctx->bm_words, 256, 32);
}

static int check_block_log2(int b)
{
    return b >= WG_MIN_BLOCK_LOG2 && b <= WG_MAX_BLOCK_LOG2;
}

// A marked frontier must be consumed with the granularity it was built at.
static int prepare_frontier_build(wg_ctx* ctx, int blog2)
{
    if (ctx->wedge_log2 >= 0 && ctx->wedge_log2 != blog2) {
        int rc = clear_wedge_state(ctx);
        if (rc)
            return rc;
    }
    ctx->wedge_log2 = blog2;
    return WG_OK;
}

extern "C" int wg_wedge_activate_single_vertex(wg_ctx* ctx, uint64_t vertex, int block_log2)
{
    if (!ctx || ctx->app == WG_APP_NONE || ctx->app == WG_APP_PR)
        return WG_E_STATE;
    if (!check_block_log2(block_log2))
        return WG_E_INVAL;
    uint32_t dv = 0;
    int rc = wg_to_device_id(ctx, vertex, &dv);
    if (rc || (rc = prepare_frontier_build(ctx, block_log2)))
        return rc;
    int slot = wg_prof_begin(ctx, WG_K_WEDGE_BUILD);
    hipLaunchKernelGGL(k_wedge_single, dim3(16), dim3(256), 0, ctx->stream, dv, ctx->csri_off, ctx->csri_pos,
                       block_log2, ctx->wedge, ctx->ctr, WG_GATE_ALWAYS);
    rc = wg_check_launch(ctx);
    wg_prof_end(ctx, WG_K_WEDGE_BUILD, slot);
    return rc;
}

// Marks the blocks of every out-edge of the vertices set in `active` (V bits)
// at granularity 2^block_log2 (no granularity bookkeeping: callers track it).
int wg_launch_wedge_build(wg_ctx* ctx, const uint64_t* active, int block_log2)
{
    int rc = wg_reset_counters(ctx, RST_HEAVY);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_wedge_build, dim3(wg_wedge_build_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream, active,
                       ctx->bm_words, ctx->csri_off, ctx->csri_pos, block_log2, ctx->wedge, ctx->heavy_list, ctx->heavy_cap, ctx->ctr,
                       WG_GATE_ALWAYS);
    if ((rc = wg_check_launch(ctx)))
        return rc;
    hipLaunchKernelGGL(k_wedge_heavy, dim3(ctx->num_cu * 2), dim3(WG_BLOCK), 0, ctx->stream, ctx->heavy_list,
                       ctx->heavy_cap, ctx->csri_off, ctx->csri_pos, block_log2, ctx->wedge, ctx->ctr, WG_GATE_ALWAYS);
    return wg_check_launch(ctx);
}

extern "C" int wg_wedge_generate_frontier(wg_ctx* ctx, int block_log2)
{
    if (!ctx || ctx->app == WG_APP_NONE || ctx->app == WG_APP_PR)
        return WG_E_STATE;
    if (!check_block_log2(block_log2))
        return WG_E_INVAL;
    int rc = prepare_frontier_build(ctx, block_log2);
    if (rc || (rc = wg_reset_counters(ctx, RST_HEAVY)))
        return rc;
    int slot = wg_prof_begin(ctx, WG_K_WEDGE_BUILD);
    hipLaunchKernelGGL(k_wedge_build, dim3(wg_wedge_build_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream, ctx->bm_next,
                       ctx->bm_words, ctx->csri_off, ctx->csri_pos, block_log2, ctx->wedge, ctx->heavy_list, ctx->heavy_cap, ctx->ctr,
                       WG_GATE_ALWAYS);
    if ((rc = wg_check_launch(ctx)))
        return rc;
    hipLaunchKernelGGL(k_wedge_heavy, dim3(ctx->num_cu * 2), dim3(WG_BLOCK), 0, ctx->stream, ctx->heavy_list,
                       ctx->heavy_cap, ctx->csri_off, ctx->csri_pos, block_log2, ctx->wedge, ctx->ctr, WG_GATE_ALWAYS);
    rc = wg_check_launch(ctx);
    wg_prof_end(ctx, WG_K_WEDGE_BUILD, slot);
    return rc;
}

static int finish_pull(wg_ctx* ctx, uint64_t* edges_activated)
{
    ctx->iter += 1;
    ctx->roots_frontier = 0;
    if (!edges_activated)
        return WG_OK;
    int rc = wg_read_counters(ctx);
    if (rc)
        return rc;
    *edges_activated = ctx->hctr->activity;
    if (ctx->app == WG_APP_SSSP && ctx->hctr->overflow)
        return WG_E_OVERFLOW;
    return WG_OK;
}

extern "C" int wg_last_pull_counts(wg_ctx* ctx, uint64_t* edges, uint64_t* vertices, uint64_t* newly_reached)
{
    if (!ctx)
        return WG_E_INVAL;
    if (edges)
        *edges = ctx->hctr->activity;
    if (vertices)
        *vertices = ctx->hctr->activated;
    if (newly_reached)
        *newly_reached = ctx->hctr->fresh;
    return WG_OK;
}

// A dense pull covers every edge: any marked blocks are dropped.
static int begin_pull(wg_ctx* ctx, int mode)
{
    int rc = wg_reset_counters(ctx, RST_ACTIVITY);
    if (rc)
        return rc;
    if (mode == WG_MODE_DENSE && ctx->wedge_log2 >= 0)
        rc = clear_wedge_state(ctx);
    return rc;
}

static int end_frontier_pull(wg_ctx* ctx)
{
    // the pull zeroed every frontier word it consumed
    ctx->wedge_log2 = -1;
    return wg_reset_counters(ctx, RST_HEAVY);
}

extern "C" int wg_edge_pull_bfs(wg_ctx* ctx, int mode, uint64_t* edges_activated)
{
    if (!ctx || ctx->app != WG_APP_BFS)
        return WG_E_STATE;
    if (mode != WG_MODE_DENSE && mode != WG_MODE_FRONTIER)
        return WG_E_INVAL;
    if (mode == WG_MODE_FRONTIER && ctx->wedge_log2 < 0)
        return WG_E_STATE;
    int rc = begin_pull(ctx, mode);
    if (rc)
        return rc;
    uint64_t* packed = (uint64_t*)ctx->prop[0];
    if (mode == WG_MODE_DENSE) {
        int slot = wg_prof_begin(ctx, WG_K_BFS_DENSE);
        hipLaunchKernelGGL(WG_CW(ctx, k_bfs_dense), dim3(wg_grid(ctx, ctx->bm_words, WG_BLOCK / WG_WAVE, 8)), dim3(WG_BLOCK), 0,
                           ctx->stream, ctx->off, ctx->src, ctx->first_src, ctx->od, ctx->bm_cur, ctx->bm_visited,
                           ctx->bm_next, packed, (uint32_t)ctx->V, ctx->perm, ctx->ctr, WG_GATE_ALWAYS);
        if ((rc = wg_check_launch(ctx)))
            return rc;
        wg_prof_end(ctx, WG_K_BFS_DENSE, slot);
        if (ctx->nbfs_heavy) {
            slot = wg_prof_begin(ctx, WG_K_BFS_HEAVY);
            hipLaunchKernelGGL(WG_CW(ctx, k_bfs_heavy), dim3(wg_grid(ctx, ctx->nbfs_heavy, WG_BLOCK / WG_WAVE, 8)), dim3(WG_BLOCK), 0,
                               ctx->stream, ctx->bfs_heavy_v, ctx->bfs_heavy_e, ctx->nbfs_heavy, ctx->off, ctx->src,
                               ctx->od, ctx->bm_cur, ctx->bm_visited, ctx->bm_next, packed, ctx->perm, ctx->ctr,
                               WG_GATE_ALWAYS);
            if ((rc = wg_check_launch(ctx)))
                return rc;
            wg_prof_end(ctx, WG_K_BFS_HEAVY, slot);
        }
    } else {
        int slot = wg_prof_begin(ctx, WG_K_BFS_BLOCKS);
        hipLaunchKernelGGL(WG_CW(ctx, k_bfs_blocks), dim3(wg_frontier_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream, ctx->ctr,
                           ctx->wedge_log2, (uint32_t)ctx->E, ctx->src, ctx->dst, ctx->od, ctx->bm_cur, ctx->bm_visited,
                           ctx->bm_next, packed, ctx->wedge, ctx->perm, WG_GATE_ALWAYS);
        if ((rc = wg_check_launch(ctx)))
            return rc;
        wg_prof_end(ctx, WG_K_BFS_BLOCKS, slot);
        if ((rc = end_frontier_pull(ctx)))
            return rc;
    }
    return finish_pull(ctx, edges_activated);
}

// Vertices whose values a dense pull reads from LDS: the first ones, if the
// vertices are in degree order.
template <class P>
static uint32_t min_hub_count(const wg_ctx* ctx)
{
    const uint64_t cap = WG_MIN_HUB_BYTES / sizeof(typename P::T);
    return (P::HUB && ctx->perm) ? (uint32_t)(ctx->V < cap ? ctx->V : cap) : 0u;
}

// Tiles of the dense pulls of policy P: 256 edges (4 per lane) if
// P::DENSE_IPT asks for them and the finer tilings exist (built by the SSSP
// init), else the default 512.
struct min_tiling {
    const uint32_t *tile_v, *tile_e;
    uint32_t ntiles;
    int ipt;
};

template <class P>
static min_tiling dense_min_tiling(const wg_ctx* ctx)
{
    if (P::DENSE_IPT == 4 && ctx->pr_alt_built) {
        const wg_tiling& t = ctx->pr_alt[0];
        return min_tiling{t.tile_v, t.tile_e, t.ntiles, 4};
    }
    return min_tiling{ctx->tile_v, ctx->tile_e, ctx->ntiles, WG_IPT};
}

template <class P>
static int min_pull(wg_ctx* ctx, int mode, void* prop)
{
    int rc;
    if (mode == WG_MODE_DENSE && wg_pb_min_on(ctx)) {
        if ((rc = wg_pb_min_prepare(ctx)))
            return rc;
        int slot = wg_prof_begin(ctx, WG_K_MIN_TILE);
        if ((rc = wg_pb_min_pull(ctx, WG_GATE_ALWAYS)))
            return rc;
        wg_prof_end(ctx, WG_K_MIN_TILE, slot);
    } else if (mode == WG_MODE_DENSE) {
        int slot = wg_prof_begin(ctx, WG_K_MIN_TILE);
        const min_tiling mt = dense_min_tiling<P>(ctx);
        if (mt.ipt == 4)
            hipLaunchKernelGGL(WG_CW(ctx, k_min_tile, P, 4), dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->src,
                               ctx->w, (typename P::T*)prop, ctx->od, mt.tile_v, mt.tile_e, mt.ntiles, ctx->bm_next,
                               ctx->ctr, &ctx->ctr->overflow, WG_GATE_ALWAYS, min_hub_count<P>(ctx));
        else
            hipLaunchKernelGGL(WG_CW(ctx, k_min_tile, P), dim3(ctx->pr_grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->src,
                               ctx->w, (typename P::T*)prop, ctx->od, mt.tile_v, mt.tile_e, mt.ntiles, ctx->bm_next,
                               ctx->ctr, &ctx->ctr->overflow, WG_GATE_ALWAYS, min_hub_count<P>(ctx));
        if ((rc = wg_check_launch(ctx)))
            return rc;
        wg_prof_end(ctx, WG_K_MIN_TILE, slot);
    } else {
        int slot = wg_prof_begin(ctx, WG_K_MIN_BLOCKS);
        hipLaunchKernelGGL(WG_CW(ctx, k_min_blocks, P), dim3(wg_frontier_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream, ctx->ctr,
                           ctx->wedge_log2, (uint32_t)ctx->E, ctx->src, ctx->dst, ctx->w, (typename P::T*)prop, ctx->od,
                           ctx->bm_next, ctx->wedge, &ctx->ctr->overflow, WG_GATE_ALWAYS);
        if ((rc = wg_check_launch(ctx)))
            return rc;
        wg_prof_end(ctx, WG_K_MIN_BLOCKS, slot);
        if ((rc = end_frontier_pull(ctx)))
            return rc;
    }
    return WG_OK;
}

extern "C" int wg_edge_pull_cc(wg_ctx* ctx, int mode, uint64_t* edges_activated)
{
    if (!ctx || ctx->app != WG_APP_CC)
        return WG_E_STATE;
    if (mode != WG_MODE_DENSE && mode != WG_MODE_FRONTIER)
        return WG_E_INVAL;
    if (mode == WG_MODE_FRONTIER && ctx->wedge_log2 < 0)
        return WG_E_STATE;
    int rc = begin_pull(ctx, mode);
    if (rc || (rc = min_pull<PolCC>(ctx, mode, ctx->prop[0])))
        return rc;
    return finish_pull(ctx, edges_activated);
}

extern "C" int wg_edge_pull_sssp(wg_ctx* ctx, int mode, uint64_t* edges_activated)
{
    if (!ctx || ctx->app != WG_APP_SSSP)
        return WG_E_STATE;
    if (mode != WG_MODE_DENSE && mode != WG_MODE_FRONTIER)
        return WG_E_INVAL;
    if (mode == WG_MODE_FRONTIER && ctx->wedge_log2 < 0)
        return WG_E_STATE;
    int rc = begin_pull(ctx, mode);
    if (rc)
        return rc;
    rc = (ctx->wkind == WG_WEIGHT_F64) ? min_pull<PolSF>(ctx, mode, ctx->prop[0]) : min_pull<PolSU>(ctx, mode, ctx->prop[0]);
    if (rc)
        return rc;
    return finish_pull(ctx, edges_activated);
}

extern "C" int wg_advance_frontier(wg_ctx* ctx, uint64_t* frontier_vertices)
{
    if (!ctx || ctx->app == WG_APP_NONE || ctx->app == WG_APP_PR)
        return WG_E_STATE;
    ctx->roots_frontier = 0;
    int rc = wg_reset_counters(ctx, RST_FRONTIER);
    if (rc)
        return rc;
    int slot = wg_prof_begin(ctx, WG_K_ADVANCE);
    const uint32_t grid = wg_advance_grid(ctx);
    if (ctx->app == WG_APP_BFS)
        hipLaunchKernelGGL(k_bfs_advance, dim3(grid), dim3(256), 0, ctx->stream, ctx->bm_visited, ctx->bm_next,
                           ctx->bm_cur, ctx->bm_words, ctx->ctr, WG_GATE_ALWAYS);
    else
        hipLaunchKernelGGL(k_count_clear, dim3(grid), dim3(256), 0, ctx->stream, ctx->bm_next, ctx->bm_words, ctx->ctr,
                           WG_GATE_ALWAYS);
    if ((rc = wg_check_launch(ctx)))
        return rc;
    wg_prof_end(ctx, WG_K_ADVANCE, slot);
    if (frontier_vertices) {
        if ((rc = wg_read_counters(ctx)))
            return rc;
        *frontier_vertices = ctx->hctr->frontier_verts;
    }
    return WG_OK;
}

static int read_prop(wg_ctx* ctx, int app, void* host, size_t esz)
{
    if (!ctx || ctx->app != app || !host)
        return WG_E_STATE;
    return wg_read_vertex_array(ctx, ctx->prop[0], host, esz);
}

extern "C" int wg_read_bfs(wg_ctx* ctx, uint64_t* host_packed)
{
    return read_prop(ctx, WG_APP_BFS, host_packed, sizeof(uint64_t));
}

extern "C" int wg_read_cc(wg_ctx* ctx, uint32_t* host_labels)
{
    return read_prop(ctx, WG_APP_CC, host_labels, sizeof(uint32_t));
}

extern "C" int wg_read_sssp(wg_ctx* ctx, void* host_dist)
{
    if (!ctx)
        return WG_E_INVAL;
    return read_prop(ctx, WG_APP_SSSP, host_dist, ctx->wkind == WG_WEIGHT_F64 ? sizeof(uint64_t) : sizeof(uint32_t));
}

extern "C" int wg_work_counts(wg_ctx* ctx, uint64_t* edges_read, uint64_t* dests_examined)
{
    if (!ctx || !ctx->count_work || (ctx->app != WG_APP_BFS && ctx->app != WG_APP_SSSP && ctx->app != WG_APP_CC))
        return WG_E_STATE;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const int rc = wg_read_counters(ctx);
    if (rc)
        return rc;
    if (edges_read)
        *edges_read = ctx->hctr->edges_read;
    if (dests_examined)
        *dests_examined = ctx->hctr->dests_examined;
    return WG_OK;
}

extern "C" int wg_reached(wg_ctx* ctx, uint64_t* vertices, uint64_t* edges)
{
    if (!ctx || (ctx->app != WG_APP_BFS && ctx->app != WG_APP_SSSP && ctx->app != WG_APP_CC))
        return WG_E_STATE;
    WG_HIP(ctx, hipMemsetAsync(&ctx->ctr->reached_verts, 0, 2 * sizeof(unsigned long long), ctx->stream));
    const uint32_t grid = wg_grid(ctx, ctx->V, 1024, 8);
    if (ctx->app == WG_APP_BFS)
        hipLaunchKernelGGL(k_reached_bfs, dim3(grid), dim3(256), 0, ctx->stream, (const uint64_t*)ctx->prop[0], ctx->od,
                           ctx->V, ctx->ctr);
    else if (ctx->app == WG_APP_CC)
        hipLaunchKernelGGL(k_reached_min<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream, (const uint32_t*)ctx->prop[0],
                           0xffffffffu, ctx->od, ctx->V, ctx->ctr);
    else if (ctx->wkind == WG_WEIGHT_F64)
        hipLaunchKernelGGL(k_reached_min<unsigned long long>, dim3(grid), dim3(256), 0, ctx->stream,
                           (const unsigned long long*)ctx->prop[0], PolSF::INF, ctx->od, ctx->V, ctx->ctr);
    else
        hipLaunchKernelGGL(k_reached_min<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream, (const uint32_t*)ctx->prop[0],
                           PolSU::INF, ctx->od, ctx->V, ctx->ctr);
    int rc = wg_check_launch(ctx);
    if (rc || (rc = wg_read_counters(ctx)))
        return rc;
    if (vertices)
        *vertices = ctx->hctr->reached_verts;
    if (edges)
        *edges = ctx->hctr->reached_edges;
    return WG_OK;
}

// One iteration of the device-resident loop: the level pull (BFS dense: with
// the hub destinations' chunks), the decision, [a fused run], the frontier
// build (gated) and the advance (with the build's deferred hub out-lists).
static int enqueue_iteration(wg_ctx* ctx)
{
    uint64_t* packed = (uint64_t*)ctx->prop[0];
    const uint32_t qgrid = wg_frontier_grid(ctx);
    const float* mlp = ctx->mlp_set ? ctx->mlp : nullptr;
    if (ctx->app == WG_APP_BFS) {
        const uint32_t dgrid = wg_grid(ctx, ctx->bm_words, WG_BLOCK / WG_WAVE, 8);
        hipLaunchKernelGGL(WG_CW(ctx, k_bfs_level), dim3(dgrid > qgrid ? dgrid : qgrid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off,
                           ctx->src, ctx->dst, ctx->od, ctx->bm_cur, ctx->bm_visited, ctx->bm_next, packed,
                           (uint32_t)ctx->V, (uint32_t)ctx->E, ctx->wedge, ctx->perm, ctx->ctr, ctx->bfs_heavy_v,
                           ctx->bfs_heavy_e, ctx->nbfs_heavy, ctx->first_src, ctx->roots, ctx->push_roots,
                           ctx->root_off, ctx->csri_off, ctx->csr_dst);
    } else {
        const uint32_t grid = ctx->pr_grid > qgrid ? ctx->pr_grid : qgrid;
        const int binned = wg_pb_min_on(ctx) ? 1 : 0; // the layout was built by wg_converge
        if (binned) {
            int rc = wg_pb_min_pull(ctx, WG_MODE_DENSE);
            if (rc)
                return rc;
        }
#define WG_MIN_LEVEL_IPT(P, T, IPT)                                                                                   \
    hipLaunchKernelGGL(WG_CW(ctx, k_min_level, P, IPT), dim3(grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->src, ctx->dst, \
                       ctx->w, (T*)ctx->prop[0], ctx->od, mt.tile_v, mt.tile_e, mt.ntiles, (uint32_t)ctx->E,            \
                       ctx->bm_next, ctx->wedge, ctx->ctr, &ctx->ctr->overflow, min_hub_count<P>(ctx), binned)
#define WG_MIN_LEVEL(P, T)                          \
    do {                                            \
        const min_tiling mt = dense_min_tiling<P>(ctx); \
        if (mt.ipt == 4)                            \
            WG_MIN_LEVEL_IPT(P, T, 4);              \
        else                                        \
            WG_MIN_LEVEL_IPT(P, T, WG_IPT);         \
    } while (0)
        if (ctx->app == WG_APP_CC)
            WG_MIN_LEVEL(PolCC, uint32_t);
        else if (ctx->wkind == WG_WEIGHT_F64)
            WG_MIN_LEVEL(PolSF, unsigned long long);
        else
            WG_MIN_LEVEL(PolSU, uint32_t);
#undef WG_MIN_LEVEL
#undef WG_MIN_LEVEL_IPT
    }
    const uint32_t run_wgs = ctx->run_wgs[ctx->app & 7];
    const bool run = run_wgs && ctx->run_q && (ctx->app == WG_APP_BFS || ctx->run_spare);
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(WG_WAVE), 0, ctx->stream, ctx->ctr, mlp, run ? ctx->run_ctl : nullptr);
    if (run) {
        const dim3 rg(run_wgs), rb(WG_BLOCK);
        uint4* qb = ctx->run_q + WG_RUN_QCAP;
        if (ctx->app == WG_APP_BFS) {
            hipLaunchKernelGGL(WG_CW(ctx, k_bfs_run), rg, rb, 0, ctx->stream, ctx->csri_off, ctx->csr_dst, ctx->bm_next,
                               ctx->bm_visited, packed, ctx->bm_words, ctx->perm, ctx->run_q, qb, ctx->run_ctl, ctx->ctr,
                               mlp);
        } else {
#define WG_MIN_RUN(P, T)                                                                                               \
    hipLaunchKernelGGL(WG_CW(ctx, k_min_run, P), rg, rb, 0, ctx->stream, ctx->csri_off, ctx->csri_pos, ctx->dst, ctx->w,            \
                       (T*)ctx->prop[0], ctx->bm_next, ctx->run_spare, ctx->bm_words, ctx->run_q, qb, ctx->run_ctl,     \
                       ctx->ctr, &ctx->ctr->overflow, mlp)
            if (ctx->app == WG_APP_CC)
                WG_MIN_RUN(PolCC, uint32_t);
            else if (ctx->wkind == WG_WEIGHT_F64)
                WG_MIN_RUN(PolSF, unsigned long long);
            else
                WG_MIN_RUN(PolSU, uint32_t);
#undef WG_MIN_RUN
        }
    }
    hipLaunchKernelGGL(k_wedge_build, dim3(wg_wedge_build_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream, ctx->bm_next,
                       ctx->bm_words, ctx->csri_off, ctx->csri_pos, 0, ctx->wedge, ctx->heavy_list, ctx->heavy_cap, ctx->ctr,
                       WG_GATE_FRONTIER);
    hipLaunchKernelGGL(k_level_advance, dim3(wg_advance_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream, ctx->bm_visited,
                       ctx->bm_next, ctx->bm_cur, ctx->bm_words, ctx->app == WG_APP_BFS ? 1 : 0, ctx->heavy_list,
                       ctx->heavy_cap, ctx->csri_off, ctx->csri_pos, ctx->wedge, ctx->ctr);
    return wg_check_launch(ctx);
}

extern "C" int wg_set_fused_levels(wg_ctx* ctx, uint32_t workgroups, uint32_t apps)
{
    if (!ctx)
        return WG_E_INVAL;
    if (workgroups > WG_RUN_MAX_WGS || workgroups > (uint32_t)ctx->num_cu)
        return WG_E_INVAL;
    if (apps & ~((1u << WG_APP_BFS) | (1u << WG_APP_CC) | (1u << WG_APP_SSSP)))
        return WG_E_INVAL;
    for (int a = 0; a < 8; ++a)
        if ((apps >> a) & 1u)
            ctx->run_wgs[a] = workgroups;
    return WG_OK;
}

extern "C" int wg_fused_level_count(wg_ctx* ctx, uint64_t* levels)
{
    if (!ctx || !levels)
        return WG_E_INVAL;
    *levels = ctx->hctr->run_levels;
    return WG_OK;
}

extern "C" int wg_converge(wg_ctx* ctx, const wg_loop_params* p, uint64_t* iterations, uint64_t* frontier_iterations)
{
    if (!ctx || !p)
        return WG_E_INVAL;
    if (ctx->app != WG_APP_BFS && ctx->app != WG_APP_CC && ctx->app != WG_APP_SSSP)
        return WG_E_STATE;
    if (p->schedule < 0 || p->schedule > 3)
        return WG_E_INVAL;
    const int blog2 = p->block_log2 ? p->block_log2 : WG_DEFAULT_BLOCK_LOG2;
    if (!check_block_log2(blog2))
        return WG_E_INVAL;
    int rc;
    if (ctx->wedge_log2 >= 0 && (rc = clear_wedge_state(ctx)))
        return rc;
    if (wg_pb_min_on(ctx) && (rc = wg_pb_min_prepare(ctx))) // dense iterations on the binned layout
        return rc;
    if (ctx->run_wgs[ctx->app & 7]) {
        if (!ctx->run_q) {
            WG_HIP(ctx, hipMalloc((void**)&ctx->run_ctl, sizeof(wg_run_ctl)));
            WG_HIP(ctx, hipMalloc((void**)&ctx->run_q, 2 * (size_t)WG_RUN_QCAP * sizeof(uint4)));
        }
        WG_HIP(ctx, hipMemsetAsync(ctx->run_ctl, 0, sizeof(wg_run_ctl), ctx->stream));
        if (ctx->app != WG_APP_BFS) {
            if (!ctx->run_spare)
                WG_HIP(ctx, hipMalloc((void**)&ctx->run_spare, ctx->bm_words * sizeof(uint64_t)));
            if ((rc = wg_clear_u64(ctx, ctx->run_spare, ctx->bm_words)))
                return rc;
        }
    }
    const unsigned long long threshold = (ctx->E * (unsigned long long)p->threshold_pct) / 100ull;
    // CC: source/execution_cc.c:95; BFS / SSSP: out-degree of the roots
    const unsigned long long act0 = ctx->app == WG_APP_CC ? ctx->E : ctx->root_activity;
    hipLaunchKernelGGL(k_loop_init, dim3(1), dim3(WG_WAVE), 0, ctx->stream, ctx->ctr, ctx->mlp_set ? ctx->mlp : nullptr,
                       (unsigned int)p->schedule, threshold, blog2, p->max_iterations, act0,
                       ctx->app == WG_APP_CC ? (unsigned long long)ctx->V : (unsigned long long)ctx->nroots,
                       ctx->app == WG_APP_CC ? 1 : 0,
                       ctx->app == WG_APP_CC ? 0ull : (unsigned long long)(ctx->V - ctx->nroots), (unsigned int)ctx->app,
                       (unsigned long long)ctx->V, (unsigned long long)ctx->E, wg_pb_min_on(ctx) ? 1u : 0u);
    // A BFS that starts from its roots (at most WG_PUSH_ROOTS_MAX, whose
    // out-edge offsets the init computed) pushes its first level along the
    // roots' out-edges if that level is frontier-filtered (k_bfs_level), so
    // no blocks are marked for it.
    ctx->push_roots =
        (ctx->app == WG_APP_BFS && ctx->roots_frontier && ctx->root_off) ? ctx->nroots : 0u;
    ctx->roots_frontier = 0;
    if (ctx->app != WG_APP_CC) { // the roots' blocks, if the first pull is frontier-filtered
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->ctr, RST_HEAVY);
        if (!ctx->push_roots) {
            hipLaunchKernelGGL(k_wedge_build, dim3(wg_wedge_build_grid(ctx)), dim3(WG_BLOCK), 0, ctx->stream,
                               ctx->bm_cur, ctx->bm_words, ctx->csri_off, ctx->csri_pos, 0, ctx->wedge, ctx->heavy_list,
                               ctx->heavy_cap, ctx->ctr, WG_GATE_FRONTIER);
            hipLaunchKernelGGL(k_wedge_heavy, dim3(ctx->num_cu * 2), dim3(WG_BLOCK), 0, ctx->stream, ctx->heavy_list,
                               ctx->heavy_cap, ctx->csri_off, ctx->csri_pos, 0, ctx->wedge, ctx->ctr, WG_GATE_FRONTIER);
        }
    }
    if ((rc = wg_check_launch(ctx)))
        return rc;
    int slot = wg_prof_begin(ctx, WG_K_LOOP);
    // Every iteration either activates a vertex or ends the loop, and a
    // vertex can only improve finitely often, but the host loop does not rely
    // on that: it stops with an error if the device stops counting iterations
    // or the count passes 2V + 64.
    const uint64_t cap = 2 * ctx->V + 64;
    uint32_t batch = 2;
    uint32_t seen = 0;
    for (;;) {
        for (uint32_t i = 0; i < batch; ++i)
            if ((rc = enqueue_iteration(ctx)))
                return rc;
        WG_HIP(ctx, hipMemcpyAsync(ctx->hctr, ctx->ctr, WG_CTR_HOST_BYTES, hipMemcpyDeviceToHost, ctx->stream));
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->hctr->run_err)
            return WG_E_STATE;
        if (ctx->hctr->done)
            break;
        if (ctx->hctr->iters <= seen || ctx->hctr->iters > cap)
            return WG_E_STATE;
        seen = ctx->hctr->iters;
        batch = batch < 64 ? batch * 2 : 64;
    }
    wg_prof_end(ctx, WG_K_LOOP, slot);
    ctx->wedge_log2 = -1;
    ctx->iter += (int)ctx->hctr->iters;
    if (iterations)
        *iterations = ctx->hctr->iters;
    if (frontier_iterations)
        *frontier_iterations = ctx->hctr->fiters;
    if (ctx->app == WG_APP_SSSP && ctx->hctr->overflow)
        return WG_E_OVERFLOW;
    return WG_OK;
}
