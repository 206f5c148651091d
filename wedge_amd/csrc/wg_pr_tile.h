// wg_pr_tile.h -- the PageRank edge pull over edge tiles, shared by the
// single-graph PageRank and the batches of member graphs.
//
// Reference: perform_edge_pull_phase_pr (source/phase_edge_pull_pr.c:28-116)
// and the merge of destinations split across work units
// (source/phases.c:69-114).
//
// IPT is the number of edges per lane: a tile holds at most 64 * IPT edges.
// The tile tables (tile_v / tile_e / heavy_v / heavy_t, built by
// wg_build_tiles) must have been built for the same IPT.
#pragma once

#include "wg_internal.h"

// XCD-aware tile order.  Workgroups are dispatched to the 8 XCDs round-robin
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  With the
// split on, XCD x walks only the tiles [start[x], start[x + 1]) (ranges of
// about E/8 edges each), so the sources its tiles gather from stay within a
// few neighbouring destinations' in-neighbourhoods -- for a batch of small
// graphs, within the members that XCD is working through, whose
// contributions fit its L2.  Off: all workgroups stride over all tiles.
#define WG_NXCD 8
struct wg_tile_split {
    uint32_t on;
    uint32_t start[WG_NXCD + 1];
};

// Deterministic block-wide sum of a block of WPB wavefronts (xor butterfly
// inside each wave, then the wave sums in wave order).  All threads receive
// the result.
template <int WPB = WG_WPB>
static __device__ __forceinline__ double block_sum_det(double v, double* s_red)
{
    v = wg_wave_sum(v);
    const int wave = threadIdx.x / WG_WAVE;
    __syncthreads();
    if ((threadIdx.x & (WG_WAVE - 1)) == 0)
        s_red[wave] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < WPB; ++i)
        r += s_red[i];
    return r;
}

template <typename T>
static __device__ __forceinline__ double load_contrib(const T* __restrict__ p, uint32_t i)
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
//
// HUB: the contributions of vertices [0, nhub) (nhub <= wg_pr_hub_max<T,
// WPB>()) are copied to LDS first and gathered from there.  After the degree
// relabelling these are the highest-degree vertices, the sources of a large
// share of the edges (rmat-s24: 11.6% for the first 2048, 15% for the first
// 4096, 21% for the first 8192), whose gathers then skip the L2.  The cache
// is 16 KB per 4 wavefronts: 16 KB for the 4-wave workgroups (four per CU;
// 32 and 64 KB there measured slower, fewer resident waves), 64 KB for a
// 16-wave workgroup, one per CU (WPB = 16).
#define WG_PR_HUB_BYTES 16384u
#define WG_PR_BIG_WPB 16 // wavefronts of the one-per-CU workgroup (shape blocks_per_cu 1)
#define WG_PR_BIG_HUB_BYTES 90112u // what the 16-wave workgroup's 160 KB of LDS leaves
template <typename T, int WPB>
__host__ __device__ constexpr uint32_t wg_pr_hub_max()
{
    return (WPB == WG_PR_BIG_WPB ? WG_PR_BIG_HUB_BYTES : WG_PR_HUB_BYTES * (WPB / WG_WPB)) / (uint32_t)sizeof(T);
}
template <typename T, int MODE, int IPT = WG_IPT, bool HUB = false, int WPB = WG_WPB>
__global__ __launch_bounds__(WPB * WG_WAVE) void k_pr_tile(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                       const T* __restrict__ cin, T* __restrict__ cout,
                                                       double* __restrict__ acc_out, const uint32_t* __restrict__ od,
                                                       const uint32_t* __restrict__ tile_v,
                                                       const uint32_t* __restrict__ tile_e, uint32_t ntiles,
                                                       const double* __restrict__ S_in, double damping, double N,
                                                       double* __restrict__ tile_part, double* __restrict__ blk_part,
                                                       wg_tile_split split = wg_tile_split{}, uint32_t nhub = 0)
{
    constexpr int NOFF = (WG_TILE_VERTS + 1 + WG_WAVE - 1) / WG_WAVE;
    constexpr int NVTX = WG_TILE_VERTS / WG_WAVE;
    __shared__ uint32_t s_off_all[WPB][WG_TILE_VERTS + 1];
    __shared__ double s_acc_all[WPB][WG_TILE_VERTS];
    __shared__ double s_fp_all[WPB][WG_WAVE];
    __shared__ double s_lp_all[WPB][WG_WAVE];
    __shared__ double s_red[WPB];
    __shared__ T s_hub[HUB ? wg_pr_hub_max<T, WPB>() : 1];
    if (HUB) {
        for (uint32_t i = threadIdx.x; i < nhub; i += WPB * WG_WAVE)
            s_hub[i] = cin[i];
        __syncthreads();
    }
    auto gather = [&](uint32_t sj) -> double {
        if (HUB && sj < nhub)
            return (double)s_hub[sj];
        return load_contrib(cin, sj);
    };

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

    uint32_t t = blockIdx.x * WPB + wave, tend = ntiles, stride = gridDim.x * WPB;
    if (split.on) { // gridDim.x is a multiple of WG_NXCD
        const uint32_t x = blockIdx.x % WG_NXCD;
        t = split.start[x] + (blockIdx.x / WG_NXCD) * WPB + wave;
        tend = split.start[x + 1];
        stride = (gridDim.x / WG_NXCD) * WPB;
    }
    for (; t < tend; t += stride) {
        const uint32_t tv = tile_v[t];
        const uint32_t e0 = tile_e[t], e1 = tile_e[t + 1];
        const uint32_t v0 = tv & ~WG_HEAVY_BIT;
        const uint32_t a = e0 + (uint32_t)lane * IPT;
        uint32_t s_idx[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            s_idx[j] = (a + j < e1) ? __builtin_nontemporal_load(&src[a + j]) : 0u;
        if (tv & WG_HEAVY_BIT) {
            // heavy tile: one destination, partial sum in a fixed order
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < IPT; ++j)
                if (a + j < e1)
                    s += gather(s_idx[j]);
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
        double g[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            g[j] = 0.0;
            if (a + j < e1)
                g[j] = gather(s_idx[j]);
        }
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
            for (int j = 0; j < IPT; ++j) {
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
            const uint32_t b = (a + IPT < e1) ? a + IPT : e1;
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
                    const uint32_t k0 = (vs - e0) / IPT, k1 = (ve - 1 - e0) / IPT;
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
    s_part = block_sum_det<WPB>(s_part, s_red);
    if (threadIdx.x == 0)
        blk_part[blockIdx.x] = s_part;
}

// Deterministic reduction of n partials by the calling block.
static __device__ double reduce_partials(const double* p, uint32_t n, double* s_red)
{
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += WG_BLOCK)
        s += __builtin_nontemporal_load(&p[i]);
    return block_sum_det(s, s_red);
}

// Heavy destinations (sum of their tile partials in tile order) and the final
// reduction of all block partials by the last block to finish.
template <typename T, int MODE, int IPT = WG_IPT>
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
        const uint32_t nch = (off[v + 1] - off[v] + (WG_WAVE * IPT) - 1) / (WG_WAVE * IPT);
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
