// wg_members.hip -- batches of independent graphs held as one disjoint union.
//
// Member g owns the vertices [first[g], first[g+1]) and its edges; no edge
// crosses members, so one pull over the union is one pull over every member.
// The traversals need nothing new (one root per member).  PageRank does: the
// reference's vertex phase (source/phase_vertex_pr.c:26-77)
//     contrib'[v] = ((acc[v] + (1 - S) / N) * d + (1 - d) / N) / od'[v]
// takes N and S (the sum of the accumulators) per graph, so here both are per
// member.  One iteration is
//   1. the edge pull over edge tiles (k_pr_tile / k_pr_span in edge-phase
//      mode, source/phase_edge_pull_pr.c:28-116): acc[v] for every vertex;
//   2. k_bm_chunk_sum: the accumulators of each chunk (<= WG_MEMBER_CHUNK
//      vertices of one member) summed in a fixed order;
//   3. k_bm_vertex: S of the chunk's member from its chunk sums (fixed order),
//      then the vertex phase of the chunk's vertices with the member's N.
// Every sum is taken in a fixed order, so runs are bit-reproducible and the
// captured HIP graph gives the same ranks as plain launches.
//
// Small graphs make this launch-bound when run one by one (a few kernels per
// iteration per graph); as one union, a whole K-iteration run is 1 + 4K
// kernels, captured once into a HIP graph and replayed as one launch.  The
// tile shape (edges per lane) and the tile kernel's workgroups per CU are
// picked by timing every candidate (wg_batch_autotune).
#include "wg_pr_tile.h"

#include <algorithm>
#include <vector>

#define WG_MEMBER_CHUNK 2048 // vertices per chunk of the per-member vertex phase

static const uint32_t k_shape_ipt[] = {4, 8, 16};
static const uint32_t k_shape_bpc[] = {2, 4, 8, 16};
static constexpr int k_n_ipt = 3;
static constexpr int k_n_bpc = 4;
static_assert(k_n_ipt * k_n_bpc * 2 == WG_BATCH_NSHAPES, "candidate list");

namespace {

struct shape_tiles {
    uint32_t ntiles = 0, nheavy = 0;
    uint32_t *tile_v = nullptr, *tile_e = nullptr, *heavy_v = nullptr, *heavy_t = nullptr;
    wg_tile_split split = {}; // edge-balanced XCD ranges of the tiles (split.on unset)
};

} // namespace

struct wg_members {
    uint32_t n = 0;
    std::vector<uint64_t> first;  // host copy, n + 1
    uint32_t* d_first = nullptr;  // n + 1
    uint32_t nchunks = 0;
    uint32_t* chunk_v = nullptr;  // nchunks + 1: chunk c holds [chunk_v[c], chunk_v[c + 1])
    uint32_t* chunk_m = nullptr;  // member of chunk c
    uint32_t* member_chunk = nullptr; // n + 1: member g holds chunks [member_chunk[g], member_chunk[g + 1])
    shape_tiles tiles[k_n_ipt];
    double* tile_part = nullptr;  // heavy-tile partial sums (largest tiling)
    double* blk_part = nullptr;   // per-block sums of the edge pull (not used for S)
    double* chunk_part = nullptr; // per-chunk sums of the accumulators
    double* acc = nullptr;        // V accumulators
    double* scratch = nullptr;    // [0]: unused S input, [1]: total of the accumulators
    void* c[2] = {nullptr, nullptr}; // contributions
    int c_fp32 = -1;              // precision c[] was allocated for
    int cur = -1;                 // buffer holding the last run's contributions
    int tuned = 0;
    wg_batch_shape best = {8, 8, 0};
    // the captured run
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t g_iters = 0;
    double g_damping = 0.0;
    int g_fp32 = -1;
    wg_batch_shape g_shape = {0, 0, 0};
};

static void free_dev_ptr(void* p)
{
    if (p)
        (void)hipFree(p);
}

static void drop_graph(wg_members* m)
{
    if (m->exec)
        (void)hipGraphExecDestroy(m->exec);
    if (m->graph)
        (void)hipGraphDestroy(m->graph);
    m->exec = nullptr;
    m->graph = nullptr;
    m->g_fp32 = -1;
}

void wg_free_members(wg_ctx* ctx)
{
    wg_members* m = ctx->members;
    if (!m)
        return;
    (void)hipStreamSynchronize(ctx->stream);
    drop_graph(m);
    free_dev_ptr(m->d_first);
    free_dev_ptr(m->chunk_v);
    free_dev_ptr(m->chunk_m);
    free_dev_ptr(m->member_chunk);
    for (auto& t : m->tiles) {
        free_dev_ptr(t.tile_v);
        free_dev_ptr(t.tile_e);
        free_dev_ptr(t.heavy_v);
        free_dev_ptr(t.heavy_t);
    }
    free_dev_ptr(m->tile_part);
    free_dev_ptr(m->blk_part);
    free_dev_ptr(m->chunk_part);
    free_dev_ptr(m->acc);
    free_dev_ptr(m->scratch);
    free_dev_ptr(m->c[0]);
    free_dev_ptr(m->c[1]);
    delete m;
    ctx->members = nullptr;
}

// ---- kernels ----------------------------------------------------------------

// Flags an edge u -> v whose source lies outside v's member.
__global__ void k_bm_check(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E,
                           const uint32_t* __restrict__ first, uint32_t n, unsigned int* __restrict__ bad)
{
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = dst[e], u = src[e];
        uint32_t lo = 0, hi = n - 1; // the member g with first[g] <= v < first[g + 1]
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (first[mid] <= v)
                lo = mid;
            else
                hi = mid - 1;
        }
        if (u < first[lo] || u >= first[lo + 1])
            atomicOr(bad, 1u);
    }
}

template <typename T>
__global__ __launch_bounds__(WG_BLOCK) void k_bm_init(const uint32_t* __restrict__ od, T* __restrict__ c,
                                                      const uint32_t* __restrict__ chunk_v,
                                                      const uint32_t* __restrict__ chunk_m,
                                                      const uint32_t* __restrict__ first, uint32_t nchunks)
{
    for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint32_t g = chunk_m[ch];
        const double N = (double)(first[g + 1] - first[g]);
        for (uint32_t v = chunk_v[ch] + threadIdx.x; v < chunk_v[ch + 1]; v += WG_BLOCK) {
            const uint32_t o = od[v];
            c[v] = (T)((1.0 / N) / (o ? (double)o : N));
        }
    }
}

// Sum of the accumulators of each chunk, thread-strided then in wave order.
__global__ __launch_bounds__(WG_BLOCK) void k_bm_chunk_sum(const double* __restrict__ acc,
                                                           const uint32_t* __restrict__ chunk_v, uint32_t nchunks,
                                                           double* __restrict__ chunk_part)
{
    __shared__ double s_red[WG_WPB];
    for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        double s = 0.0;
        for (uint32_t v = chunk_v[ch] + threadIdx.x; v < chunk_v[ch + 1]; v += WG_BLOCK)
            s += acc[v];
        s = block_sum_det(s, s_red);
        if (threadIdx.x == 0)
            chunk_part[ch] = s;
    }
}

// source/phase_vertex_pr.c:26-77 for the vertices of each chunk, with N and S
// of the chunk's member.
template <typename T>
__global__ __launch_bounds__(WG_BLOCK) void k_bm_vertex(const double* __restrict__ acc, const uint32_t* __restrict__ od,
                                                        T* __restrict__ cout, const uint32_t* __restrict__ chunk_v,
                                                        const uint32_t* __restrict__ chunk_m,
                                                        const uint32_t* __restrict__ member_chunk,
                                                        const uint32_t* __restrict__ first, uint32_t nchunks,
                                                        const double* __restrict__ chunk_part, double damping)
{
    __shared__ double s_red[WG_WPB];
    for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint32_t g = chunk_m[ch];
        const uint32_t c0 = member_chunk[g];
        const double S = reduce_partials(chunk_part + c0, member_chunk[g + 1] - c0, s_red);
        const double N = (double)(first[g + 1] - first[g]);
        const double corr = (1.0 - S) / N;
        const double omd_n = (1.0 - damping) / N;
        for (uint32_t v = chunk_v[ch] + threadIdx.x; v < chunk_v[ch + 1]; v += WG_BLOCK) {
            const uint32_t odv = od[v];
            const double t1 = acc[v] + corr;
            const double t2 = t1 * damping;
            const double rank = t2 + omd_n;
            cout[v] = (T)(rank / (odv ? (double)odv : N));
        }
        __syncthreads(); // s_red is reused by the next chunk
    }
}

template <typename T>
__global__ __launch_bounds__(WG_BLOCK) void k_bm_ranks(const uint32_t* __restrict__ od, const T* __restrict__ c,
                                                       double* __restrict__ r, const uint32_t* __restrict__ chunk_v,
                                                       const uint32_t* __restrict__ chunk_m,
                                                       const uint32_t* __restrict__ first, uint32_t nchunks)
{
    for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint32_t g = chunk_m[ch];
        const double N = (double)(first[g + 1] - first[g]);
        for (uint32_t v = chunk_v[ch] + threadIdx.x; v < chunk_v[ch + 1]; v += WG_BLOCK) {
            const uint32_t o = od[v];
            r[v] = (double)c[v] * (o ? (double)o : N);
        }
    }
}

// ---- host side --------------------------------------------------------------

template <typename T>
static int upload_vec(wg_ctx* ctx, T** d, const std::vector<T>& h)
{
    WG_HIP(ctx, hipMalloc((void**)d, (h.empty() ? 1 : h.size()) * sizeof(T)));
    if (!h.empty())
        WG_HIP(ctx, hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return WG_OK;
}

static int build_members(wg_ctx* ctx, wg_members* m, const uint64_t* first, uint32_t n)
{
    m->n = n;
    m->first.assign(first, first + n + 1);
    std::vector<uint32_t> f32(n + 1), cv, cm, mc(n + 1);
    for (uint32_t g = 0; g <= n; ++g)
        f32[g] = (uint32_t)first[g];
    for (uint32_t g = 0; g < n; ++g) {
        mc[g] = (uint32_t)cm.size();
        for (uint64_t v = first[g]; v < first[g + 1]; v += WG_MEMBER_CHUNK) {
            cv.push_back((uint32_t)v);
            cm.push_back(g);
        }
    }
    mc[n] = (uint32_t)cm.size();
    cv.push_back((uint32_t)ctx->V);
    m->nchunks = (uint32_t)cm.size();
    int rc;
    if ((rc = upload_vec(ctx, &m->d_first, f32)) || (rc = upload_vec(ctx, &m->chunk_v, cv)) ||
        (rc = upload_vec(ctx, &m->chunk_m, cm)) || (rc = upload_vec(ctx, &m->member_chunk, mc)))
        return rc;

    // edges must stay inside their member
    unsigned int* bad = nullptr;
    WG_HIP(ctx, hipMalloc((void**)&bad, sizeof(unsigned int)));
    unsigned int hbad = 0;
    hipError_t e0 = hipMemsetAsync(bad, 0, sizeof(unsigned int), ctx->stream);
    if (ctx->E && e0 == hipSuccess)
        hipLaunchKernelGGL(k_bm_check, dim3(wg_grid(ctx, ctx->E, 1024, 8)), dim3(256), 0, ctx->stream, ctx->src, ctx->dst,
                           ctx->E, m->d_first, n, bad);
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e3 = hipStreamSynchronize(ctx->stream);
    (void)hipFree(bad);
    if (e0 != hipSuccess || e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        ctx->last_hip_error = (int)(e0 != hipSuccess ? e0 : e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3);
        return WG_E_HIP;
    }
    if (hbad)
        return WG_E_FORMAT;

    // edge tiles for every candidate tile shape
    std::vector<uint32_t> hoff(ctx->V + 1);
    WG_HIP(ctx, hipMemcpy(hoff.data(), ctx->off, (ctx->V + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
    uint32_t max_tiles = 0;
    for (int i = 0; i < k_n_ipt; ++i) {
        std::vector<uint32_t> tv, te, hv, ht;
        wg_build_tiles(hoff.data(), ctx->V, WG_WAVE * k_shape_ipt[i], tv, te, hv, ht);
        shape_tiles& t = m->tiles[i];
        t.ntiles = (uint32_t)(tv.size() - 1);
        t.nheavy = (uint32_t)hv.size();
        wg_tile_split_points(te, ctx->E, t.split.start);
        max_tiles = std::max(max_tiles, t.ntiles);
        if ((rc = upload_vec(ctx, &t.tile_v, tv)) || (rc = upload_vec(ctx, &t.tile_e, te)) ||
            (rc = upload_vec(ctx, &t.heavy_v, hv)) || (rc = upload_vec(ctx, &t.heavy_t, ht)))
            return rc;
    }
    const size_t nblk = (size_t)ctx->num_cu * 16 + WG_NXCD + (size_t)ctx->num_cu * 4 + 16;
    WG_HIP(ctx, hipMalloc((void**)&m->tile_part, (max_tiles + 1) * sizeof(double)));
    WG_HIP(ctx, hipMalloc((void**)&m->blk_part, nblk * sizeof(double)));
    WG_HIP(ctx, hipMalloc((void**)&m->chunk_part, (m->nchunks + 1) * sizeof(double)));
    WG_HIP(ctx, hipMalloc((void**)&m->acc, (ctx->V ? ctx->V : 1) * sizeof(double)));
    WG_HIP(ctx, hipMalloc((void**)&m->scratch, 2 * sizeof(double)));
    WG_HIP(ctx, hipMemset(m->scratch, 0, 2 * sizeof(double)));
    return WG_OK;
}

extern "C" int wg_set_members(wg_ctx* ctx, uint32_t n, const uint64_t* first)
{
    if (!ctx || !first || n == 0)
        return WG_E_INVAL;
    if (!ctx->off)
        return WG_E_STATE;
    if (ctx->perm && n > 1) // relabelling interleaves the members
        return WG_E_STATE;
    if (first[0] != 0 || first[n] != ctx->V)
        return WG_E_INVAL;
    for (uint32_t g = 0; g < n; ++g)
        if (first[g + 1] <= first[g])
            return WG_E_INVAL; // every member holds at least one vertex
    WG_HIP(ctx, hipSetDevice(ctx->device));
    wg_free_members(ctx);
    wg_members* m = new (std::nothrow) wg_members();
    if (!m)
        return WG_E_NOMEM;
    ctx->members = m;
    const int rc = build_members(ctx, m, first, n);
    if (rc)
        wg_free_members(ctx);
    return rc;
}

extern "C" int wg_batch_shape_at(uint32_t i, wg_batch_shape* out)
{
    if (i >= WG_BATCH_NSHAPES || !out)
        return WG_E_INVAL;
    out->edges_per_lane = k_shape_ipt[(i / 2) / k_n_bpc];
    out->blocks_per_cu = k_shape_bpc[(i / 2) % k_n_bpc];
    out->xcd_split = i & 1;
    return WG_OK;
}

static int shape_index(const wg_batch_shape& s)
{
    for (int i = 0; i < k_n_ipt; ++i)
        if (k_shape_ipt[i] == s.edges_per_lane)
            return i;
    return -1;
}

static bool same_shape(const wg_batch_shape& a, const wg_batch_shape& b)
{
    return a.edges_per_lane == b.edges_per_lane && a.blocks_per_cu == b.blocks_per_cu && a.xcd_split == b.xcd_split;
}

// Queues the whole run: init, then `iters` times edge pull, chunk sums and
// vertex phase.  Only kernel launches (no copies, fills or waits), so the same
// sequence is captured into the HIP graph.
template <typename T, int IPT>
static int queue_run(wg_ctx* ctx, wg_members* m, const shape_tiles& t, uint32_t bpc, uint32_t xcd, uint32_t iters,
                     double damping)
{
    const uint32_t cgrid = wg_grid(ctx, m->nchunks, 1, 16);
    uint32_t grid = wg_grid(ctx, t.ntiles, WG_WPB, bpc);
    wg_tile_split split = t.split;
    split.on = xcd;
    if (xcd)
        grid = (grid + WG_NXCD - 1) / WG_NXCD * WG_NXCD;
    const uint32_t sgrid = wg_grid(ctx, t.nheavy, 4, 4);
    T* c[2] = {(T*)m->c[0], (T*)m->c[1]};
    hipLaunchKernelGGL(k_bm_init<T>, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream, ctx->od, c[0], m->chunk_v, m->chunk_m,
                       m->d_first, m->nchunks);
    for (uint32_t it = 0; it < iters; ++it) {
        const T* cin = c[it & 1];
        T* cout = c[(it & 1) ^ 1];
        hipLaunchKernelGGL((k_pr_tile<T, 1, IPT>), dim3(grid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, ctx->src, cin,
                           (T*)nullptr, m->acc, ctx->od, t.tile_v, t.tile_e, t.ntiles, m->scratch, damping, 1.0,
                           m->tile_part, m->blk_part, split);
        hipLaunchKernelGGL((k_pr_span<T, 1, IPT>), dim3(sgrid), dim3(WG_BLOCK), 0, ctx->stream, ctx->off, t.heavy_v,
                           t.heavy_t, t.nheavy, m->tile_part, (T*)nullptr, m->acc, ctx->od, m->scratch, damping, 1.0,
                           m->blk_part, grid, m->scratch + 1, &ctx->ctr->ticket);
        hipLaunchKernelGGL(k_bm_chunk_sum, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream, m->acc, m->chunk_v, m->nchunks,
                           m->chunk_part);
        hipLaunchKernelGGL(k_bm_vertex<T>, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream, m->acc, ctx->od, cout,
                           m->chunk_v, m->chunk_m, m->member_chunk, m->d_first, m->nchunks, m->chunk_part, damping);
    }
    return wg_check_launch(ctx);
}

template <typename T>
static int queue_run_t(wg_ctx* ctx, wg_members* m, const wg_batch_shape& s, uint32_t iters, double damping)
{
    const int i = shape_index(s);
    switch (s.edges_per_lane) {
    case 4:
        return queue_run<T, 4>(ctx, m, m->tiles[i], s.blocks_per_cu, s.xcd_split, iters, damping);
    case 8:
        return queue_run<T, 8>(ctx, m, m->tiles[i], s.blocks_per_cu, s.xcd_split, iters, damping);
    case 16:
        return queue_run<T, 16>(ctx, m, m->tiles[i], s.blocks_per_cu, s.xcd_split, iters, damping);
    default:
        return WG_E_INVAL;
    }
}

static int queue(wg_ctx* ctx, wg_members* m, const wg_batch_shape& s, uint32_t iters, double damping, int fp32)
{
    return fp32 ? queue_run_t<float>(ctx, m, s, iters, damping) : queue_run_t<double>(ctx, m, s, iters, damping);
}

static int ensure_buffers(wg_ctx* ctx, wg_members* m, int fp32)
{
    if (m->c_fp32 == fp32)
        return WG_OK;
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(m);
    free_dev_ptr(m->c[0]);
    free_dev_ptr(m->c[1]);
    m->c[0] = m->c[1] = nullptr;
    m->c_fp32 = -1;
    m->cur = -1;
    const size_t esz = fp32 ? sizeof(float) : sizeof(double);
    const size_t n = ctx->V ? ctx->V : 1;
    WG_HIP(ctx, hipMalloc(&m->c[0], n * esz));
    WG_HIP(ctx, hipMalloc(&m->c[1], n * esz));
    m->c_fp32 = fp32;
    return WG_OK;
}

// Captures the run into a HIP graph unless the kept one matches.
static int capture(wg_ctx* ctx, wg_members* m, const wg_batch_shape& s, uint32_t iters, double damping, int fp32)
{
    if (m->exec && m->g_iters == iters && m->g_damping == damping && m->g_fp32 == fp32 && same_shape(m->g_shape, s))
        return WG_OK;
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(m);
    WG_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = queue(ctx, m, s, iters, damping, fp32);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc || e != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        if (e != hipSuccess)
            ctx->last_hip_error = (int)e;
        return rc ? rc : WG_E_HIP;
    }
    hipGraphExec_t x = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
        (void)hipGraphDestroy(g);
        ctx->last_hip_error = (int)e2;
        return WG_E_HIP;
    }
    m->graph = g;
    m->exec = x;
    m->g_iters = iters;
    m->g_damping = damping;
    m->g_fp32 = fp32;
    m->g_shape = s;
    return WG_OK;
}

static bool valid_shape(const wg_batch_shape& s)
{
    return shape_index(s) >= 0 && s.blocks_per_cu >= 1 && s.blocks_per_cu <= 16 && s.xcd_split <= 1;
}

static int launch(wg_ctx* ctx, wg_members* m, const wg_batch_shape& s, uint32_t iters, double damping, int fp32,
                  int use_graph)
{
    int rc = ensure_buffers(ctx, m, fp32);
    if (rc)
        return rc;
    if (use_graph) {
        if ((rc = capture(ctx, m, s, iters, damping, fp32)))
            return rc;
        WG_HIP(ctx, hipGraphLaunch(m->exec, ctx->stream));
    } else if ((rc = queue(ctx, m, s, iters, damping, fp32))) {
        return rc;
    }
    m->cur = (int)(iters & 1);
    return WG_OK;
}

extern "C" int wg_batch_pr(wg_ctx* ctx, uint32_t iterations, double damping, int contrib_fp32,
                           const wg_batch_shape* shape, int use_graph)
{
    if (!ctx)
        return WG_E_INVAL;
    wg_members* m = ctx->members;
    if (!m)
        return WG_E_STATE;
    if (!(damping >= 0.0 && damping <= 1.0))
        return WG_E_INVAL;
    const wg_batch_shape s = shape ? *shape : m->best;
    if (!valid_shape(s))
        return WG_E_INVAL;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const int slot = wg_prof_begin(ctx, WG_K_BATCH_PR);
    const int rc = launch(ctx, m, s, iterations, damping, contrib_fp32 ? 1 : 0, use_graph);
    wg_prof_end(ctx, WG_K_BATCH_PR, slot);
    return rc;
}

extern "C" int wg_batch_autotune(wg_ctx* ctx, uint32_t iterations, double damping, int contrib_fp32,
                                 wg_batch_shape* best, float* ms_per_shape)
{
    if (!ctx)
        return WG_E_INVAL;
    wg_members* m = ctx->members;
    if (!m)
        return WG_E_STATE;
    if (!(damping >= 0.0 && damping <= 1.0) || iterations == 0)
        return WG_E_INVAL;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const int fp32 = contrib_fp32 ? 1 : 0;
    hipEvent_t ev[2];
    WG_HIP(ctx, hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return WG_E_HIP;
    }
    int rc = WG_OK;
    float best_ms = 0.0f;
    wg_batch_shape best_s = m->best;
    // Candidates are timed with plain launches: the captured graph runs the same
    // kernels (it saved nothing measurable on 256 graphs), and one capture for
    // the chosen shape replaces 24 rounds of capture, instantiation and
    // destruction.
    for (uint32_t i = 0; i < WG_BATCH_NSHAPES && rc == WG_OK; ++i) {
        wg_batch_shape s;
        wg_batch_shape_at(i, &s);
        float ms = 0.0f;
        if ((rc = launch(ctx, m, s, iterations, damping, fp32, 0))) // warm-up
            break;
        for (int rep = 0; rep < 3 && rc == WG_OK; ++rep) {
            float t = 0.0f;
            if (hipEventRecord(ev[0], ctx->stream) != hipSuccess) {
                ctx->last_hip_error = (int)hipGetLastError();
                rc = WG_E_HIP;
                break;
            }
            if ((rc = launch(ctx, m, s, iterations, damping, fp32, 0)))
                break;
            if (hipEventRecord(ev[1], ctx->stream) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess) {
                ctx->last_hip_error = (int)hipGetLastError();
                rc = WG_E_HIP;
                break;
            }
            ms = (rep == 0 || t < ms) ? t : ms;
        }
        if (rc)
            break;
        if (ms_per_shape)
            ms_per_shape[i] = ms;
        if (i == 0 || ms < best_ms) {
            best_ms = ms;
            best_s = s;
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc)
        return rc;
    m->best = best_s;
    m->tuned = 1;
    if (best)
        *best = best_s;
    return WG_OK;
}

extern "C" int wg_read_batch_ranks(wg_ctx* ctx, double* host_ranks)
{
    if (!ctx || !host_ranks)
        return WG_E_INVAL;
    wg_members* m = ctx->members;
    if (!m || m->cur < 0)
        return WG_E_STATE;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    double* r = nullptr;
    WG_HIP(ctx, hipMalloc((void**)&r, (ctx->V ? ctx->V : 1) * sizeof(double)));
    const uint32_t cgrid = wg_grid(ctx, m->nchunks, 1, 16);
    if (m->c_fp32)
        hipLaunchKernelGGL(k_bm_ranks<float>, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream, ctx->od,
                           (const float*)m->c[m->cur], r, m->chunk_v, m->chunk_m, m->d_first, m->nchunks);
    else
        hipLaunchKernelGGL(k_bm_ranks<double>, dim3(cgrid), dim3(WG_BLOCK), 0, ctx->stream, ctx->od,
                           (const double*)m->c[m->cur], r, m->chunk_v, m->chunk_m, m->d_first, m->nchunks);
    int rc = wg_check_launch(ctx);
    if (rc == WG_OK)
        rc = wg_read_vertex_array(ctx, r, host_ranks, sizeof(double));
    (void)hipFree(r);
    return rc;
}
