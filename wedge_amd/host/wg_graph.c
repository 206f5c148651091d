/*
 * wg_graph.c -- graph ingress and seeded synthetic generators (host, C11 + pthreads).
 *
 * The file format is the reference's binary edge list (source/graphdata.c:
 * header :589-642, exact size check :963-979, 2 or 3 u64 words per record
 * :188-208, records grouped by ascending destination, README.md:50).  Unlike
 * the reference, the loader verifies the destination ordering instead of
 * silently mis-packing unsorted input (the reference flushes only on a
 * destination change, source/graphdata.c:284).
 *
 * Every generator is counter-based: edge candidate i draws its random bits
 * from wg_rand64(key, counter(i)), so the graph is a pure function of the
 * parameters and the seed, independent of the thread count.
 */
#define _GNU_SOURCE
#include "wedge_host.h"

#include <errno.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "wg_parallel.h"
#include "wg_rand.h"

/* ---- random numbers ------------------------------------------------------- */

uint64_t wg_rand64(uint64_t key, uint64_t counter)
{
    return wg_rand64_inline(key, counter);
}

double wg_weight_f64(uint64_t src, uint64_t dst, uint64_t seed)
{
    const uint64_t h = wg_rand64_inline(seed ^ 0xA5A5A5A5A5A5A5A5ull, (src << 32) | (dst & 0xffffffffull));
    return (double)((h >> 11) + 1ull) * 0x1.0p-53; /* (0, 1] */
}

static uint32_t grid_weight(uint64_t u, uint64_t v, uint64_t seed, uint32_t wmax)
{
    const uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
    const uint64_t h = wg_rand64_inline(seed ^ 0x3C3C3C3C3C3C3C3Cull, (lo << 32) | hi);
    return 1u + (uint32_t)(((h >> 32) * (uint64_t)wmax) >> 32); /* [1, wmax] */
}

int wg_host_threads(void)
{
    const char* env = getenv("WEDGE_THREADS");
    if (env && atoi(env) > 0)
        return atoi(env);
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    if (n < 1)
        n = 1;
    if (n > 16)
        n = 16;
    return (int)n;
}

/* ---- small sorts for the per-destination source lists --------------------- */

static void sort_u32(uint32_t* a, uint64_t n, uint32_t* tmp)
{
    if (n <= 48) {
        for (uint64_t i = 1; i < n; ++i) {
            const uint32_t x = a[i];
            uint64_t j = i;
            while (j > 0 && a[j - 1] > x) {
                a[j] = a[j - 1];
                --j;
            }
            a[j] = x;
        }
        return;
    }
    /* LSD radix sort, 11-bit digits; passes whose digit is constant are skipped */
    uint32_t mx = 0;
    for (uint64_t i = 0; i < n; ++i)
        mx |= a[i];
    uint32_t* from = a;
    uint32_t* to = tmp;
    for (int shift = 0; shift < 32; shift += 11) {
        if ((mx >> shift) == 0)
            break;
        uint64_t cnt[2049];
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; ++i)
            cnt[((from[i] >> shift) & 2047u) + 1]++;
        for (int d = 0; d < 2048; ++d)
            cnt[d + 1] += cnt[d];
        for (uint64_t i = 0; i < n; ++i)
            to[cnt[(from[i] >> shift) & 2047u]++] = from[i];
        uint32_t* t = from;
        from = to;
        to = t;
    }
    if (from != a)
        memcpy(a, from, n * sizeof(uint32_t));
}

/* pairs (key, payload) sorted by key, stable */
static void sort_pairs_u32(uint32_t* key, uint64_t* pay, uint64_t n, uint32_t* tk, uint64_t* tp)
{
    if (n <= 48) {
        for (uint64_t i = 1; i < n; ++i) {
            const uint32_t k = key[i];
            const uint64_t p = pay[i];
            uint64_t j = i;
            while (j > 0 && key[j - 1] > k) {
                key[j] = key[j - 1];
                pay[j] = pay[j - 1];
                --j;
            }
            key[j] = k;
            pay[j] = p;
        }
        return;
    }
    uint32_t mx = 0;
    for (uint64_t i = 0; i < n; ++i)
        mx |= key[i];
    uint32_t *fk = key, *tok = tk;
    uint64_t *fp = pay, *top = tp;
    for (int shift = 0; shift < 32; shift += 11) {
        if ((mx >> shift) == 0)
            break;
        uint64_t cnt[2049];
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; ++i)
            cnt[((fk[i] >> shift) & 2047u) + 1]++;
        for (int d = 0; d < 2048; ++d)
            cnt[d + 1] += cnt[d];
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t pos = cnt[(fk[i] >> shift) & 2047u]++;
            tok[pos] = fk[i];
            top[pos] = fp[i];
        }
        uint32_t* t = fk;
        fk = tok;
        tok = t;
        uint64_t* u = fp;
        fp = top;
        top = u;
    }
    if (fk != key) {
        memcpy(key, fk, n * sizeof(uint32_t));
        memcpy(pay, fp, n * sizeof(uint64_t));
    }
}

/* ---- bucketed CSC builder over a counter-based edge source ----------------
 * Pass A: each thread counts its fixed candidate range into a private
 *         histogram of coarse destination buckets (no shared atomics; hub
 *         destinations made shared counters the bottleneck).
 * Pass B: each thread regenerates its range and writes (dst, src) pairs into
 *         slots reserved per (bucket, thread), so positions are deterministic.
 * Pass C: per bucket, counting sort by destination, sort each source list,
 *         drop self-loops and duplicates.
 * Pass D: compaction into the final CSC.                                      */

/* Emits 0, 1 or 2 directed edges for candidate i. */
typedef uint32_t (*edge_fn)(const void* ctx, uint64_t i, uint32_t src[2], uint32_t dst[2]);

typedef struct {
    const void* gctx;
    edge_fn fn;
    int nthreads;
    uint64_t ncand;
    uint64_t V;
    int bshift;          /* bucket = dst >> bshift */
    uint64_t nb;         /* number of buckets */
    uint64_t* hist;      /* [thread][bucket] counts, then slot cursors */
    uint64_t* bstart;    /* nb + 1 bucket starts in the pair array */
    uint64_t* pairs;     /* (dst << 32) | src */
    uint32_t* newcnt;    /* per-destination count after sort + dedup */
    uint64_t* bout;      /* per-bucket number of kept sources */
    int dedup;
} build_ctx;

typedef struct {
    build_ctx* c;
    int tid;
} build_arg;

static void cand_range(const build_ctx* c, int tid, uint64_t* b, uint64_t* e)
{
    *b = c->ncand * (uint64_t)tid / (uint64_t)c->nthreads;
    *e = c->ncand * (uint64_t)(tid + 1) / (uint64_t)c->nthreads;
}

static void* pass_count(void* p)
{
    build_arg* a = (build_arg*)p;
    build_ctx* c = a->c;
    uint64_t* h = c->hist + (uint64_t)a->tid * c->nb;
    uint64_t b, e;
    cand_range(c, a->tid, &b, &e);
    uint32_t s[2], d[2];
    for (uint64_t i = b; i < e; ++i) {
        const uint32_t k = c->fn(c->gctx, i, s, d);
        for (uint32_t j = 0; j < k; ++j)
            h[d[j] >> c->bshift]++;
    }
    return NULL;
}

static void* pass_scatter(void* p)
{
    build_arg* a = (build_arg*)p;
    build_ctx* c = a->c;
    uint64_t* cur = c->hist + (uint64_t)a->tid * c->nb;
    uint64_t b, e;
    cand_range(c, a->tid, &b, &e);
    uint32_t s[2], d[2];
    for (uint64_t i = b; i < e; ++i) {
        const uint32_t k = c->fn(c->gctx, i, s, d);
        for (uint32_t j = 0; j < k; ++j)
            c->pairs[cur[d[j] >> c->bshift]++] = ((uint64_t)d[j] << 32) | s[j];
    }
    return NULL;
}

static void run_threads(build_ctx* c, void* (*fn)(void*))
{
    pthread_t th[64];
    build_arg args[64];
    int started = 0;
    for (int t = 1; t < c->nthreads; ++t) {
        args[t].c = c;
        args[t].tid = t;
        if (pthread_create(&th[t], NULL, fn, &args[t]) != 0)
            break;
        started = t;
    }
    args[0].c = c;
    args[0].tid = 0;
    fn(&args[0]);
    /* threads that failed to start: run their ranges here */
    for (int t = started + 1; t < c->nthreads; ++t) {
        args[t].c = c;
        args[t].tid = t;
        fn(&args[t]);
    }
    for (int t = 1; t <= started; ++t)
        pthread_join(th[t], NULL);
}

static void pass_bucket_sort(void* p, uint64_t b0, uint64_t b1, int tid)
{
    (void)tid;
    build_ctx* c = (build_ctx*)p;
    const uint64_t width = 1ull << c->bshift;
    uint64_t* cnt = (uint64_t*)malloc((width + 1) * sizeof(uint64_t));
    uint32_t* tmp = NULL;
    uint32_t* stmp = NULL;
    uint64_t cap = 0;
    for (uint64_t bk = b0; bk < b1; ++bk) {
        const uint64_t ps = c->bstart[bk], pe = c->bstart[bk + 1], n = pe - ps;
        const uint64_t v0 = bk << c->bshift;
        const uint64_t vn = (v0 + width <= c->V) ? width : c->V - v0;
        if (n > cap) {
            free(tmp);
            free(stmp);
            cap = n + n / 2;
            tmp = (uint32_t*)malloc(cap * sizeof(uint32_t));
            stmp = (uint32_t*)malloc(cap * sizeof(uint32_t));
        }
        memset(cnt, 0, (vn + 1) * sizeof(uint64_t));
        for (uint64_t i = ps; i < pe; ++i)
            cnt[(c->pairs[i] >> 32) - v0 + 1]++;
        for (uint64_t v = 0; v < vn; ++v)
            cnt[v + 1] += cnt[v];
        for (uint64_t i = ps; i < pe; ++i) {
            const uint64_t x = c->pairs[i];
            tmp[cnt[(x >> 32) - v0]++] = (uint32_t)x;
        }
        /* cnt[v] now holds the end of group v; group v starts at cnt[v-1] */
        uint32_t* out = (uint32_t*)(c->pairs + ps); /* reuse the bucket's own region */
        uint64_t w = 0, gb = 0;
        for (uint64_t v = 0; v < vn; ++v) {
            const uint64_t ge = cnt[v];
            uint32_t* a = tmp + gb;
            const uint64_t gn = ge - gb;
            sort_u32(a, gn, stmp);
            const uint64_t w0 = w;
            for (uint64_t i = 0; i < gn; ++i) {
                if (c->dedup) {
                    if (a[i] == (uint32_t)(v0 + v))
                        continue;
                    if (w > w0 && out[w - 1] == a[i])
                        continue;
                }
                out[w++] = a[i];
            }
            c->newcnt[v0 + v] = (uint32_t)(w - w0);
            gb = ge;
        }
        c->bout[bk] = w;
    }
    free(cnt);
    free(tmp);
    free(stmp);
}

typedef struct {
    build_ctx* c;
    const uint32_t* off;
    uint32_t* src;
} compact_ctx;

static void pass_compact(void* p, uint64_t b0, uint64_t b1, int tid)
{
    (void)tid;
    compact_ctx* cc = (compact_ctx*)p;
    build_ctx* c = cc->c;
    for (uint64_t bk = b0; bk < b1; ++bk) {
        const uint64_t n = c->bout[bk];
        if (n)
            memcpy(cc->src + cc->off[bk << c->bshift], (const uint32_t*)(c->pairs + c->bstart[bk]),
                   n * sizeof(uint32_t));
    }
}

static int build_csc(wg_graph* g, uint64_t V, uint64_t ncand, edge_fn fn, const void* gctx, int dedup)
{
    memset(g, 0, sizeof(*g));
    if (V == 0 || V >= 0xffffffffull)
        return WG_E_RANGE;
    build_ctx c;
    memset(&c, 0, sizeof(c));
    c.gctx = gctx;
    c.fn = fn;
    c.dedup = dedup;
    c.V = V;
    c.ncand = ncand;
    c.nthreads = wg_host_threads();
    if (c.nthreads > 64)
        c.nthreads = 64;
    int vbits = 0;
    while ((1ull << vbits) < V)
        ++vbits;
    c.bshift = vbits > 16 ? vbits - 16 : 0;
    c.nb = (V + (1ull << c.bshift) - 1) >> c.bshift;
    c.hist = (uint64_t*)calloc((uint64_t)c.nthreads * c.nb, sizeof(uint64_t));
    c.bstart = (uint64_t*)malloc((c.nb + 1) * sizeof(uint64_t));
    c.bout = (uint64_t*)malloc(c.nb * sizeof(uint64_t));
    c.newcnt = (uint32_t*)malloc(V * sizeof(uint32_t));
    if (!c.hist || !c.bstart || !c.bout || !c.newcnt) {
        free(c.hist);
        free(c.bstart);
        free(c.bout);
        free(c.newcnt);
        return WG_E_NOMEM;
    }
    run_threads(&c, pass_count);
    /* slot cursors: bucket-major, thread-minor */
    uint64_t run = 0;
    for (uint64_t bk = 0; bk < c.nb; ++bk) {
        c.bstart[bk] = run;
        for (int t = 0; t < c.nthreads; ++t) {
            uint64_t* h = &c.hist[(uint64_t)t * c.nb + bk];
            const uint64_t n = *h;
            *h = run;
            run += n;
        }
    }
    c.bstart[c.nb] = run;
    c.pairs = (uint64_t*)malloc((run ? run : 1) * sizeof(uint64_t));
    if (!c.pairs) {
        free(c.hist);
        free(c.bstart);
        free(c.bout);
        free(c.newcnt);
        return WG_E_NOMEM;
    }
    run_threads(&c, pass_scatter);
    free(c.hist);
    c.hist = NULL;
    wg_parallel_for(c.nthreads, c.nb, 1, pass_bucket_sort, &c);

    uint32_t* off = (uint32_t*)malloc((V + 1) * sizeof(uint32_t));
    int rc = off ? WG_OK : WG_E_NOMEM;
    uint64_t E = 0;
    for (uint64_t v = 0; v < V && rc == WG_OK; ++v) {
        off[v] = (uint32_t)E;
        E += c.newcnt[v];
        if (E >= 0xffffffffull)
            rc = WG_E_RANGE;
    }
    uint32_t* src = NULL;
    if (rc == WG_OK) {
        off[V] = (uint32_t)E;
        src = (uint32_t*)malloc((E ? E : 1) * sizeof(uint32_t));
        if (!src)
            rc = WG_E_NOMEM;
    }
    if (rc == WG_OK) {
        compact_ctx cc = { &c, off, src };
        wg_parallel_for(c.nthreads, c.nb, 64, pass_compact, &cc);
    }
    free(c.pairs);
    free(c.bstart);
    free(c.bout);
    free(c.newcnt);
    if (rc != WG_OK) {
        free(off);
        free(src);
        return rc;
    }
    g->num_vertices = V;
    g->num_edges = E;
    g->offsets = off;
    g->sources = src;
    g->weights = NULL;
    g->weight_kind = WG_WEIGHT_NONE;
    return WG_OK;
}

/* ---- per-edge weights after the build ------------------------------------- */

typedef struct {
    wg_graph* g;
    uint64_t seed;
    uint32_t wmax;
} weight_ctx;

static void pass_weights_f64(void* p, uint64_t b, uint64_t e, int tid)
{
    (void)tid;
    weight_ctx* c = (weight_ctx*)p;
    double* w = (double*)c->g->weights;
    for (uint64_t v = b; v < e; ++v)
        for (uint64_t k = c->g->offsets[v]; k < c->g->offsets[v + 1]; ++k)
            w[k] = wg_weight_f64(c->g->sources[k], v, c->seed);
}

static void pass_weights_grid(void* p, uint64_t b, uint64_t e, int tid)
{
    (void)tid;
    weight_ctx* c = (weight_ctx*)p;
    uint32_t* w = (uint32_t*)c->g->weights;
    for (uint64_t v = b; v < e; ++v)
        for (uint64_t k = c->g->offsets[v]; k < c->g->offsets[v + 1]; ++k)
            w[k] = grid_weight(c->g->sources[k], v, c->seed, c->wmax);
}

/* ---- R-MAT ------------------------------------------------------------------ */

typedef struct {
    int scale;
    int symmetrize;
    uint64_t key;
    uint32_t t_a, t_ab, t_abc;
} rmat_ctx;

static uint32_t rmat_fn(const void* p, uint64_t i, uint32_t src[2], uint32_t dst[2])
{
    const rmat_ctx* r = (const rmat_ctx*)p;
    uint32_t s, d;
    wg_rmat_draw(r->key, r->scale, r->t_a, r->t_ab, r->t_abc, i, &s, &d);
    src[0] = s;
    dst[0] = d;
    if (!r->symmetrize)
        return 1;
    src[1] = d;
    dst[1] = s;
    return 2;
}

int wg_graph_rmat(wg_graph* g, int scale, int edge_factor, double a, double b, double c,
                  uint64_t seed, int symmetrize, int weight_kind)
{
    if (!g || scale < 1 || scale > 31 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return WG_E_INVAL;
    rmat_ctx r;
    r.scale = scale;
    r.symmetrize = symmetrize;
    r.key = wg_rand64(seed, 0x524D4154ull); /* "RMAT" */
    r.t_a = (uint32_t)(a * 4294967296.0);
    r.t_ab = (uint32_t)((a + b) * 4294967296.0);
    r.t_abc = (a + b + c >= 1.0) ? 0xffffffffu : (uint32_t)((a + b + c) * 4294967296.0);
    const uint64_t V = 1ull << scale;
    const uint64_t draws = (uint64_t)edge_factor << scale;
    int rc = build_csc(g, V, draws, rmat_fn, &r, 1);
    if (rc)
        return rc;
    if (weight_kind == WG_WEIGHT_F64) {
        g->weights = malloc((g->num_edges ? g->num_edges : 1) * sizeof(double));
        if (!g->weights) {
            wg_graph_free(g);
            return WG_E_NOMEM;
        }
        g->weight_kind = WG_WEIGHT_F64;
        weight_ctx wc = { g, seed, 0 };
        wg_parallel_for(wg_host_threads(), V, 1u << 12, pass_weights_f64, &wc);
    } else if (weight_kind != WG_WEIGHT_NONE) {
        wg_graph_free(g);
        return WG_E_INVAL;
    }
    return WG_OK;
}

/* ---- road-like grid --------------------------------------------------------- */

typedef struct {
    uint64_t rows, cols;
    uint64_t n_right, n_down, n_diag;
    uint32_t t_keep, t_diag;
    uint64_t key;
} grid_ctx;

static uint32_t grid_fn(const void* p, uint64_t i, uint32_t src[2], uint32_t dst[2])
{
    const grid_ctx* q = (const grid_ctx*)p;
    uint64_t u, v;
    uint32_t thr;
    if (i < q->n_right) { /* (r, c) - (r, c+1) */
        const uint64_t r = i / (q->cols - 1), c = i % (q->cols - 1);
        u = r * q->cols + c;
        v = u + 1;
        thr = q->t_keep;
    } else if (i < q->n_right + q->n_down) { /* (r, c) - (r+1, c) */
        const uint64_t k = i - q->n_right;
        u = k;
        v = k + q->cols;
        thr = q->t_keep;
    } else { /* (r, c) - (r+1, c+1) */
        const uint64_t k = i - q->n_right - q->n_down;
        const uint64_t r = k / (q->cols - 1), c = k % (q->cols - 1);
        u = r * q->cols + c;
        v = u + q->cols + 1;
        thr = q->t_diag;
    }
    if ((uint32_t)wg_rand64_inline(q->key, i) >= thr)
        return 0;
    src[0] = (uint32_t)u;
    dst[0] = (uint32_t)v;
    src[1] = (uint32_t)v;
    dst[1] = (uint32_t)u;
    return 2;
}

int wg_graph_grid(wg_graph* g, uint32_t rows, uint32_t cols, double p_keep, double p_diag,
                  uint32_t wmax, uint64_t seed)
{
    if (!g || rows < 2 || cols < 2 || p_keep < 0 || p_keep > 1 || p_diag < 0 || p_diag > 1 || wmax < 1)
        return WG_E_INVAL;
    grid_ctx q;
    q.rows = rows;
    q.cols = cols;
    q.n_right = (uint64_t)rows * (cols - 1);
    q.n_down = (uint64_t)(rows - 1) * cols;
    q.n_diag = (uint64_t)(rows - 1) * (cols - 1);
    q.t_keep = p_keep >= 1.0 ? 0xffffffffu : (uint32_t)(p_keep * 4294967296.0);
    q.t_diag = p_diag >= 1.0 ? 0xffffffffu : (uint32_t)(p_diag * 4294967296.0);
    q.key = wg_rand64(seed, 0x47524944ull); /* "GRID" */
    const uint64_t V = (uint64_t)rows * cols;
    int rc = build_csc(g, V, q.n_right + q.n_down + q.n_diag, grid_fn, &q, 1);
    if (rc)
        return rc;
    g->weights = malloc((g->num_edges ? g->num_edges : 1) * sizeof(uint32_t));
    if (!g->weights) {
        wg_graph_free(g);
        return WG_E_NOMEM;
    }
    g->weight_kind = WG_WEIGHT_U32;
    weight_ctx wc = { g, seed, wmax };
    wg_parallel_for(wg_host_threads(), V, 1u << 12, pass_weights_grid, &wc);
    return WG_OK;
}

/* ---- file I/O ---------------------------------------------------------------- */

int wg_graph_read(wg_graph* g, const char* path, int weight_kind)
{
    if (!g || !path)
        return WG_E_INVAL;
    memset(g, 0, sizeof(*g));
    FILE* f = fopen(path, "rb");
    if (!f)
        return WG_E_IO;
    uint64_t hdr[2];
    if (fread(hdr, sizeof(uint64_t), 2, f) != 2) {
        fclose(f);
        return WG_E_FORMAT;
    }
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return WG_E_IO;
    }
    const long fsize = ftell(f);
    const uint64_t V = hdr[0], E = hdr[1];
    const uint64_t words = (weight_kind == WG_WEIGHT_NONE) ? 2 : 3;
    /* exact size check, source/graphdata.c:963-979 */
    if (fsize < 0 || E > (UINT64_MAX - 16) / (8 * words) || (uint64_t)fsize != 16 + E * 8 * words) {
        fclose(f);
        return WG_E_FORMAT;
    }
    if (V == 0 || V >= 0xffffffffull || E >= 0xffffffffull) {
        fclose(f);
        return WG_E_RANGE;
    }
    if (fseek(f, 16, SEEK_SET) != 0) {
        fclose(f);
        return WG_E_IO;
    }
    uint32_t* off = (uint32_t*)calloc(V + 1, sizeof(uint32_t));
    uint32_t* src = (uint32_t*)malloc((E ? E : 1) * sizeof(uint32_t));
    void* w = NULL;
    if (weight_kind == WG_WEIGHT_F64)
        w = malloc((E ? E : 1) * sizeof(double));
    else if (weight_kind == WG_WEIGHT_U32)
        w = malloc((E ? E : 1) * sizeof(uint32_t));
    if (!off || !src || (weight_kind != WG_WEIGHT_NONE && !w)) {
        free(off);
        free(src);
        free(w);
        fclose(f);
        return WG_E_NOMEM;
    }
    const uint64_t chunk = 1u << 18;
    uint64_t* buf = (uint64_t*)malloc(chunk * words * sizeof(uint64_t));
    if (!buf) {
        free(off);
        free(src);
        free(w);
        fclose(f);
        return WG_E_NOMEM;
    }
    int rc = WG_OK;
    uint64_t prev_dst = 0;
    for (uint64_t at = 0; at < E && rc == WG_OK;) {
        const uint64_t n = (E - at < chunk) ? E - at : chunk;
        if (fread(buf, sizeof(uint64_t) * words, n, f) != n) {
            rc = WG_E_IO;
            break;
        }
        for (uint64_t k = 0; k < n; ++k) {
            const uint64_t s = buf[k * words], d = buf[k * words + 1];
            if (s >= V || d >= V) {
                rc = WG_E_RANGE;
                break;
            }
            if (d < prev_dst) {
                rc = WG_E_FORMAT; /* not grouped by ascending destination */
                break;
            }
            prev_dst = d;
            off[d + 1]++;
            src[at + k] = (uint32_t)s;
            if (weight_kind == WG_WEIGHT_F64) {
                double x;
                memcpy(&x, &buf[k * words + 2], sizeof(double));
                ((double*)w)[at + k] = x;
            } else if (weight_kind == WG_WEIGHT_U32) {
                if (buf[k * words + 2] > 0xffffffffull) {
                    rc = WG_E_RANGE;
                    break;
                }
                ((uint32_t*)w)[at + k] = (uint32_t)buf[k * words + 2];
            }
        }
        at += n;
    }
    free(buf);
    fclose(f);
    if (rc != WG_OK) {
        free(off);
        free(src);
        free(w);
        return rc;
    }
    for (uint64_t v = 0; v < V; ++v)
        off[v + 1] += off[v];
    /* sort sources within each destination (weights follow their edge) */
    uint32_t* tk = NULL;
    uint64_t* pay = NULL;
    uint64_t* tp = NULL;
    uint64_t cap = 0;
    for (uint64_t v = 0; v < V; ++v) {
        const uint64_t b = off[v], n = off[v + 1] - off[v];
        if (n < 2)
            continue;
        int sorted = 1;
        for (uint64_t k = 1; k < n && sorted; ++k)
            sorted = src[b + k - 1] <= src[b + k];
        if (sorted)
            continue;
        if (n > cap) {
            free(tk);
            free(pay);
            free(tp);
            cap = n * 2;
            tk = (uint32_t*)malloc(cap * sizeof(uint32_t));
            pay = (uint64_t*)malloc(cap * sizeof(uint64_t));
            tp = (uint64_t*)malloc(cap * sizeof(uint64_t));
            if (!tk || !pay || !tp) {
                free(tk);
                free(pay);
                free(tp);
                free(off);
                free(src);
                free(w);
                return WG_E_NOMEM;
            }
        }
        for (uint64_t k = 0; k < n; ++k) {
            if (weight_kind == WG_WEIGHT_F64)
                memcpy(&pay[k], &((double*)w)[b + k], 8);
            else if (weight_kind == WG_WEIGHT_U32)
                pay[k] = ((uint32_t*)w)[b + k];
            else
                pay[k] = 0;
        }
        sort_pairs_u32(src + b, pay, n, tk, tp);
        for (uint64_t k = 0; k < n; ++k) {
            if (weight_kind == WG_WEIGHT_F64)
                memcpy(&((double*)w)[b + k], &pay[k], 8);
            else if (weight_kind == WG_WEIGHT_U32)
                ((uint32_t*)w)[b + k] = (uint32_t)pay[k];
        }
    }
    free(tk);
    free(pay);
    free(tp);
    g->num_vertices = V;
    g->num_edges = E;
    g->offsets = off;
    g->sources = src;
    g->weights = w;
    g->weight_kind = weight_kind;
    return WG_OK;
}

int wg_graph_write(const wg_graph* g, const char* path)
{
    if (!g || !path || !g->offsets)
        return WG_E_INVAL;
    FILE* f = fopen(path, "wb");
    if (!f)
        return WG_E_IO;
    const uint64_t hdr[2] = { g->num_vertices, g->num_edges };
    int rc = fwrite(hdr, sizeof(uint64_t), 2, f) == 2 ? WG_OK : WG_E_IO;
    const uint64_t words = g->weight_kind == WG_WEIGHT_NONE ? 2 : 3;
    const uint64_t chunk = 1u << 16;
    uint64_t* buf = (uint64_t*)malloc(chunk * words * sizeof(uint64_t));
    if (!buf) {
        fclose(f);
        return WG_E_NOMEM;
    }
    uint64_t n = 0;
    for (uint64_t v = 0; v < g->num_vertices && rc == WG_OK; ++v) {
        for (uint64_t k = g->offsets[v]; k < g->offsets[v + 1]; ++k) {
            buf[n * words] = g->sources[k];
            buf[n * words + 1] = v;
            if (g->weight_kind == WG_WEIGHT_F64)
                memcpy(&buf[n * words + 2], &((const double*)g->weights)[k], 8);
            else if (g->weight_kind == WG_WEIGHT_U32)
                buf[n * words + 2] = ((const uint32_t*)g->weights)[k];
            if (++n == chunk) {
                if (fwrite(buf, sizeof(uint64_t) * words, n, f) != n)
                    rc = WG_E_IO;
                n = 0;
            }
        }
    }
    if (rc == WG_OK && n && fwrite(buf, sizeof(uint64_t) * words, n, f) != n)
        rc = WG_E_IO;
    free(buf);
    if (fclose(f) != 0 && rc == WG_OK)
        rc = WG_E_IO;
    return rc;
}

void wg_graph_free(wg_graph* g)
{
    if (!g)
        return;
    free(g->offsets);
    free(g->sources);
    free(g->weights);
    memset(g, 0, sizeof(*g));
}

typedef struct {
    const wg_graph* g;
    _Atomic uint32_t* od;
} outdeg_ctx;

static void pass_outdeg(void* p, uint64_t b, uint64_t e, int tid)
{
    (void)tid;
    outdeg_ctx* c = (outdeg_ctx*)p;
    for (uint64_t k = b; k < e; ++k)
        atomic_fetch_add_explicit(&c->od[c->g->sources[k]], 1u, memory_order_relaxed);
}

int wg_graph_outdegrees(const wg_graph* g, uint32_t* out)
{
    if (!g || !out)
        return WG_E_INVAL;
    memset(out, 0, g->num_vertices * sizeof(uint32_t));
    outdeg_ctx c = { g, (_Atomic uint32_t*)out };
    wg_parallel_for(wg_host_threads(), g->num_edges, 1u << 16, pass_outdeg, &c);
    return WG_OK;
}

uint64_t wg_graph_max_outdegree_vertex(const wg_graph* g)
{
    if (!g || !g->num_vertices)
        return 0;
    uint32_t* od = (uint32_t*)malloc(g->num_vertices * sizeof(uint32_t));
    if (!od)
        return 0;
    wg_graph_outdegrees(g, od);
    uint64_t best = 0;
    for (uint64_t v = 1; v < g->num_vertices; ++v)
        if (od[v] > od[best])
            best = v;
    free(od);
    return best;
}

int wg_graph_union(const wg_graph* members, uint32_t n, wg_graph* out, uint64_t* first)
{
    if (!members || !out || !first || n == 0)
        return WG_E_INVAL;
    memset(out, 0, sizeof(*out));
    const int kind = members[0].weight_kind;
    uint64_t V = 0, E = 0;
    for (uint32_t g = 0; g < n; ++g) {
        if (members[g].weight_kind != kind || !members[g].offsets)
            return WG_E_INVAL;
        first[g] = V;
        V += members[g].num_vertices;
        E += members[g].num_edges;
    }
    first[n] = V;
    if (V >= 0xffffffffull || E >= 0xffffffffull)
        return WG_E_RANGE;
    const size_t wsz = kind == WG_WEIGHT_F64 ? sizeof(double) : kind == WG_WEIGHT_U32 ? sizeof(uint32_t) : 0;
    uint32_t* off = (uint32_t*)malloc((V + 1) * sizeof(uint32_t));
    uint32_t* src = (uint32_t*)malloc((E ? E : 1) * sizeof(uint32_t));
    char* w = wsz ? (char*)malloc((E ? E : 1) * wsz) : NULL;
    if (!off || !src || (wsz && !w)) {
        free(off);
        free(src);
        free(w);
        return WG_E_NOMEM;
    }
    uint64_t e0 = 0;
    for (uint32_t g = 0; g < n; ++g) {
        const wg_graph* m = &members[g];
        const uint32_t v0 = (uint32_t)first[g];
        for (uint64_t v = 0; v < m->num_vertices; ++v)
            off[v0 + v] = (uint32_t)(e0 + m->offsets[v]);
        for (uint64_t e = 0; e < m->num_edges; ++e)
            src[e0 + e] = v0 + m->sources[e];
        if (wsz && m->num_edges)
            memcpy(w + e0 * wsz, m->weights, m->num_edges * wsz);
        e0 += m->num_edges;
    }
    off[V] = (uint32_t)E;
    out->num_vertices = V;
    out->num_edges = E;
    out->offsets = off;
    out->sources = src;
    out->weights = w;
    out->weight_kind = kind;
    return WG_OK;
}
