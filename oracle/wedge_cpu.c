/*
 * wedge_cpu.c -- ORACLE / TEST INFRASTRUCTURE ONLY (see wedge_cpu.h).
 *
 * CPU restatement of Wedge's pull engine.  Structure follows the reference:
 * every worker thread runs the application driver (SPMD), work is split into
 * 32 dynamically scheduled units per thread, destinations that straddle two
 * units meet again in the merge buffer, per-thread scalars go through a
 * reduce buffer, and the frontier mode walks the Wedge frontier (one bit per
 * 2^P edge vectors) built from the CSRI.  Citations are to the reference
 * sources at stanford-mast/Wedge.
 */
#define _GNU_SOURCE
#include "wedge_cpu.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* include/configuration.h:19-24: 2^2 edge vectors per frontier bit */
#define LOG2_PREC 2ull
#define PREC (1ull << LOG2_PREC)
#define UNITS_PER_THREAD 32ull
#define LANE_VALID 0x8000000000000000ull
#define LANE_SRC 0x0000ffffffffffffull
#define MERGE_INVALID 0x8000000000000000ull

typedef union {
    double d;
    uint64_t u;
} prop_t;

/* source/graphtypes.h mergeaccum_t */
typedef struct {
    uint64_t initial;
    uint64_t final;
    prop_t partial;
    uint64_t pad;
} merge_t;

struct wcpu {
    uint64_t V, E;
    uint64_t nvec;        /* edge vectors holding edges */
    uint64_t nvec_alloc;  /* padded to the frontier precision */
    int wkind;
    uint64_t* vecs;       /* 4 lanes per vector; weighted: 4 edge lanes then 4 weight lanes */
    uint64_t vstride;     /* u64 words per vector (4 or 8) */
    uint64_t vpad;        /* padded vertex count (multiple of 8, > V) */
    uint64_t* vindex;     /* first vector of each destination, UINT64_MAX if none */
    uint64_t* outdeg;     /* out-degree from valid lanes */
    uint64_t* csri_index; /* V + 1 */
    uint64_t* csri_edges; /* Wedge group index per (source, vector) occurrence */
    uint64_t wedge_words;
    uint64_t* wedge;
    uint64_t frontier_words;
    uint64_t* frontier_a;
    uint64_t* frontier_b;
    uint64_t* unconverged;
    int nthreads;

    /* per-run state */
    uint64_t* active;     /* graph_frontier_active_vertices */
    uint64_t* snapshot;   /* graph_frontier_active_vertices_snapshot */
    prop_t* props;
    prop_t* acc;
    double* od_d;
    uint32_t* depth;
    merge_t* merge;
    uint64_t nunits;
    uint64_t* reduce;
    _Atomic uint64_t sched;
    pthread_barrier_t bar;
    pthread_mutex_t gate_mu;
    pthread_cond_t gate_cv;
    int gate_state; /* 0 wait, 1 run, 2 abort */
    int kind;       /* min-propagation flavour of the current run */

    /* run parameters */
    uint32_t iterations;
    double damping;
    uint64_t root;
    uint32_t threshold_pct;
    wcpu_stats stats;
};

/* ---- helpers ---------------------------------------------------------------- */

static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC_RAW, &t);
    return (double)t.tv_sec * 1e3 + (double)t.tv_nsec * 1e-6;
}

static inline const uint64_t* vec_at(const wcpu* w, uint64_t i)
{
    return w->vecs + i * w->vstride;
}

/* destination id from the four 15-bit pieces (source/phase_edge_pull_pr.c:22) */
static inline uint64_t vec_dst(const uint64_t* l)
{
    return ((l[0] >> 48) & 0x7fffull) | (((l[1] >> 48) & 0x7fffull) << 15) |
           (((l[2] >> 48) & 0x7fffull) << 30) | (((l[3] >> 48) & 0x7ull) << 45);
}

/* lane encoding (source/graphdata.c:212-243) */
static inline uint64_t lane_encode(uint64_t dst, int lane, int valid, uint64_t src)
{
    uint64_t piece;
    switch (lane) {
    case 0: piece = dst & 0x7fffull; break;
    case 1: piece = (dst >> 15) & 0x7fffull; break;
    case 2: piece = (dst >> 30) & 0x7fffull; break;
    default: piece = (dst >> 45) & 0x7ull; break;
    }
    return ((uint64_t)(valid ? 1 : 0) << 63) | (piece << 48) | (src & LANE_SRC);
}

static inline void atomic_or_u64(uint64_t* p, uint64_t m)
{
    __atomic_fetch_or(p, m, __ATOMIC_RELAXED);
}

static inline void set_bit(uint64_t* bm, uint64_t i)
{
    atomic_or_u64(&bm[i >> 6], 1ull << (i & 63));
}

/* returns 1 if the bit was already set (source/phase_edge_pull_cc.c:37-47) */
static inline int cond_set_bit(uint64_t* bm, uint64_t i)
{
    const uint64_t m = 1ull << (i & 63);
    if (__atomic_load_n(&bm[i >> 6], __ATOMIC_RELAXED) & m)
        return 1;
    atomic_or_u64(&bm[i >> 6], m);
    return 0;
}

static inline void clear_bit(uint64_t* bm, uint64_t i)
{
    __atomic_fetch_and(&bm[i >> 6], ~(1ull << (i & 63)), __ATOMIC_RELAXED);
}

static inline int test_bit(const uint64_t* bm, uint64_t i)
{
    return (int)((__atomic_load_n(&bm[i >> 6], __ATOMIC_RELAXED) >> (i & 63)) & 1ull);
}

/* source/phase_edge_pull_bfs.c:86-103: first set bit strictly after start */
static uint64_t find_next_set_bit(const uint64_t* bm, uint64_t start, uint64_t count64)
{
    const uint64_t initmask = ~((2ull << (start & 63ull)) - 1ull);
    uint64_t cur = start >> 6;
    uint64_t bits = __atomic_load_n(&bm[cur], __ATOMIC_RELAXED) & initmask;
    while (bits == 0) {
        cur += 1;
        if (cur == count64)
            return UINT64_MAX;
        bits = __atomic_load_n(&bm[cur], __ATOMIC_RELAXED);
    }
    return (cur << 6) + (uint64_t)__builtin_ctzll(bits);
}

static void unit_range(uint64_t total, uint64_t nunits, uint64_t unit, uint64_t* b, uint64_t* e)
{
    const uint64_t per = total / nunits, extra = total % nunits;
    *b = per * unit + (unit < extra ? unit : extra);
    *e = *b + per + (unit < extra ? 1 : 0);
}

/* dynamic schedule: unit = tid first, then fetch-add on a shared counter
 * initialised to the thread count (source/phase_edge_pull_pr.c:43-50) */
static void sched_begin(wcpu* w, int tid)
{
    if (tid == 0)
        atomic_store(&w->sched, (uint64_t)w->nthreads);
    pthread_barrier_wait(&w->bar);
}

static inline uint64_t sched_next(wcpu* w)
{
    return atomic_fetch_add(&w->sched, 1ull);
}

/* ---- construction ---------------------------------------------------------- */

wcpu* wcpu_create(uint64_t V, uint64_t E, const uint32_t* off, const uint32_t* src, const void* weights,
                  int wkind, int nthreads)
{
    if (V == 0 || !off || (E && !src) || (wkind != WCPU_WEIGHT_NONE && !weights))
        return NULL;
    wcpu* w = (wcpu*)calloc(1, sizeof(wcpu));
    if (!w)
        return NULL;
    w->V = V;
    w->E = E;
    w->wkind = wkind;
    w->vstride = (wkind == WCPU_WEIGHT_NONE) ? 4 : 8;
    if (nthreads <= 0) {
        long n = sysconf(_SC_NPROCESSORS_ONLN);
        nthreads = n > 0 ? (int)n : 1;
    }
    w->nthreads = nthreads;
    pthread_mutex_init(&w->gate_mu, NULL);
    pthread_cond_init(&w->gate_cv, NULL);
    /* vectors: ceil(indegree / 4) per destination (source/graphdata.c:1032-1035) */
    uint64_t nv = 0;
    for (uint64_t v = 0; v < V; ++v)
        nv += ((uint64_t)(off[v + 1] - off[v]) + 3) >> 2;
    w->nvec = nv;
    /* round up to a multiple of the frontier precision (source/graphdata.c:1071) */
    w->nvec_alloc = nv + (PREC - (nv & (PREC - 1)));
    w->vpad = V + (8 - (V & 7));
    w->vecs = (uint64_t*)aligned_alloc(64, ((w->nvec_alloc * w->vstride * 8 + 63) / 64) * 64);
    w->vindex = (uint64_t*)malloc(w->vpad * sizeof(uint64_t));
    w->outdeg = (uint64_t*)calloc(w->vpad, sizeof(uint64_t));
    if (!w->vecs || !w->vindex || !w->outdeg) {
        wcpu_destroy(w);
        return NULL;
    }
    memset(w->vindex, 0xff, w->vpad * sizeof(uint64_t));
    uint64_t vi = 0;
    for (uint64_t v = 0; v < V; ++v) {
        const uint64_t b = off[v], n = off[v + 1] - off[v];
        if (n)
            w->vindex[v] = vi;
        for (uint64_t j = 0; j < n; j += 4, ++vi) {
            uint64_t* l = w->vecs + vi * w->vstride;
            for (int k = 0; k < 4; ++k) {
                const int valid = (j + (uint64_t)k) < n;
                const uint64_t s = valid ? src[b + j + k] : 0;
                l[k] = lane_encode(v, k, valid, s);
                if (w->vstride == 8) {
                    uint64_t wt = 0;
                    if (valid) {
                        if (wkind == WCPU_WEIGHT_F64)
                            memcpy(&wt, &((const double*)weights)[b + j + k], 8);
                        else
                            wt = ((const uint32_t*)weights)[b + j + k];
                    }
                    l[4 + k] = wt;
                }
            }
        }
    }
    /* tail padding: destination V-1, all lanes invalid (source/graphdata.c:1073-1093) */
    for (; vi < w->nvec_alloc; ++vi) {
        uint64_t* l = w->vecs + vi * w->vstride;
        for (int k = 0; k < 4; ++k) {
            l[k] = lane_encode(V - 1, k, 0, 0);
            if (w->vstride == 8)
                l[4 + k] = 0;
        }
    }
    /* out-degrees from valid lanes (source/graphdata.c:430-490) */
    for (uint64_t i = 0; i < nv; ++i) {
        const uint64_t* l = vec_at(w, i);
        for (int k = 0; k < 4; ++k)
            if (l[k] & LANE_VALID)
                w->outdeg[l[k] & LANE_SRC]++;
    }
    /* CSRI (source/graphdata.c:493-585) */
    w->csri_index = (uint64_t*)malloc((V + 1) * sizeof(uint64_t));
    w->csri_edges = (uint64_t*)malloc((nv * 4 + 1) * sizeof(uint64_t));
    if (!w->csri_index || !w->csri_edges) {
        wcpu_destroy(w);
        return NULL;
    }
    uint64_t run = 0;
    for (uint64_t v = 0; v < V; ++v) {
        w->csri_index[v] = run;
        run += w->outdeg[v];
    }
    w->csri_index[V] = run;
    for (uint64_t i = 0; i < nv; ++i) {
        const uint64_t* l = vec_at(w, i);
        for (int k = 0; k < 4; ++k)
            if (l[k] & LANE_VALID)
                w->csri_edges[w->csri_index[l[k] & LANE_SRC]++] = i >> LOG2_PREC;
    }
    run = 0;
    for (uint64_t v = 0; v < V; ++v) {
        w->csri_index[v] = run;
        run += w->outdeg[v];
    }
    w->csri_index[V] = run;
    /* Wedge frontier sizing (source/graphdata.c:810-819) */
    const uint64_t blocks = (nv >> (9 + LOG2_PREC)) + 1;
    w->wedge_words = blocks << 3;
    w->wedge = (uint64_t*)aligned_alloc(64, w->wedge_words * 8);
    /* traditional frontiers (source/graphdata.c:846-891) */
    const uint64_t fbytes = ((V >> 9) + ((V & 511) ? 1 : 0)) << 6;
    w->frontier_words = fbytes >> 3;
    w->frontier_a = (uint64_t*)aligned_alloc(64, fbytes);
    w->frontier_b = (uint64_t*)aligned_alloc(64, fbytes);
    w->unconverged = (uint64_t*)aligned_alloc(64, fbytes);
    w->nunits = (uint64_t)nthreads * UNITS_PER_THREAD;
    w->merge = (merge_t*)calloc(w->nunits, sizeof(merge_t));
    w->reduce = (uint64_t*)calloc((uint64_t)nthreads + 8, sizeof(uint64_t));
    w->props = (prop_t*)malloc(w->vpad * sizeof(prop_t));
    w->acc = (prop_t*)malloc(w->vpad * sizeof(prop_t));
    w->od_d = (double*)malloc(w->vpad * sizeof(double));
    w->depth = (uint32_t*)malloc(w->vpad * sizeof(uint32_t));
    if (!w->wedge || !w->frontier_a || !w->frontier_b || !w->unconverged || !w->merge || !w->reduce ||
        !w->props || !w->acc || !w->od_d || !w->depth) {
        wcpu_destroy(w);
        return NULL;
    }
    return w;
}

void wcpu_destroy(wcpu* w)
{
    if (!w)
        return;
    free(w->vecs);
    free(w->vindex);
    free(w->outdeg);
    free(w->csri_index);
    free(w->csri_edges);
    free(w->wedge);
    free(w->frontier_a);
    free(w->frontier_b);
    free(w->unconverged);
    free(w->merge);
    free(w->reduce);
    free(w->props);
    free(w->acc);
    free(w->od_d);
    free(w->depth);
    pthread_mutex_destroy(&w->gate_mu);
    pthread_cond_destroy(&w->gate_cv);
    free(w);
}

int wcpu_threads(const wcpu* w)
{
    return w ? w->nthreads : 0;
}

uint64_t wcpu_num_vectors(const wcpu* w)
{
    return w ? w->nvec : 0;
}

int wcpu_outdegrees(const wcpu* w, uint64_t* out)
{
    if (!w || !out)
        return -1;
    memcpy(out, w->outdeg, w->V * sizeof(uint64_t));
    return 0;
}

/* ---- SPMD runner ---------------------------------------------------------- */

typedef struct {
    wcpu* w;
    int tid;
    void (*fn)(wcpu*, int);
} spmd_arg;

static void* spmd_entry(void* p)
{
    spmd_arg* a = (spmd_arg*)p;
    wcpu* w = a->w;
    pthread_mutex_lock(&w->gate_mu);
    while (w->gate_state == 0)
        pthread_cond_wait(&w->gate_cv, &w->gate_mu);
    const int go = (w->gate_state == 1);
    pthread_mutex_unlock(&w->gate_mu);
    if (go)
        a->fn(w, a->tid);
    return NULL;
}

static int spmd_run(wcpu* w, void (*fn)(wcpu*, int))
{
    /* start the workers first; the barrier is sized only once every thread
     * exists, so a failed pthread_create can never leave a barrier unmet */
    int T = w->nthreads;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)T);
    spmd_arg* args = (spmd_arg*)malloc(sizeof(spmd_arg) * (size_t)T);
    if (!th || !args) {
        free(th);
        free(args);
        return -1;
    }
    pthread_mutex_lock(&w->gate_mu);
    w->gate_state = 0;
    pthread_mutex_unlock(&w->gate_mu);
    int started = 1;
    for (int t = 1; t < T; ++t) {
        args[t].w = w;
        args[t].tid = t;
        args[t].fn = fn;
        if (pthread_create(&th[t], NULL, spmd_entry, &args[t]) != 0)
            break;
        started = t + 1;
    }
    const int ok = (started == T);
    if (ok)
        pthread_barrier_init(&w->bar, NULL, (unsigned)T);
    pthread_mutex_lock(&w->gate_mu);
    w->gate_state = ok ? 1 : 2;
    pthread_cond_broadcast(&w->gate_cv);
    pthread_mutex_unlock(&w->gate_mu);
    if (ok)
        fn(w, 0);
    for (int t = 1; t < started; ++t)
        pthread_join(th[t], NULL);
    if (ok)
        pthread_barrier_destroy(&w->bar);
    free(th);
    free(args);
    return ok ? 0 : -1;
}

static void static_chunk(uint64_t total, int T, int tid, uint64_t* b, uint64_t* e)
{
    unit_range(total, (uint64_t)T, (uint64_t)tid, b, e);
}

/* ---- Wedge frontier (source/wedge.c) -------------------------------------- */

static void wedge_activate(wcpu* w, uint64_t v)
{
    for (uint64_t i = w->csri_index[v]; i < w->csri_index[v + 1]; ++i) {
        const uint64_t g = w->csri_edges[i];
        if (!test_bit(w->wedge, g))
            set_bit(w->wedge, g);
    }
}

static void wedge_generate(wcpu* w, int tid)
{
    const uint64_t total = w->frontier_words >> 3;
    sched_begin(w, tid);
    for (uint64_t unit = (uint64_t)tid; unit < w->nunits; unit = sched_next(w)) {
        uint64_t b, e;
        unit_range(total, w->nunits, unit, &b, &e);
        for (uint64_t i = b; i < e; ++i) {
            for (uint64_t j = 0; j < 8; ++j) {
                uint64_t bits = w->active[(i << 3) + j];
                while (bits) {
                    const uint64_t v = (((i << 3) + j) << 6) + (uint64_t)__builtin_ctzll(bits);
                    wedge_activate(w, v);
                    bits &= bits - 1;
                }
            }
        }
    }
}

/* frontier-mode scan of one 8-word block (512 bits) of the Wedge frontier:
 * calls visit(w, vector_index, ctx) for every vector of every set bit, after
 * consuming (clearing) the word (source/phase_edge_pull_cc.c:200-216) */
static inline void wedge_first_last(const wcpu* w, uint64_t fb, uint64_t* js, uint64_t* je)
{
    const uint64_t* f = &w->wedge[fb];
    *js = (f[0] | f[1] | f[2] | f[3]) ? 0 : 4;
    *je = (f[4] | f[5] | f[6] | f[7]) ? 8 : 4;
}

/* ---- PageRank ---------------------------------------------------------------- */

static inline double hsum4(const double a[4])
{
    /* _mm256_hadd_pd then add of the halves: (a0 + a1) + (a2 + a3) */
    return (a[0] + a[1]) + (a[2] + a[3]);
}

/* source/phase_edge_pull_pr.c:28-116 */
static void pr_edge_phase(wcpu* w, int tid)
{
    double total = 0.0;
    sched_begin(w, tid);
    for (uint64_t unit = (uint64_t)tid; unit < w->nunits; unit = sched_next(w)) {
        uint64_t b, e;
        unit_range(w->nvec, w->nunits, unit, &b, &e);
        if (b >= e) {
            w->merge[unit].final = MERGE_INVALID;
            continue;
        }
        double acc[4] = { 0, 0, 0, 0 };
        uint64_t prev = vec_dst(vec_at(w, b));
        w->merge[unit].initial = prev;
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t* l = vec_at(w, i);
            double g[4];
            for (int k = 0; k < 4; ++k)
                g[k] = (l[k] & LANE_VALID) ? w->props[l[k] & LANE_SRC].d : 0.0;
            const uint64_t dst = vec_dst(l);
            if (prev < dst) {
                const double part = hsum4(acc);
                w->acc[prev].d = part;
                total += part;
                acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
            }
            prev = dst;
            for (int k = 0; k < 4; ++k)
                acc[k] += g[k];
        }
        w->merge[unit].final = prev;
        const double part = hsum4(acc);
        w->merge[unit].partial.d = part;
        total += part;
    }
    memcpy(&w->reduce[tid], &total, sizeof(double));
}

/* source/phases.c:69-114 */
static void pr_merge(wcpu* w)
{
    prop_t proposed = { 0 };
    uint64_t last = UINT64_MAX;
    for (uint64_t i = 0; i < w->nunits; ++i) {
        merge_t* m = &w->merge[i];
        if (m->final & MERGE_INVALID)
            continue;
        if (last == UINT64_MAX) {
            last = m->final;
            proposed = m->partial;
        } else if (last == m->final) {
            proposed.d = proposed.d + m->partial.d;
        } else {
            if (last == m->initial)
                proposed.d = proposed.d + w->acc[last].d;
            w->acc[last] = proposed;
            last = m->final;
            proposed = m->partial;
        }
        m->final = MERGE_INVALID;
    }
    if (last != UINT64_MAX)
        w->acc[last] = proposed;
}

/* source/phase_vertex_pr.c:26-77 */
static void pr_vertex_phase(wcpu* w, int tid)
{
    const double N = (double)w->V;
    const double d = w->damping;
    const double omd_n = (1.0 - d) / N;
    double total = 0.0;
    for (int t = 0; t < w->nthreads; ++t) {
        double x;
        memcpy(&x, &w->reduce[t], sizeof(double));
        total += x;
    }
    const double corr = (1.0 - total) / N;
    const uint64_t groups = (w->V >> 2) + ((w->V & 3) ? 1 : 0);
    uint64_t b, e;
    static_chunk(groups, w->nthreads, tid, &b, &e);
    for (uint64_t gi = b; gi < e; ++gi) {
        for (uint64_t v = gi << 2; v < (gi << 2) + 4; ++v) {
            const double od = (w->od_d[v] == 0.0) ? N : w->od_d[v];
            const double t1 = w->acc[v].d + corr;
            const double t2 = t1 * d;
            const double t3 = t2 + omd_n;
            w->props[v].d = t3 / od;
        }
    }
    if (tid == 0)
        w->stats.special = total;
}

static void pr_driver(wcpu* w, int tid)
{
    double t0 = 0.0;
    for (uint32_t it = 0; it < w->iterations; ++it) {
        if (tid == 0)
            t0 = now_ms();
        pr_edge_phase(w, tid);
        pthread_barrier_wait(&w->bar);
        if (tid == 0)
            pr_merge(w);
        pthread_barrier_wait(&w->bar);
        pr_vertex_phase(w, tid);
        pthread_barrier_wait(&w->bar);
        if (tid == 0)
            w->stats.dense_pull_ms += now_ms() - t0;
    }
}

int wcpu_pr(wcpu* w, uint32_t iterations, double damping, double* ranks, wcpu_stats* st)
{
    if (!w || !ranks || iterations == 0)
        return -1;
    memset(&w->stats, 0, sizeof(w->stats));
    const double N = (double)w->V;
    /* source/execution_pr.c:73-85 */
    for (uint64_t v = 0; v < w->vpad; ++v) {
        w->od_d[v] = (double)w->outdeg[v];
        w->acc[v].d = 0.0;
        w->props[v].d = (1.0 / N) / (w->od_d[v] == 0.0 ? N : w->od_d[v]);
    }
    for (uint64_t i = 0; i < w->nunits; ++i)
        w->merge[i].final = MERGE_INVALID;
    w->iterations = iterations;
    w->damping = damping;
    if (spmd_run(w, pr_driver) != 0)
        return -1;
    double sum = 0.0;
    for (uint64_t v = 0; v < w->V; ++v) {
        ranks[v] = w->props[v].d * (w->od_d[v] == 0.0 ? N : w->od_d[v]);
        sum += ranks[v];
    }
    w->stats.iterations = iterations;
    w->stats.pull_ms = w->stats.dense_pull_ms;
    w->stats.running_ms = w->stats.pull_ms;
    w->stats.special = sum;
    if (st)
        *st = w->stats;
    return 0;
}

/* ---- BFS --------------------------------------------------------------------- */

/* per-vector bottom-up test shared by both modes (source/phase_edge_pull_bfs.c:162-188);
 * returns 1 if dst was visited (caller must skip ahead) */
static inline int bfs_vector(wcpu* w, const uint64_t* l, uint64_t dst, uint64_t* changed, uint32_t level)
{
    if (__atomic_load_n(&w->props[dst].u, __ATOMIC_RELAXED) != UINT64_MAX)
        return 1;
    int any = 0;
    uint64_t best = 0;
    for (int k = 0; k < 4; ++k) {
        if ((l[k] & LANE_VALID) && test_bit(w->snapshot, l[k] & LANE_SRC)) {
            any = 1;
            const uint64_t s = l[k] & LANE_SRC;
            if (s > best)
                best = s;
        }
    }
    if (!any)
        return 0;
    __atomic_store_n(&w->props[dst].u, best, __ATOMIC_RELAXED);
    if (__atomic_load_n(&w->depth[dst], __ATOMIC_RELAXED) == UINT32_MAX)
        __atomic_store_n(&w->depth[dst], level, __ATOMIC_RELAXED);
    set_bit(w->active, dst);
    clear_bit(w->unconverged, dst);
    *changed += w->outdeg[dst];
    return 0;
}

/* source/phase_edge_pull_bfs.c:109-210 */
static void bfs_dense(wcpu* w, int tid, uint32_t level)
{
    uint64_t changed = 0, last_found = 0;
    sched_begin(w, tid);
    for (uint64_t unit = (uint64_t)tid; unit < w->nunits; unit = sched_next(w)) {
        uint64_t b, e;
        unit_range(w->nvec, w->nunits, unit, &b, &e);
        if (last_found > e)
            continue;
        const uint64_t start = last_found > b ? last_found : b;
        for (uint64_t i = start; i < e; ++i) {
            const uint64_t* l = vec_at(w, i);
            const uint64_t dst = vec_dst(l);
            if (bfs_vector(w, l, dst, &changed, level)) {
                const uint64_t next = find_next_set_bit(w->unconverged, dst, w->frontier_words);
                if (next >= w->V) {
                    w->reduce[tid] = changed;
                    return;
                }
                last_found = w->vindex[next];
                i = last_found - 1;
            }
        }
    }
    w->reduce[tid] = changed;
}

/* source/phase_edge_pull_bfs.c:212-332 */
static void bfs_frontier(wcpu* w, int tid, uint32_t level)
{
    uint64_t changed = 0, last_found = 0;
    const uint64_t total = w->wedge_words >> 3;
    sched_begin(w, tid);
    for (uint64_t unit = (uint64_t)tid; unit < w->nunits; unit = sched_next(w)) {
        uint64_t b, e;
        unit_range(total, w->nunits, unit, &b, &e);
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t fb = i << 3;
            uint64_t js, je;
            wedge_first_last(w, fb, &js, &je);
            for (uint64_t j = js; j < je; ++j) {
                const uint64_t fi = fb + j;
                uint64_t bits = w->wedge[fi];
                if (!bits)
                    continue;
                w->wedge[fi] = 0;
                if (((fi + 64ull) << (6ull + LOG2_PREC)) < last_found)
                    continue;
                for (; bits; bits &= bits - 1) {
                    const uint64_t vidx = (fi << (6ull + LOG2_PREC)) + ((uint64_t)__builtin_ctzll(bits) << LOG2_PREC);
                    for (uint64_t k = 0; k < PREC; ++k) {
                        if (vidx + k < last_found)
                            continue;
                        const uint64_t* l = vec_at(w, vidx + k);
                        const uint64_t dst = vec_dst(l);
                        if (bfs_vector(w, l, dst, &changed, level)) {
                            const uint64_t next = find_next_set_bit(w->unconverged, dst, w->frontier_words);
                            if (next >= w->V) {
                                w->reduce[tid] = changed;
                                return;
                            }
                            last_found = w->vindex[next];
                        }
                    }
                }
            }
        }
    }
    w->reduce[tid] = changed;
}

static void clear_words(uint64_t* p, uint64_t words, int T, int tid)
{
    uint64_t b, e;
    static_chunk(words, T, tid, &b, &e);
    if (e > b)
        memset(p + b, 0, (e - b) * sizeof(uint64_t));
}

/* source/execution_bfs.c:122-365 */
static void bfs_driver(wcpu* w, int tid)
{
    const uint64_t threshold = (w->E * (uint64_t)w->threshold_pct) / 100ull;
    uint64_t active_edges = w->outdeg[w->root];
    uint32_t ctr = 0;
    double t0 = 0.0;
    for (;;) {
        const int dense = active_edges > threshold;
        ctr += 1;
        if (tid == 0)
            t0 = now_ms();
        if (dense)
            bfs_dense(w, tid, ctr);
        else
            bfs_frontier(w, tid, ctr);
        pthread_barrier_wait(&w->bar);
        if (tid == 0) {
            const double dt = now_ms() - t0;
            if (dense) {
                w->stats.dense_pull_ms += dt;
            } else {
                w->stats.frontier_pull_ms += dt;
                w->stats.frontier_iterations += 1;
            }
            w->stats.iterations = ctr;
        }
        active_edges = 0;
        for (int t = 0; t < w->nthreads; ++t)
            active_edges += w->reduce[t];
        if (active_edges == 0)
            break;
        pthread_barrier_wait(&w->bar);
        if (tid == 0)
            t0 = now_ms();
        if (active_edges <= threshold)
            wedge_generate(w, tid);
        pthread_barrier_wait(&w->bar);
        if (tid == 0) {
            w->stats.wedge_ms += now_ms() - t0;
            t0 = now_ms();
            uint64_t* tmp = w->active;
            w->active = w->snapshot;
            w->snapshot = tmp;
        }
        pthread_barrier_wait(&w->bar);
        clear_words(w->active, w->frontier_words, w->nthreads, tid);
        pthread_barrier_wait(&w->bar);
        if (tid == 0) {
            const double dt = now_ms() - t0;
            if (dense)
                w->stats.dense_pull_ms += dt;
            else
                w->stats.frontier_pull_ms += dt;
        }
    }
}

static void reset_frontiers(wcpu* w)
{
    memset(w->wedge, 0, w->wedge_words * 8);
    memset(w->frontier_a, 0, w->frontier_words * 8);
    memset(w->frontier_b, 0, w->frontier_words * 8);
    w->active = w->frontier_a;
    w->snapshot = w->frontier_b;
    for (uint64_t i = 0; i < w->nunits; ++i)
        w->merge[i].final = MERGE_INVALID;
}

int wcpu_bfs(wcpu* w, uint64_t root, uint32_t threshold_pct, uint64_t* parent, uint32_t* depth, wcpu_stats* st)
{
    if (!w || root >= w->V)
        return -1;
    memset(&w->stats, 0, sizeof(w->stats));
    reset_frontiers(w);
    /* source/execution_bfs.c:84-118 */
    memset(w->unconverged, 0xff, w->frontier_words * 8);
    for (uint64_t v = 0; v < w->vpad; ++v) {
        if (v < w->V && w->vindex[v] == UINT64_MAX)
            clear_bit(w->unconverged, v);
        w->props[v].u = UINT64_MAX;
        w->depth[v] = UINT32_MAX;
    }
    set_bit(w->snapshot, root);
    clear_bit(w->unconverged, root);
    wedge_activate(w, root);
    w->depth[root] = 0;
    w->root = root;
    w->threshold_pct = threshold_pct;
    if (spmd_run(w, bfs_driver) != 0)
        return -1;
    for (uint64_t v = 0; v < w->V; ++v) {
        if (parent)
            parent[v] = w->props[v].u;
        if (depth)
            depth[v] = w->depth[v];
    }
    w->stats.pull_ms = w->stats.dense_pull_ms + w->stats.frontier_pull_ms;
    w->stats.running_ms = w->stats.pull_ms + w->stats.wedge_ms;
    if (st)
        *st = w->stats;
    return 0;
}

/* ---- CC and SSSP (min-propagation, in place) --------------------------------- */

enum { MIN_CC = 0, MIN_SSSP_U64 = 1, MIN_SSSP_F64 = 2 };

/* value a vector proposes for its destination: min over lanes */
static inline prop_t min_vector(const wcpu* w, const uint64_t* l, int kind)
{
    prop_t r;
    if (kind == MIN_CC) {
        double g[4];
        for (int k = 0; k < 4; ++k)
            g[k] = (l[k] & LANE_VALID) ? w->props[l[k] & LANE_SRC].d : (double)INFINITY;
        /* source/phase_edge_pull_cc.c:133-134: min of halves, then of lanes */
        const double m0 = g[2] < g[0] ? g[2] : g[0];
        const double m1 = g[3] < g[1] ? g[3] : g[1];
        r.d = m0 < m1 ? m0 : m1;
    } else if (kind == MIN_SSSP_U64) {
        uint64_t m = UINT64_MAX;
        for (int k = 0; k < 4; ++k) {
            const uint64_t p = (l[k] & LANE_VALID) ? w->props[l[k] & LANE_SRC].u : UINT64_MAX;
            const uint64_t wt = (p == UINT64_MAX) ? 0 : l[4 + k]; /* :136 */
            const uint64_t s = p + wt;
            if (s < m)
                m = s;
        }
        r.u = m;
    } else {
        double m = (double)INFINITY;
        for (int k = 0; k < 4; ++k) {
            const double p = (l[k] & LANE_VALID) ? w->props[l[k] & LANE_SRC].d : (double)INFINITY;
            double wt;
            memcpy(&wt, &l[4 + k], 8);
            const double s = (p == (double)INFINITY) ? p : p + wt;
            if (s < m)
                m = s;
        }
        r.d = m;
    }
    return r;
}

static inline prop_t min_combine(prop_t a, prop_t b, int kind)
{
    if (kind == MIN_SSSP_U64)
        return a.u < b.u ? a : b;
    return a.d < b.d ? a : b; /* _mm_min_pd(a, b) */
}

static inline prop_t min_infinity(int kind)
{
    prop_t r;
    if (kind == MIN_SSSP_U64)
        r.u = UINT64_MAX;
    else
        r.d = (double)INFINITY;
    return r;
}

/* CC changes on any bit difference (source/phase_edge_pull_cc.c:113); SSSP on
 * a strictly smaller value (source/phase_edge_pull_sssp.c:116) */
static inline int min_changes(prop_t cur, prop_t proposed, int kind)
{
    if (kind == MIN_CC)
        return cur.u != proposed.u;
    if (kind == MIN_SSSP_U64)
        return proposed.u < cur.u;
    return proposed.d < cur.d;
}

static inline void min_flush(wcpu* w, uint64_t v, prop_t accv, int kind, uint64_t* changed)
{
    const prop_t cur = w->props[v];
    const prop_t proposed = (kind == MIN_CC) ? min_combine(cur, accv, kind) : accv;
    if (min_changes(cur, proposed, kind)) {
        w->props[v] = proposed;
        if (!cond_set_bit(w->active, v))
            *changed += w->outdeg[v];
    }
}

static inline void min_unit_end(wcpu* w, uint64_t unit, uint64_t v, prop_t accv, int kind, uint64_t* changed)
{
    const prop_t cur = w->props[v];
    prop_t proposed;
    if (kind == MIN_CC)
        proposed = min_combine(cur, accv, kind);
    else
        proposed = min_changes(cur, accv, kind) ? accv : cur;
    w->merge[unit].final = v;
    w->merge[unit].partial = proposed;
    if (min_changes(cur, (kind == MIN_CC) ? proposed : accv, kind)) {
        if (!cond_set_bit(w->active, v))
            *changed += w->outdeg[v];
    }
}

static void min_dense(wcpu* w, int tid, int kind)
{
    uint64_t changed = 0;
    sched_begin(w, tid);
    for (uint64_t unit = (uint64_t)tid; unit < w->nunits; unit = sched_next(w)) {
        uint64_t b, e;
        unit_range(w->nvec, w->nunits, unit, &b, &e);
        prop_t accv = min_infinity(kind);
        uint64_t prev = UINT64_MAX;
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t* l = vec_at(w, i);
            const prop_t m = min_vector(w, l, kind);
            const uint64_t dst = vec_dst(l);
            if (prev < dst) {
                min_flush(w, prev, accv, kind, &changed);
                accv = min_infinity(kind);
            }
            prev = dst;
            accv = min_combine(m, accv, kind);
        }
        if (prev != UINT64_MAX)
            min_unit_end(w, unit, prev, accv, kind, &changed);
        else
            w->merge[unit].final = MERGE_INVALID;
    }
    w->reduce[tid] = changed;
}

static void min_frontier(wcpu* w, int tid, int kind)
{
    uint64_t changed = 0;
    const uint64_t total = w->wedge_words >> 3;
    sched_begin(w, tid);
    for (uint64_t unit = (uint64_t)tid; unit < w->nunits; unit = sched_next(w)) {
        uint64_t b, e;
        unit_range(total, w->nunits, unit, &b, &e);
        prop_t accv = min_infinity(kind);
        uint64_t prev = UINT64_MAX;
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t fb = i << 3;
            uint64_t js, je;
            wedge_first_last(w, fb, &js, &je);
            for (uint64_t j = js; j < je; ++j) {
                const uint64_t fi = fb + j;
                uint64_t bits = w->wedge[fi];
                if (!bits)
                    continue;
                w->wedge[fi] = 0;
                for (; bits; bits &= bits - 1) {
                    const uint64_t vidx = (fi << (6ull + LOG2_PREC)) + ((uint64_t)__builtin_ctzll(bits) << LOG2_PREC);
                    for (uint64_t k = 0; k < PREC; ++k) {
                        const uint64_t* l = vec_at(w, vidx + k);
                        const prop_t m = min_vector(w, l, kind);
                        const uint64_t dst = vec_dst(l);
                        if (prev < dst) {
                            min_flush(w, prev, accv, kind, &changed);
                            accv = min_infinity(kind);
                        }
                        prev = dst;
                        accv = min_combine(m, accv, kind);
                    }
                }
            }
        }
        if (prev != UINT64_MAX)
            min_unit_end(w, unit, prev, accv, kind, &changed);
    }
    w->reduce[tid] = changed;
}

/* source/phases.c:54-65 */
static void min_merge(wcpu* w, int kind)
{
    for (uint64_t i = 0; i < w->nunits; ++i) {
        merge_t* m = &w->merge[i];
        if (m->final & MERGE_INVALID)
            continue;
        w->props[m->final] = min_combine(w->props[m->final], m->partial, kind);
        m->final = MERGE_INVALID;
    }
}

/* source/execution_cc.c:83-325 and source/execution_sssp.c:93-334 */
static void min_driver(wcpu* w, int tid)
{
    const int kind = w->kind;
    const uint64_t threshold = (w->E * (uint64_t)w->threshold_pct) / 100ull;
    uint64_t active_edges = (kind == MIN_CC) ? w->E : w->outdeg[w->root];
    uint64_t ctr = 0;
    double t0 = 0.0;
    for (;;) {
        const int dense = (active_edges > threshold) || (kind == MIN_CC && ctr == 0);
        ctr += 1;
        if (tid == 0)
            t0 = now_ms();
        if (dense)
            min_dense(w, tid, kind);
        else
            min_frontier(w, tid, kind);
        pthread_barrier_wait(&w->bar);
        if (tid == 0)
            min_merge(w, kind);
        pthread_barrier_wait(&w->bar);
        if (tid == 0) {
            const double dt = now_ms() - t0;
            if (dense) {
                w->stats.dense_pull_ms += dt;
            } else {
                w->stats.frontier_pull_ms += dt;
                w->stats.frontier_iterations += 1;
            }
            w->stats.iterations = ctr;
        }
        active_edges = 0;
        for (int t = 0; t < w->nthreads; ++t)
            active_edges += w->reduce[t];
        if (active_edges == 0)
            break;
        pthread_barrier_wait(&w->bar);
        if (tid == 0)
            t0 = now_ms();
        if (active_edges <= threshold) {
            wedge_generate(w, tid);
            pthread_barrier_wait(&w->bar);
            /* clear the non-zero 256-bit chunks (source/execution_cc.c:260-272) */
            uint64_t b, e;
            static_chunk(w->frontier_words >> 2, w->nthreads, tid, &b, &e);
            for (uint64_t i = b; i < e; ++i) {
                uint64_t* p = &w->active[i << 2];
                if (p[0] | p[1] | p[2] | p[3])
                    p[0] = p[1] = p[2] = p[3] = 0;
            }
        } else {
            clear_words(w->active, w->frontier_words, w->nthreads, tid);
        }
        pthread_barrier_wait(&w->bar);
        if (tid == 0)
            w->stats.wedge_ms += now_ms() - t0;
        pthread_barrier_wait(&w->bar);
    }
}

int wcpu_cc(wcpu* w, uint32_t threshold_pct, double* labels, wcpu_stats* st)
{
    if (!w || !labels)
        return -1;
    memset(&w->stats, 0, sizeof(w->stats));
    reset_frontiers(w);
    for (uint64_t v = 0; v < w->vpad; ++v)
        w->props[v].d = (double)v; /* source/execution_cc.c:77 */
    w->threshold_pct = threshold_pct;
    w->root = 0;
    w->kind = MIN_CC;
    if (spmd_run(w, min_driver) != 0)
        return -1;
    for (uint64_t v = 0; v < w->V; ++v)
        labels[v] = w->props[v].d;
    w->stats.pull_ms = w->stats.dense_pull_ms + w->stats.frontier_pull_ms;
    w->stats.running_ms = w->stats.pull_ms + w->stats.wedge_ms;
    if (st)
        *st = w->stats;
    return 0;
}

int wcpu_sssp(wcpu* w, uint64_t root, uint32_t threshold_pct, void* dist, wcpu_stats* st)
{
    if (!w || !dist || root >= w->V || w->wkind == WCPU_WEIGHT_NONE)
        return -1;
    memset(&w->stats, 0, sizeof(w->stats));
    reset_frontiers(w);
    const int kind = (w->wkind == WCPU_WEIGHT_F64) ? MIN_SSSP_F64 : MIN_SSSP_U64;
    /* source/execution_sssp.c:70-89 */
    for (uint64_t v = 0; v < w->vpad; ++v)
        w->props[v] = min_infinity(kind);
    if (kind == MIN_SSSP_F64)
        w->props[root].d = 0.0;
    else
        w->props[root].u = 0;
    wedge_activate(w, root);
    w->root = root;
    w->threshold_pct = threshold_pct;
    w->kind = kind;
    if (spmd_run(w, min_driver) != 0)
        return -1;
    for (uint64_t v = 0; v < w->V; ++v) {
        if (kind == MIN_SSSP_F64)
            ((double*)dist)[v] = w->props[v].d;
        else
            ((uint64_t*)dist)[v] = w->props[v].u;
    }
    w->stats.pull_ms = w->stats.dense_pull_ms + w->stats.frontier_pull_ms;
    w->stats.running_ms = w->stats.pull_ms + w->stats.wedge_ms;
    if (st)
        *st = w->stats;
    return 0;
}

/* ---- personalised PageRank ------------------------------------------------------ */

/* One seed, serially: the same operations in the same order whatever thread
 * runs it, so results do not depend on the thread count. */
static void ppr_seed(const wcpu* w, uint64_t s, uint32_t iterations, double damping, double* c, double* acc,
                     double* out)
{
    const uint64_t V = w->V;
    memset(c, 0, V * sizeof(double));
    c[s] = w->outdeg[s] ? 1.0 / (double)w->outdeg[s] : 0.0;
    double D = w->outdeg[s] ? 0.0 : 1.0;
    for (uint32_t t = 0; t < iterations; ++t) {
        memset(acc, 0, V * sizeof(double));
        for (uint64_t k = 0; k < w->nvec; ++k) {
            const uint64_t* l = vec_at(w, k);
            const uint64_t dst = vec_dst(l);
            for (int q = 0; q < 4; ++q)
                if (l[q] & LANE_VALID)
                    acc[dst] += c[l[q] & LANE_SRC];
        }
        double Dn = 0.0;
        for (uint64_t v = 0; v < V; ++v) {
            const double a = (v == s) ? acc[v] + D : acc[v];
            const double r = (v == s) ? a * damping + (1.0 - damping) : a * damping;
            c[v] = w->outdeg[v] ? r / (double)w->outdeg[v] : 0.0;
            if (!w->outdeg[v])
                Dn += r;
            if (t + 1 == iterations)
                out[v] = r;
        }
        D = Dn;
    }
}

typedef struct {
    const wcpu* w;
    const uint64_t* seeds;
    uint32_t nseeds, iterations, tid, nthreads;
    double damping;
    double* ranks;
    int err;
} ppr_arg;

/* Seeds tid, tid + nthreads, ... with this thread's own buffers. */
static void* ppr_worker(void* p)
{
    ppr_arg* a = (ppr_arg*)p;
    const uint64_t V = a->w->V;
    double* c = (double*)malloc(V * sizeof(double));
    double* acc = (double*)malloc(V * sizeof(double));
    if (!c || !acc) {
        a->err = 1;
    } else {
        for (uint32_t i = a->tid; i < a->nseeds; i += a->nthreads)
            ppr_seed(a->w, a->seeds[i], a->iterations, a->damping, c, acc, a->ranks + (uint64_t)i * V);
    }
    free(c);
    free(acc);
    return NULL;
}

int wcpu_ppr(wcpu* w, const uint64_t* seeds, uint32_t nseeds, uint32_t iterations, double damping, double* ranks)
{
    if (!w || !seeds || !ranks || nseeds == 0 || iterations == 0)
        return -1;
    const uint64_t V = w->V;
    for (uint32_t i = 0; i < nseeds; ++i)
        if (seeds[i] >= V)
            return -1;
    /* seeds are independent: one per thread at a time */
    uint32_t T = (uint32_t)(w->nthreads > 0 ? w->nthreads : 1);
    if (T > nseeds)
        T = nseeds;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * T);
    ppr_arg* args = (ppr_arg*)malloc(sizeof(ppr_arg) * T);
    if (!th || !args) {
        free(th);
        free(args);
        return -1;
    }
    uint32_t started = 0;
    for (uint32_t t = 0; t < T; ++t) {
        args[t] = (ppr_arg){ w, seeds, nseeds, iterations, t, T, damping, ranks, 0 };
        if (t > 0 && pthread_create(&th[t], NULL, ppr_worker, &args[t]) != 0)
            break;
        started = t + 1;
    }
    int err = 0;
    if (started < T) { /* a thread could not start: run every seed here instead */
        for (uint32_t t = 1; t < started; ++t)
            pthread_join(th[t], NULL);
        args[0].nthreads = 1;
        started = 1;
    }
    ppr_worker(&args[0]);
    for (uint32_t t = 1; t < started; ++t)
        pthread_join(th[t], NULL);
    for (uint32_t t = 0; t < started; ++t)
        err |= args[t].err;
    free(th);
    free(args);
    return err ? -1 : 0;
}
