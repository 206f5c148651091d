/*
 * wg_parallel.h -- minimal pthreads parallel-for with dynamic chunking.
 * Used by the host graph builders; the work function receives [begin, end).
 */
#ifndef WG_PARALLEL_H
#define WG_PARALLEL_H

#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>

typedef void (*wg_range_fn)(void* ctx, uint64_t begin, uint64_t end, int tid);

typedef struct {
    wg_range_fn fn;
    void* ctx;
    uint64_t n;
    uint64_t chunk;
    _Atomic uint64_t next;
} wg_pfor_state;

typedef struct {
    wg_pfor_state* st;
    int tid;
} wg_pfor_arg;

static void* wg_pfor_worker(void* p)
{
    wg_pfor_arg* a = (wg_pfor_arg*)p;
    wg_pfor_state* st = a->st;
    for (;;) {
        const uint64_t b = atomic_fetch_add_explicit(&st->next, st->chunk, memory_order_relaxed);
        if (b >= st->n)
            break;
        const uint64_t e = (st->n - b < st->chunk) ? st->n : b + st->chunk;
        st->fn(st->ctx, b, e, a->tid);
    }
    return NULL;
}

static inline void wg_parallel_for(int nthreads, uint64_t n, uint64_t chunk, wg_range_fn fn, void* ctx)
{
    if (n == 0)
        return;
    if (chunk == 0)
        chunk = 1;
    if (nthreads < 1)
        nthreads = 1;
    if ((uint64_t)nthreads > (n + chunk - 1) / chunk)
        nthreads = (int)((n + chunk - 1) / chunk);
    wg_pfor_state st;
    st.fn = fn;
    st.ctx = ctx;
    st.n = n;
    st.chunk = chunk;
    atomic_store(&st.next, 0);
    if (nthreads == 1) {
        fn(ctx, 0, n, 0);
        return;
    }
    pthread_t th[64];
    wg_pfor_arg args[64];
    if (nthreads > 64)
        nthreads = 64;
    int started = 0;
    for (int t = 1; t < nthreads; ++t) {
        args[t].st = &st;
        args[t].tid = t;
        if (pthread_create(&th[t], NULL, wg_pfor_worker, &args[t]) != 0)
            break;
        started = t;
    }
    args[0].st = &st;
    args[0].tid = 0;
    wg_pfor_worker(&args[0]);
    for (int t = 1; t <= started; ++t)
        pthread_join(th[t], NULL);
}

#endif /* WG_PARALLEL_H */
