/*
 * main_batch.c -- the front end's batched runs (no counterpart in the
 * reference, whose applications take one graph and one root):
 *   - several sources on one graph: -R v,v,... with -a bfs | sssp (one
 *     traversal for all of them), or -a ppr (personalised PageRank, seeds -R);
 *   - a batch of graphs: comma-separated -i files or -g generators, loaded as
 *     one disjoint union; PR, CC, or BFS / SSSP from vertex -r of every member.
 * Output (-o): sources mode "<vertex> <value per source>", batch mode
 * "<member> <vertex> <value>", with the per-application value formats of the
 * single runs (-1 for unreached).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "main_batch.h"

static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC_RAW, &t);
    return (double)t.tv_sec * 1e3 + (double)t.tv_nsec * 1e-6;
}

static void print_stats(double running_ms, uint64_t edges_work, uint64_t iterations, const char* what, uint64_t count)
{
    printf("\n--------------- EXECUTION STATISTICS ---------------\n");
    printf("%-28s = %.2lf msec\n", "Running Time", running_ms);
    printf("%-28s = %.4lf Bedges/sec\n", "Effective Throughput",
           running_ms > 0 ? (double)edges_work / running_ms / 1000000.0 : 0.0);
    printf("----------------------------------------------------\n");
    printf("%-28s = %llu\n", "Total Iterations", (unsigned long long)iterations);
    printf("%-28s = %llu\n", what, (unsigned long long)count);
    printf("----------------------------------------------------\n\n");
}

static void print_dist(FILE* f, int wkind, const void* col, uint64_t v)
{
    if (wkind == WG_WEIGHT_F64) {
        const double d = ((const double*)col)[v];
        if (d == (double)INFINITY)
            fprintf(f, " -1");
        else
            fprintf(f, " %.17g", d);
    } else {
        const uint32_t d = ((const uint32_t*)col)[v];
        fprintf(f, " %lld", d == 0xffffffffu ? -1ll : (long long)d);
    }
}

int wg_main_validate(wg_ctx* ctx)
{
    uint64_t bad = 0;
    const int rc = wg_validate(ctx, &bad);
    if (rc) {
        printf("Validation:  ERROR: %s\n", wg_error_string(rc));
        return 1;
    }
    if (bad) {
        printf("Validation:  FAILED (%llu violations)\n", (unsigned long long)bad);
        return 7;
    }
    printf("Validation:  passed\n");
    return 0;
}

int wg_main_sources(const cmdline_opts_t* cs, int app, wg_ctx* ctx, const wg_graph* g)
{
    const uint64_t V = g->num_vertices;
    const uint32_t n = cs->nroots;
    if (n == 0 || (app == WG_APP_PPR && n > WG_PPR_MAX_SEEDS) || (app == WG_APP_CC || app == WG_APP_PR)) {
        printf("Application: ERROR: -R takes up to %d roots for bfs/sssp and up to %d seeds for ppr\n",
               WG_MS_MAX_SOURCES, WG_PPR_MAX_SEEDS);
        return 1;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (cs->roots[i] >= V) {
            printf("Application: ERROR: root/seed %llu is out of range\n", (unsigned long long)cs->roots[i]);
            return 1;
        }
    wg_loop_params p;
    memset(&p, 0, sizeof(p));
    p.schedule = cs->schedule == WG_SCHED_LEARNED ? WG_SCHED_PLAIN : cs->schedule;
    p.threshold_pct = (uint32_t)cs->frontier_threshold_pct;
    p.block_log2 = cs->block_log2;
    printf("Application: starting execution (%u %s)\n", n, app == WG_APP_PPR ? "seeds" : "sources");
    const double t0 = now_ms();
    uint64_t iters = 0, reached = 0;
    int rc;
    if (app == WG_APP_BFS)
        rc = wg_msbfs(ctx, cs->roots, n, &p, &iters, &reached);
    else if (app == WG_APP_SSSP)
        rc = wg_mssssp(ctx, cs->roots, n, &p, &iters);
    else {
        iters = cs->num_iterations;
        rc = wg_ppr(ctx, cs->roots, n, cs->num_iterations, cs->damping);
    }
    const double running = now_ms() - t0;
    if (rc) {
        printf("Application: ERROR: %s (hip %d)\n", wg_error_string(rc), wg_last_hip_error(ctx));
        return 1;
    }
    printf("Application: completed execution\n");
    if (cs->graph_application_output_filename) {
        const size_t esz = (app == WG_APP_SSSP && g->weight_kind == WG_WEIGHT_U32) || app == WG_APP_BFS ? 4 : 8;
        char* cols = (char*)malloc((size_t)n * V * esz);
        FILE* f = cols ? fopen(cs->graph_application_output_filename, "w") : NULL;
        if (!f) {
            printf("Application: ERROR: unable to write to \"%s\"\n", cs->graph_application_output_filename);
            free(cols);
            return 1;
        }
        for (uint32_t s = 0; s < n && rc == WG_OK; ++s) {
            void* col = cols + (size_t)s * V * esz;
            rc = app == WG_APP_BFS ? wg_read_msbfs(ctx, s, (uint32_t*)col)
               : app == WG_APP_SSSP ? wg_read_mssssp(ctx, s, col)
                                    : wg_read_ppr(ctx, s, (double*)col);
        }
        for (uint64_t v = 0; v < V && rc == WG_OK && !ferror(f); ++v) {
            fprintf(f, "%llu", (unsigned long long)v);
            for (uint32_t s = 0; s < n; ++s) {
                const void* col = cols + (size_t)s * V * esz;
                if (app == WG_APP_BFS) {
                    const uint32_t d = ((const uint32_t*)col)[v];
                    fprintf(f, " %lld", d == 0xffffffffu ? -1ll : (long long)d);
                } else if (app == WG_APP_SSSP) {
                    print_dist(f, g->weight_kind, col, v);
                } else {
                    fprintf(f, cs->full_precision ? " %.17g" : " %.4le", ((const double*)col)[v]);
                }
            }
            fprintf(f, "\n");
        }
        const int bad = ferror(f) || rc;
        fclose(f);
        free(cols);
        if (bad) {
            printf("Application: ERROR: writing the output failed\n");
            return 1;
        }
    }
    print_stats(running, app == WG_APP_PPR ? g->num_edges * iters * n : reached, iters,
                app == WG_APP_PPR ? "Seeds" : "Sources", n);
    return 0;
}

/* Splits a comma-separated list in place; returns the number of items. */
static uint32_t split_list(char* s, char** items, uint32_t max)
{
    uint32_t n = 0;
    char* save = NULL;
    for (char* t = strtok_r(s, ",", &save); t && n < max; t = strtok_r(NULL, ",", &save))
        items[n++] = t;
    return n;
}

int wg_main_members(const cmdline_opts_t* cs, int app, member_loader load_one)
{
    if (app == WG_APP_PPR) {
        printf("Application: ERROR: ppr takes one graph\n");
        return 1;
    }
    char list[sizeof(cs->graph_input_filename) > sizeof(cs->generator) ? sizeof(cs->graph_input_filename)
                                                                         : sizeof(cs->generator)];
    strncpy(list, cs->generator[0] ? cs->generator : cs->graph_input_filename, sizeof(list) - 1);
    list[sizeof(list) - 1] = '\0';
    enum { MAX_MEMBERS = 4096 };
    char** items = (char**)malloc(MAX_MEMBERS * sizeof(char*));
    wg_graph* ms = (wg_graph*)calloc(MAX_MEMBERS, sizeof(wg_graph));
    uint64_t* first = (uint64_t*)malloc((MAX_MEMBERS + 1) * sizeof(uint64_t));
    uint64_t* roots = (uint64_t*)malloc(MAX_MEMBERS * sizeof(uint64_t));
    wg_graph u;
    memset(&u, 0, sizeof(u));
    wg_ctx* ctx = NULL;
    void* values = NULL;
    int ret = 1, rc = WG_OK;
    uint32_t n = 0;
    if (!items || !ms || !first || !roots)
        goto done;
    n = split_list(list, items, MAX_MEMBERS);
    printf("Graph:       batch of %u graphs\n", n);
    for (uint32_t i = 0; i < n && rc == WG_OK; ++i)
        rc = load_one(items[i], app, &ms[i]);
    if (rc || (rc = wg_graph_union(ms, n, &u, first))) {
        printf("Graph:       ERROR: %s\n", wg_error_string(rc));
        goto done;
    }
    printf("Graph:       union contains %llu vertices and %llu edges\n", (unsigned long long)u.num_vertices,
           (unsigned long long)u.num_edges);
    if ((rc = wg_open(&ctx, cs->device)) || (rc = wg_set_option(ctx, WG_OPT_RELABEL, WG_RELABEL_OFF)) ||
        (rc = wg_ctx_load_graph(ctx, &u)) || (rc = wg_set_members(ctx, n, first))) {
        printf("Graph:       ERROR: upload failed: %s\n", wg_error_string(rc));
        goto done;
    }
    const uint64_t V = u.num_vertices;
    const size_t esz = (app == WG_APP_CC || (app == WG_APP_SSSP && u.weight_kind == WG_WEIGHT_U32)) ? 4 : 8;
    values = malloc(V * esz);
    if (!values)
        goto done;
    wg_loop_params p;
    memset(&p, 0, sizeof(p));
    p.schedule = cs->schedule;
    p.threshold_pct = (uint32_t)cs->frontier_threshold_pct;
    p.block_log2 = cs->block_log2;
    for (uint32_t i = 0; i < n; ++i)
        roots[i] = first[i] + (cs->root_vertex < ms[i].num_vertices ? cs->root_vertex : 0);
    printf("Application: starting execution\n");
    uint64_t iters = 0, fiters = 0;
    double t0 = now_ms();
    if (app == WG_APP_PR) {
        iters = cs->num_iterations;
        rc = wg_batch_pr(ctx, cs->num_iterations, cs->damping, cs->contrib_fp32, NULL, 1);
        if (rc == WG_OK)
            rc = wg_sync(ctx);
    } else {
        rc = app == WG_APP_BFS ? wg_init_bfs_roots(ctx, roots, n)
           : app == WG_APP_SSSP ? wg_init_sssp_roots(ctx, roots, n)
                                : wg_init_cc(ctx);
        t0 = now_ms();
        if (rc == WG_OK)
            rc = wg_converge(ctx, &p, &iters, &fiters);
    }
    const double running = now_ms() - t0;
    if (rc == WG_OK)
        rc = app == WG_APP_PR ? wg_read_batch_ranks(ctx, (double*)values)
           : app == WG_APP_BFS ? wg_read_bfs(ctx, (uint64_t*)values)
           : app == WG_APP_CC ? wg_read_cc(ctx, (uint32_t*)values)
                              : wg_read_sssp(ctx, values);
    if (rc) {
        printf("Application: ERROR: %s (hip %d)\n", wg_error_string(rc), wg_last_hip_error(ctx));
        goto done;
    }
    printf("Application: completed execution\n");
    if (cs->graph_application_output_filename) {
        FILE* f = fopen(cs->graph_application_output_filename, "w");
        if (!f) {
            printf("Application: ERROR: unable to write to \"%s\"\n", cs->graph_application_output_filename);
            goto done;
        }
        for (uint32_t g = 0; g < n && !ferror(f); ++g)
            for (uint64_t v = first[g]; v < first[g + 1]; ++v) {
                fprintf(f, "%u %llu", g, (unsigned long long)(v - first[g]));
                if (app == WG_APP_PR) {
                    fprintf(f, cs->full_precision ? " %.17g" : " %.4le", ((const double*)values)[v]);
                } else if (app == WG_APP_BFS) {
                    const uint64_t pk = ((const uint64_t*)values)[v];
                    fprintf(f, " %lld", pk == UINT64_MAX ? -1ll : (long long)((pk & 0xffffffffull) - first[g]));
                } else if (app == WG_APP_CC) {
                    fprintf(f, " %llu", (unsigned long long)(((const uint32_t*)values)[v] - first[g]));
                } else {
                    print_dist(f, u.weight_kind, values, v);
                }
                fprintf(f, "\n");
            }
        const int bad = ferror(f);
        fclose(f);
        if (bad) {
            printf("Application: ERROR: file I/O error during write\n");
            goto done;
        }
    }
    print_stats(running, u.num_edges * (iters ? iters : 1), iters, "Graphs", n);
    ret = (cs->validate && app != WG_APP_PR) ? wg_main_validate(ctx) : 0;
done:
    free(values);
    if (ctx)
        wg_close(ctx);
    wg_graph_free(&u);
    if (ms)
        for (uint32_t i = 0; i < n; ++i)
            wg_graph_free(&ms[i]);
    free(ms);
    free(items);
    free(first);
    free(roots);
    return ret;
}
