/*
 * main.c -- wedge front end: parse options, load or generate the graph, run
 * the application on the GPU, write the optional per-vertex output and print
 * the statistics block.
 *
 * Order of operations and output formats follow source/main.c:33-185 and the
 * per-application string functions (source/execution_bfs.c:69-73,
 * source/execution_cc.c:47-51, source/execution_pr.c:31-35,
 * source/execution_sssp.c:48-52).  Unreached vertices print -1 as in the
 * reference.  BFS prints the parent; the root prints itself (the reference
 * gives the root a child as parent, an artefact of never marking it visited).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <time.h>

#include "cmdline.h"
#include "main_batch.h"
#include "wedge_host.h"

#include <unistd.h>

/* The schedule model shipped with the package: <binary dir>/../data/schedule_mlp.txt. */
static const char* default_model_path(char* buf, size_t n)
{
    const ssize_t len = readlink("/proc/self/exe", buf, n - 1);
    if (len <= 0)
        return NULL;
    buf[len] = 0;
    char* slash = strrchr(buf, '/');
    if (!slash)
        return NULL;
    *slash = 0;
    const char* rel = "/../data/schedule_mlp.txt";
    if (strlen(buf) + strlen(rel) + 1 > n)
        return NULL;
    strcat(buf, rel);
    return buf;
}

static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC_RAW, &t);
    return (double)t.tv_sec * 1e3 + (double)t.tv_nsec * 1e-6;
}

/* source/application.c:48-58 names and aliases, case-insensitive */
static int app_from_string(const char* s)
{
    static const char* bfs[] = { "bfs", "breadth_first_search", "breadthfirstsearch" };
    static const char* cc[] = { "cc", "connected_components", "connectedcomponents" };
    static const char* pr[] = { "pr", "pagerank" };
    static const char* sssp[] = { "sssp", "single_source_shortest_path", "singlesourceshortestpath" };
    static const char* ppr[] = { "ppr", "personalized_pagerank", "personalised_pagerank" };
    for (size_t i = 0; i < 3; ++i)
        if (!strcasecmp(s, bfs[i]))
            return WG_APP_BFS;
    for (size_t i = 0; i < 3; ++i)
        if (!strcasecmp(s, cc[i]))
            return WG_APP_CC;
    for (size_t i = 0; i < 2; ++i)
        if (!strcasecmp(s, pr[i]))
            return WG_APP_PR;
    for (size_t i = 0; i < 3; ++i)
        if (!strcasecmp(s, sssp[i]))
            return WG_APP_SSSP;
    for (size_t i = 0; i < 3; ++i)
        if (!strcasecmp(s, ppr[i]))
            return WG_APP_PPR;
    return WG_APP_NONE;
}

static const char* app_name(int app)
{
    switch (app) {
    case WG_APP_BFS: return "Breadth-First Search";
    case WG_APP_CC: return "Connected Components";
    case WG_APP_PR: return "PageRank";
    case WG_APP_SSSP: return "Single-Source Shortest Path";
    case WG_APP_PPR: return "Personalized PageRank";
    default: return "?";
    }
}

static int generate(const char* spec, int app, wg_graph* g)
{
    char buf[256];
    strncpy(buf, spec, sizeof(buf) - 1);
    buf[sizeof(buf) - 1] = '\0';
    char* save = NULL;
    char* kind = strtok_r(buf, ":", &save);
    if (!kind)
        return WG_E_INVAL;
    if (!strcasecmp(kind, "rmat")) {
        char* a = strtok_r(NULL, ":", &save);
        char* b = strtok_r(NULL, ":", &save);
        char* c = strtok_r(NULL, ":", &save);
        if (!a || !b || !c)
            return WG_E_INVAL;
        int sym = 0, weighted = (app == WG_APP_SSSP);
        for (char* t = strtok_r(NULL, ":", &save); t; t = strtok_r(NULL, ":", &save)) {
            if (!strcasecmp(t, "sym"))
                sym = 1;
            else if (!strcasecmp(t, "w"))
                weighted = 1;
            else
                return WG_E_INVAL;
        }
        return wg_graph_rmat(g, atoi(a), atoi(b), 0.57, 0.19, 0.19, strtoull(c, NULL, 10), sym,
                             weighted ? WG_WEIGHT_F64 : WG_WEIGHT_NONE);
    }
    if (!strcasecmp(kind, "grid")) {
        char* r = strtok_r(NULL, ":", &save);
        char* c = strtok_r(NULL, ":", &save);
        char* s = strtok_r(NULL, ":", &save);
        if (!r || !c || !s)
            return WG_E_INVAL;
        return wg_graph_grid(g, (uint32_t)atoi(r), (uint32_t)atoi(c), 0.6, 0.05, 1000, strtoull(s, NULL, 10));
    }
    return WG_E_INVAL;
}

/* One member of a batch of graphs: a generator spec (-g) or a file (-i). */
static int load_member(const char* item, int app, wg_graph* g)
{
    const cmdline_opts_t* cs = cmdline_get_current_settings();
    if (cs->generator[0])
        return generate(item, app, g);
    return wg_graph_read(g, item, app == WG_APP_SSSP ? cs->weight_kind : WG_WEIGHT_NONE);
}

int main(int argc, char* argv[])
{
    cmdline_parse_options_or_die(argc, argv);
    const cmdline_opts_t* cs = cmdline_get_current_settings();
    printf("Frontier:    granularity set to %d edges/bit\n", 1 << cs->block_log2);
    printf("Frontier:    threshold set to %llu%%\n", (unsigned long long)cs->frontier_threshold_pct);
    const int app = app_from_string(cs->application);
    if (app == WG_APP_NONE) {
        printf("Application: ERROR: unable to run \"%s\"\n", cs->application);
        return __LINE__;
    }
    printf("Application: %s on HIP device %d\n", app_name(app), cs->device);
    if (strchr(cs->generator[0] ? cs->generator : cs->graph_input_filename, ','))
        return wg_main_members(cs, app, load_member);
    if (app == WG_APP_PPR && cs->nroots == 0) {
        printf("Application: ERROR: ppr needs seeds (-R)\n");
        return __LINE__;
    }

    wg_graph g;
    double t0 = now_ms();
    int rc;
    if (cs->generator[0]) {
        printf("Graph:       generating \"%s\"\n", cs->generator);
        rc = generate(cs->generator, app, &g);
    } else {
        printf("Graph:       reading \"%s\"\n", cs->graph_input_filename);
        rc = wg_graph_read(&g, cs->graph_input_filename, app == WG_APP_SSSP ? cs->weight_kind : WG_WEIGHT_NONE);
    }
    if (rc) {
        printf("Graph:       ERROR: %s\n", wg_error_string(rc));
        return __LINE__;
    }
    printf("Graph:       contains %llu vertices and %llu edges\n", (unsigned long long)g.num_vertices,
           (unsigned long long)g.num_edges);
    printf("Graph:       loading completed in %.2lf msec\n", now_ms() - t0);

    wg_ctx* ctx = NULL;
    rc = wg_open(&ctx, cs->device);
    if (rc) {
        printf("Device:      ERROR: %s\n", wg_error_string(rc));
        wg_graph_free(&g);
        return __LINE__;
    }
    t0 = now_ms();
    rc = wg_ctx_load_graph(ctx, &g);
    if (rc) {
        printf("Graph:       ERROR: upload failed: %s (hip %d)\n", wg_error_string(rc), wg_last_hip_error(ctx));
        wg_close(ctx);
        wg_graph_free(&g);
        return __LINE__;
    }
    printf("Graph:       initialization completed in %.2lf msec\n", now_ms() - t0);

    wg_run_opts o;
    wg_run_opts_default(&o);
    o.schedule = cs->schedule;
    o.threshold_pct = (uint32_t)cs->frontier_threshold_pct;
    o.block_log2 = cs->block_log2;
    o.iterations = cs->num_iterations;
    o.damping = cs->damping;
    o.contrib_fp32 = cs->contrib_fp32;
    o.root = cs->root_vertex;
    if (o.root >= g.num_vertices) {
        /* source/main.c:82-86 */
        printf("Application: WARNING: root vertex %llu is out of range, using %d instead\n",
               (unsigned long long)o.root, CMDLINE_DEFAULT_ROOT_VERTEX);
        o.root = CMDLINE_DEFAULT_ROOT_VERTEX;
    }
    if (cs->model[0] || cs->schedule == WG_SCHED_LEARNED) {
        char path[4096];
        const char* model = cs->model[0] ? cs->model : default_model_path(path, sizeof(path));
        rc = model ? wg_mlp_load(model) : WG_E_IO;
        if (rc)
            printf("Schedule:    WARNING: cannot load model \"%s\" (%s); using the plain heuristic\n",
                   model ? model : "", wg_error_string(rc));
        else
            printf("Schedule:    model \"%s\"\n", model);
    }

    if (app == WG_APP_PPR || cs->nroots > 1) {
        const int ret = wg_main_sources(cs, app, ctx, &g);
        wg_close(ctx);
        wg_graph_free(&g);
        return ret;
    }
    if (cs->nroots == 1)
        o.root = cs->roots[0];
    const uint64_t V = g.num_vertices;
    void* values = NULL;
    size_t esz = 8;
    if (app == WG_APP_CC || (app == WG_APP_SSSP && g.weight_kind == WG_WEIGHT_U32))
        esz = 4;
    values = malloc(V * esz);
    if (!values) {
        wg_close(ctx);
        wg_graph_free(&g);
        return __LINE__;
    }
    printf("Application: starting execution\n");
    wg_run_stats st;
    memset(&st, 0, sizeof(st));
    switch (app) {
    case WG_APP_PR: rc = wg_run_pr(ctx, &o, (double*)values, &st); break;
    case WG_APP_BFS: rc = wg_run_bfs(ctx, &o, (uint64_t*)values, &st); break;
    case WG_APP_CC: rc = wg_run_cc(ctx, &o, (uint32_t*)values, &st); break;
    default: rc = wg_run_sssp(ctx, &o, values, &st); break;
    }
    if (rc) {
        printf("Application: ERROR: %s (hip %d)\n", wg_error_string(rc), wg_last_hip_error(ctx));
        free(values);
        wg_close(ctx);
        wg_graph_free(&g);
        return __LINE__;
    }
    printf("Application: completed execution\n");

    if (cs->graph_application_output_filename) {
        FILE* f = fopen(cs->graph_application_output_filename, "w");
        if (!f) {
            printf("Application: ERROR: unable to write to \"%s\"\n", cs->graph_application_output_filename);
        } else {
            printf("Application: writing output to \"%s\"\n", cs->graph_application_output_filename);
            const double tw = now_ms();
            for (uint64_t v = 0; v < V && !ferror(f); ++v) {
                switch (app) {
                case WG_APP_PR:
                    fprintf(f, cs->full_precision ? "%llu %.17g\n" : "%llu %.4le\n", (unsigned long long)v,
                            ((double*)values)[v]);
                    break;
                case WG_APP_BFS: {
                    const uint64_t pk = ((uint64_t*)values)[v];
                    fprintf(f, "%llu %lld\n", (unsigned long long)v, pk == UINT64_MAX ? -1ll : (long long)(pk & 0xffffffffull));
                    break;
                }
                case WG_APP_CC:
                    fprintf(f, "%llu %u\n", (unsigned long long)v, ((uint32_t*)values)[v]);
                    break;
                default:
                    if (g.weight_kind == WG_WEIGHT_F64) {
                        const double d = ((double*)values)[v];
                        if (d == (double)INFINITY)
                            fprintf(f, "%llu -1\n", (unsigned long long)v);
                        else
                            fprintf(f, "%llu %.17g\n", (unsigned long long)v, d);
                    } else {
                        const uint32_t d = ((uint32_t*)values)[v];
                        fprintf(f, "%llu %lld\n", (unsigned long long)v, d == 0xffffffffu ? -1ll : (long long)d);
                    }
                }
            }
            if (ferror(f))
                printf("Application: ERROR: file I/O error during write\n");
            else
                printf("Application: writing completed in %.2lf msec\n", now_ms() - tw);
            fclose(f);
        }
    }

    /* source/main.c:153-182 */
    const double running = st.running_ms;
    const uint64_t iters = st.iterations;
    printf("\n--------------- EXECUTION STATISTICS ---------------\n");
    printf("%-28s = %.2lf msec\n", "Running Time", running);
    printf("%-28s = %.4lf Bedges/sec\n", "Effective Throughput",
           running > 0 ? (double)g.num_edges * (double)iters / running / 1000000.0 : 0.0);
    if (app != WG_APP_PR && app != WG_APP_CC)
        printf("%-28s = %.4lf GTEPS\n", "Reached-Edge Throughput",
               running > 0 ? (double)st.edges_reached / running / 1000000.0 : 0.0);
    printf("----------------------------------------------------\n");
    printf("%-28s = %.2lf msec\n", "Pull Time", st.pull_ms);
    printf("%-28s = %.2lf msec\n", "Frontier Pull Time", st.frontier_pull_ms);
    printf("%-28s = %.2lf msec\n", "No Frontier Pull Time", st.dense_pull_ms);
    printf("%-28s = %.2lf msec\n", "Wedge Time", st.wedge_ms);
    printf("----------------------------------------------------\n");
    printf("%-28s = %llu\n", "Total Iterations", (unsigned long long)iters);
    printf("%-28s = %llu\n", "Iterations with Wedge", (unsigned long long)st.frontier_iterations);
    printf("%-28s = %llu\n", "Iterations without Wedge", (unsigned long long)(iters - st.frontier_iterations));
    printf("----------------------------------------------------\n");
    if (app == WG_APP_PR) {
        printf("%-28s = %.8lf\n", "PageRank Sum", st.special);
        printf("----------------------------------------------------\n");
    }
    printf("\n");
    int ret = 0;
    if (cs->validate && app != WG_APP_PR)
        ret = wg_main_validate(ctx);
    free(values);
    wg_close(ctx);
    wg_graph_free(&g);
    return ret;
}
