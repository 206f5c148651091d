/*
 * wg_mlp.c -- learned schedule on the host: the model in use, its file format,
 * and the decision for the host-driven loop.  The forward pass is the one the
 * device loop runs (wg_mlp_math.h), so both loops pick the same schedules.
 *
 * File format (text): a line "wedge-schedule-mlp 1", a line
 * "in <IN> hidden <H> actions <A>", then WG_MLP_PARAMS numbers, one per line,
 * each printed with 9 significant digits (an exact FP32 round trip).
 */
#include "wg_mlp.h"

#include <stdio.h>
#include <string.h>

#include "wg_mlp_math.h"

static float g_params[WG_MLP_PARAMS];
static int g_loaded = 0;

int wg_mlp_decide(const wg_sched_features* f, int plain_dense, int* mode, int* block_log2)
{
    if (!g_loaded || wg_mlp_out_of_range(g_params, f->app, f->num_vertices))
        return WG_E_STATE; /* the caller applies the plain rule */
    float x[WG_MLP_IN];
    float base[WG_MLP_ACTIONS];
    wg_mlp_features(f->app, f->iteration, f->num_vertices, f->num_edges, f->frontier_vertices, f->frontier_edges,
                    f->unvisited, x);
    wg_mlp_base(f->num_vertices, f->num_edges, f->frontier_edges, base);
    const int a = wg_mlp_forward(g_params, f->app, x, base, wg_mlp_plain_action(plain_dense, *block_log2), NULL);
    *mode = a == 0 ? WG_MODE_DENSE : WG_MODE_FRONTIER;
    if (a)
        *block_log2 = WG_MLP_ACTION_BLOG2(a);
    return WG_OK;
}

int wg_mlp_feature_vector(int app, uint64_t iteration, uint64_t num_vertices, uint64_t num_edges,
                          uint64_t frontier_vertices, uint64_t frontier_edges, uint64_t unvisited, float* out)
{
    if (!out)
        return WG_E_INVAL;
    wg_mlp_features(app, iteration, num_vertices, num_edges, frontier_vertices, frontier_edges, unvisited, out);
    return WG_OK;
}

int wg_mlp_baseline(uint64_t num_vertices, uint64_t num_edges, uint64_t frontier_edges, float* out)
{
    if (!out)
        return WG_E_INVAL;
    wg_mlp_base(num_vertices, num_edges, frontier_edges, out);
    return WG_OK;
}

int wg_mlp_predict(int app, const float* features, const float* base, int plain_action, float* out)
{
    if (!features)
        return WG_E_INVAL;
    if (!g_loaded)
        return WG_E_STATE;
    if (plain_action < -1 || plain_action >= WG_MLP_ACTIONS)
        return WG_E_RANGE;
    return wg_mlp_forward(g_params, app, features, base, plain_action, out);
}

int wg_mlp_set(const float* params, uint64_t count)
{
    if (!params || count != (uint64_t)WG_MLP_PARAMS)
        return WG_E_INVAL;
    memcpy(g_params, params, sizeof(g_params));
    g_loaded = 1;
    return WG_OK;
}

int wg_mlp_get(float* params, uint64_t count)
{
    if (!params || count != (uint64_t)WG_MLP_PARAMS)
        return WG_E_INVAL;
    if (!g_loaded)
        return WG_E_STATE;
    memcpy(params, g_params, sizeof(g_params));
    return WG_OK;
}

const float* wg_mlp_params(void)
{
    return g_loaded ? g_params : NULL;
}

int wg_mlp_load(const char* path)
{
    if (!path)
        return WG_E_INVAL;
    FILE* f = fopen(path, "r");
    if (!f)
        return WG_E_IO;
    static float p[WG_MLP_PARAMS];
    int version = 0, in = 0, hid = 0, act = 0;
    int rc = WG_OK;
    if (fscanf(f, " wedge-schedule-mlp %d", &version) != 1 || version != 1 ||
        fscanf(f, " in %d hidden %d actions %d", &in, &hid, &act) != 3 || in != WG_MLP_IN || hid != WG_MLP_HIDDEN ||
        act != WG_MLP_ACTIONS) {
        rc = WG_E_FORMAT;
    } else {
        for (int i = 0; i < WG_MLP_PARAMS; ++i) {
            if (fscanf(f, " %f", &p[i]) != 1) {
                rc = WG_E_FORMAT;
                break;
            }
        }
        char extra;
        if (rc == WG_OK && fscanf(f, " %c", &extra) == 1)
            rc = WG_E_FORMAT; /* trailing content */
    }
    fclose(f);
    if (rc)
        return rc;
    memcpy(g_params, p, sizeof(g_params));
    g_loaded = 1;
    return WG_OK;
}

int wg_mlp_save(const char* path)
{
    if (!path)
        return WG_E_INVAL;
    if (!g_loaded)
        return WG_E_STATE;
    FILE* f = fopen(path, "w");
    if (!f)
        return WG_E_IO;
    int ok = fprintf(f, "wedge-schedule-mlp 1\nin %d hidden %d actions %d\n", WG_MLP_IN, WG_MLP_HIDDEN, WG_MLP_ACTIONS) > 0;
    for (int i = 0; ok && i < WG_MLP_PARAMS; ++i)
        ok = fprintf(f, "%.9g\n", (double)g_params[i]) > 0;
    if (fclose(f) != 0)
        ok = 0;
    return ok ? WG_OK : WG_E_IO;
}

void wg_mlp_unload(void)
{
    g_loaded = 0;
}

int wg_mlp_loaded(void)
{
    return g_loaded;
}
