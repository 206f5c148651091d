/*
 * wg_mlp.c -- learned schedule (model management and decision).
 */
#include "wg_mlp.h"

int wg_mlp_decide(const wg_sched_features* f, int* mode, int* block_log2)
{
    (void)f;
    (void)mode;
    (void)block_log2;
    return WG_E_STATE;
}

int wg_mlp_load(const char* path)
{
    (void)path;
    return WG_E_IO;
}

int wg_mlp_save(const char* path)
{
    (void)path;
    return WG_E_STATE;
}

void wg_mlp_unload(void)
{
}

int wg_mlp_loaded(void)
{
    return 0;
}
