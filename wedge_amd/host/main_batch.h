/*
 * main_batch.h -- batched runs of the wedge front end (several sources on one
 * graph, or a batch of graphs as one disjoint union).
 */
#ifndef WG_MAIN_BATCH_H
#define WG_MAIN_BATCH_H

#include "cmdline.h"
#include "wedge_host.h"

/* -c: validates the context's BFS / CC / SSSP result on the device and prints
 * the verdict; returns 0 if valid, 7 if not, 1 on error. */
int wg_main_validate(wg_ctx* ctx);

/* Loads one member of a batch: a file name or a generator spec. */
typedef int (*member_loader)(const char* item, int app, wg_graph* out);

/* -R with bfs / sssp, or ppr: ctx already holds g.  Returns the exit code. */
int wg_main_sources(const cmdline_opts_t* cs, int app, wg_ctx* ctx, const wg_graph* g);
/* Comma-separated -i / -g: loads, runs and reports the batch.  Returns the exit code. */
int wg_main_members(const cmdline_opts_t* cs, int app, member_loader load_one);

#endif /* WG_MAIN_BATCH_H */
