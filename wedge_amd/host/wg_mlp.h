/*
 * wg_mlp.h -- learned schedule: FP32 MLP (features -> predicted time of each
 * candidate schedule), trained on the engine's own per-iteration traces.
 */
#ifndef WG_MLP_H
#define WG_MLP_H

#include "wg_sched.h"

/* Chooses the schedule with the smallest predicted time.  Returns WG_E_STATE
 * when no model is loaded (callers then use the plain heuristic). */
int wg_mlp_decide(const wg_sched_features* f, int* mode, int* block_log2);

#endif /* WG_MLP_H */
