/*
 * wg_mlp.h -- learned schedule: FP32 MLP (features -> predicted time of each
 * candidate schedule), trained on the engine's own per-iteration traces.
 */
#ifndef WG_MLP_H
#define WG_MLP_H

#include "wg_sched.h"

/* Chooses the schedule with the smallest predicted time, settled against the
 * plain rule's choice (plain_dense, and *block_log2 on input: the plain
 * granularity).  Returns WG_E_STATE when no model is loaded or the graph is
 * outside the model's range for the app (callers then use the plain rule). */
int wg_mlp_decide(const wg_sched_features* f, int plain_dense, int* mode, int* block_log2);

/* Parameters of the loaded model, NULL if none. */
const float* wg_mlp_params(void);

#endif /* WG_MLP_H */
