// wg_mlp_wave.h -- the schedule model's forward pass on one wavefront: lane j
// holds hidden unit j (WG_MLP_HIDDEN == wavefront size), lanes 0..A-1 the
// outputs.  Each unit sums its inputs in ascending order from its bias, as
// wg_mlp_forward does on the host, so both pick the same action.
#pragma once

#include "wg_internal.h"
#include "wg_mlp_math.h"

static_assert(WG_MLP_HIDDEN == WG_WAVE, "one hidden unit per lane");
static_assert(WG_MLP_ACTIONS <= WG_WAVE, "one output per lane");

// Called by all 64 lanes with the same x and base (base may be nullptr: 0).
// Returns the chosen action (the cheapest allowed for app, settled against the
// plain rule's action plain, -1 for none) on every lane; lane
// o < WG_MLP_ACTIONS writes its prediction (network output + base) to y[o] if
// y != nullptr.  keep (training) receives the lane's pre-activations,
// activations and network output without the base.
struct wg_mlp_lane {
    float a1, h1, a2, h2, y;
};

// Value of x in lane k (k the same in every lane): a scalar read of the lane's
// register, cheaper than a cross-lane permute.
__device__ __forceinline__ float wg_lane_value(float x, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}

__device__ __forceinline__ int wg_mlp_choose_wave(const float* __restrict__ P, int app, const float* x,
                                                  const float* base, int plain, float* y, wg_mlp_lane* keep = nullptr)
{
    const int lane = wg_lane();
    float z[WG_MLP_IN];
#pragma unroll
    for (int i = 0; i < WG_MLP_IN; ++i)
        z[i] = wg_mlp_input(P, x, i);
    float a1 = P[WG_MLP_OFF_B1 + lane];
#pragma unroll
    for (int i = 0; i < WG_MLP_IN; ++i)
        a1 = a1 + P[WG_MLP_OFF_W1 + i * WG_MLP_HIDDEN + lane] * z[i];
    const float h1 = wg_mlp_relu(a1);
    // a layer's activations reach every lane through a per-wave LDS row (reads
    // of one address by all lanes are broadcasts, issued 16 at a time)
    __shared__ float s_act[16][WG_MLP_HIDDEN]; // up to 16 wavefronts per workgroup
    float* act = s_act[(threadIdx.x / WG_WAVE) & 15];
    act[lane] = h1;
    wg_wave_sync();
    float a2 = P[WG_MLP_OFF_B2 + lane];
#pragma unroll 16
    for (int k = 0; k < WG_MLP_HIDDEN; ++k)
        a2 = a2 + P[WG_MLP_OFF_W2 + k * WG_MLP_HIDDEN + lane] * act[k];
    const float h2 = wg_mlp_relu(a2);
    wg_wave_sync();
    act[lane] = h2;
    wg_wave_sync();
    const int o = lane < WG_MLP_ACTIONS ? lane : WG_MLP_ACTIONS - 1;
    float a3 = P[WG_MLP_OFF_B3 + o];
#pragma unroll 16
    for (int j = 0; j < WG_MLP_HIDDEN; ++j)
        a3 = a3 + P[WG_MLP_OFF_W3 + j * WG_MLP_ACTIONS + o] * act[j];
    wg_wave_sync(); // the row is free again
    const float pred = base ? a3 + base[o] : a3;
    if (y && lane < WG_MLP_ACTIONS)
        y[lane] = pred;
    if (keep) {
        keep->a1 = a1;
        keep->h1 = h1;
        keep->a2 = a2;
        keep->h2 = h2;
        keep->y = a3;
    }
    int best = 0;
    float bv = wg_lane_value(pred, 0);
    for (int q = 1; q < WG_MLP_ACTIONS; ++q) {
        const float v = wg_lane_value(pred, q);
        if (wg_mlp_allowed(app, q) && v < bv) {
            best = q;
            bv = v;
        }
    }
    const float py = wg_lane_value(pred, plain >= 0 ? plain : 0);
    return wg_mlp_settle(P, app, best, bv, plain, py);
}
