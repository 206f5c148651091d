/*
 * wg_mlp_math.h -- the schedule model's features and forward pass, written
 * once for the host library (C) and the device decision code (HIP), so that
 * both compute the same FP32 operations in the same order and choose the same
 * action.  Both sides are compiled without contraction (-ffp-contract=off).
 *
 * The model (no counterpart in the reference; it replaces the threshold rule
 * of source/execution_bfs.c:141): x = features clamped to [lo, hi],
 * z = (x - mean) * inv_std,
 * h1 = relu(W1^T z + b1), h2 = relu(W2^T h1 + b2), y = W3^T h2 + b3 + base,
 * where y[a] is the predicted log2 time of action a and base[a] is the log2 of
 * an analytic work estimate for it (wg_mlp_base): the network learns the time
 * per unit of work, which varies far less with graph size than the time
 * itself, so a graph larger than any seen in training is priced by the
 * estimate.  Sums run over the input index in ascending order, starting from
 * the bias.
 */
#ifndef WG_MLP_MATH_H
#define WG_MLP_MATH_H

#include <stdint.h>

#include "wedge_gpu.h"

#ifdef __HIPCC__
#define WG_HD __host__ __device__
#else
#define WG_HD
#endif

/* offsets into the parameter vector (include/wedge_gpu.h) */
#define WG_MLP_OFF_LO 0
#define WG_MLP_OFF_HI (WG_MLP_OFF_LO + WG_MLP_IN)
#define WG_MLP_OFF_MEAN (WG_MLP_OFF_HI + WG_MLP_IN)
#define WG_MLP_OFF_ISTD (WG_MLP_OFF_MEAN + WG_MLP_IN)
#define WG_MLP_OFF_W1 (WG_MLP_OFF_ISTD + WG_MLP_IN)
#define WG_MLP_OFF_B1 (WG_MLP_OFF_W1 + WG_MLP_IN * WG_MLP_HIDDEN)
#define WG_MLP_OFF_W2 (WG_MLP_OFF_B1 + WG_MLP_HIDDEN)
#define WG_MLP_OFF_B2 (WG_MLP_OFF_W2 + WG_MLP_HIDDEN * WG_MLP_HIDDEN)
#define WG_MLP_OFF_W3 (WG_MLP_OFF_B2 + WG_MLP_HIDDEN)
#define WG_MLP_OFF_B3 (WG_MLP_OFF_W3 + WG_MLP_HIDDEN * WG_MLP_ACTIONS)
#define WG_MLP_OFF_MARGIN (WG_MLP_OFF_B3 + WG_MLP_ACTIONS)
#define WG_MLP_OFF_MAXLV (WG_MLP_OFF_MARGIN + 3)
#define WG_MLP_OFF_END (WG_MLP_OFF_MAXLV + 1)

/* block_log2 of the frontier actions 1..5; action 0 is the dense pull */
#define WG_MLP_ACTION_BLOG2(a) ((a) == 1 ? 3 : (a) == 2 ? 4 : (a) == 3 ? 5 : (a) == 4 ? 6 : 8)

/* log2(1 + x) within 4e-3: exponent plus a cubic in the mantissa fraction f,
 * log2(1 + f) ~ f (c1 + f (c2 + f c3)), exact at f = 0, 1/2 and 1 and in slope
 * at 0. */
static inline WG_HD float wg_mlp_log2p1(uint64_t x)
{
    const uint64_t y = x + 1ull;
    if (y == 0ull)
        return 64.0f;
    const int e = 63 - __builtin_clzll(y);
    const uint64_t m = (e >= 23) ? (y >> (e - 23)) : (y << (23 - e)); /* 24 bits, leading one */
    const float f = (float)(m & 0x7fffffull) * (1.0f / 8388608.0f);
    const float p = f * (1.4426950f + f * (-0.6483851f + f * 0.2056901f));
    return (float)e + p;
}

/* Scale-free description of a loop state, L(x) = log2(1 + x): shares and
 * ratios only, so that a graph larger than the training graphs looks like a
 * training graph in the same phase of its run; its size enters the prediction
 * through the work estimate wg_mlp_base. */
static inline WG_HD void wg_mlp_features(int app, uint64_t iteration, uint64_t V, uint64_t E, uint64_t fv, uint64_t fe,
                                         uint64_t unvisited, float* x)
{
    const float lfv = wg_mlp_log2p1(fv), lfe = wg_mlp_log2p1(fe);
    const float lv = wg_mlp_log2p1(V), le = wg_mlp_log2p1(E);
    x[0] = lfv - lv;                                                      /* frontier share of vertices */
    x[1] = lfe - le;                                                      /* frontier share of edges */
    x[2] = app == WG_APP_CC ? 0.0f : wg_mlp_log2p1(unvisited) - lv;     /* share not reached yet */
    x[3] = le - lv;                                                       /* mean degree */
    x[4] = lfe - lfv;                                                     /* mean frontier out-degree */
    x[5] = wg_mlp_log2p1(iteration) - 0.5f * lv;                          /* iterations against sqrt(V) */
    x[6] = app == WG_APP_BFS ? 1.0f : 0.0f;
    x[7] = app == WG_APP_CC ? 1.0f : 0.0f;
    x[8] = app == WG_APP_SSSP ? 1.0f : 0.0f;
    x[9] = lfe - lv;                                                      /* frontier out-edges per vertex */
    x[10] = lfv - le;                                                     /* frontier vertices per edge */
    x[11] = app == WG_APP_CC ? 0.0f : wg_mlp_log2p1(unvisited) - le;    /* not reached per edge */
}

static inline WG_HD float wg_mlp_relu(float a)
{
    return a > 0.0f ? a : 0.0f;
}

/* Normalised input i: clamped to the training range, then standardised. */
static inline WG_HD float wg_mlp_input(const float* P, const float* x, int i)
{
    const float lo = P[WG_MLP_OFF_LO + i], hi = P[WG_MLP_OFF_HI + i];
    const float v = x[i] < lo ? lo : (x[i] > hi ? hi : x[i]);
    return (v - P[WG_MLP_OFF_MEAN + i]) * P[WG_MLP_OFF_ISTD + i];
}

/* Actions the schedule may take for app (see include/wedge_gpu.h). */
static inline WG_HD int wg_mlp_allowed(int app, int a)
{
    return app == WG_APP_BFS || a == 0 || a == 2;
}

/* log2 work estimate of each action from the loop state, in edges read: a
 * dense pull reads every edge (E); a frontier pull at granularity 2^b reads at
 * most E and at most fe * 2^b edges (each activated out-edge marks one
 * block), plus scans of the vertex bitmasks (V/32); both pay a fixed launch
 * cost counted as 2^18 edges. */
static inline WG_HD void wg_mlp_base(uint64_t V, uint64_t E, uint64_t fe, float* base)
{
    base[0] = wg_mlp_log2p1(E + (1ull << 18));
    for (int a = 1; a < WG_MLP_ACTIONS; ++a) {
        const uint64_t blk = fe << WG_MLP_ACTION_BLOG2(a);
        base[a] = wg_mlp_log2p1((blk < E ? blk : E) + (V >> 5) + (1ull << 18));
    }
}

/* Action of the plain rule's choice (dense, or frontier at plain_blog2), -1
 * if that granularity is not an action. */
static inline WG_HD int wg_mlp_plain_action(int dense, int plain_blog2)
{
    if (dense)
        return 0;
    for (int a = 1; a < WG_MLP_ACTIONS; ++a)
        if (WG_MLP_ACTION_BLOG2(a) == plain_blog2)
            return a;
    return -1;
}

/* Final choice: the cheapest allowed action, unless the plain rule's action
 * is predicted to cost less than 2^margin[app] times as much (the plain rule
 * stays the default wherever the model does not see a clear gain). */
static inline WG_HD int wg_mlp_settle(const float* P, int app, int best, float best_y, int plain, float plain_y)
{
    if (plain < 0 || plain == best)
        return best;
    const int k = app == WG_APP_BFS ? 0 : app == WG_APP_CC ? 1 : 2;
    return (plain_y - best_y > P[WG_MLP_OFF_MARGIN + k]) ? best : plain;
}

/* A level whose frontier pull reads at most 1/8 of the edges even at the
 * coarsest block granularity (activity * 2^8 * 8 <= E) is one where the dense
 * pull, which reads every edge, cannot win: there the schedule takes the plain
 * rule's frontier pull without evaluating the model.  Such levels are most of
 * a traversal's levels on road-like graphs (thousands of them), and there the
 * model's evaluation costs more than it can save. */
#define WG_MLP_GATE_LOG2 11
static inline WG_HD int wg_mlp_gated(uint64_t activity, uint64_t num_edges)
{
    return activity <= (num_edges >> WG_MLP_GATE_LOG2);
}

/* Dense CC / SSSP pulls on the binned layout stream every edge's 2-byte code
 * in about the time a frontier-filtered pull, its block marking and its hub
 * chunks take for E/64 active edges (rmat-s24 FP64 SSSP: 1.42 ms per dense
 * iteration, 2.4-5.8 ms per frontier iteration at 10-20 % of E).  Above that
 * activity the learned schedule takes the dense pull without evaluating the
 * model (rmat-s24 SSSP 30.7 -> 21.5 ms; E/10, E/20, E/40, E/80, E/160:
 * 25.1, 23.1, 21.9, 21.5, 21.9 ms). */
#define WG_SCHED_BINNED_DENSE_LOG2 6
static inline WG_HD int wg_sched_binned_dense(int binned, uint64_t activity, uint64_t num_edges)
{
    return binned && activity > (num_edges >> WG_SCHED_BINNED_DENSE_LOG2);
}

/* CC and SSSP keep the plain rule on graphs with more vertices than the model
 * was trained on (L(V) above the model's bound): their schedule also decides
 * how many pulls follow, which a per-pull model cannot extrapolate.  A BFS
 * level is the same in every mode, so BFS uses the model at any size. */
static inline WG_HD int wg_mlp_out_of_range(const float* P, int app, uint64_t V)
{
    return app != WG_APP_BFS && wg_mlp_log2p1(V) > P[WG_MLP_OFF_MAXLV];
}

/* Host-order forward pass (the device runs the same sums, one unit per lane).
 * y[a] = network output + base[a] (base may be NULL: 0).  Returns the allowed
 * action for app with the smallest prediction (lowest index on ties), settled
 * against the plain rule's action plain (-1: none). */
static inline WG_HD int wg_mlp_forward(const float* P, int app, const float* x, const float* base, int plain, float* y)
{
    float yy[WG_MLP_ACTIONS];
    float z[WG_MLP_IN], h1[WG_MLP_HIDDEN], h2[WG_MLP_HIDDEN];
    for (int i = 0; i < WG_MLP_IN; ++i)
        z[i] = wg_mlp_input(P, x, i);
    for (int j = 0; j < WG_MLP_HIDDEN; ++j) {
        float a = P[WG_MLP_OFF_B1 + j];
        for (int i = 0; i < WG_MLP_IN; ++i)
            a = a + P[WG_MLP_OFF_W1 + i * WG_MLP_HIDDEN + j] * z[i];
        h1[j] = wg_mlp_relu(a);
    }
    for (int j = 0; j < WG_MLP_HIDDEN; ++j) {
        float a = P[WG_MLP_OFF_B2 + j];
        for (int k = 0; k < WG_MLP_HIDDEN; ++k)
            a = a + P[WG_MLP_OFF_W2 + k * WG_MLP_HIDDEN + j] * h1[k];
        h2[j] = wg_mlp_relu(a);
    }
    int best = 0;
    float bv = 0.0f;
    for (int o = 0; o < WG_MLP_ACTIONS; ++o) {
        float a = P[WG_MLP_OFF_B3 + o];
        for (int j = 0; j < WG_MLP_HIDDEN; ++j)
            a = a + P[WG_MLP_OFF_W3 + j * WG_MLP_ACTIONS + o] * h2[j];
        if (base)
            a = a + base[o];
        yy[o] = a;
        if (y)
            y[o] = a;
        if (o == 0 || (wg_mlp_allowed(app, o) && a < bv)) {
            best = o;
            bv = a;
        }
    }
    return wg_mlp_settle(P, app, best, bv, plain, plain >= 0 ? yy[plain] : 0.0f);
}

#endif /* WG_MLP_MATH_H */
