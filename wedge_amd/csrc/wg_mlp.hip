// wg_mlp.hip -- the learned schedule's model on the device: batch inference
// with the same wavefront forward pass the decision kernel uses, and training
// (hand-written minibatch Adam).  No counterpart in the reference.
//
// Training step over a minibatch of B samples (sample s = idx[b]):
//   k_mlp_fb    one wavefront per sample: forward pass, then the backward pass
//               of the squared error of the taken action's output; keeps the
//               per-sample inputs z, hidden activations and their gradients
//   k_mlp_grad  one thread per parameter: sums its gradient over the B samples
//               in index order (no atomics, so a run is reproducible)
//   k_mlp_adam  one thread per parameter: Adam update
#include "wg_mlp_wave.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

extern "C" int wg_mlp_action(int action, int* mode, int* block_log2)
{
    if (action < 0 || action >= WG_MLP_ACTIONS)
        return WG_E_INVAL;
    if (mode)
        *mode = action == 0 ? WG_MODE_DENSE : WG_MODE_FRONTIER;
    if (block_log2)
        *block_log2 = action == 0 ? WG_DEFAULT_BLOCK_LOG2 : WG_MLP_ACTION_BLOG2(action);
    return WG_OK;
}

extern "C" int wg_set_schedule_model(wg_ctx* ctx, const float* params, uint64_t count)
{
    if (!ctx)
        return WG_E_INVAL;
    if (!params) {
        ctx->mlp_set = 0;
        return WG_OK;
    }
    if (count != (uint64_t)WG_MLP_PARAMS)
        return WG_E_INVAL;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->mlp)
        WG_HIP(ctx, hipMalloc((void**)&ctx->mlp, WG_MLP_PARAMS * sizeof(float)));
    WG_HIP(ctx, hipMemcpyAsync(ctx->mlp, params, WG_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->mlp_set = 1;
    return WG_OK;
}

__global__ __launch_bounds__(WG_WAVE) void k_mlp_infer(const float* __restrict__ P, int app, const float* __restrict__ X,
                                                       const float* __restrict__ Bs, const int32_t* __restrict__ plain,
                                                       uint64_t n, float* __restrict__ Y, int32_t* __restrict__ act)
{
    for (uint64_t s = blockIdx.x; s < n; s += gridDim.x) {
        float x[WG_MLP_IN], base[WG_MLP_ACTIONS];
#pragma unroll
        for (int i = 0; i < WG_MLP_IN; ++i)
            x[i] = X[s * WG_MLP_IN + i];
#pragma unroll
        for (int a = 0; a < WG_MLP_ACTIONS; ++a)
            base[a] = Bs ? Bs[s * WG_MLP_ACTIONS + a] : 0.0f;
        const int a = wg_mlp_choose_wave(P, app, x, base, plain ? plain[s] : -1, Y ? Y + s * WG_MLP_ACTIONS : nullptr);
        if (act && threadIdx.x == 0)
            act[s] = a;
    }
}

extern "C" int wg_mlp_infer(wg_ctx* ctx, int app, const float* features, const float* base, const int32_t* plain,
                            uint64_t n, float* out, int32_t* actions)
{
    if (!ctx || (n && !features))
        return WG_E_INVAL;
    if (!ctx->mlp_set)
        return WG_E_STATE;
    if (n == 0)
        return WG_OK;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    float *dX = nullptr, *dY = nullptr, *dB = nullptr;
    int32_t *dA = nullptr, *dPl = nullptr;
    int rc = WG_OK;
    if (plain)
        for (uint64_t s = 0; s < n; ++s)
            if (plain[s] < -1 || plain[s] >= WG_MLP_ACTIONS)
                return WG_E_RANGE;
    if ((base && hipMalloc((void**)&dB, n * WG_MLP_ACTIONS * sizeof(float)) != hipSuccess) ||
        (plain && hipMalloc((void**)&dPl, n * sizeof(int32_t)) != hipSuccess) ||
        hipMalloc((void**)&dX, n * WG_MLP_IN * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&dY, n * WG_MLP_ACTIONS * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&dA, n * sizeof(int32_t)) != hipSuccess) {
        rc = WG_E_NOMEM;
        goto done;
    }
    if (hipMemcpyAsync(dX, features, n * WG_MLP_IN * sizeof(float), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        (base && hipMemcpyAsync(dB, base, n * WG_MLP_ACTIONS * sizeof(float), hipMemcpyHostToDevice, ctx->stream) !=
                     hipSuccess) ||
        (plain && hipMemcpyAsync(dPl, plain, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)) {
        rc = WG_E_HIP;
        goto done;
    }
    hipLaunchKernelGGL(k_mlp_infer, dim3((uint32_t)(n < 65536 ? n : 65536)), dim3(WG_WAVE), 0, ctx->stream, ctx->mlp, app,
                       dX, dB, dPl, n, dY, dA);
    if ((rc = wg_check_launch(ctx)))
        goto done;
    if ((out && hipMemcpyAsync(out, dY, n * WG_MLP_ACTIONS * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) ||
        (actions && hipMemcpyAsync(actions, dA, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = WG_E_HIP;
done:
    (void)hipFree(dX);
    (void)hipFree(dY);
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dPl);
    return rc;
}

// ---- training -----------------------------------------------------------------------

// Per-sample training buffers, [b][unit] so that the gradient threads of
// neighbouring units read neighbouring words.
struct mlp_batch {
    float* z;    // [B][IN] normalised inputs
    float* h1;   // [B][H]
    float* h2;   // [B][H]
    float* dh1;  // [B][H] d loss / d pre-activation of layer 1
    float* dh2;  // [B][H] d loss / d pre-activation of layer 2
    float* dy;   // [B]    d loss / d y[action]
    int32_t* o;  // [B]    action
};

__global__ __launch_bounds__(WG_WAVE) void k_mlp_fb(const float* __restrict__ P, const float* __restrict__ X,
                                                    const int32_t* __restrict__ A, const float* __restrict__ T,
                                                    const uint32_t* __restrict__ idx, uint32_t B, mlp_batch mb)
{
    const uint32_t b = blockIdx.x;
    if (b >= B)
        return;
    const int lane = wg_lane();
    const uint32_t s = idx[b];
    float x[WG_MLP_IN];
#pragma unroll
    for (int i = 0; i < WG_MLP_IN; ++i)
        x[i] = X[(uint64_t)s * WG_MLP_IN + i];
    wg_mlp_lane k;
    (void)wg_mlp_choose_wave(P, WG_APP_BFS, x, nullptr, -1, nullptr, &k);
    const int o = A[s];
    const float y = __shfl(k.y, o); // the taken action's prediction
    const float dy = 2.0f * (y - T[s]) / (float)B;
    // layer 2: d/d a2[lane] = dy * W3[lane][o] * [a2 > 0]
    const float d2 = k.a2 > 0.0f ? dy * P[WG_MLP_OFF_W3 + lane * WG_MLP_ACTIONS + o] : 0.0f;
    // layer 1: d/d a1[lane] = sum_j W2[lane][j] d2[j] * [a1 > 0]
    float s1 = 0.0f;
    for (int j = 0; j < WG_MLP_HIDDEN; ++j)
        s1 = s1 + P[WG_MLP_OFF_W2 + lane * WG_MLP_HIDDEN + j] * __shfl(d2, j);
    const float d1 = k.a1 > 0.0f ? s1 : 0.0f;
    const uint64_t r = (uint64_t)b * WG_MLP_HIDDEN + (uint64_t)lane;
    mb.h1[r] = k.h1;
    mb.h2[r] = k.h2;
    mb.dh1[r] = d1;
    mb.dh2[r] = d2;
    if (lane < WG_MLP_IN)
        mb.z[(uint64_t)b * WG_MLP_IN + lane] = wg_mlp_input(P, X + (uint64_t)s * WG_MLP_IN, lane);
    if (lane == 0) {
        mb.dy[b] = dy;
        mb.o[b] = o;
    }
}

// Gradient of parameter p (normalisation entries are not trained: 0).
__global__ void k_mlp_grad(uint32_t B, mlp_batch mb, float* __restrict__ G)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)WG_MLP_OFF_END)
        return;
    float g = 0.0f; // stays 0 for the clamp range, normalisation and margins
    if (p >= WG_MLP_OFF_W1 && p < WG_MLP_OFF_B1) {
        const uint32_t q = p - WG_MLP_OFF_W1, i = q / WG_MLP_HIDDEN, j = q % WG_MLP_HIDDEN;
        for (uint32_t b = 0; b < B; ++b)
            g = g + mb.z[b * WG_MLP_IN + i] * mb.dh1[b * WG_MLP_HIDDEN + j];
    } else if (p >= WG_MLP_OFF_B1 && p < WG_MLP_OFF_W2) {
        const uint32_t j = p - WG_MLP_OFF_B1;
        for (uint32_t b = 0; b < B; ++b)
            g = g + mb.dh1[b * WG_MLP_HIDDEN + j];
    } else if (p >= WG_MLP_OFF_W2 && p < WG_MLP_OFF_B2) {
        const uint32_t q = p - WG_MLP_OFF_W2, k = q / WG_MLP_HIDDEN, j = q % WG_MLP_HIDDEN;
        for (uint32_t b = 0; b < B; ++b)
            g = g + mb.h1[b * WG_MLP_HIDDEN + k] * mb.dh2[b * WG_MLP_HIDDEN + j];
    } else if (p >= WG_MLP_OFF_B2 && p < WG_MLP_OFF_W3) {
        const uint32_t j = p - WG_MLP_OFF_B2;
        for (uint32_t b = 0; b < B; ++b)
            g = g + mb.dh2[b * WG_MLP_HIDDEN + j];
    } else if (p >= WG_MLP_OFF_W3 && p < WG_MLP_OFF_B3) {
        const uint32_t q = p - WG_MLP_OFF_W3, j = q / WG_MLP_ACTIONS, o = q % WG_MLP_ACTIONS;
        for (uint32_t b = 0; b < B; ++b)
            if ((uint32_t)mb.o[b] == o)
                g = g + mb.h2[b * WG_MLP_HIDDEN + j] * mb.dy[b];
    } else if (p >= WG_MLP_OFF_B3 && p < WG_MLP_OFF_MARGIN) {
        const uint32_t o = p - WG_MLP_OFF_B3;
        for (uint32_t b = 0; b < B; ++b)
            if ((uint32_t)mb.o[b] == o)
                g = g + mb.dy[b];
    }
    G[p] = g;
}

__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M, float* __restrict__ Vv,
                           float lr_t, float beta1, float beta2, float eps)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)WG_MLP_OFF_W1 || p >= (uint32_t)WG_MLP_OFF_MARGIN) // the decision constants are not trained
        return;
    const float g = G[p];
    const float m = beta1 * M[p] + (1.0f - beta1) * g;
    const float v = beta2 * Vv[p] + (1.0f - beta2) * g * g;
    M[p] = m;
    Vv[p] = v;
    P[p] = P[p] - lr_t * m / (sqrtf(v) + eps);
}

// Squared error of every sample's taken action (final loss).
__global__ __launch_bounds__(WG_WAVE) void k_mlp_err(const float* __restrict__ P, const float* __restrict__ X,
                                                     const int32_t* __restrict__ A, const float* __restrict__ T, uint64_t n,
                                                     float* __restrict__ err)
{
    for (uint64_t s = blockIdx.x; s < n; s += gridDim.x) {
        float x[WG_MLP_IN];
#pragma unroll
        for (int i = 0; i < WG_MLP_IN; ++i)
            x[i] = X[s * WG_MLP_IN + i];
        wg_mlp_lane k;
        (void)wg_mlp_choose_wave(P, WG_APP_BFS, x, nullptr, -1, nullptr, &k);
        const float y = __shfl(k.y, A[s]);
        if (threadIdx.x == 0)
            err[s] = (y - T[s]) * (y - T[s]);
    }
}

static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" int wg_mlp_train(wg_ctx* ctx, const float* features, const int32_t* actions, const float* targets, uint64_t n,
                            const wg_mlp_train_opts* opts, float* params, float* final_loss)
{
    if (!ctx || !opts || !params || n == 0 || !features || !actions || !targets)
        return WG_E_INVAL;
    if (n >= 0xffffffffull || opts->batch == 0 || opts->batch > 65536u || opts->epochs == 0)
        return WG_E_INVAL;
    for (uint64_t s = 0; s < n; ++s)
        if (actions[s] < 0 || actions[s] >= WG_MLP_ACTIONS)
            return WG_E_RANGE;
    WG_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t B = opts->batch;
    const uint32_t NP = WG_MLP_PARAMS;
    std::vector<void*> owned;
    auto dalloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess)
            return nullptr;
        owned.push_back(p);
        return p;
    };
    int rc = WG_OK;
    float* dX = (float*)dalloc(n * WG_MLP_IN * sizeof(float));
    int32_t* dA = (int32_t*)dalloc(n * sizeof(int32_t));
    float* dT = (float*)dalloc(n * sizeof(float));
    uint32_t* dIdx = (uint32_t*)dalloc(n * sizeof(uint32_t));
    float* dP = (float*)dalloc(NP * sizeof(float));
    float* dG = (float*)dalloc(NP * sizeof(float));
    float* dM = (float*)dalloc(NP * sizeof(float));
    float* dV = (float*)dalloc(NP * sizeof(float));
    float* dErr = (float*)dalloc(n * sizeof(float));
    mlp_batch mb;
    mb.z = (float*)dalloc((size_t)B * WG_MLP_IN * sizeof(float));
    mb.h1 = (float*)dalloc((size_t)B * WG_MLP_HIDDEN * sizeof(float));
    mb.h2 = (float*)dalloc((size_t)B * WG_MLP_HIDDEN * sizeof(float));
    mb.dh1 = (float*)dalloc((size_t)B * WG_MLP_HIDDEN * sizeof(float));
    mb.dh2 = (float*)dalloc((size_t)B * WG_MLP_HIDDEN * sizeof(float));
    mb.dy = (float*)dalloc((size_t)B * sizeof(float));
    mb.o = (int32_t*)dalloc((size_t)B * sizeof(int32_t));
    std::vector<uint32_t> perm(n);
    std::vector<float> err(n);
    if (owned.size() != 16) {
        rc = WG_E_NOMEM;
        goto done;
    }
    if (hipMemcpyAsync(dX, features, n * WG_MLP_IN * sizeof(float), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dA, actions, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dT, targets, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dP, params, NP * sizeof(float), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemsetAsync(dM, 0, NP * sizeof(float), ctx->stream) != hipSuccess ||
        hipMemsetAsync(dV, 0, NP * sizeof(float), ctx->stream) != hipSuccess) {
        rc = WG_E_HIP;
        goto done;
    }
    {
        for (uint64_t s = 0; s < n; ++s)
            perm[s] = (uint32_t)s;
        uint64_t step = 0;
        const uint32_t pgrid = (NP + 255) / 256;
        for (uint32_t ep = 0; ep < opts->epochs; ++ep) {
            // Fisher-Yates with a counter-based generator: the order depends on (seed, epoch) only
            for (uint64_t s = n - 1; s > 0; --s) {
                const uint64_t r = mix64(opts->seed ^ mix64(((uint64_t)ep << 32) | s)) % (s + 1);
                const uint32_t t = perm[s];
                perm[s] = perm[r];
                perm[r] = t;
            }
            if (hipMemcpyAsync(dIdx, perm.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
                rc = WG_E_HIP;
                goto done;
            }
            for (uint64_t b0 = 0; b0 < n; b0 += B) {
                const uint32_t bb = (uint32_t)((n - b0) < B ? (n - b0) : B);
                ++step;
                const double c1 = 1.0 - pow((double)opts->beta1, (double)step);
                const double c2 = 1.0 - pow((double)opts->beta2, (double)step);
                const float lr_t = (float)((double)opts->lr * sqrt(c2) / c1);
                hipLaunchKernelGGL(k_mlp_fb, dim3(bb), dim3(WG_WAVE), 0, ctx->stream, dP, dX, dA, dT, dIdx + b0, bb, mb);
                hipLaunchKernelGGL(k_mlp_grad, dim3(pgrid), dim3(256), 0, ctx->stream, bb, mb, dG);
                hipLaunchKernelGGL(k_mlp_adam, dim3(pgrid), dim3(256), 0, ctx->stream, dP, dG, dM, dV, lr_t, opts->beta1,
                                   opts->beta2, opts->eps);
            }
            // the host reshuffles the order buffer next epoch: wait for this epoch's reads
            if (hipStreamSynchronize(ctx->stream) != hipSuccess || (rc = wg_check_launch(ctx))) {
                rc = rc ? rc : WG_E_HIP;
                goto done;
            }
        }
        hipLaunchKernelGGL(k_mlp_err, dim3((uint32_t)(n < 65536 ? n : 65536)), dim3(WG_WAVE), 0, ctx->stream, dP, dX, dA, dT,
                           n, dErr);
        if ((rc = wg_check_launch(ctx)))
            goto done;
        if (hipMemcpyAsync(params, dP, NP * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(err.data(), dErr, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = WG_E_HIP;
            goto done;
        }
        double sum = 0.0;
        for (uint64_t s = 0; s < n; ++s)
            sum += err[s];
        if (final_loss)
            *final_loss = (float)(sum / (double)n);
    }
done:
    for (void* p : owned)
        (void)hipFree(p);
    return rc;
}
