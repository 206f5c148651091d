"""The schedule model's training kernels against an independent float64
restatement (numpy): forward pass, the squared error of the taken action,
backward pass through both ReLU layers, the gradient sums over the minibatch
and Adam with bias correction.  The device runs FP32, so parameters agree to
FP32 rounding, not bit for bit.

Adam's first step moves every parameter by about lr * sign(g) whatever |g|
is, which would pin signs only.  The tests therefore run with eps = 1, where
the step is lr * g / (sqrt(v) + 1), close to linear in g: a dropped or scaled
gradient term changes the update by its own size.

The model has no counterpart in the reference; it replaces the threshold rule
of source/execution_bfs.c:141.
"""
import numpy as np
import pytest

import wedge_amd as W

IN, H, A = W.MLP_IN, W.MLP_HIDDEN, W.MLP_ACTIONS
O_LO, O_HI, O_MEAN, O_ISTD = 0, IN, 2 * IN, 3 * IN
O_W1 = 4 * IN
O_B1 = O_W1 + IN * H
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + H * A
O_END = O_B3 + A
M64 = (1 << 64) - 1


def _mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _epoch_order(perm, seed, ep):
    """The training loop's Fisher-Yates shuffle of the sample order for one
    epoch (counter-based: r = mix64(seed ^ mix64(ep << 32 | s)) mod (s + 1))."""
    n = len(perm)
    for s in range(n - 1, 0, -1):
        r = _mix64(seed ^ _mix64((ep << 32) | s)) % (s + 1)
        perm[s], perm[r] = perm[r], perm[s]
    return perm


def _forward(P, X):
    lo, hi = P[O_LO:O_HI], P[O_HI:O_MEAN]
    z = (np.clip(X, lo, hi) - P[O_MEAN:O_ISTD]) * P[O_ISTD:O_W1]
    W1 = P[O_W1:O_B1].reshape(IN, H)
    W2 = P[O_W2:O_B2].reshape(H, H)
    W3 = P[O_W3:O_B3].reshape(H, A)
    a1 = z @ W1 + P[O_B1:O_W2]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + P[O_B2:O_W3]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ W3 + P[O_B3:O_END]
    return z, a1, h1, a2, h2, y


def _grad(P, X, acts, T):
    """d/dP of mean_s (y[s, acts[s]] - T[s])^2 over the minibatch (float64)."""
    B = X.shape[0]
    z, a1, h1, a2, h2, y = _forward(P, X)
    W2 = P[O_W2:O_B2].reshape(H, H)
    W3 = P[O_W3:O_B3].reshape(H, A)
    dy = 2.0 * (y[np.arange(B), acts] - T) / B
    dY = np.zeros((B, A))
    dY[np.arange(B), acts] = dy
    d2 = (dY @ W3.T) * (a2 > 0)
    d1 = (d2 @ W2.T) * (a1 > 0)
    G = np.zeros_like(P)
    G[O_W1:O_B1] = (z.T @ d1).ravel()
    G[O_B1:O_W2] = d1.sum(0)
    G[O_W2:O_B2] = (h1.T @ d2).ravel()
    G[O_B2:O_W3] = d2.sum(0)
    G[O_W3:O_B3] = (h2.T @ dY).ravel()
    G[O_B3:O_END] = dY.sum(0)
    return G, min(np.abs(a1).min(), np.abs(a2).min())


def _train_ref(P0, X, acts, T, epochs, batch, lr, seed, b1, b2, eps):
    """Minibatch Adam as the training entry point defines it, in float64."""
    P = P0.astype(np.float64).copy()
    m = np.zeros_like(P)
    v = np.zeros_like(P)
    n = X.shape[0]
    perm = list(range(n))
    step = 0
    margin = np.inf
    for ep in range(epochs):
        perm = _epoch_order(perm, seed, ep)
        for b0 in range(0, n, batch):
            idx = np.array(perm[b0:b0 + batch])
            step += 1
            g, mg = _grad(P, X[idx].astype(np.float64), acts[idx], T[idx].astype(np.float64))
            margin = min(margin, mg)
            m = b1 * m + (1 - b1) * g
            v = b2 * v + (1 - b2) * g * g
            lr_t = lr * np.sqrt(1 - b2**step) / (1 - b1**step)
            upd = lr_t * m / (np.sqrt(v) + eps)
            P[O_W1:O_END] -= upd[O_W1:O_END]
    return P, margin


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(W.MLP_PARAMS, np.float32)
    P[O_LO:O_HI] = -3.0
    P[O_HI:O_MEAN] = 3.0
    P[O_MEAN:O_ISTD] = rng.normal(0, 0.3, IN)
    P[O_ISTD:O_W1] = rng.uniform(0.5, 1.5, IN)
    P[O_W1:O_B1] = rng.normal(0, 1 / np.sqrt(IN), IN * H)
    P[O_B1:O_W2] = rng.normal(0, 0.1, H)
    P[O_W2:O_B2] = rng.normal(0, 1 / np.sqrt(H), H * H)
    P[O_B2:O_W3] = rng.normal(0, 0.1, H)
    P[O_W3:O_B3] = rng.normal(0, 1 / np.sqrt(H), H * A)
    P[O_B3:O_END] = rng.normal(0, 0.1, A)
    P[O_END:] = [0.2, 0.5, 0.5, 30.0]  # decision margins and size bound: not trained
    X = rng.normal(0, 1.5, (n, IN)).astype(np.float32)  # some inputs clamp at +-3
    acts = rng.integers(0, A, n).astype(np.int32)
    T = rng.normal(0, 1, n).astype(np.float32)
    return P, X, acts, T


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,epochs,batch", [(96, 1, 96), (70, 2, 32), (40, 3, 7)])
def test_training_matches_float64_adam(ctx, n, epochs, batch):
    lr, seed, b1, b2, eps = 5e-2, 5, 0.9, 0.999, 1.0
    P0, X, acts, T = _problem(n, seed=n)
    want, margin = _train_ref(P0, X, acts, T, epochs, batch, lr, seed, b1, b2, eps)
    assert margin > 1e-5  # no pre-activation so close to 0 that FP32 could flip its ReLU
    got, _ = ctx.mlp_train(X, acts, T, P0, epochs=epochs, batch=batch, lr=lr, seed=seed, beta1=b1, beta2=b2, eps=eps)
    got = got.astype(np.float64)
    moved = want - P0.astype(np.float64)
    d_got = got - P0.astype(np.float64)
    err = np.abs(d_got - moved)
    print(f"n={n} epochs={epochs} batch={batch}: max |step| {np.abs(moved).max():.3e}, "
          f"max |device - float64| {err.max():.3e}")
    # untrained entries are untouched
    assert np.array_equal(got[:O_W1], P0[:O_W1].astype(np.float64))
    assert np.array_equal(got[O_END:], P0[O_END:].astype(np.float64))
    # every trained entry moved as the float64 restatement moved it, to FP32 rounding
    assert np.all(err <= 1e-6 + 1e-3 * np.abs(moved)), f"worst at {int(np.argmax(err))}"
    assert np.count_nonzero(moved[O_W1:O_END]) > 0.5 * (O_END - O_W1)


def test_planted_gradient_mistake_is_caught():
    """The comparison fails when the float64 side drops one gradient term
    (the output-bias gradient), so it can tell such a mistake."""
    P0, X, acts, T = _problem(96, seed=96)
    good, _ = _train_ref(P0, X, acts, T, 1, 96, 5e-2, 5, 0.9, 0.999, 1.0)
    g, _ = _grad(P0.astype(np.float64), X.astype(np.float64), acts, T.astype(np.float64))
    g[O_B3:O_END] = 0.0
    lr_t = 5e-2 * np.sqrt(1 - 0.999) / (1 - 0.9)
    m, v = 0.1 * g, 0.001 * g * g
    bad = P0.astype(np.float64).copy()
    bad[O_W1:O_END] -= (lr_t * m / (np.sqrt(v) + 1.0))[O_W1:O_END]
    err = np.abs(bad - good)
    assert np.any(err > 1e-6 + 1e-3 * np.abs(good - P0))
