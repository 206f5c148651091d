"""Learned schedule on the MI355X: the device decision code, the training
kernels, and that a learned schedule never changes results.

Bars: device inference picks the same action as the host forward pass and
its predictions agree to FP32 rounding (the same operations in the same
order); training on data from a known network lowers the error and is
reproducible run to run; BFS/CC/SSSP outputs under any model equal the
oracle's bit for bit.
"""
import numpy as np
import pytest

import oracle
import wedge_amd as W

from helpers import U32_MAX

pytestmark = pytest.mark.gpu


def _params(seed, scale=0.5, lo=-20.0, hi=40.0):
    rng = np.random.default_rng(seed)
    I = W.MLP_IN
    p = rng.normal(0.0, scale, W.MLP_PARAMS).astype(np.float32)
    p[-4:-1] = (0.0, 0.3, 0.3)  # margins against the plain rule
    p[-1] = 1e30                # no size bound
    p[:I] = lo
    p[I:2 * I] = hi
    p[2 * I:3 * I] = rng.normal(0.0, 1.0, I)
    p[3 * I:4 * I] = rng.uniform(0.1, 1.0, I)
    return p


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()
    W.mlp_unload()


def test_device_inference_matches_host(ctx):
    p = _params(11)
    ctx.set_schedule_model(p)
    W.mlp_set(p)
    rng = np.random.default_rng(5)
    apps = (W.APP_BFS, W.APP_CC, W.APP_SSSP)
    X = np.stack([W.mlp_features(apps[int(rng.integers(0, 3))], int(rng.integers(0, 1000)),
                                 1 << 20, 1 << 24, int(rng.integers(0, 1 << 20)), int(rng.integers(0, 1 << 24)),
                                 int(rng.integers(0, 1 << 20))) for _ in range(300)])
    X = np.concatenate([X, rng.normal(0.0, 4.0, (300, W.MLP_IN)).astype(np.float32)])
    B = np.stack([W.mlp_base(int(rng.integers(1, 2**26)), int(rng.integers(1, 2**30)), int(rng.integers(0, 2**24)))
                  for _ in range(X.shape[0])])
    plain = rng.choice([-1, 0, 2, 5], X.shape[0]).astype(np.int32)
    for app in (W.APP_BFS, W.APP_SSSP):
        Y, A = ctx.mlp_infer(X, app, B, plain)
        for i in range(X.shape[0]):
            a, y = W.mlp_predict(X[i], app, B[i], int(plain[i]))
            assert A[i] == a, i
            assert np.array_equal(Y[i], y), (i, Y[i], y)  # bit-identical: same FP32 operations in the same order
        assert len(set(A.tolist())) > 1  # the test exercises more than one action


def test_training_learns_a_known_network_and_is_reproducible(ctx):
    teacher = _params(21, scale=0.3)
    rng = np.random.default_rng(3)
    n = 4000
    X = rng.normal(0.0, 1.5, (n, W.MLP_IN)).astype(np.float32)
    W.mlp_set(teacher)
    Y = np.stack([W.mlp_predict(x)[1] for x in X])
    acts = rng.integers(0, W.MLP_ACTIONS, n).astype(np.int32)
    T = Y[np.arange(n), acts].astype(np.float32)
    p0 = _params(22, scale=0.2)
    p0[:4 * W.MLP_IN] = teacher[:4 * W.MLP_IN]  # same input clamp and normalisation
    p0[-4:] = teacher[-4:]
    _, loss0 = ctx.mlp_train(X, acts, T, p0, epochs=1, batch=n, lr=1e-9, seed=1)
    p1, loss1 = ctx.mlp_train(X, acts, T, p0, epochs=60, batch=128, lr=3e-3, seed=1)
    p2, loss2 = ctx.mlp_train(X, acts, T, p0, epochs=60, batch=128, lr=3e-3, seed=1)
    assert np.array_equal(p1, p2) and loss1 == loss2  # fixed-order sums: reproducible
    assert loss1 < 0.2 * loss0, (loss0, loss1)
    assert np.array_equal(p1[:4 * W.MLP_IN], p0[:4 * W.MLP_IN])  # clamp and normalisation are not trained
    assert np.array_equal(p1[-4:], p0[-4:])  # nor are the decision constants
    with pytest.raises(W.WedgeError):
        ctx.mlp_train(X, np.full(n, W.MLP_ACTIONS, np.int32), T, p0, epochs=1)


def _check_same(g, ctx, app, root, **kw):
    if app == "bfs":
        out, st = ctx.run_bfs(root, **kw)
        _, depth, _ = oracle.Oracle.of(g, 4).bfs(root)
        assert np.array_equal(W.bfs_depth(out), depth)
    elif app == "cc":
        out, st = ctx.run_cc(**kw)
        lab, _ = oracle.Oracle.of(g, 4).cc()
        assert np.array_equal(out, lab.astype(np.uint32))
    else:
        out, st = ctx.run_sssp(root, **kw)
        want, _ = oracle.Oracle.of(g, 4).sssp(root)
        if out.dtype == np.uint32:
            out = np.where(out == U32_MAX, np.uint64(2**64 - 1), out.astype(np.uint64))
        assert np.array_equal(out, want)
    return out, st


@pytest.mark.parametrize("seed", [31, 32, 33])
def test_learned_schedule_keeps_results(ctx, seed):
    p = _params(seed, scale=1.0)
    W.mlp_set(p)
    for g, apps in ((W.rmat(12, 16, seed=2, weights="f64"), ("bfs", "sssp")),
                    (W.rmat(11, 16, seed=3, symmetrize=True), ("cc",)),
                    (W.grid(40, 70, seed=4), ("bfs", "sssp", "cc"))):
        ctx.load(g)
        root = g.max_outdegree_vertex()
        for app in apps:
            _, st = _check_same(g, ctx, app, root, schedule="learned")               # device loop
            _, st_h = _check_same(g, ctx, app, root, schedule="learned", record_trace=1)  # host loop
            if app == "bfs":  # level-synchronous: same decisions give the same counts
                assert (st["iterations"], st["frontier_iterations"]) == (st_h["iterations"],
                                                                         st_h["frontier_iterations"])
                rows = W.trace_rows()
                for r in rows:  # the host loop's choices are the model's
                    if r["frontier_edges"] <= g.num_edges >> W.MLP_GATE_LOG2:  # gated: the plain rule's frontier pull
                        assert (r["mode"], r["block_log2"]) == (W.MODE_FRONTIER, W.DEFAULT_BLOCK_LOG2)
                        continue
                    x = W.mlp_features(W.APP_BFS, r["iteration"], g.num_vertices, g.num_edges,
                                       r["frontier_vertices"], r["frontier_edges"], r["unvisited"])
                    plain = 0 if r["frontier_edges"] > g.num_edges * 20 // 100 else 2
                    a, _ = W.mlp_predict(x, W.APP_BFS, W.mlp_base(g.num_vertices, g.num_edges, r["frontier_edges"]),
                                         plain)
                    assert W.mlp_action(a)[0] == r["mode"]
                    if r["mode"] == W.MODE_FRONTIER:
                        assert W.mlp_action(a)[1] == r["block_log2"]
    W.mlp_unload()


def test_learned_without_model_is_plain(ctx):
    W.mlp_unload()
    g = W.rmat(12, 16, seed=5)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    out_l, st_l = ctx.run_bfs(root, schedule="learned")
    out_p, st_p = ctx.run_bfs(root, schedule="plain")
    assert np.array_equal(out_l, out_p)
    assert st_l["frontier_iterations"] == st_p["frontier_iterations"]


def test_trace_rows_carry_device_times(ctx):
    g = W.grid(30, 30, seed=6)
    ctx.load(g)
    ctx.run_bfs(g.max_outdegree_vertex(), schedule="frontier", block_log2=4, record_trace=1)
    rows = W.trace_rows()
    assert len(rows) > 10
    assert all(r["gpu_pull_ms"] > 0.0 for r in rows)
    assert all(r["gpu_build_ms"] > 0.0 for r in rows[:-1])  # the last pull builds nothing


def test_cc_sssp_keep_the_plain_rule_beyond_the_model_range(ctx):
    p = _params(41, scale=1.0)
    p[-1] = 1.0  # every graph is larger than the (pretend) training graphs
    W.mlp_set(p)
    g = W.grid(40, 70, seed=4)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    _check_same(g, ctx, "sssp", root, schedule="learned", record_trace=1)
    threshold = g.num_edges * 20 // 100
    for r in W.trace_rows():  # every choice is the plain rule's
        assert r["mode"] == (W.MODE_DENSE if r["frontier_edges"] > threshold else W.MODE_FRONTIER)
        assert r["mode"] == W.MODE_DENSE or r["block_log2"] == W.DEFAULT_BLOCK_LOG2
    W.mlp_unload()
