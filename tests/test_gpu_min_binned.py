"""Dense CC and SSSP pulls on the binned layout (WG_OPT_MIN_BINNED) on the
MI355X: labels and distances bit-exact against the oracle, FP64 and integer
weights, in every schedule, in the device-resident loop and in the phase
API's host loop, relabelled and in native order.

The binned pull reads the values the previous pull left (a Jacobi step)
where the tile pull may also see values lowered earlier in the same pull, so
iteration counts may differ; the fixpoint (perform_edge_pull_phase_cc,
source/phase_edge_pull_cc.c, and perform_edge_pull_phase_sssp,
source/phase_edge_pull_sssp.c, run to convergence) is the same."""
import numpy as np
import pytest

import oracle
import wedge_amd as W

pytestmark = pytest.mark.gpu

SCHEDULES = [dict(schedule="plain"), dict(schedule="dense"), dict(schedule="frontier"),
             dict(schedule="plain", threshold_pct=2), dict(schedule="plain", record_trace=1),
             dict(schedule="dense", record_trace=1)]


@pytest.fixture(scope="module", params=[True, False], ids=["relabel", "native-order"])
def ctx(request):
    c = W.Context(0, relabel=request.param)
    c.set_option(W.OPT_MIN_BINNED, 1)
    yield c
    c.close()


def _with_u32_weights(g, seed, wmax=1000):
    rng = np.random.default_rng(seed)
    w = rng.integers(1, wmax + 1, g.num_edges).astype(np.uint32)
    return W.from_edges(g.num_vertices, g.sources, g.destinations(), weights=w)


def _hub(seed=7, V=6000):
    """A destination with more in-edges than a push tile holds, hub sources
    with long out-lists, and random edges."""
    rng = np.random.default_rng(seed)
    src = np.concatenate([rng.integers(1, V, 5000), np.full(3000, 5), rng.integers(0, V, 20000)])
    dst = np.concatenate([np.zeros(5000, np.int64), rng.integers(0, V, 3000), rng.integers(0, V, 20000)])
    return W.from_edges(V, src, dst, dedup=True)


def _graph(kind):
    if kind == "rmat_sym":
        return W.rmat(13, 16, seed=2, symmetrize=True)
    if kind == "rmat":
        return W.rmat(13, 16, seed=3)
    if kind == "rmat_w":
        return W.rmat(13, 16, seed=4, weights="f64")
    if kind == "rmat_u32":
        return _with_u32_weights(W.rmat(12, 16, seed=5), seed=5)
    if kind == "grid":
        return W.grid(40, 70, seed=6)
    if kind == "hub":
        return _hub()
    if kind == "hub_u32":
        return _with_u32_weights(_hub(seed=8), seed=8)
    raise ValueError(kind)


def _sssp_ok(g, dist, ref):
    if g.weight_kind == W.WEIGHT_U32:
        got = np.where(dist == W.UNREACHED_U32, np.uint64(2**64 - 1), dist.astype(np.uint64))
        return np.array_equal(got, ref)
    return np.array_equal(dist, ref)


@pytest.mark.parametrize("kind", ["rmat_sym", "rmat", "grid", "hub"])
def test_binned_cc_matches_oracle(ctx, kind):
    g = _graph(kind)
    ctx.load(g)
    ref, _ = oracle.Oracle.of(g, 4).cc()
    for opts in SCHEDULES:
        labels, st = ctx.run_cc(**opts)
        assert np.array_equal(labels.astype(np.float64), ref), f"{kind} {opts}"
        print(f"CC {kind} {opts}: iterations {st['iterations']} frontier {st['frontier_iterations']}")


@pytest.mark.parametrize("kind", ["rmat_w", "rmat_u32", "grid", "hub_u32"])
def test_binned_sssp_matches_oracle(ctx, kind):
    g = _graph(kind)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    ref, _ = oracle.Oracle.of(g, 4).sssp(root)
    for opts in SCHEDULES:
        dist, st = ctx.run_sssp(root, **opts)
        assert _sssp_ok(g, dist, ref), f"{kind} {opts}"
        print(f"SSSP {kind} {opts}: iterations {st['iterations']} frontier {st['frontier_iterations']}")


def test_binned_min_bins_split_over_many_workgroups():
    """Bins cut into slices of 29 pairs meet through the slot minima and the
    per-bin ticket."""
    c = W.Context(0)
    try:
        c.set_option(W.OPT_MIN_BINNED, 1)
        c.set_option(W.OPT_PR_BIN_SLICE, 29)
        g = W.rmat(13, 16, seed=9, weights="f64")
        c.load(g)
        root = g.max_outdegree_vertex()
        ref, _ = oracle.Oracle.of(g, 4).sssp(root)
        for sched in ("dense", "plain"):
            dist, _ = c.run_sssp(root, schedule=sched)
            assert _sssp_ok(g, dist, ref), sched
        gs = W.rmat(12, 16, seed=10, symmetrize=True)
        c.load(gs)
        ref, _ = oracle.Oracle.of(gs, 4).cc()
        labels, _ = c.run_cc(schedule="dense")
        assert np.array_equal(labels.astype(np.float64), ref)
    finally:
        c.close()


def test_binned_min_pulls_count_every_edge():
    """With the work counts on, a binned dense pull reads E edges and examines
    V destinations, like the tile pull it replaces."""
    c = W.Context(0)
    try:
        c.set_option(W.OPT_MIN_BINNED, 1)
        c.set_option(W.OPT_COUNT_WORK, 1)
        gs = W.rmat(12, 8, seed=8, symmetrize=True)
        c.load(gs)
        c.init_cc()
        it, _ = c.converge(W.SCHED_DENSE)
        e, d = c.work_counts()
        assert (e, d) == (it * gs.num_edges, it * gs.num_vertices)
        gw = W.rmat(12, 8, seed=9, weights="f64")
        c.load(gw)
        c.init_sssp(gw.max_outdegree_vertex())
        it, _ = c.converge(W.SCHED_DENSE)
        e, d = c.work_counts()
        assert (e, d) == (it * gw.num_edges, it * gw.num_vertices)
    finally:
        c.close()


def test_binned_and_tile_dense_pulls_reach_the_same_fixpoint_in_the_phase_api():
    """Dense pulls one at a time through wg_edge_pull_sssp / _cc until nothing
    changes, binned and on the tiles: the same distances and labels."""
    c = W.Context(0)
    try:
        g = W.rmat(12, 16, seed=11, weights="f64")
        c.load(g)
        root = g.max_outdegree_vertex()
        out = {}
        for opt in (0, 1):
            c.set_option(W.OPT_MIN_BINNED, opt)
            c.init_sssp(root)
            pulls = 0
            while c.edge_pull(W.APP_SSSP, W.MODE_DENSE):
                c.advance_frontier()
                pulls += 1
                assert pulls < 200
            out[opt] = c.read_sssp()
        assert np.array_equal(out[0], out[1])
        ref, _ = oracle.Oracle.of(g, 4).sssp(root)
        assert np.array_equal(out[1], ref)
    finally:
        c.close()


@pytest.mark.parametrize("kind", ["rmat_w", "rmat_sym"])
def test_learned_schedule_takes_binned_dense_pulls_above_e_over_64(ctx, kind):
    """With the dense pulls on the binned layout, the learned schedule (device
    loop and host loop, model loaded) pulls dense whenever the activity
    exceeds E / 64, where the plain rule keeps frontier pulls up to 20 % of E;
    the results stay the oracle's."""
    g = _graph(kind)
    ctx.load(g)
    cc = kind == "rmat_sym"
    root = g.max_outdegree_vertex()
    ref = oracle.Oracle.of(g, 4).cc()[0] if cc else oracle.Oracle.of(g, 4).sssp(root)[0]
    W.mlp_load()
    try:
        for trace in (0, 1):
            if cc:
                out, st = ctx.run_cc(schedule="learned", record_trace=trace)
                assert np.array_equal(out.astype(np.float64), ref)
            else:
                out, st = ctx.run_sssp(root, schedule="learned", record_trace=trace)
                assert _sssp_ok(g, out, ref)
            print(f"{kind} learned trace={trace}: iterations {st['iterations']} frontier {st['frontier_iterations']}")
        assert ctx.dense_pull_binned()
        rows = W.trace_rows()
        floor, plain = g.num_edges >> 6, g.num_edges * 20 // 100
        for r in rows:
            if r["frontier_edges"] > floor:
                assert r["mode"] == W.MODE_DENSE, r
        # the rule decided at least one pull the plain rule leaves frontier-filtered
        assert any(floor < r["frontier_edges"] <= plain for r in rows), [r["frontier_edges"] for r in rows]
    finally:
        W.mlp_unload()
