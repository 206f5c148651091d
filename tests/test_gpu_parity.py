"""HIP path vs the oracle on the same seeded graphs (MI355X only).

Bars: BFS depths, CC labels and SSSP distances bit-exact; BFS parents form a
valid shortest-path tree and equal the smallest in-neighbour one level up (the
HIP path's definition, identical for every schedule); PageRank within 1e-10 L1
of the FP64 oracle (FP32-stored contributions: within 2e-7 L1, error reported).
"""
import numpy as np
import pytest

import oracle
import wedge_amd as W

from helpers import U32_MAX, check_bfs_tree, min_frontier_parent

pytestmark = pytest.mark.gpu

PR_TOL = 1e-10  # L1, FP64 contributions
PR_TOL_FP32 = 2e-7  # L1, FP32-stored contributions (about 10x the measured error)


@pytest.fixture(scope="module", params=[True, False], ids=["relabel", "native-order"])
def ctx(request):
    """Every test runs with the device-side degree relabelling forced on and off."""
    c = W.Context(0, relabel=request.param)
    yield c
    c.close()


def _graph(kind, seed=1, scale=12):
    if kind == "rmat":
        return W.rmat(scale, 16, seed=seed)
    if kind == "rmat_sym":
        return W.rmat(scale, 16, seed=seed, symmetrize=True)
    if kind == "rmat_w":
        return W.rmat(scale, 16, seed=seed, weights="f64")
    if kind == "grid":
        return W.grid(48, 80, seed=seed)
    raise ValueError(kind)


def _hub_graph(V=6000, hub_in=5000, seed=7):
    """A destination with more in-edges than one tile (heavy tiles) plus a
    source hub with a long out-edge list."""
    rng = np.random.default_rng(seed)
    src = [rng.integers(1, V, hub_in)]
    dst = [np.zeros(hub_in, np.int64)]
    src.append(np.full(3000, 5, np.int64))
    dst.append(rng.integers(0, V, 3000))
    s = rng.integers(0, V, 20000)
    d = rng.integers(0, V, 20000)
    src.append(s)
    dst.append(d)
    return W.from_edges(V, np.concatenate(src), np.concatenate(dst), dedup=True)


# ---- PageRank ------------------------------------------------------------------


@pytest.mark.parametrize("scale,seed,damping", [(12, 1, 0.85), (14, 1, 0.875), (13, 5, 0.85)])
def test_pr_fused_matches_oracle(ctx, scale, seed, damping):
    g = W.rmat(scale, 16, seed=seed)
    ctx.load(g)
    ranks, st = ctx.run_pr(iterations=16, damping=damping)
    ref, _ = oracle.Oracle.of(g).pr(16, damping)
    l1 = float(np.abs(ranks - ref).sum())
    print(f"PR s{scale} d={damping}: L1 {l1:.3e} sum {ranks.sum():.15f}")
    assert l1 <= PR_TOL
    assert abs(ranks.sum() - 1.0) < 1e-9


def test_pr_phase_api_matches_fused(ctx):
    """wg_edge_pull_pr + wg_vertex_pr (the reference's two phases) and the fused
    iteration agree, also when interleaved."""
    g = W.rmat(12, 16, seed=3)
    ctx.load(g)
    ref, _ = oracle.Oracle.of(g).pr(6, 0.875)
    ctx.init_pr(0.875)
    totals = []
    for it in range(6):
        if it % 2 == 0:
            totals.append(ctx.edge_pull_pr())
            ctx.vertex_pr()
        else:
            totals.append(ctx.pr_iteration())
    ranks = ctx.read_ranks()
    l1 = float(np.abs(ranks - ref).sum())
    print(f"PR phase API: L1 {l1:.3e}, totals {totals}")
    assert l1 <= PR_TOL
    assert all(0.0 < t <= 1.0 + 1e-12 for t in totals)


def test_pr_fp32_contributions(ctx):
    g = W.rmat(13, 16, seed=1)
    ctx.load(g)
    ranks, _ = ctx.run_pr(iterations=16, damping=0.85, contrib_fp32=1)
    ref, _ = oracle.Oracle.of(g).pr(16, 0.85)
    l1 = float(np.abs(ranks - ref).sum())
    print(f"PR FP32 contributions: L1 vs FP64 oracle {l1:.3e}")
    assert l1 <= PR_TOL_FP32


def test_pr_heavy_destinations(ctx):
    g = _hub_graph()
    ctx.load(g)
    ranks, _ = ctx.run_pr(iterations=10, damping=0.85)
    ref, _ = oracle.Oracle.of(g).pr(10, 0.85)
    assert float(np.abs(ranks - ref).sum()) <= PR_TOL


def test_pr_deterministic(ctx):
    g = W.rmat(13, 16, seed=2)
    ctx.load(g)
    a, _ = ctx.run_pr(iterations=5, damping=0.85)
    b, _ = ctx.run_pr(iterations=5, damping=0.85)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [(4, 4, 0), (4, 16, 1), (16, 2, 0), (16, 8, 1), (8, 2, 1), (4, 1, 0), (16, 1, 1),
                                   (2, 1, 0), (2, 4, 1)])
def test_pr_tile_shapes_match_oracle(ctx, shape):
    """Every PageRank tile shape (edges per lane, workgroups per CU, XCD split;
    1 workgroup per CU: 16 wavefronts with a 64 KB contribution cache) on a
    power-law graph and on a graph whose hubs exceed every tile size; FP64 and
    FP32-stored contributions, fused and two-phase iterations."""
    for g, its in ((W.rmat(13, 16, seed=4), 16), (_hub_graph(), 10)):
        ctx.load(g)
        ctx.set_pr_shape(shape)
        assert ctx.pr_shape() == shape
        ref, _ = oracle.Oracle.of(g).pr(its, 0.85)
        ranks, _ = ctx.run_pr(iterations=its, damping=0.85)
        assert float(np.abs(ranks - ref).sum()) <= PR_TOL, shape
        ranks32, _ = ctx.run_pr(iterations=its, damping=0.85, contrib_fp32=1)
        assert float(np.abs(ranks32 - ref).sum()) <= PR_TOL_FP32, shape
        ctx.init_pr(0.85)
        for _ in range(its):
            ctx.edge_pull_pr()
            ctx.vertex_pr()
        assert float(np.abs(ctx.read_ranks() - ref).sum()) <= PR_TOL, shape
    ctx.set_pr_shape(None)
    assert ctx.pr_shape() == (8, 8, 0)


def test_pr_autotune_resets_the_state(ctx):
    g = W.rmat(13, 16, seed=4)
    ctx.load(g)
    ctx.init_pr(0.85)
    best, times = ctx.pr_autotune(2)
    assert len(times) == W.PR_NSHAPES and best in times
    assert times[best] == min(times.values()) and ctx.pr_shape() == best
    print("PR autotune:", best, {k: round(v, 4) for k, v in times.items()})
    for _ in range(12):
        ctx.pr_iteration()
    tuned = ctx.read_ranks()
    again, _ = ctx.run_pr(iterations=12, damping=0.85)  # fresh run, same shape
    if best[2] == W.PR_BINNED:  # its per-bin sums are reproducible to rounding only
        assert float(np.abs(tuned - again).sum()) <= 1e-13
    else:
        assert np.array_equal(tuned, again)
    ref, _ = oracle.Oracle.of(g).pr(12, 0.85)
    assert float(np.abs(tuned - ref).sum()) <= PR_TOL
    with pytest.raises(W.WedgeError):
        ctx.set_pr_shape((5, 4, 0))
    ctx.set_pr_shape(None)


# ---- BFS -------------------------------------------------------------------------

# device-resident loop (default) and the host-driven loop of the phase API
# (record_trace=1), every pull mode and several Wedge granularities
SCHEDULES = [dict(schedule="plain"), dict(schedule="dense"), dict(schedule="frontier"),
             dict(schedule="frontier", block_log2=3), dict(schedule="frontier", block_log2=6),
             dict(schedule="plain", block_log2=8), dict(schedule="plain", threshold_pct=2),
             dict(schedule="plain", record_trace=1), dict(schedule="frontier", block_log2=5, record_trace=1)]


@pytest.mark.parametrize("kind", ["rmat", "rmat_sym", "grid"])
def test_bfs_matches_oracle_all_schedules(ctx, kind):
    g = _graph(kind)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    _, depth_ref, _ = oracle.Oracle.of(g).bfs(root)
    want_parent = min_frontier_parent(g, depth_ref)
    first = None
    for opts in SCHEDULES:
        packed, st = ctx.run_bfs(root, **opts)
        depth = W.bfs_depth(packed)
        parent = W.bfs_parent(packed)
        assert np.array_equal(depth, depth_ref), f"{kind} {opts}: depths differ"
        assert check_bfs_tree(g, root, depth, parent) == 0
        reached = (depth != U32_MAX) & (np.arange(g.num_vertices) != root)
        assert np.array_equal(parent[reached], want_parent[reached]), f"{kind} {opts}: parent rule"
        assert parent[root] == root
        if first is None:
            first = packed
        assert np.array_equal(packed, first), f"{kind} {opts}: output depends on schedule"
        print(f"BFS {kind} {opts}: iters {st['iterations']} frontier {st['frontier_iterations']}")


def test_bfs_heavy_and_isolated_root(ctx):
    g = _hub_graph()
    ctx.load(g)
    for root in (0, 5, g.max_outdegree_vertex()):
        _, depth_ref, _ = oracle.Oracle.of(g).bfs(root)
        for opts in SCHEDULES[:3]:
            packed, _ = ctx.run_bfs(root, **opts)
            assert np.array_equal(W.bfs_depth(packed), depth_ref)
    # a root with no out-edges reaches nothing
    od = g.outdegrees()
    lonely = int(np.nonzero(od == 0)[0][0])
    packed, st = ctx.run_bfs(lonely)
    assert int((packed != W.UNREACHED_U64).sum()) == 1 and st["iterations"] == 1


# ---- CC ---------------------------------------------------------------------------


@pytest.mark.parametrize("kind", ["rmat_sym", "rmat", "grid"])
def test_cc_matches_oracle_all_schedules(ctx, kind):
    g = _graph(kind, seed=2)
    ctx.load(g)
    ref, _ = oracle.Oracle.of(g).cc()
    for opts in SCHEDULES:
        labels, st = ctx.run_cc(**opts)
        assert np.array_equal(labels.astype(np.float64), ref), f"{kind} {opts}"
        print(f"CC {kind} {opts}: iters {st['iterations']} frontier {st['frontier_iterations']}")


# ---- SSSP -------------------------------------------------------------------------


@pytest.mark.parametrize("kind", ["rmat_w", "grid"])
def test_sssp_matches_oracle_all_schedules(ctx, kind):
    g = _graph(kind, seed=2)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    ref, _ = oracle.Oracle.of(g).sssp(root)
    for opts in SCHEDULES:
        dist, st = ctx.run_sssp(root, **opts)
        if g.weight_kind == W.WEIGHT_U32:
            got = np.where(dist == W.UNREACHED_U32, np.uint64(2**64 - 1), dist.astype(np.uint64))
            assert np.array_equal(got, ref), f"{kind} {opts}"
        else:
            assert np.array_equal(dist, ref), f"{kind} {opts}"
        print(f"SSSP {kind} {opts}: iters {st['iterations']} frontier {st['frontier_iterations']}")


# ---- phase-level API edge cases ------------------------------------------------------


def test_single_vertex_graph(ctx):
    g = W.from_edges(1, [], [])
    ctx.load(g)
    ranks, _ = ctx.run_pr(iterations=3, damping=0.85)
    assert abs(ranks[0] - 1.0) < 1e-15
    packed, _ = ctx.run_bfs(0)
    assert packed[0] == 0
    labels, _ = ctx.run_cc()
    assert labels[0] == 0


def test_bad_arguments_are_errors(ctx):
    g = W.rmat(8, 4, seed=1)
    ctx.load(g)
    with pytest.raises(W.WedgeError):
        ctx.init_bfs(g.num_vertices)  # root out of range
    ctx.init_bfs(0)
    with pytest.raises(W.WedgeError):
        ctx.edge_pull(W.APP_BFS, W.MODE_FRONTIER)  # no frontier built yet
    with pytest.raises(W.WedgeError):
        ctx.wedge_generate_frontier(2)  # granularity below the minimum
    with pytest.raises(W.WedgeError):
        ctx.init_sssp(0)  # unweighted graph
    with pytest.raises(W.WedgeError):
        ctx.set_option(W.OPT_RELABEL, 3)  # not a relabelling mode


@pytest.mark.gpu
def test_default_relabel_choice_keeps_results():
    """The default context relabels a skewed graph and keeps a near-regular one
    in place; results are identical to the oracle either way."""
    c = W.Context(0)
    try:
        for g in (W.rmat(12, 16, seed=3), W.grid(48, 64, seed=5)):
            c.load(g)
            root = g.max_outdegree_vertex()
            packed, _ = c.run_bfs(root)
            _, depth, _ = oracle.Oracle.of(g, 4).bfs(root)
            assert np.array_equal(W.bfs_depth(packed), depth)
            if g.weights is not None:
                dist, _ = c.run_sssp(root)
                want, _ = oracle.Oracle.of(g, 4).sssp(root)
                got = np.where(dist == W.UNREACHED_U32, np.uint64(2**64 - 1), dist.astype(np.uint64))
                assert np.array_equal(got, want)
    finally:
        c.close()
