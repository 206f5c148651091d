"""Fused small-frontier BFS levels (wg_set_fused_levels) on the MI355X.

With the fused levels on, wg_converge must return the same packed (depth,
parent) words, the same iteration and frontier-iteration counts as the
level-per-launch loop (fused levels off), on graphs whose levels are small
(grids, the tails of R-MAT traversals), with several roots, on unions of
member graphs, under every schedule, for several workgroup counts, and when an
iteration cap ends the loop inside a run.  Depths are checked against the
oracle, parents against the smallest-in-neighbour rule."""
import numpy as np
import pytest

import oracle
import wedge_amd as W

from helpers import U32_MAX, min_frontier_parent

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[True, False], ids=["relabel", "native-order"])
def ctx(request):
    c = W.Context(0, relabel=request.param)
    yield c
    c.close()


@pytest.fixture(scope="module")
def model():
    W.mlp_load()
    return W.mlp_get()


def _pair(ctx, init, sched=W.SCHED_PLAIN, max_iterations=0, wgs=32):
    """(packed, (iterations, frontier iterations), fused levels) with fused levels off, then on."""
    out = []
    for n in (0, wgs):
        ctx.set_fused_levels(n)
        init()
        it = ctx.converge(sched, max_iterations=max_iterations)
        out.append((ctx.read_bfs(), it, ctx.fused_level_count()))
    ctx.set_fused_levels(32)
    return out


def _check_against_oracle(g, roots, packed):
    o = oracle.Oracle.of(g)
    depth = np.full(g.num_vertices, U32_MAX, np.uint32)
    for r in roots:
        d = o.bfs(int(r))[1]
        depth = np.minimum(depth, d)
    assert np.array_equal(W.bfs_depth(packed), depth)
    parent = min_frontier_parent(g, depth)
    par = W.bfs_parent(packed)
    nonroot = (depth != U32_MAX) & (depth != 0)
    assert np.array_equal(par[nonroot], parent[nonroot])


@pytest.mark.parametrize("shape", [(48, 80), (300, 200), (12, 3000)])
@pytest.mark.parametrize("sched", ["plain", "learned", "frontier"])
def test_grid_levels_fused_match(ctx, model, shape, sched):
    g = W.grid(*shape, seed=3)
    ctx.load(g)
    ctx.set_schedule_model(model if sched == "learned" else None)
    root = g.max_outdegree_vertex()
    (p0, it0, f0), (p1, it1, f1) = _pair(ctx, lambda: ctx.init_bfs(root), W.SCHEDULES[sched])
    assert f0 == 0
    if sched != "learned":  # the model may keep small graphs dense
        assert f1 > 0
    assert it0 == it1
    assert np.array_equal(p0, p1)
    _check_against_oracle(g, [root], p1)
    ctx.set_schedule_model(None)


@pytest.mark.parametrize("scale", [12, 14, 18])
def test_rmat_fused_match(ctx, model, scale):
    g = W.rmat(scale, 16, seed=2)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    for sched in (W.SCHED_PLAIN, W.SCHED_LEARNED, W.SCHED_FRONTIER):
        ctx.set_schedule_model(model if sched == W.SCHED_LEARNED else None)
        (p0, it0, f0), (p1, it1, f1) = _pair(ctx, lambda: ctx.init_bfs(root), sched)
        assert f0 == 0
        assert it0 == it1, sched
        assert np.array_equal(p0, p1), sched
    ctx.set_schedule_model(None)
    _check_against_oracle(g, [root], p1)


@pytest.mark.parametrize("wgs", [1, 8, 128])
def test_workgroup_counts(ctx, wgs):
    g = W.grid(120, 90, seed=5)
    ctx.load(g)
    roots = [0, 5000, 10799]
    (p0, it0, _), (p1, it1, f1) = _pair(ctx, lambda: ctx.init_bfs_roots(roots), W.SCHED_FRONTIER, wgs=wgs)
    assert f1 > 0
    assert it0 == it1
    assert np.array_equal(p0, p1)
    _check_against_oracle(g, roots, p1)


@pytest.mark.parametrize("cap", [1, 2, 7, 40])
def test_iteration_cap_inside_a_run(ctx, cap):
    g = W.grid(64, 64, seed=7)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    (p0, it0, _), (p1, it1, _) = _pair(ctx, lambda: ctx.init_bfs(root), W.SCHED_FRONTIER, max_iterations=cap)
    assert it0 == it1 == (cap, cap)
    assert np.array_equal(p0, p1)
    # the capped result is the full one cut at depth cap
    (q0, jt0, _), (q1, jt1, _) = _pair(ctx, lambda: ctx.init_bfs(root), W.SCHED_FRONTIER)
    assert jt0 == jt1 and np.array_equal(q0, q1)
    capped = W.bfs_depth(p1)
    full = W.bfs_depth(q1)
    reached = capped != U32_MAX
    assert np.array_equal(capped[reached], full[reached])
    assert int(capped[reached].max()) == cap


def test_member_union_fused():
    ms = [W.grid(30, 40, seed=1), W.grid(20, 90, seed=4), W.grid(5, 700, seed=2)]
    c = W.Context(0, relabel=False)
    first = c.load_batch(ms)
    roots = first[:-1] + np.array([m.max_outdegree_vertex() for m in ms], np.uint64)
    (p0, it0, _), (p1, it1, f1) = _pair(c, lambda: c.init_bfs_roots(roots), W.SCHED_PLAIN)
    assert f1 > 0 and it0 == it1
    assert np.array_equal(p0, p1)
    assert c.validate() == 0
    c.close()


def test_empty_frontier_and_isolated_root(ctx):
    # a root with no out-edges ends the loop at its first decision; one with a
    # single out-edge to a sink ends it inside a run
    g = W.from_edges(6, [0, 1, 2, 3], [1, 2, 3, 4])
    ctx.load(g)
    for root in (5, 4, 3, 0):
        (p0, it0, _), (p1, it1, _) = _pair(ctx, lambda: ctx.init_bfs(root), W.SCHED_FRONTIER)
        assert it0 == it1 and np.array_equal(p0, p1), root


def test_setter_bounds():
    c = W.Context(0)
    with pytest.raises(W.WedgeError):
        c.set_fused_levels(257)
    with pytest.raises(W.WedgeError):
        c.set_fused_levels(32, (W.APP_PR,))
    c.set_fused_levels(0)
    c.set_fused_levels(1, ALL_APPS)
    c.close()


def test_default_is_bfs_only():
    g = W.grid(40, 400, seed=2)
    c = W.Context(0)
    c.load(g)
    c.init_bfs(g.max_outdegree_vertex())
    c.converge(W.SCHED_FRONTIER)
    assert c.fused_level_count() > 0
    c.init_sssp(g.max_outdegree_vertex())
    c.converge(W.SCHED_FRONTIER)
    assert c.fused_level_count() == 0
    c.close()


# ---- CC and SSSP -------------------------------------------------------------------


ALL_APPS = (W.APP_BFS, W.APP_CC, W.APP_SSSP)


def _pair_values(ctx, init, read, sched, wgs=32, max_iterations=0):
    """(values, iteration counts, fused levels) with fused levels off, then on
    (for every application)."""
    out = []
    for n in (0, wgs):
        ctx.set_fused_levels(n, ALL_APPS)
        init()
        it = ctx.converge(sched, max_iterations=max_iterations)
        out.append((read(), it, ctx.fused_level_count()))
    ctx.set_fused_levels(0, (W.APP_CC, W.APP_SSSP))
    ctx.set_fused_levels(32)
    return out


def _oracle_sssp(g, roots):
    o = oracle.Oracle.of(g)
    ref = None
    for r in roots:
        d, _ = o.sssp(int(r))
        ref = d if ref is None else np.minimum(ref, d)
    if g.weight_kind == W.WEIGHT_U32:
        ref = np.where(ref == np.uint64(2**64 - 1), U32_MAX, ref).astype(np.uint32)
    return ref


MIN_GRAPHS = {
    "grid": lambda: W.grid(150, 220, seed=3),
    "grid_long": lambda: W.grid(8, 4000, seed=6),
    "rmat_w": lambda: W.rmat(13, 8, seed=4, weights="f64"),
    "rmat_sym": lambda: W.rmat(13, 8, seed=4, symmetrize=True),
}


@pytest.mark.parametrize("kind", list(MIN_GRAPHS))
@pytest.mark.parametrize("sched", ["plain", "learned", "frontier"])
def test_cc_and_sssp_fused_match(ctx, model, kind, sched):
    g = MIN_GRAPHS[kind]()
    ctx.load(g)
    ctx.set_schedule_model(model if sched == "learned" else None)
    s = W.SCHEDULES[sched]
    (c0, _, f0), (c1, _, f1) = _pair_values(ctx, ctx.init_cc, ctx.read_cc, s)
    assert f0 == 0
    assert np.array_equal(c0, c1)
    assert np.array_equal(c1, oracle.Oracle.of(g).cc()[0].astype(np.uint32))
    fused = f1
    if g.weights is not None:
        root = g.max_outdegree_vertex()
        (d0, _, f0), (d1, _, f1) = _pair_values(ctx, lambda: ctx.init_sssp(root), ctx.read_sssp, s)
        assert f0 == 0
        assert np.array_equal(d0, d1)
        assert np.array_equal(d1, _oracle_sssp(g, [root]))
        assert ctx.validate() == 0
        fused += f1
    if sched == "frontier":
        assert fused > 0
    ctx.set_schedule_model(None)


@pytest.mark.parametrize("wgs", [1, 5, 128, 256])
def test_sssp_several_roots_and_workgroup_counts(ctx, wgs):
    g = W.grid(90, 140, seed=9)
    ctx.load(g)
    roots = [3, 6000, 12599]
    (d0, _, _), (d1, _, f1) = _pair_values(ctx, lambda: ctx.init_sssp_roots(roots), ctx.read_sssp,
                                           W.SCHED_FRONTIER, wgs=wgs)
    assert f1 > 0
    assert np.array_equal(d0, d1)
    assert np.array_equal(d1, _oracle_sssp(g, roots))


def test_sssp_cap_inside_a_run(ctx):
    # an iteration cap ends the loop inside a run: the distances so far are
    # upper bounds of the final ones, and the next run from a fresh start
    # (new init) is not disturbed by the state the capped one left
    g = W.grid(60, 60, seed=11)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    ref = _oracle_sssp(g, [root])
    for cap in (1, 3, 17):
        ctx.set_fused_levels(32, ALL_APPS)
        ctx.init_sssp(root)
        it = ctx.converge(W.SCHED_FRONTIER, max_iterations=cap)
        assert it == (cap, cap)
        d = ctx.read_sssp()
        assert np.all(d >= ref) and d[root] == 0
        assert np.count_nonzero(d != U32_MAX) > 1
        ctx.init_sssp(root)
        ctx.converge(W.SCHED_FRONTIER)
        assert np.array_equal(ctx.read_sssp(), ref)
    ctx.set_fused_levels(0, (W.APP_CC, W.APP_SSSP))


def test_hub_overflows_the_queue(ctx):
    # root -> a -> hub -> 2.2M leaves: the level that reaches the hub appends
    # more entries than a queue holds; the run ends there and the loop goes
    # on with ordinary pulls
    n = (1 << 21) + 100000
    leaves = np.arange(3, n + 3, dtype=np.int64)
    src = np.concatenate([[0, 1], np.full(n, 2, np.int64)])
    dst = np.concatenate([[1, 2], leaves])
    g = W.from_edges(n + 3, src, dst)
    ctx.load(g)
    for sched in (W.SCHED_PLAIN, W.SCHED_FRONTIER):
        (p0, it0, _), (p1, it1, f1) = _pair(ctx, lambda: ctx.init_bfs(0), sched)
        assert f1 > 0 and it0 == it1
        assert np.array_equal(p0, p1)
    depth = W.bfs_depth(p1)
    assert depth[0] == 0 and depth[1] == 1 and depth[2] == 2 and np.all(depth[3:] == 3)
    assert np.all(W.bfs_parent(p1)[3:] == 2)
    (c0, _, _), (c1, _, _) = _pair_values(ctx, ctx.init_cc, ctx.read_cc, W.SCHED_FRONTIER)
    assert np.array_equal(c0, c1) and np.all(c1 == 0)


# ---- first level pushed along the roots' out-edges ---------------------------------


@pytest.mark.parametrize("nroots", [1, 3, 64, 300, 4096, 4200])
def test_first_level_push_matches_block_pull(ctx, nroots):
    """wg_converge pushes a frontier-filtered first level along the out-edges of
    up to 4096 roots instead of pulling over marked blocks; the host-driven
    loop always pulls.  Same depths and parents, both against the oracle."""
    g = W.rmat(13, 8, seed=21)
    ctx.load(g)
    rng = np.random.default_rng(nroots)
    roots = rng.choice(g.num_vertices, nroots, replace=False).astype(np.uint64)
    roots[0] = g.max_outdegree_vertex()
    ctx.init_bfs_roots(roots)
    it_dev = ctx.converge(W.SCHED_FRONTIER)
    dev = ctx.read_bfs()
    # host-driven: mark every root's blocks, then one pull per level
    ctx.init_bfs_roots(roots)
    for r in np.unique(roots):
        ctx.wedge_activate_single_vertex(int(r))
    levels = 0
    while True:
        act = ctx.edge_pull(W.APP_BFS, W.MODE_FRONTIER)
        levels += 1
        if act == 0:
            break
        ctx.wedge_generate_frontier()
        ctx.advance_frontier()
    host = ctx.read_bfs()
    assert np.array_equal(dev, host)
    assert it_dev[0] == levels
    _check_against_oracle(g, roots, dev)


def test_converge_after_a_host_pull_does_not_push(ctx):
    """After a host-driven first level the frontier is no longer the root: the
    device loop must pull its first level over the marked blocks."""
    g = W.rmat(12, 8, seed=22)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    ctx.init_bfs(root)
    ctx.converge(W.SCHED_FRONTIER)
    want = ctx.read_bfs()
    ctx.init_bfs(root)
    ctx.wedge_activate_single_vertex(root)
    ctx.edge_pull(W.APP_BFS, W.MODE_FRONTIER)
    ctx.wedge_generate_frontier()
    ctx.advance_frontier()
    ctx.converge(W.SCHED_FRONTIER)
    assert np.array_equal(ctx.read_bfs(), want)
