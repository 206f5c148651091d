"""Batches of member graphs on the MI355X.  A batch is loaded as the disjoint
union of its members; every member must give what an independent oracle run on
that member gives: BFS depths, CC labels and SSSP distances bit-exact (shifted
by the member's first vertex), BFS parents a valid tree with the smallest
parent one level up, PageRank within 1e-10 L1 per member (FP64 contributions).
The captured HIP graph must reproduce plain launches bit for bit, replay after
replay with other work in between."""
import numpy as np
import pytest

import oracle
import wedge_amd as W

from helpers import U32_MAX, check_bfs_tree, min_frontier_parent

pytestmark = pytest.mark.gpu

PR_TOL = 1e-10  # L1 per member, FP64 contributions
PR_TOL_FP32 = 2e-7  # L1 per member, FP32-stored contributions (about 10x the measured error)


def _members(seed, n, weights=None, symmetrize=False):
    """Uneven small R-MAT members plus the degenerate ones: a single vertex, a
    two-vertex path, and a star whose centre has more in-edges than the largest
    tile (a heavy destination for every tile shape)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        scale = int(rng.integers(4, 11))
        ef = int(rng.integers(2, 17))
        out.append(W.rmat(scale, ef, seed=1000 * seed + i, symmetrize=symmetrize, weights=weights))
    wf = (lambda k: np.linspace(0.25, 1.0, k)) if weights == "f64" else (lambda k: None)
    out.insert(1, W.from_edges(1, [], [], weights=wf(0)))
    src = np.array([0, 1] if symmetrize else [0]), np.array([1, 0] if symmetrize else [1])
    out.insert(3, W.from_edges(2, src[0], src[1], weights=wf(src[0].size)))
    k = 3000
    s = np.arange(1, k + 1)
    d = np.zeros(k, np.int64)
    if symmetrize:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
    out.append(W.from_edges(k + 1, s, d, weights=wf(s.size)))
    return out


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def _roots(members, first):
    return np.array([int(first[g]) + m.max_outdegree_vertex() for g, m in enumerate(members)], np.uint64)


# ---- PageRank ---------------------------------------------------------------------


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_batch_pr_matches_member_oracle(ctx, fp32):
    ms = _members(1, 14)
    first = ctx.load_batch(ms)
    refs = [oracle.Oracle.of(m).pr(16, 0.85)[0] for m in ms]
    tol = PR_TOL_FP32 if fp32 else PR_TOL
    for shape in (None, (4, 2), (8, 8), (16, 16), (16, 1), (4, 4, 1), (8, 16, 1), (16, 2, 1)):
        for use_graph in (False, True):
            ctx.batch_pr(16, 0.85, fp32, shape, use_graph)
            r = ctx.read_batch_ranks()
            worst = 0.0
            for g, ref in enumerate(refs):
                l1 = float(np.abs(r[first[g]:first[g + 1]] - ref).sum())
                worst = max(worst, l1)
                assert l1 <= tol, f"member {g} shape {shape} graph {use_graph}: L1 {l1:.3e}"
            print(f"batch PR fp32={fp32} shape {shape} graph {use_graph}: worst member L1 {worst:.3e}")


def test_batch_pr_graph_replays_equal_plain_launches(ctx):
    ms = _members(2, 10)
    first = ctx.load_batch(ms)
    ctx.batch_pr(16, 0.85, False, (8, 4), use_graph=False)
    plain = ctx.read_batch_ranks()
    roots = _roots(ms, first)
    for rep in range(4):
        ctx.batch_pr(16, 0.85, False, (8, 4), use_graph=True)  # the kept graph after the first
        got = ctx.read_batch_ranks()
        assert np.array_equal(got, plain), f"replay {rep}"
        ctx.batch_pr(16, 0.85, False, (8, 4, 1), use_graph=True)  # the XCD split sums in the same order
        assert np.array_equal(ctx.read_batch_ranks(), plain), f"split {rep}"
        # other work between replays: plain launches of another run and a traversal
        ctx.batch_pr(5, 0.5, True, (16, 2), use_graph=False)
        ctx.sync()
        ctx.init_bfs_roots(roots)
        ctx.converge(W.SCHED_PLAIN)
    # a different iteration count recaptures
    ctx.batch_pr(3, 0.85, False, (8, 4), use_graph=True)
    r3 = ctx.read_batch_ranks()
    ctx.batch_pr(3, 0.85, False, (8, 4), use_graph=False)
    assert np.array_equal(r3, ctx.read_batch_ranks())


def test_batch_autotune_picks_a_candidate(ctx):
    ms = _members(3, 12)
    first = ctx.load_batch(ms)
    best, times = ctx.batch_autotune(4, 0.85)
    assert best in times and len(times) == W.BATCH_NSHAPES
    assert all(t > 0 for t in times.values())
    assert times[best] == min(times.values())
    print("autotune:", best, {k: round(v, 4) for k, v in times.items()})
    ctx.batch_pr(16, 0.85)  # the tuned shape
    r = ctx.read_batch_ranks()
    for g, m in enumerate(ms):
        ref, _ = oracle.Oracle.of(m).pr(16, 0.85)
        assert float(np.abs(r[first[g]:first[g + 1]] - ref).sum()) <= PR_TOL


# ---- traversals -------------------------------------------------------------------


SCHEDULES = [dict(schedule=W.SCHED_PLAIN), dict(schedule=W.SCHED_DENSE),
             dict(schedule=W.SCHED_FRONTIER, block_log2=3), dict(schedule=W.SCHED_LEARNED)]


def _set_model(ctx):
    W.mlp_load(W.DEFAULT_MODEL_PATH)
    ctx.set_schedule_model(W.mlp_get())


def test_batch_bfs_one_root_per_member(ctx):
    ms = _members(4, 16)
    first = ctx.load_batch(ms)
    _set_model(ctx)
    roots = _roots(ms, first)
    refs = [oracle.Oracle.of(m).bfs(int(roots[g] - first[g]))[1] for g, m in enumerate(ms)]
    for opts in SCHEDULES:
        ctx.init_bfs_roots(roots)
        it, fit = ctx.converge(**opts)
        packed = ctx.read_bfs()
        depth, parent = W.bfs_depth(packed), W.bfs_parent(packed)
        for g, m in enumerate(ms):
            a, b = int(first[g]), int(first[g + 1])
            assert np.array_equal(depth[a:b], refs[g]), f"member {g} {opts}"
            p = parent[a:b].astype(np.int64)
            local = np.where(p == U32_MAX, U32_MAX, p - a).astype(np.uint32)
            r = int(roots[g]) - a
            assert check_bfs_tree(m, r, refs[g], local) == 0, f"member {g} {opts}"
            want = min_frontier_parent(m, refs[g])
            ok = (refs[g] != U32_MAX) & (np.arange(m.num_vertices) != r)
            assert np.array_equal(local[ok], want[ok]), f"member {g} {opts}: parent rule"
            assert local[r] == r
        print(f"batch BFS {opts}: {it} levels, {fit} frontier")


def test_batch_cc(ctx):
    ms = _members(5, 16, symmetrize=True)
    first = ctx.load_batch(ms)
    refs = [oracle.Oracle.of(m).cc()[0] for m in ms]
    for opts in SCHEDULES[:3]:
        ctx.init_cc()
        ctx.converge(**opts)
        lab = ctx.read_cc().astype(np.float64)
        for g in range(len(ms)):
            a, b = int(first[g]), int(first[g + 1])
            assert np.array_equal(lab[a:b] - a, refs[g]), f"member {g} {opts}"


def test_batch_sssp_one_root_per_member(ctx):
    ms = _members(6, 12, weights="f64")
    first = ctx.load_batch(ms)
    _set_model(ctx)
    roots = _roots(ms, first)
    refs = [oracle.Oracle.of(m).sssp(int(roots[g] - first[g]))[0] for g, m in enumerate(ms)]
    for opts in SCHEDULES:
        ctx.init_sssp_roots(roots)
        ctx.converge(**opts)
        dist = ctx.read_sssp()
        for g in range(len(ms)):
            a, b = int(first[g]), int(first[g + 1])
            assert np.array_equal(dist[a:b], refs[g]), f"member {g} {opts}"


def test_multi_root_bfs_equals_min_over_roots(ctx):
    """Several roots in one connected graph: depth = min over the roots' depths."""
    g = W.rmat(11, 8, seed=9, symmetrize=True)
    ctx.set_option(W.OPT_RELABEL, W.RELABEL_AUTO)
    ctx.load(g)
    roots = np.array([g.max_outdegree_vertex(), 5, 77, 5], np.uint64)
    o = oracle.Oracle.of(g)
    want = np.min(np.stack([o.bfs(int(r))[1] for r in roots]), axis=0)
    ctx.init_bfs_roots(roots)
    ctx.converge(W.SCHED_PLAIN)
    assert np.array_equal(W.bfs_depth(ctx.read_bfs()), want)


# ---- errors -------------------------------------------------------------------------


def test_member_errors(ctx):
    ms = _members(7, 3)
    u, first = W.union(ms)
    ctx.set_option(W.OPT_RELABEL, W.RELABEL_OFF)
    ctx.load(u)
    with pytest.raises(W.WedgeError) as e:
        ctx.batch_pr(4)
    assert e.value.code == W.WG_E_STATE  # no members yet
    bad = first.copy()
    bad[-1] -= 1
    with pytest.raises(W.WedgeError) as e:
        ctx.set_members(bad)
    assert e.value.code == W.WG_E_INVAL
    with pytest.raises(W.WedgeError) as e:
        ctx.set_members(np.array([0, 0, u.num_vertices], np.uint64))  # an empty member
    assert e.value.code == W.WG_E_INVAL
    # an edge from member 1 into member 0
    src = np.concatenate([u.sources, [first[1]]]).astype(np.uint64)
    dst = np.concatenate([u.destinations(), [0]]).astype(np.uint64)
    ctx.load(W.from_edges(u.num_vertices, src, dst))
    with pytest.raises(W.WedgeError) as e:
        ctx.set_members(first)
    assert e.value.code == W.WG_E_FORMAT
    ctx.load(u)
    ctx.set_members(first)
    with pytest.raises(W.WedgeError) as e:
        ctx.read_batch_ranks()  # nothing ran
    assert e.value.code == W.WG_E_STATE
    for shape in ((5, 4), (8, 0), (8, 17), (8, 4, 2)):
        with pytest.raises(W.WedgeError) as e:
            ctx.batch_pr(4, 0.85, False, shape)
        assert e.value.code == W.WG_E_INVAL
    # relabelled loads interleave members
    ctx.set_option(W.OPT_RELABEL, W.RELABEL_ON)
    ctx.load(u)
    with pytest.raises(W.WedgeError) as e:
        ctx.set_members(first)
    assert e.value.code == W.WG_E_STATE
    ctx.set_option(W.OPT_RELABEL, W.RELABEL_AUTO)
