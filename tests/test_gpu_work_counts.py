"""Work counts of the traversal pulls (wg_work_counts, WG_OPT_COUNT_WORK) on
the MI355X: exact counts of dense BFS levels against a count derived from the
oracle's depths and the in-lists, the edges of dense CC / SSSP pulls, bounds
for the frontier and learned schedules, and results that do not depend on
whether the pulls count.

The counts have no reference counterpart; the dense BFS level they describe
is perform_edge_pull_phase_bfs (source/phase_edge_pull_bfs.c), a scan of each
unvisited destination's in-list that stops at the first source on the
frontier."""
import numpy as np
import pytest

import oracle
import wedge_amd as W

pytestmark = pytest.mark.gpu

BFS_SEQ = 16       # in-edges one lane scans alone (WG_BFS_SEQ)
BFS_HEAVY = 2048   # in-degree above which a vertex is split into chunks (WG_BFS_HEAVY)
WAVE = 64


def dense_bfs_work(g, root, depth):
    """(pulls, edges read, destinations examined) of a BFS that pulls densely
    at every level, in the caller's vertex order: each unvisited destination
    with 1..BFS_HEAVY in-edges reads its first in-neighbour, then the next ones
    alone up to BFS_SEQ, then the rest 64 at a time with its wavefront, and
    stops after the first source on the frontier."""
    V, E = g.num_vertices, g.num_edges
    off = g.offsets.astype(np.int64)
    src = g.sources.astype(np.int64)
    deg = np.diff(off)
    dst = g.destinations().astype(np.int64)
    pos = np.arange(E, dtype=np.int64) - off[dst]
    od = g.outdegrees().astype(np.int64)
    d = np.where(depth == np.uint32(0xFFFFFFFF), np.int64(1) << 40, depth.astype(np.int64))
    assert deg.max(initial=0) <= BFS_HEAVY
    if od[root] == 0:
        return 0, 0, 0
    big = np.int64(1) << 40
    pulls = edges = dests = 0
    level = 0
    while True:
        pulls += 1
        cand = (d > level) & (deg >= 1)
        hit = d[src] == level
        first = np.full(V, big, np.int64)
        np.minimum.at(first, dst[hit], pos[hit])
        i, n = first[cand], deg[cand]
        seq = np.minimum(n, BFS_SEQ)
        chunks = (np.maximum(i, BFS_SEQ) - BFS_SEQ) // WAVE + 1
        e = np.where(i < seq, i + 1,
                     np.where(n <= BFS_SEQ, n,
                              np.where(i < n, BFS_SEQ + np.minimum(n - BFS_SEQ, chunks * WAVE), n)))
        edges += int(e.sum())
        dests += int(cand.sum())
        if int(od[d == level + 1].sum()) == 0:
            break
        level += 1
    return pulls, edges, dests


def _converge_bfs(ctx, root, schedule, model=None):
    ctx.set_schedule_model(model)
    ctx.init_bfs(root)
    it = ctx.converge(schedule)
    return it, ctx.read_bfs()


def test_work_counts_need_the_option():
    g = W.rmat(10, 8, seed=1)
    c = W.Context(0)
    try:
        c.load(g)
        c.init_bfs(g.max_outdegree_vertex())
        c.converge(W.SCHED_PLAIN)
        with pytest.raises(RuntimeError):
            c.work_counts()
        c.set_option(W.OPT_COUNT_WORK, 1)
        c.init_bfs(g.max_outdegree_vertex())
        c.converge(W.SCHED_PLAIN)
        e, d = c.work_counts()
        assert 0 < d <= e
        with pytest.raises(RuntimeError):
            c.set_option(W.OPT_COUNT_WORK, 2)
    finally:
        c.close()


@pytest.mark.parametrize("scale,ef,seed", [(12, 8, 3), (14, 4, 5)])
def test_dense_bfs_work_counts_exact(scale, ef, seed):
    g = W.rmat(scale, ef, seed=seed)
    root = g.max_outdegree_vertex()
    _, depth, _ = oracle.Oracle.of(g).bfs(root)
    want = dense_bfs_work(g, root, depth)
    c = W.Context(0, relabel=False)
    try:
        c.load(g)
        c.set_option(W.OPT_COUNT_WORK, 1)
        (it, fit), _ = _converge_bfs(c, root, W.SCHED_DENSE)
        got = (it,) + c.work_counts()
        print(f"rmat{scale}: pulls, edges, dests {got} want {want}")
        assert fit == 0
        assert got == want
    finally:
        c.close()


def test_counting_changes_no_result_and_bounds_hold():
    """Every schedule: the same parents with and without counting, repeated
    counts equal on a graph without split vertices, destinations examined <=
    edges read, and at least one edge read per vertex reached."""
    c = W.Context(0)
    try:
        was = W.mlp_loaded()
        W.mlp_load()
        model = W.mlp_get()
        if not was:
            W.mlp_unload()
        for g in (W.rmat(14, 16, seed=2), W.grid(60, 90, seed=4)):
            c.load(g)
            root = g.max_outdegree_vertex()
            _, depth, _ = oracle.Oracle.of(g).bfs(root)
            reached = int((depth != np.uint32(0xFFFFFFFF)).sum())
            for sched, m in ((W.SCHED_PLAIN, None), (W.SCHED_LEARNED, model), (W.SCHED_FRONTIER, None)):
                c.set_option(W.OPT_COUNT_WORK, 0)
                it0, out0 = _converge_bfs(c, root, sched, m)
                c.set_option(W.OPT_COUNT_WORK, 1)
                it1, out1 = _converge_bfs(c, root, sched, m)
                e1, d1 = c.work_counts()
                it2, _ = _converge_bfs(c, root, sched, m)
                e2, d2 = c.work_counts()
                print(f"V={g.num_vertices} sched {sched}: {it1} edges {e1} dests {d1}")
                assert it0 == it1 == it2
                assert np.array_equal(out0, out1)
                assert d1 <= e1 and e1 >= reached - 1
                if g.num_edges < 100000:  # grid: no destination split over wavefronts
                    assert (e1, d1) == (e2, d2)
    finally:
        c.close()


def test_dense_cc_and_sssp_read_every_edge_per_pull():
    c = W.Context(0)
    try:
        c.set_option(W.OPT_COUNT_WORK, 1)
        gs = W.rmat(12, 8, seed=8, symmetrize=True)
        c.load(gs)
        c.init_cc()
        it, _ = c.converge(W.SCHED_DENSE)
        e, d = c.work_counts()
        assert e == it * gs.num_edges and d > 0
        gw = W.rmat(12, 8, seed=9, weights="f64")
        c.load(gw)
        c.init_sssp(gw.max_outdegree_vertex())
        it, _ = c.converge(W.SCHED_DENSE)
        e, d = c.work_counts()
        assert e == it * gw.num_edges and d > 0
    finally:
        c.close()
