"""On-device validation (wg_validate, wg_validate_values) on the MI355X.

Every engine result must validate with 0 violations; results computed by the
oracle, given in the caller's numbering, must too.  Mistakes planted in a host
copy of a correct result (a depth off by one, a parent that is not an
in-neighbour, a distance nudged down or up, a label raised or lowered) must be
reported.  The planted values are ordinary data: the validator reads them like
any other result."""
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


def _reached_non_root(depth, root):
    idx = np.nonzero((depth != U32_MAX) & (np.arange(depth.size) != root))[0]
    return idx


def test_bfs_results_validate_and_mistakes_are_caught(ctx):
    g = W.rmat(12, 16, seed=2)
    ctx.load(g)
    root = g.max_outdegree_vertex()
    for sched in ("plain", "dense", "learned"):
        if sched == "learned":
            W.mlp_load()
        packed, _ = ctx.run_bfs(root, schedule=sched)
        assert ctx.validate() == 0, sched
    assert ctx.validate_values(W.APP_BFS, packed, [root]) == 0
    # the oracle's depths with the smallest-parent rule, packed the same way
    _, depth, _ = oracle.Oracle.of(g).bfs(root)
    parent = min_frontier_parent(g, depth)
    parent[root] = root
    ref = np.where(depth == U32_MAX, W.UNREACHED_U64,
                   (depth.astype(np.uint64) << np.uint64(32)) | parent.astype(np.uint64))
    assert ctx.validate_values(W.APP_BFS, ref, [root]) == 0
    rng = np.random.default_rng(1)
    v = int(rng.choice(_reached_non_root(depth, root)))
    bad = packed.copy()
    bad[v] += np.uint64(1 << 32)  # one level too deep
    assert ctx.validate_values(W.APP_BFS, bad, [root]) > 0
    bad = packed.copy()
    non_nb = int(np.setdiff1d(np.arange(g.num_vertices), g.sources[g.offsets[v]:g.offsets[v + 1]])[0])
    bad[v] = (bad[v] & ~np.uint64(0xFFFFFFFF)) | np.uint64(non_nb)  # parent is not an in-neighbour
    assert ctx.validate_values(W.APP_BFS, bad, [root]) > 0
    bad = packed.copy()
    unreached = np.nonzero(depth == U32_MAX)[0]
    if unreached.size:
        bad[int(unreached[0])] = np.uint64(5 << 32) | np.uint64(root)  # an unreachable vertex claimed
        assert ctx.validate_values(W.APP_BFS, bad, [root]) > 0
    assert ctx.validate_values(W.APP_BFS, packed, [root, 7]) > 0  # 7 is no root of this result


def test_sssp_results_validate_and_mistakes_are_caught(ctx):
    for g in (W.rmat(12, 16, seed=2, weights="f64"), W.grid(60, 90, seed=3)):
        ctx.load(g)
        root = g.max_outdegree_vertex()
        dist, _ = ctx.run_sssp(root)
        assert ctx.validate() == 0
        ref, _ = oracle.Oracle.of(g).sssp(root)
        if g.weight_kind == W.WEIGHT_U32:
            ref = np.where(ref == np.uint64(2**64 - 1), U32_MAX, ref).astype(np.uint32)
        assert ctx.validate_values(W.APP_SSSP, ref, [root]) == 0
        fin = np.nonzero((dist != (np.inf if dist.dtype == np.float64 else U32_MAX)) &
                         (np.arange(g.num_vertices) != root))[0]
        v = int(fin[len(fin) // 2])
        bad = dist.copy()
        bad[v] = np.nextafter(bad[v], 0) if bad.dtype == np.float64 else bad[v] - 1  # too short: no tight edge
        assert ctx.validate_values(W.APP_SSSP, bad, [root]) > 0
        bad = dist.copy()
        bad[v] = bad[v] * 2 + (1 if bad.dtype == np.uint32 else 0.5)  # too long: an edge into v relaxes it
        assert ctx.validate_values(W.APP_SSSP, bad, [root]) > 0


def test_cc_results_validate_and_mistakes_are_caught(ctx):
    for g in (W.rmat(11, 8, seed=4, symmetrize=True), W.rmat(11, 8, seed=4)):
        ctx.load(g)
        labels, _ = ctx.run_cc()
        assert ctx.validate() == 0
        ref, _ = oracle.Oracle.of(g).cc()
        assert ctx.validate_values(W.APP_CC, ref.astype(np.uint32), []) == 0
        v = int(np.nonzero(labels < np.arange(g.num_vertices))[0][0])
        bad = labels.copy()
        bad[v] = v  # a label above the minimum reaching v
        assert ctx.validate_values(W.APP_CC, bad, []) > 0
        bad = labels.copy()
        w = int(np.nonzero(labels == np.arange(g.num_vertices))[0][-1])
        bad[w] = max(0, int(labels[w]) - 1) if labels[w] > 0 else 0  # a label nothing justifies
        if bad[w] != labels[w]:
            assert ctx.validate_values(W.APP_CC, bad, []) > 0


def test_validate_batched_union():
    ms = [W.rmat(s, 8, seed=s) for s in (6, 8, 9, 7)]
    c = W.Context(0)
    first = c.load_batch(ms)
    roots = first[:-1] + np.array([m.max_outdegree_vertex() for m in ms], np.uint64)
    c.init_bfs_roots(roots)
    c.converge(W.SCHED_PLAIN)
    assert c.validate() == 0
    c.init_cc()
    c.converge(W.SCHED_PLAIN)
    assert c.validate() == 0
    with pytest.raises(W.WedgeError):
        c.validate_values(W.APP_PR, np.zeros(c.V), [])
    c.close()


# ---- graphs built on the device from unsorted edge lists ---------------------------


def _shuffled_edges(g, rng, dup_frac=0.1, loops=50):
    """g's edges in random order, with repeats and self-loops added."""
    src, dst = g.sources.astype(np.int64), g.destinations().astype(np.int64)
    w = g.weights
    k = int(dup_frac * src.size)
    pick = rng.integers(0, src.size, k)
    lv = rng.integers(0, g.num_vertices, loops)
    s = np.concatenate([src, src[pick], lv])
    d = np.concatenate([dst, dst[pick], lv])
    ww = None
    if w is not None:
        extra = np.full(k + loops, 7.0 if w.dtype == np.float64 else 7, w.dtype)  # repeats carry other weights
        ww = np.concatenate([w, extra])
    # originals in random order first, then the repeats and loops in random order,
    # so that dedup keeps each original's weight
    order = np.concatenate([rng.permutation(src.size), src.size + rng.permutation(k + loops)])
    return s[order], d[order], (ww[order] if ww is not None else None)


@pytest.mark.parametrize("kind", ["rmat", "rmat_f64", "grid"])
def test_load_edges_builds_the_same_graph(kind):
    rng = np.random.default_rng(7)
    g = {"rmat": lambda: W.rmat(11, 8, seed=3), "rmat_f64": lambda: W.rmat(11, 8, seed=3, weights="f64"),
         "grid": lambda: W.grid(40, 50, seed=3)}[kind]()
    s, d, w = _shuffled_edges(g, rng)
    c = W.Context(0, relabel=False)
    kept = c.load_edges(g.num_vertices, s, d, w)
    assert kept == g.num_edges
    assert np.array_equal(c.outdegrees(), g.outdegrees())
    root = g.max_outdegree_vertex()
    o = oracle.Oracle.of(g)
    packed, _ = c.run_bfs(root)
    assert np.array_equal(W.bfs_depth(packed), o.bfs(root)[1])
    if g.weights is not None:
        dist, _ = c.run_sssp(root)
        ref, _ = o.sssp(root)
        if g.weight_kind == W.WEIGHT_U32:
            ref = np.where(ref == np.uint64(2**64 - 1), U32_MAX, ref).astype(np.uint32)
        assert np.array_equal(dist, ref)  # repeats kept their first weight, not the 7s
    ranks, _ = c.run_pr(iterations=8, damping=0.85)
    assert float(np.abs(ranks - o.pr(8, 0.85)[0]).sum()) <= 1e-10
    # keeping repeats and loops: every input edge is loaded
    assert c.load_edges(g.num_vertices, s, d, w, dedup=False, drop_self_loops=False) == s.size
    assert int(c.outdegrees().sum()) == s.size
    c.close()


def test_load_edges_errors_and_empty():
    c = W.Context(0)
    with pytest.raises(W.WedgeError) as e:
        c.load_edges(10, [1, 2], [3, 10])  # endpoint out of range
    assert e.value.code == W.WG_E_RANGE
    assert c.load_edges(5, [], []) == 0
    assert c.load_edges(3, [0, 0, 1], [0, 0, 0]) == 1  # a loop, then a repeat: only 1 -> 0 stays
    c.run_cc()
    assert c.validate() == 0
    c.close()
