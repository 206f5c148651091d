"""Pins the oracle (the CPU restatement of Wedge's pull engine) before it is
trusted as the checker of the HIP path.

The reference ships no tests or vectors, so every oracle function is checked
against (a) the committed golden vectors (scipy shortest paths / components /
Dijkstra and a sparse mat-vec PageRank loop, tests/golden/make_golden.py),
(b) hand-computed known answers on tiny graphs, (c) invariants the
mathematics fixes, and (d) planted mistakes that each check must catch.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
import wedge_amd as W

from helpers import (U32_MAX, bfs_depths_scipy, cc_min_labels_undirected, cc_min_reaching, check_bfs_tree,
                     pagerank_closed_form, sssp_dijkstra)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_small.npz")


def _golden():
    z = np.load(GOLDEN, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in z.files})
    out = {}
    for n in names:
        w = z[f"{n}__weights"] if f"{n}__weights" in z.files else None
        out[n] = (W.Graph(z[f"{n}__offsets"], z[f"{n}__sources"], w), z)
    return out


GOLD = _golden()


# ---- golden vectors ------------------------------------------------------------------


@pytest.mark.parametrize("name", sorted(GOLD))
def test_oracle_matches_golden(name):
    g, z = GOLD[name]
    orc = oracle.Oracle.of(g, threads=4)
    root = int(z[f"{name}__root"][0])
    for d, key in ((0.85, "pr_d085_n16"), (0.875, "pr_d0875_n16")):
        r, st = orc.pr(16, d)
        assert np.abs(r - z[f"{name}__{key}"]).sum() <= 1e-12
        assert abs(st["special"] - 1.0) < 1e-12
    parent, depth, _ = orc.bfs(root)
    assert np.array_equal(depth, z[f"{name}__bfs_depth"])
    assert check_bfs_tree(g, root, depth, parent) == 0
    if f"{name}__cc" in z.files:
        lab, _ = orc.cc()
        assert np.array_equal(lab.astype(np.uint32), z[f"{name}__cc"])
    ppr = orc.ppr(z[f"{name}__ppr_seeds"], 16, 0.85)
    assert np.abs(ppr - z[f"{name}__ppr_d085_n16"]).sum() <= 1e-13
    assert np.allclose(ppr.sum(axis=1), 1.0, rtol=0, atol=1e-12)  # mass is conserved
    if f"{name}__sssp" in z.files:
        dist, _ = orc.sssp(root)
        assert np.array_equal(dist, z[f"{name}__sssp"])  # bit-exact, FP64 included


def test_golden_fixture_is_reproducible():
    """The generators are deterministic: regenerating a golden graph gives the
    committed arrays."""
    g, z = GOLD["rmat10_s1"]
    h = W.rmat(10, 16, seed=1)
    assert np.array_equal(h.offsets, g.offsets) and np.array_equal(h.sources, g.sources)


# ---- hand-computed known answers ---------------------------------------------------------


def _tiny(V, edges, weights=None):
    src = [e[0] for e in edges]
    dst = [e[1] for e in edges]
    return W.from_edges(V, src, dst, weights=weights)


def test_pr_known_answers():
    # directed cycle: every rank stays exactly 1/n (each vertex gives all to one)
    n = 5
    g = _tiny(n, [(i, (i + 1) % n) for i in range(n)])
    r, _ = oracle.Oracle.of(g, 2).pr(7, 0.85)
    assert np.allclose(r, 1.0 / n, rtol=0, atol=1e-15)
    # no edges: every vertex is a sink, ranks stay 1/N
    g = _tiny(4, [])
    r, _ = oracle.Oracle.of(g, 2).pr(3, 0.85)
    assert np.allclose(r, 0.25, rtol=0, atol=1e-16)
    # 0 -> 1, 1 a sink: one iteration by hand with exact fractions.
    # contrib0 = (1/2, 1/4) (od' of the sink is N = 2); acc = (0, 1/2); S = 1/2
    # rank = (acc + (1 - S)/N) * d + (1 - d)/N
    d = Fraction(85, 100)
    acc = [Fraction(0), Fraction(1, 2)]
    S = sum(acc)
    want = [(a + (1 - S) / 2) * d + (1 - d) / 2 for a in acc]
    g = _tiny(2, [(0, 1)])
    r, _ = oracle.Oracle.of(g, 1).pr(1, 0.85)
    assert np.allclose(r, [float(x) for x in want], rtol=0, atol=1e-15)
    assert float(sum(want)) == pytest.approx(1.0)


def test_bfs_cc_sssp_known_answers():
    # path 0-1-2-3 plus an isolated vertex 4 (undirected)
    und = [(0, 1), (1, 0), (1, 2), (2, 1), (2, 3), (3, 2)]
    g = _tiny(5, und)
    parent, depth, _ = oracle.Oracle.of(g, 2).bfs(0)
    assert depth.tolist() == [0, 1, 2, 3, int(U32_MAX)]
    assert parent[1] == 0 and parent[2] == 1 and parent[3] == 2 and parent[4] == np.uint64(2**64 - 1)
    # two triangles -> labels are the component minima
    tri = [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3)]
    g = _tiny(6, tri + [(b, a) for a, b in tri])
    lab, _ = oracle.Oracle.of(g, 2).cc()
    assert lab.tolist() == [0, 0, 0, 3, 3, 3]
    # weighted: 0->1 (4), 0->2 (1), 2->1 (2), 1->3 (1), 2->3 (5): dist = 0, 3, 1, 4
    e = [(0, 1), (0, 2), (2, 1), (1, 3), (2, 3)]
    w = [4, 1, 2, 1, 5]
    g = _tiny(4, e, weights=np.array(w, np.uint32))
    dist, _ = oracle.Oracle.of(g, 2).sssp(0)
    assert dist.tolist() == [0, 3, 1, 4]
    gf = _tiny(4, e, weights=np.array(w, np.float64) / 8.0)
    dist, _ = oracle.Oracle.of(gf, 2).sssp(0)
    assert dist.tolist() == [0.0, 3 / 8, 1 / 8, 4 / 8]


def test_ppr_known_answers():
    # 2-cycle, seed 0, d = 1/2: r1 = (1/2, 1/2), r2 = (3/4, 1/4)
    g = _tiny(2, [(0, 1), (1, 0)])
    orc = oracle.Oracle.of(g, 1)
    assert np.allclose(orc.ppr([0], 1, 0.5)[0], [0.5, 0.5], rtol=0, atol=1e-16)
    assert np.allclose(orc.ppr([0], 2, 0.5)[0], [0.75, 0.25], rtol=0, atol=1e-16)
    # 0 -> 1 with 1 a sink: its mass returns to the seed
    d = Fraction(85, 100)
    r1 = [1 - d, d]
    r2 = [d * d + 1 - d, d * (1 - d)]
    g = _tiny(2, [(0, 1)])
    orc = oracle.Oracle.of(g, 1)
    assert np.allclose(orc.ppr([0], 1, 0.85)[0], [float(x) for x in r1], rtol=0, atol=1e-15)
    assert np.allclose(orc.ppr([0], 2, 0.85)[0], [float(x) for x in r2], rtol=0, atol=1e-15)
    # no damping: all mass stays on the seed
    g = W.rmat(8, 8, seed=3)
    r = oracle.Oracle.of(g, 1).ppr([5, 9], 4, 0.0)
    assert r[0, 5] == 1.0 and r[1, 9] == 1.0 and r.sum() == 2.0


def test_planted_mistake_ppr_dangling_mass():
    g = W.rmat(10, 8, seed=1)
    seeds = [g.max_outdegree_vertex(), 3]
    want = oracle.Oracle.of(g, 2).ppr(seeds, 12, 0.85)
    V = g.num_vertices
    od = g.outdegrees().astype(np.float64)
    from helpers import adjacency
    At = adjacency(g).T.tocsr()
    # wrong: dangling mass spread uniformly instead of returned to the seed
    R = np.zeros((V, 2))
    R[seeds, [0, 1]] = 1.0
    E = R.copy()
    for _ in range(12):
        C = np.where(od[:, None] > 0, R / np.where(od > 0, od, 1.0)[:, None], 0.0)
        D = R[od == 0].sum(axis=0)
        R = 0.85 * (At @ C + D / V) + 0.15 * E
    assert np.abs(R.T - want).sum() > 1e-6


def test_cc_directed_is_min_reaching_label():
    """On a directed input the reference computes min({u : u reaches v} | {v})."""
    g = W.rmat(7, 4, seed=9)
    lab, _ = oracle.Oracle.of(g, 2).cc()
    assert np.array_equal(lab.astype(np.uint32), cc_min_reaching(g))


# ---- invariants and independent references at larger sizes ------------------------------


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_oracle_results_do_not_depend_on_threads(threads):
    g = W.rmat(13, 16, seed=4, symmetrize=True)
    orc = oracle.Oracle.of(g, threads)
    root = g.max_outdegree_vertex()
    _, depth, _ = orc.bfs(root)
    assert np.array_equal(depth, bfs_depths_scipy(g, root))
    lab, _ = orc.cc()
    assert np.array_equal(lab.astype(np.uint32), cc_min_labels_undirected(g))
    r, _ = orc.pr(10, 0.85)
    assert np.abs(r - pagerank_closed_form(g, 10, 0.85)).sum() <= 1e-12


@pytest.mark.parametrize("threshold", [0, 2, 20, 100])
def test_oracle_schedule_threshold_does_not_change_results(threshold):
    """Dense and Wedge-frontier pulls must agree; the threshold only picks between them."""
    g = W.rmat(12, 16, seed=2, weights="f64")
    orc = oracle.Oracle.of(g, 4)
    root = g.max_outdegree_vertex()
    _, depth, st = orc.bfs(root, threshold)
    assert np.array_equal(depth, bfs_depths_scipy(g, root))
    dist, _ = orc.sssp(root, threshold)
    assert np.array_equal(dist, sssp_dijkstra(g, root))


def test_oracle_grid_integer_sssp():
    g = W.grid(40, 60, seed=3)
    orc = oracle.Oracle.of(g, 4)
    root = g.max_outdegree_vertex()
    dist, st = orc.sssp(root)
    d = sssp_dijkstra(g, root)
    fin = ~np.isinf(d)
    assert np.array_equal(dist[fin], d[fin].astype(np.uint64))
    assert np.all(dist[~fin] == np.uint64(2**64 - 1))
    assert st["frontier_iterations"] > 0  # the Wedge frontier path was exercised


def test_oracle_outdegree_and_vector_count():
    g = W.rmat(11, 16, seed=1)
    orc = oracle.Oracle.of(g, 2)
    assert np.array_equal(orc.outdegrees(), g.outdegrees().astype(np.uint64))
    indeg = np.diff(g.offsets.astype(np.int64))
    assert orc.num_vectors == int(((indeg + 3) // 4).sum())  # ceil(indeg / 4) per destination


# ---- planted mistakes: each check must fail on a wrong answer ----------------------------


def test_planted_mistake_pagerank_damping():
    g = W.rmat(10, 16, seed=1)
    ref = pagerank_closed_form(g, 16, 0.85)
    r, _ = oracle.Oracle.of(g, 2).pr(16, 0.875)  # wrong damping
    assert np.abs(r - ref).sum() > 1e-10


def test_planted_mistake_pagerank_dropped_edge():
    g = W.rmat(10, 16, seed=1)
    ref = pagerank_closed_form(g, 16, 0.85)
    src = g.sources.copy()
    bad = W.Graph(g.offsets.copy(), src, None)
    bad.sources[5] = (int(bad.sources[5]) + 1) % g.num_vertices  # one edge re-pointed
    # keep the destination list ascending for the oracle's input contract
    v = int(np.searchsorted(g.offsets, 5, side="right") - 1)
    seg = bad.sources[g.offsets[v]:g.offsets[v + 1]]
    seg.sort()
    r, _ = oracle.Oracle.of(bad, 2).pr(16, 0.85)
    assert np.abs(r - ref).sum() > 1e-10


def test_planted_mistake_bfs_parent_check():
    g = W.rmat(10, 16, seed=1)
    root = g.max_outdegree_vertex()
    parent, depth, _ = oracle.Oracle.of(g, 2).bfs(root)
    p = parent.astype(np.int64)
    v = int(np.nonzero((depth == 2))[0][0])
    p[v] = root  # the root is an in-neighbour only at depth 1: a wrong parent
    assert check_bfs_tree(g, root, depth, p) >= 1
    d2 = depth.copy()
    d2[v] += 1
    assert not np.array_equal(d2, bfs_depths_scipy(g, root))


def test_planted_mistake_sssp_weight():
    g = W.rmat(10, 16, seed=2, weights="f64")
    root = g.max_outdegree_vertex()
    ref = sssp_dijkstra(g, root)
    w = g.weights.copy()
    first = int(g.offsets[int(g.destinations()[0])])
    w[:] = w * (1.0 + 2.0 ** -40)  # every weight perturbed in the last bits
    bad = W.Graph(g.offsets, g.sources, w)
    dist, _ = oracle.Oracle.of(bad, 2).sssp(root)
    assert not np.array_equal(dist, ref)
    assert first >= 0


def test_planted_mistake_cc_missing_symmetry():
    g = W.rmat(10, 8, seed=3, symmetrize=True)
    want = cc_min_labels_undirected(g)
    directed = W.rmat(10, 8, seed=3)  # same draws without the reverse edges
    lab, _ = oracle.Oracle.of(directed, 2).cc()
    assert not np.array_equal(lab.astype(np.uint32), want)
