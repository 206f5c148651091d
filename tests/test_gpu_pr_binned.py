"""The binned PageRank pull (shape {8, 1|2, WG_PR_BINNED}) against the oracle
on the MI355X: FP64 ranks within 1e-10 L1, FP32-stored contributions within
2e-7 L1, fused iterations and the two-phase API, relabelled and native vertex
order, graphs with hubs, empty bins and windows, and bins split over many pull
workgroups (a tiny WG_OPT_PR_BIN_SLICE).  The layout's pair count is checked
against a count of distinct (source window, destination) pairs.

Reference: perform_edge_pull_phase_pr (source/phase_edge_pull_pr.c:28-116) and
perform_vertex_phase_pr (source/phase_vertex_pr.c:26-77)."""
import numpy as np
import pytest

import oracle
import wedge_amd as W

pytestmark = pytest.mark.gpu

PR_TOL = 1e-10
PR_TOL_FP32 = 2e-7
BINNED = [(8, 1, W.PR_BINNED), (8, 2, W.PR_BINNED)]


def _hub_graph(V=6000, seed=7):
    rng = np.random.default_rng(seed)
    src = [rng.integers(1, V, 5000), np.full(3000, 5, np.int64), rng.integers(0, V, 20000)]
    dst = [np.zeros(5000, np.int64), rng.integers(0, V, 3000), rng.integers(0, V, 20000)]
    return W.from_edges(V, np.concatenate(src), np.concatenate(dst), dedup=True)


def _graphs():
    return [("rmat12", W.rmat(12, 16, seed=1)), ("rmat16", W.rmat(16, 8, seed=2)), ("hub", _hub_graph()),
            ("grid", W.grid(40, 70, seed=4)), ("sparse", W.rmat(15, 1, seed=9))]


@pytest.fixture(scope="module", params=[True, False], ids=["relabel", "native-order"])
def ctx(request):
    c = W.Context(0, relabel=request.param)
    yield c
    c.close()


@pytest.mark.parametrize("shape", BINNED, ids=["128k", "64k"])
def test_binned_matches_oracle(ctx, shape):
    for name, g in _graphs():
        ctx.load(g)
        ctx.set_pr_shape(shape)
        assert ctx.pr_shape() == shape
        ref, _ = oracle.Oracle.of(g, 8).pr(16, 0.85)
        ranks, _ = ctx.run_pr(iterations=16, damping=0.85)
        l1 = float(np.abs(ranks - ref).sum())
        r32, _ = ctx.run_pr(iterations=16, damping=0.85, contrib_fp32=1)
        l1_32 = float(np.abs(r32 - ref).sum())
        print(f"{name} {shape}: L1 FP64 {l1:.3e}, FP32-stored {l1_32:.3e}")
        assert l1 <= PR_TOL, name
        assert l1_32 <= PR_TOL_FP32, name
        # the reference's two phases, interleaved with fused iterations
        ref6, _ = oracle.Oracle.of(g, 8).pr(6, 0.875)
        ctx.init_pr(0.875)
        for it in range(6):
            if it % 2:
                ctx.pr_iteration()
            else:
                total = ctx.edge_pull_pr()
                assert 0.0 < total <= 1.0 + 1e-12
                ctx.vertex_pr()
        assert float(np.abs(ctx.read_ranks() - ref6).sum()) <= PR_TOL, name


def test_binned_split_bins_match_oracle(ctx):
    """Bins cut into many pull workgroups (at most 37 sums each): the per-bin
    tickets and the slot sums."""
    ctx.set_option(W.OPT_PR_BIN_SLICE, 37)
    try:
        for name, g in _graphs()[:3]:
            ctx.load(g)
            for shape in BINNED:
                ctx.set_pr_shape(shape)
                ref, _ = oracle.Oracle.of(g, 8).pr(10, 0.85)
                ranks, _ = ctx.run_pr(iterations=10, damping=0.85)
                info = ctx.pr_bins_info()
                print(f"{name} {shape}: {info}")
                assert info["pull_groups"] > 3 * ((g.num_vertices >> 13) + 1)
                assert float(np.abs(ranks - ref).sum()) <= PR_TOL, (name, shape)
    finally:
        ctx.set_option(W.OPT_PR_BIN_SLICE, 262144)


def test_binned_pair_count():
    """Native order: pairs = distinct (source window, destination) pairs plus
    the cuts at 512-entry tile ends that fall inside a pair."""
    c = W.Context(0, relabel=False)
    try:
        g = W.rmat(16, 16, seed=3)
        c.load(g)
        c.set_pr_shape(BINNED[0])
        c.init_pr(0.85)
        info = c.pr_bins_info()
        src = g.sources.astype(np.int64)
        dst = g.destinations().astype(np.int64)
        distinct = np.unique((src >> 14) << 32 | dst).size
        E = g.num_edges
        print(info, distinct, E)
        assert info["code_entries"] % 512 == 0 and info["code_entries"] >= E
        assert distinct <= info["pairs"] <= distinct + info["code_entries"] // 512
    finally:
        c.close()


def test_binned_tiles_over_many_bins():
    """A window whose tiles span more than 64 destination bins (3000 edges from
    the first window to destinations spread over 4M vertices, 256 or 512
    bins): the block offsets past the 64 a tile's lanes hold come from the
    wavefront's LDS row."""
    V = 1 << 22
    rng = np.random.default_rng(5)
    src = np.concatenate([rng.integers(0, 16384, 3000), rng.integers(0, V, 4000)])
    dst = np.concatenate([rng.integers(0, V, 3000), rng.integers(0, V, 4000)])
    g = W.from_edges(V, src, dst, dedup=True)
    ref, _ = oracle.Oracle.of(g, 8).pr(8, 0.85)
    c = W.Context(0, relabel=False)
    try:
        c.load(g)
        for shape in BINNED:
            c.set_pr_shape(shape)
            ranks, _ = c.run_pr(iterations=8, damping=0.85)
            print(shape, c.pr_bins_info(), float(np.abs(ranks - ref).sum()))
            assert float(np.abs(ranks - ref).sum()) <= PR_TOL
    finally:
        c.close()


def test_binned_single_vertex_and_no_edges():
    c = W.Context(0)
    try:
        for g in (W.from_edges(1, [], []), W.from_edges(40000, [], []), W.from_edges(3, [0, 1], [1, 2])):
            c.load(g)
            c.set_pr_shape(BINNED[0])
            ranks, _ = c.run_pr(iterations=4, damping=0.85)
            ref, _ = oracle.Oracle.of(g, 2).pr(4, 0.85)
            assert float(np.abs(ranks - ref).sum()) <= PR_TOL
    finally:
        c.close()
