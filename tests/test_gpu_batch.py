"""Batched sources on the MI355X: every source of a batched traversal must
equal an independent single-source run of the oracle (bit-exact depths)."""
import numpy as np
import pytest

import oracle
import wedge_amd as W

pytestmark = pytest.mark.gpu

MS_SCHEDULES = [dict(schedule="plain"), dict(schedule="dense"), dict(schedule="frontier", block_log2=3),
                dict(schedule="frontier", block_log2=6)]


@pytest.fixture(scope="module", params=[True, False], ids=["relabel", "native-order"])
def ctx(request):
    c = W.Context(0, relabel=request.param)
    yield c
    c.close()


def _graphs():
    return [("rmat", W.rmat(12, 16, seed=2)), ("rmat_sym", W.rmat(11, 16, seed=3, symmetrize=True)),
            ("grid", W.grid(40, 70, seed=4))]


def _roots(g, n, seed):
    rng = np.random.default_rng(seed)
    r = rng.integers(0, g.num_vertices, n).astype(np.uint64)
    r[0] = g.max_outdegree_vertex()
    if n > 3:
        r[3] = r[1]  # a repeated root
    return r


@pytest.mark.parametrize("nroots", [1, 5, 64])
def test_msbfs_matches_single_source_oracle(ctx, nroots):
    for name, g in _graphs():
        ctx.load(g)
        roots = _roots(g, nroots, seed=nroots)
        orc = oracle.Oracle.of(g, 4)
        want = [orc.bfs(int(r))[1] for r in roots]
        od = g.outdegrees().astype(np.uint64)
        for opts in MS_SCHEDULES:
            levels, reached = ctx.msbfs(roots, **opts)
            for i in range(nroots):
                assert np.array_equal(ctx.read_msbfs(i), want[i]), f"{name} {opts} root #{i}"
            assert levels == max(int(d[d != np.uint32(2**32 - 1)].max()) for d in want) + 1
            assert reached == sum(int(od[d != np.uint32(2**32 - 1)].sum()) for d in want)


def test_msbfs_reuses_buffers_without_stale_depths(ctx):
    """Consecutive runs on one graph keep their buffers and do not clear the
    depth matrix: entries of sources that did not reach a vertex must read as
    unreached even where an earlier run wrote a depth."""
    g = W.rmat(12, 4, seed=6)  # directed and sparse: roots reach different sets
    ctx.load(g)
    orc = oracle.Oracle.of(g, 4)
    first = _roots(g, 64, seed=11)
    ctx.msbfs(first)
    rng = np.random.default_rng(12)
    od = g.outdegrees()
    # sources with few out-edges, so that most vertices stay unreached
    low = np.flatnonzero((od > 0) & (od <= 2))
    second = rng.choice(low, 5, replace=False).astype(np.uint64)
    for roots in (second, first[:7]):
        ctx.msbfs(roots)
        for i in range(len(roots)):
            assert np.array_equal(ctx.read_msbfs(i), orc.bfs(int(roots[i]))[1]), f"root #{i}"
    ctx.init_bfs(int(first[0]))  # another application frees the buffers
    ctx.msbfs(second)
    assert np.array_equal(ctx.read_msbfs(4), orc.bfs(int(second[4]))[1])


def test_msbfs_heavy_destinations(ctx):
    rng = np.random.default_rng(7)
    V = 6000
    src = np.concatenate([rng.integers(1, V, 5000), rng.integers(0, V, 20000)])
    dst = np.concatenate([np.zeros(5000, np.int64), rng.integers(0, V, 20000)])
    g = W.from_edges(V, src, dst, dedup=True)
    ctx.load(g)
    roots = _roots(g, 64, seed=9)
    orc = oracle.Oracle.of(g, 4)
    ctx.msbfs(roots)
    for i in (0, 17, 63):
        assert np.array_equal(ctx.read_msbfs(i), orc.bfs(int(roots[i]))[1])


def test_msbfs_errors(ctx):
    g = W.rmat(8, 4, seed=1)
    ctx.load(g)
    with pytest.raises(W.WedgeError):
        ctx.msbfs(np.zeros(0, np.uint64))
    with pytest.raises(W.WedgeError):
        ctx.msbfs(np.zeros(65, np.uint64))
    with pytest.raises(W.WedgeError):
        ctx.msbfs([g.num_vertices])  # root out of range
    ctx.msbfs([0, 1])
    with pytest.raises(W.WedgeError):
        ctx.read_msbfs(2)  # only two roots
    ctx.init_bfs(0)
    with pytest.raises(W.WedgeError):
        ctx.read_msbfs(0)  # the batched state was replaced


# ---- personalised PageRank ---------------------------------------------------------------

PPR_TOL = 2e-7  # L1 per seed: FP32-stored contributions (about 10x the measured error), FP64 sums


@pytest.mark.parametrize("nseeds", [1, 7, 32])
def test_ppr_matches_oracle(ctx, nseeds):
    for name, g in _graphs() + [("hub", W.from_edges(6000, np.concatenate([np.random.default_rng(7).integers(1, 6000, 5000),
                                                                           np.random.default_rng(8).integers(0, 6000, 20000)]),
                                                     np.concatenate([np.zeros(5000, np.int64),
                                                                     np.random.default_rng(9).integers(0, 6000, 20000)]),
                                                     dedup=True))]:
        ctx.load(g)
        seeds = _roots(g, nseeds, seed=nseeds + 100)
        want = oracle.Oracle.of(g, 4).ppr(seeds, 12, 0.85)
        ctx.ppr(seeds, 12, 0.85)
        for i in range(nseeds):
            got = ctx.read_ppr(i)
            err = np.abs(got - want[i]).sum()
            assert err <= PPR_TOL, f"{name} seed #{i}: L1 {err:.3e}"
            assert abs(got.sum() - 1.0) < 1e-6
        print(f"PPR {name} {nseeds} seeds: max L1 {max(np.abs(ctx.read_ppr(i) - want[i]).sum() for i in range(nseeds)):.2e}")


def test_ppr_reuses_buffers_across_calls(ctx):
    """Calls on one graph keep their buffers: a call with fewer seeds and other
    iteration counts after a 32-seed call must still match the oracle, also
    after the PageRank tile shape was changed in between."""
    g = W.rmat(12, 16, seed=5)
    ctx.load(g)
    orc = oracle.Oracle.of(g, 4)
    ctx.ppr(_roots(g, 32, seed=21), 9, 0.85)
    for seeds, iters in ((_roots(g, 3, seed=22), 12), (_roots(g, 7, seed=23), 5)):
        ctx.ppr(seeds, iters, 0.85)
        want = orc.ppr(seeds, iters, 0.85)
        for i in range(len(seeds)):
            assert np.abs(ctx.read_ppr(i) - want[i]).sum() <= PPR_TOL, f"{len(seeds)} seeds, seed #{i}"
    ctx.set_pr_shape((16, 2, 1))
    seeds = _roots(g, 5, seed=24)
    ctx.ppr(seeds, 6, 0.85)
    want = orc.ppr(seeds, 6, 0.85)
    assert all(np.abs(ctx.read_ppr(i) - want[i]).sum() <= PPR_TOL for i in range(len(seeds)))
    ctx.set_pr_shape(None)


def test_ppr_is_reproducible_and_checks_arguments(ctx):
    g = W.rmat(12, 16, seed=2)
    ctx.load(g)
    seeds = _roots(g, 32, seed=3)
    ctx.ppr(seeds, 8, 0.85)
    a = [ctx.read_ppr(i) for i in (0, 31)]
    ctx.ppr(seeds, 8, 0.85)
    b = [ctx.read_ppr(i) for i in (0, 31)]
    assert all(np.array_equal(x, y) for x, y in zip(a, b))  # fixed-order sums
    with pytest.raises(W.WedgeError):
        ctx.ppr(np.zeros(33, np.uint64))
    with pytest.raises(W.WedgeError):
        ctx.ppr([0], iterations=0)
    with pytest.raises(W.WedgeError):
        ctx.read_ppr(32)


# ---- SSSP from many sources ----------------------------------------------------------------


@pytest.mark.parametrize("nroots", [1, 6, 64])
def test_mssssp_matches_single_source_oracle(ctx, nroots):
    for name, g in (("rmat_w", W.rmat(12, 16, seed=2, weights="f64")), ("grid", W.grid(40, 70, seed=4))):
        ctx.load(g)
        roots = _roots(g, nroots, seed=nroots + 50)
        orc = oracle.Oracle.of(g, 4)
        want = [orc.sssp(int(r))[0] for r in roots]
        for opts in (dict(schedule="plain"), dict(schedule="dense"), dict(schedule="frontier")):
            ctx.mssssp(roots, **opts)
            for i in range(nroots):
                got = ctx.read_mssssp(i)
                if got.dtype == np.uint32:
                    got = np.where(got == W.UNREACHED_U32, np.uint64(2**64 - 1), got.astype(np.uint64))
                assert np.array_equal(got, want[i]), f"{name} {opts} root #{i}"


def test_mssssp_errors(ctx):
    g = W.rmat(8, 4, seed=1)
    ctx.load(g)
    with pytest.raises(W.WedgeError):
        ctx.mssssp([0])  # unweighted graph
    g = W.grid(10, 10, seed=1)
    ctx.load(g)
    with pytest.raises(W.WedgeError):
        ctx.mssssp(np.zeros(65, np.uint64))
    ctx.mssssp([0, 5])
    with pytest.raises(W.WedgeError):
        ctx.read_mssssp(2)


@pytest.mark.parametrize("slice_pairs", [262144, 29])
def test_msbfs_binned_dense_levels_match_oracle(slice_pairs):
    """Dense levels on the binned layout (WG_OPT_MS_BINNED = 1): frontier masks
    ORed per (source window, destination) pair, then per destination bin,
    bins split over many pull workgroups with a tiny WG_OPT_PR_BIN_SLICE."""
    c = W.Context(0)
    try:
        c.set_option(W.OPT_MS_BINNED, 1)
        c.set_option(W.OPT_PR_BIN_SLICE, slice_pairs)
        for name, g in _graphs() + [("rmat16", W.rmat(16, 8, seed=5))]:
            c.load(g)
            roots = _roots(g, 64, seed=21)
            orc = oracle.Oracle.of(g, 8)
            want = [orc.bfs(int(r))[1] for r in roots]
            for opts in (dict(schedule="dense"), dict(schedule="plain")):
                c.msbfs(roots, **opts)
                for i in range(64):
                    assert np.array_equal(c.read_msbfs(i), want[i]), f"{name} {opts} root #{i}"
    finally:
        c.close()


def test_msbfs_binned_kept_levels_reuse_and_stop_early():
    """Binned dense levels keep their new-source masks and write the depth rows
    after the run: runs that follow one another on one graph, runs stopped
    after a few levels (rows of the kept levels only), and a graph with more
    dense levels than masks kept (the later levels write their rows at once)."""
    c = W.Context(0)
    INF = np.uint32(2**32 - 1)
    try:
        c.set_option(W.OPT_MS_BINNED, 1)
        g = W.rmat(12, 4, seed=6)
        c.load(g)
        orc = oracle.Oracle.of(g, 4)
        first = _roots(g, 64, seed=11)
        od = g.outdegrees()
        low = np.flatnonzero((od > 0) & (od <= 2))
        second = np.random.default_rng(12).choice(low, 5, replace=False).astype(np.uint64)
        for roots, opts in ((first, dict(schedule="dense")), (second, dict(schedule="dense")),
                            (first[:9], dict(schedule="plain")), (first, dict(schedule="dense", max_iterations=2))):
            c.msbfs(roots, **opts)
            stop = opts.get("max_iterations", 0)
            for i in range(len(roots)):
                want = orc.bfs(int(roots[i]))[1]
                if stop:
                    want = np.where((want != INF) & (want <= stop), want, INF)
                assert np.array_equal(c.read_msbfs(i), want), f"{opts} root #{i}"
        road = W.grid(30, 50, seed=5)  # about 80 levels, all dense
        c.load(road)
        orc = oracle.Oracle.of(road, 4)
        roots = _roots(road, 64, seed=13)
        levels, _ = c.msbfs(roots, schedule="dense")
        assert levels > 8
        for i in (0, 1, 30, 63):
            assert np.array_equal(c.read_msbfs(i), orc.bfs(int(roots[i]))[1]), f"grid root #{i}"
    finally:
        c.close()
