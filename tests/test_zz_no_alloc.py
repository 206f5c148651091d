"""wg_validate_noalloc, the validation entry point that allocates no device
memory (its bitmaps and counter live in a fixed device array), called through
the C-ABI and compared with wg_validate on the same results: two calls in a
row, the second on a smaller graph, with and without violations, and the
rejection of a graph above its vertex limit."""
import ctypes as C

import numpy as np
import pytest

import wedge_amd as W

pytestmark = pytest.mark.gpu


def _noalloc(ctx):
    n = C.c_uint64(12345)
    rc = ctx._lib.wg_validate_noalloc(ctx.handle, C.byref(n))
    return rc, int(n.value)


def _both(ctx):
    want = ctx.validate()
    rc, got = _noalloc(ctx)
    assert rc == W.WG_OK
    return want, got


def test_noalloc_validation_equals_wg_validate():
    ctx = W.Context(0)
    try:
        # first call: a larger graph, a BFS stopped after two levels (edges
        # from reached to unreached vertices are violations)
        big = W.rmat(15, 16, seed=3)
        ctx.load(big)
        ctx.init_bfs(big.max_outdegree_vertex())
        ctx.converge(W.SCHED_PLAIN, max_iterations=2)
        want, got = _both(ctx)
        print(f"s15 BFS capped: {want} / {got}")
        assert want > 0 and got == want
        # second call, smaller graph: complete results (0) and capped ones (> 0)
        small = W.rmat(10, 8, seed=4, weights="f64")
        ctx.load(small)
        root = small.max_outdegree_vertex()
        ctx.run_sssp(root)
        want, got = _both(ctx)
        assert want == 0 and got == 0
        ctx.init_sssp(root)
        ctx.converge(W.SCHED_PLAIN, max_iterations=1)
        want, got = _both(ctx)
        print(f"s10 SSSP capped: {want} / {got}")
        assert want > 0 and got == want
        sym = W.rmat(9, 8, seed=5, symmetrize=True)
        ctx.load(sym)
        ctx.init_cc()
        ctx.converge(W.SCHED_PLAIN, max_iterations=1)
        want, got = _both(ctx)
        print(f"s9 CC capped: {want} / {got}")
        assert got == want
        ctx.run_cc()
        assert _both(ctx) == (0, 0)
        # the larger graph again, after the smaller ones: same count as before
        ctx.load(big)
        ctx.init_bfs(big.max_outdegree_vertex())
        ctx.converge(W.SCHED_PLAIN, max_iterations=2)
        want, got = _both(ctx)
        assert want > 0 and got == want
    finally:
        ctx.close()


def test_noalloc_validation_rejects_graphs_above_its_limit():
    ctx = W.Context(0)
    try:
        V = W.NOALLOC_MAX_VERTICES + 1
        ctx.load(W.from_edges(V, [], []))
        ctx.init_bfs(0)
        ctx.converge(W.SCHED_PLAIN)
        rc, _ = _noalloc(ctx)
        assert rc == W.WG_E_RANGE
        assert ctx.validate() == 0  # the allocating entry point still serves it
    finally:
        ctx.close()
