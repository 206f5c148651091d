"""The vectorised parent-rule helper (used at full benchmark sizes) agrees with
the per-edge version on small graphs, including chunks that split nothing and
chunks of a single destination."""
import numpy as np
import pytest

import oracle
import wedge_amd as W

from helpers import min_frontier_parent, min_frontier_parent_vec


@pytest.mark.parametrize("scale,ef,sym,chunk", [(10, 8, False, 1000), (12, 8, True, 1 << 26), (9, 8, False, 1)])
def test_min_frontier_parent_vec_matches(scale, ef, sym, chunk):
    g = W.rmat(scale, ef, seed=3, symmetrize=sym)
    o = oracle.Oracle.of(g, 2)
    _, depth, _ = o.bfs(g.max_outdegree_vertex())
    o.close()
    assert np.array_equal(min_frontier_parent(g, depth), min_frontier_parent_vec(g, depth, chunk=chunk))
