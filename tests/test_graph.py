"""Host graph library: seeded generators, the binary edge-list format, errors."""
import os

import numpy as np
import pytest

import wedge_amd as W


def _check_csc(g, dedup=True):
    V = g.num_vertices
    off = g.offsets.astype(np.int64)
    assert off[0] == 0 and off[-1] == g.num_edges and np.all(np.diff(off) >= 0)
    assert np.all(g.sources < V)
    dst = g.destinations()
    key = dst.astype(np.int64) * V + g.sources
    assert np.all(np.diff(key) >= (1 if dedup else 0))  # grouped by destination, ascending sources
    if dedup:
        assert not np.any(dst == g.sources)


def test_rmat_shape_and_determinism():
    g = W.rmat(12, 16, seed=1)
    assert g.num_vertices == 4096
    _check_csc(g)
    h = W.rmat(12, 16, seed=1)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.sources, h.sources)
    k = W.rmat(12, 16, seed=2)
    assert not (np.array_equal(g.offsets, k.offsets) and np.array_equal(g.sources, k.sources))
    # counter-based generation: thread count does not matter
    os.environ["WEDGE_THREADS"] = "1"
    try:
        one = W.rmat(12, 16, seed=1)
    finally:
        del os.environ["WEDGE_THREADS"]
    assert np.array_equal(g.sources, one.sources)


def test_rmat_edge_counts_are_pinned():
    # regression pins of this generator (draws = ef * 2^scale, dedup + no self-loops)
    assert W.rmat(10, 16, seed=1).num_edges == 12096
    assert W.rmat(16, 16, seed=1).num_edges == 955415


def test_rmat_skew_matches_quadrant_probabilities():
    # P(source bit = 0) = a + b = 0.76 for every bit position, before dedup;
    # after dedup low bits of frequent sources are still mostly 0
    g = W.rmat(14, 16, seed=5)
    s = g.sources.astype(np.int64)
    top = (s >> 13) & 1
    assert 0.70 < 1.0 - top.mean() < 0.80


def test_rmat_symmetrized_is_symmetric():
    g = W.rmat(11, 8, seed=2, symmetrize=True)
    _check_csc(g)
    dst = g.destinations().astype(np.int64)
    fwd = set(zip(g.sources.tolist(), dst.tolist()))
    assert all((b, a) in fwd for a, b in fwd)


def test_rmat_f64_weights():
    g = W.rmat(10, 16, seed=2, weights="f64")
    assert g.weights.dtype == np.float64 and g.weights.shape[0] == g.num_edges
    assert np.all(g.weights > 0.0) and np.all(g.weights <= 1.0)
    dst = g.destinations()
    k = 17
    assert g.weights[k] == W.weight_f64(int(g.sources[k]), int(dst[k]), 2)


def test_grid_structure():
    R, Cc = 20, 30
    g = W.grid(R, Cc, seed=3)
    _check_csc(g)
    assert g.num_vertices == R * Cc and g.weights.dtype == np.uint32
    assert np.all((g.weights >= 1) & (g.weights <= 1000))
    dst = g.destinations().astype(np.int64)
    src = g.sources.astype(np.int64)
    dr = np.abs(dst // Cc - src // Cc)
    dc = np.abs(dst % Cc - src % Cc)
    assert np.all((dr <= 1) & (dc <= 1))  # lattice and diagonal neighbours only
    wmap = {(a, b): w for a, b, w in zip(src.tolist(), dst.tolist(), g.weights.tolist())}
    assert all(wmap[(b, a)] == w for (a, b), w in wmap.items())  # symmetric weights
    deg = np.diff(g.offsets.astype(np.int64))
    assert 1.5 < deg.mean() < 3.5


def test_file_roundtrip(tmp_path):
    for g, kind in ((W.rmat(9, 8, seed=1), None), (W.rmat(9, 8, seed=1, weights="f64"), "f64"),
                    (W.grid(8, 9, seed=3), "int")):
        p = str(tmp_path / "g.bin")
        g.write(p)
        words = 2 if kind is None else 3
        assert os.path.getsize(p) == 16 + 8 * words * g.num_edges
        h = W.read_graph(p, weights=kind)
        assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.sources, h.sources)
        if kind is not None:
            assert np.array_equal(g.weights, h.weights)


def _write_raw(path, V, recs, words=2):
    arr = np.array([V, len(recs)] + [x for r in recs for x in r], dtype=np.uint64)
    arr.tofile(path)


def test_reader_errors(tmp_path):
    p = str(tmp_path / "bad.bin")
    _write_raw(p, 4, [(0, 1), (2, 1), (1, 3)])
    g = W.read_graph(p)
    assert g.num_edges == 3
    with open(p, "ab") as f:
        f.write(b"\0" * 8)  # size no longer 16 + 16E
    with pytest.raises(W.WedgeError) as e:
        W.read_graph(p)
    assert e.value.code == W.WG_E_FORMAT
    _write_raw(p, 4, [(0, 3), (2, 1)])  # not grouped by ascending destination
    with pytest.raises(W.WedgeError) as e:
        W.read_graph(p)
    assert e.value.code == W.WG_E_FORMAT
    _write_raw(p, 4, [(7, 1)])  # source out of range
    with pytest.raises(W.WedgeError) as e:
        W.read_graph(p)
    assert e.value.code == W.WG_E_RANGE
    with pytest.raises(W.WedgeError) as e:
        W.read_graph(str(tmp_path / "missing.bin"))
    assert e.value.code == W.WG_E_IO


def test_reader_keeps_duplicates_and_sorts_sources(tmp_path):
    p = str(tmp_path / "dup.bin")
    _write_raw(p, 3, [(2, 0), (1, 0), (2, 0), (0, 1), (1, 1)])
    g = W.read_graph(p)
    assert g.offsets.tolist() == [0, 3, 5, 5]
    assert g.sources.tolist() == [1, 2, 2, 0, 1]  # duplicates and self-loops kept, sources ascending


def test_generator_argument_errors():
    with pytest.raises(W.WedgeError):
        W.rmat(0, 16, seed=1)
    with pytest.raises(W.WedgeError):
        W.rmat(10, 16, seed=1, a=0.9, b=0.2, c=0.1)
    with pytest.raises(W.WedgeError):
        W.grid(1, 5, seed=1)


def test_union_keeps_every_member_in_its_range():
    ms = [W.rmat(5, 4, seed=1), W.from_edges(1, [], []), W.rmat(7, 8, seed=2, symmetrize=True)]
    u, first = W.union(ms)
    assert first.tolist() == [0, ms[0].num_vertices, ms[0].num_vertices + 1, u.num_vertices]
    assert u.num_edges == sum(m.num_edges for m in ms)
    dst, src = u.destinations().astype(np.int64), u.sources.astype(np.int64)
    for g, m in enumerate(ms):
        a, b = int(first[g]), int(first[g + 1])
        sel = (dst >= a) & (dst < b)
        assert np.all((src[sel] >= a) & (src[sel] < b))  # no edge leaves its member
        assert np.array_equal(src[sel] - a, m.sources.astype(np.int64))
        assert np.array_equal(dst[sel] - a, m.destinations().astype(np.int64))
    with pytest.raises(ValueError):
        W.union([W.rmat(4, 4, seed=1), W.rmat(4, 4, seed=1, weights="f64")])
    with pytest.raises(ValueError):
        W.union([])


def test_batch_shapes_are_listed():
    shapes = [W.batch_shape_at(i) for i in range(W.BATCH_NSHAPES)]
    assert len(set(shapes)) == W.BATCH_NSHAPES
    assert {s[0] for s in shapes} == {4, 8, 16}
    with pytest.raises(W.WedgeError):
        W.batch_shape_at(W.BATCH_NSHAPES)


def test_pr_shapes_extend_the_batch_shapes():
    shapes = [W.pr_shape_at(i) for i in range(W.PR_NSHAPES)]
    assert len(set(shapes)) == W.PR_NSHAPES
    assert shapes[:W.BATCH_NSHAPES] == [W.batch_shape_at(i) for i in range(W.BATCH_NSHAPES)]
    assert {s[0] for s in shapes} == {2, 4, 8, 16}
    with pytest.raises(W.WedgeError):
        W.pr_shape_at(W.PR_NSHAPES)
