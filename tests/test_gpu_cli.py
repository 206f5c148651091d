"""The wedge command line on the MI355X: single runs, several sources (-R),
personalised PageRank (-a ppr) and batches of graphs (comma-separated -g),
each output file checked against the oracle on the same generated graphs.
Output values are printed with the reference's formats (%.4le for ranks), so
ranks compare to 5 significant digits; depths, labels and distances exactly."""
import subprocess

import numpy as np
import pytest

import oracle
import wedge_amd as W

from helpers import U32_MAX

pytestmark = pytest.mark.gpu


def run(*args):
    r = subprocess.run([W.CLI_PATH, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout
    return r.stdout


def table(path):
    rows = [line.split() for line in open(path)]
    return rows


def test_cli_single_bfs_and_pr(tmp_path):
    out = str(tmp_path / "bfs.txt")
    g = W.rmat(10, 8, seed=1)
    root = g.max_outdegree_vertex()
    run("-a", "bfs", "-g", "rmat:10:8:1", "-r", str(root), "-o", out)
    _, depth, _ = oracle.Oracle.of(g).bfs(root)
    parent = np.array([int(r[1]) for r in table(out)])
    assert np.array_equal(parent >= 0, depth != U32_MAX)
    out = str(tmp_path / "pr.txt")
    run("-a", "pr", "-g", "rmat:10:8:1", "-N", "16", "-D", "0.85", "-o", out)
    ref, _ = oracle.Oracle.of(g).pr(16, 0.85)
    got = np.array([float(r[1]) for r in table(out)])
    assert np.allclose(got, ref, rtol=1e-4, atol=0)
    run("-a", "pr", "-g", "rmat:10:8:1", "-N", "16", "-D", "0.85", "-P", "-o", out)  # full precision
    got = np.array([float(r[1]) for r in table(out)])
    assert float(np.abs(got - ref).sum()) <= 1e-10


@pytest.mark.parametrize("app", ["bfs", "cc", "sssp"])
def test_cli_validates_results(app):
    out = run("-a", app, "-g", "rmat:11:8:2", "-r", "0", "-c")
    assert "Validation:  passed" in out
    out = run("-a", app, "-g", "rmat:8:8:1,rmat:9:4:2", "-r", "0", "-c")
    assert "Validation:  passed" in out


def test_cli_multi_source_bfs_and_sssp(tmp_path):
    roots = [0, 5, 9, 700]
    g = W.rmat(10, 8, seed=1)
    out = str(tmp_path / "ms.txt")
    stdout = run("-a", "bfs", "-g", "rmat:10:8:1", "-R", ",".join(map(str, roots)), "-o", out)
    assert "4 sources" in stdout
    rows = np.array([[int(x) for x in r] for r in table(out)])
    o = oracle.Oracle.of(g)
    for i, r in enumerate(roots):
        d = o.bfs(r)[1].astype(np.int64)
        d[d == int(U32_MAX)] = -1
        assert np.array_equal(rows[:, 1 + i], d), f"root {r}"
    gw = W.rmat(10, 8, seed=1, weights="f64")
    out = str(tmp_path / "mss.txt")
    run("-a", "sssp", "-g", "rmat:10:8:1", "-R", ",".join(map(str, roots)), "-o", out)
    rows = table(out)
    ow = oracle.Oracle.of(gw)
    for i, r in enumerate(roots):
        ref = ow.sssp(r)[0]
        got = np.array([float(x[1 + i]) for x in rows])
        got[got == -1.0] = np.inf
        assert np.array_equal(got, ref), f"root {r}"


def test_cli_ppr(tmp_path):
    seeds = [3, 4, 100]
    g = W.rmat(10, 8, seed=1)
    out = str(tmp_path / "ppr.txt")
    run("-a", "ppr", "-g", "rmat:10:8:1", "-R", ",".join(map(str, seeds)), "-N", "10", "-D", "0.85", "-o", out)
    rows = np.array([[float(x) for x in r[1:]] for r in table(out)])
    ref = oracle.Oracle.of(g).ppr(seeds, 10, 0.85)
    assert np.allclose(rows.T, ref, rtol=1e-4, atol=1e-12)


def _members(spec):
    return [W.rmat(int(a), int(b), seed=int(c), symmetrize="sym" in rest)
            for a, b, c, *rest in (s.split(":")[1:] for s in spec.split(","))]


@pytest.mark.parametrize("app", ["pr", "cc", "bfs"])
def test_cli_batch_of_graphs(tmp_path, app):
    spec = "rmat:8:8:1:sym,rmat:9:4:2:sym,rmat:6:16:3:sym" if app == "cc" else "rmat:8:8:1,rmat:9:4:2,rmat:6:16:3"
    ms = _members(spec)
    out = str(tmp_path / f"{app}.txt")
    stdout = run("-a", app, "-g", spec, "-N", "16", "-D", "0.85", "-r", "0", "-o", out)
    assert "batch of 3 graphs" in stdout
    rows = table(out)
    for g, m in enumerate(ms):
        mine = [r for r in rows if int(r[0]) == g]
        assert len(mine) == m.num_vertices
        vals = np.array([float(r[2]) for r in mine])
        o = oracle.Oracle.of(m)
        if app == "pr":
            assert np.allclose(vals, o.pr(16, 0.85)[0], rtol=1e-4, atol=0)
        elif app == "cc":
            assert np.array_equal(vals, o.cc()[0])
        else:
            depth = o.bfs(0)[1]
            assert np.array_equal(vals >= 0, depth != U32_MAX)
            par = vals.astype(np.int64)
            for v in np.nonzero((depth != U32_MAX) & (np.arange(m.num_vertices) != 0))[0]:
                assert depth[par[v]] + 1 == depth[v]
