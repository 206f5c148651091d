"""The wedge command line reading the reference's binary edge list (-i) on the
MI355X: 2-word files (unweighted), 3-word files with integer weights and with
FP64 weights (-W f64), each run with -a pr | bfs | cc | sssp and -o, and every
output file compared with the oracle on the graph that was written.

Reference contract: the file layout and exact-size rule of
source/graphdata.c:946-1133 (size check :963-979), the options of
source/cmdline.c:65-108, the output formats of source/execution_*.c.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import wedge_amd as W

from helpers import U32_MAX, min_frontier_parent

pytestmark = pytest.mark.gpu


def run(*args):
    r = subprocess.run([W.CLI_PATH, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout
    return r.stdout


def column(path, dtype):
    rows = [line.split() for line in open(path)]
    assert [int(r[0]) for r in rows] == list(range(len(rows)))
    return np.array([dtype(r[1]) for r in rows])


def _bfs_check(g, path, root):
    _, depth, _ = oracle.Oracle.of(g).bfs(root)
    parent = column(path, int)
    reached = depth != U32_MAX
    assert np.array_equal(parent >= 0, reached)
    want = min_frontier_parent(g, depth).astype(np.int64)
    nonroot = reached & (np.arange(g.num_vertices) != root)
    assert np.array_equal(parent[nonroot], want[nonroot])


@pytest.mark.parametrize("sched", ["plain", "learned"])
def test_cli_two_word_file_pr_bfs_cc(tmp_path, sched):
    g = W.rmat(12, 8, seed=7)
    f = str(tmp_path / "g.bin")
    g.write(f)
    assert os.path.getsize(f) == 16 + 16 * g.num_edges
    out = str(tmp_path / "o.txt")
    run("-a", "pr", "-i", f, "-N", "16", "-D", "0.85", "-P", "-o", out)
    ref, _ = oracle.Oracle.of(g).pr(16, 0.85)
    assert float(np.abs(column(out, float) - ref).sum()) <= 1e-10
    run("-a", "pagerank", "-i", f, "-N", "4", "-o", out)  # reference alias, its damping 0.875, %.4le
    ref, _ = oracle.Oracle.of(g).pr(4, 0.875)
    assert np.allclose(column(out, float), ref, rtol=1e-4, atol=0)
    root = g.max_outdegree_vertex()
    run("-a", "bfs", "-i", f, "-r", str(root), "-s", sched, "-o", out)
    _bfs_check(g, out, root)
    gs = W.rmat(11, 8, seed=8, symmetrize=True)
    fs = str(tmp_path / "s.bin")
    gs.write(fs)
    run("-a", "cc", "-i", fs, "-s", sched, "-o", out)
    ref, _ = oracle.Oracle.of(gs).cc()
    assert np.array_equal(column(out, int).astype(np.float64), ref)


@pytest.mark.parametrize("sched", ["plain", "learned"])
def test_cli_three_word_files_sssp(tmp_path, sched):
    out = str(tmp_path / "o.txt")
    gi = W.grid(40, 64, seed=5)  # integer weights in [1, 1000]
    fi = str(tmp_path / "gi.bin")
    gi.write(fi)
    assert os.path.getsize(fi) == 16 + 24 * gi.num_edges
    root = gi.max_outdegree_vertex()
    run("-a", "sssp", "-i", fi, "-r", str(root), "-s", sched, "-o", out)
    ref, _ = oracle.Oracle.of(gi).sssp(root)
    got = column(out, int)
    want = np.where(ref == np.uint64(2**64 - 1), -1, ref.astype(np.int64))
    assert np.array_equal(got, want)
    gf = W.rmat(12, 8, seed=9, weights="f64")
    ff = str(tmp_path / "gf.bin")
    gf.write(ff)
    root = gf.max_outdegree_vertex()
    run("-a", "sssp", "-i", ff, "-W", "f64", "-r", str(root), "-s", sched, "-o", out)
    ref, _ = oracle.Oracle.of(gf).sssp(root)
    got = column(out, float)
    got[got == -1.0] = np.inf
    assert np.array_equal(got, ref)  # %.17g round-trips every double
    # the reference reads the weight word only for SSSP (source/application.c:171)
    # and checks the file size against that (source/graphdata.c:963-979): a
    # 3-word file given to BFS is refused as malformed
    r = subprocess.run([W.CLI_PATH, "-a", "bfs", "-i", ff, "-r", str(root)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "malformed" in r.stdout


def test_cli_input_errors(tmp_path):
    g = W.rmat(8, 4, seed=1)
    f = str(tmp_path / "g.bin")
    g.write(f)
    with open(f, "ab") as fh:
        fh.write(b"\0" * 8)  # not 16 + 16E bytes any more
    r = subprocess.run([W.CLI_PATH, "-a", "bfs", "-i", f], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    r = subprocess.run([W.CLI_PATH, "-a", "bfs", "-i", str(tmp_path / "missing.bin")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0
