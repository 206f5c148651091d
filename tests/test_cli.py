"""wedge command line: the reference's option syntax and exit codes
(source/cmdline.c:127-179: 1 unknown option, 2 invalid value, 3 missing value,
4 extraneous value, 5 missing required option, 6 incompatible options)."""
import subprocess

import pytest

import wedge_amd as W


def run(*args):
    return subprocess.run([W.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args,code", [
    (["-a", "bfs"], 5),                                  # missing -i
    (["-i", "g.bin"], 5),                                # missing -a
    (["-a", "bfs", "-i", "g.bin", "-x"], 1),             # unknown option
    (["-a", "bfs", "-i", "g.bin", "-Z", "3"], 1),
    (["-a", "bfs", "-i", "g.bin", "-N", "0"], 2),        # invalid value
    (["-a", "bfs", "-i", "g.bin", "-r", "abc"], 2),
    (["-a", "bfs", "-i", "g.bin", "-s", "sometimes"], 2),
    (["-a", "bfs", "-i", "g.bin", "-b", "1"], 2),
    (["-a", "bfs", "-i", "g.bin", "-N"], 3),             # missing value
    (["-a", "bfs", "-i", "g.bin", "-f", "yes"], 4),      # extraneous value
    (["-a", "bfs", "-i", "g.bin", "-g", "rmat:8:8:1"], 6),  # -i and -g together
    (["-a", "bfs", "-i", "g.bin", "-R", "1,,2"], 2),     # root lists
    (["-a", "bfs", "-i", "g.bin", "-R", "x"], 2),
    (["-a", "bfs", "-i", "g.bin", "-R", "3,"], 2),
    (["-a", "bfs", "-i", "g.bin", "-R", ",".join(["1"] * 65)], 2),
    (["-a", "bfs", "-i", "g.bin", "-R"], 3),
    (["-a", "bfs", "-i", "g.bin", "-c", "1"], 4),        # -c takes no value
])
def test_cli_error_codes(args, code):
    assert run(*args).returncode == code


def test_cli_help_and_version():
    r = run("-h")
    assert r.returncode == 0 and "Usage:" in r.stdout
    r = run("-V")
    assert r.returncode == 0 and "wedge" in r.stdout


def test_cli_unknown_application():
    r = run("-a", "dfs", "-g", "rmat:6:4:1")
    assert r.returncode != 0 and "unable to run" in r.stdout


def test_cli_application_aliases_are_accepted():
    for name in ("BFS", "breadth_first_search", "PageRank", "connected_components", "SingleSourceShortestPath",
                 "ppr", "personalized_pagerank"):
        r = run("-a", name, "-g", "rmat:6:4:1", "-d", "99")  # device 99 never exists: stops after ingress
        assert "unable to run" not in r.stdout


def test_cli_bad_input_file(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(b"\x01\x00\x00\x00\x00\x00\x00\x00")
    r = run("-a", "pr", "-i", str(p))
    assert r.returncode != 0 and "ERROR" in r.stdout


def test_cli_ppr_needs_seeds():
    r = run("-a", "ppr", "-g", "rmat:6:4:1")
    assert r.returncode != 0 and "needs seeds" in r.stdout


def test_cli_batch_reports_member_errors():
    r = run("-a", "bfs", "-g", "rmat:6:4:1,bogus:1", "-d", "99")
    assert r.returncode != 0 and "batch of 2 graphs" in r.stdout and "ERROR" in r.stdout
