"""Writes tests/golden/golden_small.npz: small seeded graphs and their expected
BFS depths, CC labels, SSSP distances, PageRank and personalised PageRank
ranks, computed with scipy (breadth-first shortest paths, connected
components, Dijkstra) and numpy sparse mat-vec PageRank loops -- implementations independent of both the
oracle and the HIP path.  The reference repository ships no test vectors, so
these fixtures are the committed known answers.

Run from the repository root after building: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import wedge_amd as W  # noqa: E402
from helpers import (bfs_depths_scipy, cc_min_labels_undirected, pagerank_closed_form,  # noqa: E402
                     ppr_closed_form, sssp_dijkstra)

CASES = {
    "rmat10_s1": lambda: W.rmat(10, 16, seed=1),
    "rmat11_sym_s2": lambda: W.rmat(11, 8, seed=2, symmetrize=True),
    "rmat10_w_s2": lambda: W.rmat(10, 16, seed=2, weights="f64"),
    "grid24x40_s3": lambda: W.grid(24, 40, seed=3),
}


def main():
    out = {}
    for name, make in CASES.items():
        g = make()
        out[f"{name}__offsets"] = g.offsets
        out[f"{name}__sources"] = g.sources
        if g.weights is not None:
            out[f"{name}__weights"] = g.weights
        root = g.max_outdegree_vertex()
        out[f"{name}__root"] = np.array([root], np.uint64)
        out[f"{name}__bfs_depth"] = bfs_depths_scipy(g, root)
        out[f"{name}__pr_d085_n16"] = pagerank_closed_form(g, 16, 0.85)
        out[f"{name}__pr_d0875_n16"] = pagerank_closed_form(g, 16, 0.875)
        seeds = np.array([root, 1, 7], np.uint64)
        out[f"{name}__ppr_seeds"] = seeds
        out[f"{name}__ppr_d085_n16"] = ppr_closed_form(g, seeds, 16, 0.85)
        if "sym" in name or "grid" in name:
            out[f"{name}__cc"] = cc_min_labels_undirected(g)
        if g.weights is not None:
            d = sssp_dijkstra(g, root)
            if g.weights.dtype == np.float64:
                out[f"{name}__sssp"] = d
            else:
                o = np.full(d.shape[0], np.iinfo(np.uint64).max, np.uint64)
                fin = ~np.isinf(d)
                o[fin] = d[fin].astype(np.uint64)
                out[f"{name}__sssp"] = o
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_small.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
