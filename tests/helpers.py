"""Shared checks: invariants that pin BFS/CC/SSSP/PR results independently of
any one implementation, used against both the oracle and the HIP path."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

U32_MAX = np.uint32(0xFFFFFFFF)
U64_MAX = np.uint64(0xFFFFFFFFFFFFFFFF)


def adjacency(g, weights=None):
    """scipy CSR with A[u, v] = weight of edge u -> v (1 if unweighted)."""
    dst = g.destinations().astype(np.int64)
    src = g.sources.astype(np.int64)
    data = np.ones(src.shape[0]) if weights is None else np.asarray(weights, dtype=np.float64)
    return sp.csr_matrix((data, (src, dst)), shape=(g.num_vertices, g.num_vertices))


def bfs_depths_scipy(g, root):
    d = cg.shortest_path(adjacency(g), method="D", unweighted=True, indices=root)
    out = np.full(g.num_vertices, U32_MAX, np.uint32)
    fin = ~np.isinf(d)
    out[fin] = d[fin].astype(np.uint32)
    return out


def pagerank_closed_form(g, iterations, damping):
    """The reference's PageRank loop written with a sparse mat-vec:
    contrib0 = (1/N)/od', acc = A^T contrib, S = sum(acc),
    contrib' = ((acc + (1-S)/N) * d + (1-d)/N) / od', rank = contrib * od'."""
    V = g.num_vertices
    od = g.outdegrees().astype(np.float64)
    odp = np.where(od == 0, float(V), od)
    At = adjacency(g).T.tocsr()
    c = (1.0 / V) / odp
    for _ in range(iterations):
        acc = At @ c
        S = acc.sum()
        c = ((acc + (1.0 - S) / V) * damping + (1.0 - damping) / V) / odp
    return c * odp


def ppr_closed_form(g, seeds, iterations, damping):
    """Personalised PageRank with a sparse mat-vec, one column per seed:
    C = R / od (0 where od = 0), D = mass on vertices without out-edges,
    R' = d * (A^T C + e_s D) + (1 - d) e_s, from R = e_s."""
    V = g.num_vertices
    od = g.outdegrees().astype(np.float64)
    At = adjacency(g).T.tocsr()
    Es = np.zeros((V, len(seeds)))
    Es[np.asarray(seeds, np.int64), np.arange(len(seeds))] = 1.0
    R = Es.copy()
    sink = od == 0
    for _ in range(iterations):
        C = np.where(od[:, None] > 0, R / np.where(od > 0, od, 1.0)[:, None], 0.0)
        D = R[sink].sum(axis=0)
        R = damping * (At @ C + Es * D) + (1.0 - damping) * Es
    return R.T


def cc_min_labels_undirected(g):
    n, comp = cg.connected_components(adjacency(g), directed=False)
    mins = np.full(n, np.iinfo(np.int64).max)
    np.minimum.at(mins, comp, np.arange(g.num_vertices))
    return mins[comp].astype(np.uint32)


def cc_min_reaching(g):
    """Directed input: label[v] = min({u : u reaches v} | {v}), by brute force."""
    A = adjacency(g)
    V = g.num_vertices
    lab = np.arange(V, dtype=np.int64)
    for u in range(V):
        order = cg.breadth_first_order(A, u, directed=True, return_predecessors=False)
        lab[order] = np.minimum(lab[order], u)
    return lab.astype(np.uint32)


def sssp_dijkstra(g, root):
    """FP64 Dijkstra (scipy): dist[u] + w per relaxation, the same single rounding
    per edge as the min-plus fixpoint."""
    return cg.dijkstra(adjacency(g, g.weights), indices=root)


def check_bfs_tree(g, root, depth, parent):
    """Every reached vertex except the root has a parent that is an in-neighbour
    one level up; unreached vertices have no parent.  Returns the number of bad
    vertices (0 = valid tree)."""
    V = g.num_vertices
    bad = 0
    reached = depth != U32_MAX
    off = g.offsets
    src = g.sources
    for v in np.nonzero(reached)[0]:
        if v == root:
            continue
        p = int(parent[v])
        if p >= V or depth[p] == U32_MAX or int(depth[p]) + 1 != int(depth[v]):
            bad += 1
            continue
        nb = src[off[v]:off[v + 1]]
        i = np.searchsorted(nb, p)
        if i >= nb.shape[0] or nb[i] != p:
            bad += 1
    return bad


def min_frontier_parent_vec(g, depth, chunk=1 << 26):
    """min_frontier_parent for large graphs: per destination, the smallest
    in-neighbour one level up, with segment minima over the destination-grouped
    in-lists (np.minimum.reduceat) taken in chunks of whole destinations."""
    V = g.num_vertices
    off = g.offsets.astype(np.int64)
    src = g.sources
    out = np.full(V, U32_MAX, np.uint32)
    lvl = depth.astype(np.int64)
    lvl[depth == U32_MAX] = -2  # unreached: no level matches
    v0 = 0
    while v0 < V:
        # whole destinations whose in-edges fit the chunk (at least one)
        v1 = int(np.searchsorted(off, off[v0] + chunk, side="right")) - 1
        v1 = min(max(v1, v0 + 1), V)
        a, b = int(off[v0]), int(off[v1])
        if b > a:
            s = src[a:b]
            dd = np.repeat(lvl[v0:v1], np.diff(off[v0:v1 + 1]))
            ok = (lvl[s] + 1 == dd) & (dd >= 1)
            cand = np.where(ok, s, U32_MAX).astype(np.uint32)
            nz = np.nonzero(off[v0 + 1:v1 + 1] > off[v0:v1])[0]
            out[v0 + nz] = np.minimum.reduceat(cand, off[v0 + nz] - a)
        v0 = v1
    return out


def min_frontier_parent(g, depth):
    """Smallest in-neighbour one level up, for every reached non-root vertex
    (U32_MAX elsewhere): the parent the HIP path defines."""
    V = g.num_vertices
    out = np.full(V, U32_MAX, np.uint32)
    dst = g.destinations()
    src = g.sources
    dd = depth[dst].astype(np.int64)
    ds = depth[src].astype(np.int64)
    ok = (depth[dst] != U32_MAX) & (depth[src] != U32_MAX) & (ds + 1 == dd)
    cand_v = dst[ok]
    cand_u = src[ok]
    np.minimum.at(out, cand_v, cand_u)
    return out
