#!/usr/bin/env python3
"""Benchmark of the MI355X pull engine.

Headline (one step = one PageRank iteration, edge pull + vertex phase, FP64
contributions, damping 0.85, over the whole graph with the inputs resident in
HBM): R-MAT scale 27, edge factor 16, seed 5 (a/b/c/d = .57/.19/.19/.05,
self-loops and duplicates removed, grouped by destination; 134M vertices,
2.1 billion edges).  `value` is E / (time per iteration) in GTEPS;
`roofline` prices the iteration's algorithmic bytes (4E + 24V + 4: sources,
offsets, out-degrees, contributions read once and written once) per launch of
its kernels, timed with HIP events on the engine's stream, against the 8 TB/s
HBM peak; `cpu_baseline` is the oracle (a C restatement of Wedge's pull
engine) on the same graph, all host cores, a bounded number of iterations.

The same run reports, each with the oracle's time on the host beside it:
  large        the scale-27 graph: PageRank with FP32-stored contributions
               (error against the FP64 ranks), BFS from the max-out-degree
               vertex under the plain rule and the learned schedule with the
               algorithmic bytes counted by the kernels;
  pagerank_s20 R-MAT scale 20, edge factor 16, seed 1: 16 iterations at
               d = 0.85 and at the reference's 0.875;
  apps         R-MAT scale 24, seed 2: BFS and FP64 SSSP (plain and learned),
               CC on the symmetrised graph; the 4096 x 6144 road grid (seed 3):
               BFS and integer SSSP, plain and learned;
  batched      64-source BFS and SSSP and 32-seed personalised PageRank on the
               scale-24 graph, and 256 uneven R-MAT graphs as one batch.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import wedge_amd as W  # noqa: E402

METRIC = "GTEPS (PR iter, BFS) + HBM roofline %, 1 MI355X; learned vs plain schedule"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads_used() -> int:
    """Host threads the oracle may use: the CPUs this process may run on."""
    return max(1, len(os.sched_getaffinity(0)))


def pr_algorithmic_bytes(V: int, E: int, fp32: bool) -> int:
    """Bytes one PageRank iteration must move at minimum: sources 4E, offsets
    4(V+1), out-degrees 4V, each contribution read once and written once."""
    c = 4 if fp32 else 8
    return 4 * E + 4 * (V + 1) + 4 * V + 2 * c * V


def traversal_bytes(app: str, V: int, st: dict, work: tuple[int, int], wbytes: int = 0, fused: int = 0) -> int:
    """Bytes a traversal must move at minimum, from the work the kernels
    counted: per edge read, its 4-byte source (plus its weight for SSSP); per
    destination state examined, that state (BFS depth/parent word 8 B, SSSP
    distance 8 B FP64 or 4 B integer, CC label 4 B); per iteration that is not
    a fused level (those pass their frontier in queues) one read of the
    current frontier bitmap and one write of the next (2 * V/8); BFS also
    writes 8 B per newly reached vertex."""
    edges, dests = work
    state = {"bfs": 8, "cc": 4}.get(app, 8 if wbytes == 8 else 4)
    b = (4 + wbytes) * edges + state * dests + (st["iterations"] - fused) * 2 * ((V + 7) // 8)
    if app == "bfs":
        b += 8 * st["vertices_reached"]
    return b


def measured_traffic(name: str, key: str = "traffic_bytes_per_launch"):
    """HBM bytes per launch from a committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes, scaled by a copy kernel of known size),
    or None when no summary covers the workload."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return int(line.split(":", 1)[1])
    return None


def shape_name(shape) -> str:
    return "binned" if shape[2] == W.PR_BINNED else ("tiles-xcd" if shape[2] else "tiles")


def pr_timed(ctx, damping, fp32, steps, warmup, shape=None):
    """PageRank with the autotuned shape (or the given one): wall ms per
    iteration over `steps` iterations and the iteration's kernel ms per launch
    (HIP events on the engine's stream)."""
    ctx.init_pr(damping, fp32)
    if shape is None:
        shape, times = ctx.pr_autotune(2)
    else:
        ctx.set_pr_shape(shape)
        shape, times = ctx.pr_shape(), {}
    for _ in range(warmup):
        ctx.pr_iteration()
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx._lib.wg_pr_iteration(ctx.handle, None)
    ctx.sync()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    kern = {}
    for k in ("pr_tile", "pr_span", "pb_push", "pb_pull"):
        n, ms = ctx.profile_get(k)
        if n:
            kern[k] = ms / n
    ctx.profile(False)
    return wall, kern, shape, times


def roofline_entry(alg: int, kern: dict, traffic):
    ms = sum(kern.values())
    achieved = alg / (ms * 1e-3) / 1e9
    names = {"pr_tile": "k_pr_tile", "pr_span": "k_pr_span", "pb_push": "k_pb_push", "pb_pull": "k_pb_pull"}
    return {
        "bound": "hbm",
        "kernel": " + ".join(names[k] for k in kern) + " (one PageRank iteration)",
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "algorithmic_bytes_per_launch": alg, "avg_launch_ms": round(ms, 5),
        "kernel_ms": {names[k]: round(v, 5) for k, v in kern.items()},
    }


def counted_run(ctx, run):
    """One more run with the kernels that count their work (slower, so never
    the timed one); the work it reports, and its iteration count."""
    ctx.set_option(W.OPT_COUNT_WORK, 1)
    try:
        _, st = run()
        e, d = ctx.work_counts()
    finally:
        ctx.set_option(W.OPT_COUNT_WORK, 0)
    return (e, d), st["iterations"]


def run_best(ctx, app, root, reps, **kw):
    """Best of reps timed runs, and the work of the same run counted apart."""
    run = (lambda: ctx.run_bfs(root, **kw)) if app == "bfs" else (lambda: ctx.run_sssp(root, **kw))
    best = None
    for _ in range(reps):
        res, st = run()
        if best is None or st["running_ms"] < best[1]["running_ms"]:
            best = (res, st, ctx.fused_level_count())
    work, it = counted_run(ctx, run)
    if it != best[1]["iterations"]:
        log(f"{app} {kw}: the counted run took {it} iterations, the timed one {best[1]['iterations']}")
    return best[0], best[1], work, best[2]


def traversal_entry(app, V, st, work, wbytes=0, violations=None, fused=0):
    b = traversal_bytes(app, V, st, work, wbytes, fused)
    e = {"ms": round(st["running_ms"], 3), "iterations": st["iterations"],
         "frontier_iterations": st["frontier_iterations"],
         "gteps": round(st["edges_reached"] / (st["running_ms"] * 1e-3) / 1e9, 3),
         "edges_read": work[0], "dests_examined": work[1], "algorithmic_bytes": b,
         "roofline_frac": round(b / (st["running_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if violations is not None:
        e["violations"] = violations
    return e


def schedule_pair(ctx, g, root, reps, wbytes, orc=None):
    """BFS and SSSP under the plain rule and the learned schedule: GTEPS
    (edges whose source was reached / run time), iterations, algorithmic bytes
    and roofline fraction from the kernels' work counts, on-device certificate
    violations, identical outputs, and the oracle's time on the host."""
    out = {"vertices": g.num_vertices, "edges": g.num_edges, "root": root}
    for app in ("bfs", "sssp"):
        res = {}
        for sched in ("plain", "learned"):
            r, st, work, fused = run_best(ctx, app, root, reps, schedule=sched)
            res[sched] = (r, st)
            out[f"{app}_{sched}"] = traversal_entry(app, g.num_vertices, st, work, 0 if app == "bfs" else wbytes,
                                                    ctx.validate(), fused)
            out[f"{app}_{sched}"]["fused_levels"] = fused
        out[f"{app}_learned_vs_plain"] = round(res["plain"][1]["running_ms"] / res["learned"][1]["running_ms"], 3)
        out[f"{app}_identical"] = bool(np.array_equal(res["plain"][0], res["learned"][0]))
        if orc is not None:
            t0 = time.time()
            if app == "bfs":
                _, _, ost = orc.bfs(root)
            else:
                _, ost = orc.sssp(root)
            out[f"{app}_cpu_oracle"] = {"ms": round(ost["running_ms"], 2), "iterations": int(ost["iterations"]),
                                        "threads": orc.threads, "wall_s": round(time.time() - t0, 2)}
            out[f"{app}_gpu_vs_cpu"] = round(ost["running_ms"] / res["learned"][1]["running_ms"], 1)
        log(f"{app}: plain {out[f'{app}_plain']}, learned {out[f'{app}_learned']}, "
            f"x{out[f'{app}_learned_vs_plain']}, identical {out[f'{app}_identical']}, "
            f"cpu {out.get(f'{app}_cpu_oracle')}")
    return out


def best_ms(fn, reps):
    """Smallest wall time of fn() over reps runs (fn waits for the device)."""
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        res = fn()
        dt = (time.perf_counter() - t0) * 1e3
        if best is None or dt < best[0]:
            best = (dt, res)
    return best


def batched_sources(ctx, g, pr_ms, damping, reps=3):
    """64-source BFS and SSSP and 32-seed personalised PageRank on the loaded
    graph, each against the same work done one source at a time: sources and
    seeds drawn with seed 2 among the vertices with out-degree > 0."""
    od = g.outdegrees()
    rng = np.random.default_rng(2)
    roots = rng.choice(np.nonzero(od > 0)[0], 64, replace=False).astype(np.uint64)
    seeds = rng.choice(np.nonzero(od > 0)[0], 32, replace=False).astype(np.uint64)
    out = {}
    single = {s: float(np.mean([ctx.run_bfs(int(r), schedule=s)[1]["running_ms"] for r in roots[:16]]))
              for s in ("plain", "learned")}
    ms, (levels, reached) = best_ms(lambda: ctx.msbfs(roots), reps)
    out["msbfs_64"] = {
        "ms": round(ms, 3), "levels": levels, "gteps": round(reached / (ms * 1e-3) / 1e9, 2),
        "single_ms_plain": round(single["plain"], 3), "single_ms_learned": round(single["learned"], 3),
        "speedup_vs_64_plain": round(64 * single["plain"] / ms, 2),
        "speedup_vs_64_learned": round(64 * single["learned"] / ms, 2),
    }
    log(f"MS-BFS 64: {out['msbfs_64']}")
    iters = 16
    ms, _ = best_ms(lambda: (ctx.ppr(seeds, iters, damping), ctx.sync()), reps)
    out["ppr_32"] = {
        "ms_per_iteration": round(ms / iters, 3), "iterations": iters,
        "pr_iteration_ms": round(pr_ms, 3), "speedup_vs_32_pr": round(32 * pr_ms / (ms / iters), 2),
        "gteps": round(32 * g.num_edges / (ms / iters * 1e-3) / 1e9, 2),
    }
    log(f"PPR 32: {out['ppr_32']}")
    single_sssp = {s: float(np.mean([ctx.run_sssp(int(r), schedule=s)[1]["running_ms"] for r in roots[:4]]))
                   for s in ("plain", "learned")}
    ms, it = best_ms(lambda: ctx.mssssp(roots), 1)
    out["mssssp_64"] = {"ms": round(ms, 1), "iterations": it,
                        "single_ms_plain": round(single_sssp["plain"], 2),
                        "single_ms_learned": round(single_sssp["learned"], 2),
                        "speedup_vs_64_plain": round(64 * single_sssp["plain"] / ms, 2),
                        "speedup_vs_64_learned": round(64 * single_sssp["learned"] / ms, 2)}
    log(f"MS-SSSP 64: {out['mssssp_64']}")
    return out


def batched_graphs(damping, iterations=16, reps=5):
    """256 uneven R-MAT graphs (scale and edge factor drawn per graph from
    U{10..16} and U{8..32} with seed 4, graph i generated with seed 4000 + i)
    as one disjoint union: PageRank on every member in one captured HIP graph
    with the autotuned tile shape, and BFS from every member's max-out-degree
    vertex in one traversal; against the same work run graph by graph."""
    rng = np.random.default_rng(4)
    scales = rng.integers(10, 17, 256)
    efs = rng.integers(8, 33, 256)
    members = [W.rmat(int(s), int(e), seed=4000 + i) for i, (s, e) in enumerate(zip(scales, efs))]
    E = sum(m.num_edges for m in members)
    V = sum(m.num_vertices for m in members)
    locals_ = [m.max_outdegree_vertex() for m in members]
    ctx = W.Context(0)
    first = ctx.load_batch(members)
    shape, times = ctx.batch_autotune(iterations, damping)
    ctx.batch_pr(iterations, damping)
    ctx.sync()
    ms_graph, _ = best_ms(lambda: (ctx.batch_pr(iterations, damping), ctx.sync()), reps)
    ms_plain, _ = best_ms(lambda: (ctx.batch_pr(iterations, damping, use_graph=False), ctx.sync()), reps)
    roots = first[:-1] + np.array(locals_, np.uint64)
    bfs_ms = None
    for _ in range(reps):
        ctx.init_bfs_roots(roots)
        t0 = time.perf_counter()
        ctx.converge(W.SCHED_PLAIN)
        dt = (time.perf_counter() - t0) * 1e3
        bfs_ms = dt if bfs_ms is None else min(bfs_ms, dt)
    ctx.close()
    # graph by graph: each member loaded on its own, initialisation not timed
    one = W.Context(0)
    seq_pr = seq_bfs = 0.0
    for m, r in zip(members, locals_):
        one.load(m)
        one.init_pr(damping)
        t0 = time.perf_counter()
        for _ in range(iterations):
            one._lib.wg_pr_iteration(one.handle, None)
        one.sync()
        seq_pr += (time.perf_counter() - t0) * 1e3
        one.init_bfs(r)
        t0 = time.perf_counter()
        one.converge(W.SCHED_PLAIN)
        seq_bfs += (time.perf_counter() - t0) * 1e3
    one.close()
    out = {
        "graphs": len(members), "vertices": V, "edges": E, "iterations": iterations,
        "autotuned_shape": {"edges_per_lane": shape[0], "blocks_per_cu": shape[1]},
        "autotune_ms": {f"{k[0]}x{k[1]}": round(v, 4) for k, v in times.items()},
        "pr_ms_hipgraph": round(ms_graph, 3), "pr_ms_plain_launches": round(ms_plain, 3),
        "pr_gteps": round(iterations * E / (ms_graph * 1e-3) / 1e9, 2),
        "pr_ms_one_by_one": round(seq_pr, 2), "pr_speedup": round(seq_pr / ms_graph, 2),
        "bfs_ms": round(bfs_ms, 3), "bfs_ms_one_by_one": round(seq_bfs, 2),
        "bfs_speedup": round(seq_bfs / bfs_ms, 2),
    }
    log(f"256 graphs: {out}")
    return out


def host_cpu_info():
    """CPU model and logical / physical core counts of the host."""
    model, phys = "unknown", set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k == "model name":
                    model = v
                cur[k] = v
            elif cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
        if cur:
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    return {"model": model, "logical": os.cpu_count(), "physical": len(phys) or None,
            "usable": len(os.sched_getaffinity(0))}


def oracle_of(g):
    import oracle  # test infrastructure: the checker and the CPU baseline only

    t0 = time.time()
    o = oracle.Oracle.of(g, threads=host_threads_used())
    return o, time.time() - t0


def cc_pair(scale, edge_factor, seed, reps=2, cpu=True):
    """CC on the symmetrised R-MAT (the reference's CC input) under the plain
    rule and the learned schedule: run time, iterations, reference-style
    E * iterations / time, algorithmic bytes, identical labels, and the
    oracle's time."""
    t0 = time.time()
    g = W.rmat(scale, edge_factor, seed=seed, symmetrize=True)
    log(f"rmat s{scale} ef{edge_factor} seed {seed} symmetrised: V={g.num_vertices} E={g.num_edges} "
        f"({time.time() - t0:.1f} s)")
    ctx = W.Context(0)
    ctx.load(g)
    out = {"vertices": g.num_vertices, "edges": g.num_edges}
    res = {}
    for sched in ("plain", "learned"):
        best = None
        for _ in range(reps):
            lab, st = ctx.run_cc(schedule=sched)
            if best is None or st["running_ms"] < best[1]["running_ms"]:
                best = (lab, st)
        work, _ = counted_run(ctx, lambda: ctx.run_cc(schedule=sched))
        res[sched] = best[0]
        st = best[1]
        e = traversal_entry("cc", g.num_vertices, st, work, violations=ctx.validate())
        e["bedges_per_s"] = round(g.num_edges * st["iterations"] / (st["running_ms"] * 1e-3) / 1e9, 2)
        out[f"cc_{sched}"] = e
    out["cc_learned_vs_plain"] = round(out["cc_plain"]["ms"] / out["cc_learned"]["ms"], 3)
    out["cc_identical"] = bool(np.array_equal(res["plain"], res["learned"]))
    ctx.close()
    if cpu:
        o, build_s = oracle_of(g)
        _, ost = o.cc()
        out["cc_cpu_oracle"] = {"ms": round(ost["running_ms"], 2), "iterations": int(ost["iterations"]),
                                "threads": o.threads, "layout_s": round(build_s, 1)}
        out["cc_gpu_vs_cpu"] = round(ost["running_ms"] / out["cc_plain"]["ms"], 1)
        o.close()
    log(f"cc: {out}")
    return out


def large_graph(args):
    """The headline: R-MAT scale 27 resident in HBM.  PageRank FP64 (timed
    steps, roofline), FP32-stored contributions, BFS plain and learned with
    kernel-counted bytes, and the oracle's PageRank and BFS on the host."""
    t0 = time.time()
    g = W.rmat(27, 16, seed=5)
    V, E = g.num_vertices, g.num_edges
    root = g.max_outdegree_vertex()
    log(f"rmat s27 ef16 seed 5: V={V} E={E} ({time.time() - t0:.1f} s, {W.host_threads()} host threads)")
    ctx = W.Context(0)
    t0 = time.time()
    ctx.load(g)
    log(f"uploaded + device layout built in {time.time() - t0:.1f} s")
    out = {"workload": "rmat-s27-ef16-seed5", "vertices": V, "edges": E, "root": root}
    res = {}
    for fp32 in (False, True) if not args.no_fp32 else (False,):
        if args.no_pr:
            break
        wall, kern, shape, times = pr_timed(ctx, args.damping, fp32, args.steps if not fp32 else 5, args.warmup,
                                            args.pr_shape)
        alg = pr_algorithmic_bytes(V, E, fp32)
        traffic = (measured_traffic("k_pb_rmat_s27_hbm_traffic.md",
                                    f"traffic_bytes_per_launch_{'float' if fp32 else 'double'}")
                   if shape[2] == W.PR_BINNED else None)
        key = "pr_fp32" if fp32 else "pr_fp64"
        res[key] = {"ms_per_iteration": round(wall, 4), "gteps": round(E / (wall * 1e-3) / 1e9, 3),
                    "shape": {"edges_per_lane": shape[0], "blocks_per_cu": shape[1], "kind": shape_name(shape)},
                    "roofline": roofline_entry(alg, kern, traffic),
                    "autotune_ms": {f"{s[0]}x{s[1]}x{s[2]}": round(v, 3) for s, v in times.items()}}
        if shape[2] == W.PR_BINNED:
            ctx.init_pr(args.damping, fp32)
            res[key]["binned_layout"] = ctx.pr_bins_info()
        log(f"s27 {key}: {res[key]}")
    out.update(res)
    if not args.no_fp32 and not args.no_pr:
        ranks = {}
        for fp32 in (False, True):
            ranks[fp32], _ = ctx.run_pr(iterations=16, damping=args.damping, contrib_fp32=fp32)
        out["pr_fp32_l1_vs_fp64"] = float(np.abs(ranks[True] - ranks[False]).sum())
        out["pr_iterations_compared"] = 16
        del ranks
    W.mlp_load()
    bfs = {}
    for sched in ("plain", "learned") if not args.no_bfs else ():
        r, st, work, fused = run_best(ctx, "bfs", root, 3, schedule=sched)
        bfs[sched] = r
        out[f"bfs_{sched}"] = traversal_entry("bfs", V, st, work, violations=ctx.validate(), fused=fused)
        out[f"bfs_{sched}"]["fused_levels"] = fused
        log(f"s27 bfs {sched}: {out[f'bfs_{sched}']}")
    if not args.no_bfs:
        out["bfs_learned_vs_plain"] = round(out["bfs_plain"]["ms"] / out["bfs_learned"]["ms"], 3)
        out["bfs_identical"] = bool(np.array_equal(bfs["plain"], bfs["learned"]))
    del bfs
    ctx.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, args.damping, root, args.cpu_budget)
        if not args.no_bfs:
            out["bfs_gpu_vs_cpu"] = round(cpu["bfs"]["ms"] / out["bfs_learned"]["ms"], 1)
    return g, out, cpu


def cpu_baseline(g, damping, root, budget_s):
    """The oracle's PageRank on the same graph with every host CPU this
    process may use, as many iterations as fit the budget (at least 1, at
    most 8): time per iteration -> GTEPS; and the oracle's BFS from the same
    root.  The oracle's layout build is not timed (the reference times only
    the iterations, source/main.c:153-160)."""
    o, build_s = oracle_of(g)
    _, st = o.pr(1, damping)
    per_iter = max(st["running_ms"], 1e-3) / 1e3
    iters = int(max(1, min(8, budget_s / per_iter)))
    _, st = o.pr(iters, damping)
    ms = st["running_ms"] / iters
    _, _, bst = o.bfs(root)
    threads = o.threads
    o.close()
    return {
        "value": round(g.num_edges / (ms * 1e-3) / 1e9, 4),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle PageRank on the same graph: {iters} iteration(s) on {threads} threads, "
                  f"{ms:.1f} ms per iteration (layout build {build_s:.1f} s not timed)",
        "host": host_cpu_info(),
        "bfs": {"ms": round(bst["running_ms"], 2), "iterations": int(bst["iterations"])},
    }


def pagerank_s20(damping_values=(0.85, 0.875), iterations=16, cpu=True):
    """R-MAT scale 20, edge factor 16, seed 1: 16 PageRank iterations per
    damping, GPU (autotuned shape) against the oracle, with the L1 error."""
    g = W.rmat(20, 16, seed=1)
    out = {"workload": "rmat-s20-ef16-seed1", "vertices": g.num_vertices, "edges": g.num_edges,
           "iterations": iterations}
    o = oracle_of(g)[0] if cpu else None
    with W.Context(0) as ctx:
        ctx.load(g)
        for d in damping_values:
            ctx.init_pr(d)
            shape, _ = ctx.pr_autotune(2)
            ms, _ = best_ms(lambda: ([ctx._lib.wg_pr_iteration(ctx.handle, None) for _ in range(iterations)],
                                     ctx.sync()), 1)
            ranks, _ = ctx.run_pr(iterations=iterations, damping=d)
            e = {"gpu_ms": round(ms, 3), "gpu_ms_per_iteration": round(ms / iterations, 4),
                 "gteps": round(iterations * g.num_edges / (ms * 1e-3) / 1e9, 2),
                 "shape": {"edges_per_lane": shape[0], "blocks_per_cu": shape[1], "kind": shape_name(shape)}}
            if o is not None:
                ref, st = o.pr(iterations, d)
                e.update({"cpu_ms": round(st["running_ms"], 2), "cpu_threads": o.threads,
                          "cpu_gteps": round(iterations * g.num_edges / (st["running_ms"] * 1e-3) / 1e9, 3),
                          "gpu_vs_cpu": round(st["running_ms"] / ms, 1),
                          "l1_vs_oracle": float(f"{float(np.abs(ranks - ref).sum()):.3e}")})
            out[f"d{d}"] = e
            log(f"PR s20 d={d}: {e}")
    if o is not None:
        o.close()
    return out


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--damping", type=float, default=0.85)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle PageRank on the headline graph")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle timings")
    ap.add_argument("--no-apps", action="store_true", help="skip the scale-24 and road-grid traversals")
    ap.add_argument("--no-grid", action="store_true", help="skip the 4096x6144 road grid")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sources and graphs")
    ap.add_argument("--no-s20", action="store_true", help="skip the scale-20 PageRank")
    ap.add_argument("--pr-shape", default=None,
                    help="PageRank shape edges_per_lane,blocks_per_cu,kind (kind 0 tiles, 1 tiles with XCD split, "
                         "2 binned) instead of the autotuner's choice")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-stored PageRank on the headline graph")
    ap.add_argument("--no-bfs", action="store_true", help="skip BFS on the headline graph")
    ap.add_argument("--no-pr", action="store_true", help="skip PageRank on the headline graph (profiling BFS alone)")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the headline graph (profiling the other sections alone; no headline line)")
    args = ap.parse_args()
    args.pr_shape = tuple(int(x) for x in args.pr_shape.split(",")) if args.pr_shape else None
    return args


def headline(args, copy_gbs):
    """The headline line (rmat-s27 PageRank) and its large-graph section, or
    None for a BFS-only profiling run (which prints that section alone)."""
    g27, large, cpu = large_graph(args)
    del g27
    if args.no_pr:
        print(json.dumps({"large": large}), flush=True)
        return None
    pr = large["pr_fp64"]
    out = {
        "metric": METRIC,
        "value": pr["gteps"],
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pr["ms_per_iteration"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"PageRank iteration (FP64 contributions, d={args.damping}) on rmat-s27-ef16-seed5",
            "scale": 27, "edge_factor": 16, "seed": 5,
            "vertices": large["vertices"], "edges": large["edges"], "parallelism": "single GPU",
            "pr_shape": pr["shape"],
        },
        "roofline": dict(pr["roofline"], measured_copy_GBs=copy_gbs),
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    out["large"] = large
    return out


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is designed but not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    copy_gbs = None
    with W.Context(0) as c:
        copy_gbs = round(c.measure_copy(4 << 30), 1)
    if args.no_large:  # a profiling run of the other sections: no headline line
        out = {"measured_copy_GBs": copy_gbs}
    else:
        out = headline(args, copy_gbs)
        if out is None:
            return
    if not args.no_s20:
        out["pagerank_s20"] = pagerank_s20(cpu=not args.no_cpu)

    if not args.no_apps:
        W.mlp_load()
        g = W.rmat(24, 16, seed=2, weights="f64")
        log(f"rmat s24 seed 2: V={g.num_vertices} E={g.num_edges}")
        ctx = W.Context(0)
        ctx.load(g)
        orc = None if args.no_cpu else oracle_of(g)[0]
        root = g.max_outdegree_vertex()
        out["apps"] = {"rmat_s24": schedule_pair(ctx, g, root, 3, 8, orc)}
        if orc is not None:
            orc.close()
        if not args.no_batch:
            ctx.init_pr(args.damping)
            ctx.pr_autotune(2)
            ms, _ = best_ms(lambda: (ctx.pr_iteration(), ctx.sync()), 5)
            out["batched"] = batched_sources(ctx, g, ms, args.damping)
        ctx.close()
        del g
        out["apps"]["rmat_s24_symmetrised_cc"] = cc_pair(24, 16, 2, cpu=not args.no_cpu)
        if not args.no_grid:
            road = W.grid(4096, 6144, seed=3)
            log(f"road grid 4096x6144 seed 3: V={road.num_vertices} E={road.num_edges}")
            ctx = W.Context(0)
            ctx.load(road)
            orc = None if args.no_cpu else oracle_of(road)[0]
            out["apps"]["road_grid_4096x6144"] = schedule_pair(ctx, road, road.max_outdegree_vertex(), 2, 4, orc)
            if orc is not None:
                orc.close()
            ctx.close()
    if not args.no_batch:
        out.setdefault("batched", {})["graphs_256"] = batched_graphs(args.damping)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
