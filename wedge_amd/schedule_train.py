"""Trains the learned schedule from the engine's own per-iteration traces.

For every training graph and application, the converging loop runs once per
fixed action (dense pull; frontier pull at block_log2 3, 4, 5, 6, 8),
recording per-iteration rows: the state before the pull (frontier vertices,
frontier out-edges, vertices not reached yet, iteration) and the device time
of the pull and of the frontier build it consumed.  Each row is one sample:
(features, action, log2 of the pull's own device time minus the log2 work
estimate of that action, wg_mlp_baseline).  The network learns time per unit
of estimated work; the estimate carries the graph size, so graphs larger than
the training graphs are priced sensibly.  Per-pull time is the right target
for BFS, whose levels are the same in every mode; CC and SSSP, whose
iteration count depends on the granularity, may only choose between the dense
pull and the default granularity (include/wedge_gpu.h).  With --target togo
the sample is instead the time the fixed-action run still took to converge.
The model (2x64 ReLU MLP, FP32) is trained on the GPU with the hand-written
kernels behind wg_mlp_train and written as a model file.  Training graphs: R-MAT scales 16-22 and grids
256^2-2048^2; the 4096x6144 road grid used for evaluation is not among them.

    python -m wedge_amd.schedule_train --out wedge_amd/data/schedule_mlp.txt
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

import wedge_amd as W

ACTIONS = [dict(schedule="dense")] + [dict(schedule="frontier", block_log2=b) for b in (3, 4, 5, 6, 8)]
ROWS_PER_RUN = 400
# log2 margins (BFS, CC, SSSP) by which the model must beat the plain rule's
# choice: small for BFS, whose per-pull cost is the whole story; larger for CC
# and SSSP, whose dense-or-frontier choice also changes how many pulls follow
MARGINS = (0.2, 0.5, 0.5)


def log(msg):
    print(f"[train] {msg}", flush=True)


def action_of(mode: int, blog2: int) -> int:
    if mode == W.MODE_DENSE:
        return 0
    return {3: 1, 4: 2, 5: 3, 6: 4, 8: 5}[blog2]


def training_graphs(quick: bool):
    """(name, graph, apps) triples."""
    out = []
    rmat_scales = (14, 16) if quick else (16, 18, 20, 22)
    for i, s in enumerate(rmat_scales):
        out.append((f"rmat{s}w", W.rmat(s, 16, seed=40 + i, weights="f64"), ("bfs", "sssp")))
        out.append((f"rmat{s}sym", W.rmat(s, 16, seed=50 + i, symmetrize=True), ("cc", "bfs")))
    sides = (128, 256) if quick else (256, 512, 1024, 2048)
    for i, n in enumerate(sides):
        out.append((f"grid{n}", W.grid(n, n, seed=60 + i), ("bfs", "sssp", "cc")))
    return out


def roots_of(g, k: int, seed: int):
    """The max-out-degree vertex plus k-1 random vertices with out-edges."""
    od = g.outdegrees()
    cand = np.nonzero(od > 0)[0]
    rng = np.random.default_rng(seed)
    extra = rng.choice(cand, size=min(k - 1, cand.size), replace=False) if cand.size else []
    return [g.max_outdegree_vertex()] + [int(v) for v in extra]


def run_app(ctx, app: str, root: int, **kw):
    if app == "bfs":
        return ctx.run_bfs(root, **kw)
    if app == "sssp":
        return ctx.run_sssp(root, **kw)
    return ctx.run_cc(**kw)


APP_ID = {"bfs": W.APP_BFS, "cc": W.APP_CC, "sssp": W.APP_SSSP}


def samples_from_trace(rows, app: str, V: int, E: int):
    """Per kept row: features, action, the pull's own time and the time to go (ms)."""
    cost = np.zeros(len(rows))
    prev_build = 0.0
    for i, r in enumerate(rows):
        cost[i] = r["gpu_pull_ms"] + (prev_build if r["mode"] == W.MODE_FRONTIER else 0.0)
        prev_build = r["gpu_build_ms"]
    togo = np.cumsum(cost[::-1])[::-1]
    idx = np.arange(len(rows))
    if len(rows) > ROWS_PER_RUN:
        idx = np.unique(np.linspace(0, len(rows) - 1, ROWS_PER_RUN).astype(int))
    X = [W.mlp_features(APP_ID[app], rows[i]["iteration"], V, E, rows[i]["frontier_vertices"],
                        rows[i]["frontier_edges"], rows[i]["unvisited"]) for i in idx]
    A = [action_of(rows[i]["mode"], rows[i]["block_log2"]) for i in idx]
    Bs = [W.mlp_base(V, E, rows[i]["frontier_edges"])[a] for i, a in zip(idx, A)]
    raw = [(APP_ID[app], rows[i]["iteration"], V, E, rows[i]["frontier_vertices"], rows[i]["frontier_edges"],
            rows[i]["unvisited"]) for i in idx]
    return X, A, cost[idx].tolist(), togo[idx].tolist(), Bs, raw


def collect(ctx, graphs, roots_per_graph: int):
    """Samples of every fixed-action run: features, action, step and to-go times, run number."""
    X, A, S, G, R, B, RAW = [], [], [], [], [], [], []
    run = 0
    for name, g, apps in graphs:
        ctx.load(g)
        V, E = g.num_vertices, g.num_edges
        for app in apps:
            roots = roots_of(g, roots_per_graph if app != "cc" else 1, seed=V)
            for root in roots:
                for act in ACTIONS:
                    t0 = time.time()
                    run_app(ctx, app, root, record_trace=1, **act)
                    rows = W.trace_rows()
                    x, a, st, tg, bs, raw = samples_from_trace(rows, app, V, E)
                    RAW += raw
                    X += x
                    A += a
                    S += st
                    G += tg
                    B += bs
                    R += [run] * len(x)
                    run += 1
                    log(f"{name} {app} root {root} {act}: {len(rows)} rows, {sum(r['gpu_pull_ms'] for r in rows):.2f} ms "
                        f"pulls, {len(x)} samples ({time.time() - t0:.1f} s)")
    return (np.array(X, np.float32), np.array(A, np.int32), np.array(S, np.float64), np.array(G, np.float64),
            np.array(R, np.int32), np.array(B, np.float32), np.array(RAW, np.uint64))


def features_of(raw: np.ndarray, A: np.ndarray):
    """Features and work estimates of the taken actions from raw loop states
    (app, iteration, V, E, frontier vertices, frontier edges, not reached)."""
    X = np.stack([W.mlp_features(*(int(v) for v in r)) for r in raw])
    Bs = np.array([W.mlp_base(int(r[2]), int(r[3]), int(r[5]))[a] for r, a in zip(raw, A)], np.float32)
    return X, Bs


def init_params(X: np.ndarray, T: np.ndarray, seed: int, lv_max: float) -> np.ndarray:
    """Input range and normalisation from the data; He-uniform weights; output
    bias at the mean target."""
    I, H, A = W.MLP_IN, W.MLP_HIDDEN, W.MLP_ACTIONS
    rng = np.random.default_rng(seed)
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    istd = np.where(std > 1e-3, 1.0 / np.maximum(std, 1e-3), 0.0)
    parts = [X.min(axis=0), X.max(axis=0), mean, istd,
             rng.uniform(-1, 1, (I, H)) * np.sqrt(6.0 / I), np.zeros(H),
             rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H), np.zeros(H),
             rng.uniform(-1, 1, (H, A)) * np.sqrt(6.0 / H) * 0.1, np.full(A, T.mean()),
             np.array(MARGINS), np.array([lv_max])]
    p = np.concatenate([np.ravel(q) for q in parts]).astype(np.float32)
    assert p.size == W.MLP_PARAMS
    return p


def eval_graphs(spec: str):
    """Held-out evaluation graphs, comma separated: 'grid:RxC[:seed]' (integer
    weights) or 'rmat:SCALE[:seed]' (edge factor 16, FP64 weights)."""
    out = []
    for item in filter(None, spec.split(",")):
        kind, dims, *rest = item.split(":")
        seed = int(rest[0]) if rest else 3
        if kind == "grid":
            r, c = (int(v) for v in dims.split("x"))
            out.append((f"grid{r}x{c}", W.grid(r, c, seed=seed), ("bfs", "sssp")))
        elif kind == "rmat":
            out.append((f"rmat{dims}", W.rmat(int(dims), 16, seed=seed, weights="f64"), ("bfs", "sssp")))
        else:
            raise SystemExit(f"unknown evaluation graph {item}")
    return out


def regret_table(ctx, graphs, params):
    """Plain vs learned on each training graph (device loop), outputs compared."""
    W.mlp_set(params)
    for name, g, apps in graphs:
        ctx.load(g)
        for app in apps:
            root = g.max_outdegree_vertex()
            out_p, st_p = run_app(ctx, app, root, schedule="plain")
            out_p, st_p = run_app(ctx, app, root, schedule="plain")
            out_l, st_l = run_app(ctx, app, root, schedule="learned")
            out_l, st_l = run_app(ctx, app, root, schedule="learned")
            same = np.array_equal(out_p, out_l)
            log(f"eval {name} {app}: plain {st_p['running_ms']:.3f} ms, learned {st_l['running_ms']:.3f} ms "
                f"(x{st_p['running_ms'] / max(st_l['running_ms'], 1e-9):.2f}), frontier iters "
                f"{st_p['frontier_iterations']}/{st_l['frontier_iterations']} of {st_p['iterations']}/"
                f"{st_l['iterations']}, identical outputs: {same}")
            if not same:
                raise SystemExit(f"learned schedule changed the {app} result on {name}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=W.DEFAULT_MODEL_PATH)
    ap.add_argument("--dataset", default=None, help="also write the samples (npz)")
    ap.add_argument("--from-dataset", default=None, help="train from saved samples instead of recording")
    ap.add_argument("--quick", action="store_true", help="small graphs only (smoke test)")
    ap.add_argument("--roots", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--target", choices=("step", "togo"), default="step")
    ap.add_argument("--no-eval", action="store_true")
    ap.add_argument("--eval-train", action="store_true", help="with --from-dataset: also evaluate on the training graphs")
    ap.add_argument("--eval", default="", help="held-out graphs to evaluate on, e.g. grid:4096x6144:3")
    args = ap.parse_args(argv)

    ctx = W.Context(0)
    graphs = None
    if args.from_dataset:
        z = np.load(args.from_dataset, allow_pickle=False)
        A, S, G, raw = z["A"], z["step_ms"], z["togo_ms"], z["raw"]
        X, Bs = features_of(raw, A)  # recomputed: the feature definition may have changed
        if args.eval_train:
            graphs = training_graphs(args.quick)
    else:
        t0 = time.time()
        graphs = training_graphs(args.quick)
        log(f"{len(graphs)} training graphs generated ({time.time() - t0:.1f} s)")
        X, A, S, G, R, Bs, raw = collect(ctx, graphs, args.roots)
        if args.dataset:
            np.savez_compressed(args.dataset, X=X, A=A, step_ms=S, togo_ms=G, run=R, base=Bs, raw=raw)
    T = (np.log2(np.maximum(G if args.target == "togo" else S, 1e-4)) - Bs).astype(np.float32)
    log(f"{X.shape[0]} samples; per action: {np.bincount(A, minlength=W.MLP_ACTIONS).tolist()}")
    lv_max = float(np.log2(1.0 + float(raw[:, 2].max()))) + 0.5  # half an octave above the largest training graph
    p0 = init_params(X, T, args.seed, lv_max)
    t0 = time.time()
    params, loss = ctx.mlp_train(X, A, T, p0, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    var = float(np.var(T))
    log(f"trained in {time.time() - t0:.1f} s: mean squared error {loss:.4f} (target variance {var:.4f})")
    W.mlp_set(params)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    W.mlp_save(args.out)
    log(f"model written to {args.out}")
    if graphs is not None and not args.no_eval:
        regret_table(ctx, graphs, params)
    if args.eval:
        regret_table(ctx, eval_graphs(args.eval), params)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
