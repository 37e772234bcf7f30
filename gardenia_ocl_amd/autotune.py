"""Autotuner for the plan-time knobs of the PageRank, batched-PPR and
many-graph plans.

For a given device-resident graph it times candidate configurations with the
library's own clock (HIP events around whole iterations in production mode:
hipGraph replay of the iteration chunks, no per-launch events), median of
``repeats`` runs after a warm-up, and keeps the fastest.  The search is a
coordinate descent from the library defaults (the first value of each knob's
list): one knob at a time over its values, adopting a value only when it is
faster than the current choice by more than ``margin`` (1%: below that the
difference is run-to-run noise), until a pass over all knobs changes nothing.
Every configuration gives the same scores bit for bit (the knobs change where
work runs, never the per-row order of the sums), so tuning is safe.

Results are cached in a JSON file keyed by the graph's shape (vertices,
edges, longest row, a hash of the log2 degree histogram); ``pagerank_config``
/ ``ppr_config`` / ``multigraph_config`` return the cached choice for a graph
(None when it was never tuned), which ``PageRankPlan(**cfg)``,
``PersonalizedPageRankPlan(**cfg)``, ``MultiGraphPageRank.run(chunk=...)``,
``PRSolver(config=cfg)`` and ``bench.py --tune-cache`` take as is.

Knobs (include/gardenia_hip.h):
  PageRank:     block (threads per row tile: its sort window), items_per_thread
                (tile size), hub_threshold (rows above it are whole-row exact
                hubs: the owner-thread / workgroup cut), sort_tiles, relabel,
                nontemporal (cache-hint level), persistent (pipelined grid)
  batched PPR:  k_chunk (columns per wavefront pass), hub_cols (columns per
                hub-row workgroup), relabel
  many graphs:  chunk (iterations per captured hipGraph between host checks)

Usage:
  python -m gardenia_ocl_amd.autotune --scale 24 [--ppr-k 256] [--multigraph 64]
      [--cache tune.json]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import time

import numpy as np

# first value of each list: the library default for a large graph
PR_SPACE = {
    "block": [256, 512, 1024],
    "items_per_thread": [16, 8],
    "hub_threshold": [0, 2048, 1024],
    "sort_tiles": [True, False],
    "relabel": [True, False],
    "nontemporal": [0, 1, 2],
    "persistent": [False, True],
    "stage_hubs": [-1, 0, 1],
}
PPR_SPACE = {"k_chunk": [256, 128, 64], "hub_cols": [64, 32, 16], "relabel": [True, False]}
MULTI_SPACE = {"chunk": [8, 4, 16, 32, 2]}


def graph_key(m, nnz, degrees) -> str:
    """Shape signature of a graph: sizes, longest row, degree-histogram hash."""
    d = np.asarray(degrees)
    hist = np.bincount(np.minimum(np.log2(np.maximum(d, 1)).astype(np.int64), 40), minlength=41)
    h = hashlib.sha1(hist.astype(np.int64).tobytes()).hexdigest()[:12]
    return f"m{m}-e{nnz}-max{int(d.max()) if d.size else 0}-{h}"


def _row_degrees(in_row_offsets):
    off = in_row_offsets.cpu().numpy() if hasattr(in_row_offsets, "cpu") else in_row_offsets
    off = np.asarray(off, np.int64)
    return off[1:] - off[:-1]


def pagerank_key(in_row_offsets, in_column_indices, out_degree) -> str:
    return "pr/" + graph_key(int(out_degree.shape[0]), int(in_column_indices.shape[0]),
                             _row_degrees(in_row_offsets))


def ppr_key(in_row_offsets, in_column_indices, out_degree, k) -> str:
    return f"ppr{k}/" + graph_key(int(out_degree.shape[0]), int(in_column_indices.shape[0]),
                                  _row_degrees(in_row_offsets))


def multigraph_key(graphs) -> str:
    sig = hashlib.sha1(";".join(
        graph_key(int(g[2].shape[0]), int(g[1].shape[0]), _row_degrees(g[0]))
        for g in graphs).encode()).hexdigest()[:16]
    return f"multi{len(graphs)}/{sig}"


def _load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _save(path, data):
    if path:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def _cached(cache, key):
    db = _load(cache)
    return db[key]["best"] if key in db else None


def pagerank_config(cache, in_row_offsets, in_column_indices, out_degree):
    """The tuned PageRankPlan keyword arguments for this graph, or None."""
    return _cached(cache, pagerank_key(in_row_offsets, in_column_indices, out_degree))


def ppr_config(cache, in_row_offsets, in_column_indices, out_degree, k):
    """The tuned PersonalizedPageRankPlan keyword arguments, or None."""
    return _cached(cache, ppr_key(in_row_offsets, in_column_indices, out_degree, k))


def multigraph_config(cache, graphs):
    """The tuned MultiGraphPageRank.run keyword arguments ({"chunk": c}), or None."""
    return _cached(cache, multigraph_key(graphs))


def descend(space, measure, repair=lambda cfg: cfg, margin=0.01, max_passes=3, log=None):
    """Coordinate descent over ``space`` from its first values.  ``measure(cfg)``
    returns a time (lower is better); ``repair(cfg)`` maps a candidate to the
    valid configuration it stands for (None: skip it).  Returns (best cfg, its
    time, {cfg json: time} of every configuration measured)."""
    names = list(space)
    cur = {n: space[n][0] for n in names}
    timings = {}

    def t(cfg):
        key = json.dumps(cfg, sort_keys=True)
        if key not in timings:
            timings[key] = float(measure(cfg))
            if log:
                log(f"{cfg}: {timings[key]:.4f}")
        return timings[key]

    best = t(cur)
    for _ in range(max_passes):
        changed = False
        for n in names:
            for v in space[n]:
                if v == cur[n]:
                    continue
                cand = repair(dict(cur, **{n: v}))
                if cand is None or cand == cur:
                    continue
                ms = t(cand)
                if ms < best * (1.0 - margin):
                    cur, best, changed = cand, ms, True
        if not changed:
            break
    return cur, best, timings


def _pr_repair(cfg):
    if cfg["hub_threshold"] > 256 * cfg["items_per_thread"]:
        return None
    if cfg["persistent"] and cfg["block"] != 256:  # the persistent grid is 256-wide
        return dict(cfg, block=256)
    return cfg


def _time_pr(off, col, deg, cfg, iters, repeats):
    import torch

    from .pr import PageRankPlan

    plan = PageRankPlan(off, col, deg, **cfg)
    m = plan.m
    s = torch.empty(m, dtype=torch.float32, device=off.device)
    s.fill_(1.0 / m)
    plan.run(s, tol=0.0, max_iter=2)
    times = []
    for _ in range(repeats):
        s.fill_(1.0 / m)
        st, _ = plan.run(s, tol=0.0, max_iter=iters)
        times.append(st["solve_ms"] / max(1, st["iterations"]))
    plan.close()
    return float(np.median(times))


def tune_pagerank(in_row_offsets, in_column_indices, out_degree, *, iters=10, repeats=3,
                  cache=None, space=None, log=None):
    """Returns (best config dict, {config-as-json: ms per iteration})."""
    key = pagerank_key(in_row_offsets, in_column_indices, out_degree)
    db = _load(cache)
    if key in db:
        return db[key]["best"], db[key]["timings"]
    best, _, timings = descend(
        space or PR_SPACE,
        lambda cfg: _time_pr(in_row_offsets, in_column_indices, out_degree, cfg, iters, repeats),
        _pr_repair, log=log)
    db[key] = {"best": best, "timings": timings, "tuned_at": time.time(),
               "unit": "ms per iteration"}
    _save(cache, db)
    return best, timings


def tune_ppr(in_row_offsets, in_column_indices, out_degree, k, *, iters=3, repeats=2,
             cache=None, space=None, log=None):
    """Best batched-PPR plan config for k seeds (ms per iteration of `iters`)."""
    import torch

    from .ppr import PersonalizedPageRankPlan

    key = ppr_key(in_row_offsets, in_column_indices, out_degree, k)
    db = _load(cache)
    if key in db:
        return db[key]["best"], db[key]["timings"]
    m = int(out_degree.shape[0])
    seeds = torch.from_numpy(np.random.default_rng(0).integers(0, m, k).astype(np.int32)).to(
        in_row_offsets.device)

    def measure(cfg):
        plan = PersonalizedPageRankPlan(in_row_offsets, in_column_indices, out_degree, k, **cfg)
        X = plan.one_hot(seeds)
        plan.run(seeds, X, tol=0.0, max_iter=1)
        times = []
        for _ in range(repeats):
            X = plan.one_hot(seeds)
            st, _, _ = plan.run(seeds, X, tol=0.0, max_iter=iters)
            times.append(st["solve_ms"] / iters)
        plan.close()
        del X
        return float(np.median(times))

    space = dict(space or PPR_SPACE)
    space["k_chunk"] = [c for c in space["k_chunk"] if c and k % c == 0] or [0]  # 0: auto
    best, _, timings = descend(space, measure, log=log)
    db[key] = {"best": best, "timings": timings, "tuned_at": time.time(),
               "unit": "ms per iteration"}
    _save(cache, db)
    return best, timings


def tune_multigraph(graphs, *, tol=1e-4, max_iter=100, repeats=3, cache=None, space=None,
                    log=None):
    """Best iterations-per-chunk for a set of graphs run together (ms per solve)."""
    import torch

    from .multigraph import MultiGraphPageRank

    key = multigraph_key(graphs)
    db = _load(cache)
    if key in db:
        return db[key]["best"], db[key]["timings"]
    mg = MultiGraphPageRank(graphs)
    dev = graphs[0][0].device
    scores = [torch.empty(int(g[2].shape[0]), dtype=torch.float32, device=dev) for g in graphs]

    def measure(cfg):
        times = []
        for r in range(repeats + 1):
            for s in scores:
                s.fill_(1.0 / s.shape[0])
            _, _, ms = mg.run(scores, tol=tol, max_iter=max_iter, chunk=cfg["chunk"])
            if r:
                times.append(ms)
        return float(np.median(times))

    best, _, timings = descend(space or MULTI_SPACE, measure, log=log)
    mg.close()
    db[key] = {"best": best, "timings": timings, "tuned_at": time.time(),
               "unit": "ms per solve"}
    _save(cache, db)
    return best, timings


def main():
    import torch

    from . import graph

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--ppr-k", type=int, default=0)
    ap.add_argument("--ppr-scale", type=int, default=22)
    ap.add_argument("--multigraph", type=int, default=0, help="graphs of scales 12-18")
    ap.add_argument("--cache", default="")
    args = ap.parse_args()
    dev = torch.device("cuda:0")
    log = lambda s: print(s, flush=True)  # noqa: E731
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)  # noqa: E731
    out = {}
    oc, ic = graph.rmat(args.scale, args.edgefactor, seed=args.seed)
    off, col, deg = t(ic.row_offsets), t(ic.column_indices), t(oc.degree)
    best, timings = tune_pagerank(off, col, deg, cache=args.cache or None, log=log)
    out["pagerank"] = {"graph": f"RMAT-{args.scale}", "best": best, "timings_ms": timings}
    del off, col, deg, oc, ic
    if args.ppr_k:
        oc, ic = graph.rmat(args.ppr_scale, args.edgefactor, seed=args.seed)
        off, col, deg = t(ic.row_offsets), t(ic.column_indices), t(oc.degree)
        best, timings = tune_ppr(off, col, deg, args.ppr_k, cache=args.cache or None, log=log)
        out["ppr"] = {"graph": f"RMAT-{args.ppr_scale}", "k": args.ppr_k, "best": best,
                      "timings_ms": timings}
        del off, col, deg, oc, ic
    if args.multigraph:  # the graphs bench.py's multigraph leg builds for this seed
        graphs = []
        scales = np.random.default_rng(args.seed + 5).integers(12, 19, args.multigraph)
        for i, sc in enumerate(scales):
            _, c = graph.rmat(int(sc), 16, seed=args.seed + 100 + i, want_out=False)
            od = np.bincount(c.column_indices, minlength=c.m).astype(np.int32)
            graphs.append((t(c.row_offsets), t(c.column_indices), t(od)))
        best, timings = tune_multigraph(graphs, cache=args.cache or None, log=log)
        out["multigraph"] = {"graphs": args.multigraph, "best": best, "timings_ms": timings}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
