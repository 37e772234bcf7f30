"""Autotuner for the plan-time knobs of the PageRank and batched-PPR plans.

For a given device-resident graph it builds each candidate plan, times a fixed
number of iterations with HIP events (median of ``repeats`` runs, after a
warm-up), and returns the fastest configuration.  Results are cached in a JSON
file keyed by the graph's shape (vertices, edges, longest row, a hash of the
degree histogram), so a repeated workload is tuned once.

Knobs (see include/gardenia_hip.h):
  PageRank:     relabel (L2-resident hot sources), items_per_thread (tile size,
                i.e. the hub-row threshold 256*ipt), nontemporal (cache-hint
                level), persistent (pipelined persistent grid), sort_tiles (gather
                each tile in source order)
  batched PPR:  k_chunk (columns per wavefront pass), relabel

Usage:
  python -m gardenia_ocl_amd.autotune --scale 22 [--ppr-k 256] [--cache tune.json]
"""
from __future__ import annotations

import argparse
import hashlib
import itertools
import json
import os
import time

import numpy as np

PR_SPACE = {
    "relabel": [False, True],
    "items_per_thread": [8, 16],
    "nontemporal": [0, 1],
    "persistent": [False, True],
    "sort_tiles": [False, True],
}
PPR_SPACE = {"k_chunk": [64, 128, 256], "relabel": [False, True]}


def graph_key(m, nnz, degrees) -> str:
    """Shape signature of a graph: sizes, longest row, degree-histogram hash."""
    d = np.asarray(degrees)
    hist = np.bincount(np.minimum(np.log2(np.maximum(d, 1)).astype(np.int64), 40), minlength=41)
    h = hashlib.sha1(hist.astype(np.int64).tobytes()).hexdigest()[:12]
    return f"m{m}-e{nnz}-max{int(d.max()) if d.size else 0}-{h}"


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
        st, _ = plan.run(s, tol=0.0, max_iter=iters, profile=True)
        times.append(st["pull_ms"] / max(1, st["pull_launches"]))
    plan.close()
    return float(np.median(times))


def tune_pagerank(in_row_offsets, in_column_indices, out_degree, *, iters=5, repeats=3,
                  cache=None, space=None, log=None):
    """Returns (best config dict, {config-as-json: ms per pull launch})."""
    key = "pr/" + graph_key(int(out_degree.shape[0]), int(in_column_indices.shape[0]),
                            (in_row_offsets[1:] - in_row_offsets[:-1]).cpu().numpy())
    db = _load(cache)
    if key in db:
        return db[key]["best"], db[key]["timings"]
    space = space or PR_SPACE
    names = list(space)
    timings = {}
    for vals in itertools.product(*(space[n] for n in names)):
        cfg = dict(zip(names, vals))
        ms = _time_pr(in_row_offsets, in_column_indices, out_degree, cfg, iters, repeats)
        timings[json.dumps(cfg, sort_keys=True)] = ms
        if log:
            log(f"pr {cfg}: {ms:.4f} ms")
    best = json.loads(min(timings, key=timings.get))
    db[key] = {"best": best, "timings": timings, "tuned_at": time.time()}
    _save(cache, db)
    return best, timings


def tune_ppr(in_row_offsets, in_column_indices, out_degree, k, *, iters=3, cache=None,
             space=None, log=None):
    """Best batched-PPR plan config for k seeds (times `iters` iterations)."""
    import torch

    from .ppr import PersonalizedPageRankPlan

    key = f"ppr{k}/" + graph_key(int(out_degree.shape[0]), int(in_column_indices.shape[0]),
                                 (in_row_offsets[1:] - in_row_offsets[:-1]).cpu().numpy())
    db = _load(cache)
    if key in db:
        return db[key]["best"], db[key]["timings"]
    space = space or PPR_SPACE
    names = list(space)
    m = int(out_degree.shape[0])
    seeds = torch.from_numpy(np.random.default_rng(0).integers(0, m, k).astype(np.int32)).to(
        in_row_offsets.device)
    timings = {}
    for vals in itertools.product(*(space[n] for n in names)):
        cfg = dict(zip(names, vals))
        if k % cfg["k_chunk"]:
            continue
        plan = PersonalizedPageRankPlan(in_row_offsets, in_column_indices, out_degree, k, **cfg)
        X = plan.one_hot(seeds)
        plan.run(seeds, X, tol=0.0, max_iter=1)
        X = plan.one_hot(seeds)
        st, _, _ = plan.run(seeds, X, tol=0.0, max_iter=iters)
        ms = st["solve_ms"] / iters
        timings[json.dumps(cfg, sort_keys=True)] = ms
        plan.close()
        del X
        if log:
            log(f"ppr {cfg}: {ms:.3f} ms/iteration")
    best = json.loads(min(timings, key=timings.get))
    db[key] = {"best": best, "timings": timings, "tuned_at": time.time()}
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
    ap.add_argument("--cache", default="")
    args = ap.parse_args()
    dev = torch.device("cuda:0")
    _, ic = graph.rmat(args.scale, args.edgefactor, seed=args.seed, want_out=False)
    od = np.bincount(ic.column_indices, minlength=ic.m).astype(np.int32)
    off = torch.from_numpy(ic.row_offsets).to(dev)
    col = torch.from_numpy(ic.column_indices).to(dev)
    deg = torch.from_numpy(od).to(dev)
    out = {}
    best, timings = tune_pagerank(off, col, deg, cache=args.cache or None, log=print)
    out["pagerank"] = {"best": best, "timings_ms": timings}
    if args.ppr_k:
        best, timings = tune_ppr(off, col, deg, args.ppr_k, cache=args.cache or None, log=print)
        out["ppr"] = {"best": best, "timings_ms": timings}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
