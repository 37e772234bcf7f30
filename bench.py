"""Benchmark: pull PageRank on RMAT-24 (one MI355X).

One step = one PageRank solve of 20 fixed iterations (tol 0) on the RMAT-24
in-CSR already resident in HBM, scores reset to 1/V at the start of the step.
Prints ONE JSON line:
  value        GTEPS = E * iterations * steps / timed seconds / 1e9
  roofline     the fused pull kernel: algorithmic bytes per iteration
               (4E + 24V + 4) / its average launch time (HIP events on the
               stream it runs on), against 8 TB/s HBM
  cpu_baseline the oracle's OpenMP pull loop (restatement of
               src/pr/omp_base.cc) on the same in-CSR, a bounded number of
               iterations on the host cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 24]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PR GTEPS + % HBM peak at RMAT-24; batched PPR vs looped plain PR speedup"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(in_csr, out_deg, iters_per_pr, target_s):
    """Time the oracle's OpenMP pull loop (port of src/pr/omp_base.cc) for a
    bounded number of iterations; returns GTEPS and the sample description."""
    import numpy as np

    import oracle

    m = in_csr.m
    init = np.full(m, np.float32(1.0) / np.float32(m), np.float32)
    t0 = time.perf_counter()
    oracle.pr(m, in_csr.row_offsets, in_csr.column_indices, out_deg, init, tol=0.0, max_iter=1)
    one = time.perf_counter() - t0
    n = int(max(1, min(iters_per_pr, target_s // max(one, 1e-6))))
    t0 = time.perf_counter()
    oracle.pr(m, in_csr.row_offsets, in_csr.column_indices, out_deg, init, tol=0.0, max_iter=n)
    dt = time.perf_counter() - t0
    gteps = in_csr.nnz * n / dt / 1e9
    return {
        "value": round(gteps, 4), "unit": "GTEPS", "cores": oracle.num_threads(), "kind": "port",
        "sample": (f"oracle OpenMP pull loop (restates src/pr/omp_base.cc, g++ -O3 -fopenmp) on the "
                   f"same RMAT in-CSR, {n} iterations from 1/V, tol 0, {dt:.1f} s"),
    }


def traffic_from_profile(scale):
    path = os.path.join(ROOT, "profiles", f"pr_pull_rmat{scale}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--relabel", choices=["auto", "on", "off"], default="auto")
    ap.add_argument("--ipt", type=int, default=0, help="edges per thread per tile (8/16, 0=auto)")
    ap.add_argument("--nt", type=int, default=-1, help="non-temporal level 0/1/2 (-1 auto)")
    ap.add_argument("--persistent", choices=["auto", "on", "off"], default="auto")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: the PageRank/SpMV path here has no multi-GPU "
              "build (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gardenia_ocl_amd import PageRankPlan, graph

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    _, in_csr = graph.rmat(args.scale, args.edgefactor, seed=args.seed, want_out=False)
    V, E = in_csr.m, in_csr.nnz
    # out-degree = number of times a vertex appears as a source in the in-CSR
    out_deg = np.bincount(in_csr.column_indices, minlength=V).astype(np.int32)
    log(f"RMAT-{args.scale}: V={V} E={E} built in {time.perf_counter() - t0:.1f} s")
    off = torch.from_numpy(in_csr.row_offsets).to(dev)
    col = torch.from_numpy(in_csr.column_indices).to(dev)
    deg = torch.from_numpy(out_deg).to(dev)
    scores = torch.empty(V, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    relabel = {"auto": None, "on": True, "off": False}[args.relabel]
    plan = PageRankPlan(off, col, deg, stream=stream, relabel=relabel, items_per_thread=args.ipt,
                        nontemporal=None if args.nt < 0 else args.nt,
                        persistent={"auto": None, "on": True, "off": False}[args.persistent])
    info = plan.info()
    log(f"plan: {info}")
    init = float(np.float32(1.0) / np.float32(V))

    def step(profile):
        scores.fill_(init)
        st, _ = plan.run(scores, tol=0.0, max_iter=args.iters, profile=profile)
        assert st["iterations"] == args.iters, st
        return st

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    pull_ms, pull_n = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step(True)
        pull_ms += st["pull_ms"]
        pull_n += st["pull_launches"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {dt:.3f} s")

    gteps = E * args.iters * args.steps / dt / 1e9
    avg_pull_ms = pull_ms / max(1, pull_n)
    bytes_iter = 4 * E + 24 * V + 4
    achieved = bytes_iter / (avg_pull_ms * 1e-3) / 1e9
    roof = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(args.scale),
        "kernel": "pr_pull_kernel", "avg_launch_ms": round(avg_pull_ms, 5),
        "algorithmic_bytes_per_launch": bytes_iter,
    }
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(in_csr, out_deg, args.iters, args.cpu_seconds)
    rec = {
        "metric": METRIC, "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"RMAT-{args.scale} pull PageRank, {args.iters} fixed iterations",
                   "scale": args.scale, "edgefactor": args.edgefactor, "seed": args.seed,
                   "V": V, "E": E, "iterations_per_step": args.iters,
                   "abcd": [0.57, 0.19, 0.19, 0.05], "parallelism": "single-gpu",
                   "relabel": bool(info["relabel"]), "items_per_thread": info["items_per_thread"],
                   "nontemporal": info["nontemporal"], "persistent": bool(info["persistent"]),
                   "pull_grid": info["pull_grid"],
                   "plan_ms": round(info["plan_ms"], 1)},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
