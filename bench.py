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
# the autotuner's choices for the workloads below on the MI355X
# (python -m gardenia_ocl_amd.autotune --scale 24 --ppr-k 256 --multigraph 64)
TUNED = os.path.join(ROOT, "gardenia_ocl_amd", "tuned_mi355x.json")
sys.path.insert(0, ROOT)

METRIC = "PR GTEPS + % HBM peak at RMAT-24; batched PPR vs looped plain PR speedup"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """The host CPU model and the logical CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return f"{model}, {n} logical CPUs available to the process"


def cpu_baseline(in_csr, out_deg, iters_per_pr, target_s):
    """The CPU baseline on the same in-CSR.  Where oracle/_ref/libref.so was
    built (the reference's own src/pr/omp_base.cc compiled from source), the
    reference's PRSolver itself: solves from 1/V to its compile-time EPSILON
    1e-3 (MAX_ITER 30), repeated for about target_s seconds, GTEPS from the
    iterations and runtime it prints (its own Timer, include/timer.h), median
    over the solves.  The oracle's OpenMP pull loop (the restatement) is timed
    beside it for a bounded number of fixed iterations."""
    port = cpu_baseline_port(in_csr, out_deg, iters_per_pr, target_s)
    import oracle

    if not oracle.ref_available():
        return port
    import statistics
    import tempfile

    import numpy as np

    m, E = in_csr.m, in_csr.nnz
    init = np.float32(1.0) / np.float32(m)
    runs, total = [], 0.0
    with tempfile.TemporaryDirectory() as tmp:
        log_path = os.path.join(tmp, "pr.log")
        while total < target_s and len(runs) < 10:
            sc = np.full(m, init, np.float32)
            t0 = time.perf_counter()
            its, ms, _ = oracle.ref_pr_solve(m, E, in_csr.row_offsets, in_csr.column_indices,
                                             out_deg, sc, log_path)
            total += time.perf_counter() - t0
            runs.append((its, ms))
    its = runs[0][0]
    gteps = statistics.median(E * i / (ms * 1e-3) / 1e9 for i, ms in runs)
    return {
        "value": round(gteps, 4), "unit": "GTEPS", "cores": oracle.num_threads(),
        "kind": "reference",
        "sample": (f"the reference's own OpenMP PRSolver (src/pr/omp_base.cc compiled from source, "
                   f"g++ -O3 -fopenmp) on the same RMAT in-CSR from 1/V: {its} iterations to its "
                   f"EPSILON 1e-3, runtime from its own Timer, median of {len(runs)} solves "
                   f"({total:.1f} s)"),
        "host": host_cpu(), "port": port,
    }


def cpu_baseline_port(in_csr, out_deg, iters_per_pr, target_s):
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
        "host": host_cpu(),
    }


def tol_parity(in_csr, out_deg, tol, max_iter, scores, errs, x64, rec):
    """The oracle's run to the same tolerance on the same in-CSR, beside the
    GPU's: iteration counts, the L1 values at the deciding iterations, the
    largest per-iteration L1 difference, and whether the oracle's own L1 lies
    within that difference of tol at any iteration up to its stop (only then
    could two correct evaluations stop at different counts)."""
    import numpy as np

    import oracle

    t0 = time.perf_counter()
    ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices, out_deg,
                                  tol=tol, max_iter=max_iter)
    dt = time.perf_counter() - t0
    g = rec["iterations"]
    n = min(len(errs), len(ref_errs))
    gap = float(np.max(np.abs(np.asarray(errs[:n]) - ref_errs[:n]))) if n else 0.0
    tie = bool(np.any(np.abs(ref_errs[:it] - tol) <= gap))
    rel = np.abs(scores.astype(np.float64) - ref) / np.maximum(np.abs(ref.astype(np.float64)),
                                                               1e-300)
    out = {"oracle_iterations": int(it), "iterations_match": g == it,
           "l1_at_stop": {"gpu": float(errs[g - 1]), "oracle": float(ref_errs[it - 1])},
           "l1_before_stop": {"gpu": float(errs[g - 2]) if g > 1 else None,
                              "oracle": float(ref_errs[it - 2]) if it > 1 else None},
           "max_l1_gap": gap, "oracle_l1_within_gap_of_tol": tie,
           "max_rel_vs_oracle": float(rel.max()), "oracle_solve_s": round(dt, 3)}
    if x64 is not None:
        out["oracle_l1_vs_fp64"] = float(np.abs(ref.astype(np.float64) - x64).sum())
    return out


def cpu_baseline_extras(spmv_log2, seed, verify=None, checks=None, extras=None):
    """The oracle on the secondary workloads: SpMV y += A x on the ER matrix
    (OpenMP, src/spmv/omp_base.cc restated), one SymGS sweep on the 128^3
    stencil (sequential, src/symgs/verifier.cc restated), the RMAT-16 solve,
    and (verify = (out-CSR, scores)) the PRVerifier residual (sequential,
    src/pr/verifier.cc:50-68 restated)."""
    import numpy as np

    import oracle
    from gardenia_ocl_amd import graph

    out = {}
    er = graph.er_matrix(spmv_log2, 32, seed=seed)
    x = graph.uniform_pm1(er.m, seed + 1)
    y = np.zeros(er.m, np.float32)
    y1 = oracle.spmv(er.m, er.row_offsets, er.column_indices, er.weight, x, y)
    if checks and "spmv" in checks and extras and "spmv" in extras:
        # the GPU's y = A x from y0 = 0 (both modes gave the same bits)
        extras["spmv"]["bit_exact_vs_oracle"] = bool(np.array_equal(checks["spmv"], y1))
    del y1
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.spmv(er.m, er.row_offsets, er.column_indices, er.weight, x, y)
    ms = (time.perf_counter() - t0) / reps * 1e3
    out["spmv"] = {"value": round(ms, 3), "unit": "ms per y += A x", "cores": oracle.num_threads(),
                   "kind": "port", "sample": f"ER 2^{spmv_log2} x 32, {reps} reps"}
    del er, x, y
    c = graph.stencil27(128)
    colors, nc = graph.color_first_fit(c)
    idx, co = graph.color_order(colors, nc)
    b = graph.uniform_pm1(c.m, seed + 17)
    t0 = time.perf_counter()
    oracle.symgs(c.row_offsets, c.column_indices, idx, c.weight, np.zeros(c.m, np.float32), b, co)
    ms = (time.perf_counter() - t0) * 1e3
    out["symgs"] = {"value": round(ms, 1), "unit": "ms per symmetric sweep", "cores": 1,
                    "kind": "port", "sample": "stencil27 128^3, 8 colors, one sweep"}
    del c, colors, idx, co, b
    # the small-graph workload: RMAT-16 to L1 tol 1e-4 (the OpenMP pull loop)
    o16, i16 = graph.rmat(16, 16, seed=seed)
    best, iters = float("inf"), 0
    for _ in range(10):
        t0 = time.perf_counter()
        _, iters, _ = oracle.pr(i16.m, i16.row_offsets, i16.column_indices, o16.degree,
                                tol=1e-4, max_iter=100)
        best = min(best, time.perf_counter() - t0)
    if checks and "small" in checks and extras and "small" in extras:
        ref, _, ref_errs = oracle.pr(i16.m, i16.row_offsets, i16.column_indices, o16.degree,
                                     tol=1e-4, max_iter=100)
        gs, gerrs, gits = checks["small"]
        n = min(len(gerrs), len(ref_errs))
        extras["small"].update({
            "oracle_iterations": iters, "iterations_match": gits == iters,
            "max_l1_gap": float(np.max(np.abs(gerrs[:n] - ref_errs[:n]))) if n else 0.0,
            "max_rel_vs_oracle": float(np.max(np.abs(gs - ref) / ref))})
    out["pr_rmat16"] = {"value": round(best * 1e3, 3), "unit": "ms per solve",
                        "us_per_iteration": round(best * 1e6 / max(1, iters), 2),
                        "iterations": iters, "cores": oracle.num_threads(), "kind": "port",
                        "sample": "RMAT-16 pull PageRank to L1 tol 1e-4, best of 10 solves"}
    if verify is not None:
        oc, sc = verify
        t0 = time.perf_counter()
        oracle.pr_residual(oc.m, oc.row_offsets, oc.column_indices, oc.degree, sc)
        out["verify"] = {"value": round((time.perf_counter() - t0) * 1e3, 1), "unit": "ms",
                         "cores": 1, "kind": "port",
                         "sample": "PRVerifier push residual over the RMAT out-CSR, one call"}
    return out


def bench_spmv(log2_rows, nnz_per_row, reps, seed):
    """ER 2^n x 2^n, nnz_per_row uniform columns, y += A x repeated `reps` times
    (HIP events around the reps on the plan's stream)."""
    import torch

    from gardenia_ocl_amd import SpmvPlan, graph

    dev = torch.device("cuda:0")
    er = graph.er_matrix(log2_rows, nnz_per_row, seed=seed)
    Ap = torch.from_numpy(er.row_offsets).to(dev)
    Aj = torch.from_numpy(er.column_indices).to(dev)
    Ax = torch.from_numpy(er.weight).to(dev)
    x = torch.from_numpy(graph.uniform_pm1(er.m, seed + 1)).to(dev)
    y = torch.zeros(er.m, dtype=torch.float32, device=dev)
    nnz, M = er.nnz, er.m
    byts = 8 * nnz + 16 * M + 4
    st = torch.cuda.current_stream()
    out = {"workload": f"ER 2^{log2_rows} x 2^{log2_rows}, {nnz_per_row} nnz/row, y += A x",
           "nnz": nnz, "reps": reps, "algorithmic_bytes": byts}
    # y = A x in FP64 (torch CSR product) for the L1 figure
    A64 = torch.sparse_csr_tensor(Ap, Aj, Ax.double(), (M, M))
    y64 = (A64 @ x.double().unsqueeze(1)).squeeze(1)
    del A64
    firsts = {}
    for mode in ("auto", "tiles"):
        plan = SpmvPlan(Ap, Aj, Ax, mode=mode)
        y.zero_()
        plan.run(x, y)  # y = A x from y0 = 0: checked against FP64 here, the oracle later
        firsts[mode] = y.cpu().numpy()
        for _ in range(2):
            plan.run(x, y)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(reps):
            plan.run(x, y)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        key = "" if mode == "auto" else "tiles_"
        out[key + "mode"] = plan.mode
        out[key + "ms_per_spmv"] = round(ms, 5)
        out[key + "gflops"] = round(2 * nnz / (ms * 1e-3) / 1e9, 2)
        out[key + "achieved_gbs"] = round(byts / (ms * 1e-3) / 1e9, 1)
        out[key + "frac_of_hbm_peak"] = round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        yf = torch.from_numpy(firsts[mode]).to(dev).double()
        out[key + "l1_vs_fp64"] = float((yf - y64).abs().sum())
        out[key + "max_abs_vs_fp64"] = float((yf - y64).abs().max())
        plan.close()
    out["both_modes_identical"] = bool((firsts["auto"] == firsts["tiles"]).all())
    return out, firsts["auto"]


def fp64_pagerank(off, col, outdeg, *, damping=0.85, seeds=None, tol=1e-13, max_iter=1000):
    """Independent FP64 reference on the device: power iteration
    x <- base + d * A x with torch's CSR SpMV in float64, A[v, u] = 1/outdeg[u]
    over the in-CSR (dangling mass dropped, as in src/pr/omp_base.cc).
    seeds=None: uniform teleport (1-d)/V from x = 1/V; else one column per
    seed, teleport (1-d) to the seed from the one-hot start.  Returns
    (x [V] or [V, k], iterations)."""
    import torch

    V = int(outdeg.shape[0])
    inv = 1.0 / outdeg.to(torch.float64).clamp(min=1)
    A = torch.sparse_csr_tensor(off, col, inv[col.long()], (V, V))

    def solve(base, x):  # one column: torch's CSR x vector product
        it = 0
        while it < max_iter:
            xn = base + damping * (A @ x)
            err = float((xn - x).abs().sum())
            x = xn
            it += 1
            if err < tol:
                break
        return x[:, 0], it

    if seeds is None:
        base = torch.full((V, 1), (1.0 - damping) / V, dtype=torch.float64, device=off.device)
        x = torch.full((V, 1), 1.0 / V, dtype=torch.float64, device=off.device)
        return solve(base, x)
    cols, its = [], 0
    for sd in seeds:
        base = torch.zeros((V, 1), dtype=torch.float64, device=off.device)
        base[int(sd), 0] = 1.0 - damping
        x = torch.zeros((V, 1), dtype=torch.float64, device=off.device)
        x[int(sd), 0] = 1.0
        xs, it = solve(base, x)
        cols.append(xs)
        its = max(its, it)
    return torch.stack(cols, 1), its


def bench_tol_run(plan, off, col, deg, tol, max_iter):
    """The same plan run to an L1 tolerance from 1/V; L1 distance of the FP32
    scores to the FP64 fixed point."""
    import numpy as np
    import torch

    V = int(deg.shape[0])
    s = torch.full((V,), float(np.float32(1.0) / np.float32(V)), dtype=torch.float32,
                   device=off.device)
    st, errs = plan.run(s, tol=tol, max_iter=max_iter, profile=False)
    x64, it64 = fp64_pagerank(off, col, deg, tol=1e-13)
    l1 = float((s.double() - x64).abs().sum())
    rel = float(((s.double() - x64).abs() / x64.clamp(min=1e-300)).max())
    rec = {"tol": tol, "iterations": st["iterations"], "converged": bool(st["converged"]),
           "final_l1_change": st["final_error"], "near_tie": bool(st["near_tie"]),
           "solve_ms": round(st["solve_ms"], 3), "l1_vs_fp64": l1,
           "max_rel_vs_fp64": rel, "fp64_iterations": it64}
    return rec, s.cpu().numpy(), errs, x64.cpu().numpy()


def bench_verify(scale, seed, scores_h):
    """Device PageRank verifier (gd_pr_verify) on the RMAT out-CSR: the
    reference PRVerifier's push residual (src/pr/verifier.cc:50-68) of the
    headline's final scores.  One call transposes the out-CSR, builds the
    schedule and pulls once; best of 3 calls after a warm-up.  Returns the
    record and the host out-CSR (for the CPU baseline)."""
    import numpy as np
    import torch

    from gardenia_ocl_amd import graph, verify_residual

    dev = torch.device("cuda:0")
    out_csr, _ = graph.rmat(scale, 16, seed=seed, want_in=False)
    off = torch.from_numpy(out_csr.row_offsets).to(dev)
    col = torch.from_numpy(out_csr.column_indices).to(dev)
    deg = torch.from_numpy(out_csr.degree).to(dev)
    s = torch.from_numpy(np.ascontiguousarray(scores_h, np.float32)).to(dev)
    verify_residual(off, col, deg, s)
    torch.cuda.synchronize()
    best, res = float("inf"), 0.0
    for _ in range(3):
        t0 = time.perf_counter()
        res = verify_residual(off, col, deg, s)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    rec = {"workload": f"RMAT-{scale} out-CSR ({out_csr.nnz} edges), residual of the "
                       "20-iteration scores", "ms": round(best * 1e3, 3), "residual": res,
           "verdict": "Correct" if res < 1e-3 else f"Total Error: {res:f}"}
    del off, col, deg, s
    return rec, out_csr


def bench_push(out_csr, iters, scores_pull):
    """Push PageRank (the reference's scatter variant) on the same RMAT
    out-CSR, the same fixed iterations from 1/V: FP32 atomic adds into HBM.
    Scores against the pull path's (FP32 summation-order noise)."""
    import numpy as np
    import torch

    from gardenia_ocl_amd import pagerank_push

    dev = torch.device("cuda:0")
    off = torch.from_numpy(out_csr.row_offsets).to(dev)
    col = torch.from_numpy(out_csr.column_indices).to(dev)
    m, E = out_csr.m, out_csr.nnz
    init = float(np.float32(1.0) / np.float32(m))
    s = torch.full((m,), init, dtype=torch.float32, device=dev)
    pagerank_push(off, col, s, tol=0.0, max_iter=2)  # warm-up
    s.fill_(init)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, _ = pagerank_push(off, col, s, tol=0.0, max_iter=iters)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    got = s.cpu().numpy()
    ref = np.asarray(scores_pull, np.float32)
    del off, col, s
    return {"workload": f"push PageRank over the RMAT out-CSR ({E} edges), {iters} iterations",
            "ms_per_iteration": round(st["solve_ms"] / iters, 4),
            "gteps": round(E * iters / (st["solve_ms"] * 1e-3) / 1e9, 2),
            "wall_ms": round(wall * 1e3, 2),
            "max_rel_vs_pull": float(np.max(np.abs(got - ref) / np.maximum(ref, 1e-30))),
            "l1_vs_pull": float(np.abs(got.astype(np.float64) - ref).sum())}


def bench_oneshot(in_csr, out_deg, iters):
    """The drop-in path a reference caller takes: PRSolver on host arrays
    (gd_pr_solve: copies the in-CSR to HBM, builds a plan, iterates, copies
    the scores back), timed end to end on the headline graph; the iteration
    loop alone is the library's HIP-event solve time."""
    import numpy as np

    from gardenia_ocl_amd import PRSolver

    m, E = in_csr.m, in_csr.nnz
    init = np.float32(1.0) / np.float32(m)
    s = np.full(m, init, np.float32)
    PRSolver(m, E, in_csr.row_offsets, in_csr.column_indices, None, None, out_deg, s,
             tol=0.0, max_iter=2)  # warm-up
    s[:] = init
    t0 = time.perf_counter()
    st, _ = PRSolver(m, E, in_csr.row_offsets, in_csr.column_indices, None, None, out_deg, s,
                     tol=0.0, max_iter=iters)
    wall = time.perf_counter() - t0
    return {"workload": f"PRSolver on host arrays, {iters} iterations, V={m} E={E}",
            "wall_ms": round(wall * 1e3, 2), "solve_ms": round(st["solve_ms"], 3),
            "copy_and_plan_ms": round(wall * 1e3 - st["solve_ms"], 2),
            "gteps_end_to_end": round(E * iters / wall / 1e9, 2),
            "gteps_iterations_only": round(E * iters / (st["solve_ms"] * 1e-3) / 1e9, 2)}


def bench_small(scale, tol, max_iter, seed, reps=7):
    """A whole solve on a small graph (RMAT-16, tol 1e-4, max 100: the
    reference's CPU-runnable case) with the plan reused, scores reset between
    solves; median over reps of the HIP-event solve time and of the wall time
    of the call."""
    import numpy as np
    import torch

    from gardenia_ocl_amd import PageRankPlan, graph

    dev = torch.device("cuda:0")
    _, ic = graph.rmat(scale, 16, seed=seed, want_out=False)
    V, E = ic.m, ic.nnz
    od = np.bincount(ic.column_indices, minlength=V).astype(np.int32)
    off = torch.from_numpy(ic.row_offsets).to(dev)
    col = torch.from_numpy(ic.column_indices).to(dev)
    deg = torch.from_numpy(od).to(dev)
    plan = PageRankPlan(off, col, deg)
    s = torch.empty(V, dtype=torch.float32, device=dev)
    init = float(np.float32(1.0) / np.float32(V))
    ev, wall = [], []
    for _ in range(reps):
        s.fill_(init)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, errs = plan.run(s, tol=tol, max_iter=max_iter)
        wall.append((time.perf_counter() - t0) * 1e3)
        ev.append(st["solve_ms"])
    its = st["iterations"]
    ms = float(np.median(ev))
    x64, _ = fp64_pagerank(off, col, deg, tol=1e-13)
    rec = {"workload": f"RMAT-{scale} ({E} edges) pull PageRank to L1 tol {tol}, max {max_iter}",
           "iterations": its, "solve_ms": round(ms, 4), "wall_ms": round(float(np.median(wall)), 4),
           "us_per_iteration": round(ms / max(its, 1) * 1e3, 2),
           "gteps": round(E * its / (ms * 1e-3) / 1e9, 2),
           "l1_vs_fp64": float((s.double() - x64).abs().sum())}
    plan.close()
    return rec, (s.cpu().numpy(), errs, its)


def bench_ppr(scale, k, tol, seed, fp64_seeds, hub_cols=0, tune_cache=None):
    """Batched PPR (k seeds in one SpMM-style pull) vs k looped single-seed
    runs of the plain pull kernel, both to the same per-seed L1 tolerance.
    The batched plan takes the tuned configuration for this graph and k from
    ``tune_cache`` when it holds one (hub_cols > 0 overrides it)."""
    import numpy as np
    import torch

    from gardenia_ocl_amd import PageRankPlan, PersonalizedPageRankPlan, graph

    dev = torch.device("cuda:0")
    _, ic = graph.rmat(scale, 16, seed=seed, want_out=False)
    V = ic.m
    od = np.bincount(ic.column_indices, minlength=V).astype(np.int32)
    seeds = np.random.default_rng(seed + 17).integers(0, V, k).astype(np.int32)
    off = torch.from_numpy(ic.row_offsets).to(dev)
    col = torch.from_numpy(ic.column_indices).to(dev)
    deg = torch.from_numpy(od).to(dev)
    tseeds = torch.from_numpy(seeds).to(dev)
    ppr_kw, source = {}, "library defaults"
    if tune_cache:
        from gardenia_ocl_amd import autotune

        tuned = autotune.ppr_config(tune_cache, off, col, deg, k)
        if tuned:
            ppr_kw, source = dict(tuned), f"autotuned ({os.path.relpath(tune_cache, ROOT)})"
    if hub_cols:
        ppr_kw["hub_cols"], source = hub_cols, "command line"
    log(f"batched ppr: RMAT-{scale} built, plan {ppr_kw or 'defaults'} ...")
    bp = PersonalizedPageRankPlan(off, col, deg, k, **ppr_kw)
    X = bp.one_hot(seeds)
    bp.run(tseeds, X, tol=tol, max_iter=2)  # warm-up
    X = bp.one_hot(seeds)
    torch.cuda.synchronize()
    log("batched ppr: plan and warm-up done; timed run ...")
    t0 = time.perf_counter()
    bst, _, _ = bp.run(tseeds, X, tol=tol, max_iter=300)
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    log(f"batched ppr: {t_batched:.3f} s, {bst['iterations']} iterations; single-seed plan ...")
    pp = PageRankPlan(off, col, deg)
    s = torch.zeros(V, dtype=torch.float32, device=dev)
    s[int(seeds[0])] = 1.0
    pp.run(s, tol=tol, max_iter=2, seed_vertex=int(seeds[0]))  # warm-up
    looped = []
    its = 0
    log("batched ppr: looped single-seed runs ...")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j in range(k):
        s.zero_()
        s[int(seeds[j])] = 1.0
        st, _ = pp.run(s, tol=tol, max_iter=300, seed_vertex=int(seeds[j]))
        its += st["iterations"]
        if j < fp64_seeds:
            looped.append(s.clone())
    torch.cuda.synchronize()
    t_looped = time.perf_counter() - t0
    # L1 against the FP64 personalised solution for the first seeds
    l1_b, l1_l = [], []
    if fp64_seeds > 0:
        log(f"looped: {t_looped:.3f} s; FP64 reference of {fp64_seeds} seeds ...")
        X64, _ = fp64_pagerank(off, col, deg, seeds=seeds[:fp64_seeds], tol=1e-13)
        log("FP64 reference done")
        l1_b = [float((X[:, j].double() - X64[:, j]).abs().sum()) for j in range(fp64_seeds)]
        l1_l = [float((looped[j].double() - X64[:, j]).abs().sum()) for j in range(fp64_seeds)]
    # algorithmic bytes of the batched solve (every array once per iteration,
    # the k-wide score / contrib rows of the running columns only):
    # iterations * (4(V+1) + 4E + 4V) + 16 V * sum over iterations of the running columns
    col_its = int(bp.column_iterations().sum())
    E = ic.nnz
    abytes = bst["iterations"] * (4 * (V + 1) + 4 * E + 4 * V) + 16 * V * col_its
    achieved = abytes / t_batched / 1e9
    return {"workload": f"RMAT-{scale} ({ic.nnz} edges), {k} one-hot seeds, per-seed L1 tol {tol}",
            "algorithmic_bytes": abytes, "achieved_gbs": round(achieved, 1),
            "frac_of_hbm_peak": round(achieved / HBM_PEAK_GBS, 4),
            "gather_bytes": 4 * E * col_its,
            "batched_s": round(t_batched, 4), "looped_s": round(t_looped, 4),
            "speedup": round(t_looped / t_batched, 2), "batched_iterations": bst["iterations"],
            "batched_column_iterations_total": col_its,
            "looped_iterations_total": its, "k_chunk": bp.info()["k_chunk"],
            "hub_cols": bp.info()["hub_cols"], "plan_config_source": source,
            "l1_vs_fp64_batched": l1_b, "l1_vs_fp64_looped": l1_l,
            "fp64_seeds_checked": fp64_seeds}


def bench_csr_build(scale, seed, host_s, E):
    """fill_data() on the device: the RMAT in-CSR (transposed, sorted,
    self-loops and duplicates removed -- the graph the bench runs on) from its
    raw edge list already in HBM, against the host builder's time."""
    import numpy as np
    import torch

    from gardenia_ocl_amd import graph

    src, dst = graph.rmat_edges(scale, 16, seed=seed)
    dev = torch.device("cuda:0")
    ts, td = torch.from_numpy(src).to(dev), torch.from_numpy(dst).to(dev)
    del src, dst
    graph.csr_from_edges_device(1 << scale, ts, td, transpose=True, sorted=True)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d = graph.csr_from_edges_device(1 << scale, ts, td, transpose=True, sorted=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"workload": f"RMAT-{scale} in-CSR from {int(ts.shape[0])} raw edges (fill_data flags)",
            "device_ms": round(dt * 1e3, 2), "host_rmat_graph_s": round(host_s, 2),
            "edges_kept": int(d["column_indices"].shape[0]),
            "same_edge_count_as_host_graph": int(d["column_indices"].shape[0]) == E,
            "speedup_vs_host": round(host_s / dt, 1)}


def bench_symgs(n, rmat_scale, seed, reps=20):
    """Multicolor SymGS sweeps (SymGSPlan): a 27-point stencil on an n^3 grid
    (HPCG-shaped, 8 colors) and a symmetrized RMAT (weights U(-1,1), read with
    the SymGS driver's flags) made strictly diagonally dominant, so repeated
    sweeps converge (first-fit colors).  Algorithmic bytes per
    symmetric sweep: two passes over every swept row, each reading the
    off-diagonal column, value and x[column] (12 B per entry) and the row's
    offset, diagonal, original index and b and writing x (20 B per row)."""
    import numpy as np
    import torch

    from gardenia_ocl_amd import SymGSPlan, graph

    dev = torch.device("cuda:0")
    out = {}

    def run(name, c):
        colors, nc = graph.color_first_fit(c)
        idx, co = graph.color_order(colors, nc)
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(dev)  # noqa: E731
        plan = SymGSPlan(t(c.row_offsets, np.int32), t(c.column_indices, np.int32),
                         t(c.weight, np.float32), t(idx, np.int32), co)
        info = plan.info()
        x = torch.zeros(c.m, dtype=torch.float32, device=dev)
        b = t(graph.uniform_pm1(c.m, seed + 17), np.float32)
        plan.run(x, b, 2)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.run(x, b, reps)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = 2 * (12 * info["offdiag_nnz"] + 20 * info["rows"])
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"rows": c.m, "nnz": c.nnz, "offdiag_nnz": info["offdiag_nnz"],
                     "colors": nc, "exact_rows": info["n_exact_rows"],
                     "launches_per_sweep": 2 * nc, "ms_per_sweep": round(ms, 4),
                     "algorithmic_bytes_per_sweep": nbytes, "achieved_gbs": round(gbs, 1),
                     "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                     "plan_ms": round(info["plan_ms"], 1)}
        plan.close()

    run(f"stencil27_{n}^3", graph.stencil27(n))
    src, dst = graph.rmat_edges(rmat_scale, 16, seed=seed)
    m = 1 << rmat_scale
    w = graph.uniform_pm1(len(src), seed + 3)  # off-diagonal weights of both signs
    c = graph.csr_from_edges(m, src, dst, w, symmetrize=True, transpose=False, sorted=True,
                             remove_selfloops=False, remove_redundents=True)
    del src, dst, w
    run(f"rmat{rmat_scale}_sym_dominant", graph.with_dominant_diagonal(c))
    return out


def bench_multigraph(n, lo, hi, tol, seed, tune_cache=None):
    """n RMAT graphs of scales uniform in [lo, hi]: one multi-graph plan (one
    pull + one finalize launch per iteration for all graphs, replayed from a
    hipGraph) vs one single-graph plan per graph run back to back.  The
    iterations per captured chunk come from ``tune_cache`` when it holds a
    tuned value for this set of graphs."""
    import numpy as np
    import torch

    from gardenia_ocl_amd import MultiGraphPageRank, PageRankPlan, graph

    dev = torch.device("cuda:0")
    scales = np.random.default_rng(seed + 5).integers(lo, hi + 1, n)
    gs = []
    for i, sc in enumerate(scales):
        _, ic = graph.rmat(int(sc), 16, seed=seed + 100 + i, want_out=False)
        od = np.bincount(ic.column_indices, minlength=ic.m).astype(np.int32)
        gs.append(tuple(torch.from_numpy(a).to(dev) for a in (ic.row_offsets, ic.column_indices, od)))
    E = int(sum(int(g[1].shape[0]) for g in gs))
    init = [float(np.float32(1.0) / np.float32(int(g[2].shape[0]))) for g in gs]
    scores = [torch.empty(int(g[2].shape[0]), dtype=torch.float32, device=dev) for g in gs]
    mg = MultiGraphPageRank(gs)
    run_kw, source = {}, "library defaults"
    if tune_cache:
        from gardenia_ocl_amd import autotune

        tuned = autotune.multigraph_config(tune_cache, gs)
        if tuned:
            run_kw, source = dict(tuned), f"autotuned ({os.path.relpath(tune_cache, ROOT)})"
    out = {"workload": f"{n} RMAT graphs, scales {lo}-{hi} (seeded), L1 tol {tol}",
           "scales": [int(x) for x in scales], "total_edges": E,
           "chunk": run_kw.get("chunk", 8), "plan_config_source": source}
    for use_graph in (True, False):
        for rep in range(2):  # first run captures / warms up
            for s, v in zip(scores, init):
                s.fill_(v)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            its, errs, ms = mg.run(scores, tol=tol, max_iter=100, use_graph=use_graph, **run_kw)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        key = "hipgraph" if use_graph else "stream"
        out[f"{key}_ms"] = round(wall * 1e3, 3)
        out[f"{key}_solve_ms"] = round(ms, 3)
    out["iterations_max"] = int(its.max())
    out["iterations_total"] = int(its.sum())
    l1 = []
    for (g_off, g_col, g_deg), sc in zip(gs, scores):
        x64, _ = fp64_pagerank(g_off, g_col, g_deg, tol=1e-13)
        l1.append(float((sc.double() - x64).abs().sum()))
    out["l1_vs_fp64_max"] = max(l1)
    out["l1_vs_fp64_mean"] = sum(l1) / len(l1)
    plans = [PageRankPlan(*g) for g in gs]
    for p, s, v in zip(plans, scores, init):  # warm-up
        s.fill_(v)
        p.run(s, tol=tol, max_iter=100)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its_l = 0
    for p, s, v in zip(plans, scores, init):
        s.fill_(v)
        st, _ = p.run(s, tol=tol, max_iter=100)
        its_l += st["iterations"]
    torch.cuda.synchronize()
    out["looped_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
    out["looped_iterations_total"] = its_l
    out["speedup_vs_looped"] = round(out["looped_ms"] / out["hipgraph_ms"], 2)
    out["gteps_hipgraph"] = round(int(sum(int(g[1].shape[0]) * int(i) for g, i in zip(gs, its)))
                                  / (out["hipgraph_ms"] * 1e-3) / 1e9, 2)
    # plain PR's 4E + 24V + 4 bytes per graph-iteration, summed over the run
    byts = sum((4 * int(g[1].shape[0]) + 24 * int(g[2].shape[0]) + 4) * int(i)
               for g, i in zip(gs, its))
    out["algorithmic_bytes"] = byts
    out["achieved_gbs"] = round(byts / (out["hipgraph_ms"] * 1e-3) / 1e9, 1)
    out["frac_of_hbm_peak"] = round(out["achieved_gbs"] / HBM_PEAK_GBS, 4)
    for p in plans:
        p.close()
    mg.close()
    return out


def bench_adjoint(scale, spmv_log2, seed):
    """Weighted PageRank forward + implicit adjoint (RMAT, weights U(0.5,1.5))
    and SpMV backward (dx = A^T g, dA = g[row] x[col]) on an ER matrix."""
    import numpy as np
    import torch

    from gardenia_ocl_amd import SpmvPlan, graph
    from gardenia_ocl_amd.autograd import SpmvBackwardPlan, WeightedPageRankPlan

    dev = torch.device("cuda:0")
    _, ic = graph.rmat(scale, 16, seed=seed, want_out=False)
    m, nnz = ic.m, ic.nnz
    a = np.random.default_rng(seed).uniform(0.5, 1.5, nnz).astype(np.float32)
    off = torch.from_numpy(ic.row_offsets).to(dev)
    col = torch.from_numpy(ic.column_indices).to(dev)
    wt = torch.from_numpy(a).to(dev)
    wp = WeightedPageRankPlan(off, col, wt)
    x = torch.full((m,), 1.0 / m, dtype=torch.float32, device=dev)
    w = torch.from_numpy(np.random.default_rng(seed + 1).standard_normal(m).astype(np.float32)).to(dev)
    lam = torch.empty_like(x)
    grad = torch.empty(nnz, dtype=torch.float32, device=dev)
    wp.forward(x, tol=1e-7, max_iter=300)  # warm-up
    wp.backward(x, w, lam, grad, tol=1e-7, max_iter=300)
    x.fill_(1.0 / m)
    f = wp.forward(x, tol=1e-7, max_iter=300)
    b = wp.backward(x, w, lam, grad, tol=1e-7, max_iter=300)
    out = {"workload": f"weighted PR RMAT-{scale} ({nnz} edges), tol 1e-7; SpMV backward "
                       f"ER 2^{spmv_log2} x 32",
           "forward_iterations": f["iterations"], "forward_ms": round(f["solve_ms"], 3),
           "backward_iterations": b["iterations"], "backward_ms": round(b["solve_ms"], 3),
           "forward_gteps": round(nnz * f["iterations"] / (f["solve_ms"] * 1e-3) / 1e9, 2),
           "backward_gteps": round(nnz * b["iterations"] / (b["solve_ms"] * 1e-3) / 1e9, 2)}
    # per iteration, forward and adjoint alike: plain PR's 4E + 24V + 4 bytes
    # (the adjoint pulls over the reversed graph)
    it_bytes = 4 * nnz + 24 * m + 4
    for k, r in (("forward", f), ("backward", b)):
        gbs = it_bytes * r["iterations"] / (r["solve_ms"] * 1e-3) / 1e9
        out[f"{k}_achieved_gbs"] = round(gbs, 1)
        out[f"{k}_frac_of_hbm_peak"] = round(gbs / HBM_PEAK_GBS, 4)
    wp.close()
    del off, col, wt, grad
    torch.cuda.empty_cache()
    er = graph.er_matrix(spmv_log2, 32, seed=seed)
    Ap = torch.from_numpy(er.row_offsets).to(dev)
    Aj = torch.from_numpy(er.column_indices).to(dev)
    Ax = torch.from_numpy(er.weight).to(dev)
    xs = torch.from_numpy(graph.uniform_pm1(er.m, seed + 1)).to(dev)
    g = torch.from_numpy(graph.uniform_pm1(er.m, seed + 2)).to(dev)
    dx = torch.empty_like(xs)
    dA = torch.empty_like(Ax)
    bwd = SpmvBackwardPlan(Ap, Aj)
    for _ in range(2):
        bwd.run(Ax, xs, g, dx, dA)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(st)
    for _ in range(reps):
        bwd.run(Ax, xs, g, dx, dA)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # dx pull over the CSC: 8nnz + 16M + 4; dA edge map: 4(M+1) + 4nnz + 4nnz + 8M
    byts = (8 * er.nnz + 16 * er.m + 4) + (4 * (er.m + 1) + 8 * er.nnz + 8 * er.m)
    out.update({"spmv_bwd_ms": round(ms, 4),
                "spmv_bwd_achieved_gbs": round(byts / (ms * 1e-3) / 1e9, 1)})
    bwd.close()
    return out


def traffic_from_profile(scale):
    """The pull phase's counters measured with rocprofv3 (HBM bytes from
    FETCH_SIZE / WRITE_SIZE scaled by a 1 GiB copy, L2 requests and hits from
    TCC_REQ / TCC_HIT) for this workload, or {} when none is committed."""
    path = os.path.join(ROOT, "profiles", f"pr_pull_rmat{scale}_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


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
    ap.add_argument("--sort-tiles", choices=["auto", "on", "off"], default="auto")
    ap.add_argument("--block", type=int, default=0, help="threads per row tile (256/512/1024, 0=auto)")
    ap.add_argument("--hub-threshold", type=int, default=0,
                    help="rows longer than this are whole-row exact hubs (0=auto: the tile size)")
    ap.add_argument("--tune-cache", default=TUNED,
                    help="autotuner cache (python -m gardenia_ocl_amd.autotune --cache FILE): "
                         "the headline, batched-PPR and multi-graph plans take its tuned "
                         "configuration for their graphs when it holds one (default: the "
                         "package's MI355X cache; '' for library defaults)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the tol-1e-6 run, SpMV (ER 2^22 x 32), batched-PPR (RMAT-22, "
                         "k=256), multi-graph and adjoint figures")
    ap.add_argument("--ppr-scale", type=int, default=22)
    ap.add_argument("--ppr-k", type=int, default=256)
    ap.add_argument("--ppr-tol", type=float, default=1e-6)
    ap.add_argument("--ppr-hub-cols", type=int, default=0, help="PPR hub-row column group (16/32/64)")
    ap.add_argument("--spmv-log2", type=int, default=22)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device-to-device 5 times (scales rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    ap.add_argument("--only", choices=["symgs", "csr_build", "spmv", "small", "ppr", "adjoint",
                                       "multigraph", "push"],
                    default=None,
                    help="run just this secondary figure and print its JSON (for profiling)")
    ap.add_argument("--tol-run", type=float, default=1e-6,
                    help="L1 tolerance of the extra run to convergence (with its FP64 L1)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: the PageRank/SpMV path here has no multi-GPU "
              "build (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gardenia_ocl_amd import PageRankPlan, graph

    dev = torch.device("cuda:0")
    if args.only == "symgs":
        print(json.dumps({"symgs": bench_symgs(128, 20, args.seed)}), flush=True)
        return
    if args.only == "small":
        print(json.dumps({"small": bench_small(16, 1e-4, 100, args.seed)[0]}), flush=True)
        return
    if args.only == "ppr":
        print(json.dumps({"ppr": bench_ppr(args.ppr_scale, args.ppr_k, args.ppr_tol, args.seed, 8,
                                           args.ppr_hub_cols, args.tune_cache)}), flush=True)
        return
    if args.only == "spmv":
        print(json.dumps({"spmv": bench_spmv(args.spmv_log2, 32, 50, args.seed)[0]}), flush=True)
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint": bench_adjoint(20, args.spmv_log2, args.seed)}), flush=True)
        return
    if args.only == "multigraph":
        print(json.dumps({"multigraph": bench_multigraph(64, 12, 18, 1e-4, args.seed,
                                                         args.tune_cache)}), flush=True)
        return
    if args.only == "push":
        out_csr, in_csr = graph.rmat(args.scale, 16, seed=args.seed)
        off = torch.from_numpy(in_csr.row_offsets).to(dev)
        col = torch.from_numpy(in_csr.column_indices).to(dev)
        deg = torch.from_numpy(out_csr.degree).to(dev)
        s = torch.full((in_csr.m,), float(np.float32(1.0) / np.float32(in_csr.m)),
                       dtype=torch.float32, device=dev)
        PageRankPlan(off, col, deg).run(s, tol=0.0, max_iter=args.iters)
        scores_h = s.cpu().numpy()
        del off, col, deg, s
        torch.cuda.empty_cache()
        print(json.dumps({"push": bench_push(out_csr, args.iters, scores_h)}), flush=True)
        return
    if args.only == "csr_build":
        t0 = time.perf_counter()
        _, ic = graph.rmat(args.scale, args.edgefactor, seed=args.seed, want_out=False)
        host_s = time.perf_counter() - t0
        print(json.dumps({"csr_build": bench_csr_build(args.scale, args.seed, host_s, ic.nnz)}),
              flush=True)
        return
    if args.calibrate_copy:
        a = torch.ones(256 * 2**20, dtype=torch.float32, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 1 GiB 5 times")
        return
    t0 = time.perf_counter()
    _, in_csr = graph.rmat(args.scale, args.edgefactor, seed=args.seed, want_out=False)
    V, E = in_csr.m, in_csr.nnz
    # out-degree = number of times a vertex appears as a source in the in-CSR
    out_deg = np.bincount(in_csr.column_indices, minlength=V).astype(np.int32)
    host_build_s = time.perf_counter() - t0
    log(f"RMAT-{args.scale}: V={V} E={E} built in {host_build_s:.1f} s")
    off = torch.from_numpy(in_csr.row_offsets).to(dev)
    col = torch.from_numpy(in_csr.column_indices).to(dev)
    deg = torch.from_numpy(out_deg).to(dev)
    scores = torch.empty(V, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    tri = {"auto": None, "on": True, "off": False}
    plan_kw = dict(relabel=tri[args.relabel], items_per_thread=args.ipt,
                   nontemporal=None if args.nt < 0 else args.nt,
                   persistent=tri[args.persistent], sort_tiles=tri[args.sort_tiles],
                   block=args.block, hub_threshold=args.hub_threshold)
    config_source = "library defaults and command-line knobs"
    if args.tune_cache:
        from gardenia_ocl_amd import autotune

        tuned = autotune.pagerank_config(args.tune_cache, off, col, deg)
        if tuned:
            plan_kw = tuned
            config_source = f"autotuned ({os.path.relpath(args.tune_cache, ROOT)})"
        else:
            log(f"no tuned configuration for this graph in {args.tune_cache}: defaults")
    plan = PageRankPlan(off, col, deg, stream=stream, **plan_kw)
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

    scores_h = scores.cpu().numpy()  # the 20-iteration scores (verifier input)
    gteps = E * args.iters * args.steps / dt / 1e9
    avg_pull_ms = pull_ms / max(1, pull_n)
    bytes_iter = 4 * E + 24 * V + 4
    achieved = bytes_iter / (avg_pull_ms * 1e-3) / 1e9
    prof = traffic_from_profile(args.scale)
    roof = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": prof.get("hbm_bytes_per_launch"),
        "traffic_kernels": prof.get("phase_kernels"),
        "l2_requests_per_edge": prof.get("l2_requests_per_edge"),
        "l2_hit_rate": prof.get("l2_hit_rate"),
        "kernel": (f"{'pull_persistent_kernel' if info['persistent'] else 'pull_tile_kernel'}"
                   f"<PrPolicy, {info['items_per_thread']}, {info['nontemporal']}, "
                   f"{'true' if info['sort_tiles'] else 'false'}"
                   f"{', ' + str(info['block']) if info['block'] != 256 else ''}>"),
        "avg_launch_ms": round(avg_pull_ms, 5),
        "algorithmic_bytes_per_launch": bytes_iter,
    }
    # The headline: the same steps in production mode (the plan's own run:
    # chunks of iterations replayed from captured hipGraphs, no per-launch
    # events).  The timed region above launches eagerly with HIP events around
    # every pull launch to time the pull phase for the roofline; it is
    # reported beside the headline as "profiled".
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    torch.cuda.synchronize()
    dt_prod = time.perf_counter() - t0
    log(f"timed {args.steps} production steps in {dt_prod:.3f} s")
    gteps_prod = E * args.iters * args.steps / dt_prod / 1e9
    profiled = {"ms_per_step": round(dt / args.steps * 1e3, 4), "gteps": round(gteps, 3),
                "mode": "eager launches, HIP events around every pull launch (the roofline's "
                        "kernel timing)"}
    extras = {}
    checks = {}  # GPU results the CPU leg compares with the oracle
    if not args.no_extras:
        log(f"run to tol {args.tol_run} ...")
        extras["tol_run"], tol_scores, tol_errs, tol_x64 = bench_tol_run(plan, off, col, deg,
                                                                         args.tol_run, 200)
        checks["tol_run"] = (tol_scores, tol_errs, tol_x64)
        del plan, off, col, deg, scores
        torch.cuda.empty_cache()
        log("one-shot PRSolver on host arrays ...")
        extras["oneshot"] = bench_oneshot(in_csr, out_deg, args.iters)
        torch.cuda.empty_cache()
        log("device verifier ...")
        extras["verify"], verify_csr = bench_verify(args.scale, args.seed, scores_h)
        torch.cuda.empty_cache()
        log("push PageRank ...")
        extras["push"] = bench_push(verify_csr, args.iters, scores_h)
        torch.cuda.empty_cache()
        log("small graph ...")
        extras["small"], checks["small"] = bench_small(16, 1e-4, 100, args.seed)
        log("spmv ...")
        extras["spmv"], checks["spmv"] = bench_spmv(args.spmv_log2, 32, 50, args.seed)
        torch.cuda.empty_cache()
        log("batched ppr ...")
        extras["ppr"] = bench_ppr(args.ppr_scale, args.ppr_k, args.ppr_tol, args.seed, 8,
                                  args.ppr_hub_cols, args.tune_cache)
        torch.cuda.empty_cache()
        log("device CSR build ...")
        extras["csr_build"] = bench_csr_build(args.scale, args.seed, host_build_s, E)
        torch.cuda.empty_cache()
        log("symgs ...")
        extras["symgs"] = bench_symgs(128, 20, args.seed)
        torch.cuda.empty_cache()
        log("multi-graph ...")
        extras["multigraph"] = bench_multigraph(64, 12, 18, 1e-4, args.seed, args.tune_cache)
        torch.cuda.empty_cache()
        log("adjoint ...")
        extras["adjoint"] = bench_adjoint(20, args.spmv_log2, args.seed)
        torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(in_csr, out_deg, args.iters, args.cpu_seconds)
        if not args.no_extras:
            log("oracle comparison of the tol run ...")
            extras["tol_run"].update(tol_parity(in_csr, out_deg, args.tol_run, 200,
                                                *checks["tol_run"], extras["tol_run"]))
            del in_csr
            cpu["secondary"] = cpu_baseline_extras(args.spmv_log2, args.seed,
                                                   verify=(verify_csr, scores_h), checks=checks,
                                                   extras=extras)
    rec = {
        "metric": METRIC, "value": round(gteps_prod, 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt_prod / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"RMAT-{args.scale} pull PageRank, {args.iters} fixed iterations",
                   "scale": args.scale, "edgefactor": args.edgefactor, "seed": args.seed,
                   "V": V, "E": E, "iterations_per_step": args.iters,
                   "abcd": [0.57, 0.19, 0.19, 0.05], "parallelism": "single-gpu",
                   "relabel": bool(info["relabel"]), "items_per_thread": info["items_per_thread"],
                   "nontemporal": info["nontemporal"], "persistent": bool(info["persistent"]),
                   "pull_grid": info["pull_grid"], "sort_tiles": bool(info["sort_tiles"]),
                   "block": info["block"], "hub_threshold": info["hub_threshold"],
                   "plan_config_source": config_source,
                   "plan_ms": round(info["plan_ms"], 1)},
        "mode": "hipGraph replay of the iteration chunks, no events (the plan's run)",
        "roofline": roof, "cpu_baseline": cpu, "profiled": profiled,
    }
    if "ppr" in extras:
        rec["ppr_speedup_vs_looped"] = extras["ppr"]["speedup"]
    rec.update(extras)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
