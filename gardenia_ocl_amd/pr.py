"""Pull PageRank: the reference's PRSolver interface over the HIP library.

``PRSolver`` keeps the reference signature (src/pr/pr.h:23) with host numpy
arrays, scores in/out, and prints the reference's stdout lines when asked
(src/pr/omp_base.cc:15,42,46-47).  ``PageRankPlan`` is the device-resident
form: torch tensors in HBM, a schedule built once, many runs.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _native as N

K_DAMP = 0.85        # include/common.h:18
EPSILON = 1e-3       # src/pr/pr.h:4
MAX_ITER = 30        # src/pr/pr.h:5


def make_options(damping=K_DAMP, tol=EPSILON, max_iter=MAX_ITER, seed_vertex=-1, chunk=0,
                 profile=False) -> N.PrOptions:
    o = N.PrOptions()
    N.lib().gd_pr_default_options(ctypes.byref(o))
    o.damping = float(damping)
    o.tol = float(tol)
    o.max_iter = int(max_iter)
    o.seed_vertex = int(seed_vertex)
    o.chunk = int(chunk)
    o.profile = int(bool(profile))
    return o


def _np_ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _as_i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def PRSolver(m: int, nnz: int, in_row_offsets, in_column_indices, out_row_offsets,
             out_column_indices, degree, scores: np.ndarray, *, damping=K_DAMP, tol=EPSILON,
             max_iter=MAX_ITER, seed_vertex=-1, verbose=False, variant="hip_pull",
             config=None):
    """PRSolver(m, nnz, in_off, in_col, out_off, out_col, degree, scores) on the GPU.

    ``scores`` (float32, length m) is updated in place, as in the reference.
    Returns (stats dict, per-iteration L1 errors).  Raises GardeniaError on any
    failure (the reference exit()s instead, include/cutil_subset.h:4-12).
    ``config``: PageRankPlan keyword arguments (e.g. a tuned configuration,
    ``autotune.pagerank_config``); the solve then runs on a plan built with
    them (same scores and iterations: no knob changes the arithmetic);
    None: the library's one-shot gd_pr_solve with its defaults.
    """
    if not (isinstance(scores, np.ndarray) and scores.dtype == np.float32
            and scores.flags.c_contiguous):
        raise TypeError("scores must be a contiguous float32 numpy array (updated in place)")
    if scores.shape[0] != m:
        raise ValueError("scores must have length m")
    in_off, in_col, deg = _as_i32(in_row_offsets), _as_i32(in_column_indices), _as_i32(degree)
    out_off = None if out_row_offsets is None else _as_i32(out_row_offsets)
    out_col = None if out_column_indices is None else _as_i32(out_column_indices)
    if in_off.shape[0] != m + 1 or deg.shape[0] != m or in_col.shape[0] < nnz:
        raise ValueError("CSR array lengths do not match m / nnz")
    opts = make_options(damping, tol, max_iter, seed_vertex)
    stats = N.PrStats()
    errors = np.zeros(max(1, max_iter), np.float64)
    if verbose:
        print("Launching HIP PR solver (gfx950 pull, fused contrib/L1) ...")
    if config is None:
        N.call("gd_pr_solve", m, nnz, _np_ptr(in_off), _np_ptr(in_col), _np_ptr(out_off),
               _np_ptr(out_col), _np_ptr(deg), _np_ptr(scores), ctypes.byref(opts),
               ctypes.byref(stats), _np_ptr(errors))
    else:
        import torch

        dev = torch.device("cuda", torch.cuda.current_device())
        t = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
        plan = PageRankPlan(t(in_off), t(in_col[:nnz]), t(deg), **config)
        x = t(scores)
        st, errs = plan.run(x, damping=damping, tol=tol, max_iter=max_iter,
                            seed_vertex=seed_vertex)
        scores[:] = x.cpu().numpy()
        plan.close()
        for name, _ in N.PrStats._fields_:
            setattr(stats, name, st[name])
        errors[: len(errs)] = errs
    errs = errors[: stats.iterations].copy()
    if verbose:
        for i, e in enumerate(errs):
            print(f" {i + 1:2d}    {e:f}")
        # iter+1 after the reference's loop (src/pr/omp_base.cc:23,46): MAX_ITER+1
        # when no iteration converged, 1 for max_iter 0
        printed = stats.iterations + (0 if stats.converged else 1)
        print(f"\titerations = {printed}.")
        print(f"\truntime [{variant}] = {stats.solve_ms:f} ms.")
    return stats.as_dict(), errs


def pagerank_push(out_row_offsets, out_column_indices, scores, *, damping=K_DAMP, tol=EPSILON,
                  max_iter=MAX_ITER, stream=None):
    """Push-style PageRank over the out-CSR (torch int32 CUDA tensors; scores
    float32, in/out): the reference's scatter variant (src/pr/scatter.cu:11-41),
    propagation-blocked (each edge's value binned by destination window, summed
    per window in LDS with FP32 atomic adds in the hardware's order), a
    cross-check of the pull path.
    Continues while the L1 change is > tol, at least once.  Returns (stats
    dict, per-iteration L1 changes)."""
    import torch

    for t, dt in ((out_row_offsets, torch.int32), (out_column_indices, torch.int32),
                  (scores, torch.float32)):
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
            raise TypeError("pagerank_push inputs must be contiguous CUDA tensors")
    m = int(scores.shape[0])
    if int(out_row_offsets.shape[0]) != m + 1:
        raise ValueError("out_row_offsets must have m+1 entries")
    st = stream if stream is not None else torch.cuda.current_stream()
    opts = make_options(damping, tol, max_iter)
    stats = N.PrStats()
    errors = np.zeros(max(1, max_iter), np.float64)
    N.call("gd_pr_push_solve", m, int(out_column_indices.shape[0]), out_row_offsets.data_ptr(),
           out_column_indices.data_ptr(), scores.data_ptr(), ctypes.byref(opts),
           ctypes.c_void_p(st.cuda_stream), ctypes.byref(stats), _np_ptr(errors))
    return stats.as_dict(), errors[: stats.iterations].copy()


def verify_residual(out_row_offsets, out_column_indices, degree, scores, *, damping=K_DAMP,
                    stream=None) -> float:
    """The reference verifier's push residual on the GPU (torch CUDA tensors:
    out-CSR int32, out-degree int32, scores float32): one push iteration of
    ``scores`` over the out-CSR, sum_v |base + d*sum_v - scores_v| in FP64,
    every per-vertex term bit-identical to src/pr/verifier.cc:50-68."""
    import torch

    for t, dt in ((out_row_offsets, torch.int32), (out_column_indices, torch.int32),
                  (degree, torch.int32), (scores, torch.float32)):
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
            raise TypeError("verify_residual inputs must be contiguous CUDA tensors")
    m = int(degree.shape[0])
    if int(out_row_offsets.shape[0]) != m + 1 or int(scores.shape[0]) != m:
        raise ValueError("out_row_offsets must have m+1 entries and scores m")
    st = stream if stream is not None else torch.cuda.current_stream()
    res = ctypes.c_double()
    N.call("gd_pr_verify", m, int(out_column_indices.shape[0]), out_row_offsets.data_ptr(),
           out_column_indices.data_ptr(), degree.data_ptr(), scores.data_ptr(),
           ctypes.c_float(damping), ctypes.c_void_p(st.cuda_stream), ctypes.byref(res))
    return res.value


def PRVerifier(m: int, in_row_offsets, in_column_indices, out_row_offsets, out_column_indices,
               degree, scores_to_test, target_error=EPSILON, *, damping=K_DAMP):
    """PRVerifier(m, in_off, in_col, out_off, out_col, degree, scores, target_error)
    (src/pr/pr.h:24) with the residual computed on the GPU.  Prints
    "Correct" or "Total Error: %f" as the reference does (verifier.cc:67-68)
    and returns (correct, residual).  The reference's discarded serial
    re-run of the solver (verifier.cc:11-41) is not repeated; the in-CSR
    arguments are accepted for the signature and unused."""
    import torch

    dev = torch.device("cuda:0")
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(dev)  # noqa: E731
    print("Verifying...")
    res = verify_residual(t(out_row_offsets, np.int32), t(out_column_indices, np.int32),
                          t(degree, np.int32), t(scores_to_test, np.float32), damping=damping)
    if res < target_error:
        print("Correct")
    else:
        print(f"Total Error: {res:f}")
    return res < target_error, res


class PageRankPlan:
    """Device-resident PR plan over torch int32 CUDA tensors (in-CSR + out-degree).

    The tensors are borrowed: keep them alive and unchanged while the plan lives.
    Work is issued on ``stream`` (default: torch's current stream).
    """

    def __init__(self, in_row_offsets, in_column_indices, out_degree, stream=None, *,
                 relabel=None, items_per_thread=0, nontemporal=None, persistent=None,
                 sort_tiles=None, block=0, hub_threshold=0, stage_hubs=-1):
        """relabel / nontemporal / persistent / sort_tiles: None = library default,
        True / False force it (sort_tiles: gather each tile's edges in source
        order; same results).  block: threads per row-tile workgroup (256, 512
        or 1024; 0 = library default).  hub_threshold: rows with more in-edges
        are whole-row hubs of their own workgroup (64 .. 256*items_per_thread;
        0 = library default).  stage_hubs: 1 gathers the hub rows with a kernel
        of their own before their workgroups sum them, 0 never, -1 library
        default (when the longest hub would outlast the row tiles).  Every
        setting gives the same results."""
        import torch

        for t in (in_row_offsets, in_column_indices, out_degree):
            if not (t.is_cuda and t.dtype == torch.int32 and t.is_contiguous()):
                raise TypeError("plan inputs must be contiguous int32 CUDA tensors")
        self.m = int(out_degree.shape[0])
        self.nnz = int(in_column_indices.shape[0])
        if int(in_row_offsets.shape[0]) != self.m + 1:
            raise ValueError("in_row_offsets must have m+1 entries")
        self._keep = (in_row_offsets, in_column_indices, out_degree)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        cfg = N.PrPlanConfig()
        N.lib().gd_pr_default_plan_config(ctypes.byref(cfg))
        if relabel is not None:
            cfg.relabel = 1 if relabel else 0
        cfg.items_per_thread = int(items_per_thread)
        if nontemporal is not None:
            cfg.nontemporal = int(nontemporal)
        if persistent is not None:
            cfg.persistent = 1 if persistent else 0
        if sort_tiles is not None:
            cfg.sort_tiles = 1 if sort_tiles else 0
        cfg.block = int(block)
        cfg.hub_threshold = int(hub_threshold)
        cfg.stage_hubs = int(stage_hubs)
        h = ctypes.c_void_p()
        N.call("gd_pr_plan_create_ex", self.m, self.nnz, in_row_offsets.data_ptr(),
               in_column_indices.data_ptr(), out_degree.data_ptr(),
               ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(cfg), ctypes.byref(h))
        self._h = h

    def info(self) -> dict:
        inf = N.PrPlanInfo()
        N.call("gd_pr_plan_info", self._h, ctypes.byref(inf))
        d = inf.as_dict()
        d["tiles"], d["hubs"], d["hub_segments"] = d["n_tiles"], d["n_hubs"], d["n_hub_segments"]
        return d

    def run(self, scores, *, damping=K_DAMP, tol=EPSILON, max_iter=MAX_ITER, seed_vertex=-1,
            chunk=0, profile=False):
        """Iterate on ``scores`` (float32 CUDA tensor, in/out).  Returns
        (stats dict, per-iteration L1 errors as numpy)."""
        import torch

        if not (scores.is_cuda and scores.dtype == torch.float32 and scores.is_contiguous()
                and scores.shape[0] == self.m):
            raise TypeError("scores must be a contiguous float32 CUDA tensor of length m")
        opts = make_options(damping, tol, max_iter, seed_vertex, chunk, profile)
        stats = N.PrStats()
        errors = np.zeros(max(1, max_iter), np.float64)
        N.call("gd_pr_plan_run", self._h, scores.data_ptr(), ctypes.byref(opts),
               ctypes.byref(stats), _np_ptr(errors))
        return stats.as_dict(), errors[: stats.iterations].copy()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().gd_pr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
