"""Batched personalised PageRank over the HIP library (k seeds per run).

Not in the reference (it has no PPR); the iteration is the reference's pull
loop (src/pr/omp_base.cc:23-44) with the teleport term moved to one seed per
column.  Scores are vertex-major, ``X[v, j]``, so each gathered source index
feeds k additions.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .pr import EPSILON, K_DAMP, MAX_ITER, make_options


class PersonalizedPageRankPlan:
    """Device-resident batched PPR plan over torch int32 CUDA tensors.

    ``k`` (a multiple of 64) is the number of seed columns per run.  The CSR
    tensors are borrowed: keep them alive and unchanged while the plan lives.
    """

    def __init__(self, in_row_offsets, in_column_indices, out_degree, k, stream=None, *,
                 relabel=None, k_chunk=0, column_stop=True, compact=True, hub_cols=0):
        """column_stop: each column stops at its own first iteration with L1
        change < tol (as k single-seed runs would); False: all columns run
        until the largest change is < tol.  compact: move stopped columns
        behind the running ones so they are no longer gathered (k <= 4096)."""
        import torch

        for t in (in_row_offsets, in_column_indices, out_degree):
            if not (t.is_cuda and t.dtype == torch.int32 and t.is_contiguous()):
                raise TypeError("plan inputs must be contiguous int32 CUDA tensors")
        self.m = int(out_degree.shape[0])
        self.nnz = int(in_column_indices.shape[0])
        self.k = int(k)
        self._keep = (in_row_offsets, in_column_indices, out_degree)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        cfg = N.PprPlanConfig()
        N.lib().gd_ppr_default_plan_config(ctypes.byref(cfg))
        if relabel is not None:
            cfg.relabel = 1 if relabel else 0
        cfg.k_chunk = int(k_chunk)
        cfg.column_stop = 1 if column_stop else 0
        cfg.compact = 1 if compact else 0
        cfg.hub_cols = int(hub_cols)
        h = ctypes.c_void_p()
        N.call("gd_ppr_plan_create", self.m, self.nnz, in_row_offsets.data_ptr(),
               in_column_indices.data_ptr(), out_degree.data_ptr(), self.k,
               ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(cfg), ctypes.byref(h))
        self._h = h

    def info(self) -> dict:
        inf = N.PprPlanInfo()
        N.call("gd_ppr_plan_info", self._h, ctypes.byref(inf))
        return inf.as_dict()

    def one_hot(self, seeds):
        """X0 for ``seeds`` (length k): X0[seeds[j], j] = 1."""
        import torch

        seeds = torch.as_tensor(seeds, dtype=torch.int64, device=self._keep[0].device)
        X = torch.zeros((self.m, self.k), dtype=torch.float32, device=seeds.device)
        X[seeds, torch.arange(self.k, device=seeds.device)] = 1.0
        return X

    def run(self, seeds, X, *, damping=K_DAMP, tol=EPSILON, max_iter=MAX_ITER, chunk=0):
        """Iterate on X (float32 CUDA tensor [m, k], in/out) for int32 CUDA
        ``seeds`` [k].  Returns (stats, per-iteration max-column L1, last
        per-column L1 of each column's last iteration)."""
        import torch

        if not (X.is_cuda and X.dtype == torch.float32 and X.is_contiguous()
                and tuple(X.shape) == (self.m, self.k)):
            raise TypeError("X must be a contiguous float32 CUDA tensor of shape (m, k)")
        if not (seeds.is_cuda and seeds.dtype == torch.int32 and seeds.is_contiguous()
                and seeds.shape[0] == self.k):
            raise TypeError("seeds must be a contiguous int32 CUDA tensor of length k")
        opts = make_options(damping, tol, max_iter, -1, chunk)
        stats = N.PrStats()
        errors = np.zeros(max(1, max_iter), np.float64)
        col_err = np.zeros(self.k, np.float64)
        N.call("gd_ppr_plan_run", self._h, seeds.data_ptr(), X.data_ptr(), ctypes.byref(opts),
               ctypes.byref(stats), errors.ctypes.data_as(ctypes.c_void_p),
               col_err.ctypes.data_as(ctypes.c_void_p))
        return stats.as_dict(), errors[: stats.iterations].copy(), col_err

    def column_iterations(self) -> np.ndarray:
        """Iterations each column ran in the last run (int32 [k])."""
        out = np.zeros(self.k, np.int32)
        N.call("gd_ppr_plan_column_iterations", self._h, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().gd_ppr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
