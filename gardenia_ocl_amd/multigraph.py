"""PageRank on many independent graphs in one launch per iteration
(gd_multi_pr_*): block-diagonal packing, per-graph convergence, iterations
replayed from a captured hipGraph."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .pr import EPSILON, K_DAMP, MAX_ITER, make_options


class MultiGraphPageRank:
    """``graphs``: list of (in_row_offsets, in_column_indices, out_degree) int32
    CUDA tensors.  The plan packs its own copy; the tensors may be freed after."""

    def __init__(self, graphs, stream=None):
        import torch

        self.n = len(graphs)
        self.m = np.array([int(g[2].shape[0]) for g in graphs], np.int32)
        self.nnz = np.array([int(g[1].shape[0]) for g in graphs], np.int64)
        ptrs = lambda i: (ctypes.c_void_p * self.n)(*[g[i].data_ptr() for g in graphs])  # noqa: E731
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        N.call("gd_multi_pr_plan_create", self.n, self.m.ctypes.data_as(ctypes.c_void_p),
               self.nnz.ctypes.data_as(ctypes.c_void_p), ptrs(0), ptrs(1), ptrs(2),
               ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        self._h = h

    def run(self, scores, *, damping=K_DAMP, tol=EPSILON, max_iter=MAX_ITER, use_graph=True,
            chunk=0):
        """``scores``: list of float32 CUDA tensors (in/out).  Returns
        (iterations per graph, final L1 per graph, solve ms)."""
        opts = make_options(damping, tol, max_iter)
        cfg = N.MultiPrConfig(int(bool(use_graph)), int(chunk))
        its = np.zeros(self.n, np.int32)
        errs = np.zeros(self.n, np.float64)
        ms = ctypes.c_float()
        ptrs = (ctypes.c_void_p * self.n)(*[s.data_ptr() for s in scores])
        N.call("gd_multi_pr_run", self._h, ptrs, ctypes.byref(opts), ctypes.byref(cfg),
               its.ctypes.data_as(ctypes.c_void_p), errs.ctypes.data_as(ctypes.c_void_p),
               ctypes.byref(ms))
        return its, errs, ms.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().gd_multi_pr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
