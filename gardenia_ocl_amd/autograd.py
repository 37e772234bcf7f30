"""torch.autograd bindings of the hand-written adjoints (ctypes -> C ABI).

* ``weighted_pagerank(weight, plan)``: x = (1-d)/m + d P x with
  P[v,u] = a_uv / sum_v' a_uv' for edge weights ``weight`` (in-CSR order).
  Backward: the implicit adjoint lambda = g + d P^T lambda on the reversed
  graph, then dL/da per edge (gd_wpr_backward).  No iterates are stored.
* ``spmv(values, x, plan, bwd_plan)``: y = A x with A's nonzeros ``values``;
  backward dx = A^T g, dvalues[k] = g[row(k)] x[col(k)] (gd_spmv_bwd_run).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N
from .pr import make_options


class WeightedPageRankPlan:
    """Plan over a device in-CSR and a float32 weight tensor (borrowed)."""

    def __init__(self, in_row_offsets, in_column_indices, weight, stream=None):
        for t, dt in ((in_row_offsets, torch.int32), (in_column_indices, torch.int32),
                      (weight, torch.float32)):
            if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
                raise TypeError("inputs must be contiguous CUDA tensors (int32, int32, float32)")
        self.m = int(in_row_offsets.shape[0]) - 1
        self.nnz = int(in_column_indices.shape[0])
        if int(weight.shape[0]) != self.nnz:
            raise ValueError("one weight per edge")
        self.weight = weight
        self._keep = (in_row_offsets, in_column_indices, weight)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        N.call("gd_wpr_plan_create", self.m, self.nnz, in_row_offsets.data_ptr(),
               in_column_indices.data_ptr(), weight.data_ptr(),
               ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        self._h = h

    def forward(self, x, *, damping=0.85, tol=1e-7, max_iter=200):
        opts = make_options(damping, tol, max_iter)
        st = N.PrStats()
        errs = np.zeros(max(1, max_iter), np.float64)
        N.call("gd_wpr_forward", self._h, x.data_ptr(), ctypes.byref(opts), ctypes.byref(st),
               errs.ctypes.data_as(ctypes.c_void_p))
        return st.as_dict()

    def backward(self, x, w, lam, grad, *, damping=0.85, tol=1e-7, max_iter=200):
        opts = make_options(damping, tol, max_iter)
        st = N.PrStats()
        N.call("gd_wpr_backward", self._h, x.data_ptr(), w.data_ptr(), lam.data_ptr(),
               grad.data_ptr(), ctypes.byref(opts), ctypes.byref(st))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().gd_wpr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _WeightedPageRankFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weight, plan, damping, tol, max_iter):
        if weight.data_ptr() != plan.weight.data_ptr():
            raise ValueError("weight must be the tensor the plan was built over")
        x = torch.full((plan.m,), 1.0 / plan.m, dtype=torch.float32, device=weight.device)
        plan.forward(x, damping=damping, tol=tol, max_iter=max_iter)
        ctx.save_for_backward(x)
        ctx.plan, ctx.damping, ctx.tol, ctx.max_iter = plan, damping, tol, max_iter
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        plan = ctx.plan
        lam = torch.empty_like(x)
        grad = torch.empty(plan.nnz, dtype=torch.float32, device=x.device)
        plan.backward(x, grad_x.contiguous().float(), lam, grad, damping=ctx.damping, tol=ctx.tol,
                      max_iter=ctx.max_iter)
        return grad, None, None, None, None


def weighted_pagerank(weight, plan, *, damping=0.85, tol=1e-7, max_iter=200):
    """Differentiable weighted PageRank scores for the edge weights ``weight``."""
    return _WeightedPageRankFn.apply(weight, plan, damping, tol, max_iter)


class SpmvBackwardPlan:
    """Transposed structure of a device CSR (built once) for SpMV backward."""

    def __init__(self, Ap, Aj, num_cols=None, stream=None):
        self.rows = int(Ap.shape[0]) - 1
        self.cols = int(num_cols) if num_cols is not None else self.rows
        self.nnz = int(Aj.shape[0])
        self._keep = (Ap, Aj)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        N.call("gd_spmv_bwd_plan_create", self.rows, self.cols, self.nnz, Ap.data_ptr(),
               Aj.data_ptr(), ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        self._h = h

    def run(self, values, x, g, dx=None, dA=None):
        N.call("gd_spmv_bwd_run", self._h, values.data_ptr(), x.data_ptr(), g.data_ptr(),
               None if dx is None else dx.data_ptr(), None if dA is None else dA.data_ptr())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().gd_spmv_bwd_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _SpmvFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, x, plan, bwd_plan):
        y = torch.zeros(plan.rows, dtype=torch.float32, device=x.device)
        plan.run(x.contiguous(), y)
        ctx.save_for_backward(values, x)
        ctx.bwd_plan = bwd_plan
        return y

    @staticmethod
    def backward(ctx, g):
        values, x = ctx.saved_tensors
        g = g.contiguous().float()
        dx = torch.empty(ctx.bwd_plan.cols, dtype=torch.float32, device=g.device) \
            if ctx.needs_input_grad[1] else None
        dA = torch.empty_like(values) if ctx.needs_input_grad[0] else None
        ctx.bwd_plan.run(values, x.contiguous(), g, dx, dA)
        return dA, dx, None, None


def spmv(values, x, plan, bwd_plan):
    """Differentiable y = A x; ``plan`` is an SpmvPlan built over ``values``."""
    return _SpmvFn.apply(values, x, plan, bwd_plan)
