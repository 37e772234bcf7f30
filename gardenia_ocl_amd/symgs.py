"""Multicolor symmetric Gauss-Seidel: the reference's SymGSSolver interface
over the HIP library.

``SymGSSolver`` keeps the reference signature (src/symgs/symgs.h:28) with host
arrays and x updated in place by one symmetric sweep (colors forward, then
backward; src/symgs/omp_base.cc:27-42).  ``SymGSPlan`` is the device-resident
form over torch tensors, for repeated sweeps.  Results are bit-exact with the
reference's sequential sweep (src/symgs/verifier.cc:50-77) for a proper
coloring; ``graph.color_first_fit`` + ``graph.color_order`` give one.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def SymGSSolver(num_rows: int, nnz: int, Ap, Aj, indices, Ax, x: np.ndarray, b,
                color_offsets, *, verbose=False, variant="hip_tile"):
    """One symmetric sweep on the GPU; ``x`` (float32 numpy, length num_rows)
    updated in place.  Returns the device time of the sweep in ms."""
    if not (isinstance(x, np.ndarray) and x.dtype == np.float32 and x.flags.c_contiguous):
        raise TypeError("x must be a contiguous float32 numpy array (updated in place)")
    Ap = np.ascontiguousarray(Ap, np.int32)
    Aj = np.ascontiguousarray(Aj, np.int32)
    Ax = np.ascontiguousarray(Ax, np.float32)
    idx = np.ascontiguousarray(indices, np.int32)
    b = np.ascontiguousarray(b, np.float32)
    co = np.ascontiguousarray(color_offsets, np.int32)
    if (Ap.shape[0] != num_rows + 1 or x.shape[0] != num_rows or b.shape[0] != num_rows
            or idx.shape[0] != num_rows or Aj.shape[0] < nnz or Ax.shape[0] < nnz
            or co.shape[0] < 1):
        raise ValueError("array lengths do not match num_rows / nnz")
    ms = ctypes.c_float()
    if verbose:
        print("Launching HIP SymGS solver (gfx950 LDS tiles, exact row order) ...")
    N.call("gd_symgs_solve", num_rows, nnz, _np_ptr(Ap), _np_ptr(Aj), _np_ptr(idx), _np_ptr(Ax),
           _np_ptr(x), _np_ptr(b), _np_ptr(co), int(co.shape[0]) - 1, ctypes.byref(ms))
    if verbose:
        print(f"\truntime [{variant}] = {ms.value:f} ms.")
    return ms.value


class SymGSPlan:
    """Device-resident SymGS plan.  Ap, Aj, indices: int32 CUDA tensors; Ax:
    float32 CUDA tensor; color_offsets: host int sequence (n_colors + 1).  The
    plan copies what it needs (rows in sweep order, diagonal split off)."""

    def __init__(self, Ap, Aj, Ax, indices, color_offsets, stream=None):
        import torch

        for t, dt in ((Ap, torch.int32), (Aj, torch.int32), (Ax, torch.float32),
                      (indices, torch.int32)):
            if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
                raise TypeError("SymGSPlan inputs must be contiguous CUDA tensors")
        self.m = int(Ap.shape[0]) - 1
        self.nnz = int(Aj.shape[0])
        if int(Ax.shape[0]) != self.nnz or int(indices.shape[0]) != self.m:
            raise ValueError("Aj / Ax / indices lengths do not match the matrix")
        co = np.ascontiguousarray(color_offsets, np.int32)
        if co.ndim != 1 or co.shape[0] < 1:
            raise ValueError("color_offsets needs n_colors + 1 >= 1 entries")
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        N.call("gd_symgs_plan_create", self.m, self.nnz, Ap.data_ptr(), Aj.data_ptr(),
               Ax.data_ptr(), indices.data_ptr(), _np_ptr(co), int(co.shape[0]) - 1,
               ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        self._h = h

    def info(self) -> dict:
        i = N.SymgsPlanInfo()
        N.call("gd_symgs_plan_info", self._h, ctypes.byref(i))
        return {f: getattr(i, f) for f, _ in N.SymgsPlanInfo._fields_}

    def run(self, x, b, sweeps: int = 1):
        """`sweeps` symmetric sweeps on x in place (float32 CUDA tensors of
        length m), asynchronous on the plan's stream."""
        import torch

        for t in (x, b):
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                    and int(t.shape[0]) == self.m):
                raise TypeError("x, b must be contiguous float32 CUDA tensors of length m")
        N.call("gd_symgs_plan_run", self._h, x.data_ptr(), b.data_ptr(), int(sweeps),
               ctypes.c_void_p(self.stream.cuda_stream))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().gd_symgs_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
