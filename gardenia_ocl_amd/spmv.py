"""CSR SpMV: the reference's SpmvSolver interface over the HIP library.

``SpmvSolver`` keeps the reference signature (src/spmv/spmv.h:29) with host
arrays and y in/out (y <- y + A x, src/spmv/omp_base.cc:17-28).  ``SpmvPlan``
is the device-resident form over torch tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def SpmvSolver(m: int, nnz: int, Ap, Aj, Ax, x, y: np.ndarray, *, verbose=False,
               variant="hip_tile"):
    """y <- y + A x on the GPU; ``y`` (float32 numpy, length m) updated in place.
    Returns the kernel time in ms (HIP events)."""
    if not (isinstance(y, np.ndarray) and y.dtype == np.float32 and y.flags.c_contiguous):
        raise TypeError("y must be a contiguous float32 numpy array (updated in place)")
    Ap = np.ascontiguousarray(Ap, np.int32)
    Aj = np.ascontiguousarray(Aj, np.int32)
    Ax = np.ascontiguousarray(Ax, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    if Ap.shape[0] != m + 1 or y.shape[0] != m or x.shape[0] < m or Aj.shape[0] < nnz:
        raise ValueError("array lengths do not match m / nnz")
    ms = ctypes.c_float()
    if verbose:
        print("Launching HIP SpMV solver (gfx950 LDS tiles) ...")
    N.call("gd_spmv_solve", m, nnz, _np_ptr(Ap), _np_ptr(Aj), _np_ptr(Ax), _np_ptr(x), _np_ptr(y),
           ctypes.byref(ms))
    if verbose:
        print(f"\truntime [{variant}] = {ms.value:f} ms.")
    return ms.value


class SpmvPlan:
    """Device-resident SpMV plan over torch CUDA tensors (Ap, Aj int32; Ax fp32).

    Borrowed tensors: keep them alive and unchanged while the plan lives.
    """

    MODES = {"auto": 0, "tiles": 1, "sliced": 2}

    def __init__(self, Ap, Aj, Ax, num_cols=None, stream=None, mode="tiles"):
        """mode: "tiles" (the tile engine), "sliced" (one tile-engine pass per
        column slice of x sized for an XCD's L2, sorted rows only), "auto"
        (sliced when x exceeds one XCD's L2 and rows are short)."""
        import torch

        for t, dt in ((Ap, torch.int32), (Aj, torch.int32), (Ax, torch.float32)):
            if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
                raise TypeError("SpmvPlan inputs must be contiguous CUDA tensors (int32, int32, fp32)")
        self.rows = int(Ap.shape[0]) - 1
        self.cols = int(num_cols) if num_cols is not None else self.rows
        self.nnz = int(Aj.shape[0])
        if int(Ax.shape[0]) != self.nnz:
            raise ValueError("Aj and Ax lengths differ")
        self._keep = (Ap, Aj, Ax)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        N.call("gd_spmv_plan_create_ex", self.rows, self.cols, self.nnz, Ap.data_ptr(),
               Aj.data_ptr(), Ax.data_ptr(), ctypes.c_void_p(self.stream.cuda_stream),
               self.MODES[mode], ctypes.byref(h))
        self._h = h
        self.mode = {1: "tiles", 2: "sliced"}[N.lib().gd_spmv_plan_mode(h)]

    def refresh_values(self):
        """Re-read Ax after it changed (needed by the sliced mode)."""
        N.call("gd_spmv_plan_refresh_values", self._h)

    def info(self) -> dict:
        t, hb = ctypes.c_int64(), ctypes.c_int64()
        mx = ctypes.c_int32()
        N.call("gd_spmv_plan_info", self._h, ctypes.byref(t), ctypes.byref(hb), ctypes.byref(mx))
        return {"tiles": t.value, "hubs": hb.value, "max_row_len": mx.value}

    def run(self, x, y):
        """y <- y + A x (float32 CUDA tensors), asynchronous on the plan's stream."""
        import torch

        if not (x.is_cuda and y.is_cuda and x.dtype == torch.float32 and y.dtype == torch.float32
                and x.is_contiguous() and y.is_contiguous()):
            raise TypeError("x, y must be contiguous float32 CUDA tensors")
        if int(x.shape[0]) < self.cols or int(y.shape[0]) != self.rows:
            raise ValueError("x / y lengths do not match the matrix")
        N.call("gd_spmv_plan_run", self._h, x.data_ptr(), y.data_ptr())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().gd_spmv_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
