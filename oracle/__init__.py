"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/liboracle.so, a C++ restatement of the reference's
OpenMP PageRank / SpMV path (src/pr/omp_base.cc, src/spmv/omp_base.cc,
src/pr/verifier.cc, src/spmv/verifier.cc, include/graph_io.h).  Only tests/,
``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py may import
this package, and only as the checker.  The product (gardenia_ocl_amd) never
imports it.

``load_ref()`` loads oracle/_ref/libref.so, the reference's own sources built
from /root/reference by ``make -C oracle ref``; it exists only where the
reference is mounted and is used to write golden vectors (tests/golden/).

Every function here is pinned by a test (see the header of oracle.cpp);
none is "parity unpinned".
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref.so")

_P = ctypes.c_void_p
_i, _i64, _f, _d = ctypes.c_int, ctypes.c_longlong, ctypes.c_float, ctypes.c_double

_SIG = {
    "orc_csr_from_edges": (_P, [_i, _i64, _P, _P, _P, _i, _i, _i, _i, _i, _i]),
    "orc_csr_from_mtx": (_P, [ctypes.c_char_p, _i, _i, _i, _i, _i]),
    "orc_csr_info": (None, [_P, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "orc_csr_copy": (None, [_P, _P, _P, _P, _P]),
    "orc_csr_free": (None, [_P]),
    "orc_pr": (_i, [_i, _P, _P, _P, _P, _f, _d, _i, _P]),
    "orc_ppr": (_i, [_i, _P, _P, _P, _i, _P, _f, _d, _i, _P]),
    "orc_pr_residual": (_d, [_i, _P, _P, _P, _P, _f]),
    "orc_pr_push": (_i, [_i, _P, _P, _P, _f, _d, _i, _P]),
    "orc_pr_fp64": (_i, [_i, _P, _P, _P, _P, _d, _d, _i]),
    "orc_ppr_fp64": (_i, [_i, _P, _P, _P, _i, _P, _P, _d, _d, _i]),
    "orc_spmv": (None, [_i, _P, _P, _P, _P, _P]),
    "orc_max_rel_err": (_f, [_P, _P, _i64]),
    "orc_spmv_bwd_fp64": (None, [_i, _i, _P, _P, _P, _P, _P, _P, _P]),
    "orc_wpr_fp64": (_i, [_i, _P, _P, _P, _P, _d, _d, _i, _i]),
    "orc_wpr_grad_fp64": (_i, [_i, _P, _P, _P, _P, _P, _d, _d, _i, _P, _P]),
    "orc_symgs": (None, [_P, _P, _P, _P, _P, _P, _P, _i, _i]),
    "orc_num_threads": (_i, []),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(LIB_PATH)
        for n, (r, a) in _SIG.items():
            fn = getattr(h, n)
            fn.restype, fn.argtypes = r, a
        _lib = h
    return _lib


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


class OCsr:
    def __init__(self, m, off, col, w, deg):
        self.m, self.row_offsets, self.column_indices, self.weight, self.degree = m, off, col, w, deg

    @property
    def nnz(self):
        return int(self.column_indices.shape[0])


def _take(h) -> OCsr:
    if not h:
        raise RuntimeError("oracle CSR build failed")
    m, nnz = _i(), _i()
    lib().orc_csr_info(h, ctypes.byref(m), ctypes.byref(nnz))
    off = np.zeros(m.value + 1, np.int32)
    col = np.zeros(nnz.value, np.int32)
    w = np.zeros(nnz.value, np.float32)
    deg = np.zeros(m.value, np.int32)
    lib().orc_csr_copy(h, _p(off), _p(col), _p(w), _p(deg))
    lib().orc_csr_free(h)
    return OCsr(m.value, off, col, w, deg)


def csr_from_edges(m, src, dst, wt=None, *, symmetrize=False, transpose=False, sorted=True,
                   remove_selfloops=True, remove_redundents=True, mirror_selfloops=False) -> OCsr:
    """fill_data() restatement (include/graph_io.h:21-130)."""
    src, dst = _i32(src), _i32(dst)
    wt = None if wt is None else np.ascontiguousarray(wt, np.float32)
    h = lib().orc_csr_from_edges(m, len(src), _p(src), _p(dst), _p(wt), int(symmetrize),
                                 int(transpose), int(sorted), int(remove_selfloops),
                                 int(remove_redundents), int(mirror_selfloops))
    return _take(h)


def csr_from_mtx(path, *, symmetrize=False, transpose=False, sorted=True, remove_selfloops=True,
                 remove_redundents=True) -> OCsr:
    """mtx2csr() restatement (include/graph_io.h:216-261)."""
    h = lib().orc_csr_from_mtx(os.fsencode(path), int(symmetrize), int(transpose), int(sorted),
                               int(remove_selfloops), int(remove_redundents))
    return _take(h)


def pr(m, in_off, in_col, degree, scores=None, *, damping=0.85, tol=1e-3, max_iter=30):
    """src/pr/omp_base.cc:9-49.  Returns (scores, iterations, errors)."""
    in_off, in_col, degree = _i32(in_off), _i32(in_col), _i32(degree)
    s = (np.full(m, np.float32(1.0) / np.float32(m), np.float32) if scores is None
         else np.array(scores, np.float32, copy=True))
    errs = np.zeros(max(1, max_iter), np.float64)
    it = lib().orc_pr(m, _p(in_off), _p(in_col), _p(degree), _p(s), damping, tol, max_iter,
                      _p(errs))
    return s, it, errs[:it].copy()


def ppr(m, in_off, in_col, degree, seed, *, damping=0.85, tol=1e-3, max_iter=30):
    """FP32 personalised PR (teleport to one seed), init one-hot."""
    in_off, in_col, degree = _i32(in_off), _i32(in_col), _i32(degree)
    s = np.zeros(m, np.float32)
    s[seed] = 1.0
    errs = np.zeros(max(1, max_iter), np.float64)
    it = lib().orc_ppr(m, _p(in_off), _p(in_col), _p(degree), int(seed), _p(s), damping, tol,
                       max_iter, _p(errs))
    return s, it, errs[:it].copy()


def pr_push(m, out_off, out_col, scores=None, *, damping=0.85, tol=1e-3, max_iter=30):
    """src/pr/scatter.cu:11-41 (push over the out-CSR, additions in out-CSR
    order).  Returns (scores, iterations, errors)."""
    out_off, out_col = _i32(out_off), _i32(out_col)
    s = (np.full(m, np.float32(1.0) / np.float32(m), np.float32) if scores is None
         else np.array(scores, np.float32, copy=True))
    errs = np.zeros(max(1, max_iter), np.float64)
    it = lib().orc_pr_push(m, _p(out_off), _p(out_col), _p(s), damping, tol, max_iter, _p(errs))
    return s, it, errs[:it].copy()


def pr_residual(m, out_off, out_col, degree, scores, damping=0.85) -> float:
    """PRVerifier's push residual (src/pr/verifier.cc:50-68)."""
    out_off, out_col, degree = _i32(out_off), _i32(out_col), _i32(degree)
    scores = np.ascontiguousarray(scores, np.float32)
    return lib().orc_pr_residual(m, _p(out_off), _p(out_col), _p(degree), _p(scores), damping)


def pr_fp64(m, in_off, in_col, degree, *, damping=0.85, tol=1e-14, max_iter=1000):
    in_off, in_col, degree = _i32(in_off), _i32(in_col), _i32(degree)
    x = np.zeros(m, np.float64)
    it = lib().orc_pr_fp64(m, _p(in_off), _p(in_col), _p(degree), _p(x), damping, tol, max_iter)
    return x, it


def ppr_fp64(m, in_off, in_col, degree, seeds, *, damping=0.85, tol=1e-13, max_iter=1000):
    in_off, in_col, degree = _i32(in_off), _i32(in_col), _i32(degree)
    seeds = _i32(seeds)
    k = len(seeds)
    X = np.zeros((m, k), np.float64)
    it = lib().orc_ppr_fp64(m, _p(in_off), _p(in_col), _p(degree), k, _p(seeds), _p(X), damping,
                            tol, max_iter)
    return X, it


def spmv(m, Ap, Aj, Ax, x, y):
    """src/spmv/omp_base.cc:17-28; returns a new y."""
    Ap, Aj = _i32(Ap), _i32(Aj)
    Ax, x = np.ascontiguousarray(Ax, np.float32), np.ascontiguousarray(x, np.float32)
    out = np.array(y, np.float32, copy=True)
    lib().orc_spmv(m, _p(Ap), _p(Aj), _p(Ax), _p(x), _p(out))
    return out


def symgs(Ap, Aj, indices, Ax, x, b, color_offsets, sweeps=1):
    """src/symgs/verifier.cc:50-77 (gs_serial over the colors forward, then
    backward); returns a new x."""
    Ap, Aj, indices, co = _i32(Ap), _i32(Aj), _i32(indices), _i32(color_offsets)
    Ax, b = np.ascontiguousarray(Ax, np.float32), np.ascontiguousarray(b, np.float32)
    out = np.array(x, np.float32, copy=True)
    lib().orc_symgs(_p(Ap), _p(Aj), _p(indices), _p(Ax), _p(out), _p(b), _p(co), len(co) - 1,
                    sweeps)
    return out


def max_rel_err(a, b) -> float:
    """maximum_relative_error<float> (src/spmv/verifier.cc:9-22)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().orc_max_rel_err(_p(a), _p(b), len(a))


def spmv_bwd_fp64(rows, cols, Ap, Aj, Ax, x, g):
    Ap, Aj = _i32(Ap), _i32(Aj)
    Ax, x, g = (np.ascontiguousarray(v, np.float32) for v in (Ax, x, g))
    dx = np.zeros(cols, np.float64)
    dA = np.zeros(len(Aj), np.float64)
    lib().orc_spmv_bwd_fp64(rows, cols, _p(Ap), _p(Aj), _p(Ax), _p(x), _p(g), _p(dx), _p(dA))
    return dx, dA


def wpr_fp64(m, in_off, in_col, a, *, damping=0.85, tol=1e-14, max_iter=2000, x0=None):
    """FP64 weighted PR (Jacobi); x0: start vector (default 1/m)."""
    in_off, in_col = _i32(in_off), _i32(in_col)
    a = np.ascontiguousarray(a, np.float64)
    x = np.zeros(m, np.float64) if x0 is None else np.array(x0, np.float64)
    it = lib().orc_wpr_fp64(m, _p(in_off), _p(in_col), _p(a), _p(x), damping, tol, max_iter,
                            0 if x0 is None else 1)
    return x, it


def wpr_grad_fp64(m, in_off, in_col, a, x, w, *, damping=0.85, tol=1e-14, max_iter=2000):
    in_off, in_col = _i32(in_off), _i32(in_col)
    a, x, w = (np.ascontiguousarray(v, np.float64) for v in (a, x, w))
    lam = np.zeros(m, np.float64)
    grad = np.zeros(len(in_col), np.float64)
    it = lib().orc_wpr_grad_fp64(m, _p(in_off), _p(in_col), _p(a), _p(x), _p(w), damping, tol,
                                 max_iter, _p(lam), _p(grad))
    return lam, grad, it


def num_threads() -> int:
    return lib().orc_num_threads()


def ref_available() -> bool:
    """True where oracle/_ref/libref.so (the reference's own OpenMP sources,
    compiled here by ``make -C oracle ref``) has been built."""
    return os.path.exists(REF_LIB_PATH)


def ref_pr_solve(m, nnz, in_off, in_col, degree, scores, log_path):
    """The reference's own PRSolver (src/pr/omp_base.cc:9-49, built from
    source with its compile-time EPSILON 1e-3 and MAX_ITER 30) on host arrays;
    ``scores`` (float32) is updated in place and the reference's stdout goes to
    ``log_path``.  Returns (printed iteration count, printed runtime in ms from
    the reference's own Timer, printed per-iteration L1 values).  The out-CSR
    arguments are unused by that solver and passed as NULL."""
    import re

    h = ctypes.CDLL(REF_LIB_PATH)
    fn = h.ref_pr_solve
    fn.restype = None
    fn.argtypes = [_i, _i, _P, _P, _P, _P, _P, _P, ctypes.c_char_p]
    in_off, in_col, degree = _i32(in_off), _i32(in_col), _i32(degree)
    assert scores.dtype == np.float32 and scores.flags.c_contiguous
    fn(m, nnz, _p(in_off), _p(in_col), None, None, _p(degree), _p(scores), os.fsencode(log_path))
    with open(log_path) as f:
        text = f.read()
    its = int(re.findall(r"iterations = (\d+)\.", text)[-1])
    ms = float(re.findall(r"runtime \[[^\]]*\] = ([0-9.]+) ms\.", text)[-1])
    errs = [float(e) for e in re.findall(r"^ *\d+ +([0-9.]+)$", text, flags=re.M)]
    return its, ms, errs
