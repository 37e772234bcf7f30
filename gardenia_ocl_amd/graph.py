"""Host graph layer: CSR arrays as numpy, built by the native library.

Mirrors the reference's ``read_graph`` (include/graph_io.h:280-295) and adds
the seeded synthetic inputs the benchmarks use (RMAT, Erdos-Renyi).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native as N

# RMAT quadrant probabilities of the benchmark configurations (d = 0.05).
RMAT_A, RMAT_B, RMAT_C = 0.57, 0.19, 0.19


@dataclass
class Csr:
    """int32 CSR as read_graph() returns it; ``degree[i]`` = row length."""

    m: int
    row_offsets: np.ndarray
    column_indices: np.ndarray
    degree: np.ndarray
    weight: Optional[np.ndarray] = None

    @property
    def nnz(self) -> int:
        return int(self.column_indices.shape[0])


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _take(h: N.HostCsr) -> Csr:
    m, nnz = int(h.m), int(h.nnz)
    try:
        off = np.ctypeslib.as_array(h.row_offsets, shape=(m + 1,)).copy()
        col = (np.ctypeslib.as_array(h.column_indices, shape=(nnz,)).copy() if nnz
               else np.zeros(0, np.int32))
        deg = (np.ctypeslib.as_array(h.degree, shape=(m,)).copy() if m
               else np.zeros(0, np.int32))
        w = None
        if bool(h.weight) and nnz:
            w = np.ctypeslib.as_array(h.weight, shape=(nnz,)).copy()
        elif bool(h.weight):
            w = np.zeros(0, np.float32)
    finally:
        N.lib().gd_csr_free(ctypes.byref(h))
    return Csr(m, off, col, deg, w)


def rmat(scale: int, edgefactor: int = 16, a: float = RMAT_A, b: float = RMAT_B,
         c: float = RMAT_C, seed: int = 1, want_out: bool = True, want_in: bool = True):
    """Canonical RMAT graph -> (out_csr, in_csr) (either may be None)."""
    ho, hi = N.HostCsr(), N.HostCsr()
    N.call("gd_rmat_graph", scale, edgefactor, a, b, c, seed,
           ctypes.byref(ho) if want_out else None, ctypes.byref(hi) if want_in else None)
    return (_take(ho) if want_out else None, _take(hi) if want_in else None)


def rmat_edges(scale: int, edgefactor: int = 16, a: float = RMAT_A, b: float = RMAT_B,
               c: float = RMAT_C, seed: int = 1):
    """Raw RMAT edge list (before self-loop/duplicate removal) as (src, dst)."""
    n = edgefactor << scale
    src = np.empty(n, np.int32)
    dst = np.empty(n, np.int32)
    N.call("gd_rmat_edges", scale, edgefactor, a, b, c, seed, 0, n, _ptr(src), _ptr(dst))
    return src, dst


def er_matrix(log2_rows: int, nnz_per_row: int, seed: int = 1) -> Csr:
    """Erdos-Renyi matrix with exactly nnz_per_row sorted distinct columns/row and
    weight ~ U(-1,1)."""
    h = N.HostCsr()
    N.call("gd_er_matrix", log2_rows, nnz_per_row, seed, ctypes.byref(h))
    return _take(h)


def uniform_pm1(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    N.call("gd_fill_uniform_pm1", seed, n, _ptr(out))
    return out


def csr_from_edges(m: int, src: np.ndarray, dst: np.ndarray, wt: Optional[np.ndarray] = None, *,
                   symmetrize=False, transpose=False, sorted=True, remove_selfloops=True,
                   remove_redundents=True, mirror_selfloops=False) -> Csr:
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    if wt is not None:
        wt = np.ascontiguousarray(wt, np.float32)
    h = N.HostCsr()
    N.call("gd_csr_from_edges", m, src.shape[0], _ptr(src), _ptr(dst), _ptr(wt), int(symmetrize),
           int(transpose), int(sorted), int(remove_selfloops), int(remove_redundents),
           int(mirror_selfloops), ctypes.byref(h))
    return _take(h)


def read_mtx(path: str, *, symmetrize=False, transpose=False, sorted=True, remove_selfloops=True,
             remove_redundents=True) -> Csr:
    """mtx2csr() + degree, flags as read_graph() (include/graph_io.h:280)."""
    h = N.HostCsr()
    N.call("gd_read_mtx", os.fsencode(path), int(symmetrize), int(transpose), int(sorted),
           int(remove_selfloops), int(remove_redundents), ctypes.byref(h))
    return _take(h)


def read_pr_graph(path: str, is_directed: bool = True):
    """The two read_graph() calls of the PR driver (src/pr/main.cc:23-24):
    out-CSR then in-CSR, unsorted, self-loops and adjacent duplicates removed."""
    sym = not is_directed
    out_csr = read_mtx(path, symmetrize=sym, transpose=False, sorted=False)
    in_csr = read_mtx(path, symmetrize=sym, transpose=True, sorted=False)
    return out_csr, in_csr


def read_spmv_matrix(path: str, is_directed: bool = True) -> Csr:
    """The SpMV driver's read_graph() flags (src/spmv/main.cc:24): sorted,
    self-loops kept, adjacent duplicates removed."""
    return read_mtx(path, symmetrize=not is_directed, transpose=False, sorted=True,
                    remove_selfloops=False, remove_redundents=True)


def write_mtx(path: str, m: int, src: np.ndarray, dst: np.ndarray) -> None:
    """Write a 1-based pattern .mtx in the given edge order."""
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n")
        f.write(f"{m} {m} {len(src)}\n")
        data = np.stack([np.asarray(src, np.int64) + 1, np.asarray(dst, np.int64) + 1], axis=1)
        np.savetxt(f, data, fmt="%d")
