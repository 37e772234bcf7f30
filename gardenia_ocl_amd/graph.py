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


def csr_from_edges_device(m: int, src, dst, wt=None, *, symmetrize=False, transpose=False,
                          sorted=True, remove_selfloops=True, remove_redundents=True,
                          mirror_selfloops=False, stream=None) -> dict:
    """fill_data() on the GPU for an edge list already in HBM (int32 CUDA
    tensors src/dst, optional float32 weights): the flags and result of
    csr_from_edges.  Returns torch CUDA tensors row_offsets, column_indices,
    weight, degree (the caller's; the library's arrays are released)."""
    import torch

    for t, dt in ((src, torch.int32), (dst, torch.int32)) + (((wt, torch.float32),) if wt is not None else ()):
        if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
            raise TypeError("edge arrays must be contiguous CUDA tensors (int32 src/dst, float32 wt)")
    n = int(src.shape[0])
    if int(dst.shape[0]) != n or (wt is not None and int(wt.shape[0]) != n):
        raise ValueError("src, dst (and wt) must have the same length")
    st = stream if stream is not None else torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    c = N.DeviceCsr()
    N.call("gd_csr_from_edges_device", m, n, src.data_ptr(), dst.data_ptr(),
           None if wt is None else wt.data_ptr(), int(symmetrize), int(transpose), int(sorted),
           int(remove_selfloops), int(remove_redundents), int(mirror_selfloops), sp, ctypes.byref(c))
    try:
        dev = src.device
        out = {"row_offsets": torch.empty(m + 1, dtype=torch.int32, device=dev),
               "column_indices": torch.empty(c.nnz, dtype=torch.int32, device=dev),
               "weight": torch.empty(c.nnz, dtype=torch.float32, device=dev),
               "degree": torch.empty(m, dtype=torch.int32, device=dev)}
        N.call("gd_device_csr_copy", ctypes.byref(c), out["row_offsets"].data_ptr(),
               out["column_indices"].data_ptr(), out["weight"].data_ptr(),
               out["degree"].data_ptr(), sp)
        st.synchronize()
    finally:
        N.lib().gd_device_csr_free(ctypes.byref(c))
    return out


def stencil27(n: int) -> Csr:
    """27-point stencil matrix on an n x n x n grid (rows in x-fastest order,
    sorted columns): diagonal 26.5, off-diagonal 0.25 + a symmetric function of
    the two endpoints in [0, 1) -- a symmetric, diagonally dominant SymGS input
    shaped like HPCG's."""
    m = n ** 3
    z, y, x = (a.ravel() for a in np.meshgrid(np.arange(n), np.arange(n), np.arange(n),
                                               indexing="ij"))
    keep, delta = [], []
    for dz in (-1, 0, 1):  # (dz, dy, dx) lexicographic = increasing column
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                keep.append((z + dz >= 0) & (z + dz < n) & (y + dy >= 0) & (y + dy < n)
                            & (x + dx >= 0) & (x + dx < n))
                delta.append(dz * n * n + dy * n + dx)
    keep = np.stack(keep, axis=1)  # (m, 27)
    r = np.broadcast_to(np.arange(m, dtype=np.int64)[:, None], keep.shape)[keep]
    c = (np.arange(m, dtype=np.int64)[:, None] + np.asarray(delta, np.int64)[None, :])[keep]
    lo, hi = np.minimum(r, c), np.maximum(r, c)
    v = np.where(r == c, 26.5, 0.25 + ((lo * 7919 + hi * 104729) % 1000) / 1000.0)
    deg = keep.sum(axis=1).astype(np.int32)
    off = np.zeros(m + 1, np.int32)
    np.cumsum(deg, out=off[1:])
    return Csr(m, off, c.astype(np.int32), deg, v.astype(np.float32))


def with_dominant_diagonal(csr: Csr, margin: float = 1.0) -> Csr:
    """The matrix with every row's diagonal set to margin + the sum of the
    row's off-diagonal |weights| (existing diagonal entries replaced, missing
    ones inserted in column order): a strictly diagonally dominant SymGS input,
    whose sweeps converge instead of overflowing."""
    m = csr.m
    rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(csr.row_offsets))
    col = csr.column_indices.astype(np.int64)
    w = (csr.weight if csr.weight is not None else np.ones(len(col), np.float32)).astype(np.float32)
    off_diag = col != rows
    rows, col, w = rows[off_diag], col[off_diag], w[off_diag]
    dval = (np.bincount(rows, weights=np.abs(w).astype(np.float64), minlength=m) + margin)
    rows = np.concatenate([rows, np.arange(m, dtype=np.int64)])
    col = np.concatenate([col, np.arange(m, dtype=np.int64)])
    w = np.concatenate([w, dval.astype(np.float32)])
    # stable: each row keeps its column order, the diagonal lands in place when
    # the columns are sorted, after the row's entries otherwise
    order = np.lexsort((col, rows)) if _sorted_rows(csr) else np.argsort(rows, kind="stable")
    rows, col, w = rows[order], col[order], w[order]
    deg = np.bincount(rows, minlength=m).astype(np.int32)
    off = np.zeros(m + 1, np.int32)
    np.cumsum(deg, out=off[1:])
    return Csr(m, off, col.astype(np.int32), deg, w)


def _sorted_rows(csr: Csr) -> bool:
    rows = np.repeat(np.arange(csr.m, dtype=np.int64), np.diff(csr.row_offsets))
    key = rows * csr.m + csr.column_indices
    return bool(np.all(np.diff(key) >= 0))


def color_first_fit(csr: Csr):
    """Deterministic first-fit coloring over the CSR and its transpose
    (gd_color_first_fit).  Returns (colors int32[m], n_colors)."""
    off = np.ascontiguousarray(csr.row_offsets, np.int32)
    col = np.ascontiguousarray(csr.column_indices, np.int32)
    colors = np.empty(csr.m, np.int32)
    nc = ctypes.c_int32(0)
    N.call("gd_color_first_fit", csr.m, _ptr(off), _ptr(col), _ptr(colors), ctypes.byref(nc))
    return colors, int(nc.value)


def color_order(colors: np.ndarray, n_colors: int):
    """The SymGS driver's ordering (src/symgs/main.cc:38-49): (indices,
    color_offsets) with the vertices sorted by color, stable."""
    colors = np.ascontiguousarray(colors, np.int32)
    indices = np.empty(len(colors), np.int32)
    co = np.empty(n_colors + 1, np.int32)
    N.call("gd_color_order", len(colors), _ptr(colors), n_colors, _ptr(indices), _ptr(co))
    return indices, co


def read_mtx(path: str, *, symmetrize=False, transpose=False, sorted=True, remove_selfloops=True,
             remove_redundents=True) -> Csr:
    """mtx2csr() + degree, flags as read_graph() (include/graph_io.h:280)."""
    h = N.HostCsr()
    N.call("gd_read_mtx", os.fsencode(path), int(symmetrize), int(transpose), int(sorted),
           int(remove_selfloops), int(remove_redundents), ctypes.byref(h))
    return _take(h)


def _read(fn: str, path: str, symmetrize, transpose, sorted, remove_selfloops,
          remove_redundents) -> Csr:
    h = N.HostCsr()
    N.call(fn, os.fsencode(path), int(symmetrize), int(transpose), int(sorted),
           int(remove_selfloops), int(remove_redundents), ctypes.byref(h))
    return _take(h)


def read_gr(path: str, *, symmetrize=False, transpose=False, sorted=True, remove_selfloops=True,
            remove_redundents=True) -> Csr:
    """gr2csr() + degree (include/graph_io.h:133-174): DIMACS "a src dst"
    arcs, 1-based, weights ignored."""
    return _read("gd_read_gr", path, symmetrize, transpose, sorted, remove_selfloops,
                 remove_redundents)


def read_dotgraph(path: str, *, symmetrize=False, transpose=False, sorted=True,
                  remove_selfloops=True, remove_redundents=True) -> Csr:
    """graph2csr() + degree (include/graph_io.h:177-213): one line of 1-based
    neighbours per vertex."""
    return _read("gd_read_dotgraph", path, symmetrize, transpose, sorted, remove_selfloops,
                 remove_redundents)


def read_graph(path: str, *, symmetrize=False, transpose=False, sorted=True,
               remove_selfloops=True, remove_redundents=True) -> Csr:
    """read_graph() (include/graph_io.h:280-295): .mtx, .graph or .gr chosen
    from the file name, in that order."""
    return _read("gd_read_graph", path, symmetrize, transpose, sorted, remove_selfloops,
                 remove_redundents)


def read_pr_graph(path: str, is_directed: bool = True):
    """The two read_graph() calls of the PR driver (src/pr/main.cc:23-24):
    out-CSR then in-CSR, unsorted, self-loops and adjacent duplicates removed."""
    sym = not is_directed
    out_csr = read_graph(path, symmetrize=sym, transpose=False, sorted=False)
    in_csr = read_graph(path, symmetrize=sym, transpose=True, sorted=False)
    return out_csr, in_csr


def read_spmv_matrix(path: str, is_directed: bool = True) -> Csr:
    """The SpMV driver's read_graph() flags (src/spmv/main.cc:24): sorted,
    self-loops kept, adjacent duplicates removed."""
    return read_graph(path, symmetrize=not is_directed, transpose=False, sorted=True,
                      remove_selfloops=False, remove_redundents=True)


def write_mtx(path: str, m: int, src: np.ndarray, dst: np.ndarray) -> None:
    """Write a 1-based pattern .mtx in the given edge order."""
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n")
        f.write(f"{m} {m} {len(src)}\n")
        data = np.stack([np.asarray(src, np.int64) + 1, np.asarray(dst, np.int64) + 1], axis=1)
        np.savetxt(f, data, fmt="%d")
