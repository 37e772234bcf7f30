"""The FP64 reference bench.py reports L1 distances against (torch CSR SpMV in
float64) is pinned to the oracle's FP64 power iteration and to a dense solve."""
import os
import sys

import numpy as np

import oracle
from gardenia_ocl_amd import graph

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _tensors(scale, seed):
    import torch

    _, ic = graph.rmat(scale, 16, seed=seed, want_out=False)
    od = np.bincount(ic.column_indices, minlength=ic.m).astype(np.int32)
    t = torch.from_numpy
    return ic, od, t(ic.row_offsets), t(ic.column_indices), t(od)


def test_fp64_reference_matches_oracle_pr_fp64():
    ic, od, off, col, deg = _tensors(10, 3)
    x, it = bench.fp64_pagerank(off, col, deg, tol=1e-14)
    x64, _ = oracle.pr_fp64(ic.m, ic.row_offsets, ic.column_indices, od, tol=1e-14)
    assert np.abs(x.numpy() - x64).max() <= 1e-15 + 1e-12 * np.abs(x64).max()


def test_fp64_reference_dense_solve():
    """x = (1-d)/V (I - d A D^-1)^-1 1 on a small RMAT (dangling mass dropped)."""
    ic, od, off, col, deg = _tensors(8, 5)
    V = ic.m
    A = np.zeros((V, V))
    for v in range(V):
        for e in range(ic.row_offsets[v], ic.row_offsets[v + 1]):
            u = ic.column_indices[e]
            A[v, u] += 1.0 / od[u]
    xd = np.linalg.solve(np.eye(V) - 0.85 * A, np.full(V, 0.15 / V))
    x, _ = bench.fp64_pagerank(off, col, deg, tol=1e-15)
    np.testing.assert_allclose(x.numpy(), xd, rtol=1e-10, atol=1e-16)


def test_fp64_reference_personalised_matches_oracle():
    ic, od, off, col, deg = _tensors(10, 4)
    seeds = np.array([0, 7, 513], np.int32)
    X, _ = bench.fp64_pagerank(off, col, deg, seeds=seeds, tol=1e-14)
    X64, _ = oracle.ppr_fp64(ic.m, ic.row_offsets, ic.column_indices, od, seeds, tol=1e-14)
    assert np.abs(X.numpy() - X64).max() <= 1e-12
