"""Pins of the oracle that do not depend on running the reference: closed forms,
dense solves, finite differences and invariants.  Each one is chosen so that a
plausible mistake (transposed CSR, in-degree for out-degree, dropped teleport
term, FMA-contracted update) fails it."""
import numpy as np
import pytest

import oracle
from conftest import dataset
from gardenia_ocl_amd import graph


def dense_pull(m, in_off, in_col, outdeg):
    """M[v,u] = (#edges u->v) / outdeg[u] from an in-CSR."""
    M = np.zeros((m, m))
    for v in range(m):
        for e in range(in_off[v], in_off[v + 1]):
            u = in_col[e]
            M[v, u] += 1.0 / outdeg[u]
    return M


def pr_dense(m, in_off, in_col, outdeg, d=0.85):
    M = dense_pull(m, in_off, in_col, outdeg)
    return np.linalg.solve(np.eye(m) - d * M, np.full(m, (1 - d) / m))


def test_pr_fixed_point_test_pr_closed_form():
    o, i = graph.read_pr_graph(dataset("test_pr.mtx"))
    x, _ = oracle.pr_fp64(o.m, i.row_offsets, i.column_indices, o.degree, tol=1e-15)
    expect = np.array([0.364153956, 0.196839976, 0.192265431, 0.246740637])
    np.testing.assert_allclose(x, expect, atol=2e-9)
    np.testing.assert_allclose(x, pr_dense(o.m, i.row_offsets, i.column_indices, o.degree),
                               rtol=1e-12)
    assert abs(x.sum() - 1.0) < 1e-9  # no dangling vertices: mass conserved


@pytest.mark.parametrize("tol,iters", [(1e-3, 11), (1e-4, 15), (1e-6, 23)])
def test_pr_fp32_iteration_counts_test_pr(tol, iters):
    o, i = graph.read_pr_graph(dataset("test_pr.mtx"))
    s, it, errs = oracle.pr(o.m, i.row_offsets, i.column_indices, o.degree, tol=tol, max_iter=100)
    assert it == iters
    assert errs[-1] < tol <= errs[-2]
    x = pr_dense(o.m, i.row_offsets, i.column_indices, o.degree)
    assert np.abs(s - x).sum() < 10 * tol


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pr_fp64_matches_dense_solve_random(seed):
    rng = np.random.default_rng(seed)
    m = 300
    src = rng.integers(0, m, 2500).astype(np.int32)
    dst = (src + 1 + rng.integers(0, m - 1, 2500)).astype(np.int32) % m  # no self-loops
    o = graph.csr_from_edges(m, src, dst)
    i = graph.csr_from_edges(m, src, dst, transpose=True)
    x, _ = oracle.pr_fp64(m, i.row_offsets, i.column_indices, o.degree, tol=1e-15)
    np.testing.assert_allclose(x, pr_dense(m, i.row_offsets, i.column_indices, o.degree),
                               rtol=1e-10)
    # planted mistakes are caught by the dense pin:
    xt, _ = oracle.pr_fp64(m, o.row_offsets, o.column_indices, o.degree, tol=1e-15)  # transposed
    assert np.abs(xt - x).max() > 1e-6
    xi, _ = oracle.pr_fp64(m, i.row_offsets, i.column_indices, i.degree, tol=1e-15)  # in-degree
    assert np.abs(xi - x).max() > 1e-6
    # FP32 oracle converges to the same point
    s, it, _ = oracle.pr(m, i.row_offsets, i.column_indices, o.degree, tol=1e-7, max_iter=300)
    np.testing.assert_allclose(s, x, rtol=2e-5)


def test_pr_fp32_update_is_not_fma_contracted():
    """base + d*sum must be two roundings (the reference's x86-64 build has no
    FMA).  Replaying one iteration in numpy float32 reproduces the oracle bit
    for bit; a fused multiply-add differs for some vertex."""
    o, i = graph.rmat(10, 16, seed=7)
    m = o.m
    s0 = np.full(m, np.float32(1) / np.float32(m), np.float32)
    s1, _, _ = oracle.pr(m, i.row_offsets, i.column_indices, o.degree, s0, tol=0.0, max_iter=1)
    contrib = (s0 / o.degree.astype(np.float32)).astype(np.float32)
    base = np.float32((np.float32(1) - np.float32(0.85)) / np.float32(m))
    d = np.float32(0.85)
    sums = np.zeros(m, np.float32)
    for v in range(m):
        acc = np.float32(0)
        for e in range(i.row_offsets[v], i.row_offsets[v + 1]):
            acc = np.float32(acc + contrib[i.column_indices[e]])
        sums[v] = acc
    two_round = (base + (d * sums).astype(np.float32)).astype(np.float32)
    np.testing.assert_array_equal(s1, two_round)
    fused = (np.float64(base) + np.float64(d) * sums.astype(np.float64)).astype(np.float32)
    assert np.any(fused != two_round)


def test_pr_residual_zero_at_fp64_fixed_point():
    o, i = graph.read_pr_graph(dataset("chesapeake.mtx"), is_directed=False)
    x, _ = oracle.pr_fp64(o.m, i.row_offsets, i.column_indices, o.degree, tol=1e-15)
    r = oracle.pr_residual(o.m, o.row_offsets, o.column_indices, o.degree, x.astype(np.float32))
    assert r < 1e-6
    r_bad = oracle.pr_residual(o.m, o.row_offsets, o.column_indices, o.degree,
                               np.full(o.m, 1.0 / o.m, np.float32))
    assert r_bad > 1e-3


def test_spmv_nvgraph_example_closed_form():
    # Inputs of src/nvGRAPH/spmv.cu:29-35 (5x5 CSR, x = 1.1..5.5, y0 = 0).
    Ap = np.array([0, 2, 4, 7, 9, 10], np.int32)
    Aj = np.array([0, 1, 1, 2, 0, 3, 4, 2, 4, 2], np.int32)
    Ax = np.array([1.0, 4.0, 2.0, 3.0, 5.0, 7.0, 8.0, 9.0, 6.0, 1.5], np.float32)
    x = np.array([1.1, 2.2, 3.3, 4.4, 5.5], np.float32)
    y = oracle.spmv(5, Ap, Aj, Ax, x, np.zeros(5, np.float32))
    np.testing.assert_allclose(y, [9.9, 14.3, 80.3, 62.7, 4.95], rtol=1e-6)


def test_spmv_order_and_rounding():
    """Sequential FP32 order from y[i]: a row crafted so that any other
    association gives a different float."""
    Ap = np.array([0, 3], np.int32)
    Aj = np.array([0, 1, 2], np.int32)
    Ax = np.array([1.0, 1.0, 1.0], np.float32)
    x = np.array([1.0, 1e-8, 1e-8], np.float32)
    y = oracle.spmv(1, Ap, Aj, Ax, x, np.array([0.0], np.float32))
    assert y[0] == np.float32(1.0)  # (0+1)+1e-8+1e-8 rounds back to 1 twice
    Aj2 = np.array([1, 2, 0], np.int32)
    y2 = oracle.spmv(1, Ap, Aj2, Ax, x, np.array([0.0], np.float32))
    assert y2[0] == np.float32(np.float32(2e-8) + np.float32(1.0))


def test_max_rel_err_formula():
    eps = np.sqrt(np.finfo(np.float32).eps)
    a = np.array([1.0, 2.0, 0.0, -3.0], np.float32)
    b = np.array([1.0, 2.5, 1e-4, -3.0], np.float32)
    got = oracle.max_rel_err(a, b)
    expect = max(np.float32(0.5) / np.float32(2.0 + 2.5 + eps),
                 np.float32(1e-4) / np.float32(1e-4 + eps))
    assert got == pytest.approx(float(expect), rel=1e-6)
    assert oracle.max_rel_err(a, a) == 0.0


def test_ppr_fp64_dense_and_linearity():
    o, i = graph.rmat(8, 8, seed=3)
    m = o.m
    seeds = np.arange(m, dtype=np.int32)
    X, _ = oracle.ppr_fp64(m, i.row_offsets, i.column_indices, o.degree, seeds, tol=1e-15)
    M = dense_pull(m, i.row_offsets, i.column_indices, np.maximum(o.degree, 1))
    Xd = np.linalg.solve(np.eye(m) - 0.85 * M, 0.15 * np.eye(m))
    np.testing.assert_allclose(X, Xd, atol=1e-12)
    # uniform teleport PR = average of all personalised vectors
    x, _ = oracle.pr_fp64(m, i.row_offsets, i.column_indices, o.degree, tol=1e-15)
    np.testing.assert_allclose(X.mean(axis=1), x, atol=1e-12)


def test_ppr_fp32_matches_fp64():
    o, i = graph.rmat(10, 16, seed=2)
    seed = int(np.argmax(o.degree))
    s, it, _ = oracle.ppr(o.m, i.row_offsets, i.column_indices, o.degree, seed, tol=1e-8,
                          max_iter=300)
    X, _ = oracle.ppr_fp64(o.m, i.row_offsets, i.column_indices, o.degree,
                           np.array([seed], np.int32))
    assert np.abs(s - X[:, 0]).sum() < 1e-5


def _weighted_graph(scale=7, seed=5):
    o, i = graph.rmat(scale, 8, seed=seed)
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.5, i.nnz)
    return i, a


def test_wpr_fp64_dense():
    i, a = _weighted_graph()
    m = i.m
    s = np.zeros(m)
    np.add.at(s, i.column_indices, a)
    M = np.zeros((m, m))
    for v in range(m):
        for e in range(i.row_offsets[v], i.row_offsets[v + 1]):
            M[v, i.column_indices[e]] += a[e] / s[i.column_indices[e]]
    xd = np.linalg.solve(np.eye(m) - 0.85 * M, np.full(m, 0.15 / m))
    x, _ = oracle.wpr_fp64(m, i.row_offsets, i.column_indices, a, tol=1e-16)
    np.testing.assert_allclose(x, xd, rtol=1e-10)


def test_wpr_gradient_central_differences():
    i, a = _weighted_graph()
    m = i.m
    w = np.random.default_rng(0).standard_normal(m)
    x, _ = oracle.wpr_fp64(m, i.row_offsets, i.column_indices, a, tol=1e-16)
    lam, grad, _ = oracle.wpr_grad_fp64(m, i.row_offsets, i.column_indices, a, x, w, tol=1e-16)
    rng = np.random.default_rng(1)
    for e in rng.choice(i.nnz, 24, replace=False):
        h = 1e-6
        ap, am = a.copy(), a.copy()
        ap[e] += h
        am[e] -= h
        xp, _ = oracle.wpr_fp64(m, i.row_offsets, i.column_indices, ap, tol=1e-16)
        xm, _ = oracle.wpr_fp64(m, i.row_offsets, i.column_indices, am, tol=1e-16)
        fd = (w @ xp - w @ xm) / (2 * h)
        assert grad[e] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_spmv_bwd_fp64_dense():
    rng = np.random.default_rng(3)
    rows, cols = 40, 30
    A = (rng.random((rows, cols)) < 0.2) * rng.uniform(-1, 1, (rows, cols))
    Ap = np.concatenate([[0], np.cumsum((A != 0).sum(1))]).astype(np.int32)
    Aj = np.nonzero(A)[1].astype(np.int32)
    Ax = A[A != 0].astype(np.float32)
    x = rng.uniform(-1, 1, cols).astype(np.float32)
    g = rng.uniform(-1, 1, rows).astype(np.float32)
    dx, dA = oracle.spmv_bwd_fp64(rows, cols, Ap, Aj, Ax, x, g)
    Af = np.zeros_like(A)
    Af[A != 0] = Ax
    np.testing.assert_allclose(dx, Af.T @ g.astype(np.float64), rtol=1e-12, atol=1e-12)
    r, c = np.nonzero(A)
    np.testing.assert_allclose(dA, g[r].astype(np.float64) * x[c], rtol=1e-15)


def test_pr_push_hand_computed_iteration():
    """One push iteration on 0->1, 0->2, 1->2, 2->0, 3->0 from 1/4 each, in FP32
    and the reference's order (next = base, then each source's kDamp*score/deg
    added to its out-neighbours in out-CSR order)."""
    f = np.float32
    off = np.array([0, 2, 3, 4, 5], np.int32)
    col = np.array([1, 2, 2, 0, 0], np.int32)
    d, q = f(0.85), f(0.25)
    base = (f(1.0) - d) / f(4)
    v0, v1, v2, v3 = d * q / f(2), d * q / f(1), d * q / f(1), d * q / f(1)
    expect = np.array([(base + v2) + v3, base + v0, (base + v0) + v1, base], np.float32)
    s, it, errs = oracle.pr_push(4, off, col, np.full(4, q, np.float32), tol=0.0, max_iter=1)
    assert it == 1
    np.testing.assert_array_equal(s, expect)
    assert errs[0] == float(np.abs(expect.astype(np.float64) - 0.25).sum())


@pytest.mark.parametrize("name,directed", [("test_pr.mtx", True), ("chesapeake.mtx", False)])
def test_pr_push_matches_pull_oracle_and_fixed_point(name, directed):
    """Push and pull are the same iteration up to the FP32 summation order: the
    scores agree to order noise, the stop iteration is the same (the push
    loop continues while the change is > tol, the pull loop stops once it is
    < tol), and both approach the FP64 fixed point.  A push over the in-CSR,
    or with the in-degree, fails the fixed point."""
    o, i = graph.read_pr_graph(dataset(name), is_directed=directed)
    for tol in (1e-3, 1e-6):
        sp, itp, ep = oracle.pr_push(o.m, o.row_offsets, o.column_indices, tol=tol, max_iter=200)
        sq, itq, eq = oracle.pr(o.m, i.row_offsets, i.column_indices, o.degree, tol=tol,
                                max_iter=200)
        assert itp == itq
        np.testing.assert_allclose(sp, sq, rtol=2e-6, atol=1e-9)
    x, _ = oracle.pr_fp64(o.m, i.row_offsets, i.column_indices, o.degree, tol=1e-15)
    sp, _, _ = oracle.pr_push(o.m, o.row_offsets, o.column_indices, tol=0.0, max_iter=300)
    assert np.abs(sp - x).sum() < 1e-5
    if o.m > 4:  # transposed operator: not the fixed point
        wrong, _, _ = oracle.pr_push(o.m, i.row_offsets, i.column_indices, tol=0.0, max_iter=300)
        assert directed is False or np.abs(wrong - x).sum() > 1e-3


def test_pr_push_rmat_matches_pull_oracle():
    o, i = graph.rmat(12, 16, seed=5)
    sp, itp, ep = oracle.pr_push(o.m, o.row_offsets, o.column_indices, tol=1e-6, max_iter=200)
    sq, itq, eq = oracle.pr(o.m, i.row_offsets, i.column_indices, o.degree, tol=1e-6,
                            max_iter=200)
    assert itp == itq
    assert np.max(np.abs(sp - sq) / sq) < 1e-5
    # per-iteration L1: FP32 order noise of ~1e-10 per vertex, summed over 4096
    np.testing.assert_allclose(ep, eq, rtol=1e-5, atol=1e-6)
