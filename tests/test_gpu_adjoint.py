"""GPU adjoints: weighted PageRank forward/backward and SpMV backward against
FP64 references (oracle), central finite differences and the plain kernels."""
import numpy as np
import pytest

import oracle
from gardenia_ocl_amd import PageRankPlan, SpmvPlan, graph

pytestmark = pytest.mark.gpu


def _weighted(scale, seed, device):
    import torch

    out_csr, in_csr = graph.rmat(scale, 16, seed=seed)
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.5, in_csr.nnz).astype(np.float32)
    t = lambda v: torch.from_numpy(v).to(device)  # noqa: E731
    return out_csr, in_csr, a, t(in_csr.row_offsets), t(in_csr.column_indices), t(a)


def test_wpr_unit_weights_equal_plain_pagerank(device):
    """a == 1: s_u = out-degree and 1.0f * c == c, so the weighted forward is
    the reference iteration, bit for bit, with the same tiles."""
    import torch

    from gardenia_ocl_amd.autograd import WeightedPageRankPlan

    out_csr, in_csr = graph.rmat(14, 16, seed=2)
    m = in_csr.m
    off = torch.from_numpy(in_csr.row_offsets).to(device)
    col = torch.from_numpy(in_csr.column_indices).to(device)
    ones = torch.ones(in_csr.nnz, dtype=torch.float32, device=device)
    wp = WeightedPageRankPlan(off, col, ones)
    x = torch.full((m,), 1.0 / m, dtype=torch.float32, device=device)
    st = wp.forward(x, tol=1e-4, max_iter=100)
    deg = torch.from_numpy(out_csr.degree).to(device)
    pr = PageRankPlan(off, col, deg, relabel=False, items_per_thread=8)
    s = torch.full((m,), 1.0 / m, dtype=torch.float32, device=device)
    st2, _ = pr.run(s, tol=1e-4, max_iter=100)
    np.testing.assert_array_equal(x.cpu().numpy(), s.cpu().numpy())
    assert st["iterations"] == st2["iterations"]


def test_wpr_forward_matches_fp64(device):
    import torch

    from gardenia_ocl_amd.autograd import WeightedPageRankPlan

    out_csr, in_csr, a, off, col, ta = _weighted(14, 5, device)
    m = in_csr.m
    wp = WeightedPageRankPlan(off, col, ta)
    x = torch.full((m,), 1.0 / m, dtype=torch.float32, device=device)
    st = wp.forward(x, tol=1e-7, max_iter=300)
    assert st["converged"]
    x64, _ = oracle.wpr_fp64(m, in_csr.row_offsets, in_csr.column_indices, a.astype(np.float64))
    xg = x.cpu().numpy().astype(np.float64)
    assert np.abs(xg - x64).sum() < 5e-6
    assert np.max(np.abs(xg - x64) / x64) < 1e-4


def test_wpr_gradient_matches_fp64_reference_and_fd(device):
    """dL/da for L = sum_v w_v x_v: FP64 reference gradient over all edges and
    FP64 central differences (eps = 1e-3) on 32 random edges."""
    import torch

    from gardenia_ocl_amd.autograd import WeightedPageRankPlan

    out_csr, in_csr, a, off, col, ta = _weighted(12, 7, device)
    m, nnz = in_csr.m, in_csr.nnz
    rng = np.random.default_rng(3)
    w = rng.standard_normal(m)
    wp = WeightedPageRankPlan(off, col, ta)
    x = torch.full((m,), 1.0 / m, dtype=torch.float32, device=device)
    wp.forward(x, tol=1e-9, max_iter=500)
    lam = torch.empty_like(x)
    grad = torch.empty(nnz, dtype=torch.float32, device=device)
    tw = torch.from_numpy(w.astype(np.float32)).to(device)
    st = wp.backward(x, tw, lam, grad, tol=1e-9, max_iter=500)
    assert st["converged"]
    g = grad.cpu().numpy().astype(np.float64)
    a64 = a.astype(np.float64)
    x64, _ = oracle.wpr_fp64(m, in_csr.row_offsets, in_csr.column_indices, a64)
    lam64, g64, _ = oracle.wpr_grad_fp64(m, in_csr.row_offsets, in_csr.column_indices, a64, x64,
                                         w.astype(np.float32).astype(np.float64))
    scale = np.abs(g64).max()
    assert np.abs(g - g64).max() <= 2e-4 * scale
    np.testing.assert_allclose(lam.cpu().numpy(), lam64, rtol=2e-4, atol=2e-5 * np.abs(lam64).max())
    eps = 1e-3
    for e in rng.choice(nnz, 32, replace=False):
        ap, am = a64.copy(), a64.copy()
        ap[e] += eps
        am[e] -= eps
        xp, _ = oracle.wpr_fp64(m, in_csr.row_offsets, in_csr.column_indices, ap)
        xm, _ = oracle.wpr_fp64(m, in_csr.row_offsets, in_csr.column_indices, am)
        fd = (w @ xp - w @ xm) / (2 * eps)
        assert abs(g[e] - fd) <= 2e-3 * abs(fd) + 2e-4 * scale, (e, g[e], fd)


def test_wpr_autograd_function(device):
    import torch

    from gardenia_ocl_amd.autograd import WeightedPageRankPlan, weighted_pagerank

    out_csr, in_csr, a, off, col, ta = _weighted(11, 9, device)
    weight = torch.nn.Parameter(ta)
    plan = WeightedPageRankPlan(off, col, weight.data)
    x = weighted_pagerank(weight, plan, tol=1e-9, max_iter=500)
    w = torch.randn(in_csr.m, device=device)
    loss = (w * x).sum()
    loss.backward()
    g = weight.grad.cpu().numpy().astype(np.float64)
    a64 = a.astype(np.float64)
    x64, _ = oracle.wpr_fp64(in_csr.m, in_csr.row_offsets, in_csr.column_indices, a64)
    _, g64, _ = oracle.wpr_grad_fp64(in_csr.m, in_csr.row_offsets, in_csr.column_indices, a64,
                                     x64, w.cpu().numpy().astype(np.float64))
    assert np.abs(g - g64).max() <= 2e-4 * np.abs(g64).max()


@pytest.mark.parametrize("kind", ["er", "rmat"])
def test_spmv_backward_matches_fp64(device, kind):
    import torch

    from gardenia_ocl_amd.autograd import SpmvBackwardPlan, spmv

    if kind == "er":
        csr = graph.er_matrix(14, 32, seed=2)
        vals = csr.weight
    else:
        _, csr = graph.rmat(14, 16, seed=4)
        vals = graph.uniform_pm1(csr.nnz, 5)
    m = csr.m
    x = graph.uniform_pm1(m, 6)
    gv = graph.uniform_pm1(m, 7)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(device)  # noqa: E731
    Ap, Aj = t(csr.row_offsets), t(csr.column_indices)
    values = torch.nn.Parameter(t(vals))
    xs = t(x).requires_grad_(True)
    plan = SpmvPlan(Ap, Aj, values.data)
    bwd = SpmvBackwardPlan(Ap, Aj)
    y = spmv(values, xs, plan, bwd)
    y.backward(t(gv))
    dx64, dA64 = oracle.spmv_bwd_fp64(m, m, csr.row_offsets, csr.column_indices, vals, x, gv)
    dx = xs.grad.cpu().numpy().astype(np.float64)
    dA = values.grad.cpu().numpy().astype(np.float64)
    # dA[k] = g * x is one FP32 product: exact up to its rounding
    np.testing.assert_allclose(dA, dA64, rtol=1e-7, atol=0)
    assert np.abs(dx - dx64).max() <= 1e-5 * max(1.0, np.abs(dx64).max())
    # forward value too
    y_ref = oracle.spmv(m, csr.row_offsets, csr.column_indices, vals, x, np.zeros(m, np.float32))
    np.testing.assert_array_equal(y.detach().cpu().numpy()[csr.degree <= 64],
                                  y_ref[csr.degree <= 64])


def test_wpr_gradient_rmat20_fp64_and_fd(device):
    """Full size: RMAT-20, weights U(0.5,1.5) (row-normalised per source by the
    model), L = w.x with w ~ N(0,1).  The GPU edge-weight gradient against the
    FP64 reference gradient on every edge, and against FP64 central
    differences (eps = 1e-3) on 32 random edges (each perturbed solve starts
    from the unperturbed FP64 fixed point)."""
    import torch

    from gardenia_ocl_amd.autograd import WeightedPageRankPlan

    out_csr, in_csr, a, off, col, ta = _weighted(20, 11, device)
    m, nnz = in_csr.m, in_csr.nnz
    rng = np.random.default_rng(13)
    w = rng.standard_normal(m)
    wp = WeightedPageRankPlan(off, col, ta)
    x = torch.full((m,), 1.0 / m, dtype=torch.float32, device=device)
    assert wp.forward(x, tol=1e-9, max_iter=500)["converged"]
    lam = torch.empty_like(x)
    grad = torch.empty(nnz, dtype=torch.float32, device=device)
    tw = torch.from_numpy(w.astype(np.float32)).to(device)
    assert wp.backward(x, tw, lam, grad, tol=1e-7, max_iter=500)["converged"]
    g = grad.cpu().numpy().astype(np.float64)
    a64 = a.astype(np.float64)
    w64 = w.astype(np.float32).astype(np.float64)
    x64, _ = oracle.wpr_fp64(m, in_csr.row_offsets, in_csr.column_indices, a64, tol=1e-15)
    lam64, g64, _ = oracle.wpr_grad_fp64(m, in_csr.row_offsets, in_csr.column_indices, a64, x64,
                                         w64)
    scale = np.abs(g64).max()
    assert np.abs(g - g64).max() <= 2e-4 * scale, np.abs(g - g64).max() / scale
    eps = 1e-3
    for e in rng.choice(nnz, 32, replace=False):
        ap, am = a64.copy(), a64.copy()
        ap[e] += eps
        am[e] -= eps
        xp, _ = oracle.wpr_fp64(m, in_csr.row_offsets, in_csr.column_indices, ap, tol=1e-15,
                                x0=x64)
        xm, _ = oracle.wpr_fp64(m, in_csr.row_offsets, in_csr.column_indices, am, tol=1e-15,
                                x0=x64)
        fd = (w64 @ (xp - xm)) / (2 * eps)
        assert abs(g64[e] - fd) <= 1e-4 * abs(fd) + 1e-6 * scale, (e, g64[e], fd)
        assert abs(g[e] - fd) <= 2e-3 * abs(fd) + 2e-4 * scale, (e, g[e], fd)


def test_spmv_backward_er22(device):
    """SpMV backward at the ER 2^22 x 2^22, 32 nnz/row shape: dA = g[row]*x[col]
    is one FP32 product per nonzero (bit-exact), dx = A^T g against FP64."""
    import torch

    from gardenia_ocl_amd.autograd import SpmvBackwardPlan

    csr = graph.er_matrix(22, 32, seed=3)
    m = csr.m
    x = graph.uniform_pm1(m, 4)
    gv = graph.uniform_pm1(m, 5)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(device)  # noqa: E731
    Ap, Aj, Ax = t(csr.row_offsets), t(csr.column_indices), t(csr.weight)
    bwd = SpmvBackwardPlan(Ap, Aj)
    dx = torch.empty(m, dtype=torch.float32, device=device)
    dA = torch.empty(csr.nnz, dtype=torch.float32, device=device)
    bwd.run(Ax, t(x), t(gv), dx, dA)
    rows = np.repeat(np.arange(m), np.diff(csr.row_offsets))
    np.testing.assert_array_equal(dA.cpu().numpy(), gv[rows] * x[csr.column_indices])
    dx64, _ = oracle.spmv_bwd_fp64(m, m, csr.row_offsets, csr.column_indices, csr.weight, x, gv)
    assert np.abs(dx.cpu().numpy() - dx64).max() <= 1e-5 * max(1.0, np.abs(dx64).max())


def _host_csc_dx(rows, cols, off, col, vals, g):
    """dx = A^T g summed per column in increasing row order (a stable
    transpose of the CSR, entries of one row in CSR order): the oracle's
    sequential FP32 loop over that CSC."""
    order = np.argsort(col, kind="stable")
    row_of = np.repeat(np.arange(rows, dtype=np.int32), np.diff(off))
    csc_off = np.zeros(cols + 1, np.int32)
    csc_off[1:] = np.cumsum(np.bincount(col, minlength=cols))
    return oracle.spmv(cols, csc_off, np.ascontiguousarray(row_of[order]),
                       np.ascontiguousarray(vals[order]), g, np.zeros(cols, np.float32))


@pytest.mark.parametrize("kind", ["square_er21", "rect", "dups"])
def test_spmv_backward_dx_bit_exact(device, kind):
    """dx = A^T g from the backward plan equals the sequential sum over the
    stable CSC bit for bit.  square_er21 and rect (2^21 x 2^18, 16 per row)
    have g of 8 MB, so dx runs in the column-sliced mode with the values
    permuted straight into its slice-major copy; dups (repeated entries in
    rows, so CSC rows are not strictly sorted) falls back to the tile mode.
    dA = g[row] * x[col], one FP32 product, in every case."""
    import torch

    from gardenia_ocl_amd.autograd import SpmvBackwardPlan

    rows = 1 << 21
    if kind == "square_er21":
        csr = graph.er_matrix(21, 32, seed=8)
        cols, off, col, vals = rows, csr.row_offsets, csr.column_indices, csr.weight
    else:
        rng = np.random.default_rng(9)
        cols = (1 << 18) if kind == "rect" else rows
        n = rows * 16
        src = rng.integers(0, rows, n).astype(np.int32)
        dst = rng.integers(0, cols, n).astype(np.int32)
        if kind == "dups":
            dst[1::2] = dst[0::2]  # every other entry repeats its neighbour's column
            src[1::2] = src[0::2]
        m = max(rows, cols)
        c = graph.csr_from_edges(m, src, dst, sorted=True, remove_selfloops=False,
                                 remove_redundents=(kind != "dups"))
        off = np.ascontiguousarray(c.row_offsets[:rows + 1])
        col = np.ascontiguousarray(c.column_indices[:int(off[-1])])
        assert col.max() < cols
        vals = graph.uniform_pm1(col.size, 10)
    x = graph.uniform_pm1(cols, 11)
    gv = graph.uniform_pm1(rows, 12)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(device)  # noqa: E731
    bwd = SpmvBackwardPlan(t(off), t(col), num_cols=cols)
    dx = torch.empty(cols, dtype=torch.float32, device=device)
    dA = torch.empty(col.size, dtype=torch.float32, device=device)
    for _ in range(2):  # the second run overwrites the first's results
        bwd.run(t(vals), t(x), t(gv), dx, dA)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dx.cpu().numpy(), _host_csc_dx(rows, cols, off, col, vals, gv))
    row_of = np.repeat(np.arange(rows), np.diff(off))
    np.testing.assert_array_equal(dA.cpu().numpy(), gv[row_of] * x[col])
