"""Push PageRank (the reference's scatter variant, src/pr/scatter.cu:11-41) on
the GPU, through the C ABI, against the oracle's restatement of that loop
(additions in out-CSR order): the same iteration count and the scores to FP32
summation-order noise (the GPU's atomic adds take another order), and the
reference verifier's verdict on the result."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import dataset
from gardenia_ocl_amd import GardeniaError, graph, pagerank_push


def _t(device, a, dt=np.int32):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dt)).to(device)


def _run(device, o, tol, max_iter):
    scores = _t(device, np.full(o.m, np.float32(1.0) / np.float32(o.m)), np.float32)
    st, errs = pagerank_push(_t(device, o.row_offsets), _t(device, o.column_indices), scores,
                             tol=tol, max_iter=max_iter)
    return scores.cpu().numpy(), st, errs


def _same_iterations(it, ref_it, ref_errs, tol):
    # the stop decision can only differ where the oracle's change sits at tol
    near = any(abs(e - tol) <= 1e-4 * tol for e in ref_errs[-2:])
    assert it == ref_it or (near and abs(it - ref_it) == 1), (it, ref_it, ref_errs[-2:])


@pytest.mark.gpu
@pytest.mark.parametrize("name,directed", [("test_pr.mtx", True), ("chesapeake.mtx", False),
                                           ("4.mtx", True)])
def test_push_datasets_vs_oracle(device, name, directed):
    o, _ = graph.read_pr_graph(dataset(name), is_directed=directed)
    for tol in (1e-3, 1e-6):
        s, st, errs = _run(device, o, tol, 100)
        ref, it, ref_errs = oracle.pr_push(o.m, o.row_offsets, o.column_indices, tol=tol,
                                           max_iter=100)
        _same_iterations(st["iterations"], it, ref_errs, tol)
        np.testing.assert_allclose(s, ref, rtol=2e-6, atol=1e-9)
        # per-iteration L1: sums of |change| whose terms differ in their last bits
        n = min(len(errs), len(ref_errs))
        np.testing.assert_allclose(errs[:n], ref_errs[:n], rtol=1e-5, atol=1e-7)
        assert st["converged"] == (not errs[-1] > tol)


@pytest.mark.gpu
@pytest.mark.parametrize("scale", [12, 16])
def test_push_rmat_vs_oracle_and_verifier(device, scale):
    """RMAT graphs with rows of thousands of in-edges, i.e. thousands of atomic
    adds into one vertex per iteration: scores within 1e-4 relative of the
    oracle's sequential order (typically ~1e-6), and "Correct" from the
    reference verifier's residual at tol 1e-4."""
    o, i = graph.rmat(scale, 16, seed=9)
    s, st, errs = _run(device, o, 1e-4, 100)
    ref, it, ref_errs = oracle.pr_push(o.m, o.row_offsets, o.column_indices, tol=1e-4,
                                       max_iter=100)
    _same_iterations(st["iterations"], it, ref_errs, 1e-4)
    assert np.max(np.abs(s - ref) / ref) < 1e-4
    assert oracle.pr_residual(o.m, o.row_offsets, o.column_indices, o.degree, s) < 1e-3
    # the push fixed point is the pull path's: same scores to order noise
    pull, _, _ = oracle.pr(o.m, i.row_offsets, i.column_indices, o.degree, tol=1e-4,
                           max_iter=100)
    assert np.max(np.abs(s - pull) / pull) < 1e-4


@pytest.mark.gpu
def test_push_one_iteration_and_bad_csr(device):
    o, _ = graph.rmat(10, 8, seed=2)
    s, st, errs = _run(device, o, 0.0, 1)
    ref, it, _ = oracle.pr_push(o.m, o.row_offsets, o.column_indices, tol=0.0, max_iter=1)
    assert st["iterations"] == it == 1
    np.testing.assert_allclose(s, ref, rtol=2e-6)
    bad = np.array(o.row_offsets, np.int32)
    bad[3], bad[4] = bad[4], bad[3] - 1  # not a prefix sum
    scores = _t(device, np.full(o.m, 1.0 / o.m), np.float32)
    with pytest.raises(GardeniaError):
        pagerank_push(_t(device, bad), _t(device, o.column_indices), scores, tol=0.0, max_iter=1)


def test_push_rejects_seed_vertex_before_device():
    from gardenia_ocl_amd import _native as N
    from gardenia_ocl_amd.pr import make_options

    opts = make_options(seed_vertex=3)
    st = N.PrStats()
    rc = N.lib().gd_pr_push_solve(0, 0, None, None, None, ctypes.byref(opts), None,
                                  ctypes.byref(st), None)
    assert rc == N.GD_ERR_INVALID
    assert b"seed_vertex" in N.lib().gd_last_error()


@pytest.mark.gpu
def test_push_window_edge_cases(device):
    """The additions are summed per destination window of 32,768 vertices: an
    empty window between full ones, a partial last window, a window whose
    edges (more than 2^18) are split over several workgroups, and a graph
    without edges; scores and per-iteration L1 against the oracle's push."""
    rng = np.random.default_rng(5)
    W = 1 << 15
    m = 3 * W + 777
    n = 600_000
    src = rng.integers(0, m, n)
    # no destination in window 1; about 400 K edges into window 0
    dst = np.where(rng.random(n) < 0.67, rng.integers(0, W, n), rng.integers(2 * W, m, n))
    o = graph.csr_from_edges(m, src, dst)
    assert np.sum(o.column_indices < W) > (1 << 18)
    assert not np.any((o.column_indices >= W) & (o.column_indices < 2 * W))
    s, st, errs = _run(device, o, 0.0, 3)
    ref, it, ref_errs = oracle.pr_push(o.m, o.row_offsets, o.column_indices, tol=0.0, max_iter=3)
    assert st["iterations"] == it == 3
    np.testing.assert_allclose(s, ref, rtol=2e-6, atol=1e-12)
    np.testing.assert_allclose(errs, ref_errs, rtol=1e-5, atol=1e-9)
    # window 1 receives nothing: base only
    base = (np.float32(1.0) - np.float32(0.85)) / np.float32(m)  # scatter.cu:57
    assert np.all(s[W:2 * W] == base) and np.all(ref[W:2 * W] == base)

    empty = graph.Csr(m=5, row_offsets=np.zeros(6, np.int32),
                      column_indices=np.zeros(0, np.int32), degree=np.zeros(5, np.int32))
    s, st, errs = _run(device, empty, 0.0, 2)
    ref, it, _ = oracle.pr_push(5, empty.row_offsets, empty.column_indices, tol=0.0, max_iter=2)
    assert st["iterations"] == it == 2
    np.testing.assert_array_equal(s, ref)
