"""GPU batched personalised PageRank against the oracle's scalar FP32
personalised loop (same per-column arithmetic) and the FP64 solution."""
import numpy as np
import pytest

import oracle
from gardenia_ocl_amd import PageRankPlan, PersonalizedPageRankPlan, graph

pytestmark = pytest.mark.gpu


def _plan(device, in_csr, outdeg, k, **kw):
    import torch

    off = torch.from_numpy(in_csr.row_offsets).to(device)
    col = torch.from_numpy(in_csr.column_indices).to(device)
    deg = torch.from_numpy(np.ascontiguousarray(outdeg, np.int32)).to(device)
    return PersonalizedPageRankPlan(off, col, deg, k, **kw), (off, col, deg)


def _seeds(m, k, rng_seed):
    return np.random.default_rng(rng_seed).integers(0, m, k).astype(np.int32)


@pytest.mark.parametrize("k,kc,relabel", [(64, 64, False), (128, 64, True), (256, 256, False),
                                          (256, 128, True)])
def test_ppr_columns_bit_exact_vs_scalar_oracle(device, k, kc, relabel):
    """max in-degree <= 4096: every column is the scalar loop's arithmetic."""
    import torch

    out_csr, in_csr = graph.rmat(12, 16, seed=3)
    assert in_csr.degree.max() <= 4096
    m = in_csr.m
    seeds = _seeds(m, k, 1)
    plan, keep = _plan(device, in_csr, out_csr.degree, k, relabel=relabel, k_chunk=kc)
    X = plan.one_hot(seeds)
    st, errs, col_err = plan.run(torch.from_numpy(seeds).to(device), X, tol=0.0, max_iter=12)
    assert st["iterations"] == 12
    Xh = X.cpu().numpy()
    for j in [0, 1, k // 2, k - 1]:
        ref, it, _ = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                                int(seeds[j]), tol=0.0, max_iter=12)
        np.testing.assert_array_equal(Xh[:, j], ref)
    assert plan.info()["k_chunk"] == kc


def test_ppr_hub_rows_and_fp64_accuracy(device):
    import torch

    out_csr, in_csr = graph.rmat(16, 16, seed=1)
    assert in_csr.degree.max() > 4096
    m, k = in_csr.m, 64
    seeds = _seeds(m, k, 2)
    plan, keep = _plan(device, in_csr, out_csr.degree, k)
    assert plan.info()["n_hubs"] > 0
    X = plan.one_hot(seeds)
    st, errs, col_err = plan.run(torch.from_numpy(seeds).to(device), X, tol=1e-6, max_iter=200)
    assert st["converged"] and col_err.max() < 1e-6
    Xh = X.cpu().numpy().astype(np.float64)
    X64, _ = oracle.ppr_fp64(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                             seeds[:8])
    for j in range(8):
        ref32, _, _ = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                                 int(seeds[j]), tol=1e-6, max_iter=200)
        l1_gpu = np.abs(Xh[:, j] - X64[:, j]).sum()
        l1_cpu = np.abs(ref32.astype(np.float64) - X64[:, j]).sum()
        assert l1_gpu <= 2 * l1_cpu + 2e-6


def test_ppr_matches_looped_plain_kernel(device):
    """Batched column j == the plain pull kernel personalised to seeds[j]."""
    import torch

    out_csr, in_csr = graph.rmat(13, 16, seed=4)
    m, k = in_csr.m, 64
    seeds = _seeds(m, k, 3)
    plan, keep = _plan(device, in_csr, out_csr.degree, k, relabel=False)
    X = plan.one_hot(seeds)
    plan.run(torch.from_numpy(seeds).to(device), X, tol=0.0, max_iter=10)
    Xh = X.cpu().numpy()
    off, col, deg = keep
    pr = PageRankPlan(off, col, deg, relabel=False)
    short = in_csr.degree <= 64
    for j in range(0, k, 9):
        s = torch.zeros(m, dtype=torch.float32, device=device)
        s[int(seeds[j])] = 1.0
        pr.run(s, tol=0.0, max_iter=10, seed_vertex=int(seeds[j]))
        sh = s.cpu().numpy()
        np.testing.assert_allclose(Xh[:, j], sh, rtol=1e-5, atol=1e-9)
        if in_csr.degree.max() <= 64:
            np.testing.assert_array_equal(Xh[short, j], sh[short])


def test_ppr_rejects_bad_arguments(device):
    import torch

    from gardenia_ocl_amd import GardeniaError

    out_csr, in_csr = graph.rmat(10, 8, seed=1)
    with pytest.raises(GardeniaError, match="multiple of 64"):
        _plan(device, in_csr, out_csr.degree, 100)
    plan, keep = _plan(device, in_csr, out_csr.degree, 64)
    bad = torch.full((64,), in_csr.m, dtype=torch.int32, device=device)
    X = torch.zeros((in_csr.m, 64), dtype=torch.float32, device=device)
    with pytest.raises(GardeniaError, match="seed vertex out of range"):
        plan.run(bad, X)


def _mixed_seeds(in_csr, out_deg, k, rng_seed):
    """Seeds whose columns stop at very different iterations: vertices with no
    out-edges (their mass stays put: 2 iterations) mixed with random ones."""
    rng = np.random.default_rng(rng_seed)
    sinks = np.flatnonzero(out_deg == 0)
    rand = rng.integers(0, in_csr.m, k)
    pick = rng.random(k) < 0.4
    rand[pick] = rng.choice(sinks, int(pick.sum()))
    return rand.astype(np.int32)


@pytest.mark.parametrize("k,kc,relabel,compact", [(128, 64, False, True), (128, 64, False, False),
                                                  (256, 256, True, True), (256, 128, False, True),
                                                  (192, 64, True, True)])
def test_ppr_column_stop_matches_single_seed_runs(device, k, kc, relabel, compact):
    """Each column stops after its own first iteration with L1 change < tol,
    the same iteration the scalar single-seed loop stops at, with the same
    scores bit for bit (max in-degree <= 4096: sequential per-column sums).
    Moving stopped columns behind the running ones (compact) changes nothing."""
    import torch

    out_csr, in_csr = graph.rmat(12, 16, seed=3)
    m = in_csr.m
    seeds = _mixed_seeds(in_csr, out_csr.degree, k, 5)
    plan, keep = _plan(device, in_csr, out_csr.degree, k, relabel=relabel, k_chunk=kc,
                       compact=compact)
    X = plan.one_hot(seeds)
    st, errs, col_err = plan.run(torch.from_numpy(seeds).to(device), X, tol=1e-6, max_iter=100,
                                 chunk=2)
    its = plan.column_iterations()
    assert st["converged"] and st["iterations"] == its.max()
    assert its.min() <= 3 < 20 < its.max()  # columns really stop at different times
    Xh = X.cpu().numpy()
    for j in range(k):
        ref, it, ref_errs = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices,
                                       out_csr.degree, int(seeds[j]), tol=1e-6, max_iter=100)
        assert its[j] == it, (j, its[j], it)
        np.testing.assert_array_equal(Xh[:, j], ref)
        assert abs(col_err[j] - ref_errs[it - 1]) <= 1e-12 + 1e-9 * ref_errs[it - 1]


def test_ppr_all_columns_mode_runs_to_largest_change(device):
    """column_stop=False: every column runs until the largest change < tol."""
    import torch

    out_csr, in_csr = graph.rmat(12, 16, seed=3)
    m, k = in_csr.m, 64
    seeds = _mixed_seeds(in_csr, out_csr.degree, k, 6)
    plan, keep = _plan(device, in_csr, out_csr.degree, k, column_stop=False)
    X = plan.one_hot(seeds)
    st, errs, col_err = plan.run(torch.from_numpy(seeds).to(device), X, tol=1e-6, max_iter=100)
    its = plan.column_iterations()
    assert np.all(its == st["iterations"]) and errs[-1] < 1e-6 <= errs[-2]
    Xh = X.cpu().numpy()
    for j in (0, 7, 63):
        ref, it, _ = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                                int(seeds[j]), tol=0.0, max_iter=st["iterations"])
        np.testing.assert_array_equal(Xh[:, j], ref)
