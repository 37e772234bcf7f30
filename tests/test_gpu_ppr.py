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
    """Every column is the scalar loop's arithmetic (rows of <= 4096 in-edges:
    one wavefront per row, each lane adding its columns in CSR order)."""
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


@pytest.mark.parametrize("k", [64, 192])
def test_ppr_hub_rows_bit_exact_and_fp64_accuracy(device, k):
    """Rows of more than 4096 in-edges (hubs) are added in edge order per
    column by their own workgroups: every column is the scalar loop's, bit for
    bit, with its iteration count; L1 to the FP64 solution as the reference's."""
    import torch

    out_csr, in_csr = graph.rmat(16, 16, seed=1)
    assert in_csr.degree.max() > 4096
    m = in_csr.m
    seeds = _seeds(m, k, 2)
    hub = int(np.argmax(in_csr.degree))
    seeds[1] = in_csr.column_indices[in_csr.row_offsets[hub]]  # gathers the hub at once
    plan, keep = _plan(device, in_csr, out_csr.degree, k)
    assert plan.info()["n_hubs"] > 0
    X = plan.one_hot(seeds)
    st, errs, col_err = plan.run(torch.from_numpy(seeds).to(device), X, tol=1e-6, max_iter=200)
    assert st["converged"] and col_err.max() < 1e-6
    its = plan.column_iterations()
    Xh = X.cpu().numpy()
    X64, _ = oracle.ppr_fp64(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                             seeds[:8])
    for j in range(8):
        ref32, it, _ = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                                  int(seeds[j]), tol=1e-6, max_iter=200)
        np.testing.assert_array_equal(Xh[:, j], ref32)
        assert its[j] == it
        l1_gpu = np.abs(Xh[:, j].astype(np.float64) - X64[:, j]).sum()
        l1_cpu = np.abs(ref32.astype(np.float64) - X64[:, j]).sum()
        assert l1_gpu == l1_cpu


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
    for j in range(0, k, 9):
        s = torch.zeros(m, dtype=torch.float32, device=device)
        s[int(seeds[j])] = 1.0
        pr.run(s, tol=0.0, max_iter=10, seed_vertex=int(seeds[j]))
        # both sum every row of every column in CSR order: the same bits
        np.testing.assert_array_equal(Xh[:, j], s.cpu().numpy())


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
    scores bit for bit (every column summed sequentially in CSR order).
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


def test_ppr_rmat22_k256_config_size_vs_oracle(device):
    """The batched-PPR workload at its measured size: RMAT-22, 256 one-hot
    seeds drawn with the bench's generator, per-column L1 tol 1e-6.  Eight
    columns against the oracle's scalar personalised loop (src/pr/omp_base.cc
    arithmetic, teleport to the seed): four of the bench's seeds and four
    in-neighbours of the longest row, whose columns gather that row (97K
    in-edges, far past 4096) from the first iteration.  Every column is the
    scalar loop's bit for bit, with the same per-column iteration count and
    last L1 value."""
    import torch

    from conftest import L1_GAP

    seed, k, tol = 1, 256, 1e-6
    out_csr, in_csr = graph.rmat(22, 16, seed=seed)
    m = in_csr.m
    seeds = np.random.default_rng(seed + 17).integers(0, m, k).astype(np.int32)
    hub = int(np.argmax(in_csr.degree))
    assert in_csr.degree[hub] > 4096
    srcs = in_csr.column_indices[in_csr.row_offsets[hub]:in_csr.row_offsets[hub + 1]]
    pick = [0, 1, 2, 3]
    for q, j in enumerate((100, 101, 102, 103)):
        seeds[j] = srcs[(q * 7919) % len(srcs)]
        pick.append(j)
    plan, keep = _plan(device, in_csr, out_csr.degree, k)
    assert plan.info()["n_hubs"] > 0
    X = plan.one_hot(seeds)
    st, errs, col_err = plan.run(torch.from_numpy(seeds).to(device), X, tol=tol, max_iter=300)
    its = plan.column_iterations()
    Xh = X.cpu().numpy()
    del X
    for j in pick:
        ref, it, ref_errs = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices,
                                       out_csr.degree, int(seeds[j]), tol=tol, max_iter=300)
        print(f"column {j} seed {seeds[j]}: iterations {its[j]} (oracle {it}), last L1 "
              f"{col_err[j]:.6e} (oracle {ref_errs[it - 1]:.6e}), hub score {Xh[hub, j]:.6e}")
        np.testing.assert_array_equal(Xh[:, j], ref)
        assert its[j] == it, (j, its[j], it)
        assert abs(col_err[j] - ref_errs[it - 1]) <= L1_GAP
