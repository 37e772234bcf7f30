"""Many graphs in one launch per iteration: every graph equals a single-graph
plan (same tiles), and hipGraph replays equal eager launches on every replay."""
import numpy as np
import pytest

import oracle
from gardenia_ocl_amd import MultiGraphPageRank, PageRankPlan, graph

pytestmark = pytest.mark.gpu


def _graphs(device, n, lo, hi, seed):
    import torch

    rng = np.random.default_rng(seed)
    scales = rng.integers(lo, hi + 1, n)
    gs, host = [], []
    for i, sc in enumerate(scales):
        out_csr, in_csr = graph.rmat(int(sc), 16, seed=seed * 1000 + i)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(device)  # noqa: E731
        gs.append((t(in_csr.row_offsets), t(in_csr.column_indices), t(out_csr.degree)))
        host.append((out_csr, in_csr))
    return gs, host


def _init(device, host):
    import torch

    return [torch.full((h[1].m,), 1.0 / h[1].m, dtype=torch.float32, device=device) for h in host]


@pytest.mark.parametrize("use_graph", [False, True])
def test_multigraph_equals_single_graph_plans(device, use_graph):
    gs, host = _graphs(device, 12, 8, 14, 3)
    mg = MultiGraphPageRank(gs)
    scores = _init(device, host)
    its, errs, ms = mg.run(scores, tol=1e-5, max_iter=100, use_graph=use_graph)
    for (off, col, deg), (out_csr, in_csr), s, it in zip(gs, host, scores, its):
        pr = PageRankPlan(off, col, deg, relabel=False, items_per_thread=16)
        import torch

        ref = torch.full((in_csr.m,), 1.0 / in_csr.m, dtype=torch.float32, device=device)
        st, _ = pr.run(ref, tol=1e-5, max_iter=100)
        np.testing.assert_array_equal(s.cpu().numpy(), ref.cpu().numpy())
        assert it == st["iterations"]
    # and against the oracle for one graph
    out_csr, in_csr = host[0]
    r, it, _ = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                         tol=1e-5, max_iter=100)
    np.testing.assert_array_equal(scores[0].cpu().numpy(), r)


def test_multigraph_graph_replays_match_eager(device):
    """Replay the captured graph on several runs, with eager work and syncs in
    between, and compare every replay with an eager run."""
    import torch

    gs, host = _graphs(device, 64, 12, 16, 5)
    mg = MultiGraphPageRank(gs)
    eager = _init(device, host)
    its_e, _, _ = mg.run(eager, tol=1e-6, max_iter=60, use_graph=False, chunk=4)
    ref = [s.cpu().numpy() for s in eager]
    for rep in range(3):
        scores = _init(device, host)
        junk = torch.randn(1 << 20, device=device)
        junk.mul_(2.0).add_(1.0)
        torch.cuda.synchronize()
        its_g, _, _ = mg.run(scores, tol=1e-6, max_iter=60, use_graph=True, chunk=4)
        np.testing.assert_array_equal(its_g, its_e)
        for a, b in zip(scores, ref):
            np.testing.assert_array_equal(a.cpu().numpy(), b)
    assert its_e.min() > 1 and len(set(its_e.tolist())) > 1  # uneven graphs converge unevenly


def test_multigraph_max_iter_cap_and_empty_graph(device):
    import torch

    gs, host = _graphs(device, 3, 6, 9, 7)
    m = 100
    empty = graph.csr_from_edges(m, np.zeros(0, np.int32), np.zeros(0, np.int32))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(device)  # noqa: E731
    gs.append((t(empty.row_offsets), torch.zeros(0, dtype=torch.int32, device=device),
               t(np.zeros(m, np.int32))))
    mg = MultiGraphPageRank(gs)
    scores = _init(device, host) + [torch.full((m,), 1.0 / m, device=device)]
    its, errs, _ = mg.run(scores, tol=0.0, max_iter=7)
    assert list(its) == [7, 7, 7, 7]
    base = (np.float32(1.0) - np.float32(0.85)) / np.float32(m)
    np.testing.assert_array_equal(scores[-1].cpu().numpy(), np.full(m, base, np.float32))


def test_multigraph_64_rmat_12_to_18_every_graph_vs_oracle(device):
    """The multi-graph workload at its measured size: 64 RMAT graphs of scales
    drawn uniformly from 12-18 with the bench's seeds, one plan, tol 1e-4.
    Every graph against the oracle's restatement of src/pr/omp_base.cc: the
    same iteration count (unless the oracle's own L1 lies within L1_GAP of tol
    at an iteration up to its stop), the final L1 within L1_GAP, every score
    bit-identical."""
    import torch

    from conftest import L1_GAP

    seed, n, tol = 1, 64, 1e-4
    scales = np.random.default_rng(seed + 5).integers(12, 19, n)
    assert scales.min() == 12 and scales.max() == 18
    gs, host = [], []
    for i, sc in enumerate(scales):
        out_csr, in_csr = graph.rmat(int(sc), 16, seed=seed + 100 + i)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(device)  # noqa: E731
        gs.append((t(in_csr.row_offsets), t(in_csr.column_indices), t(out_csr.degree)))
        host.append((out_csr, in_csr))
    mg = MultiGraphPageRank(gs)
    scores = _init(device, host)
    its, final, _ = mg.run(scores, tol=tol, max_iter=100)
    worst, ties = 0.0, 0
    for g, ((out_csr, in_csr), s) in enumerate(zip(host, scores)):
        ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                      out_csr.degree, tol=tol, max_iter=100)
        tie = bool(np.any(np.abs(ref_errs[:it] - tol) <= L1_GAP))
        ties += tie
        if not tie:
            assert its[g] == it, (g, int(scales[g]), int(its[g]), it)
            assert abs(final[g] - ref_errs[-1]) <= L1_GAP, (g, final[g], ref_errs[-1])
        rel = np.max(np.abs(s.cpu().numpy() - ref) / ref)
        worst = max(worst, rel)
        np.testing.assert_array_equal(s.cpu().numpy(), ref)
    print(f"64 graphs: max rel {worst:.2e}, oracle ties {ties}, iterations {its.min()}-{its.max()}")
    mg.close()
