"""The autotuner searches the plan knobs by coordinate descent from the library
defaults, caches the winner per graph shape, and the tuned configurations --
used through PageRankPlan, the one-shot PRSolver, the batched-PPR plan and the
many-graph run -- give the oracle's scores and iteration counts bit for bit."""
import json

import numpy as np
import pytest

import oracle
from gardenia_ocl_amd import (MultiGraphPageRank, PageRankPlan, PersonalizedPageRankPlan,
                              PRSolver, autotune, graph)


def _t(device, a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(device)


def test_descent_walks_from_defaults():
    """Host logic of the search: starts at the first values, adopts a value only
    when it is faster by more than the margin, stops when a pass changes nothing."""
    space = {"a": [0, 1, 2], "b": [0, 1]}
    cost = {(0, 0): 10.0, (1, 0): 9.95, (2, 0): 8.0, (2, 1): 6.0, (0, 1): 12.0, (1, 1): 7.0}
    seen = []

    def measure(cfg):
        seen.append((cfg["a"], cfg["b"]))
        return cost[(cfg["a"], cfg["b"])]

    best, ms, timings = autotune.descend(space, measure)
    assert seen[0] == (0, 0)                   # the defaults first
    assert best == {"a": 2, "b": 1} and ms == 6.0
    assert len(seen) == len(set(seen))         # each configuration timed once
    best, _, _ = autotune.descend({"a": [0, 1]}, lambda c: [10.0, 9.95][c["a"]])
    assert best == {"a": 0}                    # 0.5% faster: within the margin


@pytest.mark.gpu
def test_autotune_pagerank_tuned_plan_and_prsolver_match_oracle(device, tmp_path):
    out_csr, in_csr = graph.rmat(15, 16, seed=3)
    off, col, deg = _t(device, in_csr.row_offsets), _t(device, in_csr.column_indices), \
        _t(device, out_csr.degree)
    cache = str(tmp_path / "tune.json")
    best, timings = autotune.tune_pagerank(off, col, deg, iters=4, repeats=2, cache=cache)
    defaults = {n: v[0] for n, v in autotune.PR_SPACE.items()}
    assert json.dumps(defaults, sort_keys=True) in timings
    assert json.dumps(best, sort_keys=True) in timings and all(v > 0 for v in timings.values())
    assert timings[json.dumps(best, sort_keys=True)] <= min(timings.values()) * 1.0101
    # every knob was tried at least once away from its default
    for n, vals in autotune.PR_SPACE.items():
        assert any(json.loads(k)[n] != vals[0] for k in timings), n
    # cached: a second call does not re-time, and the lookup returns the winner
    assert autotune.tune_pagerank(off, col, deg, cache=cache) == (best, timings)
    assert autotune.pagerank_config(cache, off, col, deg) == best
    assert autotune.pagerank_config(str(tmp_path / "none.json"), off, col, deg) is None

    ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                  out_csr.degree, tol=1e-5, max_iter=50)
    import torch

    a = torch.full((in_csr.m,), 1.0 / in_csr.m, device=device)
    st, errs = PageRankPlan(off, col, deg, **best).run(a, tol=1e-5, max_iter=50)
    np.testing.assert_array_equal(a.cpu().numpy(), ref)
    assert st["iterations"] == it
    # the one-shot drop-in with the tuned configuration
    s = np.full(in_csr.m, np.float32(1.0) / np.float32(in_csr.m), np.float32)
    st2, errs2 = PRSolver(in_csr.m, in_csr.nnz, in_csr.row_offsets, in_csr.column_indices,
                          out_csr.row_offsets, out_csr.column_indices, out_csr.degree, s,
                          tol=1e-5, max_iter=50, config=best)
    np.testing.assert_array_equal(s, ref)
    assert st2["iterations"] == it and np.allclose(errs2, ref_errs[:it], rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_autotune_ppr_tuned_plan_matches_oracle(device, tmp_path):
    import torch

    out_csr, in_csr = graph.rmat(12, 16, seed=3)
    off, col, deg = _t(device, in_csr.row_offsets), _t(device, in_csr.column_indices), \
        _t(device, out_csr.degree)
    cache = str(tmp_path / "tune.json")
    k = 128
    best, timings = autotune.tune_ppr(off, col, deg, k, iters=2, repeats=1, cache=cache)
    assert best["k_chunk"] in (128, 64) and best["hub_cols"] in (64, 32, 16)
    assert autotune.ppr_config(cache, off, col, deg, k) == best
    seeds = np.random.default_rng(4).integers(0, in_csr.m, k).astype(np.int32)
    plan = PersonalizedPageRankPlan(off, col, deg, k, **best)
    X = plan.one_hot(seeds)
    plan.run(torch.from_numpy(seeds).to(device), X, tol=1e-6, max_iter=100)
    its = plan.column_iterations()
    Xh = X.cpu().numpy()
    for j in (0, 5, 77, 127):
        ref, it, _ = oracle.ppr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                out_csr.degree, int(seeds[j]), tol=1e-6, max_iter=100)
        assert its[j] == it
        np.testing.assert_array_equal(Xh[:, j], ref)


@pytest.mark.gpu
def test_autotune_multigraph_chunk_matches_oracle(device, tmp_path):
    import torch

    graphs, host = [], []
    for i, sc in enumerate((10, 12, 11, 9)):
        o, c = graph.rmat(sc, 16, seed=20 + i)
        host.append((o, c))
        graphs.append((_t(device, c.row_offsets), _t(device, c.column_indices),
                       _t(device, o.degree)))
    cache = str(tmp_path / "tune.json")
    best, timings = autotune.tune_multigraph(graphs, tol=1e-5, repeats=1, cache=cache)
    assert best["chunk"] in autotune.MULTI_SPACE["chunk"]
    assert len(timings) == len(autotune.MULTI_SPACE["chunk"])
    assert autotune.multigraph_config(cache, graphs) == best
    scores = [torch.full((int(g[2].shape[0]),), 1.0 / int(g[2].shape[0]), device=device)
              for g in graphs]
    its, _, _ = MultiGraphPageRank(graphs).run(scores, tol=1e-5, max_iter=100, **best)
    for (o, c), s, n in zip(host, scores, its):
        ref, it, _ = oracle.pr(c.m, c.row_offsets, c.column_indices, o.degree, tol=1e-5,
                               max_iter=100)
        assert n == it
        np.testing.assert_array_equal(s.cpu().numpy(), ref)
