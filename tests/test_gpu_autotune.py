"""The autotuner times every candidate plan configuration, caches the winner,
and the tuned plan gives the same scores as the default one."""
import json

import numpy as np
import pytest

from gardenia_ocl_amd import PageRankPlan, graph

pytestmark = pytest.mark.gpu


def test_autotune_pagerank_and_ppr(device, tmp_path):
    import torch

    from gardenia_ocl_amd import autotune

    out_csr, in_csr = graph.rmat(15, 16, seed=3)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(device)  # noqa: E731
    off, col, deg = t(in_csr.row_offsets), t(in_csr.column_indices), t(out_csr.degree)
    cache = str(tmp_path / "tune.json")
    best, timings = autotune.tune_pagerank(off, col, deg, iters=3, repeats=2, cache=cache)
    n_cfg = int(np.prod([len(v) for v in autotune.PR_SPACE.values()]))
    assert len(timings) == n_cfg and all(v > 0 for v in timings.values())
    assert json.dumps(best, sort_keys=True) in timings
    # cached: second call does not re-time
    best2, timings2 = autotune.tune_pagerank(off, col, deg, cache=cache)
    assert best2 == best and timings2 == timings
    # the tuned plan reproduces the default plan's scores (relabel, tile size,
    # cache hints, persistence and sorted tiles never change the arithmetic)
    cfg = dict(best)
    a = torch.full((in_csr.m,), 1.0 / in_csr.m, device=device)
    b = torch.full((in_csr.m,), 1.0 / in_csr.m, device=device)
    PageRankPlan(off, col, deg, **cfg).run(a, tol=1e-5, max_iter=50)
    PageRankPlan(off, col, deg).run(b, tol=1e-5, max_iter=50)
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    pbest, ptimings = autotune.tune_ppr(off, col, deg, 128, iters=2, cache=cache)
    assert pbest["k_chunk"] in (64, 128) and len(ptimings) == 4
