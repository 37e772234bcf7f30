"""Device PR verifier (the reference PRVerifier's push residual) against the
oracle's restatement of src/pr/verifier.cc:50-68 and the reference's own
printed verdicts.  Every per-vertex term is computed as the reference does
(sources in increasing order, sequential FP32), so the FP64 residuals agree
to the rounding of their different summation orders."""
import numpy as np
import pytest

import oracle
from conftest import dataset
from gardenia_ocl_amd import PRVerifier, graph, verify_residual

pytestmark = pytest.mark.gpu

DATASETS = ["test_pr.mtx", "4.mtx", "4w.mtx", "chesapeake.mtx"]


def _key(ds):
    return ds.replace(".mtx", "").replace(".", "_")


def _t(a, dt, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dt)).to(device)


def _gpu_residual(device, oc, scores):
    return verify_residual(_t(oc.row_offsets, np.int32, device), _t(oc.column_indices, np.int32, device),
                           _t(oc.degree, np.int32, device), _t(scores, np.float32, device))


@pytest.mark.parametrize("ds", DATASETS)
@pytest.mark.parametrize("mode", ["directed", "undirected"])
def test_verifier_datasets_match_reference_verdicts(golden, device, ds, mode, capsys):
    arrays, meta = golden
    oc, _ = graph.read_pr_graph(dataset(ds), is_directed=(mode == "directed"))  # main.cc:23-24
    s = arrays[f"{_key(ds)}/pr_{mode}/scores"]
    res = _gpu_residual(device, oc, s)
    ref = oracle.pr_residual(oc.m, oc.row_offsets, oc.column_indices, oc.degree, s)
    assert abs(res - ref) <= 1e-12 * max(1.0, ref)
    verdict = meta["datasets"][ds][f"pr_{mode}"]["verdict"]
    ok, res2 = PRVerifier(oc.m, None, None, oc.row_offsets, oc.column_indices, oc.degree, s,
                          meta["epsilon"])
    out = capsys.readouterr().out.strip().splitlines()
    assert out[0] == "Verifying..." and out[-1] == verdict
    assert ok == (verdict == "Correct") and res2 == res


def test_verifier_long_rows_exact(device):
    """A destination with 140,000 sources (more than 2^17: the binade scan) and
    one with 3,000 (the one-lane chain): each push sum is sequential over
    equal-order terms, where a tree sum would differ."""
    m = 140_002
    leaves = np.arange(2, m, dtype=np.int32)
    src = np.concatenate([leaves, leaves[:3000], np.zeros(m - 2, np.int32)])
    dst = np.concatenate([np.zeros(m - 2, np.int32), np.ones(3000, np.int32), leaves])
    oc = graph.csr_from_edges(m, src, dst, sorted=True)
    rng = np.random.default_rng(3)
    s = (rng.random(m).astype(np.float32) / m).astype(np.float32)
    res = _gpu_residual(device, oc, s)
    ref = oracle.pr_residual(oc.m, oc.row_offsets, oc.column_indices, oc.degree, s)
    assert abs(res - ref) <= 1e-12 * max(1.0, ref)


def test_verifier_rmat_after_gpu_solve(device):
    import torch

    from gardenia_ocl_amd import PageRankPlan

    out_csr, in_csr = graph.rmat(16, 16, seed=1)
    plan = PageRankPlan(_t(in_csr.row_offsets, np.int32, device),
                        _t(in_csr.column_indices, np.int32, device),
                        _t(out_csr.degree, np.int32, device))
    s = torch.full((in_csr.m,), float(np.float32(1.0) / np.float32(in_csr.m)), device=device)
    plan.run(s, tol=1e-4, max_iter=100)
    sh = s.cpu().numpy()
    res = _gpu_residual(device, out_csr, sh)
    ref = oracle.pr_residual(out_csr.m, out_csr.row_offsets, out_csr.column_indices,
                             out_csr.degree, sh)
    assert abs(res - ref) <= 1e-12 * max(1.0, ref) and res < 1e-3


@pytest.mark.parametrize("bad", ["decreasing", "last", "first", "column"])
def test_verifier_rejects_invalid_csr(device, bad):
    """Offsets and columns are checked on the device before the transpose
    reads them: a bad CSR is an error with build_schedule's text, not a fault."""
    from gardenia_ocl_amd import GardeniaError

    off = np.array([0, 2, 3, 5, 6], np.int32)
    col = np.array([1, 2, 0, 3, 1, 0], np.int32)
    deg = np.diff(off).astype(np.int32)
    if bad == "decreasing":
        off[2] = 1  # 0 2 1 5 6
        msg = "non-decreasing (row 1)"
    elif bad == "last":
        off[4] = 5
        msg = "row_offsets[m] must equal nnz"
    elif bad == "first":
        off[0] = 1
        msg = "row_offsets[0] must be 0"
    else:
        col[3] = 4
        msg = "column index out of range"
    s = np.full(4, 0.25, np.float32)
    with pytest.raises(GardeniaError, match=msg.replace("[", r"\[").replace("]", r"\]")
                       .replace("(", r"\(").replace(")", r"\)")):
        verify_residual(_t(off, np.int32, device), _t(col, np.int32, device),
                        _t(deg, np.int32, device), _t(s, np.float32, device))
