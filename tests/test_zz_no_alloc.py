"""gd_pr_run_noalloc, the entry point that allocates nothing, against
gd_pr_plan_run (the PageRankPlan path) on the same device-resident inputs,
through the C ABI: two calls in a row, the second on a smaller graph (the
static scratch keeps the first call's contents, so the second proves that
every call starts from its own state).  Bar: the plan's iteration count and
scores bit for bit -- both sum every row in CSR order with sequential FP32
rounding -- and so the oracle's."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import assert_same_iterations
from gardenia_ocl_amd import PageRankPlan, _native as N, graph
from gardenia_ocl_amd.pr import make_options

pytestmark = pytest.mark.gpu


def _noalloc(device, in_csr, outdeg, tol, max_iter):
    import torch

    m = in_csr.m
    off = torch.from_numpy(in_csr.row_offsets).to(device)
    col = torch.from_numpy(in_csr.column_indices).to(device)
    deg = torch.from_numpy(np.ascontiguousarray(outdeg, np.int32)).to(device)
    s = torch.full((m,), float(np.float32(1.0) / np.float32(m)), dtype=torch.float32,
                   device=device)
    opts = make_options(tol=tol, max_iter=max_iter)
    st = N.PrStats()
    errs = np.zeros(max(1, max_iter), np.float64)
    stream = torch.cuda.current_stream()
    N.call("gd_pr_run_noalloc", m, in_csr.nnz, off.data_ptr(), col.data_ptr(), deg.data_ptr(),
           s.data_ptr(), ctypes.byref(opts), ctypes.c_void_p(stream.cuda_stream),
           ctypes.byref(st), errs.ctypes.data_as(ctypes.c_void_p))
    torch.cuda.synchronize()
    return s.cpu().numpy(), st.as_dict(), errs[: st.iterations].copy()


def _plan(device, in_csr, outdeg, tol, max_iter):
    import torch

    m = in_csr.m
    off = torch.from_numpy(in_csr.row_offsets).to(device)
    col = torch.from_numpy(in_csr.column_indices).to(device)
    deg = torch.from_numpy(np.ascontiguousarray(outdeg, np.int32)).to(device)
    s = torch.full((m,), float(np.float32(1.0) / np.float32(m)), dtype=torch.float32,
                   device=device)
    plan = PageRankPlan(off, col, deg)
    st, errs = plan.run(s, tol=tol, max_iter=max_iter)
    plan.close()
    return s.cpu().numpy(), st, errs


def test_noalloc_matches_plan_two_calls_second_smaller(device):
    # RMAT-18 has rows longer than 64 and than a tile (the hub chain path);
    # then RMAT-12, smaller in every dimension
    for scale in (18, 12):
        out_csr, in_csr = graph.rmat(scale, 16, seed=9)
        assert in_csr.degree.max() > 4096 or scale == 12
        s, st, errs = _noalloc(device, in_csr, out_csr.degree, 1e-4, 100)
        p, pst, perrs = _plan(device, in_csr, out_csr.degree, 1e-4, 100)
        assert st["converged"] and st["iterations"] == pst["iterations"], (st, pst)
        np.testing.assert_array_equal(s, p)
        ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                      out_csr.degree, tol=1e-4, max_iter=100)
        np.testing.assert_array_equal(s, ref)
        assert st["iterations"] == it
        assert_same_iterations(pst["iterations"], perrs, it, ref_errs, 1e-4)
        np.testing.assert_allclose(errs, ref_errs, rtol=1e-12, atol=0)


def test_noalloc_rejects_bad_csr_and_sizes(device):
    import torch

    from gardenia_ocl_amd import GardeniaError

    off = torch.tensor([0, 2, 3], dtype=torch.int32, device=device)
    col = torch.tensor([1, 5, 0], dtype=torch.int32, device=device)  # 5 >= m
    deg = torch.tensor([1, 1], dtype=torch.int32, device=device)
    s = torch.full((2,), 0.5, dtype=torch.float32, device=device)
    opts = make_options(tol=1e-3, max_iter=10)
    st = N.PrStats()
    with pytest.raises(GardeniaError, match="column index out of range"):
        N.call("gd_pr_run_noalloc", 2, 3, off.data_ptr(), col.data_ptr(), deg.data_ptr(),
               s.data_ptr(), ctypes.byref(opts), None, ctypes.byref(st), None)
    bad_off = torch.tensor([0, 3, 2], dtype=torch.int32, device=device)
    with pytest.raises(GardeniaError, match="prefix sum"):
        N.call("gd_pr_run_noalloc", 2, 2, bad_off.data_ptr(), col.data_ptr(), deg.data_ptr(),
               s.data_ptr(), ctypes.byref(opts), None, ctypes.byref(st), None)
    np.testing.assert_array_equal(s.cpu().numpy(), [0.5, 0.5])  # untouched
