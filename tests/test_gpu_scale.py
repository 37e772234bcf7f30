"""Parity at full size: RMAT-24 (16.8 M vertices, 263 M edges), the workload
the throughput figure is quoted on.  The HIP plan against the oracle's
restatement of src/pr/omp_base.cc on the same in-CSR: one iteration, 20 fixed
iterations and a run to tol 1e-6, every score bit-identical, the same L1
values up to the order of their FP64 sums and the same iteration count."""
import numpy as np
import pytest

import oracle
from conftest import L1_GAP, assert_same_iterations
from gardenia_ocl_amd import PageRankPlan, graph

pytestmark = pytest.mark.gpu

SCALE = 24


@pytest.fixture(scope="module")
def rmat24(device):
    import torch

    _, ic = graph.rmat(SCALE, 16, seed=1, want_out=False)
    od = np.bincount(ic.column_indices, minlength=ic.m).astype(np.int32)
    off = torch.from_numpy(ic.row_offsets).to(device)
    col = torch.from_numpy(ic.column_indices).to(device)
    deg = torch.from_numpy(od).to(device)
    plan = PageRankPlan(off, col, deg)
    yield ic, od, plan
    plan.close()


def _run(plan, m, device, tol, max_iter):
    import torch

    s = torch.full((m,), float(np.float32(1.0) / np.float32(m)), dtype=torch.float32,
                   device=device)
    st, errs = plan.run(s, tol=tol, max_iter=max_iter)
    return s.cpu().numpy(), st, errs


def test_rmat24_first_iteration(rmat24, device):
    ic, od, plan = rmat24
    assert plan.info()["relabel"] == 1 and plan.info()["hubs"] > 0
    s, st, errs = _run(plan, ic.m, device, 0.0, 1)
    ref, it, ref_errs = oracle.pr(ic.m, ic.row_offsets, ic.column_indices, od, tol=0.0,
                                  max_iter=1)
    # every row in the reference's order and rounding: owner threads (<= 64
    # in-edges), wavefront exact scans (up to a tile), hub chains and scans
    np.testing.assert_array_equal(s, ref)
    assert abs(errs[0] - ref_errs[0]) <= L1_GAP


def test_rmat24_twenty_fixed_iterations(rmat24, device):
    ic, od, plan = rmat24
    s, st, errs = _run(plan, ic.m, device, 0.0, 20)
    assert st["iterations"] == 20 and not st["converged"]
    ref, it, ref_errs = oracle.pr(ic.m, ic.row_offsets, ic.column_indices, od, tol=0.0,
                                  max_iter=20)
    np.testing.assert_array_equal(s, ref)
    assert_same_iterations(st["iterations"], errs, it, ref_errs, 0.0)


def test_rmat24_tol_1e6(rmat24, device):
    ic, od, plan = rmat24
    s, st, errs = _run(plan, ic.m, device, 1e-6, 200)
    ref, it, ref_errs = oracle.pr(ic.m, ic.row_offsets, ic.column_indices, od, tol=1e-6,
                                  max_iter=200)
    assert st["converged"]
    # the oracle's own L1 in the iteration before its stop lies only ~5.5e-10
    # above tol: bit-identical scores decide it the reference's way
    tie, gap = assert_same_iterations(st["iterations"], errs, it, ref_errs, 1e-6)
    assert not tie and st["iterations"] == it
    np.testing.assert_array_equal(s, ref)
    rel = np.abs(s - ref) / ref
    x64, _ = oracle.pr_fp64(ic.m, ic.row_offsets, ic.column_indices, od, tol=1e-12)
    l1_gpu = np.abs(s.astype(np.float64) - x64).sum()
    l1_ref = np.abs(ref.astype(np.float64) - x64).sum()
    print(f"RMAT-24 tol 1e-6: {st['iterations']} iterations (oracle {it}), L1 at the stop "
          f"{errs[it - 1]:.6e} (oracle {ref_errs[it - 1]:.6e}), before it {errs[it - 2]:.6e} "
          f"(oracle {ref_errs[it - 2]:.6e}), max L1 gap {gap:.2e}, max rel {rel.max():.3e}, "
          f"L1 vs FP64 gpu {l1_gpu:.3e} reference {l1_ref:.3e}, oracle tie {tie}")
    assert l1_gpu <= 1.05 * l1_ref + 1e-7
