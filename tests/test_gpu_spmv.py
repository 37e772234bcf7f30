"""GPU parity of the HIP CSR SpMV against the reference (golden vectors) and the
oracle.  All calls go through the C ABI (ctypes)."""
import numpy as np
import pytest

import oracle
from conftest import dataset
from gardenia_ocl_amd import SpmvPlan, SpmvSolver, graph

pytestmark = pytest.mark.gpu

DATASETS = ["test_pr.mtx", "4.mtx", "4w.mtx", "chesapeake.mtx"]
REL_TOL = 1e-6  # maximum_relative_error bar where the summation order is the reference's


def _fp64_spmv(csr, Ax, x, y0):
    prod = np.asarray(Ax, np.float64) * np.asarray(x, np.float64)[csr.column_indices]
    rows = np.repeat(np.arange(csr.m), np.diff(csr.row_offsets))
    out = np.asarray(y0, np.float64).copy()
    np.add.at(out, rows, prod)
    return out


def _check_long_rows(csr, y, ref, Ax, x, y0):
    """Every row bit-identical: rows of <= 64 nonzeros summed by their owner
    thread, longer rows by one lane of their own workgroup, all in CSR order.
    The reference's own distance to the FP64 product is what remains."""
    np.testing.assert_array_equal(y, ref)
    assert oracle.max_rel_err(y, ref) == 0.0
    y64 = _fp64_spmv(csr, Ax, x, y0)
    assert np.abs(y - y64).max() == np.abs(ref - y64).max()


def _key(ds):
    return ds.replace(".mtx", "").replace(".", "_")


def _gpu_spmv(device, csr, Ax, x, y0, reps=1):
    import torch

    Ap = torch.from_numpy(csr.row_offsets).to(device)
    Aj = torch.from_numpy(csr.column_indices).to(device)
    tAx = torch.from_numpy(np.ascontiguousarray(Ax, np.float32)).to(device)
    plan = SpmvPlan(Ap, Aj, tAx, num_cols=len(x))
    tx = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(device)
    ty = torch.from_numpy(np.array(y0, np.float32)).to(device)
    for _ in range(reps):
        plan.run(tx, ty)
    torch.cuda.synchronize()
    info = plan.info()
    plan.close()
    return ty.cpu().numpy(), info


@pytest.mark.parametrize("ds", DATASETS)
def test_spmv_datasets_bit_exact_vs_reference(golden, device, ds):
    arrays, _ = golden
    k = _key(ds)
    csr = graph.read_spmv_matrix(dataset(ds))
    y, _ = _gpu_spmv(device, csr, arrays[f"{k}/spmv/Ax"], arrays[f"{k}/spmv/x"],
                     arrays[f"{k}/spmv/y0"])
    np.testing.assert_array_equal(y, arrays[f"{k}/spmv/y"])


def test_spmv_host_solver_interface(golden):
    arrays, _ = golden
    csr = graph.read_spmv_matrix(dataset("chesapeake.mtx"))
    y = np.array(arrays["chesapeake/spmv/y0"], np.float32)
    ms = SpmvSolver(csr.m, csr.nnz, csr.row_offsets, csr.column_indices,
                    arrays["chesapeake/spmv/Ax"], arrays["chesapeake/spmv/x"], y)
    np.testing.assert_array_equal(y, arrays["chesapeake/spmv/y"])
    assert ms >= 0.0


def test_spmv_er_bit_exact_vs_reference(golden, device):
    arrays, meta = golden
    md = meta["spmv"]["er12"]
    er = graph.er_matrix(12, 32, seed=md["seed_matrix"])
    sd = md["seed_values"]
    Ax, x, y0 = (graph.uniform_pm1(n, sd + k) for k, n in ((0, er.nnz), (1, er.m), (2, er.m)))
    y, info = _gpu_spmv(device, er, Ax, x, y0)
    np.testing.assert_array_equal(y, arrays["er12/spmv/y"])
    assert info["hubs"] == 0


def test_spmv_rmat_vs_reference(golden, device):
    arrays, meta = golden
    _, ic = graph.rmat(12, 16, seed=1)
    sd = meta["spmv"]["rmat12_in"]["seed"]
    Ax, x, y0 = (graph.uniform_pm1(n, sd + k) for k, n in ((0, ic.nnz), (1, ic.m), (2, ic.m)))
    y, _ = _gpu_spmv(device, ic, Ax, x, y0)
    _check_long_rows(ic, y, arrays["rmat12_in/spmv/y"], Ax, x, y0)


@pytest.mark.parametrize("scale", [16, 18])
def test_spmv_power_law_long_rows_vs_oracle(device, scale):
    _, ic = graph.rmat(scale, 16, seed=9)
    Ax, x, y0 = (graph.uniform_pm1(n, 40 + k) for k, n in ((0, ic.nnz), (1, ic.m), (2, ic.m)))
    y, info = _gpu_spmv(device, ic, Ax, x, y0)
    ref = oracle.spmv(ic.m, ic.row_offsets, ic.column_indices, Ax, x, y0)
    assert info["hubs"] > 0
    _check_long_rows(ic, y, ref, Ax, x, y0)


def test_spmv_er_2_20_bit_exact_vs_oracle(device):
    er = graph.er_matrix(20, 32, seed=11)
    Ax = er.weight
    x = graph.uniform_pm1(er.m, 12)
    y0 = np.zeros(er.m, np.float32)
    y, _ = _gpu_spmv(device, er, Ax, x, y0, reps=3)
    ref = y0
    for _ in range(3):
        ref = oracle.spmv(er.m, er.row_offsets, er.column_indices, Ax, x, ref)
    np.testing.assert_array_equal(y, ref)


def test_spmv_empty_rows_and_matrix(device):
    m = 5000
    rng = np.random.default_rng(1)
    src = rng.integers(0, m, 3000).astype(np.int32)
    dst = rng.integers(0, m, 3000).astype(np.int32)
    csr = graph.csr_from_edges(m, src, dst, sorted=True, remove_selfloops=False)
    Ax = graph.uniform_pm1(csr.nnz, 1)
    x = graph.uniform_pm1(m, 2)
    y0 = graph.uniform_pm1(m, 3)
    y, _ = _gpu_spmv(device, csr, Ax, x, y0)
    np.testing.assert_array_equal(y, oracle.spmv(m, csr.row_offsets, csr.column_indices, Ax, x, y0))
    empty = graph.csr_from_edges(m, np.zeros(0, np.int32), np.zeros(0, np.int32))
    y, _ = _gpu_spmv(device, empty, np.zeros(0, np.float32), x, y0)
    np.testing.assert_array_equal(y, y0)


def _gpu_spmv_mode(device, csr, Ax, x, y0, mode, reps=1):
    import torch

    Ap = torch.from_numpy(csr.row_offsets).to(device)
    Aj = torch.from_numpy(csr.column_indices).to(device)
    tAx = torch.from_numpy(np.ascontiguousarray(Ax, np.float32)).to(device)
    plan = SpmvPlan(Ap, Aj, tAx, num_cols=len(x), mode=mode)
    tx = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(device)
    ty = torch.from_numpy(np.array(y0, np.float32)).to(device)
    for _ in range(reps):
        plan.run(tx, ty)
    torch.cuda.synchronize()
    return ty.cpu().numpy(), plan, (Ap, Aj, tAx, tx, ty)


def test_spmv_sliced_er_bit_exact(golden, device):
    arrays, meta = golden
    md = meta["spmv"]["er12"]
    er = graph.er_matrix(12, 32, seed=md["seed_matrix"])
    sd = md["seed_values"]
    Ax, x, y0 = (graph.uniform_pm1(n, sd + k) for k, n in ((0, er.nnz), (1, er.m), (2, er.m)))
    y, plan, _ = _gpu_spmv_mode(device, er, Ax, x, y0, "sliced")
    assert plan.mode == "sliced"
    np.testing.assert_array_equal(y, arrays["er12/spmv/y"])
    big = graph.er_matrix(21, 32, seed=13)
    x = graph.uniform_pm1(big.m, 14)
    y, plan, _ = _gpu_spmv_mode(device, big, big.weight, x, np.zeros(big.m, np.float32), "auto",
                                reps=3)
    assert plan.mode == "sliced"  # x = 8 MB > one XCD's L2, rows of 32
    ref = np.zeros(big.m, np.float32)
    for _ in range(3):
        ref = oracle.spmv(big.m, big.row_offsets, big.column_indices, big.weight, x, ref)
    np.testing.assert_array_equal(y, ref)


def test_spmv_sliced_long_rows_are_sequential(golden, device):
    """Each slice pass adds a row's entries of that slice onto y in column
    order, so rows longer than 64 are bit-identical to the reference too
    (RMAT-12 in-CSR, max row > 1000)."""
    arrays, meta = golden
    _, ic = graph.rmat(12, 16, seed=1)
    assert ic.degree.max() > 64
    sd = meta["spmv"]["rmat12_in"]["seed"]
    Ax, x, y0 = (graph.uniform_pm1(n, sd + k) for k, n in ((0, ic.nnz), (1, ic.m), (2, ic.m)))
    y, plan, _ = _gpu_spmv_mode(device, ic, Ax, x, y0, "sliced")
    np.testing.assert_array_equal(y, arrays["rmat12_in/spmv/y"])


def test_spmv_sliced_refresh_values_and_fallback(device):
    import torch

    er = graph.er_matrix(14, 32, seed=21)
    x = graph.uniform_pm1(er.m, 22)
    y, plan, keep = _gpu_spmv_mode(device, er, er.weight, x, np.zeros(er.m, np.float32), "sliced")
    Ap, Aj, tAx, tx, ty = keep
    tAx.mul_(-0.5)
    plan.refresh_values()
    ty.zero_()
    plan.run(tx, ty)
    torch.cuda.synchronize()
    ref = oracle.spmv(er.m, er.row_offsets, er.column_indices, tAx.cpu().numpy(), x,
                      np.zeros(er.m, np.float32))
    np.testing.assert_array_equal(ty.cpu().numpy(), ref)
    # unsorted rows: forced sliced mode is rejected, auto falls back to tiles
    from gardenia_ocl_amd import GardeniaError

    rng = np.random.default_rng(0)
    m = 3000
    src = rng.integers(0, m, 20000).astype(np.int32)
    dst = rng.integers(0, m, 20000).astype(np.int32)
    un = graph.csr_from_edges(m, src, dst, sorted=False, remove_redundents=False)
    Ax = graph.uniform_pm1(un.nnz, 3)
    xx = graph.uniform_pm1(m, 4)
    with pytest.raises(GardeniaError, match="sorted columns"):
        _gpu_spmv_mode(device, un, Ax, xx, np.zeros(m, np.float32), "sliced")
    y, plan, _ = _gpu_spmv_mode(device, un, Ax, xx, np.zeros(m, np.float32), "auto")
    assert plan.mode == "tiles"
    np.testing.assert_array_equal(
        y, oracle.spmv(m, un.row_offsets, un.column_indices, Ax, xx, np.zeros(m, np.float32)))


def test_spmv_sliced_wide_matrix_many_slices(device):
    """2^16 rows over 2^24 columns (64 MB of x in 8 column slices of 8 MB),
    rows of 0..64 sorted columns spread over the slices, some empty: every row
    is the oracle's sequential sum, y carried from slice to slice."""
    rng = np.random.default_rng(31)
    rows, cols = 1 << 16, 1 << 24
    lens = rng.integers(0, 65, rows)
    lens[rng.random(rows) < 0.05] = 0
    off = np.zeros(rows + 1, np.int64)
    off[1:] = np.cumsum(lens)
    col = np.empty(int(off[-1]), np.int32)
    for r in range(rows):
        n = int(lens[r])
        if n:
            col[off[r]:off[r + 1]] = np.sort(rng.integers(0, cols - 64, n)) + np.arange(n)
    csr = graph.Csr(rows, off.astype(np.int32), col, lens.astype(np.int32))
    Ax = graph.uniform_pm1(csr.column_indices.shape[0], 32)
    x = graph.uniform_pm1(cols, 33)
    y0 = graph.uniform_pm1(rows, 34)
    y, plan, _ = _gpu_spmv_mode(device, csr, Ax, x, y0, "auto", reps=2)
    assert plan.mode == "sliced"
    ref = y0
    for _ in range(2):
        ref = oracle.spmv(rows, csr.row_offsets, csr.column_indices, Ax, x, ref)
    np.testing.assert_array_equal(y, ref)


def _csr_from_rows(rows, cols, row_cols):
    """CSR with the given sorted column lists (missing rows are empty)."""
    lens = np.zeros(rows, np.int64)
    for r, c in row_cols.items():
        lens[r] = len(c)
    off = np.zeros(rows + 1, np.int64)
    off[1:] = np.cumsum(lens)
    col = np.empty(int(off[-1]), np.int32)
    for r, c in row_cols.items():
        col[off[r]:off[r + 1]] = np.sort(np.asarray(c, np.int64))
    assert col.size == 0 or col.max() < cols
    return graph.Csr(rows, off.astype(np.int32), col, lens.astype(np.int32))


@pytest.mark.parametrize("case", ["empty_slices", "long_in_every_slice"])
def test_spmv_sliced_schedule_edge_cases(device, case):
    """Forced sliced mode, 2^13 rows x 2^16 columns (8 slices of 8192
    columns).  empty_slices: entries only in slices 0 and 5, one row of 5000
    entries inside slice 0 and one of 6000 inside slice 5 (longer than a
    2048-entry tile: the per-slice long-row chain), most rows empty.
    long_in_every_slice: one row of 3000 entries in each of the 8 slices next
    to short rows in every slice.  Every row is the oracle's sequential sum,
    two runs in a row (y carried through every pass of both)."""
    rng = np.random.default_rng(41)
    rows, cols, width = 1 << 13, 1 << 16, 1 << 13
    rc = {}
    if case == "empty_slices":
        rc[7] = rng.choice(width, 5000, replace=False)
        rc[4000] = 5 * width + rng.choice(width, 6000, replace=False)
        for r in rng.choice(rows, 300, replace=False):
            if r in (7, 4000):
                continue
            q = (0, 5)[int(r) % 2]
            rc[int(r)] = q * width + rng.choice(width, int(rng.integers(1, 40)), replace=False)
    else:
        rc[123] = np.concatenate([q * width + rng.choice(width, 3000, replace=False)
                                  for q in range(8)])
        for r in range(0, rows, 3):
            if r != 123:
                rc[r] = rng.choice(cols, int(rng.integers(1, 50)), replace=False)
    csr = _csr_from_rows(rows, cols, rc)
    Ax = graph.uniform_pm1(csr.column_indices.shape[0], 42)
    x = graph.uniform_pm1(cols, 43)
    y0 = graph.uniform_pm1(rows, 44)
    y, plan, _ = _gpu_spmv_mode(device, csr, Ax, x, y0, "sliced", reps=2)
    assert plan.mode == "sliced"
    ref = y0
    for _ in range(2):
        ref = oracle.spmv(rows, csr.row_offsets, csr.column_indices, Ax, x, ref)
    np.testing.assert_array_equal(y, ref)


@pytest.mark.parametrize("mode", ["auto", "tiles"])
def test_spmv_er22_config_size_bit_exact(device, mode):
    """The SpMV workload at its measured size: ER 2^22 x 2^22, 32 uniform
    columns per row (134,217,728 nonzeros), Ax and x ~ U(-1,1), y0 = 0, the
    bench's seeds; y += A x bit-identical to the oracle's restatement of
    src/spmv/omp_base.cc:17-28 in both modes (auto picks the XCD-sliced mode
    here)."""
    er = graph.er_matrix(22, 32, seed=1)
    assert er.nnz == 32 << 22
    x = graph.uniform_pm1(er.m, 2)
    y, plan, _ = _gpu_spmv_mode(device, er, er.weight, x, np.zeros(er.m, np.float32), mode,
                                reps=2)
    assert plan.mode == ("sliced" if mode == "auto" else "tiles")
    ref = np.zeros(er.m, np.float32)
    for _ in range(2):
        ref = oracle.spmv(er.m, er.row_offsets, er.column_indices, er.weight, x, ref)
    np.testing.assert_array_equal(y, ref)
