"""The reference's link-time plug-in boundary on the GPU: the C++ PRSolver and
SpmvSolver symbols of libgardenia_hip.so (src/pr/pr.h:23, src/spmv/spmv.h:29),
called through their mangled names exactly as a driver linked against the
library calls them.  Their stdout must carry the reference's lines
(src/pr/omp_base.cc:42,46-47; src/spmv/omp_base.cc:30): the per-iteration
" %2d    %lf" L1 values of the golden run, "iterations = N." with the
reference's iter+1 rule, and the runtime line."""
import ctypes
import os
import re
import sys
import tempfile

import numpy as np
import pytest

from conftest import dataset
from gardenia_ocl_amd import LIB_PATH, graph

pytestmark = pytest.mark.gpu

PR_SYM = "_Z8PRSolveriiPiS_S_S_S_Pf"
SPMV_SYM = "_Z10SpmvSolveriiPiS_PfS0_S0_"
DATASETS = ["test_pr.mtx", "4.mtx", "4w.mtx", "chesapeake.mtx"]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _captured(fn, *args):
    """Runs fn(*args) with the process's fd 1 sent to a file; returns the text
    the C code printed."""
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tf:
        os.dup2(tf.fileno(), 1)
        try:
            fn(*args)
        finally:
            libc.fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
        tf.seek(0)
        return tf.read().decode()


def _pr_symbol():
    lib = ctypes.CDLL(LIB_PATH)
    f = getattr(lib, PR_SYM)
    f.restype = None
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6
    return f


def _run_pr(in_csr, out_csr, env):
    m = in_csr.m
    scores = np.full(m, np.float32(1.0) / np.float32(m), np.float32)
    arrs = [np.ascontiguousarray(a, np.int32) for a in
            (in_csr.row_offsets, in_csr.column_indices, out_csr.row_offsets,
             out_csr.column_indices, out_csr.degree)]
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        text = _captured(_pr_symbol(), m, in_csr.nnz, *[_p(a) for a in arrs], _p(scores))
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    lines = re.findall(r"^ *(\d+) +([0-9.]+)$", text, flags=re.M)
    its = re.findall(r"iterations = (\d+)\.", text)
    assert re.search(r"runtime \[hip_pull\] = [0-9.]+ ms\.", text), text
    return scores, [(int(a), float(b)) for a, b in lines], int(its[-1]), text


@pytest.mark.parametrize("ds", DATASETS)
def test_cpp_prsolver_prints_the_reference_lines(golden, device, ds):
    arrays, meta = golden
    out_csr, in_csr = graph.read_pr_graph(dataset(ds), is_directed=True)
    key = ds.replace(".mtx", "").replace(".", "_")
    scores, lines, printed, text = _run_pr(in_csr, out_csr, {})
    ref_errs = arrays[f"{key}/pr_directed/errors_printed"]
    assert printed == meta["datasets"][ds]["pr_directed"]["iterations_printed"], text
    assert [i for i, _ in lines] == list(range(1, len(ref_errs) + 1)), text
    # %lf prints 6 decimals
    np.testing.assert_allclose([e for _, e in lines], ref_errs, atol=5.1e-7)
    np.testing.assert_array_equal(scores, arrays[f"{key}/pr_directed/scores"])


def test_cpp_prsolver_iteration_rule_when_not_converged(device):
    """No iteration converges: the reference prints MAX_ITER+1 after its loop
    (omp_base.cc:23,46), and 1 when MAX_ITER is 0."""
    out_csr, in_csr = graph.read_pr_graph(dataset("chesapeake.mtx"), is_directed=True)
    _, lines, printed, text = _run_pr(in_csr, out_csr, {"GD_PR_TOL": "0", "GD_PR_MAX_ITER": "5"})
    assert [i for i, _ in lines] == [1, 2, 3, 4, 5] and printed == 6, text
    _, lines, printed, text = _run_pr(in_csr, out_csr, {"GD_PR_MAX_ITER": "0"})
    assert lines == [] and printed == 1, text


@pytest.mark.parametrize("ds", DATASETS)
def test_cpp_spmvsolver_matches_golden(golden, device, ds):
    """y += A x through the mangled SpmvSolver, bit-identical to the reference's
    golden y (every row summed in CSR order)."""
    arrays, _ = golden
    key = ds.replace(".mtx", "").replace(".", "_")
    off = np.ascontiguousarray(arrays[f"{key}/spmv/row_offsets"], np.int32)
    col = np.ascontiguousarray(arrays[f"{key}/spmv/column_indices"], np.int32)
    Ax = np.ascontiguousarray(arrays[f"{key}/spmv/Ax"], np.float32)
    x = np.ascontiguousarray(arrays[f"{key}/spmv/x"], np.float32)
    y = np.array(arrays[f"{key}/spmv/y0"], np.float32)
    lib = ctypes.CDLL(LIB_PATH)
    f = getattr(lib, SPMV_SYM)
    f.restype = None
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
    m = off.shape[0] - 1
    text = _captured(f, m, int(off[-1]), _p(off), _p(col), _p(Ax), _p(x), _p(y))
    assert re.search(r"runtime \[hip_tile\] = [0-9.]+ ms\.", text), text
    np.testing.assert_array_equal(y, arrays[f"{key}/spmv/y"])
