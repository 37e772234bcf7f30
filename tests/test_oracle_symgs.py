"""The SymGS oracle (oracle.symgs, a restatement of src/symgs/verifier.cc:50-77)
pinned against the reference's own SymGSSolver compiled from source
(tests/golden/ref_symgs.npz), a hand-computed sweep, and convergence of
repeated sweeps to the dense solve of an SPD system."""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN_DIR
from gardenia_ocl_amd import graph

CASES = ["stencil6", "rmat10_diag", "4w", "chesapeake"]


@pytest.fixture(scope="module")
def sg():
    arrays = np.load(os.path.join(GOLDEN_DIR, "ref_symgs.npz"), allow_pickle=False)
    with open(os.path.join(GOLDEN_DIR, "ref_symgs.json")) as f:
        meta = json.load(f)
    return arrays, meta


def case(arrays, name):
    return {k: arrays[f"{name}/{k}"] for k in ("row_offsets", "column_indices", "weight",
                                               "indices", "color_offsets", "x0", "b", "x1", "x3")}


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_symgs(sg, name):
    arrays, meta = sg
    c = case(arrays, name)
    args = (c["row_offsets"], c["column_indices"], c["indices"], c["weight"])
    x1 = oracle.symgs(*args, c["x0"], c["b"], c["color_offsets"])
    np.testing.assert_array_equal(x1, c["x1"])
    x3 = oracle.symgs(*args, c["x0"], c["b"], c["color_offsets"], sweeps=3)
    np.testing.assert_array_equal(x3, c["x3"])
    assert meta[name]["verdict_x1"] == "Correct"


def test_rows_without_diagonal_are_left_alone(sg):
    arrays, meta = sg
    c = case(arrays, "chesapeake")
    assert meta["chesapeake"]["rows_with_diag"] == 0
    np.testing.assert_array_equal(c["x1"], c["x0"])


def test_hand_computed_sweep():
    # A = [[4,1,0],[1,4,1],[0,1,4]], b = (1,2,3), x0 = 0; colors {0,2}, {1}.
    # forward: x0 = 1/4, x2 = 3/4, x1 = (2 - 1/4 - 3/4)/4 = 1/4
    # backward: x1 = 1/4 again, x0 = (1 - 1/4)/4 = 3/16, x2 = (3 - 1/4)/4 = 11/16
    Ap = np.array([0, 2, 5, 7], np.int32)
    Aj = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    Ax = np.array([4, 1, 1, 4, 1, 1, 4], np.float32)
    x = oracle.symgs(Ap, Aj, np.array([0, 2, 1], np.int32), Ax, np.zeros(3, np.float32),
                     np.array([1, 2, 3], np.float32), np.array([0, 2, 3], np.int32))
    np.testing.assert_array_equal(x, np.array([3 / 16, 1 / 4, 11 / 16], np.float32))


def test_repeated_sweeps_converge_to_dense_solve(sg):
    arrays, _ = sg
    c = case(arrays, "stencil6")
    m = len(c["x0"])
    A = np.zeros((m, m))
    rows = np.repeat(np.arange(m), np.diff(c["row_offsets"]))
    A[rows, c["column_indices"]] = c["weight"]
    assert np.allclose(A, A.T)
    x = oracle.symgs(c["row_offsets"], c["column_indices"], c["indices"], c["weight"], c["x0"],
                     c["b"], c["color_offsets"], sweeps=60)
    ref = np.linalg.solve(A, c["b"].astype(np.float64))
    assert np.max(np.abs(x - ref)) < 1e-5 * np.max(np.abs(ref)) + 1e-6


def test_first_fit_coloring_is_proper_and_ordered(sg):
    arrays, _ = sg
    for name in CASES:
        c = case(arrays, name)
        m = len(c["x0"])
        csr = graph.Csr(m, c["row_offsets"], c["column_indices"], np.diff(c["row_offsets"]))
        colors, nc = graph.color_first_fit(csr)
        rows = np.repeat(np.arange(m), np.diff(c["row_offsets"]))
        off = rows != c["column_indices"]
        assert np.all(colors[rows[off]] != colors[c["column_indices"][off]])
        idx, co = graph.color_order(colors, nc)
        np.testing.assert_array_equal(idx, c["indices"])
        np.testing.assert_array_equal(co, c["color_offsets"])
        assert np.all(np.diff(colors[idx]) >= 0)


def test_with_dominant_diagonal():
    # the stencil is already dominant with a diagonal in place: same structure
    c = graph.stencil27(5)
    d = graph.with_dominant_diagonal(c, margin=1.0)
    np.testing.assert_array_equal(d.row_offsets, c.row_offsets)
    np.testing.assert_array_equal(d.column_indices, c.column_indices)
    rows = np.repeat(np.arange(d.m), d.degree)
    diag = d.column_indices == rows
    off = np.bincount(rows[~diag], weights=np.abs(d.weight[~diag]).astype(np.float64), minlength=d.m)
    np.testing.assert_allclose(d.weight[diag], off + 1.0, rtol=1e-6)
    np.testing.assert_array_equal(d.weight[~diag], c.weight[c.column_indices != rows])
    # rows without a diagonal get one, in column order; self-loops are replaced
    src = np.array([0, 1, 2, 2, 3], np.int32)
    dst = np.array([1, 2, 2, 0, 1], np.int32)
    w = np.array([0.5, -2.0, 7.0, 1.5, -0.25], np.float32)
    e = graph.csr_from_edges(4, src, dst, w, symmetrize=True, sorted=True, remove_selfloops=False)
    f = graph.with_dominant_diagonal(e, margin=0.5)
    rows = np.repeat(np.arange(4), f.degree)
    key = rows * 4 + f.column_indices
    assert np.all(np.diff(key) > 0)
    assert np.all(np.bincount(rows[f.column_indices == rows], minlength=4) == 1)
    # row 2: |-2| (from 1) + |1.5| (to 0) + 0.5; the self-loop 7.0 is gone
    assert f.weight[(rows == 2) & (f.column_indices == 2)][0] == np.float32(4.0)
