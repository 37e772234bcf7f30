"""SymGS on the GPU (gd_symgs_*) against the reference's own SymGSSolver
(golden vectors, tests/golden/ref_symgs.npz) and the oracle's sequential sweep
(src/symgs/verifier.cc:50-77): bit-exact, including rows of more than 64 and
more than 2048 off-diagonal entries (the exact-order workgroup scan), several
sweeps, empty colors, partial colorings, and the coloring checks."""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN_DIR
from gardenia_ocl_amd import GardeniaError, SymGSPlan, SymGSSolver, graph

pytestmark = pytest.mark.gpu

CASES = ["stencil6", "rmat10_diag", "4w", "chesapeake"]


@pytest.fixture(scope="module")
def sg():
    arrays = np.load(os.path.join(GOLDEN_DIR, "ref_symgs.npz"), allow_pickle=False)
    with open(os.path.join(GOLDEN_DIR, "ref_symgs.json")) as f:
        meta = json.load(f)
    return arrays, meta


def _t(device, a, dt):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dt)).to(device)


def _plan(device, off, col, val, idx, co):
    return SymGSPlan(_t(device, off, np.int32), _t(device, col, np.int32),
                     _t(device, val, np.float32), _t(device, idx, np.int32), co)


def _sweeps(device, plan, x0, b, sweeps):
    import torch

    x = _t(device, x0, np.float32)
    plan.run(x, _t(device, b, np.float32), sweeps)
    torch.cuda.synchronize()
    return x.cpu().numpy()


@pytest.mark.parametrize("name", CASES)
def test_symgs_matches_reference_golden(device, sg, name):
    arrays, _ = sg
    g = {k: arrays[f"{name}/{k}"] for k in ("row_offsets", "column_indices", "weight", "indices",
                                            "color_offsets", "x0", "b", "x1", "x3")}
    m = len(g["x0"])
    x = g["x0"].copy()
    SymGSSolver(m, len(g["column_indices"]), g["row_offsets"], g["column_indices"], g["indices"],
                g["weight"], x, g["b"], g["color_offsets"])
    np.testing.assert_array_equal(x, g["x1"])
    p = _plan(device, g["row_offsets"], g["column_indices"], g["weight"], g["indices"],
              g["color_offsets"])
    np.testing.assert_array_equal(_sweeps(device, p, g["x0"], g["b"], 3), g["x3"])
    np.testing.assert_array_equal(_sweeps(device, p, g["x0"], g["b"], 0), g["x0"])


def _rmat_symgs(scale, seed):
    """RMAT with self-loops kept, symmetrized, weights > 0, a diagonal entry on
    every row divisible by 3 (the SymGS driver's read flags)."""
    src, dst = graph.rmat_edges(scale, 16, seed=seed)
    m = 1 << scale
    d = np.arange(0, m, 3, dtype=np.int32)
    s = np.concatenate([src, d])
    t = np.concatenate([dst, d])
    w = np.abs(graph.uniform_pm1(len(s), seed + 3)) + np.float32(0.0625)
    w[len(src):] = 40.0
    return graph.csr_from_edges(m, s, t, w, symmetrize=True, transpose=False, sorted=True,
                                remove_selfloops=False, remove_redundents=True)


@pytest.mark.parametrize("scale,seed", [(14, 1), (16, 2)])
def test_symgs_rmat_hubs_bit_exact(device, scale, seed):
    c = _rmat_symgs(scale, seed)
    colors, nc = graph.color_first_fit(c)
    idx, co = graph.color_order(colors, nc)
    x0 = graph.uniform_pm1(c.m, seed + 10)
    b = graph.uniform_pm1(c.m, seed + 11)
    p = _plan(device, c.row_offsets, c.column_indices, c.weight, idx, co)
    info = p.info()
    assert info["n_exact_rows"] > 0 and info["max_row_len"] > 2048, info
    assert info["n_colors"] == nc and info["rows"] == c.m
    for sweeps in (1, 2):
        want = oracle.symgs(c.row_offsets, c.column_indices, idx, c.weight, x0, b, co, sweeps)
        np.testing.assert_array_equal(_sweeps(device, p, x0, b, sweeps), want)


def test_symgs_dominant_rmat_many_sweeps(device):
    """Weights of both signs, a dominant diagonal: sweeps converge, every
    sweep bit-exact (long rows included)."""
    src, dst = graph.rmat_edges(15, 16, seed=8)
    w = graph.uniform_pm1(len(src), 9)
    c = graph.csr_from_edges(1 << 15, src, dst, w, symmetrize=True, transpose=False, sorted=True,
                             remove_selfloops=False, remove_redundents=True)
    c = graph.with_dominant_diagonal(c)
    colors, nc = graph.color_first_fit(c)
    idx, co = graph.color_order(colors, nc)
    b = graph.uniform_pm1(c.m, 10)
    x0 = np.zeros(c.m, np.float32)
    p = _plan(device, c.row_offsets, c.column_indices, c.weight, idx, co)
    assert p.info()["n_exact_rows"] > 0
    want = oracle.symgs(c.row_offsets, c.column_indices, idx, c.weight, x0, b, co, 5)
    got = _sweeps(device, p, x0, b, 5)
    np.testing.assert_array_equal(got, want)
    assert np.all(np.isfinite(got)) and np.abs(got).max() < 2.0


def test_symgs_stencil_bit_exact_and_converges(device):
    c = graph.stencil27(32)
    colors, nc = graph.color_first_fit(c)
    idx, co = graph.color_order(colors, nc)
    assert nc == 8
    x0 = np.zeros(c.m, np.float32)
    b = graph.uniform_pm1(c.m, 5)
    p = _plan(device, c.row_offsets, c.column_indices, c.weight, idx, co)
    want = oracle.symgs(c.row_offsets, c.column_indices, idx, c.weight, x0, b, co, 4)
    got = _sweeps(device, p, x0, b, 4)
    np.testing.assert_array_equal(got, want)
    # residual falls sweep over sweep (diagonally dominant)
    rows = np.repeat(np.arange(c.m), c.degree)

    def res(x):
        r = b.astype(np.float64).copy()
        np.subtract.at(r, rows, c.weight.astype(np.float64) * x[c.column_indices])
        return np.linalg.norm(r)

    assert res(got) < 0.05 * res(x0)


def test_symgs_empty_colors_and_partial_coloring(device):
    c = _rmat_symgs(12, 4)
    colors, nc = graph.color_first_fit(c)
    idx, co = graph.color_order(colors, nc)
    # empty colors in front, between and behind; the first color left out
    co2 = np.concatenate([[co[1]], [co[1]], co[2:3], co[2:3], co[3:], [co[-1]]]).astype(np.int32)
    x0 = graph.uniform_pm1(c.m, 1)
    b = graph.uniform_pm1(c.m, 2)
    p = _plan(device, c.row_offsets, c.column_indices, c.weight, idx, co2)
    want = oracle.symgs(c.row_offsets, c.column_indices, idx, c.weight, x0, b, co2, 2)
    np.testing.assert_array_equal(_sweeps(device, p, x0, b, 2), want)
    skipped = idx[: co[1]]
    np.testing.assert_array_equal(_sweeps(device, p, x0, b, 2)[skipped], x0[skipped])


def test_symgs_diagonal_only_is_b_over_diag(device):
    m = 1000
    off = np.arange(m + 1, dtype=np.int32)
    col = np.arange(m, dtype=np.int32)
    val = np.abs(graph.uniform_pm1(m, 9)) + np.float32(0.5)
    val[::7] = 0.0  # zero diagonal: row left alone
    b = graph.uniform_pm1(m, 8)
    x0 = graph.uniform_pm1(m, 7)
    p = _plan(device, off, col, val, np.arange(m, dtype=np.int32), [0, m])
    got = _sweeps(device, p, x0, b, 1)
    want = np.where(val != 0, b / np.where(val != 0, val, 1), x0).astype(np.float32)
    np.testing.assert_array_equal(got, want)


def test_symgs_rejects_bad_colorings(device):
    c = _rmat_symgs(10, 6)
    idx = np.arange(c.m, dtype=np.int32)
    with pytest.raises(GardeniaError, match="not proper"):
        _plan(device, c.row_offsets, c.column_indices, c.weight, idx, [0, c.m])
    colors, nc = graph.color_first_fit(c)
    idx, co = graph.color_order(colors, nc)
    dup = idx.copy()
    dup[1] = dup[0]
    with pytest.raises(GardeniaError, match="more than once"):
        _plan(device, c.row_offsets, c.column_indices, c.weight, dup, co)
    bad = idx.copy()
    bad[3] = c.m + 5
    with pytest.raises(GardeniaError, match="outside"):
        _plan(device, c.row_offsets, c.column_indices, c.weight, bad, co)
    with pytest.raises(GardeniaError, match="non-decreasing"):
        _plan(device, c.row_offsets, c.column_indices, c.weight, idx, [0, 5, 3, c.m])
