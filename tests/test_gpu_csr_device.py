"""fill_data() on the device (gd_csr_from_edges_device) against the host
builder (gd_csr_from_edges, pinned to the reference's read_graph goldens by
the oracle tests): every flag combination on edge lists with self-loops,
adjacent and non-adjacent duplicates; the transpose flag cleared by the first
mirrored edge; an RMAT-20 edge list."""
import itertools

import numpy as np
import pytest

from gardenia_ocl_amd import graph

pytestmark = pytest.mark.gpu


def _both(device, m, src, dst, wt, **flags):
    import torch

    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(device)  # noqa: E731
    d = graph.csr_from_edges_device(m, t(src, np.int32), t(dst, np.int32),
                                    None if wt is None else t(wt, np.float32), **flags)
    h = graph.csr_from_edges(m, src, dst, wt, **flags)
    return {k: v.cpu().numpy() for k, v in d.items()}, h


def _check(d, h, weights=True):
    np.testing.assert_array_equal(d["row_offsets"], h.row_offsets)
    np.testing.assert_array_equal(d["column_indices"], h.column_indices)
    np.testing.assert_array_equal(d["degree"], h.degree)
    if weights and h.weight is not None:
        np.testing.assert_array_equal(d["weight"], h.weight)


FLAGS = list(itertools.product([0, 1], repeat=6))


@pytest.mark.parametrize("sym,tr,srt,rs,rr,mir", FLAGS)
def test_device_build_matches_host_every_flag(device, sym, tr, srt, rs, rr, mir):
    rng = np.random.default_rng(7)
    m, n = 97, 900
    src = rng.integers(0, m, n).astype(np.int32)
    dst = rng.integers(0, m, n).astype(np.int32)
    src[::9] = dst[::9]                      # self-loops
    src[1::7], dst[1::7] = src[0::7][: len(src[1::7])], dst[0::7][: len(dst[1::7])]  # adjacent dups
    # weights a symmetric function of the endpoints, so every record with the
    # same (row, column) -- duplicates and mirrored edges alike -- has the same
    # weight (std::sort's order of equal keys is unspecified)
    lo, hi = np.minimum(src, dst).astype(np.int64), np.maximum(src, dst).astype(np.int64)
    wt = ((lo * 131 + hi) % 1009).astype(np.float32) / 7.0
    d, h = _both(device, m, src, dst, wt, symmetrize=bool(sym), transpose=bool(tr),
                 sorted=bool(srt), remove_selfloops=bool(rs), remove_redundents=bool(rr),
                 mirror_selfloops=bool(mir))
    _check(d, h)


def test_transpose_cleared_by_first_mirrored_edge(device):
    """symmetrise + transpose: self-loops before the first mirrored edge are
    placed transposed, everything after the mirrored edge is not."""
    m = 6
    src = np.array([1, 2, 3, 4, 0, 5, 3], np.int32)
    dst = np.array([1, 2, 0, 4, 5, 1, 3], np.int32)
    for mir in (False, True):
        d, h = _both(device, m, src, dst, None, symmetrize=True, transpose=True, sorted=False,
                     remove_selfloops=False, remove_redundents=False, mirror_selfloops=mir)
        _check(d, h, weights=False)


def test_empty_and_unweighted(device):
    import torch

    d = graph.csr_from_edges_device(5, torch.zeros(0, dtype=torch.int32, device=device),
                                    torch.zeros(0, dtype=torch.int32, device=device))
    np.testing.assert_array_equal(d["row_offsets"].cpu().numpy(), np.zeros(6, np.int32))
    assert d["column_indices"].numel() == 0 and d["degree"].cpu().numpy().sum() == 0


@pytest.mark.parametrize("transpose,srt", [(True, False), (False, False), (True, True)])
def test_device_build_rmat20(device, transpose, srt):
    src, dst = graph.rmat_edges(20, 16, seed=3)
    d, h = _both(device, 1 << 20, src, dst, None, transpose=transpose, sorted=srt)
    _check(d, h, weights=False)


def test_device_build_rejects_out_of_range(device):
    import torch

    from gardenia_ocl_amd import GardeniaError

    s = torch.tensor([0, 5], dtype=torch.int32, device=device)
    d = torch.tensor([1, 2], dtype=torch.int32, device=device)
    with pytest.raises(GardeniaError, match="out of range"):
        graph.csr_from_edges_device(5, s, d)
