"""The library's host graph layer (gardenia_ocl_amd.graph) against the
reference's CSR builder: golden arrays from read_graph() and the oracle's
fill_data() restatement, bit-exact."""
import hashlib
import itertools

import numpy as np
import pytest

import oracle
from conftest import dataset
from gardenia_ocl_amd import graph

DATASETS = ["test_pr.mtx", "4.mtx", "4w.mtx", "chesapeake.mtx"]
FLAGS = {
    "pr_out": (0, 0, 0, 1, 1), "pr_in": (0, 1, 0, 1, 1), "spmv": (0, 0, 1, 0, 1),
    "sym_out": (1, 0, 0, 1, 1), "sym_in": (1, 1, 0, 1, 1),
}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("ds", DATASETS)
@pytest.mark.parametrize("variant", list(FLAGS))
def test_read_mtx_matches_reference(golden, ds, variant):
    arrays, _ = golden
    s, t, so, rs, rr = FLAGS[variant]
    c = graph.read_mtx(dataset(ds), symmetrize=bool(s), transpose=bool(t), sorted=bool(so),
                       remove_selfloops=bool(rs), remove_redundents=bool(rr))
    k = ds.replace(".mtx", "").replace(".", "_") + "/" + variant
    np.testing.assert_array_equal(c.row_offsets, arrays[k + "/row_offsets"])
    np.testing.assert_array_equal(c.column_indices, arrays[k + "/column_indices"])
    np.testing.assert_array_equal(c.weight, arrays[k + "/weight"])
    np.testing.assert_array_equal(c.degree, arrays[k + "/degree"])


@pytest.mark.parametrize("scale", [12, 16])
def test_rmat_csr_matches_reference(golden, scale):
    _, meta = golden
    h = meta["rmat"][str(scale)]["sha256"]
    o, i = graph.rmat(scale, 16, seed=1)
    for f in ("row_offsets", "column_indices", "degree"):
        assert sha(getattr(o, f)) == h["out_" + f]
        assert sha(getattr(i, f)) == h["in_" + f]
    assert meta["rmat"][str(scale)]["nnz"] == o.nnz == i.nnz


def test_rmat_invariants():
    o, i = graph.rmat(14, 16, seed=3)
    m = o.m
    for c in (o, i):
        assert c.row_offsets[0] == 0 and c.row_offsets[-1] == c.nnz
        assert np.all(np.diff(c.row_offsets) >= 0)
        np.testing.assert_array_equal(c.degree, np.diff(c.row_offsets))
        rows = np.repeat(np.arange(m), c.degree)
        assert np.all(c.column_indices != rows)  # no self-loops
        key = rows.astype(np.int64) * m + c.column_indices
        assert np.all(np.diff(key) > 0)  # sorted, unique
    # in-CSR is the transpose of the out-CSR
    ro = np.repeat(np.arange(m), o.degree)
    ri = np.repeat(np.arange(m), i.degree)
    a = np.sort(ro.astype(np.int64) * m + o.column_indices)
    b = np.sort(i.column_indices.astype(np.int64) * m + ri)
    np.testing.assert_array_equal(a, b)
    # out-degree from the in-CSR columns
    np.testing.assert_array_equal(np.bincount(i.column_indices, minlength=m), o.degree)


def test_rmat_is_thread_count_and_call_independent():
    a, _ = graph.rmat(12, 8, seed=9, want_in=False)
    b, _ = graph.rmat(12, 8, seed=9, want_in=False)
    np.testing.assert_array_equal(a.column_indices, b.column_indices)
    src, dst = graph.rmat_edges(12, 8, seed=9)
    key = np.unique((src.astype(np.int64) << 12) | dst)
    key = key[(key >> 12) != (key & 4095)]
    np.testing.assert_array_equal((key & 4095).astype(np.int32), a.column_indices)
    c, _ = graph.rmat(12, 8, seed=10, want_in=False)
    assert not np.array_equal(a.column_indices[:1000], c.column_indices[:1000])


def test_rmat_skew_matches_quadrant_probabilities():
    src, dst = graph.rmat_edges(16, 16, seed=1)
    # P(top bit of src == 0) = a + b = 0.76, P(top bit of dst == 0) = a + c = 0.76
    assert abs(np.mean((src >> 15) == 0) - 0.76) < 0.003
    assert abs(np.mean((dst >> 15) == 0) - 0.76) < 0.003
    both = np.mean(((src >> 15) == 0) & ((dst >> 15) == 0))
    assert abs(both - 0.57) < 0.003


def test_er_matrix_invariants():
    er = graph.er_matrix(14, 32, seed=5)
    assert er.nnz == 32 << 14
    np.testing.assert_array_equal(er.degree, 32)
    cols = er.column_indices.reshape(-1, 32)
    assert np.all(np.diff(cols, axis=1) > 0)  # sorted and distinct per row
    assert cols.min() >= 0 and cols.max() < er.m
    assert er.weight.min() >= -1.0 and er.weight.max() < 1.0
    assert abs(er.weight.mean()) < 0.01
    # uniform columns
    hist = np.bincount(cols.ravel() >> 10, minlength=16)
    assert hist.min() > 0.9 * hist.mean()


@pytest.mark.parametrize("seed", range(6))
def test_csr_from_edges_matches_oracle_all_flags(seed):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(5, 40))
    n = int(rng.integers(0, 200))
    src = rng.integers(0, m, n).astype(np.int32)
    dst = rng.integers(0, m, n).astype(np.int32)
    # plenty of duplicates and self-loops, in random order
    wt = rng.uniform(-2, 2, n).astype(np.float32)
    for flags in itertools.product([0, 1], repeat=6):
        sym, tr, so, rs, rr, mir = flags
        kw = dict(symmetrize=bool(sym), transpose=bool(tr), sorted=bool(so),
                  remove_selfloops=bool(rs), remove_redundents=bool(rr), mirror_selfloops=bool(mir))
        a = graph.csr_from_edges(m, src, dst, np.abs(wt), **kw)
        b = oracle.csr_from_edges(m, src, dst, np.abs(wt), **kw)
        np.testing.assert_array_equal(a.row_offsets, b.row_offsets, err_msg=str(kw))
        np.testing.assert_array_equal(a.column_indices, b.column_indices, err_msg=str(kw))
        np.testing.assert_array_equal(a.weight, b.weight, err_msg=str(kw))
        np.testing.assert_array_equal(a.degree, b.degree, err_msg=str(kw))


def test_csr_from_edges_rejects_out_of_range():
    from gardenia_ocl_amd import GardeniaError

    with pytest.raises(GardeniaError, match="GD_ERR_INVALID"):
        graph.csr_from_edges(4, np.array([0, 5], np.int32), np.array([1, 2], np.int32))


def test_read_mtx_missing_file():
    from gardenia_ocl_amd import GardeniaError

    with pytest.raises(GardeniaError, match="GD_ERR_IO"):
        graph.read_mtx("/nonexistent/file.mtx")


def test_write_then_read_mtx_roundtrip(tmp_path):
    o, i = graph.rmat(10, 8, seed=4)
    src = np.repeat(np.arange(o.m, dtype=np.int32), o.degree)
    p = str(tmp_path / "g.mtx")
    graph.write_mtx(p, o.m, src, o.column_indices)
    o2, i2 = graph.read_pr_graph(p)
    np.testing.assert_array_equal(o2.column_indices, o.column_indices)
    np.testing.assert_array_equal(i2.row_offsets, i.row_offsets)
    np.testing.assert_array_equal(i2.column_indices, i.column_indices)


# --- .graph (graph2csr) and .gr (gr2csr) readers --------------------------------

FORMAT_DATASETS = ["4.graph", "test1.graph", "test2.graph", "test3.graph", "test4.graph",
                   "test5.graph", "test6.graph", "rmat8_raw.gr"]
FORMAT_FLAGS = {
    "pr_out": (0, 0, 0, 1, 1), "pr_in": (0, 1, 0, 1, 1), "spmv": (0, 0, 1, 0, 1),
    "sym_out": (1, 0, 0, 1, 1), "sym_in": (1, 1, 0, 1, 1), "raw": (0, 0, 0, 0, 0),
    "sym_raw": (1, 0, 0, 0, 0),
}


@pytest.fixture(scope="module")
def formats_golden():
    import os

    from conftest import GOLDEN_DIR

    return np.load(os.path.join(GOLDEN_DIR, "ref_formats.npz"), allow_pickle=False)


@pytest.mark.parametrize("ds", FORMAT_DATASETS)
@pytest.mark.parametrize("variant", list(FORMAT_FLAGS))
def test_read_graph_formats_match_reference(formats_golden, ds, variant):
    s, t, so, rs, rr = FORMAT_FLAGS[variant]
    kw = dict(symmetrize=bool(s), transpose=bool(t), sorted=bool(so), remove_selfloops=bool(rs),
              remove_redundents=bool(rr))
    direct = graph.read_gr if ds.endswith(".gr") else graph.read_dotgraph
    k = ds.replace(".", "_") + "/" + variant
    for c in (graph.read_graph(dataset(ds), **kw), direct(dataset(ds), **kw)):
        np.testing.assert_array_equal(c.row_offsets, formats_golden[k + "/row_offsets"])
        np.testing.assert_array_equal(c.column_indices, formats_golden[k + "/column_indices"])
        np.testing.assert_array_equal(c.weight, formats_golden[k + "/weight"])
        np.testing.assert_array_equal(c.degree, formats_golden[k + "/degree"])


def test_gr_reader_matches_oracle_fill_data():
    # gr2csr = the raw arcs through fill_data with self-loops mirrored and unit weights
    src, dst = graph.rmat_edges(8, 8, seed=5)
    for flags in itertools.product([0, 1], repeat=5):
        kw = dict(zip(("symmetrize", "transpose", "sorted", "remove_selfloops",
                       "remove_redundents"), map(bool, flags)))
        a = graph.read_gr(dataset("rmat8_raw.gr"), **kw)
        b = oracle.csr_from_edges(256, src, dst, None, mirror_selfloops=True, **kw)
        np.testing.assert_array_equal(a.row_offsets, b.row_offsets, err_msg=str(kw))
        np.testing.assert_array_equal(a.column_indices, b.column_indices, err_msg=str(kw))
        np.testing.assert_array_equal(a.weight, b.weight, err_msg=str(kw))


def test_dotgraph_short_file_and_stray_tokens(tmp_path):
    p = tmp_path / "short.graph"
    # vertex 2's line stops at the first non-integer token; vertices 4 and 5 have no line
    p.write_text("5 4\n2 3\n1 x 3\n1 2\n")
    c = graph.read_graph(str(p), sorted=False)
    np.testing.assert_array_equal(c.row_offsets, [0, 2, 3, 5, 5, 5])
    np.testing.assert_array_equal(c.column_indices, [1, 2, 0, 0, 1])
    np.testing.assert_array_equal(c.weight, np.ones(5, np.float32))


def test_readers_reject_what_the_reference_cannot_read(tmp_path):
    from gardenia_ocl_amd import GardeniaError

    # 4.gr numbers its vertices from 0 and 4w.graph interleaves weights with
    # neighbours: both give ids outside [1, m], which the reference indexes
    # out of bounds with; here they are errors.
    with pytest.raises(GardeniaError, match="outside"):
        graph.read_graph(dataset("4.gr"))
    with pytest.raises(GardeniaError, match="outside"):
        graph.read_graph(dataset("4w.graph"))
    bad = tmp_path / "bad.gr"
    bad.write_text("p sp 3 2\na 1 2\nc comment inside the arcs\n")
    with pytest.raises(GardeniaError, match="arc line"):
        graph.read_gr(str(bad))
    with pytest.raises(GardeniaError, match="Unrecognizable"):
        graph.read_graph(str(tmp_path / "x.edges"))
    with pytest.raises(GardeniaError, match="GD_ERR_IO"):
        graph.read_graph("/nonexistent/file.graph")
