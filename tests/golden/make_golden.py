"""Write the golden vectors of tests/golden/ from the reference compiled from source.

Run in a container where the reference is mounted at /root/reference:

    make -C oracle ref            # builds oracle/_ref/libref.so from the reference sources
    python tests/golden/make_golden.py

It calls the reference's own read_graph() (include/graph_io.h), PRSolver()
(src/pr/omp_base.cc), PRVerifier() (src/pr/verifier.cc), SpmvSolver()
(src/spmv/omp_base.cc) and SpmvVerifier() (src/spmv/verifier.cc) through
oracle/ref_harness.cc and records inputs and outputs:

  * reference datasets (datasets/*.mtx, copied to tests/golden/datasets/):
    CSR arrays under the PR driver's flags (src/pr/main.cc:23-24), the SpMV
    driver's flags (src/spmv/main.cc:24) and symmetrized; PR scores, printed
    iteration count and per-iteration L1; verifier verdict; SpMV y.
  * canonical RMAT-12 / RMAT-16 edge lists written as .mtx (from this
    package's generator): CSR hashes, PR scores (full at scale 12, hash and a
    sample at 16), iteration count, per-iteration L1, verifier verdict.
  * SpMV on an Erdos-Renyi 2^12 x 2^12 matrix (32 nnz/row) and on RMAT-12.

`python tests/golden/make_golden.py formats` records read_graph() on the
reference's .graph datasets (graph2csr) and on a 1-based .gr of raw RMAT-8
edges (gr2csr) under seven flag sets: tests/golden/ref_formats.npz / .json.

`python tests/golden/make_golden.py symgs` records the reference's
SymGSSolver (src/symgs/omp_base.cc, one symmetric sweep; three in a row) and
SymGSVerifier verdict on a 27-point stencil, a weighted RMAT-10 with and
without diagonals, 4w.mtx and chesapeake.mtx, read with the SymGS driver's
flags: tests/golden/ref_symgs.npz / .json.

The reference's compile-time constants apply: EPSILON 1e-3, MAX_ITER 30
(src/pr/pr.h:4-5).  Output: tests/golden/ref_golden.npz and ref_golden.json.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import re
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gardenia_ocl_amd import graph  # noqa: E402
import oracle  # noqa: E402  (only for the path of the _ref build)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DATASETS = ["test_pr.mtx", "4.mtx", "4w.mtx", "chesapeake.mtx"]

_P = ctypes.c_void_p


def load_ref():
    if not os.path.exists(oracle.REF_LIB_PATH):
        raise SystemExit("oracle/_ref/libref.so missing: run `make -C oracle ref` first")
    ref = ctypes.CDLL(oracle.REF_LIB_PATH)
    I, PI = ctypes.c_int, ctypes.POINTER(ctypes.c_int)
    PPI = ctypes.POINTER(ctypes.POINTER(ctypes.c_int))
    PPF = ctypes.POINTER(ctypes.POINTER(ctypes.c_float))
    ref.ref_read_graph.argtypes = [ctypes.c_char_p, I, I, I, I, I, PI, PI, PPI, PPI, PPF, PPI,
                                   ctypes.c_char_p]
    ref.ref_free.argtypes = [_P]
    ref.ref_pr_solve.argtypes = [I, I, _P, _P, _P, _P, _P, _P, ctypes.c_char_p]
    ref.ref_pr_verify.argtypes = [I, _P, _P, _P, _P, _P, _P, ctypes.c_double, ctypes.c_char_p]
    ref.ref_spmv_solve.argtypes = [I, I, _P, _P, _P, _P, _P, ctypes.c_char_p]
    ref.ref_spmv_verify.argtypes = [I, I, _P, _P, _P, _P, _P, _P, ctypes.c_char_p]
    ref.ref_max_iter.restype = I
    ref.ref_epsilon.restype = ctypes.c_double
    return ref


def p(a):
    return a.ctypes.data_as(_P)


def read_graph(ref, path, sym, tr, srt, rsl, rrd, log):
    m, nnz = ctypes.c_int(), ctypes.c_int()
    off = ctypes.POINTER(ctypes.c_int)()
    col = ctypes.POINTER(ctypes.c_int)()
    deg = ctypes.POINTER(ctypes.c_int)()
    w = ctypes.POINTER(ctypes.c_float)()
    ref.ref_read_graph(os.fsencode(path), sym, tr, srt, rsl, rrd, ctypes.byref(m),
                       ctypes.byref(nnz), ctypes.byref(off), ctypes.byref(col), ctypes.byref(w),
                       ctypes.byref(deg), os.fsencode(log))
    M, E = m.value, nnz.value
    out = dict(
        row_offsets=np.ctypeslib.as_array(off, (M + 1,)).copy(),
        column_indices=np.ctypeslib.as_array(col, (E,)).copy() if E else np.zeros(0, np.int32),
        weight=np.ctypeslib.as_array(w, (E,)).copy() if E else np.zeros(0, np.float32),
        degree=np.ctypeslib.as_array(deg, (M,)).copy(),
    )
    for q in (off, col, w, deg):
        ref.ref_free(ctypes.cast(q, _P))
    return M, out


def parse_pr_log(text):
    errs = [float(x) for _, x in re.findall(r"^\s*(\d+)\s+([0-9.]+)\s*$", text, re.M)]
    it = int(re.search(r"iterations = (\d+)\.", text).group(1))
    return it, errs


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_pr(ref, tmp, out_csr, in_csr, m):
    scores = np.full(m, np.float32(1.0) / np.float32(m), np.float32)
    log = os.path.join(tmp, "pr.log")
    ref.ref_pr_solve(m, int(in_csr["row_offsets"][-1]), p(in_csr["row_offsets"]),
                     p(in_csr["column_indices"]), p(out_csr["row_offsets"]),
                     p(out_csr["column_indices"]), p(out_csr["degree"]), p(scores),
                     os.fsencode(log))
    it, errs = parse_pr_log(open(log).read())
    vlog = os.path.join(tmp, "ver.log")
    ref.ref_pr_verify(m, p(in_csr["row_offsets"]), p(in_csr["column_indices"]),
                      p(out_csr["row_offsets"]), p(out_csr["column_indices"]),
                      p(out_csr["degree"]), p(scores), ref.ref_epsilon(), os.fsencode(vlog))
    vtext = open(vlog).read()
    verdict = "Correct" if "Correct" in vtext else re.search(r"Total Error: ([0-9.]+)", vtext).group(0)
    return scores, it, errs, verdict


def run_spmv(ref, tmp, m, csr, seed):
    nnz = int(csr["row_offsets"][-1])
    Ax = graph.uniform_pm1(nnz, seed)
    x = graph.uniform_pm1(m, seed + 1)
    y0 = graph.uniform_pm1(m, seed + 2)
    y = y0.copy()
    ref.ref_spmv_solve(m, nnz, p(csr["row_offsets"]), p(csr["column_indices"]), p(Ax), p(x), p(y),
                       os.fsencode(os.path.join(tmp, "spmv.log")))
    return Ax, x, y0, y


def main():
    ref = load_ref()
    arrays, meta = {}, {"max_iter": ref.ref_max_iter(), "epsilon": ref.ref_epsilon(),
                        "datasets": {}, "rmat": {}, "spmv": {}}
    with tempfile.TemporaryDirectory() as tmp:
        log = os.path.join(tmp, "read.log")
        for ds in DATASETS:
            path = os.path.join(GOLDEN, "datasets", ds)
            key = ds.replace(".mtx", "").replace(".", "_")
            md = {}
            variants = {
                "pr_out": (0, 0, 0, 1, 1), "pr_in": (0, 1, 0, 1, 1),
                "spmv": (0, 0, 1, 0, 1),
                "sym_out": (1, 0, 0, 1, 1), "sym_in": (1, 1, 0, 1, 1),
            }
            csrs = {}
            for vname, flags in variants.items():
                m, c = read_graph(ref, path, *flags, log)
                csrs[vname] = c
                for f, a in c.items():
                    arrays[f"{key}/{vname}/{f}"] = a
            md["m"] = m
            for mode, (o, i) in {"directed": ("pr_out", "pr_in"),
                                 "undirected": ("sym_out", "sym_in")}.items():
                s, it, errs, verdict = run_pr(ref, tmp, csrs[o], csrs[i], m)
                arrays[f"{key}/pr_{mode}/scores"] = s
                arrays[f"{key}/pr_{mode}/errors_printed"] = np.array(errs, np.float64)
                md[f"pr_{mode}"] = {"iterations_printed": it, "verdict": verdict}
            Ax, x, y0, y = run_spmv(ref, tmp, m, csrs["spmv"], seed=100)
            for n_, a in (("Ax", Ax), ("x", x), ("y0", y0), ("y", y)):
                arrays[f"{key}/spmv/{n_}"] = a
            meta["datasets"][ds] = md

        for scale in (12, 16):
            out_csr, in_csr = graph.rmat(scale, 16, seed=1)
            src = np.repeat(np.arange(out_csr.m, dtype=np.int32), out_csr.degree)
            path = os.path.join(tmp, f"rmat{scale}.mtx")
            graph.write_mtx(path, out_csr.m, src, out_csr.column_indices)
            m, ro = read_graph(ref, path, 0, 0, 0, 1, 1, log)
            _, ri = read_graph(ref, path, 0, 1, 0, 1, 1, log)
            s, it, errs, verdict = run_pr(ref, tmp, ro, ri, m)
            md = {
                "m": m, "nnz": int(ro["row_offsets"][-1]), "seed": 1, "edgefactor": 16,
                "iterations_printed": it, "verdict": verdict,
                "errors_printed": errs,
                "sha256": {f"out_{f}": sha(ro[f]) for f in ("row_offsets", "column_indices", "degree")}
                | {f"in_{f}": sha(ri[f]) for f in ("row_offsets", "column_indices", "degree")}
                | {"scores": sha(s)},
            }
            if scale == 12:
                arrays["rmat12/scores"] = s
            else:
                idx = np.arange(0, m, 97, dtype=np.int64)
                arrays["rmat16/sample_idx"] = idx
                arrays["rmat16/sample_scores"] = s[idx]
            meta["rmat"][str(scale)] = md
            if scale == 12:
                Ax, x, y0, y = run_spmv(ref, tmp, m, ri, seed=200)
                arrays["rmat12_in/spmv/y"] = y
                meta["spmv"]["rmat12_in"] = {"seed": 200, "sha256_y": sha(y)}

        er = graph.er_matrix(12, 32, seed=3)
        erc = {"row_offsets": er.row_offsets, "column_indices": er.column_indices}
        Ax, x, y0, y = run_spmv(ref, tmp, er.m, erc, seed=300)
        arrays["er12/spmv/y"] = y
        meta["spmv"]["er12"] = {"seed_matrix": 3, "seed_values": 300, "nnz_per_row": 32,
                                "sha256_row_offsets": sha(er.row_offsets),
                                "sha256_column_indices": sha(er.column_indices),
                                "sha256_y": sha(y)}

    np.savez_compressed(os.path.join(GOLDEN, "ref_golden.npz"), **arrays)
    with open(os.path.join(GOLDEN, "ref_golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", len(arrays), "arrays")


FORMAT_DATASETS = ["4.graph", "test1.graph", "test2.graph", "test3.graph", "test4.graph",
                   "test5.graph", "test6.graph", "rmat8_raw.gr"]
FORMAT_FLAGS = {
    "pr_out": (0, 0, 0, 1, 1), "pr_in": (0, 1, 0, 1, 1), "spmv": (0, 0, 1, 0, 1),
    "sym_out": (1, 0, 0, 1, 1), "sym_in": (1, 1, 0, 1, 1), "raw": (0, 0, 0, 0, 0),
    "sym_raw": (1, 0, 0, 0, 0),
}


def write_raw_gr(path, scale=8, edgefactor=8, seed=5):
    """A 1-based DIMACS .gr of the raw RMAT edges (self-loops and duplicates
    kept) from this package's generator, with a comment line and weights."""
    src, dst = graph.rmat_edges(scale, edgefactor, seed=seed)
    m = 1 << scale
    with open(path, "w") as f:
        f.write(f"c raw RMAT-{scale} edges, edgefactor {edgefactor}, seed {seed}\n")
        f.write(f"p sp {m} {len(src)}\n")
        for i, (a, b) in enumerate(zip(src.tolist(), dst.tolist())):
            f.write(f"a {a + 1} {b + 1} {(i * 37) % 100 + 1}\n")


def formats():
    """read_graph() on the .graph datasets (graph2csr) and a .gr file (gr2csr),
    every flag set above -> tests/golden/ref_formats.npz / .json."""
    ref = load_ref()
    gr = os.path.join(GOLDEN, "datasets", "rmat8_raw.gr")
    if not os.path.exists(gr):
        write_raw_gr(gr)
    arrays, meta = {}, {}
    with tempfile.TemporaryDirectory() as tmp:
        log = os.path.join(tmp, "read.log")
        for ds in FORMAT_DATASETS:
            path = os.path.join(GOLDEN, "datasets", ds)
            key = ds.replace(".", "_")
            for vname, flags in FORMAT_FLAGS.items():
                m, c = read_graph(ref, path, *flags, log)
                for f, a in c.items():
                    arrays[f"{key}/{vname}/{f}"] = a
            meta[ds] = {"m": m}
    np.savez_compressed(os.path.join(GOLDEN, "ref_formats.npz"), **arrays)
    with open(os.path.join(GOLDEN, "ref_formats.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", len(arrays), "arrays")


def write_weighted_mtx(path, m, rows, cols, vals):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{m} {m} {len(rows)}\n")
        for r, c, v in zip(rows.tolist(), cols.tolist(), vals.tolist()):
            f.write(f"{r + 1} {c + 1} {v:.9g}\n")


def stencil27(n):
    """graph.stencil27(n) as (m, rows, cols, vals) for the .mtx writer."""
    c = graph.stencil27(n)
    rows = np.repeat(np.arange(c.m, dtype=np.int64), c.degree)
    return c.m, rows, c.column_indices.astype(np.int64), c.weight


def rmat_with_diag(scale, seed):
    """Raw RMAT edges (self-loops kept) with weights, plus a diagonal entry on
    every even row: rows with and without a diagonal."""
    src, dst = graph.rmat_edges(scale, 8, seed=seed)
    m = 1 << scale
    w = np.abs(graph.uniform_pm1(len(src), seed + 7)) + np.float32(0.125)
    ev = np.arange(0, m, 2, dtype=np.int64)
    rows = np.concatenate([src.astype(np.int64), ev])
    cols = np.concatenate([dst.astype(np.int64), ev])
    vals = np.concatenate([w, np.full(len(ev), 24.0, np.float32)])
    return m, rows, cols, vals


def load_ref_symgs():
    path = os.path.join(os.path.dirname(oracle.REF_LIB_PATH), "libref_symgs.so")
    if not os.path.exists(path):
        raise SystemExit("oracle/_ref/libref_symgs.so missing: run `make -C oracle ref` first")
    ref = ctypes.CDLL(path)
    I = ctypes.c_int
    ref.ref_symgs_solve.argtypes = [I, I, _P, _P, _P, _P, _P, _P, _P, I, ctypes.c_char_p]
    ref.ref_symgs_verify.argtypes = [I, _P, _P, _P, _P, _P, _P, _P, _P, I, ctypes.c_char_p]
    return ref


SYMGS_FLAGS = (1, 0, 1, 0, 1)  # src/symgs/main.cc:18: symmetrize, sorted, self-loops kept


def symgs():
    """SymGSSolver (src/symgs/omp_base.cc) on matrices read with the SymGS
    driver's flags, colored by gd_color_first_fit and ordered as the driver
    orders them -> tests/golden/ref_symgs.npz / .json."""
    ref = load_ref()
    rs = load_ref_symgs()
    arrays, meta = {}, {}
    with tempfile.TemporaryDirectory() as tmp:
        log = os.path.join(tmp, "log")
        cases = {}
        for name, (m, r, c, v) in {"stencil6": stencil27(6),
                                   "rmat10_diag": rmat_with_diag(10, 11)}.items():
            path = os.path.join(tmp, name + ".mtx")
            write_weighted_mtx(path, m, r, c, v)
            cases[name] = path
        cases["4w"] = os.path.join(GOLDEN, "datasets", "4w.mtx")
        cases["chesapeake"] = os.path.join(GOLDEN, "datasets", "chesapeake.mtx")
        for seed, (name, path) in enumerate(cases.items()):
            m, c = read_graph(ref, path, *SYMGS_FLAGS, log)
            csr = graph.Csr(m, c["row_offsets"], c["column_indices"], c["degree"], c["weight"])
            colors, nc = graph.color_first_fit(csr)
            idx, co = graph.color_order(colors, nc)
            x0 = graph.uniform_pm1(m, 500 + seed)
            b = graph.uniform_pm1(m, 600 + seed)
            x = x0.copy()
            nnz = int(c["row_offsets"][-1])
            solve = lambda xx: rs.ref_symgs_solve(  # noqa: E731
                m, nnz, p(c["row_offsets"]), p(c["column_indices"]), p(idx), p(c["weight"]),
                p(xx), p(b), p(co), nc, os.fsencode(log))
            solve(x)
            x1 = x.copy()
            solve(x)
            solve(x)
            rs.ref_symgs_verify(m, p(c["row_offsets"]), p(c["column_indices"]), p(idx),
                                p(c["weight"]), p(x1), p(x0), p(b), p(co), nc, os.fsencode(log))
            verdict = "Correct" if "Correct" in open(log).read() else "POSSIBLE FAILURE"
            for f in ("row_offsets", "column_indices", "weight"):
                arrays[f"{name}/{f}"] = c[f]
            for k_, a in (("indices", idx), ("color_offsets", co), ("x0", x0), ("b", b),
                          ("x1", x1), ("x3", x)):
                arrays[f"{name}/{k_}"] = a
            meta[name] = {"m": m, "nnz": nnz, "n_colors": nc, "verdict_x1": verdict,
                          "rows_with_diag": int(sum(
                              np.any(c["column_indices"][c["row_offsets"][i]:c["row_offsets"][i + 1]] == i)
                              for i in range(m)))}
    np.savez_compressed(os.path.join(GOLDEN, "ref_symgs.npz"), **arrays)
    with open(os.path.join(GOLDEN, "ref_symgs.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", len(arrays), "arrays", meta)


if __name__ == "__main__":
    if sys.argv[1:] == ["formats"]:
        formats()
    elif sys.argv[1:] == ["symgs"]:
        symgs()
    else:
        main()
