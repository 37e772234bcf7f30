"""GPU parity of the HIP pull PageRank against the reference (golden vectors)
and the oracle.  All calls go through the C ABI (ctypes)."""
import numpy as np
import pytest

import oracle
from conftest import L1_GAP, assert_same_iterations, dataset
from gardenia_ocl_amd import PRSolver, PageRankPlan, graph

pytestmark = pytest.mark.gpu

DATASETS = ["test_pr.mtx", "4.mtx", "4w.mtx", "chesapeake.mtx"]


def exact_len(info):
    """Rows longer than one tile (256 threads x items_per_thread edges) are
    hub rows, summed in CSR order with the reference's FP32 sequential rounding."""
    return 256 * info["items_per_thread"]


def _key(ds):
    return ds.replace(".mtx", "").replace(".", "_")


def _gpu_pr(device, in_csr, outdeg, tol, max_iter, seed_vertex=-1, init=None, chunk=0,
            relabel=None, ipt=0, persistent=None, block=0, sort_tiles=None, hub_threshold=0,
            stage_hubs=-1):
    import torch

    m = in_csr.m
    off = torch.from_numpy(in_csr.row_offsets).to(device)
    col = torch.from_numpy(in_csr.column_indices).to(device)
    deg = torch.from_numpy(np.ascontiguousarray(outdeg, np.int32)).to(device)
    if init is None:
        init = np.full(m, np.float32(1.0) / np.float32(m), np.float32)
    scores = torch.from_numpy(np.array(init, np.float32)).to(device)
    plan = PageRankPlan(off, col, deg, relabel=relabel, items_per_thread=ipt,
                        persistent=persistent, block=block, sort_tiles=sort_tiles,
                        hub_threshold=hub_threshold, stage_hubs=stage_hubs)
    stats, errs = plan.run(scores, tol=tol, max_iter=max_iter, seed_vertex=seed_vertex,
                           chunk=chunk)
    out = scores.cpu().numpy()
    info = plan.info()
    plan.close()
    return out, stats, errs, info


@pytest.mark.parametrize("ds", DATASETS)
@pytest.mark.parametrize("directed", [True, False])
def test_pr_datasets_bit_exact_vs_reference(golden, device, ds, directed):
    arrays, meta = golden
    out_csr, in_csr = graph.read_pr_graph(dataset(ds), is_directed=directed)
    mode = "directed" if directed else "undirected"
    s, stats, errs, _ = _gpu_pr(device, in_csr, out_csr.degree, meta["epsilon"], meta["max_iter"])
    ref = arrays[f"{_key(ds)}/pr_{mode}/scores"]
    # every row is summed in CSR order with sequential FP32 rounding (the
    # symmetrized 4.mtx keeps duplicate edges, so some rows exceed 64):
    # bit-identical to the reference
    np.testing.assert_array_equal(s, ref)
    assert stats["iterations"] == meta["datasets"][ds][f"pr_{mode}"]["iterations_printed"]
    # printed with %lf (6 decimals)
    np.testing.assert_allclose(errs, arrays[f"{_key(ds)}/pr_{mode}/errors_printed"], atol=5.1e-7)


def test_pr_host_solver_interface(golden):
    """PRSolver(m, nnz, in_off, in_col, out_off, out_col, degree, scores) -- host arrays."""
    arrays, meta = golden
    out_csr, in_csr = graph.read_pr_graph(dataset("test_pr.mtx"))
    m = in_csr.m
    scores = np.full(m, np.float32(1.0) / np.float32(m), np.float32)
    stats, errs = PRSolver(m, in_csr.nnz, in_csr.row_offsets, in_csr.column_indices,
                           out_csr.row_offsets, out_csr.column_indices, out_csr.degree, scores)
    np.testing.assert_array_equal(scores, arrays["test_pr/pr_directed/scores"])
    assert stats["iterations"] == 11


@pytest.mark.parametrize("scale", [12, 16])
def test_pr_rmat_vs_reference(golden, device, scale):
    arrays, meta = golden
    md = meta["rmat"][str(scale)]
    out_csr, in_csr = graph.rmat(scale, 16, seed=md["seed"])
    s, stats, errs, info = _gpu_pr(device, in_csr, out_csr.degree, meta["epsilon"],
                                   meta["max_iter"])
    ref, it, _ = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                           tol=meta["epsilon"], max_iter=meta["max_iter"])
    assert stats["iterations"] == md["iterations_printed"] == it
    np.testing.assert_allclose(errs, md["errors_printed"], atol=5.1e-7)
    np.testing.assert_array_equal(s, ref)
    if scale == 12:
        np.testing.assert_allclose(s, arrays["rmat12/scores"], rtol=1e-5, atol=0)
    # rows longer than a tile are whole-row hubs summed in the reference's order
    assert info["hubs"] == int(np.sum(in_csr.degree > exact_len(info)))


@pytest.mark.parametrize("relabel,ipt", [(False, 8), (True, 16), (False, 16)])
def test_pr_first_iteration_short_and_hub_rows_bit_exact(device, relabel, ipt):
    """After one iteration every row of <= 64 in-edges (owner-thread sum) and
    every row longer than a tile (whole-row sequential-order sum) is
    bit-identical to the reference arithmetic, and so is every row in between
    (summed by its wavefront with the sequential order's exact rounding)."""
    out_csr, in_csr = graph.rmat(16, 16, seed=1)
    s, stats, _, info = _gpu_pr(device, in_csr, out_csr.degree, 0.0, 1, relabel=relabel, ipt=ipt)
    ref, it, _ = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                           tol=0.0, max_iter=1)
    middle = (in_csr.degree > 64) & (in_csr.degree <= exact_len(info))
    assert info["hubs"] == int(np.sum(in_csr.degree > exact_len(info))) > 0 and middle.any()
    np.testing.assert_array_equal(s, ref)


@pytest.mark.parametrize("scale,tol,relabel,ipt", [
    (16, 1e-4, False, 8), (16, 1e-4, True, 16), (18, 1e-4, True, 8), (18, 1e-6, False, 16),
    (20, 1e-4, True, 8), (20, 1e-4, False, 16)])
def test_pr_rmat_vs_oracle(device, scale, tol, relabel, ipt):
    out_csr, in_csr = graph.rmat(scale, 16, seed=1)
    s, stats, errs, info = _gpu_pr(device, in_csr, out_csr.degree, tol, 100, relabel=relabel,
                                   ipt=ipt)
    assert info["relabel"] == int(relabel) and info["items_per_thread"] == ipt
    ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                  out_csr.degree, tol=tol, max_iter=100)
    tie, gap = assert_same_iterations(stats["iterations"], errs, it, ref_errs, tol)
    print(f"RMAT-{scale} tol {tol}: {stats['iterations']} iterations (oracle {it}), "
          f"max L1 gap {gap:.2e}, tie {tie}")
    np.testing.assert_array_equal(s, ref)
    # L1 error against the FP64 fixed point of the same operator
    x64, _ = oracle.pr_fp64(in_csr.m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree)
    l1_gpu = np.abs(s.astype(np.float64) - x64).sum()
    l1_cpu = np.abs(ref.astype(np.float64) - x64).sum()
    assert l1_gpu <= 1.05 * l1_cpu + 1e-7
    assert info["hubs"] > 0


def test_pr_deterministic_and_chunk_invariant(device):
    out_csr, in_csr = graph.rmat(16, 16, seed=5)
    a, sa, ea, _ = _gpu_pr(device, in_csr, out_csr.degree, 1e-4, 100, chunk=1)
    b, sb, eb, _ = _gpu_pr(device, in_csr, out_csr.degree, 1e-4, 100, chunk=7)
    c, sc, ec, _ = _gpu_pr(device, in_csr, out_csr.degree, 1e-4, 100)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(ea, eb)
    assert sa["iterations"] == sb["iterations"] == sc["iterations"]


@pytest.mark.parametrize("relabel", [False, True])
def test_pr_graph_replays_match_direct_launches(device, relabel):
    """Chunks of iterations are replayed from captured hipGraphs (two chunk
    lengths here: 8 and a tail of 5); the profiled run launches directly.
    Every replay, with eager work and synchronisations between, gives the
    direct launches' scores, L1 values and iteration count, bit for bit."""
    import torch

    out_csr, in_csr = graph.rmat(14, 16, seed=7)
    m = in_csr.m
    off = torch.from_numpy(in_csr.row_offsets).to(device)
    col = torch.from_numpy(in_csr.column_indices).to(device)
    deg = torch.from_numpy(out_csr.degree).to(device)
    plan = PageRankPlan(off, col, deg, relabel=relabel)
    init = float(np.float32(1.0) / np.float32(m))
    results = []
    for rep in range(4):
        s = torch.full((m,), init, dtype=torch.float32, device=device)
        junk = torch.randn(1 << 16, device=device).sum()  # eager work between replays
        torch.cuda.synchronize()
        st, errs = plan.run(s, tol=0.0, max_iter=21, chunk=8, profile=(rep == 0))
        results.append((s.cpu().numpy(), errs, st["iterations"]))
        assert float(junk) == float(junk)
    for r in results[1:]:
        np.testing.assert_array_equal(r[0], results[0][0])
        np.testing.assert_array_equal(r[1], results[0][1])
        assert r[2] == results[0][2] == 21
    # converging runs (tol > 0): the replayed chunks after convergence return at once
    for rep in range(3):
        s = torch.full((m,), init, dtype=torch.float32, device=device)
        st, errs = plan.run(s, tol=1e-4, max_iter=100, profile=(rep == 0))
        if rep == 0:
            ref = (s.cpu().numpy(), errs, st["iterations"])
        else:
            np.testing.assert_array_equal(s.cpu().numpy(), ref[0])
            np.testing.assert_array_equal(errs, ref[1])
            assert st["iterations"] == ref[2]
    plan.close()


def test_pr_fixed_iterations_not_converged(device):
    out_csr, in_csr = graph.rmat(14, 16, seed=2)
    s, stats, errs, _ = _gpu_pr(device, in_csr, out_csr.degree, 0.0, 20)
    ref, it, _ = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                           tol=0.0, max_iter=20)
    assert stats["iterations"] == it == 20 and not stats["converged"]
    assert np.max(np.abs(s - ref) / ref) <= 1e-5


def test_pr_max_iter_zero_leaves_scores(device):
    out_csr, in_csr = graph.rmat(10, 8, seed=3)
    init = np.random.default_rng(0).random(in_csr.m).astype(np.float32)
    s, stats, errs, _ = _gpu_pr(device, in_csr, out_csr.degree, 1e-3, 0, init=init)
    assert stats["iterations"] == 0
    np.testing.assert_array_equal(s, init)


def _star(n_leaves, hub=0):
    """Every leaf points at the hub and the hub points back at every leaf."""
    m = n_leaves + 1
    leaves = np.arange(1, m, dtype=np.int32)
    src = np.concatenate([leaves, np.full(n_leaves, hub, np.int32)])
    dst = np.concatenate([np.full(n_leaves, hub, np.int32), leaves])
    out_csr = graph.csr_from_edges(m, src, dst, sorted=True)
    in_csr = graph.csr_from_edges(m, src, dst, transpose=True, sorted=True)
    return out_csr, in_csr


@pytest.mark.parametrize("n_leaves", [1, 63, 64, 65, 2047, 2048, 2049, 4096, 4097, 100_000,
                                      131_072, 131_073, 140_000])
def test_pr_star_hub_boundaries(device, n_leaves):
    """Hub of n in-edges, all addends equal: the reference's sequential FP32 sum
    drifts (its L1 change plateaus at ~3e-6 for n=65, ~7e-5 for n=2048 and the
    hub score is 3.5e-5 off the FP64 fixed point).  Every hub row is summed in
    the reference's order and rounding -- by its owner thread up to 64
    in-edges, by its wavefront's exact scan up to a tile, by one lane's chain
    up to 2^17 = 131,072 in-edges and by the workgroup's binade scan beyond --
    so the scores, L1 values and iteration count are the reference's, drift
    included."""
    out_csr, in_csr = _star(n_leaves)
    s, stats, errs, info = _gpu_pr(device, in_csr, out_csr.degree, 1e-6, 100)
    ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                  out_csr.degree, tol=1e-6, max_iter=100)
    np.testing.assert_array_equal(s, ref)
    assert_same_iterations(stats["iterations"], errs, it, ref_errs, 1e-6)
    assert stats["iterations"] == it
    assert info["hubs"] == (1 if n_leaves > exact_len(info) else 0)


@pytest.mark.parametrize("mix", ["ties", "ties_and_crossings"])
@pytest.mark.parametrize("n", [20_000, 140_000])
def test_pr_hub_row_of_ties_bit_exact(device, mix, n):
    """A hub of n in-edges whose addends are mostly exact ties (odd multiples
    of half an ulp of the running sum): each rounds to even, so the sum depends
    on every step's parity.  20,000 in-edges take the one-lane chain, 140,000
    (more than 2^17) the binade scan, which resolves the ties as maps; the
    first iteration must equal the reference's sequential sum either way."""
    rng = np.random.default_rng((11 if mix == "ties" else 12) + n)
    m = n + 1
    src = np.arange(1, m, dtype=np.int32)
    dst = np.zeros(n, np.int32)
    out_csr = graph.csr_from_edges(m, src, dst)
    in_csr = graph.csr_from_edges(m, src, dst, transpose=True)
    u = np.float32(2.0 ** -23)  # ulp of [1, 2)
    init = np.zeros(m, np.float32)
    init[0] = 1.0
    init[1] = 1.0  # the first addend puts the sum in [1, 2)
    k = rng.integers(0, 4, n - 1).astype(np.float32)
    init[2:] = (k + np.float32(0.5)) * u
    if mix == "ties_and_crossings":
        big = rng.random(n - 1) < 0.01  # push the sum across binades now and then
        init[2:][big] = np.float32(2.0 ** -6)
        regular = rng.random(n - 1) < 0.2
        init[2:][regular] = (k[regular] + np.float32(0.25)) * u
    s, stats, _, info = _gpu_pr(device, in_csr, out_csr.degree, 0.0, 1, init=init)
    ref, it, _ = oracle.pr(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                           init, tol=0.0, max_iter=1)
    assert info["hubs"] == 1
    np.testing.assert_array_equal(s, ref)


def test_pr_empty_and_edgeless(device):
    import torch

    # m = 0: nothing to do; one (empty) iteration as in the reference loop
    z = torch.zeros(1, dtype=torch.int32, device=device)
    e = torch.zeros(0, dtype=torch.int32, device=device)
    plan = PageRankPlan(z, e, e)
    stats, errs = plan.run(torch.zeros(0, dtype=torch.float32, device=device))
    assert stats["iterations"] == 1 and stats["converged"]
    # no edges: every score becomes (1-d)/m after one iteration
    m = 1000
    in_csr = graph.csr_from_edges(m, np.zeros(0, np.int32), np.zeros(0, np.int32))
    s, stats, errs, _ = _gpu_pr(device, in_csr, np.zeros(m, np.int32), 1e-3, 30)
    ref, it, _ = oracle.pr(m, in_csr.row_offsets, in_csr.column_indices, np.zeros(m, np.int32))
    np.testing.assert_array_equal(s, ref)
    assert stats["iterations"] == it


def test_pr_personalised_seed_matches_oracle(device):
    out_csr, in_csr = graph.rmat(14, 16, seed=4)
    m = in_csr.m
    seed = int(np.argmax(out_csr.degree))
    init = np.zeros(m, np.float32)
    init[seed] = 1.0
    s, stats, errs, _ = _gpu_pr(device, in_csr, out_csr.degree, 1e-7, 200, seed_vertex=seed,
                                init=init)
    ref, it, ref_errs = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                                   seed, tol=1e-7, max_iter=200)
    assert_same_iterations(stats["iterations"], errs, it, ref_errs, 1e-7)
    np.testing.assert_array_equal(s, ref)


def test_pr_invalid_csr_rejected(device):
    import torch

    from gardenia_ocl_amd import GardeniaError

    off = torch.tensor([0, 2, 3], dtype=torch.int32, device=device)
    col = torch.tensor([1, 5, 0], dtype=torch.int32, device=device)  # 5 >= m
    deg = torch.tensor([1, 1], dtype=torch.int32, device=device)
    with pytest.raises(GardeniaError, match="column index out of range"):
        PageRankPlan(off, col, deg)
    off_bad = torch.tensor([0, 3, 2], dtype=torch.int32, device=device)
    with pytest.raises(GardeniaError):
        PageRankPlan(off_bad, torch.tensor([0, 1], dtype=torch.int32, device=device), deg)


@pytest.mark.parametrize("ipt", [8, 16])
def test_pr_relabel_is_bit_identical(device, ipt):
    """Relabelling keeps every row's edges in order and the reductions depend
    only on positions within a row, so scores are bit-identical."""
    out_csr, in_csr = graph.rmat(17, 16, seed=6)
    a, sa, ea, ia = _gpu_pr(device, in_csr, out_csr.degree, 1e-5, 60, relabel=False, ipt=ipt)
    b, sb, eb, ib = _gpu_pr(device, in_csr, out_csr.degree, 1e-5, 60, relabel=True, ipt=ipt)
    assert ia["relabel"] == 0 and ib["relabel"] == 1
    np.testing.assert_array_equal(a, b)
    assert sa["iterations"] == sb["iterations"]
    np.testing.assert_allclose(ea, eb, rtol=1e-12)


def test_pr_relabel_personalised_seed(device):
    out_csr, in_csr = graph.rmat(14, 16, seed=8)
    m = in_csr.m
    seed = 12345
    init = np.zeros(m, np.float32)
    init[seed] = 1.0
    a, sa, _, _ = _gpu_pr(device, in_csr, out_csr.degree, 1e-7, 200, seed_vertex=seed, init=init,
                          relabel=False)
    b, sb, _, _ = _gpu_pr(device, in_csr, out_csr.degree, 1e-7, 200, seed_vertex=seed, init=init,
                          relabel=True)
    np.testing.assert_array_equal(a, b)


def test_pr_kernel_variants_bit_identical(device):
    """Persistent/pipelined vs one-workgroup-per-tile, every cache-hint level
    and gathering in source order (sorted tiles) run the same arithmetic in the
    same order, for each tile size (RMAT-16 has rows longer than both tile
    sizes, so the exact hub kernels run too)."""
    import torch

    out_csr, in_csr = graph.rmat(16, 16, seed=11)
    m = in_csr.m
    off = torch.from_numpy(in_csr.row_offsets).to(device)
    col = torch.from_numpy(in_csr.column_indices).to(device)
    deg = torch.from_numpy(out_csr.degree).to(device)
    results = []
    for persistent in (False, True):
        for nt in (0, 1, 2):
            for ipt in (8, 16):
                for srt in (False, True):
                    if srt and nt == 2:
                        continue
                    plan = PageRankPlan(off, col, deg, relabel=True, items_per_thread=ipt,
                                        nontemporal=nt, persistent=persistent, sort_tiles=srt)
                    inf = plan.info()
                    assert inf["persistent"] == int(persistent) and inf["nontemporal"] == nt
                    assert inf["sort_tiles"] == int(srt)
                    s = torch.full((m,), 1.0 / m, dtype=torch.float32, device=device)
                    st, errs = plan.run(s, tol=1e-5, max_iter=50)
                    results.append((ipt, s.cpu().numpy(), st["iterations"]))
                    plan.close()
    for ipt in (8, 16):
        rs = [r for r in results if r[0] == ipt]
        for r in rs[1:]:
            np.testing.assert_array_equal(r[1], rs[0][1])
            assert r[2] == rs[0][2]


@pytest.mark.parametrize("relabel", [False, True])
def test_pr_sorted_tiles_bit_identical(device, relabel):
    """Tiles gathered in source order give the CSR-order results bit for bit,
    with and without the relabelled private copy (sorted in place)."""
    out_csr, in_csr = graph.rmat(15, 16, seed=12)
    a, sa, ea, ia = _gpu_pr(device, in_csr, out_csr.degree, 1e-6, 80, relabel=relabel)
    import torch

    off = torch.from_numpy(in_csr.row_offsets).to(device)
    col = torch.from_numpy(in_csr.column_indices).to(device)
    col_before = col.clone()
    deg = torch.from_numpy(out_csr.degree).to(device)
    for srt in (False, True):
        plan = PageRankPlan(off, col, deg, relabel=relabel, sort_tiles=srt)
        s = torch.full((in_csr.m,), float(np.float32(1.0) / np.float32(in_csr.m)),
                       dtype=torch.float32, device=device)
        st, errs = plan.run(s, tol=1e-6, max_iter=80)
        np.testing.assert_array_equal(s.cpu().numpy(), a)
        assert st["iterations"] == sa["iterations"]
        plan.close()
    assert torch.equal(col, col_before)  # the caller's column array is never reordered


def test_pr_persistent_tiny_and_empty_tail(device):
    """Graphs whose last tiles have no edges (the prefetch clamps to the last
    valid edge) and graphs with no edges at all."""
    import torch

    m = 5000
    src = np.arange(0, 100, dtype=np.int32)
    dst = (src + 1).astype(np.int32)
    out_csr = graph.csr_from_edges(m, src, dst)
    in_csr = graph.csr_from_edges(m, src, dst, transpose=True)
    ref, it, _ = oracle.pr(m, in_csr.row_offsets, in_csr.column_indices, out_csr.degree,
                           tol=1e-6, max_iter=100)
    for persistent in (True, False):
        s, stats, errs, info = _gpu_pr(device, in_csr, out_csr.degree, 1e-6, 100,
                                       persistent=persistent)
        assert info["persistent"] == int(persistent)
        np.testing.assert_array_equal(s, ref)
        assert stats["iterations"] == it
    empty = graph.csr_from_edges(m, np.zeros(0, np.int32), np.zeros(0, np.int32))
    for persistent in (True, False):
        s, stats, errs, _ = _gpu_pr(device, empty, np.zeros(m, np.int32), 1e-3, 30,
                                    persistent=persistent)
        base = (np.float32(1.0) - np.float32(0.85)) / np.float32(m)  # (1.0f - kDamp) / m
        np.testing.assert_array_equal(s, np.full(m, base, np.float32))


@pytest.mark.parametrize("hub_threshold,ipt", [(64, 16), (1024, 16), (2048, 8)])
def test_pr_hub_threshold_bit_exact(device, hub_threshold, ipt):
    """Any cut between row-tile rows and whole-row exact hubs gives the
    oracle's scores and iteration count (rows above the cut are summed by the
    workgroup's binade scan, below it by their owner thread)."""
    out_csr, in_csr = graph.rmat(16, 16, seed=5)
    s, stats, errs, info = _gpu_pr(device, in_csr, out_csr.degree, 1e-6, 100, ipt=ipt,
                                   hub_threshold=hub_threshold)
    assert info["hub_threshold"] == hub_threshold
    deg = np.diff(in_csr.row_offsets)
    assert info["tiles"] > 0 and (deg > hub_threshold).sum() > 0
    ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                  out_csr.degree, tol=1e-6, max_iter=100)
    np.testing.assert_array_equal(s, ref)
    assert_same_iterations(stats["iterations"], errs, it, ref_errs, 1e-6)


@pytest.mark.parametrize("hub_threshold,seed,launch", [
    (0, -1, {}), (1024, -1, {}), (0, "hub", {}), (0, -1, {"persistent": True}),
    (0, -1, {"block": 512})])
def test_pr_staged_hubs_bit_exact(device, hub_threshold, seed, launch):
    """Hub rows gathered ahead by the staging kernel (stage_hubs=1) give the
    oracle's scores and iteration count, as the gathering form does: RMAT-18
    (over 512 workgroups per iteration, so the staged launch runs), the
    default cut, a 1024-edge cut (hubs of one scan chunk staged too), a
    personalised start on a hub row's source, and the persistent-grid and
    512-thread row-tile launches beside the staged hubs."""
    out_csr, in_csr = graph.rmat(18, 16, seed=9)
    m = in_csr.m
    sv, init = -1, None
    if seed == "hub":
        deg = np.diff(in_csr.row_offsets)
        r = int(np.argmax(deg))
        sv = int(in_csr.column_indices[in_csr.row_offsets[r]])
        init = np.zeros(m, np.float32)
        init[sv] = 1.0
    s, stats, errs, info = _gpu_pr(device, in_csr, out_csr.degree, 1e-6, 100, seed_vertex=sv,
                                   init=init, hub_threshold=hub_threshold, stage_hubs=1, **launch)
    assert info["staged_hubs"] > 0 and info["staged_hubs"] == info["hubs"]
    if sv < 0:
        ref, it, ref_errs = oracle.pr(m, in_csr.row_offsets, in_csr.column_indices,
                                      out_csr.degree, tol=1e-6, max_iter=100)
    else:
        ref, it, ref_errs = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices,
                                       out_csr.degree, sv, tol=1e-6, max_iter=100)
    np.testing.assert_array_equal(s, ref)
    assert_same_iterations(stats["iterations"], errs, it, ref_errs, 1e-6)
    s0, st0, _, info0 = _gpu_pr(device, in_csr, out_csr.degree, 1e-6, 100, seed_vertex=sv,
                                init=init, hub_threshold=hub_threshold, stage_hubs=0, **launch)
    assert info0["staged_hubs"] == 0
    np.testing.assert_array_equal(s0, s)
    assert st0["iterations"] == stats["iterations"]


def test_pr_staging_forced_on_small_graph_runs_unstaged(device):
    """stage_hubs=1 on a graph small enough for the one-launch iterations:
    the plan keeps no staging (the fused launch gathers its hubs itself) and
    the scores and iteration count are the oracle's."""
    out_csr, in_csr = graph.rmat(14, 16, seed=3)
    s, stats, errs, info = _gpu_pr(device, in_csr, out_csr.degree, 1e-6, 100, stage_hubs=1,
                                   hub_threshold=512)
    assert info["hubs"] > 0 and info["staged_hubs"] == 0
    ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                  out_csr.degree, tol=1e-6, max_iter=100)
    np.testing.assert_array_equal(s, ref)
    assert_same_iterations(stats["iterations"], errs, it, ref_errs, 1e-6)


def test_pr_personalized_exact_hubs_power_of_two_sums(device):
    """One-hot personalised starts on a graph where every row above 64
    in-edges is an exact hub: a hub row gathering the seed holds the sum
    1/outdeg(seed) -- an exact power of two for outdeg 1, 2, 4, ... -- across
    long runs of zero terms.  Scores and iteration counts are the oracle's
    scalar personalised loop's, bit for bit."""
    import torch

    out_csr, in_csr = graph.rmat(16, 16, seed=7)
    m = in_csr.m
    od = out_csr.degree
    deg = np.diff(in_csr.row_offsets)
    hubs = np.nonzero(deg > 64)[0]
    picks = []
    for want in (1, 2, 4, 8, 3):  # powers of two, and one that is not
        for h in hubs:
            srcs = in_csr.column_indices[in_csr.row_offsets[h]:in_csr.row_offsets[h + 1]]
            cand = srcs[od[srcs] == want]
            if cand.size:
                picks.append(int(cand[0]))
                break
    assert len(picks) >= 4
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(device)  # noqa: E731
    plan = PageRankPlan(t(in_csr.row_offsets), t(in_csr.column_indices), t(od), hub_threshold=64)
    assert plan.info()["hub_threshold"] == 64
    for seed in picks:
        s = torch.zeros(m, dtype=torch.float32, device=device)
        s[seed] = 1.0
        st, errs = plan.run(s, tol=1e-6, max_iter=100, seed_vertex=seed)
        ref, it, ref_errs = oracle.ppr(m, in_csr.row_offsets, in_csr.column_indices, od, seed,
                                       tol=1e-6, max_iter=100)
        np.testing.assert_array_equal(s.cpu().numpy(), ref)
        assert st["iterations"] == it
    plan.close()


@pytest.mark.parametrize("block", [512, 1024])
@pytest.mark.parametrize("scale,relabel,ipt,sort", [(16, True, 16, True), (18, True, 16, True),
                                                    (18, False, 8, False), (20, True, 16, True)])
def test_pr_wide_tiles_bit_exact(device, block, scale, relabel, ipt, sort):
    """Row tiles on 512- and 1024-thread workgroups (tiles of block*ipt edges
    and 4*block rows, sorted by source over the whole tile): every score and L1
    value is the oracle's, as with 256-thread tiles."""
    out_csr, in_csr = graph.rmat(scale, 16, seed=3)
    s, stats, errs, info = _gpu_pr(device, in_csr, out_csr.degree, 1e-5, 100, relabel=relabel,
                                   ipt=ipt, block=block, sort_tiles=sort)
    assert info["block"] == block and info["sort_tiles"] == int(sort)
    ref, it, ref_errs = oracle.pr(in_csr.m, in_csr.row_offsets, in_csr.column_indices,
                                  out_csr.degree, tol=1e-5, max_iter=100)
    np.testing.assert_array_equal(s, ref)
    assert_same_iterations(stats["iterations"], errs, it, ref_errs, 1e-5)
