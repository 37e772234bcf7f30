import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
DATASET_DIR = os.path.join(GOLDEN_DIR, "datasets")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


@pytest.fixture(scope="session")
def golden():
    arrays = np.load(os.path.join(GOLDEN_DIR, "ref_golden.npz"), allow_pickle=False)
    with open(os.path.join(GOLDEN_DIR, "ref_golden.json")) as f:
        meta = json.load(f)
    return arrays, meta


@pytest.fixture(scope="session")
def device():
    """torch cuda:0 for GPU tests; fails loudly if the HIP path cannot run."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    from gardenia_ocl_amd import _native

    assert _native.lib().gd_device_count() >= 1
    return torch.device("cuda:0")


def dataset(name):
    return os.path.join(DATASET_DIR, name)


# Largest difference, in one iteration, between the L1 change of the HIP plan
# and of the oracle (the reference's arithmetic) that the parity tests allow.
# Every row of the PageRank plan is summed in the reference's order with its
# FP32 rounding, so the scores are bit-identical and the L1 values, sums of
# the same |new - old| in FP64, differ only by the order of that FP64 sum.
L1_GAP = 1e-12


def assert_same_iterations(iterations, errs, ref_iterations, ref_errs, tol, gap=L1_GAP):
    """Iteration-count parity with the reference (src/pr/omp_base.cc:38-43).

    Every common iteration's L1 change must agree with the oracle's within
    `gap`.  Then the GPU takes the oracle's stop decision at every iteration
    unless one of the oracle's own L1 values before its stop lies within that
    measured difference of tol; only then is a different count allowed.
    Returns (tie, measured difference)."""
    errs = np.asarray(errs, np.float64)
    ref_errs = np.asarray(ref_errs, np.float64)
    n = min(len(errs), len(ref_errs))
    d = float(np.max(np.abs(errs[:n] - ref_errs[:n]))) if n else 0.0
    assert d <= gap, f"per-iteration L1 differs from the oracle's by {d:.3e} > {gap:.1e}"
    tie = bool(tol > 0 and np.any(np.abs(ref_errs[:ref_iterations] - tol) <= d))
    if not tie:
        assert iterations == ref_iterations, (iterations, ref_iterations, d)
    return tie, d
