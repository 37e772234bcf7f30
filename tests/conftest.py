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
