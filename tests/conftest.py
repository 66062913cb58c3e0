import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(ROOT, "tests", "golden", "genfft_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests exercise the HIP library only: fail loudly without it."""
    import torch

    import genfft_amd
    from genfft_amd import _capi

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a GPU")
    _capi.lib()  # raises if libgenfft_amd.so is missing
    return torch.device("cuda", 0)
