import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden():
    """Loader of the committed golden fixtures (written by tests/golden/make_golden.py)."""
    import numpy as np

    def load(tag):
        return np.load(os.path.join(ROOT, "tests", "golden", f"ref_{tag}.npz"))
    return load


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch
