import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Session guard for GPU tests: the HIP library must load and a GPU must
    be visible.  Missing native code is an error, never a skip."""
    if not _gpu_available():
        pytest.fail("gpu test selected but no GPU is visible")
    from natrix_amd import _lib
    return _lib.lib()
