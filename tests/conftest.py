import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU fixture: a gpu-marked test that runs without a GPU fails loudly."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("gpu-marked test run without a visible GPU")
    from cpabdiffeo_amd import _lib
    _lib.lib()
    return torch
