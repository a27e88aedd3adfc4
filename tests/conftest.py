"""Test configuration: the `gpu` marker selects tests that need the MI355X (run with -m gpu)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and the built libspecann.so")


@pytest.fixture(scope="session")
def gpu():
    """The HIP device; GPU tests fail (not skip) when the library cannot be loaded on a GPU host."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from specann_amd import _lib
    _lib.load()
    return torch.device("cuda", 0)
