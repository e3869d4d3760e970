import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libxadv.so")
    config.addinivalue_line("markers", "slow: long CPU runs of the oracle")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but torch.cuda.is_available() is False")
    from xadvcool_amd import _lib
    _lib.load()  # raises if libxadv.so is missing: GPU tests never fall back
    return torch.device("cuda:0")
