import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libamx.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import fv
    fv.build()
    return fv


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a visible GPU")
    from astronomix_amd import _native
    _native.load_library()
    return torch.device("cuda:0")
