import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on cuda:0; fails loudly if either is missing."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    import gallatin_amd
    gallatin_amd.lib()
    torch.cuda.init()
    return gallatin_amd
