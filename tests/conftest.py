import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmrfg.so")


@pytest.fixture(scope="session")
def gpu():
    """The product library on a GPU; a missing GPU or library is an error,
    never a silent CPU fallback (creating a context fails without a device)."""
    import multiregionfoam_amd as m
    m.lib()
    m.CoupledFvMatrix(0).close()
    return m
