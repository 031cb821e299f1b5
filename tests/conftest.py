import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def hip():
    """The product library; GPU tests fail loudly when it is not built."""
    import alsvinn_amd
    alsvinn_amd._lib.load()
    return alsvinn_amd
