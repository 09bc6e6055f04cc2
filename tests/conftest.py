import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libexactclosure kernels)")


@pytest.fixture(scope="session")
def O():
    import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def ec():
    import __graft_entry__ as ge

    return ge.load_package()


@pytest.fixture(scope="session")
def gpu(ec):
    """The package with a live context; fails (not skips) when the native
    library or the GPU is missing, so a GPU run never passes on a fallback."""
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    ec._abi.context(0)
    return ec
