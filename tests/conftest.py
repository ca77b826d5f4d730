import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def bpx():
    import backprojector_amd as b
    return b


@pytest.fixture(scope="session")
def gpu(bpx):
    """The product library with a visible gfx950 device (GPU tests fail loudly otherwise)."""
    n = bpx.device_count()
    assert n >= 1, "no gfx950 device visible: GPU tests need the HIP path"
    return bpx
