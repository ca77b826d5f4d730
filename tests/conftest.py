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


@pytest.fixture
def hip_streams(gpu):
    """Two HIP streams of the caller (default priority), created through the HIP runtime that
    libbpx is linked against (already loaded, so the same runtime instance)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    streams = []
    for _ in range(2):
        s = ctypes.c_void_p()
        assert hip.hipStreamCreate(ctypes.byref(s)) == 0
        streams.append(s.value)
    yield streams
    for s in streams:
        assert hip.hipStreamSynchronize(s) == 0
        assert hip.hipStreamDestroy(s) == 0
