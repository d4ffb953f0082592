import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


_GPU = None


def gpu_available() -> bool:
    global _GPU
    if _GPU is None:
        try:
            import ctypes
            hip = ctypes.CDLL("libamdhip64.so")
            n = ctypes.c_int(0)
            _GPU = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
        except OSError:
            _GPU = False
    return _GPU


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle():
    from oracle import load_oracle
    return load_oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import Reference, reference_available
    if not reference_available():
        pytest.skip("oracle/_ref (the reference's own sources, compiled) is not built here")
    return Reference()
