import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libucy.so)")


_GPU = None


def gpu_available() -> bool:
    """A context on device 0 through libucy itself (torch is not needed by
    the library, and a second HIP runtime in the process is avoided)."""
    global _GPU
    if _GPU is None:
        try:
            from unicycle_amd import _lib
            L = _lib.load()
            h = L.ucy_create(0, 0)
            _GPU = bool(h)
            if h:
                L.ucy_destroy(h)
        except Exception:
            _GPU = False
    return _GPU


@pytest.fixture(scope="session")
def engine_factory():
    if not gpu_available():
        pytest.skip("no GPU")
    from unicycle_amd import Engine
    engines = []

    def make(device=0):
        e = Engine(device)
        engines.append(e)
        return e
    yield make
    for e in engines:
        e.close()
