import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available() -> bool:
    try:
        from vehicleroutingproblem_amd import _native as N
        import ctypes as C
        lib = N.load()
        ctx = C.c_void_p()
        rc = lib.aco_hip_create(C.byref(ctx), 0)
        if rc == 0:
            lib.aco_hip_destroy(ctx)
            return True
    except Exception:
        return False
    return False


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when a gpu-marked test runs without a usable GPU:
    the GPU path has no fallback."""
    if not gpu_available():
        pytest.fail("no usable HIP device for a gpu-marked test")
    return True
