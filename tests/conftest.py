import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_available():
    from particlesolver_amd import capi
    lib = capi.load()
    import ctypes
    p = capi.default_params()
    h = ctypes.c_void_p()
    rc = lib.ps_create(ctypes.byref(p), ctypes.byref(h))
    if rc != 0:
        pytest.fail("GPU test selected but ps_create failed: %s"
                    % lib.ps_last_error(None).decode())
    lib.ps_destroy(h)
    return True
