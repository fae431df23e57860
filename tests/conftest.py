import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP path)")
    lib = os.path.join(ROOT, "oracle", "bin", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def box1():
    from skid_amd import gen
    return gen.config1()


def have_gpu() -> bool:
    try:
        import ctypes
        from skid_amd import skg
        P = skg.make_params(0.01, 32, period=1.0)
        h = ctypes.c_void_p()
        rc = skg.lib().skg_create(ctypes.byref(P), ctypes.byref(h))
        if rc == 0:
            skg.lib().skg_destroy(h)
        return rc == 0
    except Exception:
        return False
