"""The C ABI (include/skg.h): the library loads, exports every declared entry
point, validates parameters and reports errors by status code.  No compute."""
import ctypes as C
import os
import re
import subprocess

import pytest

from skid_amd import skg
from conftest import ROOT, have_gpu

HEADER = os.path.join(ROOT, "include", "skg.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(skg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = skg.lib()
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name


def test_parameter_validation_without_compute():
    L = skg.lib()
    h = C.c_void_p()
    P = skg.make_params(0.01, nSmooth=1, period=1.0)
    assert L.skg_create(C.byref(P), C.byref(h)) == -2          # SKG_E_PARAM
    P = skg.make_params(0.0, nSmooth=32, period=1.0)
    assert L.skg_create(C.byref(P), C.byref(h)) == -2
    assert L.skg_create(None, C.byref(h)) == -2
    assert L.skg_strerror(-6).decode().startswith("nSmooth")
    assert L.skg_last_error().decode() == ""                    # no HIP failure on this thread
    cur, peak = skg.device_memory()
    assert 0 <= cur <= peak                                      # process-wide byte count and high-water mark
    assert L.skg_destroy(None) == 0
    for fn in ("skg_density", "skg_init_move", "skg_fof", "skg_unbind", "skg_centers"):
        assert getattr(L, fn)(None, *([None] * (len(getattr(L, fn).argtypes) - 1))) == -2


def test_no_device_is_a_status_not_an_abort():
    if have_gpu():
        pytest.skip("a GPU is present")
    L = skg.lib()
    h = C.c_void_p()
    P = skg.make_params(0.01, 32, period=1.0)
    assert L.skg_create(C.byref(P), C.byref(h)) == -1          # SKG_E_NODEVICE
    assert not h.value


def test_driver_cli_usage():
    exe = os.path.join(ROOT, "skid_amd", "bin", "skid")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    r = subprocess.run([exe], input=b"", capture_output=True)
    assert r.returncode == 1 and b"-tau" in r.stderr        # main.c:339 usage()
    r = subprocess.run([exe, "-tau", "0.1", "-bogus"], input=b"", capture_output=True)
    assert r.returncode == 1
    r = subprocess.run([exe, "-tau", "0.1"], input=b"\x00" * 10, capture_output=True)
    assert r.returncode == 1 and b"tipsy" in r.stderr


def test_bench_refuses_several_gpus():
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       timeout=120)
    assert r.returncode == 2 and b"one GPU" in r.stderr and r.stdout == b""
