"""The driver's parallel text writers (skid_amd/driver/textio.c) give the bytes
of one fprintf per line (.grp "%d", .den "%.10g", .ray "%g" with NaN as 0):
compared with Python's printf-style formatting of the same values, with one
thread and with several, across the block boundary of 2^22 lines."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "skid_amd", "driver", "textio.c")


@pytest.fixture(scope="module")
def tio(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("tio") / "libtio.so")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-pthread", SRC, "-o", so], check=True)
    L = C.CDLL(so)
    L.tio_write_ints.argtypes = [C.c_char_p, C.c_long, C.c_void_p, C.c_int]
    L.tio_write_floats.argtypes = [C.c_char_p, C.c_long, C.c_long, C.c_void_p, C.c_char_p, C.c_int, C.c_int]
    return L


def _ref_text(header, lines):
    return (f"{header}\n" + "".join(lines)).encode()


@pytest.mark.parametrize("nthreads", [1, 7])
def test_ints(tio, tmp_path, nthreads):
    rng = np.random.default_rng(1)
    n = (1 << 22) + 12345                      # crosses the block boundary
    v = rng.integers(0, 60000, n).astype(np.int32)
    v[:6] = [0, -1, 2147483647, -2147483648, 10, -10]
    path = str(tmp_path / "g.grp")
    assert tio.tio_write_ints(path.encode(), n, v.ctypes.data, nthreads) == 0
    got = open(path, "rb").read()
    want = _ref_text(n, (f"{x}\n" for x in v.tolist()))
    assert got == want


@pytest.mark.parametrize("fmt,nan_zero,nthreads", [("%.10g", 0, 1), ("%.10g", 0, 5), ("%g", 1, 3)])
def test_floats(tio, tmp_path, fmt, nan_zero, nthreads):
    rng = np.random.default_rng(2)
    n = 300001
    v = (rng.standard_normal(n) * 10.0 ** rng.integers(-40, 38, n)).astype(np.float32)
    v[:8] = np.array([0.0, -0.0, 1e-45, -1.17549435e-38, 3.4028235e38, np.inf, -np.inf, 1.0], np.float32)
    if nan_zero:
        v[100::1000] = np.nan
    path = str(tmp_path / "f.txt")
    header = 777
    assert tio.tio_write_floats(path.encode(), header, n, v.ctypes.data, fmt.encode(), nan_zero, nthreads) == 0
    got = open(path, "rb").read()
    lines = ("0\n" if (nan_zero and np.isnan(x)) else (fmt % float(x)) + "\n" for x in v)
    assert got == _ref_text(header, lines)


@pytest.mark.parametrize("fmt,nan_zero", [("%.10g", 0), ("%g", 1)])
def test_floats_exact_rounding(tio, tmp_path, fmt, nan_zero):
    """The writer's own "%.10g" / "%g" formatter (no printf) against correctly
    rounded formatting: random bit patterns over all finite floats (denormals
    and values below ~1e-29 go through snprintf), the float neighbours of every
    power of ten in range, decimal-looking values and halfway cases."""
    rng = np.random.default_rng(3)
    n = 200000
    v = rng.integers(0, 0x7F800000, n, dtype=np.int64).astype(np.uint32).view(np.float32).copy()
    v[: n // 2] *= -1
    near = []
    for e in range(-30, 39):
        f = np.float32(10.0 ** e)
        near += [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    dec = (rng.integers(1, 10 ** 7, 20000) * 10.0 ** rng.integers(-20, 20, 20000)).astype(np.float32)
    ties = np.array([0.5, 1.5, 2.5, 123456.5, 1234567.5, 9999999999.5, 8388607.5, 1e-5, 1e-4, 99999.95], np.float32)
    v = np.concatenate([v, np.array(near, np.float32), dec, ties])
    path = str(tmp_path / "e.txt")
    assert tio.tio_write_floats(path.encode(), 3, len(v), v.ctypes.data, fmt.encode(), nan_zero, 4) == 0
    got = open(path, "rb").read()
    assert got == _ref_text(3, ((fmt % float(x)) + "\n" for x in v))
