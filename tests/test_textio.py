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


def _reader(tio):
    tio.tio_read_ints.argtypes = [C.c_char_p, C.c_long, C.c_void_p, C.POINTER(C.c_long), C.c_int]
    tio.tio_read_ints.restype = C.c_int

    def read(path, n, nthreads):
        out = np.full(max(n, 1), -7, np.int32)
        hdr = C.c_long(-1)
        rc = tio.tio_read_ints(str(path).encode(), n, out.ctypes.data, C.byref(hdr), nthreads)
        return rc, hdr.value, out[:n]
    return read


@pytest.mark.parametrize("nthreads", [1, 7, 16])
def test_read_ints_round_trip(tio, tmp_path, nthreads):
    """kdInGroup's .grp (a count, then one fscanf("%d") per particle): the
    parallel reader returns what the writer wrote, across the 2^22-line blocks."""
    rng = np.random.default_rng(4)
    n = (1 << 22) + 4321
    v = rng.integers(0, 60000, n).astype(np.int32)
    v[:4] = [0, 2147483647, -2147483648, -1]
    path = tmp_path / "r.grp"
    assert tio.tio_write_ints(str(path).encode(), n, v.ctypes.data, 3) == 0
    rc, hdr, got = _reader(tio)(path, n, nthreads)
    assert (rc, hdr) == (0, n)
    assert np.array_equal(got, v)


def _fscanf_ref(data, n):
    """What the fscanf("%d") loop reads from data when every one of the first
    n + 1 tokens is a plain int (the reader's rc 0 / 1 cases); None otherwise."""
    import re
    toks = data.split()                           # bytes.split(): the C locale's isspace set
    if len(toks) < n + 1 or not all(re.fullmatch(rb"[+-]?[0-9]+", t) for t in toks[:n + 1]):
        return None
    vals = [int(t) for t in toks[:n + 1]]
    if any(not -2 ** 31 <= x < 2 ** 31 for x in vals):
        return None
    return vals[0], vals[1:]


@pytest.mark.parametrize("nthreads", [1, 3, 64])
def test_read_ints_forms(tio, tmp_path, nthreads):
    """Whitespace of every kind (tabs, CR LF, runs, none after the last value),
    signs and leading zeros, tokens past the n-th (read by no fscanf, so
    ignored even when they are not numbers), a short file, a wrong count, and
    tokens fscanf("%d") would split or reject (rc 2: the driver then reads the
    file with its own fscanf loop)."""
    read = _reader(tio)
    cases = [
        (b"3\n1\n2\n3\n", 3, 0),
        (b"  3\t\t+1\r\n002\v\f-3", 3, 0),
        (b"3 1 2 3 4 5 junk 0x10", 3, 0),
        (b"0\n", 0, 0),
        (b"0", 0, 0),
        (b"2 -2147483648 2147483647\n", 2, 0),
        (b"4\n1\n2\n3\n", 3, 1),
        (b"5\n1\n2\n3\n", 4, 1),
        (b"3\n1\n2\n", 3, 2),
        (b"3\n1\n2abc\n3\n", 3, 2),
        (b"3\n1\n0x10\n3\n", 3, 2),
        (b"3\n1\n-\n3\n", 3, 2),
        (b"3\n1\n2147483648\n3\n", 3, 2),
        (b"3\n1 2-3\n", 3, 2),
        (b"", 0, 2),
        (b"   \n", 1, 2),
        (b"x3\n1\n2\n3\n", 3, 2),
        (b"3x\n1\n2\n3\n", 3, 2),
        (b"3\n1\x002\n3\n", 3, 2),
    ]
    for k, (data, n, want) in enumerate(cases):
        path = tmp_path / f"c{k}.grp"
        path.write_bytes(data)
        rc, hdr, got = read(path, n, nthreads)
        assert rc == want, (data, rc)
        ref = _fscanf_ref(data, n)
        if want == 0:
            assert ref is not None and hdr == n and got.tolist() == ref[1], data
        elif want == 1:
            assert hdr == int(data.split()[0]), data
        else:
            assert ref is None or len(data.split()) < n + 1, data


def test_read_ints_missing_file(tio, tmp_path):
    rc, _, _ = _reader(tio)(tmp_path / "none.grp", 3, 2)
    assert rc == -1


def _libc():
    libc = C.CDLL(None)
    libc.fopen.restype = libc.popen.restype = C.c_void_p
    libc.fopen.argtypes = libc.popen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fread.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p]
    libc.fgetc.argtypes = libc.fclose.argtypes = libc.pclose.argtypes = [C.c_void_p]
    return libc


@pytest.mark.parametrize("src,nthreads", [("file", 1), ("file", 7), ("pipe", 7)])
def test_fread_parallel(tio, tmp_path, src, nthreads):
    """The tipsy reader's bulk read (kd.c:122-222 reads the records after the
    32-byte header): from a regular file by threads at their own offsets, from a
    pipe by fread; the bytes equal the file's and the stream is left just after
    them (the next byte is the one that follows), and a short file fails."""
    libc = _libc()
    tio.tio_fread.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rng = np.random.default_rng(5)
    head, n = 32, (20 << 20) + 12345
    data = rng.integers(0, 256, head + n + 1, dtype=np.uint8)
    path = tmp_path / "t.bin"
    path.write_bytes(data.tobytes())
    fp = libc.fopen(str(path).encode(), b"rb") if src == "file" else libc.popen(f"cat {path}".encode(), b"r")
    hbuf = np.zeros(head, np.uint8)
    assert libc.fread(hbuf.ctypes.data, 1, head, fp) == head
    out = np.zeros(n, np.uint8)
    assert tio.tio_fread(fp, out.ctypes.data, n, nthreads) == 0
    assert np.array_equal(out, data[head:head + n])
    assert libc.fgetc(fp) == int(data[-1])
    (libc.fclose if src == "file" else libc.pclose)(fp)
    fp = libc.fopen(str(path).encode(), b"rb")
    assert libc.fread(hbuf.ctypes.data, 1, head, fp) == head
    big = np.zeros(n + 2, np.uint8)
    assert tio.tio_fread(fp, big.ctypes.data, n + 2, nthreads) == -1
    libc.fclose(fp)


def test_parallel_for_shares(tio):
    """The decode's parallel loop covers [0, n) once, in contiguous shares."""
    seen = []
    CB = C.CFUNCTYPE(None, C.c_void_p, C.c_long, C.c_long)
    cb = CB(lambda arg, lo, hi: seen.append((lo, hi)))
    tio.tio_parallel_for.argtypes = [C.c_long, C.c_int, CB, C.c_void_p]
    for n, t in [(1000003, 7), (65536, 16), (100, 8), (0, 4)]:
        seen.clear()
        tio.tio_parallel_for(n, t, cb, None)
        s = sorted(seen)
        assert (s[0][0] if s else 0) == 0 and (s[-1][1] if s else 0) == n
        assert all(a[1] == b[0] for a, b in zip(s, s[1:]))
        assert len(s) <= max(t, 1)


class _Snap(C.Structure):
    _fields_ = [("time", C.c_double), ("nbodies", C.c_int), ("ndim", C.c_int), ("nsph", C.c_int),
                ("ndark", C.c_int), ("nstar", C.c_int), ("n", C.c_int),
                ("pos", C.POINTER(C.c_float)), ("vel", C.POINTER(C.c_float)), ("mass", C.POINTER(C.c_float)),
                ("soft", C.POINTER(C.c_float)), ("temp", C.POINTER(C.c_float))]


def _tipsy_bytes(rng, ns, nd, nst, xdr, t=0.375):
    """A tipsy file (kd.c:122-222's layout) and the arrays its reader must give."""
    e = ">" if xdr else "<"
    head = np.array([(t, ns + nd + nst, 3, ns, nd, nst, 0)],
                    dtype=[("t", e + "f8")] + [(f, e + "i4") for f in ("n", "d", "s", "k", "r", "p")]).tobytes()
    recs = [rng.standard_normal((c, w)).astype(np.float32) for c, w in ((ns, 12), (nd, 9), (nst, 11))]
    body = b"".join(r.astype(e + "f4").tobytes() for r in recs)
    g, d, s = recs
    want = {
        "mass": np.concatenate([g[:, 0], d[:, 0], s[:, 0]]),
        "pos": np.concatenate([g[:, 1:4], d[:, 1:4], s[:, 1:4]]),
        "vel": np.concatenate([g[:, 4:7], d[:, 4:7], s[:, 4:7]]),
        "soft": np.concatenate([g[:, 9], d[:, 7], s[:, 9]]),      # hsmooth, eps, eps
        "temp": np.concatenate([g[:, 8], np.zeros(nd + nst, np.float32)]),
    }
    return head + body, want


@pytest.mark.parametrize("xdr,src,sizes", [(0, "file", (3000, 500000, 2000)), (1, "file", (300000, 1, 200000)),
                                           (1, "pipe", (1000, 2000, 3000)), (0, "file", (0, 7, 0))])
def test_read_tipsy(tio, tmp_path, xdr, src, sizes):
    """The driver's tipsy reader against the file's own fields: native and XDR
    byte order, mixed species, files above and below the threaded-read size,
    from a file and from a pipe, the stream left after the records; a short
    file and a negative count are refused."""
    libc = _libc()
    tio.tio_read_tipsy.argtypes = [C.c_void_p, C.c_int, C.POINTER(_Snap), C.c_int]
    tio.tio_free_tipsy.argtypes = [C.POINTER(_Snap)]
    rng = np.random.default_rng(6 + xdr)
    data, want = _tipsy_bytes(rng, *sizes, xdr)
    path = tmp_path / "s.bin"
    path.write_bytes(data + b"\x5a")
    fp = libc.fopen(str(path).encode(), b"rb") if src == "file" else libc.popen(f"cat {path}".encode(), b"r")
    s = _Snap()
    assert tio.tio_read_tipsy(fp, xdr, C.byref(s), 5) == 0
    n = sum(sizes)
    assert (s.n, s.nsph, s.ndark, s.nstar, s.nbodies, s.ndim, s.time) == (n, *sizes, n, 3, 0.375)
    for f, w in want.items():
        got = np.ctypeslib.as_array(getattr(s, f), (n * (3 if f in ("pos", "vel") else 1),))
        assert np.array_equal(got.view(np.uint32), np.ascontiguousarray(w).reshape(-1).view(np.uint32)), f
    assert libc.fgetc(fp) == 0x5A
    tio.tio_free_tipsy(C.byref(s))
    (libc.fclose if src == "file" else libc.pclose)(fp)
    for bad in (data[:-5], data[:8] + np.array([-1], ">i4" if xdr else "<i4").tobytes() * 6 + data[32:]):
        path.write_bytes(bad)
        fp = libc.fopen(str(path).encode(), b"rb")
        assert tio.tio_read_tipsy(fp, xdr, C.byref(s), 5) == -1
        assert not s.pos
        libc.fclose(fp)


@pytest.mark.parametrize("nthreads", [1, 6])
def test_ray_vectors(tio, nthreads):
    """kdOutVector's displacements (kd.c:1550-1606) in float32: moved minus
    initial position on each axis, one period off above half a period, one on
    at or below minus half, all x then all y then all z; NaN (did not move)
    stays NaN for the writer to print as 0.  Includes the two boundaries."""
    tio.tio_ray_vectors.argtypes = [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    rng = np.random.default_rng(7)
    n = 200003
    per = np.array([1.0, 2.5, 0.75], np.float32)
    h = np.float32(0.5) * per
    pos = (rng.random((n, 3), dtype=np.float32) - np.float32(0.5)) * per
    moved = pos + rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32) * per
    moved[::7] = np.nan
    moved[1, 0], pos[1, 0] = np.float32(0.25), np.float32(-0.25)          # d = +h exactly: kept
    moved[2, 0], pos[2, 0] = np.float32(-0.25), np.float32(0.25)          # d = -h exactly: one period on
    out = np.empty(3 * n, np.float32)
    tio.tio_ray_vectors(n, moved.ctypes.data, pos.ctypes.data, per.ctypes.data, out.ctypes.data, nthreads)
    want = np.empty((3, n), np.float32)
    for j in range(3):
        d = moved[:, j] - pos[:, j]
        hj, p2 = h[j], np.float32(2) * h[j]
        d = np.where(d > hj, d - p2, d).astype(np.float32)
        d = np.where(d <= -hj, d + p2, d).astype(np.float32)
        want[j] = d
    assert want[0, 1] == h[0] and want[0, 2] == h[0]
    assert np.array_equal(out.view(np.uint32), want.reshape(-1).view(np.uint32))
