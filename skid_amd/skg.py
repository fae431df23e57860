"""ctypes mirror of the libskg C ABI (include/skg.h).

The host side of the GPU path: a thin Python view of the same stage calls the
C driver (skid_amd/driver/skid.c) makes, used by tests and bench.py.  It loads
the in-tree skid_amd/libskg.so and raises if that library is missing -- there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libskg.so")

SKG_OK = 0
SKG_STAT_N = 19
SKG_E_NODEVICE, SKG_E_PARAM, SKG_E_NOMEM, SKG_E_STATE, SKG_E_HIP, SKG_E_NSMOOTH = -1, -2, -3, -4, -5, -6
SKG_DARK, SKG_GAS, SKG_STAR = 1, 2, 4
SKG_PLUMMER, SKG_SPLINE = 1, 2
SKG_ACC_PLAIN, SKG_ACC_COMPENSATED = 0, 1
SKG_TUNE_BUCKET, SKG_TUNE_MAPPING, SKG_TUNE_POT_TILE, SKG_TUNE_KNN, SKG_TUNE_ALL = 1, 2, 4, 8, 15
SKG_NOALLOC_MAX_POINTS, SKG_NOALLOC_SCRATCH_BYTES = 524288, 64 << 20
STAGES = ("tree", "knn", "density", "move", "fof", "group", "unbind", "spare")
KERNELS = ("knn", "cover_density", "cover_move", "fof", "unbind_potential", "unbind_loop", "unbind_big", "cover_build")
FLT_MAX = 3.4028234663852886e38


class SkgParams(C.Structure):
    _fields_ = [
        ("nSmooth", C.c_int), ("nBucket", C.c_int),
        ("fPeriod", C.c_float * 3), ("fCenter", C.c_float * 3), ("bPeriodic", C.c_int),
        ("fTau", C.c_float), ("fCvg", C.c_float), ("fStep", C.c_float), ("fScoop", C.c_float),
        ("fDensMin", C.c_float), ("fTempMax", C.c_float), ("fMassMax", C.c_float),
        ("nMembers", C.c_int), ("nMaxMembers", C.c_int),
        ("bGasAndDark", C.c_int), ("bGasOnly", C.c_int), ("bForceInitialCut", C.c_int),
        ("bNoUnbind", C.c_int), ("iSoftType", C.c_int),
        ("G", C.c_float), ("z", C.c_float), ("fShift", C.c_float), ("fCosmo", C.c_float),
        ("accumulate", C.c_int), ("device", C.c_int),
    ]


class SkgError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SkgError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, ip, fp = C.c_void_p, C.POINTER(C.c_int), C.c_void_p
        L.skg_default_params.argtypes = [C.POINTER(SkgParams)]
        L.skg_create.argtypes = [C.POINTER(SkgParams), C.POINTER(C.c_void_p)]
        L.skg_destroy.argtypes = [vp]
        L.skg_strerror.argtypes = [C.c_int]
        L.skg_strerror.restype = C.c_char_p
        L.skg_last_error.argtypes = []
        L.skg_last_error.restype = C.c_char_p
        L.skg_device_memory.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.c_int]
        L.skg_load.argtypes = [vp, C.c_int, C.c_int, C.c_int, fp, fp, fp, fp, fp]
        L.skg_load_batch.argtypes = [vp, C.c_int, fp, fp, fp, fp, fp, fp, fp, fp, fp]
        L.skg_box_counts.argtypes = [vp, fp, fp]
        L.skg_density.argtypes = [vp, fp, fp, ip]
        L.skg_neighbors.argtypes = [vp, fp, fp]
        L.skg_init_move.argtypes = [vp, ip]
        L.skg_move_first.argtypes = [vp, ip, ip]
        L.skg_move_block.argtypes = [vp, ip, ip]
        L.skg_fof.argtypes = [vp, fp, ip]
        L.skg_microsteps.argtypes = [vp, fp]
        L.skg_get_moved.argtypes = [vp, fp]
        L.skg_centers.argtypes = [vp]
        L.skg_unbind.argtypes = [vp, ip]
        L.skg_too_small.argtypes = [vp]
        L.skg_get_groups.argtypes = [vp, fp, ip]
        L.skg_get_catalogue.argtypes = [vp, fp, fp, fp, fp, fp, fp]
        L.skg_density_adjoint.argtypes = [vp, fp, C.c_int, fp, fp]
        L.skg_field_gradient.argtypes = [vp, C.c_int, fp, fp]
        L.skg_density_dlnk.argtypes = [vp, fp]
        L.skg_set_groups.argtypes = [vp, fp, fp, fp]
        L.skg_group_energies.argtypes = [vp, fp, fp]
        L.skg_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
        L.skg_knn_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.skg_fof_points.argtypes = [vp, C.c_int, fp, C.c_float, fp, ip]
        L.skg_fof_points_noalloc.argtypes = [vp, C.c_int, vp, C.c_float, vp, ip]
        L.skg_stage_times.argtypes = [vp, C.POINTER(C.c_double)]
        L.skg_set_profiling.argtypes = [vp, C.c_int]
        L.skg_device_copy.argtypes = [vp, C.c_longlong, C.c_int]
        L.skg_host_alloc.argtypes = [C.POINTER(C.c_void_p), C.c_longlong]
        L.skg_host_free.argtypes = [vp]
        L.skg_kernel_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                       C.POINTER(C.c_double)]
        _lib = L
    return _lib


def device_memory(reset_peak: bool = False) -> tuple[int, int]:
    """(bytes the library holds on the device now, high-water mark), process-wide."""
    cur, peak = C.c_longlong(), C.c_longlong()
    _check(lib().skg_device_memory(C.byref(cur), C.byref(peak), int(reset_peak)), "skg_device_memory")
    return cur.value, peak.value


def pinned_empty(shape, dtype=np.float32) -> np.ndarray:
    """An uninitialised array in page-locked host memory (skg_host_alloc),
    freed when the array and its views are gone.  Copies between it and the
    device run as DMA, with no staging and no first-touch page faults."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    ptr = C.c_void_p()
    _check(lib().skg_host_alloc(C.byref(ptr), max(n, 1)), "skg_host_alloc")
    buf = (C.c_char * max(n, 1)).from_address(ptr.value)
    weakref.finalize(buf, lib().skg_host_free, ptr.value)
    return np.frombuffer(buf, dt, count=n // dt.itemsize).reshape(shape)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int, what: str):
    if rc != SKG_OK:
        msg = f"{what}: {lib().skg_strerror(rc).decode()} ({rc})"
        if rc == SKG_E_HIP:
            msg += ": " + lib().skg_last_error().decode()
        raise SkgError(msg)


def make_params(tau: float, nSmooth: int = 64, nBucket: int = 16, period: float | None = None,
                center=(0.0, 0.0, 0.0), cvg: float | None = None, scoop: float | None = None,
                accumulate: int = SKG_ACC_COMPENSATED, **kw) -> SkgParams:
    """skid's command-line semantics (main.c:82-345) as a parameter block."""
    P = SkgParams()
    lib().skg_default_params(C.byref(P))
    P.nSmooth, P.nBucket, P.fTau = nSmooth, nBucket, tau
    if period is not None:
        P.fPeriod[:] = [period] * 3
        P.bPeriodic = 1
    P.fCenter[:] = list(center)
    P.fCvg = cvg if cvg is not None else 0.5 * tau
    P.fScoop = scoop if scoop is not None else 2.0 * tau
    P.fStep = 0.5 * np.float32(P.fCvg)
    P.accumulate = accumulate
    for k, v in kw.items():
        setattr(P, k, v)
    return P


class Context:
    """One libskg context (skg_ctx) with particles loaded."""

    def __init__(self, box, params: SkgParams):
        self.L = lib()
        self.box = box
        h = C.c_void_p()
        _check(self.L.skg_create(C.byref(params), C.byref(h)), "skg_create")
        self.h = h
        self.n = box.n
        self.k = params.nSmooth
        b = box
        self._keep = [np.ascontiguousarray(a, np.float32) for a in
                      (b.pos.reshape(-1), b.vel.reshape(-1), b.mass, b.soft, b.temp)]
        _check(self.L.skg_load(h, b.n_gas, b.n_dark, b.n_star, *[_p(a) for a in self._keep]), "skg_load")

    def close(self):
        if getattr(self, "h", None):
            self.L.skg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def keep_outputs(self, on: bool = True, pin_inputs: bool = True):
        """Return per-particle outputs in page-locked arrays that the context
        keeps (pinned_empty): each call of a method then overwrites the arrays
        that method returned the previous time, and its copy runs as DMA.  For
        callers that repeat a pass (a step loop) and read each pass's results
        before the next.  pin_inputs moves the host arrays that reload()
        uploads into page-locked memory too."""
        self._outs = {} if on else None
        if on and pin_inputs and hasattr(self, "_keep"):
            pinned = []
            for a in self._keep:
                if a is None or a.size == 0:
                    pinned.append(a)
                    continue
                b = pinned_empty(a.shape, a.dtype)
                b[...] = a
                pinned.append(b)
            self._keep = pinned

    def _out(self, name, shape, dtype=np.float32):
        outs = getattr(self, "_outs", None)
        if outs is None:
            return np.empty(shape, dtype)
        a = outs.get(name)
        if a is None or a.shape != tuple(np.atleast_1d(shape)) or a.dtype != np.dtype(dtype):
            a = outs[name] = pinned_empty(shape, dtype)
        return a

    def density(self):
        b2 = self._out("b2", self.n)
        rho = self._out("rho", self.n)
        ne = C.c_int()
        _check(self.L.skg_density(self.h, _p(b2), _p(rho), C.byref(ne)), "skg_density")
        return b2, rho, ne.value

    def neighbors(self):
        nb = np.empty(self.n * self.k, np.int32)
        d2 = np.empty(self.n * self.k, np.float32)
        _check(self.L.skg_neighbors(self.h, _p(nb), _p(d2)), "skg_neighbors")
        return nb.reshape(self.n, self.k), d2.reshape(self.n, self.k)

    def init_move(self) -> int:
        n = C.c_int()
        _check(self.L.skg_init_move(self.h, C.byref(n)), "skg_init_move")
        return n.value

    def move_first(self):
        a, s = C.c_int(), C.c_int()
        _check(self.L.skg_move_first(self.h, C.byref(a), C.byref(s)), "skg_move_first")
        return a.value, s.value

    def move_block(self):
        a, s = C.c_int(), C.c_int()
        _check(self.L.skg_move_block(self.h, C.byref(a), C.byref(s)), "skg_move_block")
        return a.value, s.value

    def move_loop(self, max_blocks: int = 100000):
        """The Denmax loop of main.c:393-419; returns the Ittr trace (nActive, nScatter).

        max_blocks bounds the number of 5-step blocks (the reference loops until
        every mover has converged); exceeding it raises instead of spinning."""
        self.init_move()
        tr = [self.move_first()]
        while tr[-1][0]:
            if len(tr) > max_blocks:
                raise SkgError(f"move loop did not converge in {max_blocks} blocks")
            tr.append(self.move_block())
        return np.array(tr, np.int32)

    def fof(self):
        g = self._out("fof", self.n, np.int32)
        ng = C.c_int()
        _check(self.L.skg_fof(self.h, _p(g), C.byref(ng)), "skg_fof")
        return g, ng.value

    def microsteps(self):
        """The 5 micro-steps (main.c:431-438); returns the nScatter of each
        Microstep line."""
        s = np.zeros(5, np.int32)
        _check(self.L.skg_microsteps(self.h, _p(s)), "skg_microsteps")
        return s

    def moved(self):
        r = np.empty(3 * self.n, np.float32)
        _check(self.L.skg_get_moved(self.h, _p(r)), "skg_get_moved")
        return r.reshape(self.n, 3)

    def set_groups(self, grp, center=None, vcm=None):
        """Groups from outside (the -unbind re-entry); returns nGroup."""
        g = np.ascontiguousarray(grp, np.int32)
        c = None if center is None else np.ascontiguousarray(center, np.float32).reshape(-1)
        v = None if vcm is None else np.ascontiguousarray(vcm, np.float32).reshape(-1)
        _check(self.L.skg_set_groups(self.h, _p(g), _p(c), _p(v)), "skg_set_groups")
        return int(g.max()) + 1 if g.size else 1

    def group_energies(self, ngroup: int):
        """(W, T) per group before unbinding, FP64."""
        W = np.zeros(ngroup, np.float64)
        T = np.zeros(ngroup, np.float64)
        _check(self.L.skg_group_energies(self.h, _p(W), _p(T)), "skg_group_energies")
        return W, T

    def centers(self):
        _check(self.L.skg_centers(self.h), "skg_centers")

    def unbind(self) -> int:
        n = C.c_int()
        _check(self.L.skg_unbind(self.h, C.byref(n)), "skg_unbind")
        return n.value

    def group_stats(self, ngroup: int):
        """skg_group_stats: (ngroup, SKG_STAT_N) per-group .stat quantities."""
        out = np.zeros((ngroup, SKG_STAT_N), np.float32)
        _check(self.L.skg_group_stats(self.h, _p(out)), "skg_group_stats")
        return out

    def unbind_info(self):
        """(groups unbound cooperatively, refused cooperative launches, full passes)."""
        a, b, f = C.c_int(), C.c_int(), C.c_longlong()
        _check(self.L.skg_unbind_info(self.h, C.byref(a), C.byref(b), C.byref(f)), "skg_unbind_info")
        return a.value, b.value, f.value

    def too_small(self):
        _check(self.L.skg_too_small(self.h), "skg_too_small")

    def groups(self):
        g = self._out("groups", self.n, np.int32)
        ng = C.c_int()
        _check(self.L.skg_get_groups(self.h, _p(g), C.byref(ng)), "skg_get_groups")
        return g, ng.value

    def catalogue(self, ngroup: int):
        out = dict(mass=np.empty(ngroup, np.float32), center=np.empty(3 * ngroup, np.float32),
                   vcm=np.empty(3 * ngroup, np.float32), rbound=np.empty(3 * ngroup, np.float32),
                   radius=np.empty(ngroup, np.float32), nmembers=np.empty(ngroup, np.int32))
        _check(self.L.skg_get_catalogue(self.h, *[_p(out[k]) for k in
                                                  ("mass", "center", "vcm", "rbound", "radius", "nmembers")]),
               "skg_get_catalogue")
        for k in ("center", "vcm", "rbound"):
            out[k] = out[k].reshape(-1, 3)
        return out

    def adjoint(self, gbar, parts: int = 3):
        g = np.ascontiguousarray(gbar, np.float32)
        dx = self._out("adj_dx", 3 * self.n)
        dh = self._out("adj_dh", self.n)
        _check(self.L.skg_density_adjoint(self.h, _p(g), parts, _p(dx), _p(dh)), "skg_density_adjoint")
        return dx.reshape(self.n, 3), dh

    def density_dlnk(self):
        """d rho / d ln nSmooth per particle (frozen neighbour sets)."""
        d = self._out("dlnk", self.n)
        _check(self.L.skg_density_dlnk(self.h, _p(d)), "skg_density_dlnk")
        return d

    def field_gradient(self, q):
        q = np.ascontiguousarray(q, np.float32).reshape(-1)
        a = np.empty_like(q)
        _check(self.L.skg_field_gradient(self.h, q.size // 3, _p(q), _p(a)), "skg_field_gradient")
        return a.reshape(-1, 3)

    def fof_points(self, pos, tau: float):
        p = np.ascontiguousarray(pos, np.float32).reshape(-1)
        g = np.empty(p.size // 3, np.int32)
        ng = C.c_int()
        _check(self.L.skg_fof_points(self.h, p.size // 3, _p(p), tau, _p(g), C.byref(ng)), "skg_fof_points")
        return g, ng.value

    def autotune(self, what: int = SKG_TUNE_ALL) -> dict:
        """Time launch shapes on the loaded particles (skg_autotune); the path
        must then be run again from density()."""
        b, m, t = C.c_int(), C.c_longlong(), C.c_int()
        _check(self.L.skg_autotune(self.h, what, C.byref(b), C.byref(m), C.byref(t)), "skg_autotune")
        kb, kw = C.c_int(), C.c_int()
        _check(self.L.skg_knn_shape(self.h, C.byref(kb), C.byref(kw)), "skg_knn_shape")
        return {"nBucket": b.value, "sub_max": m.value, "pot_tile": t.value, "knn_bucket": kb.value,
                "knn_waves": kw.value}

    def stage_times(self) -> dict:
        t = (C.c_double * 8)()
        _check(self.L.skg_stage_times(self.h, t), "skg_stage_times")
        return {STAGES[i]: t[i] for i in range(8)}

    def device_copy(self, nbytes: int, reps: int = 3):
        """skg_device_copy: a plain device copy of known size (counter calibration)."""
        _check(self.L.skg_device_copy(self.h, int(nbytes), int(reps)), "skg_device_copy")

    def set_profiling(self, on: bool = True):
        _check(self.L.skg_set_profiling(self.h, int(on)), "skg_set_profiling")

    def kernel_stats(self) -> dict:
        out = {}
        for i, name in enumerate(KERNELS):
            ms, n, w = C.c_double(), C.c_longlong(), C.c_double()
            _check(self.L.skg_kernel_stats(self.h, i, C.byref(ms), C.byref(n), C.byref(w)), "skg_kernel_stats")
            out[name] = dict(ms=ms.value, launches=n.value, work=w.value)
        return out

    def run(self):
        """main.c:374-471 through the GPU stages; returns (trace, groups, nGroup)."""
        self.density()
        tr = self.move_loop()
        self.fof()
        self.microsteps()
        self.centers()
        self.unbind()
        self.too_small()
        g, ng = self.groups()
        return tr, g, ng


def nsmooth_report(rho, drho, groups, dens_min=0.0, rel_k=0.25, sample=1 << 20):
    """Membership sensitivity to nSmooth from d rho / d ln nSmooth.

    A particle's density moves by about drho * ln(1 + rel_k) when nSmooth
    changes by the factor 1 + rel_k.  Particles whose density would cross the
    mover threshold dens_min (-d, main.c:114-115) in either direction change
    the moving set, and with it the groups they can join.  Returns the
    distribution of d ln rho / d ln nSmooth over scatterers (quantiles from a
    fixed subsample of at most `sample` scatterers), the number at risk, and
    per group (id > 0) arrays: size, mean sensitivity, fraction at risk.
    """
    rho = np.asarray(rho, np.float32)
    drho = np.asarray(drho, np.float32)
    groups = np.asarray(groups)
    sc = rho > 0
    s = np.divide(drho, rho, out=np.zeros_like(rho), where=sc)
    up, dn = np.float32(np.log1p(rel_k)), np.float32(np.log1p(-rel_k)) if rel_k < 1 else np.float32(-np.inf)
    a, b = rho + drho * up, rho + drho * dn
    at_risk = sc & (np.minimum(a, b) < dens_min) & (np.maximum(a, b) >= dens_min)
    idx = np.flatnonzero(sc)
    if idx.size > sample:
        idx = idx[np.linspace(0, idx.size - 1, sample).astype(np.int64)]
    q = np.quantile(s[idx], [0.05, 0.5, 0.95]) if idx.size else [np.nan] * 3
    g = groups.astype(np.int64)
    n = np.bincount(g)
    n[0] = 0
    with np.errstate(invalid="ignore", divide="ignore"):
        mean_s = np.bincount(g, weights=s) / n
        risk = np.bincount(g, weights=at_risk) / n
    return {"scatterers": int(sc.sum()), "dlnrho_dlnk": {"p5": float(q[0]), "median": float(q[1]), "p95": float(q[2])},
            "at_risk": int(at_risk.sum()), "rel_k": rel_k, "dens_min": dens_min,
            "group_size": n, "group_mean_dlnrho_dlnk": mean_s, "group_frac_at_risk": risk}


class BatchContext(Context):
    """Many independent boxes in one context (skg_load_batch): every stage call
    runs all boxes in the same launches.  taus (per box) sets each box's tau,
    fCvg, fStep and fScoop as main.c:343-345 does from -tau."""

    def __init__(self, boxes, params: SkgParams, taus=None):
        self.L = lib()
        self.boxes = list(boxes)
        h = C.c_void_p()
        _check(self.L.skg_create(C.byref(params), C.byref(h)), "skg_create")
        self.h = h
        self.k = params.nSmooth
        self.nbox = len(self.boxes)
        self.sizes = np.array([b.n for b in self.boxes], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        cat = lambda f: np.ascontiguousarray(np.concatenate([f(b) for b in self.boxes]), np.float32)
        self._keep = [cat(lambda b: b.pos.reshape(-1)), cat(lambda b: b.vel.reshape(-1)),
                      cat(lambda b: b.mass), cat(lambda b: b.soft), cat(lambda b: b.temp)]
        ng = np.array([b.n_gas for b in self.boxes], np.int32)
        nd = np.array([b.n_dark for b in self.boxes], np.int32)
        ns = np.array([b.n_star for b in self.boxes], np.int32)
        tau = None if taus is None else np.ascontiguousarray(taus, np.float32)
        self._keep += [ng, nd, ns, tau]
        _check(self.L.skg_load_batch(h, self.nbox, _p(ng), _p(nd), _p(ns), _p(tau),
                                     *[_p(a) for a in self._keep[:5]]), "skg_load_batch")

    def reload(self):
        """skg_load_batch again from the host arrays kept at construction (the
        upload from host memory); the path then runs again from density()."""
        _check(self.L.skg_load_batch(self.h, self.nbox, *[_p(a) for a in self._keep[5:9]],
                                     *[_p(a) for a in self._keep[:5]]), "skg_load_batch")

    def split(self, a):
        """Per-box views of an array indexed by the concatenated iOrder."""
        return [a[self.offsets[i]:self.offsets[i + 1]] for i in range(self.nbox)]

    def box_counts(self):
        a = np.empty(self.nbox, np.int32)
        s = np.empty(self.nbox, np.int32)
        _check(self.L.skg_box_counts(self.h, _p(a), _p(s)), "skg_box_counts")
        return a, s

    def move_loop(self, max_blocks: int = 100000):
        """The Denmax loop for every box at once; returns per-box Ittr traces,
        each ending where its own loop would (the first line with nActive 0)."""
        self.init_move()
        self.move_first()
        rows = [self.box_counts()]
        while rows[-1][0].sum():
            if len(rows) > max_blocks:
                raise SkgError(f"move loop did not converge in {max_blocks} blocks")
            self.move_block()
            rows.append(self.box_counts())
        act = np.stack([r[0] for r in rows], 1)
        sca = np.stack([r[1] for r in rows], 1)
        out = []
        for b in range(self.nbox):
            z = np.flatnonzero(act[b] == 0)
            end = (z[0] + 1) if z.size else act.shape[1]
            out.append(np.stack([act[b, :end], sca[b, :end]], 1).astype(np.int32))
        return out

    def box_groups(self, g):
        """Per-box group ids renumbered 1.. within each box (0 stays 0)."""
        out = []
        for gb in self.split(g):
            gb = gb.copy()
            pos = gb > 0
            if pos.any():
                gb[pos] -= gb[pos].min() - 1
            out.append(gb)
        return out
