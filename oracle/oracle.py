"""ctypes bindings for the CPU oracle (oracle/bin/liboracle.so) and, when it has
been built, the real reference (oracle/_ref/libskidref.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker -- never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "bin", "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libskidref.so")

_f = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_d = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OrParams(C.Structure):
    _fields_ = [
        ("nSmooth", C.c_int), ("nBucket", C.c_int),
        ("fPeriod", C.c_float), ("fCenter", C.c_float * 3), ("bPeriodic", C.c_int),
        ("fTau", C.c_float), ("fCvg", C.c_float), ("fScoop", C.c_float),
        ("fDensMin", C.c_float), ("fTempMax", C.c_float), ("fMassMax", C.c_float),
        ("nMembers", C.c_int), ("nMaxMembers", C.c_int),
        ("bGasAndDark", C.c_int), ("bGasOnly", C.c_int), ("bForceInitialCut", C.c_int),
        ("bNoUnbind", C.c_int), ("iSoftType", C.c_int),
        ("G", C.c_float), ("Omega0", C.c_float), ("Lambda", C.c_float),
        ("fQuintess", C.c_float), ("H0", C.c_float), ("z", C.c_float),
        ("fEps", C.c_float), ("bOutDens", C.c_int), ("bQuiet", C.c_int),
    ]


class RefParams(C.Structure):
    _fields_ = [
        ("nSmooth", C.c_int), ("nBucket", C.c_int),
        ("fPeriod", C.c_float), ("fCenter", C.c_float * 3), ("bPeriodic", C.c_int),
        ("fTau", C.c_float), ("fCvg", C.c_float), ("fScoop", C.c_float),
        ("fDensMin", C.c_float), ("fTempMax", C.c_float), ("fMassMax", C.c_float),
        ("nMembers", C.c_int), ("nMaxMembers", C.c_int),
        ("bGasAndDark", C.c_int), ("bGasOnly", C.c_int), ("bForceInitialCut", C.c_int),
        ("bNoUnbind", C.c_int), ("iSoftType", C.c_int),
        ("G", C.c_float), ("Omega0", C.c_float), ("Lambda", C.c_float),
        ("fQuintess", C.c_float), ("H0", C.c_float), ("z", C.c_float),
        ("fEps", C.c_float), ("bOutDens", C.c_int),
    ]


class RefOut(C.Structure):
    _fields_ = [
        ("fBall2", C.c_void_p), ("fDensity", C.c_void_p), ("rMoved", C.c_void_p),
        ("rMicro", C.c_void_p), ("iFofGroup", C.c_void_p), ("iGroup", C.c_void_p),
        ("nExtraScat", C.c_int), ("nMove", C.c_int), ("nFofGroup", C.c_int),
        ("nGroup", C.c_int), ("nIttr", C.c_int), ("trace", C.c_void_p), ("maxTrace", C.c_int),
        ("grpMass", C.c_void_p), ("grpCenter", C.c_void_p), ("grpVcm", C.c_void_p),
        ("maxGroup", C.c_int), ("micro", C.c_int * 5), ("gtpPath", C.c_char_p),
    ]


def params(**kw) -> dict:
    """SKID defaults (main.c:82-136) overridden by kw; keys as in or_params."""
    p = dict(nSmooth=64, nBucket=16, fPeriod=3.4028234663852886e38, fCenter=(0.0, 0.0, 0.0),
             bPeriodic=0, fTau=0.0, fCvg=0.0, fScoop=0.0, fDensMin=0.0,
             fTempMax=3.4028234663852886e38, fMassMax=3.4028234663852886e38, nMembers=8,
             nMaxMembers=2147483647, bGasAndDark=0, bGasOnly=0, bForceInitialCut=0, bNoUnbind=0,
             iSoftType=2, G=1.0, Omega0=1.0, Lambda=0.0, fQuintess=0.0, H0=0.0, z=0.0,
             fEps=0.0, bOutDens=0, bQuiet=1)
    for k, v in kw.items():
        if k not in p:
            raise KeyError(k)
        p[k] = v
    if "fPeriod" in kw and "bPeriodic" not in kw:
        p["bPeriodic"] = 1
    return p


def _fill(struct, p: dict):
    s = struct()
    for name, _ in struct._fields_:
        v = p[name]
        if name == "fCenter":
            s.fCenter[:] = list(v)
        else:
            setattr(s, name, v)
    return s


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_ORACLE):
            raise RuntimeError(f"oracle not built: {LIB_ORACLE} (run `make -C oracle`)")
        L = C.CDLL(LIB_ORACLE)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(OrParams)]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_load.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _f, _f, _f, _f, _f]
        for fn in ("or_density",):
            getattr(L, fn).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_init_move.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.or_move_first.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_move_block.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_first_gradient.argtypes = [C.c_void_p, _f]
        L.or_fof.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        L.or_microsteps.argtypes = [C.c_void_p, C.c_void_p]
        L.or_get_moved.argtypes = [C.c_void_p, _f]
        L.or_centers.argtypes = [C.c_void_p]
        L.or_unbind.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.or_too_small.argtypes = [C.c_void_p]
        L.or_get_groups.argtypes = [C.c_void_p, _i, C.POINTER(C.c_int)]
        L.or_get_catalogue.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.or_run.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.or_fof_points.argtypes = [C.c_int, _f, C.c_float, C.c_float, C.c_int, C.c_int, _i]
        L.or_knn_brute.argtypes = [C.c_int, _f, C.c_int, C.c_float, C.c_int, C.c_void_p, C.c_void_p]
        L.or_fof_brute.argtypes = [C.c_int, _f, C.c_float, C.c_float, C.c_int, _i]
        L.or_density_f64.argtypes = [C.c_int, _f, _f, _f, C.c_float, C.c_int, _d]
        L.or_gradient_f64.argtypes = [C.c_int, _f, _f, _f, C.c_int, _f, C.c_float, C.c_int, _d]
        L.or_potential_f64.argtypes = [C.c_int, _f, _f, _f, C.c_float, C.c_int, _d]
        L.or_spline_pot.argtypes = [C.c_float, C.c_float]
        L.or_spline_pot.restype = C.c_float
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(LIB_REF)


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError("reference harness not built (make -f oracle/ref.mk)")
        R = C.CDLL(LIB_REF)
        R.ref_density.argtypes = [C.POINTER(RefParams), C.c_int, C.c_int, C.c_int, _f, _f, _f, _f,
                                  _f, _f, C.POINTER(C.c_int)]
        R.ref_pipeline.argtypes = [C.POINTER(RefParams), C.c_int, C.c_int, C.c_int, _f, _f, _f,
                                   _f, _f, C.POINTER(RefOut)]
        _i = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        R.ref_unbind.argtypes = [C.POINTER(RefParams), C.c_int, C.c_int, C.c_int, _f, _f, _f, _f, _f, _i, _i,
                                 C.POINTER(C.c_int)]
        _ref = R
    return _ref


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Oracle:
    """One oracle context (or_ctx) over a loaded box."""

    def __init__(self, box, **kw):
        self.p = params(**kw)
        self.L = lib()
        self.box = box
        self.ctx = self.L.or_create(C.byref(_fill(OrParams, self.p)))
        b = box
        rc = self.L.or_load(self.ctx, b.n_gas, b.n_dark, b.n_star, b.pos.reshape(-1), b.vel.reshape(-1),
                            b.mass, b.soft, b.temp)
        if rc:
            raise RuntimeError("or_load failed")

    def __del__(self):
        if getattr(self, "ctx", None):
            self.L.or_destroy(self.ctx)
            self.ctx = None

    def density(self):
        n = self.box.n
        b2 = np.empty(n, np.float32)
        rho = np.empty(n, np.float32)
        ne = C.c_int()
        rc = self.L.or_density(self.ctx, _ptr(b2), _ptr(rho), C.byref(ne))
        if rc:
            raise RuntimeError(f"or_density failed {rc}")
        return b2, rho, ne.value

    def run(self, max_trace=100000):
        tr = np.zeros(2 * max_trace, np.int32)
        nt = C.c_int()
        rc = self.L.or_run(self.ctx, _ptr(tr), max_trace, C.byref(nt))
        if rc:
            raise RuntimeError(f"or_run failed {rc}")
        return tr[: 2 * nt.value].reshape(-1, 2)

    def groups(self):
        g = np.empty(self.box.n, np.int32)
        ng = C.c_int()
        self.L.or_get_groups(self.ctx, g, C.byref(ng))
        return g, ng.value

    def move_loop(self):
        """main.c:393-419 after density(); returns the Ittr trace."""
        a, s = C.c_int(), C.c_int()
        self.L.or_init_move(self.ctx, C.byref(a))
        self.L.or_move_first(self.ctx, C.byref(a), C.byref(s))
        tr = [(a.value, s.value)]
        while tr[-1][0]:
            self.L.or_move_block(self.ctx, C.byref(a), C.byref(s))
            tr.append((a.value, s.value))
        return np.array(tr, np.int32)

    def first_gradient(self):
        """The reference's density gradient a of the first Denmax step (after density())."""
        a = np.empty(3 * self.box.n, np.float32)
        if self.L.or_first_gradient(self.ctx, a):
            raise RuntimeError("or_first_gradient failed")
        return a.reshape(-1, 3)

    def moved(self):
        r = np.empty(3 * self.box.n, np.float32)
        self.L.or_get_moved(self.ctx, r)
        return r.reshape(-1, 3)

    def fof(self):
        g = np.empty(self.box.n, np.int32)
        ng = C.c_int()
        self.L.or_fof(self.ctx, _ptr(g), C.byref(ng))
        return g, ng.value

    def microsteps(self):
        """The 5 micro-steps (main.c:431-438) after fof(); returns the nScatter
        of each Microstep line."""
        m = np.zeros(5, np.int32)
        if self.L.or_microsteps(self.ctx, _ptr(m)):
            raise RuntimeError("or_microsteps failed")
        return m

    def finish(self):
        """micro-steps, centres, unbinding and kdTooSmall after fof()."""
        s = C.c_int()
        self.micro = self.microsteps()
        self.L.or_centers(self.ctx)
        self.L.or_unbind(self.ctx, C.byref(s))
        self.L.or_too_small(self.ctx)
        return self.groups()

    def catalogue(self):
        """Per final group 0..nGroup-1 (kdWriteGroup's .gtp records, kd.c:1612-1687):
        mass, rCenter, vcm, rBound, radius, nMembers."""
        ng = self.groups()[1]
        out = dict(mass=np.zeros(ng, np.float32), center=np.zeros(3 * ng, np.float32),
                   vcm=np.zeros(3 * ng, np.float32), rbound=np.zeros(3 * ng, np.float32),
                   radius=np.zeros(ng, np.float32), nmembers=np.zeros(ng, np.int32))
        self.L.or_get_catalogue(self.ctx, *[_ptr(out[k]) for k in
                                            ("mass", "center", "vcm", "rbound", "radius", "nmembers")])
        for k in ("center", "vcm", "rbound"):
            out[k] = out[k].reshape(-1, 3)
        return out


def ref_density(box, **kw):
    p = params(**kw)
    R = ref()
    n = box.n
    b2 = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    ne = C.c_int()
    R.ref_density(C.byref(_fill(RefParams, p)), box.n_gas, box.n_dark, box.n_star, box.pos.reshape(-1),
                  box.mass, box.soft, box.temp, b2, rho, C.byref(ne))
    return b2, rho, ne.value


def ref_pipeline(box, max_trace=100000, max_group=1 << 20, gtp_path=None, **kw):
    """The reference's main.c:374-471 on box; with gtp_path, its kdWriteGroup
    writes the .gtp there (native byte order)."""
    p = params(**kw)
    R = ref()
    n = box.n
    out = {k: np.empty(n, np.float32) for k in ("fBall2", "fDensity")}
    out.update({k: np.empty(3 * n, np.float32) for k in ("rMoved", "rMicro")})
    out.update({k: np.empty(n, np.int32) for k in ("iFofGroup", "iGroup")})
    tr = np.zeros(2 * max_trace, np.int32)
    gm = np.zeros(max_group, np.float32)
    gc = np.zeros(3 * max_group, np.float32)
    gv = np.zeros(3 * max_group, np.float32)
    o = RefOut()
    for k in ("fBall2", "fDensity", "rMoved", "rMicro", "iFofGroup", "iGroup"):
        setattr(o, k, _ptr(out[k]))
    o.trace = _ptr(tr)
    o.maxTrace = max_trace
    o.grpMass, o.grpCenter, o.grpVcm = _ptr(gm), _ptr(gc), _ptr(gv)
    o.maxGroup = max_group
    o.gtpPath = gtp_path.encode() if gtp_path else None
    R.ref_pipeline(C.byref(_fill(RefParams, p)), box.n_gas, box.n_dark, box.n_star,
                   box.pos.reshape(-1), box.vel.reshape(-1), box.mass, box.soft, box.temp, C.byref(o))
    ng = o.nGroup
    out.update(nExtraScat=o.nExtraScat, nMove=o.nMove, nFofGroup=o.nFofGroup, nGroup=ng,
               trace=tr[: 2 * o.nIttr].reshape(-1, 2), grpMass=gm[:ng], grpCenter=gc[: 3 * ng].reshape(-1, 3),
               grpVcm=gv[: 3 * ng].reshape(-1, 3), micro=np.array(o.micro[:], np.int32))
    out["rMoved"] = out["rMoved"].reshape(-1, 3)
    out["rMicro"] = out["rMicro"].reshape(-1, 3)
    return out


def ref_unbind(box, grp, **kw):
    """The reference's -unbind path on caller groups (centres of mass); returns
    the final group ids after kdUnbind + kdTooSmall."""
    p = params(**kw)
    R = ref()
    out = np.empty(box.n, np.int32)
    ng = C.c_int()
    R.ref_unbind(C.byref(_fill(RefParams, p)), box.n_gas, box.n_dark, box.n_star, box.pos.reshape(-1),
                 box.vel.reshape(-1), box.mass, box.soft, box.temp, np.ascontiguousarray(grp, np.int32), out,
                 C.byref(ng))
    return out


def knn_brute(pos, k, period=1.0, periodic=True, lists=False):
    n = pos.shape[0]
    b2 = np.empty(n, np.float32)
    nb = np.empty(n * k, np.int32) if lists else None
    rc = lib().or_knn_brute(n, np.ascontiguousarray(pos, np.float32).reshape(-1), k, period, int(periodic),
                            _ptr(b2), _ptr(nb))
    if rc:
        raise RuntimeError("knn_brute failed")
    return (b2, nb.reshape(n, k)) if lists else b2


def fof_brute(pos, tau, period=1.0, periodic=True):
    n = pos.shape[0]
    g = np.empty(n, np.int32)
    ng = lib().or_fof_brute(n, np.ascontiguousarray(pos, np.float32).reshape(-1), tau, period, int(periodic), g)
    return g, ng


def fof_points(pos, tau, period=1.0, periodic=True, nbucket=16):
    n = pos.shape[0]
    g = np.empty(n, np.int32)
    ng = lib().or_fof_points(n, np.ascontiguousarray(pos, np.float32).reshape(-1), tau, period, int(periodic),
                             nbucket, g)
    return g, ng


def density_f64(pos, mass, ball2, period=1.0, periodic=True):
    n = pos.shape[0]
    rho = np.empty(n, np.float64)
    lib().or_density_f64(n, np.ascontiguousarray(pos, np.float32).reshape(-1), np.ascontiguousarray(mass, np.float32),
                         np.ascontiguousarray(ball2, np.float32), period, int(periodic), rho)
    return rho


def gradient_f64(spos, smass, sball2, qpos, period=1.0, periodic=True):
    nq = qpos.shape[0]
    acc = np.empty(3 * nq, np.float64)
    lib().or_gradient_f64(spos.shape[0], np.ascontiguousarray(spos, np.float32).reshape(-1),
                          np.ascontiguousarray(smass, np.float32), np.ascontiguousarray(sball2, np.float32), nq,
                          np.ascontiguousarray(qpos, np.float32).reshape(-1), period, int(periodic), acc)
    return acc.reshape(nq, 3)


def potential_f64(pos, mass, soft, G=1.0, soft_type=2):
    n = pos.shape[0]
    phi = np.empty(n, np.float64)
    lib().or_potential_f64(n, np.ascontiguousarray(pos, np.float32).reshape(-1), np.ascontiguousarray(mass, np.float32),
                           np.ascontiguousarray(soft, np.float32), G, soft_type, phi)
    return phi


def relabel_min_index(labels: np.ndarray) -> np.ndarray:
    """Renumber nonzero labels 1..G in order of each group's smallest member index."""
    lab = np.asarray(labels)
    out = np.zeros_like(lab)
    nz = np.nonzero(lab)[0]
    if nz.size == 0:
        return out
    uniq, first = np.unique(lab[nz], return_index=True)
    order = np.argsort(nz[first], kind="stable")
    newid = np.empty(uniq.size, np.int64)
    newid[order] = np.arange(1, uniq.size + 1)
    out[nz] = newid[np.searchsorted(uniq, lab[nz])]
    return out


def overlap_pairing(a: np.ndarray, b: np.ndarray) -> dict:
    """Greedy mutual-best pairing of group labels by member overlap, 0 <-> 0.

    Pairs (label_a, label_b) are taken in order of decreasing overlap count,
    each label used at most once; group 0 (not in a group) always maps to 0.
    Returns {label_a: label_b}.
    """
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    m = (a != 0) & (b != 0)
    key = a[m] * (int(b.max()) + 1) + b[m]
    uk, cnt = np.unique(key, return_counts=True)
    la, lb = uk // (int(b.max()) + 1), uk % (int(b.max()) + 1)
    order = np.lexsort((lb, la, -cnt))
    pa, used_b = {0: 0}, {0}
    for k in order:
        x, y = int(la[k]), int(lb[k])
        if x in pa or y in used_b:
            continue
        pa[x] = y
        used_b.add(y)
    return pa


def match_fraction(a: np.ndarray, b: np.ndarray) -> float:
    """(1/N) #{p : pi(a_p) == b_p} under overlap_pairing(a, b); unpaired labels never match."""
    pa = overlap_pairing(a, b)
    a = np.asarray(a)
    mapped = np.array([pa.get(int(x), -1) for x in np.unique(a)])
    lut = dict(zip(np.unique(a).tolist(), mapped.tolist()))
    am = np.vectorize(lut.get, otypes=[np.int64])(a) if a.size else a
    return float(np.mean(am == np.asarray(b))) if a.size else 1.0


def same_partition(a: np.ndarray, b: np.ndarray) -> bool:
    """True iff a and b are the same labelling up to renaming (0 fixed to 0)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape or not np.array_equal(a == 0, b == 0):
        return False
    return np.array_equal(relabel_min_index(a), relabel_min_index(b))


def unbind_f64(pos, vel, mass, soft, grp, centres, period=1.0, G=1.0, soft_type=2, fshift=1.0, fcosmo=0.0,
               z=0.0, scoop=0.0, sub_pot=True, max_members=None):
    """FP64 restatement of kdUnbind's removal (kd.c:1330-1447) for checking:
    per group, coordinates relative to the member with the largest index
    (wrapped), FP64 centre of mass, potentials from or_potential_f64 plus the
    scoop (group-0 particles within `scoop` of the group centre add G m K to
    every member, grav.c:63-135), then E = 1/2|fshift (v - vcm) + fcosmo (r - rcm)|^2
    - phi (1+z); while max E >= 0 remove the argmax (ties: smallest index), downdate
    the centre of mass and, if sub_pot, subtract the removed particle's potential.
    Returns the group ids after unbinding (removed -> 0).  Everything in float64."""
    from scipy.spatial import cKDTree
    pos = np.asarray(pos, np.float64)
    vel = np.asarray(vel, np.float64)
    mass = np.asarray(mass, np.float64)
    soft = np.asarray(soft, np.float32)
    out = np.array(grp, np.int32, copy=True)
    z0 = np.flatnonzero(out == 0)
    tree = cKDTree(np.mod(pos[z0] + 0.5 * period, period), boxsize=period) if (scoop > 0 and z0.size) else None
    order = np.argsort(out, kind="stable")
    bounds = np.searchsorted(out[order], np.arange(out.max() + 2))
    for g in range(1, out.max() + 1):
        mem = order[bounds[g]:bounds[g + 1]]
        n = mem.size
        if n == 0 or (max_members is not None and n >= max_members):
            continue
        rel = pos[mem.max()]
        d = pos[mem] - rel
        d -= period * np.round(d / period)
        phi = potential_f64(d, mass[mem], soft[mem], G=G, soft_type=soft_type)
        if tree is not None:
            c = np.asarray(centres[g], np.float64)
            hits = tree.query_ball_point(np.mod(c + 0.5 * period, period), scoop)
            for j in z0[hits]:
                dj = pos[j] - rel
                dj -= period * np.round(dj / period)
                phi += G * mass[j] * _kernel_f64(np.sqrt(((d - dj) ** 2).sum(1)), soft[mem] + soft[j], soft_type)
        m = mass[mem]
        v = vel[mem]
        alive = np.ones(n, bool)
        M = m.sum()
        rcm = (m[:, None] * d).sum(0) / M
        vcm = (m[:, None] * v).sum(0) / M
        zf = 1.0 + z
        while alive.any():
            u = fshift * (v - vcm) + fcosmo * (d - rcm)
            E = 0.5 * (u * u).sum(1) - phi * zf
            E[~alive] = -np.inf
            Emax = E.max()
            if not (Emax >= 0.0):
                break
            cand = np.flatnonzero(E == Emax)
            b = cand[np.argmin(mem[cand])]
            alive[b] = False
            out[mem[b]] = 0
            M2 = M - m[b]
            if M2 <= 0:
                break
            rcm = (rcm * M - m[b] * d[b]) / M2
            vcm = (vcm * M - m[b] * v[b]) / M2
            M = M2
            if sub_pot:
                rb = np.sqrt(((d - d[b]) ** 2).sum(1))
                phi -= G * m[b] * _kernel_f64(rb, soft[mem] + soft[mem[b]], soft_type)
    return out


def _kernel_f64(r, twoh, soft_type):
    """The softened 1/r of grav.h:11-31 (spline) or grav.c:25-27 (Plummer) in FP64."""
    r = np.asarray(r, np.float64)
    twoh = np.asarray(twoh, np.float64)
    if soft_type == 1:
        return 1.0 / np.sqrt(r * r + 0.25 * twoh * twoh)
    dih = 2.0 / twoh
    u = r * dih
    with np.errstate(divide="ignore", invalid="ignore"):
        inner = dih * (1.4 - (2.0 / 3.0) * u ** 2 + 0.3 * u ** 4 - 0.1 * u ** 5)
        mid = -1.0 / (15.0 * r) + dih * (1.6 - (4.0 / 3.0) * u ** 2 + u ** 3 - 0.3 * u ** 4 + u ** 5 / 30.0)
        far = 1.0 / r
    return np.where(r >= twoh, far, np.where(u < 1.0, inner, mid))


def group_stats_f32(pos, vel, mass, soft, grp, centres, vcms, ptype=None, rho=None, temp=None, G=1.0,
                    fexp=1.0, fexphub=0.0, period=1.0, dens_min=0.0, temp_max=3.4028235e38):
    """kdOutStats (kd.c:1703-1839) restated in the reference's own fp32/fp64 mix,
    for checking: per group g >= 1, members relative to centres[g] (wrapped at
    half a period), ordered by r^2 = ((x^2) + y^2) + z^2 in fp32 (ties by
    index, where the reference's qsort leaves them unordered), fp32 running
    masses, then the ten computed columns of the .stat line in the layout of
    skg_group_stats: total mass, gas mass, star mass, Vmax^2, V(r_half)^2,
    V(r_out)^2, r at Vmax, r_half, r_out^2, 1-D dispersion.  ptype by index
    (2 gas, 4 star, else dark).  Returns an (nGroup, 10) float64 array, row 0
    zero."""
    f = np.float32
    pos = np.asarray(pos, f).reshape(-1, 3)
    vel = np.asarray(vel, f).reshape(-1, 3)
    mass = np.asarray(mass, f)
    soft = np.asarray(soft, f)
    grp = np.asarray(grp)
    ng = int(grp.max()) + 1 if grp.size else 1
    hx = f(0.5) * f(period)
    Gf, fe, fh = f(G), f(fexp), f(fexphub)
    order = np.argsort(grp, kind="stable")
    bounds = np.searchsorted(grp[order], np.arange(ng + 1))
    out = np.zeros((ng, 10), np.float64)
    for g in range(1, ng):
        mem = order[bounds[g]:bounds[g + 1]]
        n = mem.size
        if n == 0:
            continue
        d = pos[mem] - np.asarray(centres[g], f)
        d = np.where(d > hx, d - f(2) * hx, d)
        d = np.where(d <= -hx, d + f(2) * hx, d)
        r2 = ((f(0) + d[:, 0] * d[:, 0]) + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        srt = np.argsort(r2, kind="stable")
        mem, d, r2 = mem[srt], d[srt], r2[srt]
        m = mass[mem]
        half = np.cumsum(f(0.5) * m, dtype=f)[-1]
        M = np.cumsum(m, dtype=f)
        r = np.sqrt(r2.astype(np.float64))
        sf = soft[mem].astype(np.float64)
        v = (Gf * M).astype(np.float64) / np.where(r > 0, r, np.inf)
        ok = (r2.astype(np.float64) > 4.0 * sf * sf) & (v > 0)
        vcirc = f(0)
        rvmax = f(0)
        for j in np.flatnonzero(ok):                              # the running test against the stored fp32 value
            if v[j] > np.float64(vcirc):
                rvmax = f(r[j])
                vcirc = f(Gf * M[j]) / rvmax
        jh = np.flatnonzero(M > half)
        rhalf = vhalf = f(0)
        if jh.size:
            j = int(jh[0])
            rhalf = f(r[j])
            vhalf = f(Gf * M[j]) / rhalf
        vout = f(np.float64(Gf * M[-1]) / r[-1]) if r[-1] > 0 else f(0)
        if vcirc == 0:
            vcirc = vout
            rvmax = f(r[-1])
        pt = np.asarray(ptype)[mem] if ptype is not None else np.ones(n, np.int32)
        gas_sel = pt == 2
        if rho is not None:
            gas_sel &= np.asarray(rho, f)[mem] >= f(dens_min)
        if temp is not None:
            gas_sel &= np.asarray(temp, f)[mem] <= f(temp_max)
        gas = np.cumsum(m[gas_sel], dtype=f)[-1] if gas_sel.any() else f(0)
        star = np.cumsum(m[pt == 4], dtype=f)[-1] if (pt == 4).any() else f(0)
        dv = fe * (vel[mem] - np.asarray(vcms[g], f)) + fh * d
        vd = np.cumsum((dv * dv).reshape(-1), dtype=f)[-1]
        out[g] = [M[-1], gas, star, vcirc, vhalf, vout, rvmax, rhalf, r2[-1], f(np.sqrt(np.float64(vd) / (3.0 * n)))]
    return out


def read_stat(path):
    """A .stat file (kd.c:1827-1836) as an (nGroup - 1, 21) float64 array."""
    rows = [list(map(float, line.split())) for line in open(path) if line.strip()]
    return np.array(rows, np.float64).reshape(-1, 21)


def read_gtp(path, xdr=True):
    """A .gtp file (kdWriteGroup, kd.c:1612-1687): per group (mass, rCenter,
    vcm, radius) from the tipsy star records after the 32-byte header."""
    raw = np.fromfile(path, np.uint8)
    rec = np.frombuffer(raw[32:].tobytes(), ">f4" if xdr else "<f4").reshape(-1, 11).astype(np.float32)
    return dict(mass=rec[:, 0], pos=rec[:, 1:4], vel=rec[:, 4:7], radius=rec[:, 9])   # metals, tform, eps, phi


def read_grp(path):
    """A .grp file (kdOutGroup, kd.c:1502-1524): group id by iOrder."""
    with open(path) as f:
        n = int(f.readline())
        g = np.loadtxt(f, dtype=np.int64, max_rows=n)
    return g.reshape(-1)
