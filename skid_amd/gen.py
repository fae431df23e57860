"""Seeded synthetic particle boxes for the SKID group-finding workloads.

Every box is a periodic unit cube centred on 0 (positions in (-0.5, 0.5]) with
equal-mass particles of mass 1/N, so the mean density is 1.  Halos are NFW
spheres truncated at r200 (mean enclosed density 200), sampled by the exact
inverse of the NFW mass profile mu(x) = ln(1+x) - x/(1+x), which has the closed
form x = -1/W0(-exp(-1-p)) - 1 (W0 = principal Lambert W).  Halo members get
isotropic Gaussian velocities with sigma = sqrt(G*M/(2*r200)), G = 1; the
uniform background is at rest.  Softening is 0.0288 * N**(-1/3), which is
dark.std's 9e-4 at N = 32768.

Choices the workload description leaves open, fixed here:
  * halo particle counts: the drawn halo masses are normalised so that the halos
    hold the stated fraction of N; M = n/N is then used for r200 and sigma;
  * gas (two-species box) traces each halo with a cored isothermal profile
    rho ~ 1/(r^2 + rc^2), rc = 0.1*r200, truncated at r200, and the same
    velocity dispersion; gas temperature is log-normal(ln 1e4, 1);
  * the multi-box workload uses tau = 0.16 * N_b**(-1/3) and nSmooth = 32.
"""
from __future__ import annotations

import dataclasses
import numpy as np
from scipy.special import lambertw

G_DEFAULT = 1.0


@dataclasses.dataclass
class Box:
    """Particles in tipsy species order: gas first, then dark, then stars."""
    pos: np.ndarray           # (N, 3) float32
    vel: np.ndarray           # (N, 3) float32
    mass: np.ndarray          # (N,) float32
    soft: np.ndarray          # (N,) float32
    temp: np.ndarray          # (N,) float32 (0 for dark)
    n_gas: int
    n_dark: int
    n_star: int = 0
    halo_id: np.ndarray | None = None   # (N,) int32, -1 for background

    @property
    def n(self) -> int:
        return self.n_gas + self.n_dark + self.n_star


def _wrap(x: np.ndarray) -> np.ndarray:
    """Wrap float64 coordinates into (-0.5, 0.5] and round to float32."""
    x = x - np.floor(x + 0.5)          # [-0.5, 0.5)
    x = np.where(x <= -0.5, x + 1.0, x)
    x32 = x.astype(np.float32)
    x32 = np.where(x32 <= np.float32(-0.5), x32 + np.float32(1.0), x32)
    x32 = np.where(x32 > np.float32(0.5), x32 - np.float32(1.0), x32)
    return x32.astype(np.float32)


def _iso_dirs(rng: np.random.Generator, n: int) -> np.ndarray:
    cz = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    sz = np.sqrt(1.0 - cz * cz)
    return np.stack([sz * np.cos(ph), sz * np.sin(ph), cz], axis=1)


def nfw_radii(rng: np.random.Generator, n: int, c: float) -> np.ndarray:
    """Radii in units of r200 for an NFW halo of concentration c (truncated at r200)."""
    mu_c = np.log1p(c) - c / (1.0 + c)
    p = rng.uniform(0.0, 1.0, n) * mu_c
    w = lambertw(-np.exp(-1.0 - p), 0).real
    x = -1.0 / w - 1.0
    return np.clip(x / c, 0.0, 1.0)


def cored_iso_radii(rng: np.random.Generator, n: int, rc: float = 0.1) -> np.ndarray:
    """Radii in units of r200 for rho ~ 1/(r^2+rc^2) truncated at 1."""
    grid = np.linspace(0.0, 1.0, 4097)
    m = grid - rc * np.arctan(grid / rc)
    m /= m[-1]
    return np.interp(rng.uniform(0.0, 1.0, n), m, grid)


def r200_of(m_frac: np.ndarray | float) -> np.ndarray:
    """r200 for a halo of mass fraction m of a unit-mass unit box (mean density 1)."""
    return np.cbrt(3.0 * np.asarray(m_frac, dtype=np.float64) / (800.0 * np.pi))


def _halo_counts(masses: np.ndarray, n_halo_total: int) -> np.ndarray:
    share = masses / masses.sum() * n_halo_total
    cnt = np.floor(share).astype(np.int64)
    rem = n_halo_total - int(cnt.sum())
    if rem > 0:
        order = np.argsort(-(share - cnt), kind="stable")
        cnt[order[:rem]] += 1
    return cnt


def nfw_box(n: int, n_halos: int, bg_frac: float, seed: int, *,
            mass_range=(1e-4, 1e-2), mass_slope: float | None = None,
            conc_range=(5.0, 15.0), G: float = G_DEFAULT) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Dark-matter NFW + uniform background box.

    mass_slope None draws halo masses log-uniform over mass_range; otherwise from
    dN/dM ~ M**mass_slope.  Returns (pos float32 (n,3), vel float32 (n,3), halo_id int32).
    """
    rng = np.random.default_rng(seed)
    lo, hi = mass_range
    if mass_slope is None:
        masses = np.exp(rng.uniform(np.log(lo), np.log(hi), n_halos))
    else:
        a = mass_slope + 1.0
        u = rng.uniform(0.0, 1.0, n_halos)
        masses = (lo ** a + u * (hi ** a - lo ** a)) ** (1.0 / a)
    n_bg = int(round(bg_frac * n))
    counts = _halo_counts(masses, n - n_bg)
    conc = rng.uniform(conc_range[0], conc_range[1], n_halos)
    centres = rng.uniform(-0.5, 0.5, (n_halos, 3))
    pos = np.empty((n, 3), np.float64)
    vel = np.zeros((n, 3), np.float64)
    hid = np.full(n, -1, np.int32)
    k = 0
    for h in range(n_halos):
        m = int(counts[h])
        if m == 0:
            continue
        mf = m / n
        r200 = float(r200_of(mf))
        r = nfw_radii(rng, m, float(conc[h])) * r200
        pos[k:k + m] = centres[h] + _iso_dirs(rng, m) * r[:, None]
        sig = np.sqrt(G * mf / (2.0 * r200))
        vel[k:k + m] = rng.normal(0.0, sig, (m, 3))
        hid[k:k + m] = h
        k += m
    pos[k:] = rng.uniform(-0.5, 0.5, (n - k, 3))
    perm = rng.permutation(n)
    return _wrap(pos[perm]), vel[perm].astype(np.float32), hid[perm]


def _finish(pos, vel, hid, n_gas, n_dark, temp=None) -> Box:
    n = pos.shape[0]
    mass = np.full(n, np.float32(1.0 / n), np.float32)
    soft = np.full(n, np.float32(0.0288 * n ** (-1.0 / 3.0)), np.float32)
    t = np.zeros(n, np.float32) if temp is None else temp.astype(np.float32)
    return Box(pos=np.ascontiguousarray(pos, np.float32), vel=np.ascontiguousarray(vel, np.float32),
               mass=mass, soft=soft, temp=t, n_gas=n_gas, n_dark=n_dark, halo_id=hid)


def config1(n: int = 32768, seed: int = 1) -> Box:
    """Plain SKID small box: 70% in 40 NFW halos, 30% background."""
    pos, vel, hid = nfw_box(n, 40, 0.30, seed)
    return _finish(pos, vel, hid, 0, n)


def config2(n: int = 1 << 20, seed: int = 2) -> Box:
    """Density/adjoint box: 500 NFW halos, 20% background."""
    pos, vel, hid = nfw_box(n, 500, 0.20, seed)
    return _finish(pos, vel, hid, 0, n)


def config3(n_each: int = 1 << 24, seed: int = 3, n_halos: int = 8000) -> Box:
    """Two-species box: n_each gas + n_each dark, 8000 halos, dN/dM ~ M^-1.9 on [1e-7, 1e-3]."""
    pos_d, vel_d, hid_d = nfw_box(n_each, n_halos, 0.30, seed, mass_range=(1e-7, 1e-3), mass_slope=-1.9)
    # Gas traces the same halos: re-draw the halo catalogue with the same seed.
    rng = np.random.default_rng(seed)
    a = -0.9
    u = rng.uniform(0.0, 1.0, n_halos)
    masses = ((1e-7) ** a + u * ((1e-3) ** a - (1e-7) ** a)) ** (1.0 / a)
    rng.uniform(5.0, 15.0, n_halos)
    centres = rng.uniform(-0.5, 0.5, (n_halos, 3))
    grng = np.random.default_rng(seed + 7919)
    n_bg = int(round(0.30 * n_each))
    counts = _halo_counts(masses, n_each - n_bg)
    n_tot = 2 * n_each
    pos_g = np.empty((n_each, 3), np.float64)
    vel_g = np.zeros((n_each, 3), np.float64)
    hid_g = np.full(n_each, -1, np.int32)
    k = 0
    for h in range(n_halos):
        m = int(counts[h])
        if m == 0:
            continue
        mf = 2.0 * m / n_tot
        r200 = float(r200_of(mf))
        r = cored_iso_radii(grng, m) * r200
        pos_g[k:k + m] = centres[h] + _iso_dirs(grng, m) * r[:, None]
        sig = np.sqrt(mf / (2.0 * r200))
        vel_g[k:k + m] = grng.normal(0.0, sig, (m, 3))
        hid_g[k:k + m] = h
        k += m
    pos_g[k:] = grng.uniform(-0.5, 0.5, (n_each - k, 3))
    perm = grng.permutation(n_each)
    pos_g, vel_g, hid_g = _wrap(pos_g[perm]), vel_g[perm].astype(np.float32), hid_g[perm]
    temp = np.concatenate([np.exp(grng.normal(np.log(1e4), 1.0, n_each)), np.zeros(n_each)])
    pos = np.concatenate([pos_g, pos_d])
    vel = np.concatenate([vel_g, vel_d])
    hid = np.concatenate([hid_g, hid_d])
    return _finish(pos, vel, hid, n_each, n_each, temp)


def config4_sizes(n_boxes: int = 256, seed: int = 4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(2048), np.log(524288), n_boxes)).astype(np.int64)


def config4_box(n: int, b: int) -> Box:
    """One of the independent boxes of the multi-box workload (seed 1000+b)."""
    pos, vel, hid = nfw_box(int(n), max(1, int(n) // 4000), 0.25, 1000 + b)
    return _finish(pos, vel, hid, 0, int(n))


def config4_tau(n: int) -> float:
    return 0.16 * float(n) ** (-1.0 / 3.0)


def write_tipsy(path: str, box: Box, time: float = 1.0, xdr: bool = False) -> None:
    """Write a tipsy binary (native little-endian, or XDR big-endian with xdr=True)."""
    n = box.n
    end = ">" if xdr else "<"
    with open(path, "wb") as f:
        if xdr:
            hdr = np.array([(time, n, 3, box.n_gas, box.n_dark, box.n_star, 0)],
                           dtype=[("t", ">f8"), ("n", ">i4"), ("d", ">i4"), ("g", ">i4"),
                                  ("k", ">i4"), ("s", ">i4"), ("pad", ">i4")])
        else:
            hdr = np.array([(time, n, 3, box.n_gas, box.n_dark, box.n_star, 0)],
                           dtype=[("t", "<f8"), ("n", "<i4"), ("d", "<i4"), ("g", "<i4"),
                                  ("k", "<i4"), ("s", "<i4"), ("pad", "<i4")])
        f.write(hdr.tobytes())
        g = slice(0, box.n_gas)
        if box.n_gas:
            rec = np.zeros((box.n_gas, 12), np.float32)
            rec[:, 0] = box.mass[g]
            rec[:, 1:4] = box.pos[g]
            rec[:, 4:7] = box.vel[g]
            rec[:, 8] = box.temp[g]
            rec[:, 9] = box.soft[g]
            f.write(rec.astype(end + "f4").tobytes())
        d = slice(box.n_gas, box.n_gas + box.n_dark)
        if box.n_dark:
            rec = np.zeros((box.n_dark, 9), np.float32)
            rec[:, 0] = box.mass[d]
            rec[:, 1:4] = box.pos[d]
            rec[:, 4:7] = box.vel[d]
            rec[:, 7] = box.soft[d]
            f.write(rec.astype(end + "f4").tobytes())
        s = slice(box.n_gas + box.n_dark, n)
        if box.n_star:
            rec = np.zeros((box.n_star, 11), np.float32)
            rec[:, 0] = box.mass[s]
            rec[:, 1:4] = box.pos[s]
            rec[:, 4:7] = box.vel[s]
            rec[:, 9] = box.soft[s]
            f.write(rec.astype(end + "f4").tobytes())


def read_tipsy(path: str, xdr: bool = False) -> Box:
    """Read a tipsy binary written by SKID tools (native or XDR)."""
    raw = open(path, "rb").read()
    e = ">" if xdr else "<"
    t = np.frombuffer(raw[:8], e + "f8")[0]
    nb, nd, ng, nk, ns = np.frombuffer(raw[8:28], e + "i4")
    off = 32 if not xdr else 32
    ng, nk, ns = int(ng), int(nk), int(ns)
    gas = np.frombuffer(raw[off:off + 48 * ng], e + "f4").reshape(ng, 12).astype(np.float32)
    off += 48 * ng
    dk = np.frombuffer(raw[off:off + 36 * nk], e + "f4").reshape(nk, 9).astype(np.float32)
    off += 36 * nk
    st = np.frombuffer(raw[off:off + 44 * ns], e + "f4").reshape(ns, 11).astype(np.float32)
    pos = np.concatenate([gas[:, 1:4], dk[:, 1:4], st[:, 1:4]])
    vel = np.concatenate([gas[:, 4:7], dk[:, 4:7], st[:, 4:7]])
    mass = np.concatenate([gas[:, 0], dk[:, 0], st[:, 0]])
    soft = np.concatenate([gas[:, 9], dk[:, 7], st[:, 9]])
    temp = np.concatenate([gas[:, 8], np.zeros(nk + ns, np.float32)])
    del t
    return Box(pos=np.ascontiguousarray(pos), vel=np.ascontiguousarray(vel), mass=mass.copy(),
               soft=soft.copy(), temp=temp.astype(np.float32), n_gas=ng, n_dark=nk, n_star=ns)


def config5(n: int = 1 << 27, n_groups: int = 60000, seed: int = 5, n_range=(8, 2_000_000),
            slope: float = -1.9, q_range=(0.5, 2.0), conc_range=(5.0, 15.0), G: float = G_DEFAULT):
    """Unbinding + precision stress: n_groups NFW spheres with membership drawn
    from dN/dn ~ n**slope on n_range, velocities isotropic Gaussian with
    sigma^2 = Q G M / (2 r200) for a virial ratio Q uniform on q_range (so a
    real fraction is unbound), and the rest of n a uniform background (group 0).
    Returns (box, grp): grp[i] is particle i's group id (1..n_groups) or 0."""
    rng = np.random.default_rng(seed)
    lo, hi = n_range
    a = slope + 1.0
    u = rng.uniform(0.0, 1.0, n_groups)
    sizes = np.floor((lo ** a + u * (hi ** a - lo ** a)) ** (1.0 / a)).astype(np.int64)
    if sizes.sum() > n:
        sizes = np.maximum(lo, (sizes * (n / sizes.sum())).astype(np.int64))
    conc = rng.uniform(conc_range[0], conc_range[1], n_groups)
    q = rng.uniform(q_range[0], q_range[1], n_groups)
    centres = rng.uniform(-0.5, 0.5, (n_groups, 3))
    pos = np.empty((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32)
    grp = np.zeros(n, np.int32)
    k = 0
    for g in range(n_groups):
        m = int(sizes[g])
        mf = m / n
        r200 = float(r200_of(mf))
        r = nfw_radii(rng, m, float(conc[g])) * r200
        pos[k:k + m] = _wrap(centres[g] + _iso_dirs(rng, m) * r[:, None])
        sig = np.sqrt(q[g] * G * mf / (2.0 * r200))
        vel[k:k + m] = rng.normal(0.0, sig, (m, 3))
        grp[k:k + m] = g + 1
        k += m
    pos[k:] = rng.uniform(-0.5, 0.5, (n - k, 3)).astype(np.float32)
    pos[k:] = np.where(pos[k:] <= np.float32(-0.5), np.float32(0.5), pos[k:])
    mass = np.full(n, 1.0 / n, np.float32)
    soft = np.full(n, 0.0288 * n ** (-1.0 / 3.0), np.float32)
    box = Box(pos=pos, vel=vel, mass=mass, soft=soft, temp=np.zeros(n, np.float32), n_gas=0, n_dark=n,
              halo_id=grp - 1)
    return box, grp


def config5_cosmology(z: float = 0.5, omega0: float = 0.3, lam: float = 0.7, h0: float = 2.8944):
    """fShift = 1/(1+z) and fCosmo = fShift * H(fShift) for a flat Lambda model,
    H(a) = H0 sqrt(Om a + Ok a^2 + OL a^4) / a^2 in FP64 (cosmo.c:46-58, no
    radiation or quintessence)."""
    a = 1.0 / (1.0 + z)
    ok = 1.0 - omega0 - lam
    H = h0 * np.sqrt(omega0 * a + ok * a * a + lam * a ** 4) / (a * a)
    return a, a * H
