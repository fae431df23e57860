"""Pins of the oracle by what mathematics fixes (closed forms, symmetry, brute
force on tiny inputs), chosen so that a dropped term, a wrong sign or index or
a transposed operand fails one of them.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen


def lattice(n_side: int, jitter: float = 0.0, seed: int = 0):
    """Cubic lattice in the periodic unit box; spacing 1/n_side (exact in f32)."""
    g = (np.arange(n_side, dtype=np.float64) + 0.5) / n_side - 0.5
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    if jitter:
        pos += np.random.default_rng(seed).uniform(-jitter, jitter, pos.shape)
    return pos.astype(np.float32)


def box_of(pos, mass=None):
    n = pos.shape[0]
    m = np.full(n, 1.0 / n, np.float32) if mass is None else mass.astype(np.float32)
    return gen.Box(pos=pos, vel=np.zeros_like(pos), mass=m, soft=np.full(n, 1e-3, np.float32),
                   temp=np.zeros(n, np.float32), n_gas=0, n_dark=n)


def m4(q):
    return np.where(q < 1, 1 - 1.5 * q * q + 0.75 * q ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))


@pytest.mark.parametrize("k,shells", [(7, [(6, 1.0)]), (19, [(6, 1.0), (12, math.sqrt(2))])])
def test_lattice_density_closed_form(k, shells):
    """On a lattice the k-th neighbour distance is a shell radius, so fBall2 and
    rho = sum over shells of (count * m W(r/h) / (pi h^3)) (gather + scatter halves)."""
    ns = 8
    pos = lattice(ns)
    d = 1.0 / ns
    b = box_of(pos)
    b2, rho, _ = O.Oracle(b, nSmooth=k, fTau=0.01, fPeriod=1.0).density()
    r_k = shells[-1][1] * d
    assert np.allclose(b2, np.float32(r_k) ** 2, rtol=2e-7)
    h = r_k / 2.0
    m = 1.0 / pos.shape[0]
    expect = m / (math.pi * h ** 3) * (1.0 + sum(c * float(m4(np.array(r * d / h))) for c, r in shells))
    assert np.allclose(rho, expect, rtol=2e-6)
    # and the FP64 oracle density from the same radii
    assert np.allclose(O.density_f64(pos, b.mass, b2), expect, rtol=1e-10)


def test_isolated_pair_density():
    """Two particles, k=2: h = d/2, the partner sits at q=2 (W=0): rho = m/(pi h^3)."""
    pos = np.array([[-0.1, 0.0, 0.0], [0.15, 0.05, 0.0]], np.float32)
    b = box_of(pos, mass=np.array([0.3, 0.7]))
    b2, rho, _ = O.Oracle(b, nSmooth=2, fTau=0.01).density()
    d2 = float(np.sum((pos[0].astype(np.float64) - pos[1]) ** 2))
    assert np.allclose(b2, d2, rtol=1e-6)
    h = math.sqrt(d2) / 2
    assert np.allclose(rho, b.mass / (math.pi * h ** 3), rtol=1e-5)


def test_gradient_vanishes_on_lattice():
    """By symmetry the frozen-field gradient is zero at lattice points."""
    pos = lattice(8)
    b = box_of(pos)
    b2, _, _ = O.Oracle(b, nSmooth=19, fTau=0.01, fPeriod=1.0).density()
    a = O.gradient_f64(pos, b.mass, b2, pos[:50])
    one = O.gradient_f64(pos, b.mass, b2, pos[:1] + np.float32([0.01, 0, 0]))
    assert np.abs(a).max() < 1e-9 * np.abs(one).max()
    # off-lattice the field pulls back towards the nearest point (a points away)
    assert one[0, 0] > 0


def test_gradient_matches_density_derivative():
    """a(x) = -grad rho_scatter(x): compare with central differences of the FP64 field."""
    b = gen.config1(n=2048, seed=7)
    b2 = O.knn_brute(b.pos, 32)
    x0 = b.pos[:20].astype(np.float64) + 1e-3
    a = O.gradient_f64(b.pos, b.mass, b2, x0.astype(np.float32))

    def field(x):
        d = b.pos.astype(np.float64)[None, :, :] - x[:, None, :]
        d -= np.round(d)
        r = np.sqrt((d ** 2).sum(-1))
        h = 0.5 * np.sqrt(b2.astype(np.float64))
        return (b.mass * m4(r / h) / (math.pi * h ** 3) * (r * r < b2)).sum(-1)

    eps = 1e-6
    for c in range(3):
        e = np.zeros(3)
        e[c] = eps
        g = (field(x0 + e) - field(x0 - e)) / (2 * eps)
        assert np.allclose(-g, a[:, c], rtol=1e-4, atol=1e-3 * np.abs(a).max())


def test_spline_potential_identities():
    """Hernquist & Katz spline: equals 1/r for r >= 2h, continuous at u=1 and u=2,
    and finite 7/(5h)*... at r=0 (value dih*7/5)."""
    twoh = np.float32(0.02)
    for r in (0.02, 0.03, 0.5):
        assert math.isclose(O.lib().or_spline_pot(np.float32(r) ** 2, twoh), 1.0 / r, rel_tol=1e-6)
    for r in (0.01, 0.02):   # u = 1 and u = 2
        lo = O.lib().or_spline_pot(np.float32(r * (1 - 1e-6)) ** 2, twoh)
        hi = O.lib().or_spline_pot(np.float32(r * (1 + 1e-6)) ** 2, twoh)
        assert math.isclose(lo, hi, rel_tol=1e-4)
    assert math.isclose(O.lib().or_spline_pot(0.0, twoh), 2.0 / 0.02 * 1.4, rel_tol=1e-6)


@pytest.mark.parametrize("soft_type", [1, 2])
def test_uniform_sphere_binding_energy(soft_type):
    """W = 1/2 sum m phi of a uniform sphere is -3GM^2/(5R) (tiny softening)."""
    rng = np.random.default_rng(3)
    n = 3000
    u = rng.uniform(0, 1, n) ** (1 / 3)
    v = rng.normal(size=(n, 3))
    pos = (v / np.linalg.norm(v, axis=1)[:, None] * u[:, None] * 0.1).astype(np.float32)
    m = np.full(n, 1.0 / n, np.float32)
    phi = O.potential_f64(pos, m, np.full(n, 1e-5, np.float32), 1.0, soft_type)
    W = -0.5 * float(np.sum(m * phi))
    assert math.isclose(W, -3.0 / (5.0 * 0.1), rel_tol=0.03)


def test_plummer_potential_of_a_pair():
    pos = np.array([[0, 0, 0], [0.3, 0.4, 0]], np.float32)
    m = np.array([2.0, 3.0], np.float32)
    eps = np.array([0.1, 0.2], np.float32)
    phi = O.potential_f64(pos, m, eps, 1.0, 1)
    expect = 1.0 / math.sqrt(0.25 + 0.25 * 0.3 ** 2)
    assert np.allclose(phi, [3.0 * expect, 2.0 * expect], rtol=1e-6)


def test_brute_knn_tiny():
    pos = np.array([[0, 0, 0], [0.1, 0, 0], [0.3, 0, 0], [-0.45, 0, 0], [0.45, 0, 0]], np.float32)
    b2, nb = O.knn_brute(pos, 3, period=1.0, periodic=True, lists=True)
    # through the boundary particle 3 (-0.45) is 0.1 from 4 (0.45) and 0.25 from 2 (0.3)
    assert list(nb[3]) == [3, 4, 2]
    assert math.isclose(b2[3], 0.25 ** 2, rel_tol=1e-5)
    assert list(nb[0]) == [0, 1, 2]
    b2o, nbo = O.knn_brute(pos, 3, periodic=False, lists=True)
    assert list(nbo[3]) == [3, 0, 1]


def test_fof_brute_chains():
    tau = 0.01
    x = np.arange(20) * 0.9 * tau - 0.09
    pos = np.stack([x, np.zeros(20), np.zeros(20)], 1).astype(np.float32)
    g, ng = O.fof_brute(pos, tau)
    assert ng == 1 and (g == 1).all()
    x2 = np.arange(20) * 1.1 * tau
    g2, ng2 = O.fof_brute(np.stack([x2, 0 * x2, 0 * x2], 1).astype(np.float32), tau)
    assert ng2 == 20 and list(g2) == list(range(1, 21))
    # a chain through the periodic boundary joins the two ends
    x3 = np.array([0.496, 0.4995, -0.4995, -0.496], np.float32)
    g3, ng3 = O.fof_brute(np.stack([x3, 0 * x3, 0 * x3], 1), 0.004, period=1.0)
    assert ng3 == 1
    g4, ng4 = O.fof_brute(np.stack([x3, 0 * x3, 0 * x3], 1), 0.004, periodic=False)
    assert ng4 == 2


def test_fof_points_matches_brute():
    b = gen.config1(n=4096, seed=9)
    for tau in (0.004, 0.012):
        g1, _ = O.fof_points(b.pos, tau)
        g2, _ = O.fof_brute(b.pos, tau)
        assert O.same_partition(g1, g2)


def test_match_fraction():
    a = np.array([0, 1, 1, 2, 2, 2, 3, 3])
    assert O.match_fraction(a, a) == 1.0
    b = np.array([0, 7, 7, 4, 4, 4, 9, 9])
    assert O.match_fraction(a, b) == 1.0
    c = np.array([0, 7, 4, 4, 4, 4, 9, 9])
    assert O.match_fraction(a, c) == 7 / 8


def test_generator_deterministic():
    a = gen.config1(n=4096, seed=1)
    b = gen.config1(n=4096, seed=1)
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.vel, b.vel)
    assert a.pos.min() > -0.5 and a.pos.max() <= 0.5
    assert a.mass.dtype == np.float32 and math.isclose(a.mass.sum(), 1.0, rel_tol=1e-5)


# ---- the FP64 unbinding restatement (oracle.unbind_f64) -------------------------
def test_unbind_f64_hand_cases():
    m = np.array([0.5, 0.5, 0.5], np.float32)
    sf = np.full(3, 1e-4, np.float32)
    pos = np.array([[0, 0, 0], [0.01, 0, 0], [0.3, 0, 0]], np.float32)
    g = np.array([1, 1, 1], np.int32)
    # at rest: everything bound
    assert np.array_equal(O.unbind_f64(pos, np.zeros((3, 3), np.float32), m, sf, g, {}), [1, 1, 1])
    # the distant member flies off: removed first and alone, the close pair stays
    vel = np.array([[0, 0, 0], [0, 0, 0], [0, 20, 0]], np.float32)
    assert np.array_equal(O.unbind_f64(pos, vel, m, sf, g, {}), [1, 1, 0])
    # E exactly as kd.c:1394-1399 with fshift / fcosmo: a pure Hubble flow
    # fcosmo * (r - rcm) unbinds the wide pair but not the close one
    vel0 = np.zeros((3, 3), np.float32)
    out = O.unbind_f64(pos, vel0, m, sf, g, {}, fcosmo=30.0)
    assert out[2] == 0 and out[0] == out[1] == 1


def test_unbind_f64_virial_spheres():
    rng = np.random.default_rng(2)
    n = 600
    r = gen.nfw_radii(rng, n, 10.0) * 0.05
    d = rng.normal(size=(n, 3))
    pos = (d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]).astype(np.float32)
    m = np.full(n, 1.0 / n, np.float32)
    sf = np.full(n, 1e-3, np.float32)
    g = np.ones(n, np.int32)
    sig = np.sqrt(1.0 / (2 * 0.05))
    fr = []
    for q in (0.5, 1.0, 2.0, 4.0):
        vel = rng.normal(0, np.sqrt(q) * sig, (n, 3)).astype(np.float32)
        fr.append((O.unbind_f64(pos, vel, m, sf, g, {}) > 0).mean())
    # bound fraction falls with the virial ratio: all but a few bound at 0.5,
    # a real fraction unbound at 2, nearly all at 4
    assert fr[0] > 0.98 and fr[0] >= fr[1] >= fr[2] >= fr[3], fr
    assert 0.2 < fr[2] < 0.9 and fr[3] < 0.1, fr


def test_unbind_f64_matches_reference_oracle(box1):
    """On config 1 (no scoop: every particle is in a group) the FP64 removal
    gives the catalogue of the reference-faithful oracle (bit-exact to the
    reference) to >= 99.9% of particles."""
    o = O.Oracle(box1, nSmooth=32, fTau=0.005, fPeriod=1.0)
    o.density()
    o.move_loop()
    fg, _ = o.fof()
    og, _ = o.finish()
    u = O.unbind_f64(box1.pos, box1.vel, box1.mass, box1.soft, fg, {})
    cnt = np.bincount(u)
    small = np.flatnonzero(cnt < 8)
    u[np.isin(u, small[small > 0])] = 0          # kdTooSmall with nMembers 8
    frac = O.match_fraction(u, og)
    print(f"FP64 unbinding vs reference oracle: match {frac:.5f}")
    assert frac >= 0.999
