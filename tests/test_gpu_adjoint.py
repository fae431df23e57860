"""Adjoint of the M4 density (skg_density_adjoint) and the frozen-field
gradient (skg_field_gradient): central finite differences in FP64 and the
reference's own first-step gradient (smooth1.c:447-459)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen, skg

pytestmark = pytest.mark.gpu


def m4(q):
    return np.where(q < 1, 1 - 1.5 * q * q + 0.75 * q ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))


def frozen_L(x, h, mass, gbar, nbr, parts=3):
    """L = sum_j gbar_j rho_j with the kNN sets frozen (head excluded), FP64."""
    i = np.repeat(np.arange(x.shape[0]), nbr.shape[1])
    j = nbr.reshape(-1)
    d = x[i] - x[j]
    d -= np.round(d)
    r = np.sqrt((d * d).sum(1))
    W = m4(r / h[i]) / (math.pi * h[i] ** 3)
    coef = (gbar[i] * mass[j] if parts & 1 else 0.0) + (gbar[j] * mass[i] if parts & 2 else 0.0)
    return float(np.sum(0.5 * W * coef))


@pytest.mark.parametrize("parts", [1, 2, 3])
def test_adjoint_central_differences(parts):
    b = gen.config1(n=4096, seed=13)
    k = 32
    ctx = skg.Context(b, skg.make_params(0.01, k, period=1.0))
    b2, rho, _ = ctx.density()
    rng = np.random.default_rng(parts)
    gbar = rng.normal(size=b.n).astype(np.float32)
    dx, dh = ctx.adjoint(gbar, parts)
    _, nb = O.knn_brute(b.pos, k, lists=True)
    nbr = nb[:, :-1]
    x0 = b.pos.astype(np.float64)
    h = 0.5 * np.sqrt(b2.astype(np.float64))
    m = b.mass.astype(np.float64)
    g = gbar.astype(np.float64)
    scale_x = np.abs(dx).max()
    scale_h = np.abs(dh).max()
    errs = []
    for t in range(12):
        p, c = int(rng.integers(b.n)), int(rng.integers(3))
        st = 1e-3 * h[p]
        xp, xm = x0.copy(), x0.copy()
        xp[p, c] += st
        xm[p, c] -= st
        fd = (frozen_L(xp, h, m, g, nbr, parts) - frozen_L(xm, h, m, g, nbr, parts)) / (2 * st)
        errs.append(abs(fd - dx[p, c]) / scale_x)
        assert abs(fd - dx[p, c]) <= 2e-5 * scale_x + 1e-4 * abs(fd), (p, c, fd, dx[p, c])
    for t in range(6):
        p = int(rng.integers(b.n))
        st = 1e-3 * h[p]
        hp, hm = h.copy(), h.copy()
        hp[p] += st
        hm[p] -= st
        fd = (frozen_L(x0, hp, m, g, nbr, parts) - frozen_L(x0, hm, m, g, nbr, parts)) / (2 * st)
        assert abs(fd - dh[p]) <= 2e-5 * scale_h + 1e-4 * abs(fd), (p, fd, dh[p])
    print(f"parts={parts}: max |FD - adjoint| / max|dx| = {max(errs):.2e}")


def test_field_gradient_matches_reference_gradient(box1):
    """a(x) at the scatterers equals the reference's first-step gradient."""
    ctx = skg.Context(box1, skg.make_params(0.005, 32, period=1.0, accumulate=skg.SKG_ACC_PLAIN))
    b2, _, _ = ctx.density()
    a = ctx.field_gradient(box1.pos)
    o = O.Oracle(box1, nSmooth=32, fTau=0.005, fPeriod=1.0)
    o.density()
    ar = o.first_gradient()
    a64 = O.gradient_f64(box1.pos, box1.mass, b2, box1.pos)
    nrm = np.linalg.norm(a64, axis=1)
    e_ref = np.linalg.norm(a - ar, axis=1) / nrm
    e_64 = np.linalg.norm(a - a64, axis=1) / nrm
    r_64 = np.linalg.norm(ar - a64, axis=1) / nrm
    print(f"|a_gpu - a_ref|/|a|: median {np.median(e_ref):.2e} p99.9 {np.percentile(e_ref, 99.9):.2e} max {e_ref.max():.2e}")
    print(f"|a_gpu - a_f64|/|a|: median {np.median(e_64):.2e} p99.9 {np.percentile(e_64, 99.9):.2e} max {e_64.max():.2e}")
    print(f"|a_ref - a_f64|/|a|: median {np.median(r_64):.2e} p99.9 {np.percentile(r_64, 99.9):.2e} max {r_64.max():.2e}")
    assert np.median(e_ref) < 1e-6 and np.percentile(e_ref, 99) < 1e-5
    assert np.median(e_64) < 1e-6 and np.percentile(e_64, 99) < 1e-5


def test_adjoint_target_term_is_minus_half_reference_gradient():
    """One-hot gbar at m, scatter half only: d rho_m / d x_m = -a(x_m)/2."""
    b = gen.config1(n=8192, seed=17)
    ctx = skg.Context(b, skg.make_params(0.008, 32, period=1.0))
    ctx.density()
    a = ctx.field_gradient(b.pos)
    rng = np.random.default_rng(0)
    for m in rng.integers(0, b.n, 8):
        g = np.zeros(b.n, np.float32)
        g[m] = 1.0
        dx, _ = ctx.adjoint(g, 2)
        # only the scatter half's dependence on x_m through rho_m (pass 2) lands on m
        assert np.allclose(dx[m], -0.5 * a[m], rtol=1e-4, atol=1e-6 * np.abs(a[m]).max())


def test_adjoint_2pow20_gradient_parity():
    """Config-2 scale: the field gradient at 2^20 scatterers against FP64 on a sample."""
    b = gen.config2()
    ctx = skg.Context(b, skg.make_params(0.002, 64, period=1.0))
    b2, _, _ = ctx.density()
    idx = np.random.default_rng(2).choice(b.n, 256, replace=False)
    a = ctx.field_gradient(b.pos[idx])
    a64 = O.gradient_f64(b.pos, b.mass, b2, b.pos[idx])
    e = np.linalg.norm(a - a64, axis=1) / np.linalg.norm(a64, axis=1)
    print(f"2^20 k=64 field gradient vs FP64: median {np.median(e):.2e} max {e.max():.2e}")
    assert np.median(e) < 1e-6 and e.max() < 1e-4
    g = np.ones(b.n, np.float32)
    dx, dh = ctx.adjoint(g, 3)
    assert np.isfinite(dx).all() and np.isfinite(dh).all()


def test_adjoint_2pow20_central_differences():
    """Config 2 (2^20 particles, 500 NFW halos, nSmooth 64): the adjoint dx and
    dh of L = sum_j gbar_j rho_j against FP64 central differences (step 1e-3 h)
    per sampled particle, with the GPU's own frozen neighbour sets.  Moving x_p
    changes only the pairs (i, j in N_i) with i = p or j = p, and h_p only those
    with i = p, so each difference sums those pairs alone."""
    b = gen.config2()
    k = 64
    ctx = skg.Context(b, skg.make_params(0.002, k, period=1.0))
    b2, _, _ = ctx.density()
    rng = np.random.default_rng(20)
    gbar = rng.normal(size=b.n).astype(np.float32)
    dx, dh = ctx.adjoint(gbar, 3)
    nb, _ = ctx.neighbors()
    ctx.close()
    nbr = nb[:, :-1].astype(np.int64)                       # the head (largest d2) has weight 0
    n = b.n
    x0 = b.pos.astype(np.float64)
    h = 0.5 * np.sqrt(b2.astype(np.float64))
    m = b.mass.astype(np.float64)
    g = gbar.astype(np.float64)
    ii = np.repeat(np.arange(n), k - 1)
    jj = nbr.reshape(-1)
    by_j = np.argsort(jj, kind="stable")
    jstart = np.searchsorted(jj[by_j], np.arange(n + 1))

    def pairs_of(p):
        rows = np.arange(p * (k - 1), (p + 1) * (k - 1))
        return np.concatenate([rows, by_j[jstart[p]:jstart[p + 1]]])

    def L_pairs(x, hh, rows):
        i, j = ii[rows], jj[rows]
        d = x[i] - x[j]
        d -= np.round(d)
        r = np.sqrt((d * d).sum(1))
        W = m4(r / hh[i]) / (math.pi * hh[i] ** 3)
        return float(np.sum(0.5 * W * (g[i] * m[j] + g[j] * m[i])))

    scale_x, scale_h = np.abs(dx).max(), np.abs(dh).max()
    ex, eh = [], []
    for p in rng.choice(n, 24, replace=False):
        rows = pairs_of(int(p))
        for c in range(3):
            st = 1e-3 * h[p]
            xp, xm = x0.copy(), x0.copy()
            xp[p, c] += st
            xm[p, c] -= st
            fd = (L_pairs(xp, h, rows) - L_pairs(xm, h, rows)) / (2 * st)
            ex.append(abs(fd - dx[p, c]) / scale_x)
            assert abs(fd - dx[p, c]) <= 2e-5 * scale_x + 1e-4 * abs(fd), (p, c, fd, dx[p, c])
        rows_h = np.arange(p * (k - 1), (p + 1) * (k - 1))
        st = 1e-3 * h[p]
        hp, hm = h.copy(), h.copy()
        hp[p] += st
        hm[p] -= st
        fd = (L_pairs(x0, hp, rows_h) - L_pairs(x0, hm, rows_h)) / (2 * st)
        eh.append(abs(fd - dh[p]) / scale_h)
        assert abs(fd - dh[p]) <= 2e-5 * scale_h + 1e-4 * abs(fd), (p, fd, dh[p])
    print(f"2^20 k=64 adjoint vs central differences over 24 particles: dx max {max(ex):.2e} of max|dx|, "
          f"dh max {max(eh):.2e} of max|dh|")


def frozen_rho(x, h, mass, nbr):
    """rho_i with the kNN sets frozen (head excluded), FP64: both halves."""
    n, k = nbr.shape
    i = np.repeat(np.arange(n), k)
    j = nbr.reshape(-1)
    d = x[i] - x[j]
    d -= np.round(d)
    r = np.sqrt((d * d).sum(1))
    W = 0.5 * m4(r / h[i]) / (math.pi * h[i] ** 3)
    return np.bincount(i, weights=W * mass[j], minlength=n) + np.bincount(j, weights=W * mass[i], minlength=n)


def test_density_dlnk_central_differences():
    """d rho / d ln nSmooth against central differences of the FP64 density
    with frozen sets and every h scaled by exp(eps/3)."""
    b = gen.config1(n=4096, seed=17)
    k = 32
    ctx = skg.Context(b, skg.make_params(0.01, k, period=1.0))
    b2, rho, _ = ctx.density()
    d = ctx.density_dlnk()
    _, nb = O.knn_brute(b.pos, k, lists=True)
    nbr = nb[:, :-1]
    x = b.pos.astype(np.float64)
    h = 0.5 * np.sqrt(b2.astype(np.float64))
    m = b.mass.astype(np.float64)
    r0 = frozen_rho(x, h, m, nbr)
    assert np.allclose(r0, rho, rtol=1e-5)
    eps = 1e-4
    fd = (frozen_rho(x, h * math.exp(eps / 3), m, nbr) - frozen_rho(x, h * math.exp(-eps / 3), m, nbr)) / (2 * eps)
    err = np.abs(fd - d) / r0
    print(f"d rho/d ln k: max |FD - GPU|/rho = {err.max():.2e}; d ln rho/d ln k median {np.median(d / r0):.3f}")
    assert err.max() <= 1e-5
    # once asked for, the next skg_density forms both halves in its own
    # kernels (kNN epilogue, scatter cover) and skg_density_dlnk combines them
    b2b, rhob, _ = ctx.density()
    assert np.array_equal(b2b, b2) and np.array_equal(rhob, rho)
    d2 = ctx.density_dlnk()
    err2 = np.abs(fd - d2) / r0
    print(f"fused: max |FD - GPU|/rho = {err2.max():.2e}; max |fused - separate|/rho = {np.max(np.abs(d2 - d) / r0):.2e}")
    assert err2.max() <= 1e-5
    assert np.max(np.abs(d2 - d) / r0) <= 1e-6
    # densities fall as the smoothing volume grows: the sensitivity is
    # negative for most particles
    assert np.median(d / r0) < 0
    rep = skg.nsmooth_report(rho, d, np.zeros(b.n, np.int32), dens_min=float(np.median(rho)), rel_k=0.25)
    assert rep["scatterers"] == b.n and rep["at_risk"] > 0
