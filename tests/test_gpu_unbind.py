"""The -unbind re-entry (skg_set_groups), FP64 per-group energies and the
unbinding of power-law groups with cosmology and scoop, against the FP64
restatement of kdUnbind (oracle.unbind_f64, pinned on the CPU against the
reference-faithful oracle)."""
import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen, skg

pytestmark = pytest.mark.gpu


def test_set_groups_reproduces_pipeline_unbinding(box1):
    """Groups handed in from outside unbind exactly as the pipeline's own
    (config 1: every particle is in a group, so no scoop and no centre
    dependence)."""
    P = skg.make_params(0.005, 32, period=1.0)
    a = skg.Context(box1, P)
    a.density()
    a.move_loop(max_blocks=2000)
    fg, _ = a.fof()
    a.microsteps()
    a.centers()
    a.unbind()
    a.too_small()
    g1, _ = a.groups()
    b = skg.Context(box1, P)
    b.set_groups(fg)
    b.unbind()
    b.too_small()
    g2, _ = b.groups()
    assert np.array_equal(g1, g2)


def _c5_small(seed=11):
    n = 1 << 18
    box, grp = gen.config5(n=n, n_groups=300, seed=seed, n_range=(8, 12000))
    tau = 0.16 * n ** (-1.0 / 3.0)
    fshift, fcosmo = gen.config5_cosmology()
    P = skg.make_params(tau, 32, period=1.0, z=0.5, fShift=fshift, fCosmo=fcosmo)
    return box, grp, P, tau, fshift, fcosmo


def test_c5_group_energies_vs_fp64():
    box, grp, P, tau, fshift, fcosmo = _c5_small()
    ctx = skg.Context(box, P)
    ng = ctx.set_groups(grp)
    W, T = ctx.group_energies(ng)
    order = np.argsort(grp, kind="stable")
    bounds = np.searchsorted(grp[order], np.arange(ng + 1))
    errW, errT = [], []
    for g in range(1, ng):
        mem = order[bounds[g]:bounds[g + 1]]
        if mem.size < 2:
            continue
        rel = box.pos[mem.max()].astype(np.float64)
        d = box.pos[mem].astype(np.float64) - rel
        d -= np.round(d)
        m = box.mass[mem].astype(np.float64)
        phi = O.potential_f64(d, m, box.soft[mem])
        w = -0.5 * (m * phi).sum()
        M = m.sum()
        u = fshift * (box.vel[mem] - (m[:, None] * box.vel[mem]).sum(0) / M) + \
            fcosmo * (d - (m[:, None] * d).sum(0) / M)
        t = 0.5 * (m * (u * u).sum(1)).sum()
        errW.append(abs(W[g] - w) / abs(w))
        errT.append(abs(T[g] - t) / abs(t))
    print(f"{len(errW)} groups: |dW|/|W| max {max(errW):.2e} median {np.median(errW):.2e}; "
          f"|dT|/T max {max(errT):.2e}")
    assert max(errW) <= 2e-6 and max(errT) <= 1e-6


def test_c5_unbinding_vs_fp64():
    """Power-law groups with virial ratios in [0.5, 2], flat Lambda at z = 0.5
    and the scoop of the uniform background: the FP32-potential / FP64-test
    unbinding keeps the membership of the all-FP64 removal."""
    box, grp, P, tau, fshift, fcosmo = _c5_small(seed=12)
    ctx = skg.Context(box, P)
    ng = ctx.set_groups(grp)
    nunb = ctx.unbind()
    g, _ = ctx.groups()
    cat = ctx.catalogue(ng)
    centres = {i: cat["center"][i] for i in range(1, ng)}
    ref = O.unbind_f64(box.pos, box.vel, box.mass, box.soft, grp, centres, fshift=fshift, fcosmo=fcosmo, z=0.5,
                       scoop=2.0 * tau, sub_pot=True)
    in_groups = grp > 0
    same = ((g > 0) == (ref > 0))[in_groups].mean()
    print(f"unbound gpu {nunb} fp64 {int(((ref == 0) & in_groups).sum())} of {int(in_groups.sum())} members; "
          f"same bound/unbound status {same:.6f}")
    assert same >= 0.999
    assert O.match_fraction(g, ref) >= 0.999
