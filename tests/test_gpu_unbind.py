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


_COOP_CHILD = r"""
import os, sys
import numpy as np
sys.path.insert(0, os.getcwd())
from skid_amd import gen, skg
coop_min, path = sys.argv[1], sys.argv[2]
n = 1 << 18
box, grp = gen.config5(n=n, n_groups=300, seed=12, n_range=(8, 12000))
fs, fc = gen.config5_cosmology()
P = skg.make_params(0.16 * n ** (-1.0 / 3.0), 32, period=1.0, z=0.5, fShift=fs, fCosmo=fc)
out = {}
for tag, env in (("cand", {}), ("coopc", {"SKG_UNBIND_COOP_MIN": coop_min}),
                 ("coopscan", {"SKG_UNBIND_COOP_MIN": coop_min, "SKG_UNBIND_COOP": "scan"})):
    for k in ("SKG_UNBIND_COOP_MIN", "SKG_UNBIND_COOP"):
        os.environ.pop(k, None)
    os.environ.update(env)
    ctx = skg.Context(box, P)
    ng = ctx.set_groups(grp)
    nunb = ctx.unbind()
    info = ctx.unbind_info()
    g, _ = ctx.groups()
    out[tag + "_g"] = g
    out[tag + "_n"] = np.array([nunb, *info])
    ctx.close()
np.savez(path, grp=grp, **out)
"""


@pytest.mark.parametrize("coop_min", ["2000", "300"])
def test_c5_cooperative_removal_paths_agree(coop_min, tmp_path):
    """Power-law groups with kdSubPot: the cooperative kernels (bounded
    candidates, the default, and one scan per removal) unbind the groups routed
    to them as the one-workgroup candidate kernel does.  Run in a child process:
    a process that has made cooperative launches can keep the device's
    cooperative resources, and this suite's other children need them."""
    import os
    import subprocess
    import sys
    path = str(tmp_path / "coop.npz")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([sys.executable, "-c", _COOP_CHILD, coop_min, path], check=True, timeout=240, cwd=root)
    r = np.load(path)
    sizes = np.bincount(r["grp"])[1:]
    print(f"groups >= {coop_min}: {(sizes >= int(coop_min)).sum()}; [unbound, coop groups, refused, full passes] "
          + ", ".join(f"{k} {r[k + '_n'].tolist()}" for k in ("cand", "coopc", "coopscan")))
    assert r["cand_n"][1] == 0
    for tag in ("coopc", "coopscan"):
        nunb, nbig, refused, _ = r[tag + "_n"]
        assert nbig == min(128, (sizes >= int(coop_min)).sum()) and refused == 0
        assert O.match_fraction(r[tag + "_g"], r["cand_g"]) >= 0.9999
        assert abs(nunb - r["cand_n"][0]) <= max(2, r["cand_n"][0] // 1000)
