"""The oracle's restatement against the real reference (oracle/_ref, built from
/root/reference by oracle/ref.mk) on identical inputs.  CPU only; skipped where
the reference harness is not built."""
import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen
from option_cases import OPTION_CASES, option_box

pytestmark = pytest.mark.skipif(not O.have_ref(), reason="reference harness not built")


@pytest.mark.parametrize("k,per", [(32, True), (64, True), (32, False)])
def test_density_bitexact(box1, k, per):
    kw = dict(nSmooth=k, fTau=0.005)
    if per:
        kw["fPeriod"] = 1.0
    rb2, rrho, rne = O.ref_density(box1, **kw)
    ob2, orho, one = O.Oracle(box1, **kw).density()
    assert rne == one
    assert np.array_equal(rb2, ob2)
    assert np.array_equal(rrho, orho)


def test_density_two_species():
    """-gd and gas-only scatterer selection (kd.c:600-627) on a small gas+dark box."""
    b = gen.config3(n_each=4096, seed=5, n_halos=20)
    for gd in (0, 1):
        kw = dict(nSmooth=32, fTau=0.01, fPeriod=1.0, bGasAndDark=gd)
        rb2, rrho, _ = O.ref_density(b, **kw)
        ob2, orho, _ = O.Oracle(b, **kw).density()
        assert np.array_equal(rb2, ob2)
        assert np.array_equal(rrho, orho)


def test_brute_knn_matches_reference_radii(box1):
    """Nearest-image fp32 d2 with index ties reproduces the reference's fBall2."""
    rb2, _, _ = O.ref_density(box1, nSmooth=32, fTau=0.005, fPeriod=1.0)
    assert np.array_equal(O.knn_brute(box1.pos, 32), rb2)


def test_pipeline_bitexact(box1, tmp_path):
    """Whole path main.c:374-471: Ittr trace, FOF labels, the Microstep counts,
    the final groups and the .gtp catalogue (mass, rCenter, vcm and the radius
    of kdWriteGroup, kd.c:1612-1687) identical."""
    kw = dict(nSmooth=32, fTau=0.005, fPeriod=1.0)
    gtp = str(tmp_path / "ref.gtp")
    r = O.ref_pipeline(box1, gtp_path=gtp, **kw)
    o = O.Oracle(box1, **kw)
    o.density()
    tr = o.move_loop()
    assert np.array_equal(tr, r["trace"])
    moved = o.moved()
    assert np.array_equal(moved, r["rMoved"], equal_nan=True)
    fg, nfg = o.fof()
    assert nfg == r["nFofGroup"]
    assert np.array_equal(fg, r["iFofGroup"])
    g, ng = o.finish()
    assert np.array_equal(o.micro, r["micro"])
    assert ng == r["nGroup"]
    assert np.array_equal(g, r["iGroup"])
    _check_catalogue(o, gtp, ng)


def _check_catalogue(o, gtp_path, ng):
    """The oracle's per-group catalogue equals the reference's .gtp records."""
    cat = o.catalogue()
    ref = O.read_gtp(gtp_path, xdr=False)
    assert ref["mass"].shape[0] == ng - 1
    for mine, theirs in (("mass", "mass"), ("center", "pos"), ("vcm", "vel"), ("radius", "radius")):
        assert np.array_equal(cat[mine][1:], ref[theirs]), mine
    assert (cat["radius"][1:] > 0).all()


@pytest.mark.parametrize("case", sorted(OPTION_CASES))
def test_pipeline_options_bitexact(case, tmp_path):
    """The oracle against the reference under each option: Ittr trace,
    Microstep counts, moved positions, FOF and final labels, .gtp catalogue."""
    bname, kw = OPTION_CASES[case]
    b = option_box(bname)
    gtp = str(tmp_path / "ref.gtp")
    r = O.ref_pipeline(b, gtp_path=gtp, **kw)
    o = O.Oracle(b, **kw)
    o.density()
    tr = o.move_loop()
    assert np.array_equal(tr, r["trace"])
    assert np.array_equal(o.moved(), r["rMoved"], equal_nan=True)
    fg, nfg = o.fof()
    assert nfg == r["nFofGroup"] and np.array_equal(fg, r["iFofGroup"])
    g, ng = o.finish()
    assert np.array_equal(o.micro, r["micro"])
    assert ng == r["nGroup"] and np.array_equal(g, r["iGroup"])
    _check_catalogue(o, gtp, ng)
    # each case exercises what it names
    if "densmin" in case:
        assert len(set(tr[1:, 1].tolist())) > 1          # ScatterCut after iteration 0
    if case == "massmax":
        assert r["nMove"] < b.n
    if case in ("nounbind", "maxgroup0"):
        # nothing is unbound: the final groups are the FOF groups of at least nMembers
        sizes = np.bincount(fg)[1:]
        assert ng - 1 == int((sizes >= 8).sum())


def test_pipeline_scoop_and_den_order():
    """-d > 0 leaves group 0 non-empty, so the scoop potential (grav.c:63-135,
    always on: kd.c:1379-1384) and -den's reordering (kd.c:1487-1499) are exercised."""
    b = gen.config1(n=8192, seed=3)
    kw = dict(nSmooth=32, fTau=0.008, fPeriod=1.0, fDensMin=40.0, bOutDens=1)
    r = O.ref_pipeline(b, **kw)
    o = O.Oracle(b, **kw)
    tr = o.run()
    g, ng = o.groups()
    assert np.array_equal(tr, r["trace"])
    assert ng == r["nGroup"]
    assert np.array_equal(g, r["iGroup"])


def test_fof_bfs_partition_equals_brute(box1):
    """kdFoF's BFS on the reference's converged positions: same partition as
    brute-force connected components (the link rule is symmetric here)."""
    r = O.ref_pipeline(box1, nSmooth=32, fTau=0.005, fPeriod=1.0)
    pos = r["rMoved"]
    bg, _ = O.fof_brute(pos, 0.005)
    assert O.same_partition(r["iFofGroup"], bg)


def test_fp64_unbinding_matches_reference_unbind_path():
    """The -unbind re-entry of the real reference (centres of mass, scoop of
    the background, flat Lambda at z = 0.5) against the FP64 restatement that
    pins the GPU unbinding: the same bound/unbound status for >= 99.9% of the
    members of a power-law group draw."""
    n = 1 << 16
    box, grp = gen.config5(n=n, n_groups=150, seed=7, n_range=(8, 3000))
    tau = 0.16 * n ** (-1.0 / 3.0)
    kw = dict(nSmooth=32, fTau=tau, fPeriod=1.0, z=0.5, Omega0=0.3, Lambda=0.7, H0=2.8944, nMembers=1)
    rg = O.ref_unbind(box, grp, **kw)
    fs, fc = gen.config5_cosmology()
    # centres of mass as kdReadCenter forms them: relative to the largest-iOrder member
    cen = {}
    for g in range(1, grp.max() + 1):
        mem = np.flatnonzero(grp == g)
        rel = box.pos[mem.max()].astype(np.float64)
        d = box.pos[mem].astype(np.float64) - rel
        d -= np.round(d)
        m = box.mass[mem].astype(np.float64)
        c = rel + (m[:, None] * d).sum(0) / m.sum()
        cen[g] = c - np.round(c)
    u = O.unbind_f64(box.pos, box.vel, box.mass, box.soft, grp, cen, fshift=fs, fcosmo=fc, z=0.5,
                     scoop=2.0 * tau, sub_pot=True)
    inm = grp > 0
    same = ((rg > 0) == (u > 0))[inm].mean()
    print(f"reference -unbind vs FP64: unbound {int(((rg == 0) & inm).sum())} vs {int(((u == 0) & inm).sum())}, "
          f"same status {same:.5f}")
    assert same >= 0.999


def test_group_stats_restatement_matches_reference_stat(tmp_path):
    """The reference binary's -stats output (kdOutStats) against the fp32
    restatement given the reference's own groups and centres (.grp/.gtp):
    every column of every group, z = 0.5 flat Lambda so the Hubble term of the
    dispersion is on, and gas with a density floor so the gas mass is tested."""
    import os
    import subprocess
    ref = os.path.join(O.HERE, "_ref", "skid")
    if not os.path.exists(ref):
        pytest.skip("reference binary not built")
    box = gen.config3(n_each=4096, seed=5, n_halos=20)
    inp = tmp_path / "b.std"
    gen.write_tipsy(str(inp), box, xdr=True)
    dmin = 30.0
    args = [ref, "-tau", "0.01", "-s", "32", "-m", "8", "-p", "1", "-std", "-z", "0.5", "-H", "2.8944",
            "-O", "0.3", "-Lambda", "0.7", "-d", str(dmin), "-den", "-stats", "-o", "ref"]
    with open(inp, "rb") as f:
        r = subprocess.run(args, stdin=f, capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    st = O.read_stat(str(tmp_path / "ref.stat"))
    grp = O.read_grp(str(tmp_path / "ref.grp"))
    gtp = O.read_gtp(str(tmp_path / "ref.gtp"))
    rho = np.loadtxt(str(tmp_path / "ref.den"), skiprows=1).astype(np.float32)
    ng = int(grp.max()) + 1
    assert st.shape[0] == ng - 1 and gtp["pos"].shape[0] == ng - 1
    cen = np.vstack([np.zeros((1, 3), np.float32), gtp["pos"]])
    vcm = np.vstack([np.zeros((1, 3), np.float32), gtp["vel"]])
    ptype = np.where(np.arange(box.n) < box.n_gas, 2, 1)
    fs, fc = gen.config5_cosmology()
    ours = O.group_stats_f32(box.pos, box.vel, box.mass, box.soft, grp, cen, vcm, ptype=ptype, rho=rho,
                             temp=box.temp, fexp=fs, fexphub=fc, dens_min=dmin)
    counts = np.bincount(grp, minlength=ng)
    want = np.column_stack([np.arange(1, ng), counts[1:], ours[1:, 0:3], np.sqrt(ours[1:, 3:6]),
                            ours[1:, 6:8], np.sqrt(ours[1:, 8]), ours[1:, 9], cen[1:], vcm[1:]])
    got = st[:, :18]
    # %g keeps 6 significant digits
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    bad = (rel > 6e-6) & (np.abs(got - want) > 1e-9)
    print(f"{ng - 1} groups; gas mass in {int((ours[1:, 1] > 0).sum())}; worst column rel "
          f"{np.max(np.where(np.abs(want) > 0, rel, 0), axis=0).round(8).tolist()}")
    assert not bad.any(), np.argwhere(bad)[:10]
