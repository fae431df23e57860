"""The oracle's restatement against the real reference (oracle/_ref, built from
/root/reference by oracle/ref.mk) on identical inputs.  CPU only; skipped where
the reference harness is not built."""
import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen

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


def test_pipeline_bitexact(box1):
    """Whole path main.c:374-471: Ittr trace, FOF labels and final groups identical."""
    kw = dict(nSmooth=32, fTau=0.005, fPeriod=1.0)
    r = O.ref_pipeline(box1, **kw)
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
    assert ng == r["nGroup"]
    assert np.array_equal(g, r["iGroup"])


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
