"""Batched boxes (skg_load_batch): many independent periodic boxes of uneven
size, each with its own tau, run through every stage in the same launches.
Each box must come out as it does alone -- alone on the GPU, and in the
oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen, skg

pytestmark = pytest.mark.gpu

SIZES = [2048, 2600, 3500, 5200, 8000, 12000]


def _boxes(sizes=SIZES):
    boxes = [gen.config4_box(n, b) for b, n in enumerate(sizes)]
    taus = np.array([gen.config4_tau(n) for n in sizes], np.float32)
    return boxes, taus


def _params(tau=0.01, **kw):
    return skg.make_params(tau, 32, period=1.0, **kw)


def _run_all(ctx):
    b2, rho, _ = ctx.density()
    tr = ctx.move_loop(max_blocks=2000)
    fg, _ = ctx.fof()
    ctx.microsteps()
    ctx.centers()
    ctx.unbind()
    ctx.too_small()
    g, _ = ctx.groups()
    return b2, rho, tr, fg, g


@pytest.mark.parametrize("sub_max", [None, "500"])
def test_batch_matches_each_box_alone(sub_max, monkeypatch):
    # sub_max 500: dense steps on the lane-per-mover kernel with leaf lists
    if sub_max:
        monkeypatch.setenv("SKG_COVER_SUB_MAX", sub_max)
    boxes, taus = _boxes()
    bc = skg.BatchContext(boxes, _params(), taus)
    b2, rho, trs, fg, g = _run_all(bc)
    nb, _ = bc.neighbors()
    for i, (box, tau) in enumerate(zip(boxes, taus)):
        lo, hi = bc.offsets[i], bc.offsets[i + 1]
        # neighbour lists never leave the box
        assert nb[lo:hi].min() >= lo and nb[lo:hi].max() < hi
        one = skg.Context(box, _params(float(tau)))
        sb2, srho, str_, sfg, sg = _run_all(one)
        assert np.array_equal(b2[lo:hi], sb2), i                     # kNN radii bit-exact
        rel = np.abs(rho[lo:hi] - srho) / srho
        assert rel.max() <= 1e-6, (i, rel.max())
        t = trs[i]
        print(f"box {i} n={box.n} trace batch {len(t)} alone {len(str_)}; first {t[:2].tolist()} {str_[:2].tolist()}")
        assert abs(len(t) - len(str_)) <= 1
        m = min(len(t), len(str_))
        assert np.all(np.abs(t[:m, 0] - str_[:m, 0]) <= 2)
        ff = O.match_fraction(bc.split(fg)[i], sfg)
        fr = O.match_fraction(bc.split(g)[i], sg)
        print(f"   fof match {ff:.5f} final match {fr:.5f}")
        assert ff >= 0.999 and fr >= 0.999
        one.close()


def test_batch_groups_are_contiguous_per_box():
    boxes, taus = _boxes(SIZES[:4])
    bc = skg.BatchContext(boxes, _params(), taus)
    _, _, _, _, g = _run_all(bc)
    last = 0
    for gb in bc.split(g):
        ids = np.unique(gb[gb > 0])
        if ids.size:
            assert ids[0] == last + 1 and ids[-1] == last + ids.size
            last = ids[-1]


@pytest.mark.parametrize("acc", [skg.SKG_ACC_COMPENSATED, skg.SKG_ACC_PLAIN])
def test_batch_against_oracle(acc):
    boxes, taus = _boxes(SIZES[:3])
    bc = skg.BatchContext(boxes, _params(accumulate=acc), taus)
    b2, rho, trs, fg, g = _run_all(bc)
    for i, (box, tau) in enumerate(zip(boxes, taus)):
        o = O.Oracle(box, nSmooth=32, fTau=float(tau), fPeriod=1.0)
        ob2, orho, _ = o.density()
        assert np.array_equal(bc.split(b2)[i], ob2)
        rel = np.abs(bc.split(rho)[i] - orho) / orho
        assert rel.max() <= 1e-5, (i, rel.max())
        otr = o.move_loop()
        ofg, _ = o.fof()
        og, _ = o.finish()
        ff = O.match_fraction(bc.split(fg)[i], ofg)
        fr = O.match_fraction(bc.split(g)[i], og)
        print(f"box {i} acc={acc}: trace {len(trs[i])} vs {len(otr)}, fof {ff:.5f}, final {fr:.5f}")
        assert abs(len(trs[i]) - len(otr)) <= 3
        assert ff >= 0.999 and fr >= 0.999


def test_batch_two_species_and_adjoint():
    """Gas+dark boxes with -gd; the adjoint of a batched context equals the
    per-box adjoints."""
    boxes = [gen.config3(n_each=n, seed=40 + i, n_halos=4) for i, n in enumerate((1024, 1500, 2048))]
    taus = np.array([0.012, 0.010, 0.009], np.float32)
    bc = skg.BatchContext(boxes, _params(bGasAndDark=1), taus)
    b2, rho, _ = bc.density()
    rng = np.random.default_rng(3)
    gbar = rng.normal(size=bc.n).astype(np.float32)
    dx, dh = bc.adjoint(gbar, 3)
    for i, (box, tau) in enumerate(zip(boxes, taus)):
        one = skg.Context(box, _params(float(tau), bGasAndDark=1))
        sb2, srho, _ = one.density()
        assert np.array_equal(bc.split(b2)[i], sb2)
        sdx, sdh = one.adjoint(bc.split(gbar)[i], 3)
        scale = np.abs(sdx).max()
        assert np.abs(bc.split(dx)[i] - sdx).max() <= 1e-5 * scale
        assert np.abs(bc.split(dh)[i] - sdh).max() <= 1e-5 * np.abs(sdh).max()
        one.close()
    tr = bc.move_loop(max_blocks=2000)
    bc.fof()
    bc.microsteps()
    bc.centers()
    bc.unbind()
    bc.too_small()
    g, _ = bc.groups()
    for i, (box, tau) in enumerate(zip(boxes, taus)):
        o = O.Oracle(box, nSmooth=32, fTau=float(tau), fPeriod=1.0, bGasAndDark=1)
        o.run()
        og, _ = o.groups()
        fr = O.match_fraction(bc.split(g)[i], og)
        print(f"two-species box {i}: final match {fr:.5f}")
        assert fr >= 0.999


def test_batch_rejects_mixed_species():
    a = gen.config4_box(2048, 0)
    b = gen.config3(n_each=1024, seed=9, n_halos=2)
    with pytest.raises(skg.SkgError):
        skg.BatchContext([a, b], _params(), None)
