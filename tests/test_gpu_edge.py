"""Edge cases of batched and large inputs: empty and undersized boxes are
refused with a status, a box of exactly nSmooth particles and very uneven
(ragged) batches run through the whole path, and at 2^22 particles the kNN
radii agree with an FP64 k-d tree on a sample (a size-independent check)."""
import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen, skg

pytestmark = pytest.mark.gpu


def _uniform(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    return gen.Box(pos=pos, vel=np.zeros_like(pos), mass=np.full(n, 1.0 / max(n, 1), np.float32),
                   soft=np.full(n, 1e-3, np.float32), temp=np.zeros(n, np.float32), n_gas=0, n_dark=n)


def test_batch_refuses_empty_and_undersized_boxes():
    P = skg.make_params(0.01, 32, period=1.0)
    with pytest.raises(skg.SkgError):
        skg.BatchContext([_uniform(2048, 1), _uniform(0, 2)], P, [0.01, 0.01])
    bc = skg.BatchContext([_uniform(2048, 1), _uniform(20, 2)], P, [0.01, 0.01])
    with pytest.raises(skg.SkgError, match="nSmooth"):
        bc.density()
    bc.close()


def test_ragged_batch_with_a_box_of_exactly_nsmooth():
    sizes = [32, 33, 100, 2048, 30000]
    boxes = [_uniform(32, 3), _uniform(33, 4), _uniform(100, 5), gen.config4_box(2048, 6), gen.config4_box(30000, 7)]
    taus = np.array([0.2, 0.2, 0.1, gen.config4_tau(2048), gen.config4_tau(30000)], np.float32)
    bc = skg.BatchContext(boxes, skg.make_params(0.01, 32, period=1.0), taus)
    b2, rho, _ = bc.density()
    for i, box in enumerate(boxes):
        got = bc.split(b2)[i]
        assert np.array_equal(got, O.knn_brute(box.pos, 32)), sizes[i]
        assert np.all(bc.split(rho)[i] > 0)
    bc.move_loop(max_blocks=2000)
    bc.fof()
    bc.microsteps()
    bc.centers()
    bc.unbind()
    bc.too_small()
    g, ng = bc.groups()
    # groups never span boxes, and the big boxes' catalogues match their own runs
    for i, gb in enumerate(bc.split(g)):
        ids = set(np.unique(gb[gb > 0]).tolist())
        for j, gj in enumerate(bc.split(g)):
            if j != i:
                assert not ids & set(np.unique(gj[gj > 0]).tolist())
    for i in (3, 4):
        one = skg.Context(boxes[i], skg.make_params(float(taus[i]), 32, period=1.0))
        _, sg, _ = one.run()
        assert O.match_fraction(bc.split(g)[i], sg) >= 0.999
        one.close()
    bc.close()


def test_knn_radii_at_2pow22_against_fp64_tree():
    from scipy.spatial import cKDTree
    n = 1 << 22
    box = gen.config4_box(n, 11)
    ctx = skg.Context(box, skg.make_params(gen.config4_tau(n), 32, period=1.0))
    b2, rho, _ = ctx.density()
    ctx.close()
    assert np.all(np.isfinite(b2)) and np.all(b2 > 0) and np.all(rho > 0)
    rng = np.random.default_rng(0)
    idx = rng.choice(n, 4000, replace=False)
    pos = box.pos.astype(np.float64) + 0.5
    tree = cKDTree(np.mod(pos, 1.0), boxsize=1.0)
    d, _ = tree.query(np.mod(pos[idx], 1.0), k=32)
    ref = d[:, -1] ** 2
    err = np.abs(b2[idx].astype(np.float64) - ref)
    # fp32 d2: ~1e-7 relative, plus, for a query within a ball of a face, the
    # query-side image x -/+ L rounded to fp32 (half an ulp of 1 per axis) --
    # the reference's own arithmetic, which the GPU reproduces bit for bit
    bound = 1e-6 * ref + 4.0 * np.sqrt(ref) * 2.0 ** -24
    w = int(np.argmax(err / bound))
    face = 0.5 - np.abs(box.pos[idx[w]]).max()
    print(f"2^22: kNN radius^2 vs FP64 tree on {idx.size} queries: max rel {(err / ref).max():.2e}, "
          f"worst err/bound {(err / bound)[w]:.2f} at {np.sqrt(ref[w]):.2e} from a face {face:.2e} away")
    assert np.all(err <= bound)


def test_device_memory_is_accounted_and_released():
    cur0, _ = skg.device_memory(reset_peak=True)
    n = 1 << 16
    box = gen.config4_box(n, 12)
    ctx = skg.Context(box, skg.make_params(gen.config4_tau(n), 32, period=1.0))
    ctx.run()
    cur, peak = skg.device_memory()
    # at least the particle arrays (pos, vel, mass, soft, rho, ball2, group: 40 B each)
    assert cur - cur0 >= 40 * n and peak >= cur
    ctx.close()
    assert skg.device_memory()[0] == cur0


def test_keep_outputs_pinned_arrays_match():
    """keep_outputs(): per-pass outputs in page-locked arrays the context
    reuses; their values equal those of freshly allocated arrays, pass by pass
    (the second pass forms d rho / d ln k inside the density, as a context does
    once it has been asked for it)."""
    box = gen.config1(n=4096, seed=3)
    P = skg.make_params(0.01, 32, period=1.0)
    a = skg.Context(box, P)
    ref = []
    for _ in range(2):
        b2, rho, _ = a.density()
        ref.append((b2, rho, a.density_dlnk()))
    a.close()
    c = skg.Context(box, P)
    c.keep_outputs(True)
    for b2a, rhoa, dla in ref:
        b2c, rhoc, _ = c.density()
        dlc = c.density_dlnk()
        assert np.array_equal(b2a, b2c) and np.array_equal(rhoa, rhoc) and np.array_equal(dla, dlc)
    b2d, _, _ = c.density()
    assert b2d is b2c                     # the same kept array, overwritten in place
    c.close()
