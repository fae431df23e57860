"""GPU move loop, FOF and unbinding against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen, skg

pytestmark = pytest.mark.gpu


def test_fof_points_equals_brute_force():
    rng = np.random.default_rng(4)
    b = gen.config1(n=8192, seed=21)
    sets = [b.pos,
            rng.uniform(-0.5, 0.5, (5000, 3)).astype(np.float32),
            # clumps collapsed well inside tau (contained-node path) and points on the faces
            np.concatenate([rng.normal(0.0, 1e-4, (3000, 3)), rng.normal(0.49, 1e-3, (500, 3)),
                            rng.uniform(-0.5, 0.5, (500, 3))]).astype(np.float32)]
    sets[2] = (sets[2] - np.floor(sets[2] + 0.5)).astype(np.float32)
    ctx = skg.Context(b, skg.make_params(0.005, 32, period=1.0))
    for pos in sets:
        for tau in (0.002, 0.01):
            g, ng = ctx.fof_points(pos, tau)
            bg, bng = O.fof_brute(pos, tau)
            assert ng == bng + 1
            assert np.array_equal(g, bg)        # same numbering: by smallest index


def test_fof_on_reference_converged_positions(box1):
    """The FOF stage on the reference's own converged positions reproduces the
    reference's partition (numbering differs: here by smallest iOrder)."""
    o = O.Oracle(box1, nSmooth=32, fTau=0.005, fPeriod=1.0)
    o.density()
    o.move_loop()
    pos = o.moved()
    fg, _ = o.fof()
    ctx = skg.Context(box1, skg.make_params(0.005, 32, period=1.0))
    g, _ = ctx.fof_points(pos, 0.005)
    assert O.same_partition(g, fg)
    assert np.array_equal(g, O.relabel_min_index(fg))


def test_move_trace_and_catalogue_config1(box1):
    P = skg.make_params(0.005, 32, period=1.0)
    ctx = skg.Context(box1, P)
    ctx.density()
    tr = ctx.move_loop(max_blocks=2000)
    fg, nfg = ctx.fof()
    ctx.microsteps()
    ctx.centers()
    nunb = ctx.unbind()
    ctx.too_small()
    g, ng = ctx.groups()
    o = O.Oracle(box1, nSmooth=32, fTau=0.005, fPeriod=1.0)
    o.density()
    otr = o.move_loop()
    ofg, onfg = o.fof()
    og, ong = o.finish()
    print("trace gpu", tr[:4].tolist(), len(tr), "oracle", otr[:4].tolist(), len(otr))
    print("fof groups", nfg, onfg, "final", ng, ong, "unbound", nunb)
    ffrac = O.match_fraction(fg, ofg)
    frac = O.match_fraction(g, og)
    print(f"fof match {ffrac:.5f} final match {frac:.5f}")
    assert abs(len(tr) - len(otr)) <= 3
    assert np.all(np.abs(tr[:5, 0] - otr[:5, 0]) <= 0.002 * otr[0, 0] + 2)
    assert ffrac >= 0.999
    assert frac >= 0.999


def test_unbinding_paths_agree(box1):
    """The four removal kernels -- one wavefront per small group, the bounded
    candidate kernel (one workgroup per group), and the cooperative all-GPU
    kernels with bounded candidates and with one scan per removal -- give the
    same catalogue (SKG_UNBIND_* route groups between them)."""
    import os
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, '.');"
        "from skid_amd import gen, skg;"
        "b = gen.config1();"
        "ctx = skg.Context(b, skg.make_params(0.005, 32, period=1.0));"
        "tr, g, ng = ctx.run(); np.save(sys.argv[1], g);"
        "big, refused, _ = ctx.unbind_info();"
        "assert refused == 0 and (big > 0) == ('SKG_UNBIND_COOP_MIN' in __import__('os').environ), (big, refused)"
    )
    outs = []
    for tag, env_add in (("default", {}), ("cand", {"SKG_UNBIND_SMALL_MAX": "1"}),
                         ("coop", {"SKG_UNBIND_COOP_MIN": "100"}),
                         ("coopscan", {"SKG_UNBIND_COOP_MIN": "100", "SKG_UNBIND_COOP": "scan"})):
        env = dict(os.environ)
        env.update(env_add)
        path = f"/tmp/skg_unb_{tag}.npy"
        subprocess.run([sys.executable, "-c", code, path], check=True, env=env, timeout=240,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        outs.append(np.load(path))
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])
    assert np.array_equal(outs[0], outs[3])


def test_two_species_catalogue():
    """Gas+dark with -gd (no potential updates during removal, kd.c:1441)."""
    b = gen.config3(n_each=8192, seed=5, n_halos=30)
    tau = 0.001 * (2 ** ((24 - 13) / 3))
    ctx = skg.Context(b, skg.make_params(tau, 64, period=1.0, bGasAndDark=1))
    tr, g, ng = ctx.run()
    o = O.Oracle(b, nSmooth=64, fTau=tau, fPeriod=1.0, bGasAndDark=1)
    o.run()
    og, ong = o.groups()
    frac = O.match_fraction(g, og)
    print(f"two species: groups gpu {ng - 1} oracle {ong - 1} match {frac:.5f}")
    assert frac >= 0.999


def test_move_loop_launch_paths_agree():
    """The move loop gives the same trace and groups however its blocks are
    launched: direct launches (default), one hipGraph per block, with per-tile
    entity lists built by each block's first step (default) or by a pass of
    their own, leaf lists or none, and with a list capacity so small that most
    tiles fall back to the walk.  A low SKG_COVER_SUB_MAX makes the dense steps
    use the lane-per-mover kernel (and the lists) and the sparse tail the
    sub-wave kernel."""
    import os
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, '.');"
        "from skid_amd import gen, skg;"
        "b = gen.config1();"
        "ctx = skg.Context(b, skg.make_params(0.005, 32, period=1.0));"
        "tr, g, ng = ctx.run(); np.save(sys.argv[1], np.concatenate([tr.reshape(-1), g]))"
    )
    outs = []
    base = {"SKG_COVER_SUB_MAX": "2000"}
    for tag, env_add in (("direct", {}), ("graph", {"SKG_GRAPH": "1"}), ("nolists", {"SKG_TILE_LISTS": "0"}),
                         ("leaflists", {"SKG_TILE_LISTS": "2"}), ("smallcap", {"SKG_TILE_CAP": "64"}),
                         ("sepbuild", {"SKG_FUSED_BUILD": "0"}),
                         ("sepbuild_smallcap", {"SKG_FUSED_BUILD": "0", "SKG_TILE_CAP": "64"}),
                         ("graph_nolists", {"SKG_GRAPH": "1", "SKG_TILE_LISTS": "0"}),
                         ("stamped", {"SKG_FUSED_THR": "0"}), ("scan_cut", {"SKG_SORTED_CUT": "0"})):
        env = dict(os.environ)
        env.update(base)
        env.update(env_add)
        path = f"/tmp/skg_paths_{tag}.npy"
        subprocess.run([sys.executable, "-c", code, path], check=True, env=env, timeout=240,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        outs.append(np.load(path))
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


def test_scattercut_threshold_paths_agree_batched():
    """ScatterCut with the per-box threshold taken inside the move kernels
    (per-lane minima, atomics over per-box slots) and the cut over
    density-sorted entities equals the stamp-and-scan form (SKG_FUSED_THR=0,
    SKG_SORTED_CUT=0) on a batch of uneven boxes, for both move kernels: the
    per-box Ittr traces and the group ids are identical."""
    import os
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, '.');"
        "from skid_amd import gen, skg;"
        "sizes = [2500, 9000, 40000, 120000];"
        "boxes = [gen.config4_box(n, 30 + i) for i, n in enumerate(sizes)];"
        "taus = np.array([gen.config4_tau(n) for n in sizes], np.float32);"
        "bc = skg.BatchContext(boxes, skg.make_params(0.01, 32, period=1.0), taus);"
        "bc.density(); tr = bc.move_loop(); bc.fof(); bc.microsteps(); bc.centers(); bc.unbind(); bc.too_small();"
        "g, ng = bc.groups();"
        "np.save(sys.argv[1], np.concatenate([np.concatenate([np.asarray(t).reshape(-1) for t in tr]), g]))"
    )
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for sub_max in ("0", "100000000"):
        outs = []
        for tag, fused, cut in (("fused", "1", "1"), ("stamped", "0", "0")):
            env = dict(os.environ)
            env.update({"SKG_FUSED_THR": fused, "SKG_SORTED_CUT": cut, "SKG_COVER_SUB_MAX": sub_max})
            path = f"/tmp/skg_thr_{tag}_{sub_max}.npy"
            subprocess.run([sys.executable, "-c", code, path], check=True, env=env, timeout=240, cwd=root)
            outs.append(np.load(path))
        assert np.array_equal(outs[0], outs[1]), sub_max


def test_autotune_keeps_results(box1):
    """The autotuner picks valid shapes, and the path run with them gives the
    catalogue of the default shapes (kNN radii exactly; groups >= 99.9%)."""
    P = skg.make_params(0.005, 32, period=1.0)
    ref = skg.Context(box1, P)
    rb2, _, _ = ref.density()
    rtr, rg, rng = ref.run()
    ctx = skg.Context(box1, P)
    t = ctx.autotune(skg.SKG_TUNE_ALL)
    print("tuned", t)
    assert t["nBucket"] in (8, 16, 32, 64) and t["sub_max"] >= 0
    assert t["knn_bucket"] in (8, 16, 32, 64) and t["knn_waves"] in (1, 2)
    b2, _, _ = ctx.density()
    assert np.array_equal(b2, rb2)
    tr, g, ng = ctx.run()
    t2 = ctx.autotune(0)
    assert t2["pot_tile"] in (64, 128, 256)
    frac = O.match_fraction(g, rg)
    print(f"tuned {t2}: groups {ng} vs {rng}, match {frac:.5f}")
    assert frac >= 0.999


def test_leaves_of_128():
    """Leaves of 128 points (nBucket 128): the walks that stage a leaf's records
    through LDS do so in chunks of 64, which smaller leaves never reach.  FOF
    on such a tree equals brute force; the whole path's groups match the oracle
    run with the same leaf size; the dx adjoint agrees with the one on leaves of
    16 (the same sums in another order)."""
    b = gen.config1(n=8192, seed=21)
    ctx = skg.Context(b, skg.make_params(0.005, 32, nBucket=128, period=1.0))
    for tau in (0.002, 0.01):
        g, ng = ctx.fof_points(b.pos, tau)
        bg, bng = O.fof_brute(b.pos, tau)
        assert ng == bng + 1
        assert np.array_equal(g, bg)
    box = gen.config1()
    P = skg.make_params(0.005, 32, nBucket=128, period=1.0)
    c2 = skg.Context(box, P)
    tr, g, ng = c2.run()
    o = O.Oracle(box, nSmooth=32, nBucket=128, fTau=0.005, fPeriod=1.0)
    o.density()
    otr = o.move_loop()
    o.fof()
    og, ong = o.finish()
    frac = O.match_fraction(g, og)
    print(f"nBucket 128: groups gpu {ng - 1} oracle {ong - 1} match {frac:.5f}, trace {len(tr)} vs {len(otr)}")
    assert frac >= 0.999
    rng = np.random.default_rng(5)
    gbar = rng.normal(size=box.n).astype(np.float32)
    out = []
    for B in (16, 128):
        c3 = skg.Context(box, skg.make_params(0.005, 32, nBucket=B, period=1.0))
        c3.density()
        out.append(c3.adjoint(gbar, 3))
    for a, r in zip(out[1], out[0]):
        assert np.allclose(a, r, rtol=1e-4, atol=1e-5 * np.abs(r).max())
