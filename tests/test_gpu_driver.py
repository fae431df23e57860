"""The C driver (skid_amd/bin/skid) end to end: main.c's command line, tipsy
in, .grp/.gtp out, against the reference binary built from source
(oracle/_ref/skid), including the -unbind re-entry."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from skid_amd import gen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SKID = os.path.join(ROOT, "skid_amd", "bin", "skid")
REF = os.path.join(ROOT, "oracle", "_ref", "skid")


def read_grp(path):
    with open(path) as f:
        n = int(f.readline())
        g = np.loadtxt(f, dtype=np.int64, max_rows=n)
    assert g.size == n
    return g


def run(exe, args, stdin_path, cwd):
    with open(stdin_path, "rb") as f:
        r = subprocess.run([exe] + args, stdin=f, capture_output=True, text=True, timeout=300, cwd=cwd)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_driver_matches_reference_and_unbind_reentry(tmp_path):
    box = gen.config1()
    inp = tmp_path / "c1.std"
    gen.write_tipsy(str(inp), box, xdr=True)
    args = ["-tau", "0.005", "-s", "32", "-m", "8", "-p", "1", "-std"]
    out = run(SKID, args + ["-o", "gpu"], inp, tmp_path)
    assert "Number of Groups:" in out
    g = read_grp(tmp_path / "gpu.grp")
    assert os.path.getsize(tmp_path / "gpu.gtp") == 32 + 44 * (g.max())
    if os.path.exists(REF):
        run(REF, args + ["-o", "ref"], inp, tmp_path)
        rg = read_grp(tmp_path / "ref.grp")
        frac = O.match_fraction(g, rg)
        print(f"driver vs reference binary: groups {g.max()} vs {rg.max()}, match {frac:.5f}")
        assert frac >= 0.999
    # -unbind re-entry on our own output: the groups are already bound, so
    # unbinding them again changes nothing
    out = run(SKID, ["-unbind", "gpu.grp", "-tau", "0.005", "-m", "8", "-p", "1", "-std", "-o", "gpu_u"], inp, tmp_path)
    assert "Number of particles Unbound:0" in out
    gu = read_grp(tmp_path / "gpu_u.grp")
    assert O.same_partition(gu, g)
    if os.path.exists(REF):
        run(REF, ["-unbind", "gpu", "-tau", "0.005", "-m", "8", "-p", "1", "-std", "-o", "ref_u"], inp, tmp_path)
        ru = read_grp(tmp_path / "ref_u.grp")
        assert O.match_fraction(gu, ru) >= 0.999
