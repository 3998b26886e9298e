"""CPU tests of the engine's host side (no GPU calls): the library loads and
exports every symbol the headers in include/ declare, the host derivation equals
the oracle's, and the feed / Tecplot / deck / restart-header interfaces
behave as the reference's."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import gpucode_amd as G
from gpucode_amd import _capi as A
from gpucode_amd import workloads as W
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            txt = open(os.path.join(inc, h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b((?:dem|sur)_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = G.lib()
    for name in header_symbols():
        assert hasattr(L, name), name
    assert sorted(A.declared_symbols()) == header_symbols()


def test_struct_layouts_match_oracle():
    assert C.sizeof(A.DemInput) == C.sizeof(O.OrcInput)
    assert C.sizeof(A.DemSetup) == C.sizeof(O.OrcSetup)
    for (a, _), (b, _) in zip(A.DemSetup._fields_, O.OrcSetup._fields_):
        assert a == b and getattr(A.DemSetup, a).offset == getattr(O.OrcSetup, b).offset


def setups_equal(S1, S2):
    for name, typ in A.DemSetup._fields_:
        a, b = getattr(S1, name), getattr(S2, name)
        if hasattr(a, "_fields_"):
            for f, _ in a._fields_:
                assert getattr(a, f) == getattr(b, f), f"{name}.{f}"
        elif hasattr(a, "__len__"):
            la = [list(x) if hasattr(x, "__len__") else x for x in a]
            lb = [list(x) if hasattr(x, "__len__") else x for x in b]
            assert la == lb, name
        else:
            assert a == b, name


@pytest.mark.parametrize("make", [
    lambda: W.dry_column(),
    lambda: W.small_bed(n=2000),
    lambda: W.wet_poly(n=3000, R=8e-3, H=40e-3, nLevels=3),
    lambda: W.wet_poly(n=3000, R=8e-3, H=40e-3, nLevels=2, reference_law=True),
])
def test_host_derivation_equals_oracle(make):
    w = make()
    S1, r1, l1, i1 = G.derive(w["inp"], w["dia"])
    S2, r2, l2, i2 = O.derive(w["inp"], w["dia"])
    setups_equal(S1, S2)
    assert np.array_equal(r1, r2) and np.array_equal(l1, l2) and np.array_equal(i1, i2)


def test_derive_rejects_unassigned_level():
    w = W.wet_poly(n=2000, R=8e-3, H=40e-3, nLevels=2)
    inp = w["inp"]
    inp.sizeLevelInput = 1
    inp.levelDia[0] = 0.3e-3
    inp.levelDia[1] = 0.5e-3  # too small for the largest spheres
    with pytest.raises(G.DemError) as e:
        G.derive(inp, w["dia"])
    assert e.value.code == -4


def test_create_refuses_grid_clamp_quirk():
    # kernel.cu:627-629's clamp to gridSize is refused, before any device call
    import ctypes as C

    from gpucode_amd import _capi as A
    w = W.small_bed(n=200)
    inp = w["inp"]
    inp.gridClampQuirk = 1
    ctx = C.c_void_p()
    rc = A.lib().dem_create(C.byref(inp), None, 0, C.byref(ctx))
    assert rc == -1 and not ctx.value
    assert b"gridClampQuirk" in A.lib().dem_last_error()


def test_feed_law_and_clearance():
    w = W.dry_column(n=2000)
    S, rad, _, _ = G.derive(w["inp"])
    pos = np.zeros((2000, 3))
    n, seed = G.particle_feed(S, rad, pos, 0, 12345)
    assert n == 2000
    p = pos[:n]
    rho = np.hypot(p[:, 0] - S.cirX, p[:, 1] - S.cirY)
    assert np.all(rho <= S.R - rad[:n] + 1e-12)
    assert np.all((p[:, 2] >= S.feedLmin) & (p[:, 2] <= S.feedLmax))
    from scipy.spatial import cKDTree
    t = cKDTree(p)
    close = t.query_pairs(1.002 - 1e-12)
    assert len(close) == 0
    # uniform over the disc: the radius fraction has density 2x on [0, 1]
    q = rho / (S.R - 0.5)
    assert abs(np.mean(q) - 2 / 3) < 0.03 and abs(np.mean(q ** 2) - 0.5) < 0.03
    # the same seed gives the same packing; continuing the feed never overlaps
    pos2 = np.zeros((2000, 3))
    n2, _ = G.particle_feed(S, rad, pos2, 0, 12345)
    assert np.array_equal(pos2, pos)


def test_feed_matches_bruteforce_overlap_rule():
    # the grid-accelerated test accepts exactly what an O(N) scan accepts
    inp = W.make_input(400, 1e-3, 3e-3, 8e-3, 0.3)
    S, rad, _, _ = G.derive(inp)
    pos = np.zeros((400, 3))
    n, _ = G.particle_feed(S, rad, pos, 0, 99)
    p = pos[:n]
    d = np.sqrt(((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)) + np.eye(n) * 10
    assert d.min() >= 1.002
    assert n < 400  # the dense column forces the 500-reject give-up rule


def test_write_tec_format(tmp_path):
    path = tmp_path / "particle_info.dat"
    pos = np.array([[1.0, 2.0, 3.0], [-1.5, 0.25, 10.125]])
    rad = np.array([0.5, 0.25])
    vel = np.array([[0.1, 0.2, 0.3], [0.0, 0.0, -1.0]])
    w = np.array([[0.0, 3.0, 4.0], [1.0, 0.0, 0.0]])
    F = np.array([[0.0, 0.0, -1.0], [2.0, 0.0, 0.0]])
    G.write_tec(path, 0, 0.001234, pos, rad, vel, w, F)
    G.write_tec(path, 1, 0.002, pos, rad, vel, w, F)
    lines = open(path).read().split("\n")
    assert lines[0] == 'TITLE = " PARTICLE INFORMATION " '
    assert lines[1] == "VARIABLES = X   Y   Z   R   VX   VY   VZ   W   F"
    assert lines[2] == 'ZONE T= "     0.001234 s " '
    row = "%11.5f   %11.5f   %11.5f   %11.5f  %11.5f   %11.5f   %11.5f   %11.5f   %11.5f" % (
        1.0, 2.0, 3.0, 0.5, 0.1, 0.2, 0.3, 5.0, 1.0)
    assert lines[3] == row
    assert lines[5] == 'ZONE T= "     0.002000 s " '
    assert len(lines) == 9


def test_read_deck_and_load_para(tmp_path):
    deck = tmp_path / "heyi.in"
    deck.write_text("""NUMPARTICLE
2000
SIZELEVEL
1
LOADMODE
1
packing_TIME
0.5
FoceCapillary
No
REALSIZE
0.001
CUTGAPFORNEIGHBORLIST
0.0714285714
EXPANDSIZE
2.0
ContainerType
Cylinder
CylinderBC
0.0 0.0 0.0 0.01 0.06
FeedZone
0.0005 0.0595
MATERIAL
2500 7e10 0.25 0.3 0.01 2.26e-7 9.5e9
2500 7e10 0.25 0.3 0.01 2.26e-7 9.5e9
LIQUID
0.072 0.01 1e-6
SIMULATION
0.6 10
SIZELEVEL_INPUT
No
""")
    inp = G.read_deck(deck)
    assert inp.NP == 2000 and inp.nLevels == 1 and inp.forceCap == 0
    assert inp.R == 0.01 and inp.H == 0.06 and inp.mat[1][3] == 0.3
    assert inp.dtFactor == 0.6 and inp.nOutPacking == 10 and inp.sizeLevelInput == 0
    S, rad, _, _ = G.derive(inp)
    assert S.levelNNebors[0] == 34
    dump = tmp_path / "heyi_dump.dat"
    dump.write_text("junk\nSYSTEM Info\n2000 3 1500\n")
    assert G.load_para(dump) == (2000, 3, 1500)
    bad = tmp_path / "bad.in"
    bad.write_text("NUMPARTICLE\n10\n")
    with pytest.raises(G.DemError):
        G.read_deck(bad)


def test_run_log_format(tmp_path):
    # main.cpp:93-139, 591-618, 631-634
    from gpucode_amd import _capi as A
    path = str(tmp_path / "heyi_Log.dat").encode()
    L = A.lib()
    assert L.dem_log_header(path, b"heyi", 0, 2000, 1, 500) == 0
    assert L.dem_log_output(path, 0, 0.0125, 812.5, 3, 500, 2000, 12500) == 0
    assert L.dem_log_output(path, 1, 0.025, 1600.25, 1, 2000, 2000, 25000) == 0
    assert L.dem_log_total(path, 1.6) == 0
    t = open(path.decode()).read()
    assert t.startswith("==================|*--- PROJECT---*|===================\n<File>: new\n<Time>: ")
    assert "<Name>: heyi\n" in t and "<Particle>: 2000 \n" in t
    assert ("0    \truntime:0.0125 \telapsed:812.50000           ms\trebuild:3\t"
            "particle:500\tIterate:12500\n") in t
    assert ("1    \truntime:0.0250 \telapsed:1600.25000          ms\trebuild:1\t"
            "Iterate:25000\n") in t
    assert t.endswith("\nGPU DEM Time used: 1.60         s \n\n")
    assert L.dem_log_header(path, b"heyi", 1, 2000, 3, 1500) == 0
    t = open(path.decode()).read()
    assert "<File>: restarted" in t and "<Feed batch>: 3 \t<Current particle>: 1500 \n" in t


def test_calibration_round_trip():
    from gpucode_amd.calibrate import collision_restitution, restitution_to_dmpn
    d = restitution_to_dmpn(0.5, 1e-3, 2500, 70e9, 0.25, v=0.5)
    assert collision_restitution(d, 1e-3, 2500, 70e9, 0.25, 0.5) == pytest.approx(0.5, abs=1e-4)
    assert collision_restitution(0.0, 1e-3, 2500, 70e9, 0.25, 0.5) == pytest.approx(1.0, abs=1e-3)


def test_bench_refuses_more_than_one_gpu():
    # the engine steps one packing on one GPU: asked for more, the bench
    # exits with status 2 and says so before touching a device
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and r.stdout == ""
    assert "one GPU only" in r.stderr
