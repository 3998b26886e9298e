"""The reference's driver loop as a compiled C++ program against the C ABI
(gpucode_amd/examples/dem_main.cpp: deck, feed, steps to the packing time,
Tecplot output and the run log of main.cpp:65-640), run on a small deck and
checked against the same sequence driven from Python, bit for bit."""
import os
import re
import subprocess

import numpy as np
import pytest

import gpucode_amd as G
from gpucode_amd import build as B
from gpucode_amd.engine import particle_feed

DECK = """NUMPARTICLE
1500
SIZELEVEL
1
LOADMODE
1
packing_TIME
2e-4
FoceCapillary
No
REALSIZE
0.001
CUTGAPFORNEIGHBORLIST
0.1
EXPANDSIZE
2.0
ContainerType
Cylinder
CylinderBC
0.0 0.0 0.0 0.008 0.04
FeedZone
0.0005 0.0395
MATERIAL
2500 7e10 0.25 0.3 0.01 2.26e-7 9.5e9
2500 7e10 0.25 0.3 0.01 2.26e-7 9.5e9
LIQUID
0.072 0.01 1e-6
SIMULATION
1.0 4
SIZELEVEL_INPUT
No
TIMESTEP
1e-7
"""


@pytest.mark.gpu
def test_compiled_driver_matches_python(tmp_path):
    deck = tmp_path / "heyi.in"
    deck.write_text(DECK)
    gen = str(tmp_path / "heyi")
    state = tmp_path / "state.bin"
    assert os.path.exists(B.EXAMPLE), "build the library (and its example) first"
    r = subprocess.run([B.EXAMPLE, str(deck), gen, str(state)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    # the same run from Python: deck, seeded feed, the same chunks of steps
    inp = G.read_deck(deck)
    S, rad, _, _ = G.derive(inp)
    pos = np.zeros((inp.NP, 3))
    total, seed = 0, 1
    while total < inp.NP:
        n, seed = particle_feed(S, rad, pos, total, seed)
        if n <= total:
            break
        total = n
    e = G.Engine(inp)
    e.set_particles(pos[:total])
    e.store_forces(True)
    per_out = int(np.ceil(S.tPacking / S.dt / inp.nOutPacking))
    outs = 0
    while e.ctime < S.tPacking:
        e.step(per_out)
        outs += 1
    got = np.fromfile(state, np.float64).reshape(3, total, 3)
    for a, b in zip(got, e.particles()):
        assert np.array_equal(a, b)
    # one log record per output and the total (main.cpp:591-634)
    log = open(gen + "_Log.dat").read()
    recs = re.findall(r"^(\d+)\s*\truntime:[\d.]+\s*\telapsed:[\d.]+\s*ms\trebuild:\d+\tIterate:(\d+)$",
                      log, re.M)
    assert len(recs) == outs and recs[-1][1] == str(e.iter)
    assert "GPU DEM Time used:" in log and "<File>: new" in log
    # one Tecplot zone per output
    assert open(gen + ".dat").read().count("ZONE") == outs
    e.close()
