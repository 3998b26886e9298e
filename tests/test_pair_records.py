"""Shared pair records of the wave kernel against the directed records.

The reference keeps one history record per directed slot (kernel.cu:413-552
evaluates (i, j) and (j, i) separately).  For a pair of equal radii the two
records stay exact mirrors, so the wave kernel may keep one per pair; these
tests hold that this changes nothing, bit for bit: trajectories, stored
forces, exported histories, across rebuilds, layout refreshes, exports in
the middle of a run and kernel switches, in FP64 and mixed precision, and
that a packing whose records are not mirrors falls back to the directed
records.
"""
import numpy as np
import pytest

import gpucode_amd as G
from gpucode_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def assert_same(a, b, hist=True):
    for u, v in zip(a.particles(), b.particles()):
        assert np.array_equal(u, v)
    for u, v in zip(a.forces(), b.forces()):
        assert np.array_equal(u, v)
    if hist:
        ha, hb = a.hist_pp(), b.hist_pp()
        for f in ha:
            assert np.array_equal(ha[f], hb[f]), f


def engines(w, precision, vel=None, relayout=64):
    a = G.Engine(w["inp"], w["dia"], precision)
    b = G.Engine(w["inp"], w["dia"], precision)
    for x in (a, b):
        x.set_kernel(G.Engine.KERNEL_WAVE)
        x.set_particles(w["pos"], vel)
        x.store_forces(True)
        x.set_relayout(relayout)
    b.set_pair_records(False)
    return a, b


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_MIXED])
def test_pair_records_equal_directed_records(precision):
    # random velocities: contacts forming and ending all the time, a
    # rebuild every few hundred steps and refreshes every 64 steps between
    w = W.small_bed(n=3000)
    vel = np.random.default_rng(4).normal(size=(3000, 3)) * 20.0
    a, b = engines(w, precision, vel)
    assert a.pair_records_used and not b.pair_records_used
    a.step(640)
    b.step(640)
    assert a.pair_records_used
    assert a.n_rebuild == b.n_rebuild and a.n_rebuild > 1
    assert a.n_relayout >= 3
    assert_same(a, b)


def test_pair_records_settling_bed_matches_oracle():
    # a dense bed settling under gravity: long-lived contacts, the
    # reference's law in FP64 against the oracle with zero difference
    w = W.small_bed(n=3000)
    a = G.Engine(w["inp"])
    o = O.Oracle(w["inp"])
    a.set_particles(w["pos"])
    o.set_particles(w["pos"])
    a.set_relayout(50)
    a.step(400)
    o.step(400)
    assert a.pair_records_used and a.step_kernel_name() == "k_step_wave"
    for u, v in zip(a.particles(), o.particles()):
        assert np.array_equal(u, v)
    ha, ho = a.hist_pp(), o.hist_pp()
    for f in ha:
        assert np.array_equal(ha[f], ho[f]), f


@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_MIXED])
def test_plastic_contacts_store_their_plastic_fields(precision):
    # yield at 0.05 % of d: most contacts deform plastically, so many pair
    # records keep plstRad/plstDfm explicitly, and unloading brings others
    # back to the implied values; the run still equals the directed one
    # and, in FP64, the oracle
    d = 1.0e-3
    n, R_d = 3000, 8.0
    H = d * (n / (np.pi * R_d * R_d * 0.5) + 6)
    inp = W.make_input(n, d, R_d * d, H, 0.5, cutGap=0.1, yield_frac=5e-4)
    w = dict(inp=inp, dia=None, pos=W.pocket_bed(n, R_d, 1.0, jitter=0.03, seed=7))
    vel = np.random.default_rng(6).normal(size=(n, 3)) * 20.0
    a, b = engines(w, precision, vel)
    a.step(400)
    b.step(400)
    assert a.pair_records_used
    ha = a.hist_pp()
    plastic = np.count_nonzero(ha["plstDfm"])
    assert plastic > 100, plastic
    assert_same(a, b)
    if precision == G.PREC_FP64:
        o = O.Oracle(inp)
        o.set_particles(w["pos"], vel)
        o.step(400)
        for u, v in zip(a.particles(), o.particles()):
            assert np.array_equal(u, v)


def test_export_and_kernel_switch_in_the_middle_of_a_run():
    # every export writes the pair records back and the next chunk makes
    # them again; a switch to the loop kernel and back does the same
    w = W.small_bed(n=3000)
    vel = np.random.default_rng(9).normal(size=(3000, 3)) * 20.0
    a, b = engines(w, G.PREC_FP64, vel, relayout=0)
    for k in range(6):
        a.step(37)
        b.step(37)
        ha, hb = a.hist_pp(), b.hist_pp()
        for f in ha:
            assert np.array_equal(ha[f], hb[f]), (k, f)
        if k == 2:
            a.set_kernel(G.Engine.KERNEL_LOOP)
        if k == 3:
            a.set_kernel(G.Engine.KERNEL_WAVE)
    st = a.stats()
    assert st["n_contact"] > 0
    assert_same(a, b)


def test_unequal_radii_keep_one_record_per_direction():
    # a dry polydisperse bed on one size level: pairs of unequal radii own
    # a record on each side, pairs of equal radii share one
    w = W.wet_bed(n=3000, poly=True, nLevels=1)
    w["inp"].forceCap = 0
    rng = np.random.default_rng(2)
    n = len(w["pos"])
    # three sizes: many neighbours of equal and of unequal radius
    w["dia"] = 1.0e-3 * rng.choice([0.8, 0.9, 1.0], n)
    w["dia"][0] = 1.0e-3
    vel = rng.normal(size=(n, 3)) * 20.0
    a, b = engines(w, G.PREC_FP64, vel)
    assert a.step_kernel_name() == "k_step_wave"
    a.step(300)
    b.step(300)
    assert a.pair_records_used
    assert_same(a, b)


def test_frequent_rebuilds_keep_the_directed_records():
    # a skin of 0.01 d: a rebuild every few dozen steps, too often for a
    # pack to pay; after the first rebuild interval is known the engine
    # keeps to the directed records, with the same results
    d = 1.0e-3
    n, R_d = 3000, 8.0
    H = d * (n / (np.pi * R_d * R_d * 0.5) + 6)
    inp = W.make_input(n, d, R_d * d, H, 0.5, cutGap=0.01)
    w = dict(inp=inp, dia=None, pos=W.pocket_bed(n, R_d, 1.0, jitter=0.03, seed=7))
    vel = np.random.default_rng(8).normal(size=(n, 3)) * 20.0
    a, b = engines(w, G.PREC_FP64, vel)
    assert a.pair_records_used
    a.step(300)
    b.step(300)
    assert a.n_rebuild == b.n_rebuild and a.n_rebuild >= 4
    assert not a.pair_records_used
    assert_same(a, b)


def test_records_that_are_not_mirrors_fall_back(tmp_path):
    # a restart whose history breaks the mirror property of one contact:
    # the pair records are refused and the directed records give the same
    # run as an engine that never used pair records
    w = W.small_bed(n=2000)
    src = G.Engine(w["inp"])
    src.set_particles(w["pos"])
    src.step(200)
    path = tmp_path / "dump.dat"
    src.dump(path)
    lines = path.read_text().splitlines()
    start = lines.index("NEIGHBOR LIST") + 2
    done = False
    for li in range(start, start + 2000):
        f = lines[li].split()
        num = int(f[0])
        for k in range(num):
            base = 1 + 10 * k
            if int(f[base + 8]) != 0 and float(f[base + 4]) > 0.0:  # a contact's nDsp
                f[base + 4] = repr(float(f[base + 4]) * (1.0 + 1e-9))
                lines[li] = " ".join(f)
                done = True
                break
        if done:
            break
    assert done
    path.write_text("\n".join(lines) + "\n")
    a = G.Engine(w["inp"])
    b = G.Engine(w["inp"])
    b.set_pair_records(False)
    for x in (a, b):
        x.load(path)
        x.store_forces(True)
    a.step(100)
    b.step(100)
    assert not a.pair_records_used
    assert_same(a, b)


def test_wave_kernel_steps_every_particle_once_at_100k():
    # 100,000 spheres: 1,563 chunks of 64, about 195 per XCD, of which the
    # last 48 run as half-waves and 8 of those as quarter-waves; the wave
    # kernel (with its XCD ranges, alternating direction and split tail)
    # must equal the one-thread loop kernel bit for bit
    w = W.small_bed(n=100_000, R_d=40.0)
    vel = np.random.default_rng(3).normal(size=(100_000, 3)) * 5.0
    a = G.Engine(w["inp"])
    b = G.Engine(w["inp"])
    b.set_kernel(G.Engine.KERNEL_LOOP)
    for x in (a, b):
        x.set_particles(w["pos"], vel)
        x.store_forces(True)
    a.step(51)
    b.step(51)
    assert a.step_kernel_name() == "k_step_wave" and a.pair_records_used
    assert_same(a, b, hist=False)
