"""Batched multi-instance stepping and the packing analysis.

GPU bars:
  * a batch of uneven instances (dry and wet, mono- and polydisperse, with
    rebuilds) steps every instance bit-identically to the same instance
    stepped alone (same kernels, same per-instance rebuild steps);
  * phi, Z and the bed top from the analysis kernel equal an independent
    CPU evaluation (scipy pair search for Z, fine midpoint quadrature of the
    sphere/core-region intersection for phi) within 1e-6 relative.
The CPU reference of the intersection volume is itself pinned here against
closed forms (full sphere, hemisphere, spherical cap, disc-disc lens limits).
"""
import math

import numpy as np
import pytest

import gpucode_amd as G
from gpucode_amd import workloads as W


# --- CPU reference of the analysis ------------------------------------------
def lens_area(s, Rc, d):
    """Area of disc(radius s, centre at distance d) inside disc(Rc, origin)."""
    s = np.asarray(s, float)
    out = np.zeros_like(s)
    inside = d <= np.abs(Rc - s)
    out[inside] = np.pi * np.minimum(s[inside], Rc) ** 2
    part = (~inside) & (d < s + Rc) & (s > 0)
    sp = s[part]
    c1 = np.clip((d * d + sp * sp - Rc * Rc) / (2 * d * sp), -1, 1)
    c2 = np.clip((d * d + Rc * Rc - sp * sp) / (2 * d * Rc), -1, 1)
    k = (-d + sp + Rc) * (d + sp - Rc) * (d - sp + Rc) * (d + sp + Rc)
    out[part] = sp * sp * np.arccos(c1) + Rc * Rc * np.arccos(c2) - 0.5 * np.sqrt(np.maximum(k, 0))
    return out


def core_volume_ref(rho, z, r, Rc, zlo, zhi, nslice=20000):
    a, b = max(z - r, zlo), min(z + r, zhi)
    if a >= b or rho - r >= Rc:
        return 0.0
    t = a + (np.arange(nslice) + 0.5) * (b - a) / nslice
    s = np.sqrt(np.maximum(r * r - (t - z) ** 2, 0.0))
    return float(lens_area(s, Rc, rho).sum() * (b - a) / nslice)


def analyze_ref(pos, rad, S, Tw):
    from scipy.spatial import cKDTree

    n = len(pos)
    tree = cKDTree(pos)
    pairs = np.array(sorted(tree.query_pairs(2 * rad.max())), dtype=np.int64).reshape(-1, 2)
    if len(pairs):
        d = np.linalg.norm(pos[pairs[:, 0]] - pos[pairs[:, 1]], axis=1)
        pairs = pairs[d < rad[pairs[:, 0]] + rad[pairs[:, 1]]]
    touch = np.zeros(n, bool)
    touch[pairs.ravel()] = True
    rho = np.hypot(pos[:, 0] - S.cirX, pos[:, 1] - S.cirY)
    touch |= pos[:, 2] - S.Bw - rad < 0          # floor
    touch |= (S.R - rho) - rad < 0                # lateral wall
    touch |= (Tw - pos[:, 2]) - rad < 0           # lid
    if not touch.any():
        return 0.0, 2.0 * len(pairs) / n, -1.0e300
    top = (pos[:, 2] + rad)[touch].max()
    Rc, zlo, zhi = S.R - 2.0, S.Bw + 2.0, top - 2.0
    vol = sum(core_volume_ref(rho[i], pos[i, 2], rad[i], Rc, zlo, zhi) for i in range(n))
    phi = vol / (math.pi * Rc * Rc * (zhi - zlo))
    return phi, 2.0 * len(pairs) / n, top


def test_volume_reference_closed_forms():
    r = 0.5
    full = 4 / 3 * math.pi * r ** 3
    assert abs(core_volume_ref(0.0, 5.0, r, 3.0, 0.0, 10.0) - full) < 1e-7 * full
    # cut through the centre by the slab's lower face: a hemisphere
    assert abs(core_volume_ref(0.0, 5.0, r, 3.0, 5.0, 10.0) - full / 2) < 1e-7 * full
    # a cap of height h = 0.2: pi h^2 (3r - h) / 3
    h = 0.2
    cap = math.pi * h * h * (3 * r - h) / 3
    assert abs(core_volume_ref(0.0, 5.0, r, 3.0, 5.0 + r - h, 10.0) - cap) < 1e-6 * cap
    # outside the cylinder: nothing; centred on a huge cylinder's wall: about half
    assert core_volume_ref(3.6, 5.0, r, 3.0, 0.0, 10.0) == 0.0
    v = core_volume_ref(1000.0, 5.0, r, 1000.0, 0.0, 10.0)
    assert abs(v - full / 2) < 1e-3 * full
    # lens limits: disjoint, contained, and two equal discs at distance 0
    assert lens_area(np.array([1.0]), 2.0, 3.5)[0] == 0.0
    assert abs(lens_area(np.array([0.5]), 2.0, 1.0)[0] - math.pi * 0.25) < 1e-15
    assert abs(lens_area(np.array([1.0]), 1.0, 0.0)[0] - math.pi) < 1e-15
    # equal discs at distance s: 2 s^2 acos(1/2) - s^2 sqrt(3)/2
    s = 1.0
    want = 2 * s * s * math.acos(0.5) - 0.5 * s * s * math.sqrt(3.0)
    assert abs(lens_area(np.array([s]), s, s)[0] - want) < 1e-14


# --- GPU ----------------------------------------------------------------------
def _instances():
    rng = np.random.default_rng(11)
    out = []
    w = W.small_bed(n=2000, seed=3)
    out.append((w, rng.normal(size=(2000, 3)) * 50.0))
    w = W.small_bed(n=1500, seed=4, mu=0.2)
    out.append((w, rng.normal(size=(1500, 3)) * 20.0))
    w = W.wet_bed(n=1200, poly=True, reference_law=False, theta_deg=20.0, liquid=0.2)
    out.append((w, None))
    w = W.dry_column(n=2000)
    out.append((w, None))
    w = W.small_bed(n=300, R_d=4.0, seed=5)
    out.append((w, rng.normal(size=(300, 3)) * 30.0))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["device", "host"])
@pytest.mark.parametrize("precision", [G.PREC_FP64, G.PREC_MIXED])
def test_batch_equals_separate_runs(precision, schedule):
    # device: rebuild requests acted on by the kernels at rebuild slots
    # every 2 steps, no host round trip per rebuild; host: chunks of steps
    # with the requests read back.  Either gives every instance exactly its
    # own single-instance sequence of steps and rebuilds.
    cases = _instances()
    b = G.Batch([w["inp"] for w, _ in cases], [w["dia"] for w, _ in cases], precision)
    b.set_schedule(device=schedule == "device", slot_every=2)
    singles = []
    for k, (w, vel) in enumerate(cases):
        e = G.Engine(w["inp"], w["dia"], precision)
        if k == 2:  # the wet instance on the one-thread loop kernel: both kernels run
            b[k].set_kernel(G.Engine.KERNEL_LOOP)
            e.set_kernel(G.Engine.KERNEL_LOOP)
        b[k].set_particles(w["pos"], vel)
        e.set_particles(w["pos"], vel)
        singles.append(e)
    kernels = [e.step_kernel_name() for e in singles]
    assert any(k.startswith("k_step_wave") for k in kernels) and "k_step_fused" in kernels
    steps = 300
    b.step(steps)
    for e in singles:
        e.step(steps)
    for k, e in enumerate(singles):
        print(f"instance {k}: {kernels[k]}, rebuilds {e.n_rebuild}")
        assert b[k].iter == e.iter == steps
        assert b[k].n_rebuild == e.n_rebuild
        assert b[k].ctime == e.ctime
        for u, v in zip(b[k].particles(), e.particles()):
            assert np.array_equal(u, v), k
        hb, he = b[k].hist_pp(), e.hist_pp()
        for f in hb:
            assert np.array_equal(hb[f], he[f]), (k, f)
        assert np.array_equal(b[k].lqdcnt(), e.lqdcnt())
    assert sum(e.n_rebuild for e in singles) > 2 * len(singles)
    t = b.timing()
    assert t["launches"] >= steps and t["n_rebuild"] >= len(singles)
    st = b.schedule_stats()
    if schedule == "device":
        # one rebuild slot per 2 steps; the host looks only at the end of
        # a call (after each block once the steps are issued) and every 256
        # blocks, not per rebuild
        assert st["slots"] >= steps // 2 and st["host_syncs"] <= 2 * 4 + st["slots"] // 256
        assert st["host_syncs"] < t["n_rebuild"]
    else:
        assert st["slots"] == 0
    # a second call continues from where the first stopped
    b.step(50)
    for e in singles:
        e.step(50)
    for k, e in enumerate(singles):
        for u, v in zip(b[k].particles(), e.particles()):
            assert np.array_equal(u, v), k
    b.close()


@pytest.mark.gpu
def test_batch_reports_the_overflowing_instance():
    # three instances rebuilt by one batched rebuild; the middle one has 60
    # spheres crammed together, more neighbours than its lists hold: the
    # step fails with the overflow status and names that instance, and
    # nothing is written past the lists
    ws = [W.small_bed(n=200, seed=s) for s in (3, 4, 5)]
    pos = [w["pos"].copy() for w in ws]
    pos[1][:60] = pos[1][0] + np.random.default_rng(0).normal(size=(60, 3)) * 0.01
    b = G.Batch([w["inp"] for w in ws])
    for k in range(3):
        b[k].set_particles(pos[k])
    with pytest.raises(G.DemError) as err:
        b.step(1)
    assert err.value.code == -3 and "instance 1:" in str(err.value)
    # a second step does not run on the truncated lists either
    with pytest.raises(G.DemError) as err:
        b.step(1)
    assert err.value.code == -3 and "instance 1:" in str(err.value)
    b.close()


def converging_cluster(n=150, rad=3.9, steps_to_centre=20.0, seed=1):
    """n spheres in a ball of radius rad (sphere diameters) in the middle of
    small_bed's container, each moving to the centre in steps_to_centre
    steps: the lists are rebuilt every step and, about 14 steps in, a
    rebuild finds more neighbours than the lists hold (the oracle run in
    test_converging_cluster_overflows_in_the_oracle shows where)."""
    from oracle import pyoracle as O

    w = W.small_bed(n=300, seed=3)
    S = O.derive(w["inp"])[0]
    rng = np.random.default_rng(seed)
    c = np.array([S.cirX, S.cirY, S.Bw + 0.5 * (S.Tw - S.Bw)])
    pts = []
    while len(pts) < n:
        p = c + rng.uniform(-rad, rad, 3)
        if np.linalg.norm(p - c) <= rad and all(np.linalg.norm(p - q) > 1.05 for q in pts):
            pts.append(p)
    pos = np.array(pts)
    return w, pos, -(pos - c) / (steps_to_centre * S.dt)


def test_converging_cluster_overflows_in_the_oracle():
    from oracle import pyoracle as O

    w, pos, vel = converging_cluster()
    o = O.Oracle(w["inp"])
    o.set_particles(pos, vel)
    with pytest.raises(ValueError, match="overflow"):
        o.step(40)
    assert 5 <= o.iter < 40 and o.n_rebuild == o.iter + 1
    o.close()


@pytest.mark.gpu
def test_device_schedule_reports_overflow_at_a_rebuild_slot():
    # instance 1's lists overflow at a rebuild the kernels make in the middle
    # of the run (its first rebuild, made by the host, is fine): the step
    # reports that instance, which halts at that rebuild, as the single
    # instance does; the other instances run all their steps unharmed
    wc, pc, vc = converging_cluster()
    cases = [(W.small_bed(n=400, seed=4), None), (wc, (pc, vc)),
             (W.wet_bed(n=600, poly=True, reference_law=False, theta_deg=20.0, liquid=0.2), None)]
    b = G.Batch([w["inp"] for w, _ in cases], [w["dia"] for w, _ in cases])
    b.set_schedule(device=True, slot_every=2)
    singles = []
    for k, (w, pv) in enumerate(cases):
        pos, vel = pv if pv is not None else (w["pos"], None)
        b[k].set_particles(pos, vel)
        e = G.Engine(w["inp"], w["dia"])
        e.set_particles(pos, vel)
        singles.append(e)
    with pytest.raises(G.DemError) as err:
        b.step(40)
    assert err.value.code == -3 and "instance 1:" in str(err.value)
    with pytest.raises(G.DemError) as err1:
        singles[1].step(40)
    assert err1.value.code == -3
    assert 5 <= b[1].iter == singles[1].iter < 40
    assert b[1].n_rebuild == singles[1].n_rebuild
    for k in (0, 2):
        singles[k].step(40)
        assert b[k].iter == singles[k].iter == 40
        for u, v in zip(b[k].particles(), singles[k].particles()):
            assert np.array_equal(u, v), k
    assert b.schedule_stats()["slots"] >= 15
    with pytest.raises(G.DemError) as err:
        b.step(1)
    assert err.value.code == -3 and "instance 1:" in str(err.value)
    for e in singles:
        e.close()
    b.close()


@pytest.mark.gpu
def test_analysis_matches_cpu_reference():
    # a bed settled for a while: contacts, wall contacts and a free top
    w = W.small_bed(n=3000, seed=9)
    e = G.Engine(w["inp"], w["dia"])
    e.set_particles(w["pos"])
    e.step(400)
    phi, Z, top = e.analyze()
    pos = e.particles()[0]
    rad = np.full(len(pos), 0.5)
    S = e.setup
    phi_r, Z_r, top_r = analyze_ref(pos, rad, S, S.Tw)
    print(phi, phi_r, Z, Z_r, top, top_r)
    assert top == top_r
    assert Z == Z_r
    assert abs(phi - phi_r) <= 1e-6 * phi_r
    assert 0.45 < phi < 0.75 and 3.0 < Z < 8.0


@pytest.mark.gpu
def test_batch_analysis_equals_single():
    cases = _instances()[:3]
    b = G.Batch([w["inp"] for w, _ in cases], [w["dia"] for w, _ in cases])
    for k, (w, vel) in enumerate(cases):
        b[k].set_particles(w["pos"])
    b.step(100)
    phi, Z, top = b.analyze()
    for k in range(len(cases)):
        p1, z1, t1 = b[k].analyze()
        assert (phi[k], Z[k], top[k]) == (p1, z1, t1)
    for k in (0, 2):
        pos = b[k].particles()[0]
        S = b[k].setup
        _, rad, _, _ = G.derive(cases[k][0]["inp"], cases[k][0]["dia"])
        phi_r, Z_r, top_r = analyze_ref(pos, rad, S, S.Tw)
        assert Z[k] == Z_r and top[k] == top_r, k
        assert abs(phi[k] - phi_r) <= 1e-6 * max(phi_r, 1e-300), k
    # the polydisperse bed sits on lattice sites spaced for the largest
    # size: after 100 steps nothing touches yet, so there is no bed top
    assert top[2] == -1.0e300 and phi[2] == 0.0
    b.close()


@pytest.mark.gpu
def test_autotune_times_candidates_and_keeps_the_physics():
    w = W.small_bed(n=2000, seed=13)
    rng = np.random.default_rng(2)
    vel = rng.normal(size=(2000, 3)) * 20.0
    best, table = G.autotune(w["inp"], w["pos"], vel=vel, steps=100, cut_gaps=[0.1, 0.3])
    print(best, table)
    assert best["n_tried"] == 12 and (table > 0).all()
    k = {1: 0, 2: 1, 3: 2}[best["kernel"]]
    assert table.min() == best["ms_per_step"] == table[[0.1, 0.3].index(best["cutGap"]), k,
                                                       best["morton"]]
    assert best["ms_default"] > 0
    # every candidate steps the same physics, to rounding: the cell size and
    # list contents change only the summation order.  (Checked on the bed
    # settling quietly: with the fast collisions used for timing, rounding
    # differences grow chaotically.)
    runs = []
    for cg, kern, mo in ((0.1, G.Engine.KERNEL_AUTO, True), (0.3, G.Engine.KERNEL_LOOP, False),
                         (best["cutGap"], best["kernel"], bool(best["morton"]))):
        w["inp"].cutGap = cg
        e = G.Engine(w["inp"])
        e.set_kernel(kern)
        e.set_order(mo)
        e.set_particles(w["pos"])
        e.step(100)
        runs.append(e.particles()[0])
    w["inp"].cutGap = 0.1
    assert np.abs(runs[0] - runs[1]).max() < 1e-9
    assert np.abs(runs[0] - runs[2]).max() < 1e-9


@pytest.mark.gpu
def test_autotune_skips_skins_shorter_than_capillary_bridges():
    w = W.wet_bed(n=1500, liquid=0.05)
    best, table = G.autotune(w["inp"], w["pos"], w["dia"], steps=50, cut_gaps=[0.01, 0.5])
    # rupture distance (1 + 0) * 0.99 * (2 * 0.05 / 12)^0.33 ~ 0.2 d: the 0.01 skin is out
    assert (table[0] == -1).all()
    # wet packings run the one-thread loop and the wave kernel, not the
    # block-compaction kernel
    assert (table[1, 0] > 0).all() and (table[1, 1] == -1).all() and (table[1, 2] > 0).all()
    assert best["cutGap"] == 0.5
    assert best["kernel"] in (G.Engine.KERNEL_LOOP, G.Engine.KERNEL_WAVE)


@pytest.mark.gpu
def test_autotuned_batched_instances_equal_the_oracle():
    # four instances of the batched uneven set at 1,500 spheres each (dry
    # monodisperse, dry three-level polydisperse, wet at theta 54 deg, wet
    # polydisperse at liquid 3 %), each with the skin, step kernel and
    # storage order the autotuner picks, stepped together in FP64 for 1,000
    # steps of their pour (rebuilds included) by the device-driven schedule:
    # every instance follows the oracle run with the same input (lists
    # identical, states within FP64 rounding of the oracle's)
    from oracle import pyoracle as O

    ks = (30, 16, 34, 5)
    ws = [W.c4_instance(k, n=1500) for k in ks]
    assert ws[0]["inp"].forceCap == 0 and ws[1]["inp"].forceCap == 0 and ws[1]["inp"].nLevels == 3
    assert ws[2]["inp"].forceCap == 1 and ws[3]["inp"].forceCap == 1
    choices = G.tune_instances(ws, G.PREC_FP64, steps=50)
    b = G.Batch.from_instances(ws, choices, G.PREC_FP64)
    b.set_schedule(device=True, slot_every=4)
    steps = 1000
    b.step(steps)
    for k, w in enumerate(ws):
        inp = G.DemInput.from_buffer_copy(w["inp"])
        inp.cutGap = choices[k]["cutGap"]
        o = O.Oracle(inp, w["dia"])
        o.set_particles(w["pos"], w["vel"])
        o.step(steps)
        e = b[k]
        print(ks[k], choices[k], "rebuilds", e.n_rebuild, o.n_rebuild)
        assert e.iter == o.iter == steps and e.n_rebuild == o.n_rebuild
        p1, v1, w1 = e.particles()
        p2, v2, w2 = o.particles()
        assert np.abs(p1 - p2).max() <= 1e-9
        assert np.abs(v1 - v2).max() <= 1e-7 * max(np.abs(v2).max(), 1e-12)
        assert np.abs(w1 - w2).max() <= 1e-6 * max(np.abs(w2).max(), 1e-12)
        n1, l1 = e.nebor()
        n2, l2 = o.nebor()
        assert np.array_equal(n1, n2) and np.array_equal(l1, l2)
        if w["inp"].forceCap == 1:
            assert np.array_equal(e.lqdcnt(), o.lqdcnt())
        o.close()
    b.close()
