"""FP32 pair arithmetic with FP64 state (DEM_PREC_MIXED) against the FP64
oracle, at the sizes of the two workloads that ask for it.

  * 2,000,000 dry spheres (d = 0.3 mm, cylinder radius 100 d, skin 0.1 d,
    mu 0.5): the throughput bed, on the path the benchmark times with
    --precision mixed (k_step_wave_mixed, shared pair records);
  * 200,000 wet spheres with lognormal diameters (median 0.5 mm, sigma 0.25,
    clipped 0.25-1.0 mm), three size levels, liquid at 1 % of the mean
    particle volume, gamma 0.072 N/m, contact angle 20 degrees, rupture at
    (1 + theta/2) V^(1/3): poured from the seeded feed and settled on the
    GPU (FP64) first, so that contacts, plastic contacts and capillary
    bridges all act.  The spheres are soft (E = 0.1 MPa) so the pour settles
    in 30,000 steps; the contact law, the liquid and the size levels are
    those of the wet workload, on the one-thread loop kernel.

Bars (the contact law of kernel.cu:89-228 and the capillary block of
kernel.cu:509-550 evaluated in FP32, against the same law in FP64):
  * after one step, per particle, |F - F_64| / (sum_j |f_ij| + m g) <= 1e-5,
    and the moment |M - M_64| / ((sum_j |f_ij| + m g) r) <= 1e-5;
  * the stored history after that step equals FP64 bit for bit: every value
    that is stored or decides a branch is computed in FP64;
  * over N steps trajectories part: the reference law takes branches on
    last-bit comparisons (the load/unload verdict of a resting contact,
    kernel.cu:104-111; stick or slip at the Coulomb cap, kernel.cu:190-217),
    so any perturbation grows.  The bar: the mixed run drifts from the
    oracle at most SENSITIVITY_FACTOR times as far as the FP64 law drifts
    from its own state moved by one ulp, in positions, velocities, spins,
    springs, overlaps, load modes, lists, bridge flags and counts.

With DEM_ERROR_TABLE set to a path, every test appends its measured errors
to that JSON-lines file.
"""
import json
import math
import os

import numpy as np
import pytest

import gpucode_amd as G
from gpucode_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

FORCE_BAR = 1e-5
TRAJ_STEPS_2M = (1, 10, 40)
TRAJ_STEPS_WET = (1, 10, 50, 200)
# Over many steps the mixed run may drift from the oracle at most this many
# times as far as the FP64 law drifts from a copy of its state moved by one
# ulp (below these floors the drift is not counted).
SENSITIVITY_FACTOR = 10.0
DRIFT_FLOORS = {"x_abs_d": 1e-12, "v_rel": 1e-6, "w_rel": 1e-6, "dispt_rel": 1e-6,
                "nDsp_abs_d": 1e-12, "stage_mismatch_frac": 1e-5}


def record(name, d):
    path = os.environ.get("DEM_ERROR_TABLE")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"case": name, **d}) + "\n")
    print(name, json.dumps(d), flush=True)


def force_scale(o, rad, chunk=200_000):
    """sum_j |f_ij| + m g per particle, bounded from the oracle's state: the
    Hertz force of each overlap times (1 + mu), 2 pi R gamma for each bridged
    slot, and the weight (2r)^3 (kernel.cu:439)."""
    S = o.S
    pos = o.particles()[0]
    num, lst = o.nebor()
    h = o.hist_pp() if S.forceCap == 1 else None
    n, K = lst.shape
    out = np.empty(n)
    kk = np.arange(K)[None, :]
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m = kk < num[a:b, None]
        j = np.where(m, lst[a:b], 0)
        d = np.linalg.norm(pos[a:b, None, :] - pos[j], axis=2)
        ri, rj = rad[a:b, None], rad[j]
        rs = ri * rj / (ri + rj)
        ov = np.where(m, ri + rj - d, 0.0)
        f = np.where(ov > 0, 4.0 / 3.0 * S.pp.emod * np.sqrt(rs) * np.maximum(ov, 0) ** 1.5
                     * (1.0 + S.pp.sfrc), 0.0)
        if h is not None:
            brg = (h["lqdBrg"][a:b] == 1) & m
            f = f + np.where(brg, 2.0 * math.pi * (2.0 * rs) * S.stension, 0.0)
        out[a:b] = f.sum(1) + (2.0 * rad[a:b]) ** 3
    return out


def contact_counts(x, rad, chunk=200_000):
    """p-p contacts (gap < 0) and bridged slots of a run, from its exported
    positions, lists and bridge flags."""
    pos = x.particles()[0]
    num, lst = x.nebor()
    brg = x.hist_pp()["lqdBrg"]
    n, K = lst.shape
    kk = np.arange(K)[None, :]
    nc = 0
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m = kk < num[a:b, None]
        j = np.where(m, lst[a:b], 0)
        d = np.linalg.norm(pos[a:b, None, :] - pos[j], axis=2)
        nc += int((m & (d < rad[a:b, None] + rad[j])).sum())
    return nc, int((brg == 1).sum())


def one_step_errors(e, o, rad):
    F1, M1, Mr1 = e.forces()
    F2, M2, Mr2 = o.forces()
    sc = force_scale(o, rad)
    eF = np.linalg.norm(F1 - F2, axis=1) / sc
    eM = np.linalg.norm(M1 - M2, axis=1) / (sc * rad)
    eMr = np.abs(Mr1 - Mr2) / max(np.abs(Mr2).max(), 1e-300)
    return {"F": float(eF.max()), "F_mean": float(eF.mean()), "M": float(eM.max()),
            "Mr_rel_max": float(eMr.max())}


def state_errors(e, o):
    p1, v1, w1 = e.particles()
    p2, v2, w2 = o.particles()
    r = {"x_abs_d": float(np.abs(p1 - p2).max()),
         "v_rel": float(np.abs(v1 - v2).max() / max(np.abs(v2).max(), 1e-300)),
         "w_rel": float(np.abs(w1 - w2).max() / max(np.abs(w2).max(), 1e-300))}
    del p1, v1, w1, p2, v2, w2
    n1, l1 = e.nebor()
    n2, l2 = o.nebor()
    r["list_mismatch_frac"] = float(np.mean(np.any(l1 != l2, axis=1) | (n1 != n2)))
    same = np.array_equal(l1, l2)
    del l1, l2
    h1, h2 = e.hist_pp(), o.hist_pp()
    act = h2["stage"] > 0
    nact = max(int(act.sum()), 1)
    r["active_slots"] = int(act.sum())
    r["stage_mismatch_frac"] = float((h1["stage"] != h2["stage"]).sum() / nact)
    r["lqdBrg_mismatch_frac"] = float((h1["lqdBrg"] != h2["lqdBrg"]).sum() / max(
        int((h2["lqdBrg"] > 0).sum()), 1))
    both = act & (h1["stage"] > 0)
    for f in ("nDsp", "nDsp_mx", "plstDfm"):
        r[f + "_abs_d"] = float(np.abs(h1[f][both] - h2[f][both]).max()) if both.any() else 0.0
    pr = h2["plstRad"][both]
    r["plstRad_rel"] = float(np.abs(h1["plstRad"][both] - pr).max() / max(np.abs(pr).max(), 1e-300)) \
        if both.any() else 0.0
    d1, d2 = h1["dispt"][both], h2["dispt"][both]
    r["dispt_rel"] = float(np.abs(d1 - d2).max() / max(np.abs(d2).max(), 1e-300)) if both.any() else 0.0
    r["lists_equal"] = bool(same)
    return r


def check_drift(rm, rf):
    """The mixed run's drift rm against the FP64 law's own drift rf from a
    one-ulp perturbation, quantity by quantity."""
    for k, floor in DRIFT_FLOORS.items():
        assert rm[k] <= max(SENSITIVITY_FACTOR * rf[k], floor), (k, rm[k], rf[k])
    assert rm["list_mismatch_frac"] <= max(SENSITIVITY_FACTOR * rf["list_mismatch_frac"], 1e-5)
    assert rm["lqdBrg_mismatch_frac"] <= max(SENSITIVITY_FACTOR * rf["lqdBrg_mismatch_frac"], 1e-5)


def nudged(st):
    """The same restart state with every coordinate moved up by one ulp."""
    p, v, w = st["p"]
    return dict(st, p=(np.nextafter(p, np.inf), v, w))


def export_state(e):
    """Everything a restart carries (DumpLoad): state, lists, history, bridge
    counts."""
    return dict(p=e.particles(), nb=e.nebor(), hp=e.hist_pp(), wl=e.walls(), hw=e.hist_pw(),
                lq=e.lqdcnt())


def restart(x, st):
    x.set_particles(*st["p"])
    x.set_contacts(*st["nb"], st["hp"], *st["wl"], st["hw"])
    x.set_lqdcnt(st["lq"])


def restarted_pair(inp, dia, st, precision=G.PREC_MIXED):
    O.threads(16)
    e = G.Engine(inp, dia, precision)
    o = O.Oracle(inp, dia)
    restart(e, st)
    restart(o, st)
    return e, o


# ---------------------------------------------------------------------------
# the 2,000,000-sphere bed, compacted for 3,000 steps (FP64) first, so that
# its contacts carry load and history
# ---------------------------------------------------------------------------
COMPACT_STEPS = 3000


@pytest.fixture(scope="module")
def compacted_2M():
    w = W.packed_bed(n=2_000_000)
    e = G.Engine(w["inp"])
    e.set_particles(w["pos"])
    e.step(COMPACT_STEPS)
    st = export_state(e)
    stats = e.stats()
    e.close()
    return dict(w=w, st=st, stats=stats)


def test_mixed_two_million_one_step_forces(compacted_2M):
    b = compacted_2M
    assert b["stats"]["n_contact"] > 6_000_000, b["stats"]
    e, o = restarted_pair(b["w"]["inp"], None, b["st"])
    e.store_forces(True)
    e.step(1)
    o.step(1)
    assert e.step_kernel_name() == "k_step_wave_mixed"
    r = one_step_errors(e, o, np.full(e.total, 0.5))
    r.update({"compact_" + k: int(v) for k, v in b["stats"].items()})
    record("mixed_2M_one_step", r)
    assert r["F"] <= FORCE_BAR and r["M"] <= FORCE_BAR
    e.close()


def test_mixed_two_million_trajectory(compacted_2M):
    b = compacted_2M
    e, o = restarted_pair(b["w"]["inp"], None, b["st"])
    f = G.Engine(b["w"]["inp"])
    restart(f, nudged(b["st"]))
    done = 0
    for n in TRAJ_STEPS_2M:
        for x in (e, o, f):
            x.step(n - done)
        done = n
        rm, rf = state_errors(e, o), state_errors(f, o)
        record(f"mixed_2M_{n}_steps", rm)
        record(f"fp64_nudged_2M_{n}_steps", rf)
    assert e.step_kernel_name() == "k_step_wave_mixed" and e.pair_records_used
    assert e.n_rebuild == o.n_rebuild
    check_drift(rm, rf)
    e.close()
    f.close()


# ---------------------------------------------------------------------------
# the wet lognormal bed
# ---------------------------------------------------------------------------
SETTLE_STEPS = 30_000


@pytest.fixture(scope="module")
def settled_wet_bed():
    w = W.wet_poly(n=200_000, E=1e5)
    e = G.Engine(w["inp"], w["dia"])
    e.set_particles(w["pos"])
    e.step(SETTLE_STEPS)
    st = export_state(e)
    stats = e.stats()
    e.close()
    S, rad, _, _ = G.derive(w["inp"], w["dia"])
    return dict(w=w, st=st, rad=rad[:len(st["p"][0])], stats=stats)


def test_fp64_restart_equals_oracle_restart(settled_wet_bed):
    # the restart path itself (DumpLoad's state: lists, history, bridge
    # counts) on both sides, bit for bit
    b = settled_wet_bed
    e, o = restarted_pair(b["w"]["inp"], b["w"]["dia"], b["st"], G.PREC_FP64)
    e.step(50)
    o.step(50)
    for u, v in zip(e.particles(), o.particles()):
        assert np.array_equal(u, v)
    assert np.array_equal(e.lqdcnt(), o.lqdcnt())
    he, ho = e.hist_pp(), o.hist_pp()
    for f in he:
        assert np.array_equal(he[f], ho[f]), f
    e.close()


def test_mixed_wet_lognormal_one_step_forces(settled_wet_bed):
    b = settled_wet_bed
    st = b["stats"]
    assert st["n_contact"] > 100_000 and st["n_bridge"] > 100_000, st
    e, o = restarted_pair(b["w"]["inp"], b["w"]["dia"], b["st"])
    e.store_forces(True)
    e.step(1)
    o.step(1)
    assert e.step_kernel_name() == "k_step_wave_wet_mixed"
    r = one_step_errors(e, o, b["rad"])
    h = o.hist_pp()
    r["bridged_slots"] = int((h["lqdBrg"] == 1).sum())
    r["yielded_slots"] = int((h["plstDfm"] > 0).sum())
    r.update({"settle_" + k: int(v) for k, v in st.items()})
    record("mixed_wet_lognormal_200k_one_step", r)
    assert r["F"] <= FORCE_BAR and r["M"] <= FORCE_BAR
    assert np.array_equal(e.lqdcnt(), o.lqdcnt())
    e.close()


def test_mixed_wet_lognormal_trajectory(settled_wet_bed):
    b = settled_wet_bed
    e, o = restarted_pair(b["w"]["inp"], b["w"]["dia"], b["st"])
    f = G.Engine(b["w"]["inp"], b["w"]["dia"])
    restart(f, nudged(b["st"]))
    done = 0
    for n in TRAJ_STEPS_WET:
        for x in (e, o, f):
            x.step(n - done)
        done = n
        rm, rf = state_errors(e, o), state_errors(f, o)
        lo = o.lqdcnt()
        rm["lqdCnt_mismatch_frac"] = float(np.mean(e.lqdcnt() != lo))
        rf["lqdCnt_mismatch_frac"] = float(np.mean(f.lqdcnt() != lo))
        (co, bo), (ce, be) = contact_counts(o, b["rad"]), contact_counts(e, b["rad"])
        rm["n_contact_rel"] = abs(ce - co) / co
        rm["n_bridge_rel"] = abs(be - bo) / bo
        record(f"mixed_wet_lognormal_200k_{n}_steps", rm)
        record(f"fp64_nudged_wet_lognormal_200k_{n}_steps", rf)
    assert e.n_rebuild == o.n_rebuild
    check_drift(rm, rf)
    assert rm["lqdCnt_mismatch_frac"] <= max(SENSITIVITY_FACTOR * rf["lqdCnt_mismatch_frac"], 1e-5)
    # the bulk state: contacts and bridges
    assert rm["n_contact_rel"] <= 1e-3 and rm["n_bridge_rel"] <= 1e-3
    e.close()
    f.close()
