"""The screened sweep: its packings, feed schedule, points-per-hour report,
and (GPU) that a batched sweep packing equals its plain run.

CPU: the sweep packings follow their inputs (particle count from the
radius ratio, size levels from sigma, skins that reach the capillary
rupture distance, a feed of ten equal batches spanning the feed fraction of
the run); the report's formulas on synthetic numbers.
GPU: packings run alone and the same packings in one batch (their decks as
they are) give identical phi and Z, with the feeds made mid-run; the tuned
generation runs and yields packings.
"""
import math

import numpy as np
import pytest

from gpucode_amd import sweep as S
from gpucode_amd import workloads as W


def test_sweep_packing_follows_its_inputs():
    rng = np.random.default_rng(4)
    for q in range(20):
        x = W.sweep_point(rng)
        lo, hi = S._ranges()
        assert np.all(x >= lo) and np.all(x <= hi)
        w = W.sweep_instance(x, 100 + q, steps=30000)
        w2 = W.sweep_instance(x, 100 + q, steps=30000)
        assert w["n"] == w2["n"] and w["schedule"] == w2["schedule"]
        if w["dia"] is not None:
            assert np.array_equal(w["dia"], w2["dia"])
        # N from the radius ratio: the bed, 8 largest diameters tall, at phi 0.6
        d = W.SWEEP_D if w["dia"] is None else w["dia"].max()
        vol = math.pi / 6 * (W.SWEEP_D ** 3 if w["dia"] is None else np.mean(w["dia"] ** 3))
        bed = 0.6 * math.pi * w["inp"].R ** 2 * 8.0 * d / vol
        assert 0.8 * bed < w["n"] < 1.25 * bed  # (the bed height uses a 4,000-sphere probe's d_max)
        assert abs(w["inp"].R - 0.5 * x[6] * W.SWEEP_D) < 1e-12
        # ten batches summing to N, the last one at feed_frac of the run
        counts = [c for _, c in w["schedule"]]
        assert len(counts) == W.SWEEP_BATCHES and sum(counts) == w["n"]
        assert max(counts) - min(counts) <= 1
        last = w["schedule"][-1][0]
        assert abs(last - x[7] * 30000 * (W.SWEEP_BATCHES - 1) / W.SWEEP_BATCHES) <= 10
        # wet packings: lists reach the rupture distance
        assert (w["inp"].forceCap == 1) == (x[2] > 1e-4)
        assert w["inp"].cutGap >= W.min_cut_gap(w["inp"], w["dia"]) - 1e-12
        assert abs(w["inp"].mat[0][1] - 10.0 ** x[5]) < 1e-6 * 10.0 ** x[5]
        assert abs(w["inp"].theta - math.radians(x[3])) < 1e-12


def test_report_formulas():
    n = 10
    sc = dict(kind=np.array([0, 0, 1, 1, 1, 1, 2, 2, 2, 2]), y_pick=np.zeros((n, 2)),
              n_screen=1000, n_train=64, t_train=2.0, t_screen=0.5, n_out_of_range=1,
              members=4, train_loss=0.01, spread_pick=np.linspace(0, 1, n))
    phi = np.full(n, 0.6)
    sc["y_pick"][:, 0] = phi * np.array([1.001, 1.02, 1.0, 1.005, 1.02, 0.999, 1, 1, 1, 1])
    sc["y_pick"][:, 1] = 5.0
    r = S.report(sc, t_gen=10.0, phi_p=phi, Z_p=np.full(n, 5.0), t_judge=7.5,
                 t_plain=np.array([2.0, 4.0]), plain_equal=True)
    assert r["plain_points_per_hr"] == pytest.approx(3600 / 3.0)
    assert r["sweep_points_per_hr"] == pytest.approx(1000 * 3600 / 20.0)
    assert r["screen_points_per_hr"] == pytest.approx(1000 * 3600 / 8.0)
    assert r["phi_err_random"]["n"] == 4
    assert r["phi_err_random"]["within_1pct"] == pytest.approx(0.75)
    assert r["phi_err_top"]["within_1pct"] == pytest.approx(0.5)
    # flagged points are simulated; the rest count at the random sample's rate
    n_ok = 4 + (1000 - 4) * 0.75
    assert r["points_within_1pct_per_hr"] == pytest.approx(n_ok * 3600 / 20.0)


@pytest.mark.gpu
def test_batched_sweep_packings_equal_plain_runs():
    cfg = S.SweepConfig(steps=8000, bed_d=3.0, chunk=100)
    X, ss = S.plain_sample_points(3, seed=5)
    plain = [S.run_plain(cfg, x, s) for x, s in zip(X, ss)]
    phi, Z, _, _ = S.run_batched(cfg, X, ss, tune=False)
    for q in range(3):
        print(q, plain[q][:2], phi[q], Z[q])
        assert phi[q] == plain[q][0] and Z[q] == plain[q][1]
        assert 0.0 <= phi[q] < 0.8
    # beds have formed (phi's core region, 2 d_max inside the bed, is still
    # empty in beds this shallow)
    assert (Z > 1.0).all()


@pytest.mark.gpu
def test_tuned_generation_yields_packings():
    cfg = S.SweepConfig(steps=6000, bed_d=3.0, chunk=100)
    X, ss = S.training_points(8, seed=9)
    phi, Z, t, ntuned = S.run_batched(cfg, X, ss)
    print(phi, Z, t, ntuned)
    assert ntuned >= 1
    assert np.all(np.isfinite(phi)) and np.all(phi >= 0.0) and np.all(phi < 0.8)
