"""Parameter sweeps of pour-and-settle packings: batched generation, the MLP
surrogate ensemble, screening, and the plain simulation as fallback and
judge.

A sweep point is the 8 inputs of workloads.SWEEP_PARAMS (sliding friction,
restitution, liquid fraction, contact angle, lognormal sigma, Young's
modulus, cylinder/particle radius ratio, feed rate); its outcome is (phi, Z)
of the packing after cfg.steps steps, from the engine's analysis kernel
(dem_analyze: phi in the core cylinder, Z the mean contact number).  Each
packing is fed in batches during the run (workloads.sweep_instance); the
runs advance in chunks of cfg.chunk steps and feed a batch at the first
chunk boundary at or after its step, so a packing run alone (plain) and the
same packing in a batch take exactly the same steps and feeds.

  plain      one packing per run, the deck as it is (the reference's way):
             Engine + feeds + analyze
  batched    many packings per launch (Batch), the same feeds
  generation the training packings, batched, each packing with the skin,
             step kernel and storage order the autotuner picked for its
             class (wet or dry, size levels, size)
  surrogate  an ensemble of MLPs (different seeds) trained on the generated
             outcomes screens many points; points where the members disagree
             most (or whose prediction leaves the training range) are
             simulated instead
  judge      the best predicted points (top_frac), a random sample
             (rand_frac) and the disagreeing points are simulated as plain
             runs (computed by the batched engine, which reproduces plain
             runs bit for bit: checked on the independent plain sample)

Points per hour:
  plain                 3600 / (mean time of an independent sample of plain
                        runs, none of them a screened or training point)
  sweep                 n_screen / (t_gen + t_train + t_screen + t_judge)
  sweep, screen only    n_screen / (t_screen + t_judge)
"""
import time

import numpy as np

from . import workloads as W
from ._capi import PREC_MIXED
from .engine import Batch, Engine, tune_instances
from .surrogate import MLP, Normalizer


class SweepConfig:
    def __init__(self, steps=30000, bed_d=8.0, precision=PREC_MIXED, batch_size=256, chunk=100,
                 tune=True):
        self.steps = steps
        self.bed_d = bed_d
        self.precision = precision
        self.batch_size = batch_size
        self.chunk = chunk
        self.tune = tune

    def instance(self, x, seed):
        return W.sweep_instance(x, int(seed), steps=self.steps, bed_d=self.bed_d)


def points(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([W.sweep_point(rng) for _ in range(n)])


def _ranges():
    lo = np.array([W.SWEEP_RANGES[k][0] for k in W.SWEEP_PARAMS])
    hi = np.array([W.SWEEP_RANGES[k][1] for k in W.SWEEP_PARAMS])
    return lo, hi


def _run_chunks(cfg, ws, engines, step):
    """Advance packings ws (engine k = engines[k]) by cfg.steps steps in
    chunks; at every chunk boundary each packing is fed up to the spheres
    its schedule has made due (a batch that did not fit is completed
    later); step(k) advances every engine by k steps."""
    state = [dict() for _ in ws]
    t = 0
    while t < cfg.steps:
        for k, w in enumerate(ws):
            due = W.feed_due(w, t)
            if engines[k].total < due:
                W.feed_batch(engines[k], w, due, state[k])
        c = min(cfg.chunk, cfg.steps - t)
        step(c)
        t += c


def run_plain(cfg, x, seed):
    """One packing alone, its deck as it is; returns (phi, Z, seconds)."""
    t0 = time.perf_counter()
    w = cfg.instance(x, seed)
    e = Engine(w["inp"], w["dia"], cfg.precision)
    e.set_particles(w["pos"])
    _run_chunks(cfg, [w], [e], e.step)
    phi, Z, _ = e.analyze()
    el = time.perf_counter() - t0
    e.close()
    return phi, Z, el


def _class_of(w):
    n = w["n"]
    return (int(w["inp"].forceCap), int(w["inp"].nLevels), 0 if n < 2500 else (1 if n < 5000 else 2))


def tune_classes(cfg, ws, log=None):
    """The autotuner on one packing per class (wet or dry, size levels,
    size), started from its first feed batch at rest; every packing of the
    class takes that choice (its skin no smaller than its own capillary
    bridges need).  Returns (choices, classes tuned)."""
    reps = {}
    for k, w in enumerate(ws):
        reps.setdefault(_class_of(w), []).append(k)
    choice = {}
    for c, ks in reps.items():
        ks = sorted(ks, key=lambda k: ws[k]["n"])
        w = ws[ks[len(ks) // 2]]
        e = Engine(w["inp"], w["dia"], cfg.precision)
        e.set_particles(w["pos"])
        W.feed_batch(e, w, W.feed_due(w, 0), {})
        pos, vel, _ = e.particles()
        e.close()
        probe = dict(inp=w["inp"], dia=w["dia"], pos=pos, vel=vel)
        choice[c] = tune_instances([probe], cfg.precision, steps=200)[0]
    out = []
    for w in ws:
        ch = dict(choice[_class_of(w)])
        ch["cutGap"] = max(ch["cutGap"], W.min_cut_gap(w["inp"], w["dia"]))
        out.append(ch)
    if log:
        log(f"tuned {len(choice)} classes")
    return out, len(choice)


def run_batched(cfg, X, seeds, log=None, tune=None):
    """All points in batches of cfg.batch_size; tune (default cfg.tune):
    each batch's packings take the autotuner's choice for their class,
    otherwise their decks as they are (then every packing equals its plain
    run bit for bit).  Returns (phi, Z, seconds, classes tuned)."""
    tune = cfg.tune if tune is None else tune
    phi = np.zeros(len(X))
    Z = np.zeros(len(X))
    ntuned = 0
    t0 = time.perf_counter()
    for b0 in range(0, len(X), cfg.batch_size):
        sl = slice(b0, min(len(X), b0 + cfg.batch_size))
        ws = [cfg.instance(x, s) for x, s in zip(X[sl], seeds[sl])]
        choices = None
        if tune:
            choices, nc = tune_classes(cfg, ws)
            ntuned += nc
        b = Batch.from_instances(ws, choices, cfg.precision)
        _run_chunks(cfg, ws, b.inst, b.step)
        p, z, _ = b.analyze()
        phi[sl], Z[sl] = p, z
        b.close()
        if log:
            log(f"batched {sl.stop}/{len(X)} points, {time.perf_counter() - t0:.1f}s")
    return phi, Z, time.perf_counter() - t0, ntuned


def training_points(n_train, seed=0):
    """The training points and their seeds."""
    return points(n_train, seed), np.arange(n_train) + 1_000_000 * (seed + 1)


def screen_points(n_screen, seed=0):
    return points(n_screen, seed + 7), np.arange(n_screen) + 2_000_000 * (seed + 1)


def plain_sample_points(n, seed=0):
    """An independent sample: no training or screened point."""
    return points(n, seed + 13), np.arange(n) + 3_000_000 * (seed + 1)


def train_ensemble(Xt, Yt, members=4, epochs=400, seed=0):
    """members MLPs trained from different seeds; returns (models, norm)."""
    lo, hi = _ranges()
    norm = Normalizer(lo, hi, Yt)
    ms, losses = [], []
    for q in range(members):
        m = MLP(seed=seed * 100 + q)
        losses.append(m.train(norm.x(Xt), norm.y(Yt), epochs=epochs, batch=256, lr=1e-3,
                              seed=seed * 100 + q))
        ms.append(m)
    return ms, norm, float(np.mean(losses))


def predict(ms, norm, X):
    """Ensemble mean (phi, Z) and the members' relative spread of phi."""
    ys = np.stack([norm.y_inv(m.forward(norm.x(X))) for m in ms])
    mean = ys.mean(0)
    spread = ys[:, :, 0].std(0) / np.maximum(np.abs(mean[:, 0]), 1e-12)
    return mean, spread


def phase_screen(cfg, gen, n_screen=100_000, top_frac=0.01, rand_frac=0.01, flag_frac=0.01,
                 members=4, epochs=400, seed=0, log=None):
    """Train the ensemble on the generated outcomes gen = (X, phi, Z), screen
    n_screen points, and pick the points to simulate: the top_frac best by
    predicted phi, a rand_frac random sample, and the flag_frac whose
    members disagree most or whose prediction leaves the training range of
    phi (those are answered by simulation, not by the surrogate)."""
    Xt, phi_t, Z_t = gen
    Yt = np.stack([phi_t, Z_t], 1)
    t0 = time.perf_counter()
    ms, norm, loss = train_ensemble(Xt, Yt, members, epochs, seed)
    t_train = time.perf_counter() - t0
    Xs, ss = screen_points(n_screen, seed)
    t0 = time.perf_counter()
    Ys, spread = predict(ms, norm, Xs)
    t_screen = time.perf_counter() - t0
    for m in ms:
        m.close()
    nflag = int(round(flag_frac * n_screen))
    out_of_range = (Ys[:, 0] < phi_t.min()) | (Ys[:, 0] > phi_t.max())
    order_spread = np.argsort(-spread, kind="stable")
    flagged = np.zeros(n_screen, bool)
    flagged[order_spread[:nflag]] = True
    flagged |= out_of_range
    conf = np.flatnonzero(~flagged)
    ntop, nrand = int(round(top_frac * n_screen)), int(round(rand_frac * n_screen))
    top = conf[np.argsort(-Ys[conf, 0], kind="stable")[:ntop]]
    rng = np.random.default_rng(seed + 11)
    rest = np.setdiff1d(conf, top)
    rnd = rng.choice(rest, min(nrand, len(rest)), replace=False) if nrand else np.array([], int)
    flg = np.flatnonzero(flagged)
    pick = np.concatenate([top, rnd, flg]).astype(int)
    kind = np.concatenate([np.zeros(len(top), int), np.ones(len(rnd), int),
                           np.full(len(flg), 2, int)])
    if log:
        log(f"screen: loss {loss:.4f}, {len(top)} top + {len(rnd)} random + {len(flg)} flagged "
            f"({int(out_of_range.sum())} out of the training range)")
    return dict(n_train=len(Xt), train_loss=loss, t_train=t_train, t_screen=t_screen,
                n_screen=n_screen, pick=pick, kind=kind, x_pick=Xs[pick], seed_pick=ss[pick],
                y_pick=Ys[pick], spread_pick=spread[pick], n_flagged=len(flg),
                n_out_of_range=int(out_of_range.sum()), members=members)


def report(sc, t_gen, phi_p, Z_p, t_judge, t_plain, plain_equal=None):
    """Errors of the surrogate on the judged points and points per hour."""
    kind = np.asarray(sc["kind"])
    y = np.asarray(sc["y_pick"])
    err = np.abs(y[:, 0] - phi_p) / np.maximum(phi_p, 1e-12)
    zerr = np.abs(y[:, 1] - Z_p) / np.maximum(Z_p, 1e-12)
    top, rnd, flg = kind == 0, kind == 1, kind == 2

    def stats(m):
        e = err[m]
        if not len(e):
            return None
        return dict(n=int(len(e)), mean=float(e.mean()), median=float(np.median(e)),
                    p95=float(np.percentile(e, 95)), max=float(e.max()),
                    within_1pct=float(np.mean(e <= 0.01)))

    n_screen = int(sc["n_screen"])
    t_all = float(t_gen) + float(sc["t_train"]) + float(sc["t_screen"]) + float(t_judge)
    t_plain_mean = float(np.mean(t_plain))
    plain_ph = 3600.0 / t_plain_mean
    sweep_ph = n_screen * 3600.0 / t_all
    screen_ph = n_screen * 3600.0 / (float(sc["t_screen"]) + float(t_judge))
    r_ok = stats(rnd)
    # points whose answer is within 1 %: the flagged ones are simulated, the
    # rest are estimated from the random sample's rate
    n_ok = int(flg.sum()) + (n_screen - int(flg.sum())) * (r_ok["within_1pct"] if r_ok else 0.0)
    return dict(
        n_train=int(sc["n_train"]), n_screen=n_screen, n_judged=int(len(phi_p)),
        n_flagged=int(flg.sum()), n_out_of_range=int(sc["n_out_of_range"]),
        members=int(sc["members"]), train_loss=float(sc["train_loss"]),
        t_gen=float(t_gen), t_train=float(sc["t_train"]), t_screen=float(sc["t_screen"]),
        t_judge=float(t_judge), t_plain_mean=t_plain_mean, n_plain_sample=int(len(t_plain)),
        plain_equal_batched=plain_equal,
        plain_points_per_hr=plain_ph, sweep_points_per_hr=sweep_ph,
        screen_points_per_hr=screen_ph, speedup=sweep_ph / plain_ph,
        speedup_screen_only=screen_ph / plain_ph,
        points_within_1pct_per_hr=n_ok * 3600.0 / t_all,
        speedup_within_1pct=n_ok * 3600.0 / t_all / plain_ph,
        phi_err_top=stats(top), phi_err_random=stats(rnd), phi_err_flagged=stats(flg),
        phi_err_mean=float(err[~flg].mean()) if (~flg).any() else None,
        z_err_mean=float(zerr[~flg].mean()) if (~flg).any() else None,
        spread_flagged_min=float(np.min(np.asarray(sc["spread_pick"])[flg])) if flg.any() else None,
    )


def phase_plain(cfg, n=24, seed=0, log=None):
    """The independent plain sample: each point run alone (timed) and again
    in one untuned batch; returns (phi, Z, seconds, batch equals plain)."""
    X, ss = plain_sample_points(n, seed)
    phi, Z, t = np.zeros(n), np.zeros(n), np.zeros(n)
    for q in range(n):
        phi[q], Z[q], t[q] = run_plain(cfg, X[q], ss[q])
        if log:
            log(f"plain run {q + 1}/{n}: {t[q]:.2f}s")
    pb, zb, _, _ = run_batched(cfg, X, ss, tune=False)
    equal = bool(np.array_equal(pb, phi) and np.array_equal(zb, Z))
    return phi, Z, t, equal


def surrogate_sweep(cfg, n_train=4096, n_screen=100_000, n_plain=24, top_frac=0.01,
                    rand_frac=0.01, flag_frac=0.01, members=4, epochs=400, seed=0, log=None):
    """The whole sweep in one process."""
    Xt, st = training_points(n_train, seed)
    phi_t, Z_t, t_gen, _ = run_batched(cfg, Xt, st, log)
    sc = phase_screen(cfg, (Xt, phi_t, Z_t), n_screen, top_frac, rand_frac, flag_frac, members,
                      epochs, seed, log)
    phi_p, Z_p, t_judge, _ = run_batched(cfg, sc["x_pick"], sc["seed_pick"], log, tune=False)
    _, _, t_plain, equal = phase_plain(cfg, n_plain, seed, log)
    return report(sc, t_gen, phi_p, Z_p, t_judge, t_plain, equal)


def _main(argv=None):
    """python -m gpucode_amd.sweep gen|screen|judge|plain|report ...: the sweep
    in phases that each fit a bounded run, the state passed through .npz files:
      gen --part k/m --out G_k       batched, tuned outcomes of a slice of the training points
      screen --gen G_0 .. --out S    train the ensemble, screen, pick the points to simulate
      judge --state S --part k/m --out J_k   the picked points as plain runs (batched engine)
      plain --out P                  the independent plain sample, timed (and checked batched)
      report --state S --judge J_0 .. --plain P   one JSON line
    """
    import argparse
    import json
    import sys

    def log(m):
        print(f"[sweep] {m}", file=sys.stderr, flush=True)

    ap = argparse.ArgumentParser()
    ap.add_argument("phase", choices=["gen", "screen", "judge", "plain", "report"])
    ap.add_argument("--out")
    ap.add_argument("--state")
    ap.add_argument("--gen", nargs="*", default=[])
    ap.add_argument("--judge", nargs="*", default=[])
    ap.add_argument("--plain")
    ap.add_argument("--part", default="0/1")
    ap.add_argument("--n-train", type=int, default=4096)
    ap.add_argument("--n-screen", type=int, default=100_000)
    ap.add_argument("--n-plain", type=int, default=24)
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--bed-d", type=float, default=8.0)
    ap.add_argument("--epochs", type=int, default=400)
    a = ap.parse_args(argv)
    k, m = (int(v) for v in a.part.split("/"))
    cfg = SweepConfig(steps=a.steps, bed_d=a.bed_d)
    if a.phase == "gen":
        Xt, st = training_points(a.n_train)
        sl = slice(k * a.n_train // m, (k + 1) * a.n_train // m)
        phi, Z, t, nc = run_batched(cfg, Xt[sl], st[sl], log)
        np.savez(a.out, start=sl.start, stop=sl.stop, phi=phi, Z=Z, t=t, classes=nc,
                 steps=a.steps, bed_d=a.bed_d)
        log(f"gen {sl.start}:{sl.stop} done in {t:.1f}s")
    elif a.phase == "screen":
        parts = [np.load(f) for f in a.gen]
        n_train = max(int(p["stop"]) for p in parts)
        assert sum(int(p["stop"]) - int(p["start"]) for p in parts) == n_train
        phi, Z, t = np.zeros(n_train), np.zeros(n_train), 0.0
        for p in parts:
            phi[int(p["start"]):int(p["stop"])] = p["phi"]
            Z[int(p["start"]):int(p["stop"])] = p["Z"]
            t += float(p["t"])
        Xt, _ = training_points(n_train)
        sc = phase_screen(cfg, (Xt, phi, Z), a.n_screen, epochs=a.epochs, log=log)
        np.savez(a.out, t_gen=t, steps=a.steps, bed_d=a.bed_d, **sc)
    elif a.phase == "judge":
        sc = np.load(a.state)
        n = len(sc["pick"])
        sl = slice(k * n // m, (k + 1) * n // m)
        phi, Z, t, _ = run_batched(cfg, sc["x_pick"][sl], sc["seed_pick"][sl], log, tune=False)
        np.savez(a.out, start=sl.start, stop=sl.stop, phi=phi, Z=Z, t=t)
        log(f"judge {sl.start}:{sl.stop} done in {t:.1f}s")
    elif a.phase == "plain":
        phi, Z, t, equal = phase_plain(cfg, a.n_plain, log=log)
        np.savez(a.out, phi=phi, Z=Z, t=t, equal=equal)
        log(f"plain sample: {t.mean():.2f}s per run, batched equals plain: {equal}")
    else:
        sc = dict(np.load(a.state))
        n = len(sc["pick"])
        phi, Z = np.full(n, np.nan), np.full(n, np.nan)
        t_judge = 0.0
        for f in a.judge:
            p = np.load(f)
            sl = slice(int(p["start"]), int(p["stop"]))
            phi[sl], Z[sl] = p["phi"], p["Z"]
            t_judge += float(p["t"])
        if np.isnan(phi).any():
            raise SystemExit(f"{int(np.isnan(phi).sum())} picked points have no judged run")
        pl = np.load(a.plain)
        print(json.dumps(report(sc, float(sc["t_gen"]), phi, Z, t_judge, pl["t"],
                                bool(pl["equal"]))))


if __name__ == "__main__":
    _main()
