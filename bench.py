"""Throughput of the DEM time step on one MI355X.

Default workload (the 2M-sphere throughput case): 2,000,000 monodisperse
dry spheres, d = 0.3 mm, in a cylinder of radius 30 mm (100 d) with a
floor, neighbour-list skin 0.1 d, mu = 0.5, rho 2500, E 70 GPa, nu 0.25,
e 0.7, yield at 1 % d; FP32 pair forces with FP64 positions, velocities
and contact history (--precision fp64: all FP64).  The seeded feed is a layer-by-layer
pocket deposition (about 54 d tall); --settle steps (default 40,000) settle
it, brought to rest every --quench steps (default 2,000; the quench keeps
the contact history), under the block-compaction step kernel (identical
results; the timed kernel's profile then holds only the timed launches);
the settled state goes through a restart dump (dem_dump_write_binary)
and a fresh context loads it (dem_dump_read), and from there W warmup steps
run and K steps are timed, rebuilds included.  A "step" is one pass of the
time step over all particles (forces, walls, integration; a neighbour
rebuild when due).  After the timed region every rebuild stage is timed on
the same state (dem_profile_rebuild).

--workload pour: the same 2,000,000 spheres and material as a loose cloud
(jittered lattice, spacing 1.5 d) falling at 2 m/s onto the floor of a
cylinder 75 mm tall: the unsettled variant, whose lists are rebuilt every
few hundred steps, timed with its rebuilds.

--workload wet_poly: 200,000 wet polydisperse spheres (lognormal d, median
0.5 mm, sigma 0.25, clipped 0.25-1.0 mm, three size levels; capillary
bridges with liquid 1 % of the mean particle volume, theta 20 deg) poured
from a seeded random feed into a cylinder of radius 25 mm; by default the
whole 50,000-step pour is timed, from the fed state, with no warmup.

--workload wet_bed: the same 200,000 lognormal wet spheres with soft
spheres (E = 0.1 MPa), poured and settled for --settle steps (default
30,000, in FP64) and restarted from their dump in the timed precision, so
contacts, plastic contacts and capillary bridges all act while it is timed.

Prints ONE JSON line on stdout.  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(st, n):
    """Compulsory HBM bytes per particle-step of the fused step kernel:
    read x,v,omega (72) + r (8), write x,v,omega (72), path length r/w (16),
    list counts (8), 4 B per neighbour index, and 64 B read + 64 B write of
    contact history (dispt 24, nDsp/nDsp_mx/plstRad/plstDfm 32, stage 4,
    lqdBrg 4) per contacting p-p or wall slot.  Neighbour state re-reads
    count as cache hits; forces are never stored."""
    n_nb = st["n_nebor"] / n
    n_c = st["n_contact"] / n
    n_w = st["n_wall_contact"] / n
    return 176.0 + 4.0 * n_nb + 128.0 * (n_c + n_w), n_nb, n_c, n_w


def load_traffic(kernel, tag):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(engine, inp, threads, budget_s=12.0, dia=None, start=None):
    """The FP64 C oracle timed on the host.  Without start: force + wall +
    integrate steps of the engine's current bed after one untimed rebuild.
    With start (positions, velocities, spins at the beginning of a pour):
    whole steps, rebuilds included, from that state and from the engine's
    current (end) state, half the budget each; the rate is the particle-steps
    of both samples over their summed time."""
    from oracle import pyoracle as O

    O.threads(threads)
    pos, vel, angv = engine.particles()
    n = len(pos)
    if start is None:
        o = O.Oracle(inp, dia)
        o.set_particles(pos, vel, angv)
        o.rebuild()
        t1 = o.time_force_steps(1)
        steps = int(max(1, min(20, budget_s / max(t1, 1e-6))))
        t = o.time_force_steps(steps)
        o.close()
        return dict(value=n * steps / t, unit="particle-steps/s", cores=threads, kind="port",
                    sample=f"FP64 C oracle (OpenMP, gather form): {steps} force+wall+integrate "
                           f"steps of the same {n}-sphere bed after one untimed rebuild")
    total_steps, total_t, parts = 0, 0.0, []
    for label, state in (("start", start), ("end", (pos, vel, angv))):
        o = O.Oracle(inp, dia)
        o.set_particles(*state)
        o.rebuild()
        t0 = time.perf_counter()
        o.step(10)
        t10 = time.perf_counter() - t0
        steps = int(max(10, min(5000, 0.5 * budget_s / max(t10 / 10, 1e-6))))
        t0 = time.perf_counter()
        o.step(steps)
        t = time.perf_counter() - t0
        o.close()
        total_steps += steps
        total_t += t
        parts.append(f"{steps} at the {label}")
    return dict(value=n * total_steps / total_t, unit="particle-steps/s", cores=threads,
                kind="port",
                sample=f"FP64 C oracle (OpenMP, gather form): whole steps, rebuilds included, of "
                       f"the same {n}-sphere pour: {' and '.join(parts)} of the timed run")


def sweep_main(args):
    """Points per hour of the batched + surrogate sweep against plain runs
    (one packing per run), on pour-and-settle packings fed in batches."""
    from gpucode_amd import sweep as SW

    cfg = SW.SweepConfig(steps=args.sweep_steps, bed_d=args.sweep_bed)
    t0 = time.time()
    r = SW.surrogate_sweep(cfg, n_train=args.sweep_train, n_screen=args.sweep_screen,
                           n_plain=args.sweep_plain, epochs=args.sweep_epochs, seed=0, log=log)
    log(f"sweep done in {time.time() - t0:.0f}s")
    out = {
        "metric": "sweep points/hr vs plain at <=1% phi error",
        "value": r["sweep_points_per_hr"],
        "unit": "points/hr",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f32-pair/f64-state; MLP bf16 x bf16 -> f32",
        "data": "synthetic",
        "config": {"workload": f"pour-and-settle packings fed in 10 batches, 8 inputs (mu, "
                               f"restitution, liquid, contact angle, lognormal sigma, E, R/r, "
                               f"feed rate), bed {cfg.bed_d} d, {cfg.steps} steps",
                   "n_train": r["n_train"], "n_screen": r["n_screen"], "n_judged": r["n_judged"],
                   "parallelism": "single-gpu"},
        **{k: v for k, v in r.items() if k not in ("n_train", "n_screen", "n_judged")},
    }
    print(json.dumps(out), flush=True)


def c4_cpu_baseline(instances, threads, budget_s=12.0):
    """The FP64 C oracle on a bounded sample of the batched set: whole steps
    (rebuilds included) of the first instances from their fed states, about
    budget_s of host time; particle-steps over time."""
    from oracle import pyoracle as O

    O.threads(threads)
    total_ps, total_t, parts = 0, 0.0, []
    for w in instances:
        if total_t >= budget_s:
            break
        o = O.Oracle(w["inp"], w.get("dia"))
        o.set_particles(w["pos"], w.get("vel"))
        t0 = time.perf_counter()
        o.step(10)
        t10 = time.perf_counter() - t0
        steps = int(max(10, min(3000, (budget_s / 3) / max(t10 / 10, 1e-6))))
        t0 = time.perf_counter()
        o.step(steps)
        t = time.perf_counter() - t0
        o.close()
        n = len(w["pos"])
        total_ps += n * steps
        total_t += t
        parts.append(f"{w['name']} ({n} spheres, {steps} steps)")
    return dict(value=total_ps / total_t, unit="particle-steps/s", cores=threads, kind="port",
                sample="FP64 C oracle (OpenMP, gather form): whole steps, rebuilds included, from "
                       "the fed state of " + ", ".join(parts))


def c4_main(args):
    """Batched uneven instances in one launch per step: instances from
    workloads.c4_instance, each autotuned (skin x block shape x storage
    order), then `steps` steps of all of them from the feed (pour and
    settle), timed whole.  Prints its own JSON line."""
    import numpy as np
    import gpucode_amd as G
    from gpucode_amd import workloads as W

    # default: FP32 pair forces with FP64 state
    prec = G.PREC_FP64 if args.precision == "fp64" else G.PREC_MIXED
    steps = args.steps if args.steps is not None else 30000
    t0 = time.time()
    ws = [W.c4_instance(k) for k in range(args.c4_instances)]
    t_gen = time.time() - t0
    n_tot = sum(len(w["pos"]) for w in ws)
    log(f"{len(ws)} instances, {n_tot} spheres, fed in {t_gen:.1f}s")
    t0 = time.time()
    choices = None
    if args.c4_tune_steps > 0:
        choices = G.tune_instances(ws, prec, steps=args.c4_tune_steps, log=log,
                                   kernels=args.c4_kernels.split(","),
                                   cut_gaps=tuple(float(g) for g in args.c4_skins.split(",")))
    t_tune = time.time() - t0
    b = G.Batch.from_instances(ws, choices, prec)
    b.set_schedule(device=True, slot_every=args.slot_every)
    log(f"tuned in {t_tune:.1f}s; batch up")
    n_k = np.array([len(w["pos"]) for w in ws], float)

    def bytes_per_ps():
        out = np.zeros(len(ws))
        for k in range(len(ws)):
            st = b[k].stats()
            B, _, _, _ = algorithmic_bytes(st, st["n_particles"])
            if ws[k]["inp"].forceCap == 1:
                # lqdCnt old/new; lqdBrg r/w of bridged slots that do not touch
                B += 8.0 + 8.0 * max(st["n_bridge"] - st["n_contact"], 0) / st["n_particles"]
            out[k] = B
        return out

    nseg = max(1, steps // args.c4_segment)
    seg = [steps // nseg + (1 if i < steps % nseg else 0) for i in range(nseg)]
    B_prev = bytes_per_ps()
    el = bytes_alg = 0.0
    phases = []
    tm0 = b.timing()
    for i, k in enumerate(seg):
        t0 = time.perf_counter()
        b.step(k)
        b.flush_rebuilds()  # inside the clock: the sampling below would do them
        dt = time.perf_counter() - t0
        el += dt
        B_now = bytes_per_ps()
        seg_bytes = float(np.sum(n_k * k * 0.5 * (B_prev + B_now)))
        bytes_alg += seg_bytes
        phases.append(dict(steps=k, seconds=dt, particle_steps_per_s=n_tot * k / dt,
                           gbs_algorithmic=seg_bytes / dt / 1e9,
                           bytes_per_particle_step=seg_bytes / (n_tot * k)))
        log(f"segment {i + 1}/{nseg}: {k} steps in {dt:.2f}s, {n_tot * k / dt:.3e} "
            f"particle-steps/s, {seg_bytes / (n_tot * k):.1f} B/particle-step")
        B_prev = B_now
    tm = b.timing()
    sch = b.schedule_stats()
    phi, Z, _ = b.analyze()
    value = n_tot * steps / el
    achieved = bytes_alg / el / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cpu = c4_cpu_baseline(ws, threads)
        cpu["nproc"] = os.cpu_count()
        log(f"cpu baseline {cpu['value']:.3e} particle-steps/s on {threads} threads")
    kinds = {}
    if choices:
        wave = sum(c["kernel"] == G.Engine.KERNEL_WAVE for c in choices)
        morton = sum(c["morton"] for c in choices)
        gaps = np.array([c["cutGap"] for c in choices])
        kinds = {"wave": wave, "loop": len(choices) - wave, "morton": morton,
                 "row_major": len(choices) - morton,
                 "cutGap_quartiles": [float(q) for q in np.percentile(gaps, [0, 25, 50, 75, 100])]}
    out = {
        "metric": "batched particle-steps/s (uneven instances, one launch per step)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": el / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == G.PREC_FP64 else "f32-pair/f64-state",
        "data": "synthetic",
        "config": {"workload": f"{len(ws)} cylinder packings, N~U[1000,20000] "
                               f"(total {n_tot}), mu U[0.1,0.8], liquid 0 (p=1/4) or "
                               f"U(0,3%], theta U[0,60] deg, lognormal sigma U[0,0.4]; "
                               f"{steps} steps from the feed, rebuilds included",
                   "n_instances": len(ws), "n_particles": n_tot,
                   "slot_every": args.slot_every, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "note": "algorithmic bytes of every instance's steps (per-instance "
                             "counters sampled between segments) over the whole timed run, "
                             "rebuild slots included"},
        "cpu_baseline": cpu,
        "tuning": {"seconds": t_tune, "steps_per_candidate": args.c4_tune_steps,
                   "block_shapes_allowed": args.c4_kernels, "choices": kinds,
                   "ms_per_step_default_sum": sum(c["ms_default"] for c in choices) if choices else None,
                   "ms_per_step_tuned_sum": sum(c["ms_per_step"] for c in choices) if choices else None},
        "schedule": {"global_launches": tm["launches"] - tm0["launches"], "slots": sch["slots"],
                     "host_syncs": sch["host_syncs"], "rebuilds": tm["n_rebuild"] - tm0["n_rebuild"]},
        "phases": phases,
        "packing": {"phi_mean": float(np.mean(phi)), "phi_min": float(np.min(phi)),
                    "phi_max": float(np.max(phi)), "Z_mean": float(np.mean(Z))},
    }
    b.close()
    print(json.dumps(out), flush=True)


REBUILD_BYTES = {
    # algorithmic bytes per particle of each rebuild stage (n_nb neighbours,
    # n_c contacts, c_per_p cells per particle, passes of the radix sort)
    "save": lambda q: 0.0,
    "hash": lambda q: 24.0 + 8.0,
    "sort": lambda q: 16.0 * q["passes"],
    "permute": lambda q: 2.0 * 108.0,
    "cell_range": lambda q: 8.0 + 8.0 * q["c_per_p"],
    "neighbor_find": lambda q: 32.0 + 4.0 * q["n_nb"],
    "active_slots": lambda q: 12.0 + 32.0 * q["n_nb"],
    "wall_list": lambda q: 32.0 + 16.0,
    "hist_remap": lambda q: 128.0 * q["n_c"] + 8.0 * q["n_nb"],
}


def prepare(args, G, W):
    """The workload's engine, ready for the timed region, and its labels."""
    t0 = time.time()
    prec = G.PREC_FP64 if args.precision == "fp64" else G.PREC_MIXED
    kern = {"auto": G.Engine.KERNEL_AUTO, "wave": G.Engine.KERNEL_WAVE,
            "compact": G.Engine.KERNEL_COMPACT, "loop": G.Engine.KERNEL_LOOP}[args.kernel]
    info = {}
    if args.workload in ("packed_bed", "pour"):
        n_req = args.n or 2_000_000
        nm = f"{n_req / 1e6:g}M"
        if args.workload == "packed_bed":
            w = W.packed_bed(n=n_req)
            settle = 40000 if args.settle is None else args.settle
            desc = (f"packed_bed_{nm} (dry, monodisperse d=0.3mm, R=30mm, skin 0.1d, mu 0.5; "
                    f"pocket-deposited feed settled {settle} steps, brought to rest every "
                    f"{args.quench}, restarted from its dump)")
            tag = f"packed_bed_{nm}"
        else:
            w = W.pour_2M(n=n_req)
            settle = 0
            desc = (f"pour_{nm} (dry, monodisperse d=0.3mm, R=30mm, skin 0.1d, mu 0.5; a loose "
                    f"lattice falling at 2 m/s onto the floor)")
            tag = f"pour_{nm}"
        metric = "particle-steps/s at 2M spheres"
        steps = 10000 if args.steps is None else args.steps
        warmup = 300 if args.warmup is None else args.warmup
        nseg = 1 if args.workload == "packed_bed" else 10
        seed = 0xC3
    elif args.workload == "wet_poly":
        n_req = args.n or 200_000
        w = W.wet_poly(n=n_req)
        settle = 0
        steps = w["steps"] if args.steps is None else args.steps
        warmup = 0 if args.warmup is None else args.warmup
        desc = (f"wet_poly_{n_req // 1000}k (lognormal d 0.25-1.0mm, 3 size levels, liquid 1%, "
                f"theta 20deg, R=25mm; the {steps}-step pour from the seeded feed)")
        metric = f"particle-steps/s at {n_req // 1000}k wet polydisperse spheres"
        tag = f"wet_poly_{n_req // 1000}k"
        nseg = 10
        seed = 0xC2
    else:
        n_req = args.n or 200_000
        w = W.wet_poly(n=n_req, E=1e5)
        settle = 30000 if args.settle is None else args.settle
        steps = 5000 if args.steps is None else args.steps
        warmup = 100 if args.warmup is None else args.warmup
        desc = (f"wet_bed_{n_req // 1000}k (lognormal d 0.25-1.0mm, 3 size levels, liquid 1%, "
                f"theta 20deg, R=25mm, soft spheres E=0.1MPa; the pour settled {settle} steps "
                f"in FP64 and restarted from its dump)")
        metric = f"particle-steps/s at {n_req // 1000}k wet polydisperse spheres"
        tag = f"wet_bed_{n_req // 1000}k"
        nseg = 5
        seed = 0xC2
    log(f"{args.workload} of {n_req} spheres built in {time.time() - t0:.1f}s")
    # the wet bed settles in FP64 (a different kernel from the timed one,
    # so the timed kernel's profile holds only the timed launches) and is
    # restarted from its dump in the timed precision
    settle_prec = G.PREC_FP64 if (settle and args.workload == "wet_bed") else prec
    e = G.Engine(w["inp"], w["dia"], settle_prec)
    if settle and args.workload == "packed_bed":
        # settle under another step kernel (results are identical), so the
        # timed kernel's profile covers only the warmup and timed launches
        e.set_kernel(G.Engine.KERNEL_LOOP if kern == G.Engine.KERNEL_COMPACT
                     else G.Engine.KERNEL_COMPACT)
    else:
        e.set_kernel(kern)
    e.set_particles(w["pos"], w.get("vel"))
    if settle:
        t0 = time.time()
        q = args.quench if args.workload == "packed_bed" else 0
        done = 0
        while done < settle:
            k = min(q, settle - done) if q else settle
            e.step(k)
            done += k
            if q:
                e.scale_velocities(0.0)  # a quench to rest keeps the contact history
        e.sync()
        st = e.stats()
        info["settle_steps"] = settle
        info["settle_s"] = round(time.time() - t0, 2)
        log(f"settled {settle} steps in {info['settle_s']:.1f}s, contacts/particle "
            f"{st['n_contact'] / st['n_particles']:.2f}")
        if args.workload == "packed_bed" or settle_prec != prec:
            # restart through the dump, as a run resumed from its DumpLoad file
            import tempfile
            d = tempfile.mkdtemp(prefix="dem_bench_")
            path = os.path.join(d, "heyi_dump.bin")
            t0 = time.time()
            e.dump(path, binary=True)
            size = os.path.getsize(path)
            t1 = time.time()
            e.close()
            e = G.Engine(w["inp"], w["dia"], prec)
            e.set_kernel(kern)
            e.load(path)
            t2 = time.time()
            os.remove(path)
            os.rmdir(d)
            if args.workload == "packed_bed":
                info["quench_every"] = args.quench
            else:
                info["settle_precision"] = "fp64"
            info["dump_bytes"] = size
            info["dump_write_s"] = round(t1 - t0, 2)
            info["dump_read_s"] = round(t2 - t1, 2)
            log(f"dump {size / 2**20:.0f} MiB written in {t1 - t0:.1f}s, read into a fresh "
                f"context in {t2 - t1:.1f}s")
    args.steps, args.warmup = steps, warmup
    return dict(e=e, w=w, desc=desc, metric=metric, tag=tag, nseg=nseg, seed=seed, info=info,
                prec=prec, settle=settle)


def rebuild_profile(e, n, reps, interval):
    """Rebuild stage times on the current state, their algorithmic bytes and
    the rebuild cost per step at the measured interval."""
    st = e.stats()
    S = e.setup
    q = {"n_nb": st["n_nebor"] / n, "n_c": st["n_contact"] / n,
         "c_per_p": S.sysNCells / n, "passes": max(1, (e.key_bits() + 7) // 8)}
    ms = e.profile_rebuild(reps)
    stages = {}
    for k, t in ms.items():
        b = REBUILD_BYTES[k](q) * n
        stages[k] = {"ms": round(t, 4), "bytes": round(b),
                     "gbs": round(b / (t * 1e-3) / 1e9, 1) if t > 0 else None}
    total = sum(ms.values())
    out = {"ms": round(total, 4), "stages": stages, "reps": reps}
    if interval:
        out["interval_steps"] = round(interval, 1)
        out["amortized_ms_per_step"] = round(total / interval, 5)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 10,000 for packed_bed; 50,000, the whole pour, for "
                         "wet_poly)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default 300 for packed_bed, 0 for wet_poly)")
    ap.add_argument("--n", type=int, default=None,
                    help="particles (default 2,000,000 for packed_bed, 200,000 for wet_poly)")
    ap.add_argument("--workload", choices=["packed_bed", "pour", "wet_poly", "wet_bed"],
                    default="packed_bed",
                    help="packed_bed: the settled 2M-sphere dry bed (the headline); pour: the "
                         "same spheres falling (rebuilds); wet_poly: the 200k polydisperse wet "
                         "pour; wet_bed: that pour settled (contacts and bridges)")
    ap.add_argument("--settle", type=int, default=None,
                    help="untimed settling steps before the restart dump (packed_bed, default "
                         "20,000) or the timed region (wet_bed, default 30,000)")
    ap.add_argument("--quench", type=int, default=2000,
                    help="packed_bed settling: velocities and spins are set to zero every this "
                         "many steps (0: never)")
    ap.add_argument("--rebuild-reps", type=int, default=3,
                    help="rebuilds of the final state timed stage by stage")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="mixed (the default: FP32 pair forces, FP64 positions, velocities "
                         "and contact history) or fp64")
    ap.add_argument("--kernel", choices=["auto", "wave", "compact", "loop"], default="auto",
                    help="fused step kernel: per-wave contact compaction (auto, wave), "
                         "block-wide compaction (compact) or the one-thread loop (loop)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep", action="store_true",
                    help="instead of the 2M-sphere step: the batched + surrogate parameter sweep "
                         "against plain runs (prints its own JSON line)")
    ap.add_argument("--sweep-train", type=int, default=4096)
    ap.add_argument("--sweep-screen", type=int, default=100_000)
    ap.add_argument("--sweep-steps", type=int, default=30000)
    ap.add_argument("--sweep-bed", type=float, default=8.0,
                    help="settled bed height of the sweep packings, in median diameters")
    ap.add_argument("--sweep-plain", type=int, default=24,
                    help="independent plain runs timed for the plain rate")
    ap.add_argument("--sweep-epochs", type=int, default=400)
    ap.add_argument("--c4", action="store_true",
                    help="instead of the 2M-sphere step: the batched uneven instances, "
                         "autotuned, stepped together (prints its own JSON line)")
    ap.add_argument("--c4-instances", type=int, default=256)
    ap.add_argument("--c4-skins", default="0.1,0.2,0.3,0.5",
                    help="candidate skins (cutGap) of the C4 autotuner, comma-separated")
    ap.add_argument("--c4-kernels", default="wave",
                    help="block shapes the C4 autotuner may choose (comma-separated: loop, "
                         "wave); one shape steps the batch in one launch per step (measured "
                         "3 %% faster over the run than letting instances mix the two)")
    ap.add_argument("--c4-tune-steps", type=int, default=200,
                    help="steps per autotuner candidate (0: no tuning)")
    ap.add_argument("--c4-segment", type=int, default=5000,
                    help="steps per timed segment (counters are sampled between segments)")
    ap.add_argument("--slot-every", type=int, default=4,
                    help="global steps between rebuild slots of the batched schedule")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}); the engine steps a packing on one GPU and does not split it "
              f"across GPUs.", file=sys.stderr)
        sys.exit(2)

    import gpucode_amd as G
    from gpucode_amd import workloads as W

    if args.sweep:
        return sweep_main(args)
    if args.c4:
        return c4_main(args)

    P = prepare(args, G, W)
    e, w, desc, metric, tag, nseg = P["e"], P["w"], P["desc"], P["metric"], P["tag"], P["nseg"]
    prec = P["prec"]
    steps, warmup = args.steps, args.warmup
    n = e.total
    log(f"engine up: {e.device_bytes() / 2**30:.2f} GiB on device")
    if P["settle"] and args.workload == "packed_bed" and e.iter == 0:
        raise RuntimeError("the restart did not carry the clock")
    # The attainable copy rate is measured here, right before the warmup:
    # its half second of HBM traffic also brings the GPU out of the idle
    # state the host-side restart left it in, so that the first timed
    # launches run at the clocks of the later ones.  It touches no state of
    # the engine.
    copy_gbs = G.engine.stream_copy(1 << 30, 20)
    t0 = time.time()
    e.step(warmup)
    e.sync()
    log(f"warmup {warmup} steps in {time.time() - t0:.1f}s, rebuilds {e.n_rebuild}")
    start = e.particles() if nseg > 1 else None
    rb0 = e.n_rebuild
    # The timed steps run in nseg segments; between segments (outside the
    # timed region) the bed's neighbour and contact counts are sampled, so
    # the algorithmic bytes follow a bed whose contacts change as it pours.
    seg = [steps // nseg + (1 if i < steps % nseg else 0) for i in range(nseg)]
    st_prev, tm_prev = e.stats(), e.timing()
    st_first = st_prev
    el = 0.0
    bytes_alg = fused_ms = 0.0
    launches = executed = 0
    for i, k in enumerate(seg):
        t0 = time.perf_counter()
        e.step(k)
        # a rebuild the last step asked for belongs to this run (the counter
        # sampling below would otherwise do it outside the clock)
        e.flush_rebuild()
        e.sync()
        el += time.perf_counter() - t0
        st, tm = e.stats(), e.timing()
        stm = {key: 0.5 * (st_prev[key] + st[key]) for key in st}
        B_seg, _, _, _ = algorithmic_bytes(stm, n)
        if w["inp"].forceCap == 1:
            # lqdCnt old/new, and lqdBrg r/w of the bridged slots
            B_seg += 8.0 + 8.0 * stm["n_bridge"] / n
        ex = tm["fused_steps"] - tm_prev["fused_steps"]
        bytes_alg += B_seg * n * ex
        executed += ex
        launches += tm["fused_launches"] - tm_prev["fused_launches"]
        fused_ms += tm["fused_ms"] - tm_prev["fused_ms"]
        if nseg > 1:
            log(f"segment {i + 1}/{nseg}: {k} steps, {el:.2f}s so far, rebuilds "
                f"{e.n_rebuild - rb0}, contacts/particle {st['n_contact'] / n:.2f}")
        st_prev, tm_prev = st, tm
    rb_timed = e.n_rebuild - rb0
    log(f"timed {steps} steps in {el:.3f}s, rebuilds {rb_timed}, contacts/particle "
        f"{st_first['n_contact'] / n:.2f} -> {st_prev['n_contact'] / n:.2f}")

    st = {k: 0.5 * (st_first[k] + st_prev[k]) for k in st_first}
    _, n_nb, n_c, n_w = algorithmic_bytes(st, n)
    B = bytes_alg / max(executed, 1) / n
    kern_ms = fused_ms / max(launches, 1)
    # algorithmic bytes of the steps that ran over the time of all launches
    # (a launch after a rebuild became due returns at once and moves nothing)
    achieved = bytes_alg / (fused_ms * 1e-3) / 1e9
    kname = e.step_kernel_name()
    traffic = traffic_nc = None
    tr = load_traffic(kname, tag)
    if tr and tr.get("settle_steps", 0) == P["settle"] and tr.get("n_particles") == n:
        traffic = tr.get("bytes_per_launch")
        traffic_nc = tr.get("n_contact")
    interval = executed / rb_timed if rb_timed > 0 else None
    reb = rebuild_profile(e, n, args.rebuild_reps, interval) if args.rebuild_reps > 0 else None
    if reb:
        log(f"rebuild {reb['ms']:.3f} ms: " + ", ".join(
            f"{k} {v['ms']:.3f}" for k, v in reb["stages"].items()))

    log(f"plain copy {copy_gbs:.0f} GB/s; {kname} {kern_ms:.4f} ms, {achieved:.0f} GB/s "
        f"algorithmic ({B:.1f} B/particle-step)")

    base = None
    if not args.no_cpu_baseline:
        # every core the process may use, at most the CPU share of one GPU
        # (16 host cores per GPU on these machines)
        aff = len(os.sched_getaffinity(0))
        thr = min(16, aff)
        t0 = time.time()
        base = cpu_baseline(e, w["inp"], thr, dia=w["dia"], start=start)
        base["nproc"] = os.cpu_count()
        base["affinity_cores"] = aff
        log(f"cpu baseline {base['value']:.3e} particle-steps/s on {thr} threads "
            f"(nproc {os.cpu_count()}, {aff} usable; {time.time() - t0:.1f}s)")

    value = n * args.steps / el
    out = {
        "metric": metric,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == G.PREC_FP64 else "f32-pair/f64-state",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "n_particles": n, "seed": P["seed"], **P["info"],
            "parallelism": "single-gpu",
            "rebuilds_timed": rb_timed,
            "launches_timed": launches, "steps_executed": executed,
            "n_nb": round(n_nb, 3), "n_contact": round(n_c, 3), "n_wall_contact": round(n_w, 4),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_n_contact": traffic_nc,
            "kernel": kname, "kernel_ms": kern_ms, "bytes_per_particle": B,
            "attainable_copy_gbs": copy_gbs,
        },
        "cpu_baseline": base,
        "rebuild": reb,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
