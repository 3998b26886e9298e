"""Throughput of the DEM time step on one MI355X.

Workload (the 2M-sphere throughput case): 2,000,000 monodisperse dry
spheres, d = 0.3 mm, in a cylinder of radius 30 mm (100 d) with a floor,
neighbour-list skin 0.1 d, mu = 0.5, rho 2500, E 70 GPa, nu 0.25, e 0.7,
yield at 1 % d, FP64 state.  The bed is a seeded layer-by-layer pocket
deposition (about 54 d tall); W warmup steps settle it, then K steps are
timed, rebuilds included.  A "step" is one pass of the time step over all
particles (forces, walls, integration; a neighbour rebuild when due).

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


def load_traffic(kernel):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_packed_bed_2M.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(engine, inp, threads, budget_s=12.0):
    from oracle import pyoracle as O

    O.threads(threads)
    pos, vel, angv = engine.particles()
    o = O.Oracle(inp)
    o.set_particles(pos, vel, angv)
    o.rebuild()
    t1 = o.time_force_steps(1)
    steps = int(max(1, min(20, budget_s / max(t1, 1e-6))))
    t = o.time_force_steps(steps)
    n = len(pos)
    o.close()
    return dict(value=n * steps / t, unit="particle-steps/s", cores=threads, kind="port",
                sample=f"FP64 C oracle (OpenMP, gather form): {steps} force+wall+integrate steps "
                       f"of the same {n}-sphere bed after one untimed rebuild")


def sweep_main(args):
    """Points per hour of the batched + surrogate sweep against plain runs
    (one instance per run), on pour-and-settle packings."""
    from gpucode_amd import sweep as SW

    cfg = SW.SweepConfig(n_min=1000, n_max=args.sweep_nmax, steps=args.sweep_steps, fill=2.5)
    t0 = time.time()
    r = SW.surrogate_sweep(cfg, n_train=args.sweep_train, n_screen=args.sweep_screen,
                           epochs=args.sweep_epochs, seed=0, log=log)
    log(f"sweep done in {time.time() - t0:.0f}s: {r}")
    plain_ph = 3600.0 / r["t_plain_mean"]
    surr_ph = r["n_screen"] * 3600.0 / (r["t_screen"] + r["t_rerun"])
    surr_all_ph = r["n_screen"] * 3600.0 / (r["t_gen"] + r["t_train"] + r["t_screen"] + r["t_rerun"])
    batched_ph = r["n_train"] * 3600.0 / r["t_gen"]
    out = {
        "metric": "sweep points/hr vs plain at <=1% phi error",
        "value": surr_ph,
        "unit": "points/hr",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f32-pair/f64-state; MLP bf16 x bf16 -> f32",
        "data": "synthetic",
        "config": {"workload": f"pour-and-settle sweep: 8 parameters, N~U[1000,{cfg.n_max}], "
                               f"{cfg.steps} steps, fill {cfg.fill}",
                   "n_train": r["n_train"], "n_screen": r["n_screen"], "n_rerun": r["n_rerun"],
                   "parallelism": "single-gpu"},
        "plain_points_per_hr": plain_ph,
        "batched_points_per_hr": batched_ph,
        "surrogate_points_per_hr_incl_training": surr_all_ph,
        "speedup_vs_plain": surr_ph / plain_ph,
        "speedup_vs_plain_incl_training": surr_all_ph / plain_ph,
        "batched_speedup_vs_plain": batched_ph / plain_ph,
        "phi_err": {"mean": r["phi_err_mean"], "median": r["phi_err_median"],
                    "p95": r["phi_err_p95"], "max": r["phi_err_max"],
                    "top": r["phi_err_top"], "random": r["phi_err_random"]},
        "z_err_mean": r["z_err_mean"],
        "times_s": {k: r[k] for k in ("t_gen", "t_train", "t_screen", "t_rerun", "t_plain_mean")},
        "train_loss": r["train_loss"],
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--kernel", choices=["auto", "wave", "compact", "loop"], default="auto",
                    help="fused step kernel: per-wave contact compaction (auto, wave), "
                         "block-wide compaction (compact) or the one-thread loop (loop)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep", action="store_true",
                    help="instead of the 2M-sphere step: the batched + surrogate parameter sweep "
                         "against plain runs (prints its own JSON line)")
    ap.add_argument("--sweep-train", type=int, default=4096)
    ap.add_argument("--sweep-screen", type=int, default=100_000)
    ap.add_argument("--sweep-steps", type=int, default=15000)
    ap.add_argument("--sweep-nmax", type=int, default=4000)
    ap.add_argument("--sweep-epochs", type=int, default=400)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}); a packing is not split across GPUs in this build.", file=sys.stderr)
        sys.exit(2)

    import gpucode_amd as G
    from gpucode_amd import workloads as W

    if args.sweep:
        return sweep_main(args)

    t0 = time.time()
    w = W.packed_bed(n=args.n)
    log(f"bed of {args.n} spheres built in {time.time() - t0:.1f}s")
    prec = G.PREC_FP64 if args.precision == "fp64" else G.PREC_MIXED
    e = G.Engine(w["inp"], None, prec)
    e.set_kernel({"auto": G.Engine.KERNEL_AUTO, "wave": G.Engine.KERNEL_WAVE,
                  "compact": G.Engine.KERNEL_COMPACT, "loop": G.Engine.KERNEL_LOOP}[args.kernel])
    e.set_particles(w["pos"])
    log(f"engine up: {e.device_bytes() / 2**30:.2f} GiB on device")
    t0 = time.time()
    e.step(args.warmup)
    e.sync()
    log(f"warmup {args.warmup} steps in {time.time() - t0:.1f}s, rebuilds {e.n_rebuild}")
    st0 = e.stats()
    tm0 = e.timing()
    rb0 = e.n_rebuild
    t0 = time.perf_counter()
    e.step(args.steps)
    e.sync()
    el = time.perf_counter() - t0
    tm1 = e.timing()
    st1 = e.stats()
    n = e.total
    log(f"timed {args.steps} steps in {el:.3f}s, rebuilds {e.n_rebuild - rb0}")

    st = {k: 0.5 * (st0[k] + st1[k]) for k in st0}
    B, n_nb, n_c, n_w = algorithmic_bytes(st, n)
    launches = tm1["fused_launches"] - tm0["fused_launches"]
    kern_ms = (tm1["fused_ms"] - tm0["fused_ms"]) / max(launches, 1)
    achieved = B * n / (kern_ms * 1e-3) / 1e9
    kname = e.step_kernel_name()
    traffic = None
    tr = load_traffic(kname)
    if tr:
        traffic = tr.get("bytes_per_launch")

    copy_gbs = G.engine.stream_copy(1 << 30, 20)
    log(f"plain copy {copy_gbs:.0f} GB/s; {kname} {kern_ms:.4f} ms, {achieved:.0f} GB/s "
        f"algorithmic ({B:.1f} B/particle-step)")

    base = None
    if not args.no_cpu_baseline:
        thr = min(16, os.cpu_count() or 1)
        t0 = time.time()
        base = cpu_baseline(e, w["inp"], thr)
        log(f"cpu baseline {base['value']:.3e} particle-steps/s on {thr} threads "
            f"({time.time() - t0:.1f}s)")

    value = n * args.steps / el
    out = {
        "metric": "particle-steps/s at 2M spheres",
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
            "workload": "packed_bed_2M (dry, monodisperse d=0.3mm, R=30mm, skin 0.1d, mu 0.5)",
            "n_particles": n, "seed": 0xC3, "parallelism": "single-gpu",
            "rebuilds_timed": e.n_rebuild - rb0,
            "n_nb": round(n_nb, 3), "n_contact": round(n_c, 3), "n_wall_contact": round(n_w, 4),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": kname, "kernel_ms": kern_ms, "bytes_per_particle": B,
            "attainable_copy_gbs": copy_gbs,
        },
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
