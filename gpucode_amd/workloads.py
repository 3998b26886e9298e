"""Seeded synthetic workloads of the DEM engine.

Each builder returns a dict with the input deck (DemInput), the diameters
[m] (or None for monodisperse), initial positions in reduced units and the
step count of the workload:

  dry_column   2,000 monodisperse spheres d = 1 mm, rho 2500, E 70 GPa,
               nu 0.25, mu 0.3, e 0.7, yield at 1 % d, cylinder R 10 mm with
               a floor, random non-overlapping start in a 10 x 60 mm column,
               dt 1e-7 s, 20,000 steps, cell 1.1 d
  wet_poly     lognormal diameters (median 0.5 mm, sigma 0.25, clipped
               0.25-1.0 mm), capillary bridges at 1 % liquid, gamma 0.072,
               theta 20 deg, rupture (1 + theta/2) V^(1/3), cylinder R 25 mm
  packed_bed   a dense bed of monodisperse spheres (2,000,000 of d = 0.3 mm
               in R 30 mm, skin 0.1 d, mu 0.5 for the throughput case)

Positions come from the engine's own ParticleFeed (seeded) or from
pocket_bed(), a layer-by-layer deposition onto triangular pockets.
"""
import math

import numpy as np

from . import _capi as A
from .calibrate import restitution_to_dmpn, yield_for_overlap
from .engine import derive, particle_feed


def make_input(NP, d_max, R, H, mu, *, E=70e9, nu=0.25, rho=2500.0, e=0.7, yield_frac=0.01,
               rfrc=0.01, cutGap=0.1, expand=2.0, dt=None, dtFactor=1.0, g=9.81, nLevels=1,
               forceCap=0, liquid=0.0, tension=0.072, gapmn=1e-6, theta=0.0, rupture_mode=0,
               mu_wall=None, feed=None, impact_speed=0.5):
    inp = A.DemInput()
    inp.NP = NP
    inp.nLevels = nLevels
    inp.loadStage = 1
    inp.forceCap = forceCap
    inp.tPacking = 1.0
    inp.realSize = d_max
    inp.cutGap = cutGap
    inp.expandSize = expand
    inp.cirX, inp.cirY, inp.Bw, inp.R, inp.H = 0.0, 0.0, 0.0, R, H
    lo, hi = feed if feed else (0.5 * d_max, H - 0.5 * d_max)
    inp.feedLmin, inp.feedLmax = lo, hi
    dmpn = restitution_to_dmpn(e, d_max, rho, E, nu, impact_speed)
    yldp = yield_for_overlap(E, nu, yield_frac)
    mw = mu if mu_wall is None else mu_wall
    for m, row in enumerate(((rho, E, nu, mu, rfrc, dmpn, yldp), (rho, E, nu, mw, rfrc, dmpn, yldp))):
        for k, v in enumerate(row):
            inp.mat[m][k] = v
    inp.stension = tension
    inp.vol = liquid
    inp.gapmn = gapmn
    inp.dtFactor = dtFactor
    inp.nOutPacking = 1
    inp.gravity = g
    inp.dtOverride = dt or 0.0
    inp.theta = theta
    inp.ruptureMode = rupture_mode
    return inp


def min_cut_gap(inp, dia=None, margin=1.02):
    """Smallest cutGap whose neighbour search reaches the largest capillary
    rupture distance (two of the largest spheres, each holding the liquid of
    12 bridges), times margin; 0 for a dry packing.  The search gap is
    linear in cutGap (searchGap = cutGap x the smallest diameter)."""
    if inp.forceCap != 1:
        return 0.0
    probe = A.DemInput.from_buffer_copy(inp)
    probe.cutGap = 1.0
    S = derive(probe, dia)[0]
    V = 2.0 * S.vol / 12.0
    rup = (1.0 + 0.5 * S.theta) * V ** (1.0 / 3.0) if S.ruptureMode == 1 else 0.99 * V ** 0.33
    return margin * rup / S.searchGap


def feed_positions(inp, dia, n, seed):
    """Initial positions from the engine's ParticleFeed, called until n
    particles are placed (one call places at most 60,000, as in the
    reference) or a call places none."""
    S, rad, _, _ = derive(inp, dia)
    pos = np.zeros((inp.NP, 3))
    count = 0
    while count < n:
        new, seed = particle_feed(S, rad, pos, count, seed)
        if new <= count:
            break
        count = new
    if count < n:
        raise RuntimeError(f"feed placed only {count} of {n} particles")
    return pos[:n].copy()


def dry_column(n=2000, seed=0xC1, steps=20000):
    """2,000 dry monodisperse spheres in a 10 x 60 mm column."""
    d = 1.0e-3
    inp = make_input(n, d, 10e-3, 60e-3, 0.3, cutGap=0.1 / 1.4, dt=1e-7)
    return dict(name="dry_column", inp=inp, dia=None, pos=feed_positions(inp, None, n, seed),
                steps=steps)


def lognormal_diameters(n, seed, median=0.5e-3, sigma=0.25, lo=0.25e-3, hi=1.0e-3):
    rng = np.random.default_rng(seed)
    return np.clip(median * np.exp(sigma * rng.standard_normal(n)), lo, hi)


def wet_poly(n=200_000, seed=0xC2, R=25e-3, H=60e-3, steps=50_000, theta_deg=20.0,
             reference_law=False, nLevels=3, dt=None, E=70e9):
    """Wet polydisperse packing; reference_law selects theta = 0 with the
    reference's rupture rule (kernel.cu:522).  A smaller E gives soft
    spheres whose time step lets the pour settle in tens of thousands of
    steps."""
    dia = lognormal_diameters(n, seed)
    th = 0.0 if reference_law else math.radians(theta_deg)
    inp = make_input(n, float(dia.max()), R, H, 0.3, E=E, cutGap=0.1, nLevels=nLevels, forceCap=1,
                     liquid=0.01, theta=th, rupture_mode=0 if reference_law else 1, dt=dt)
    return dict(name="wet_poly", inp=inp, dia=dia, pos=feed_positions(inp, dia, n, seed + 1),
                steps=steps)


def _trilaterate(p1, p2, p3, d):
    """Centres at distance d from three centres (rows), on the upper side;
    NaN where no such point exists."""
    ex = p2 - p1
    d12 = np.linalg.norm(ex, axis=1)
    ex = ex / d12[:, None]
    t = p3 - p1
    i = (ex * t).sum(1)
    ey = t - i[:, None] * ex
    nj = np.linalg.norm(ey, axis=1)
    ey = ey / nj[:, None]
    ez = np.cross(ex, ey)
    flip = ez[:, 2] < 0
    ez[flip] = -ez[flip]
    j = (ey * t).sum(1)
    x = d12 / 2
    y = (i * i + j * j) / (2 * j) - i * x / j
    z2 = d * d - x * x - y * y
    z = np.sqrt(np.where(z2 > 0, z2, np.nan))
    return p1 + x[:, None] * ex + y[:, None] * ey + z[:, None] * ez


def pocket_bed(n, R, d=1.0, a=1.10, jitter=0.03, seed=0, z0=0.0, pockets=True):
    """Dense bed of n spheres (diameter d, reduced units) in a cylinder of
    radius R centred at the origin, floor at z0.  The floor layer is a
    triangular lattice of spacing a with every site jittered in the plane.
    Each further layer sits in the pockets of the one below (ABC order): a
    sphere is placed touching the three spheres nearest to its lattice site
    (pockets=True), so every sphere rests on three supports and the bed
    starts with about six contacts per sphere; where three supports are not
    available (at the wall) it is lowered onto the highest sphere below."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    r = d / 2
    s3 = math.sqrt(3.0)
    nmax = int(2 * R / a) + 4
    ii, jj = np.meshgrid(np.arange(-nmax, nmax + 1), np.arange(-nmax, nmax + 1), indexing="ij")
    ii, jj = ii.ravel().astype(float), jj.ravel().astype(float)
    out = []
    total = 0
    prev = tree = None
    L = 0
    h = math.sqrt(max(d * d - a * a / 3.0, 0.25 * d * d))
    while total < n:
        ox, oy = (L % 3) * a / 2, (L % 3) * a / (2 * s3)
        x = ii * a + jj * a / 2 + ox
        y = jj * a * s3 / 2 + oy
        if L == 0 or not pockets:
            x = x + jitter * a * rng.uniform(-1, 1, x.shape)
            y = y + jitter * a * rng.uniform(-1, 1, y.shape)
        keep = x * x + y * y <= (R - r - 1e-9) ** 2
        x, y = x[keep], y[keep]
        if L == 0:
            layer = np.stack([x, y, np.full(x.shape, z0 + r)], 1)
        else:
            xy = np.stack([x, y], 1)
            dist, idx = tree.query(xy, k=8, distance_upper_bound=d)
            z = np.full(x.shape, -np.inf)
            for c in range(dist.shape[1]):
                ok = np.isfinite(dist[:, c])
                cand = np.full(x.shape, -np.inf)
                jn = idx[ok, c]
                rho = dist[ok, c]
                cand[ok] = prev[jn, 2] + np.sqrt(np.maximum(d * d - rho * rho, 0.0))
                z = np.maximum(z, cand)
            nominal = z0 + r + L * h
            z = np.where(np.isfinite(z), z, nominal)
            layer = np.stack([x, y, z], 1)
            if pockets:
                ok3 = np.isfinite(dist[:, 2])
                sel = np.where(ok3)[0]
                if len(sel):
                    q = _trilaterate(prev[idx[sel, 0]], prev[idx[sel, 1]], prev[idx[sel, 2]], d)
                    good = np.isfinite(q).all(1)
                    good &= q[:, 0] ** 2 + q[:, 1] ** 2 <= (R - r - 1e-9) ** 2
                    # no overlap with the other spheres below (the three
                    # supports sit at exactly d)
                    if good.any():
                        d4, _ = tree3.query(q[good], k=4)
                        ok = d4[:, 3] >= d * (1 - 1e-12)
                        gi = np.where(good)[0][ok]
                        layer[sel[gi]] = q[gi]
                # keep the layer free of in-plane overlaps: drop the later of
                # any two spheres closer than d
                lt = cKDTree(layer)
                bad = np.zeros(len(layer), bool)
                for a_, b_ in lt.query_pairs(d * (1 - 1e-12)):
                    bad[max(a_, b_)] = True
                layer = layer[~bad]
        if total + len(layer) > n:
            pick = np.sort(rng.choice(len(layer), n - total, replace=False))
            layer = layer[pick]
        out.append(layer)
        total += len(layer)
        prev = layer
        tree = cKDTree(layer[:, :2])
        tree3 = cKDTree(layer)
        L += 1
    return np.concatenate(out)[:n]


def packed_bed(n=2_000_000, d=0.3e-3, R=30e-3, H=30e-3, mu=0.5, skin=0.1, seed=0xC3,
               steps=10_000, dt=None, jitter=0.03):
    """Dense monodisperse bed (the 2M-sphere throughput case by default)."""
    inp = make_input(n, d, R, H, mu, cutGap=skin, dt=dt)
    pos = pocket_bed(n, R / d, 1.0, jitter=jitter, seed=seed)
    if pos[:, 2].max() + 0.5 > H / d:
        raise ValueError("bed taller than the cylinder")
    return dict(name="packed_bed", inp=inp, dia=None, pos=pos, steps=steps)


def pour_lattice(n, R, a=1.5, jitter=0.1, seed=0, z0=0.0):
    """n spheres of diameter 1 (reduced units) on a jittered square lattice of
    spacing a inside a cylinder of radius R, layers from the floor up: a
    loose cloud to be dropped."""
    rng = np.random.default_rng(seed)
    m = int(R / a) + 1
    g = np.arange(-m, m + 1) * a
    x, y = np.meshgrid(g, g, indexing="ij")
    x, y = x.ravel(), y.ravel()
    # the x and y jitter together move a site up to sqrt(2) a jitter outward
    keep = x * x + y * y <= (R - 0.5 - 1.5 * a * jitter) ** 2
    x, y = x[keep], y[keep]
    per = len(x)
    layers = (n + per - 1) // per
    out = []
    for L in range(layers):
        j = rng.uniform(-jitter * a, jitter * a, (per, 3))
        out.append(np.stack([x + j[:, 0], y + j[:, 1], z0 + 0.5 + a * (L + 0.5) + j[:, 2]], 1))
    return np.concatenate(out)[:n]


def pour_2M(n=2_000_000, d=0.3e-3, R=30e-3, H=75e-3, mu=0.5, skin=0.1, v0=2.0, seed=0xC3,
            steps=10_000):
    """The throughput bed's spheres and material as a falling cloud: a loose
    lattice (spacing 1.5 d) dropped at v0 [m/s] onto the floor, so the lists
    are rebuilt every few hundred steps."""
    inp = make_input(n, d, R, H, mu, cutGap=skin)
    pos = pour_lattice(n, R / d, seed=seed)
    if pos[:, 2].max() + 0.5 > H / d:
        raise ValueError("cloud taller than the cylinder")
    S, _, _, _ = derive(inp, None)
    rng = np.random.default_rng(seed + 1)
    vel = rng.normal(size=(n, 3)) * 0.05 * v0 / S.rvunit
    vel[:, 2] -= v0 / S.rvunit
    return dict(name="pour_2M", inp=inp, dia=None, pos=pos, vel=vel, steps=steps)


def wet_bed(n=3000, R_d=8.0, poly=False, theta_deg=0.0, reference_law=True, nLevels=None,
            seed=17, liquid=0.01):
    """A dense wet bed for capillary parity tests (d_max = 1 mm).  poly draws
    diameters uniform in [0.7, 1.0] mm on the monodisperse pocket sites, so
    no pair overlaps and the size levels and cross-level search are used."""
    d = 1.0e-3
    rng = np.random.default_rng(seed)
    dia = None
    if poly:
        dia = d * rng.uniform(0.7, 1.0, n)
        dia[0] = d
    H = d * (n / (math.pi * R_d * R_d * 0.5) + 6)
    th = 0.0 if reference_law else math.radians(theta_deg)
    inp = make_input(n, d, R_d * d, H, 0.3, cutGap=0.1, forceCap=1, liquid=liquid, theta=th,
                     rupture_mode=0 if reference_law else 1,
                     nLevels=nLevels or (3 if poly else 1))
    pos = pocket_bed(n, R_d, 1.0, seed=seed)
    return dict(name="wet_bed", inp=inp, dia=dia, pos=pos, steps=200)


def small_bed(n=3000, R_d=8.0, mu=0.5, seed=7, dt=None, jitter=0.03):
    """A small dense bed for contact-heavy parity tests (d = 1 mm)."""
    d = 1.0e-3
    H = d * (n / (math.pi * R_d * R_d * 0.5) + 6)
    inp = make_input(n, d, R_d * d, H, mu, cutGap=0.1, dt=dt)
    pos = pocket_bed(n, R_d, 1.0, jitter=jitter, seed=seed)
    return dict(name="small_bed", inp=inp, dia=None, pos=pos, steps=2000)


# --- parameter sweeps (batched instances) -------------------------------------
# The eight inputs of the screened sweep: sliding friction, restitution,
# liquid fraction, contact angle, lognormal sigma of the diameters, Young's
# modulus, cylinder-to-particle radius ratio and feed rate.
SWEEP_PARAMS = ("mu", "restitution", "liquid", "theta_deg", "sigma", "log10_E", "R_over_r",
                "feed_frac")
SWEEP_RANGES = {
    "mu": (0.1, 0.8),            # sliding friction (particle-particle and wall)
    "restitution": (0.3, 0.9),   # normal restitution at 0.5 m/s
    "liquid": (0.0, 0.03),       # liquid volume per particle volume
    "theta_deg": (0.0, 60.0),    # contact angle
    "sigma": (0.0, 0.4),         # lognormal sigma of the diameters
    "log10_E": (5.0, 6.0),       # Young's modulus 0.1-1 MPa (soft spheres)
    "R_over_r": (10.0, 30.0),    # cylinder radius / median particle radius
    "feed_frac": (0.1, 0.5),     # the feed spans this fraction of the run (rate = N / (f T))
}
SWEEP_D = 2.0e-3        # median diameter [m]
SWEEP_BATCHES = 10      # feed batches per packing


def sweep_point(rng):
    """One seeded sweep point: the 8 parameters in SWEEP_PARAMS order."""
    return np.array([rng.uniform(*SWEEP_RANGES[k]) for k in SWEEP_PARAMS])


def sweep_instance(x, seed, *, steps=30000, bed_d=8.0, batches=SWEEP_BATCHES, tension=0.072):
    """The pour-and-settle packing of sweep point x (SWEEP_PARAMS order).

    Spheres of median diameter 2 mm, lognormal with the point's sigma and
    clipped to 1-4 mm (one to three size levels), in a cylinder of radius
    R_over_r median radii; N fills a settled bed about bed_d largest
    diameters tall (phi 0.6), and the cylinder is 1.8 times that tall.  The
    spheres are fed as the reference feeds them (main.cpp:436-460,
    ParticleFeed): placed at random, at rest, in the upper half of the
    cylinder, in `batches` equal batches, one every feed_frac * steps /
    batches steps, so the feed spans feed_frac of the run (feed rate N /
    (feed_frac T) spheres per second).  A batch that does not fit in the
    feed zone (the spheres fed before have not yet fallen out of it) is
    completed at the next chunk boundaries: a fast feed is choked by the
    zone, as a real hopper's would be.  Capillary bridges (liquid fraction,
    contact angle, gamma 0.072 N/m, rupture (1 + theta/2) V^(1/3)) when the
    liquid fraction is above 1e-4.  Returns the deck, diameters, the feed
    schedule (step, count) and the feed seed; positions come from
    feed_batch as the run reaches each batch."""
    mu, e, liquid, theta_deg, sigma, lg_E, R_over_r, feed_frac = [float(v) for v in x]
    rng = np.random.default_rng(seed)
    dm = SWEEP_D
    R = 0.5 * R_over_r * dm
    # the bed is bed_d largest diameters tall, so that the analysis's core
    # region (2 d_max inside the bed at each end) holds 6 d_max of it
    rel3, d_est = 1.0, dm
    if sigma > 1e-3:
        probe = lognormal_diameters(4000, int(rng.integers(1 << 31)), median=dm, sigma=sigma,
                                    lo=0.5 * dm, hi=2.0 * dm)
        rel3, d_est = float(np.mean((probe / dm) ** 3)), float(probe.max())
    h = bed_d * d_est
    n = int(round(0.6 * 6.0 * (R / dm) ** 2 * (h / dm) / rel3))
    if sigma > 1e-3:
        dia = lognormal_diameters(n, int(rng.integers(1 << 31)), median=dm, sigma=sigma,
                                  lo=0.5 * dm, hi=2.0 * dm)
        d = float(dia.max())
        ratio = d / float(dia.min())
    else:
        dia, d, ratio = None, dm, 1.0
    nLevels = 1 if ratio <= 1.5 else (2 if ratio <= 2.5 else 3)
    H = 1.8 * h + 2.0 * d
    wet = liquid > 1e-4
    inp = make_input(n, d, R, H, mu, E=10.0 ** lg_E, e=e, cutGap=0.2, nLevels=nLevels,
                     forceCap=1 if wet else 0, liquid=liquid if wet else 0.0, tension=tension,
                     theta=math.radians(theta_deg), rupture_mode=1,
                     feed=(0.5 * H, H - 0.5 * d))
    if dia is not None:
        dmin = float(dia.min())
        inp.sizeLevelInput = 1
        for l in range(nLevels):
            inp.levelDia[l] = d if l == nLevels - 1 else dmin + (d - dmin) / nLevels * (l + 1)
    inp.cutGap = max(inp.cutGap, min_cut_gap(inp, dia))
    every = max(1, int(round(feed_frac * steps / batches)))
    counts = [n * (q + 1) // batches - n * q // batches for q in range(batches)]
    schedule = [(q * every, c) for q, c in enumerate(counts)]
    return dict(name="sweep", inp=inp, dia=dia, pos=np.zeros((0, 3)), schedule=schedule,
                feed_seed=int(rng.integers(1, 1 << 62)), x=np.asarray(x, float), n=n,
                steps=steps)


def feed_due(w, t):
    """Spheres of sweep instance w due by step t (the batches scheduled at or
    before t)."""
    return sum(c for s, c in w["schedule"] if s <= t)


def feed_batch(engine, w, want, state):
    """Feed sweep instance w into engine (a single packing or a batch
    instance) up to `want` spheres: the current positions are read back,
    ParticleFeed places new spheres in the feed zone clear of them (as many
    as fit), and they are appended at rest.  state carries the feed seed
    from call to call; the placement depends only on the instance's own
    state, so a batched run and a plain run feed identically.  Returns the
    number placed."""
    S, rad, _, _ = derive(w["inp"], w["dia"])
    count = engine.total
    if want <= count:
        return 0
    pos = np.zeros((want, 3))
    if count:
        pos[:count] = engine.particles()[0]
    seed = state.get("seed", w["feed_seed"])
    new, seed = particle_feed(S, rad[:want], pos, count, seed)
    state["seed"] = seed
    if new > count:
        engine.feed(pos[count:new])
    return new - count


# --- batched uneven instances (the autotuned multi-instance launch) ----------
C4_RANGES = {
    "n": (1000, 20000),          # particles, uniform integers
    "mu": (0.1, 0.8),            # sliding friction (particle-particle and wall)
    "liquid": (0.0, 0.03),       # liquid volume per particle volume
    "theta_deg": (0.0, 60.0),    # contact angle
    "sigma": (0.0, 0.4),         # lognormal sigma of the diameters
}


def c4_instance(k, seed=0xC4, *, n=None, d_median=0.5e-3, E=1.0e6, e=0.7, tension=0.02, v0=0.5,
                bed_height_d=25.0, fill=4.0, dry_fraction=0.25):
    """Instance k of the batched uneven set: N ~ U[1,000, 20,000] spheres
    and seeded parameters drawn from C4_RANGES.  Liquid fraction: 0 (a dry
    packing) with probability dry_fraction, else uniform in (0, 3 %], so
    both step kernels are exercised.  Diameters lognormal with median 0.5
    mm and the instance's sigma, clipped to 0.25-1.0 mm (sigma = 0:
    monodisperse); one size level up to a diameter ratio of 1.5, two up to
    2.5, else three.  Soft spheres (E = 1 MPa, restitution 0.7) give a
    time step of 2e-6 to 4.5e-6 s.  Surface tension 0.02 N/m (an alcohol's;
    with soft spheres, water's 0.072 N/m would press bridged spheres into
    each other by several per cent of d, while 0.02 keeps that below 1 %).
    The cylinder is sized for a settled bed about bed_height_d mean
    diameters tall (radius clamped to 5-15 of them) and is fill times that
    tall (the random sequential feed needs that room); the fed spheres
    start downward at v0 m/s, so they land within about 0.05 s and the
    30,000 steps pour and settle.  Capillary bridges rupture at
    (1 + theta/2) V^(1/3).  n, if given, replaces the drawn particle count
    (the other parameters stay those of instance k)."""
    rng = np.random.default_rng([seed, k])
    lo, hi = C4_RANGES["n"]
    n_drawn = int(rng.integers(lo, hi + 1))
    n = n_drawn if n is None else int(n)
    mu = float(rng.uniform(*C4_RANGES["mu"]))
    liquid = 0.0 if rng.random() < dry_fraction else float(rng.uniform(1e-4, C4_RANGES["liquid"][1]))
    theta_deg = float(rng.uniform(*C4_RANGES["theta_deg"]))
    sigma = float(rng.uniform(*C4_RANGES["sigma"]))
    dseed = int(rng.integers(1 << 31))
    if sigma < 1e-3:
        dia, d = None, d_median
        ratio = 1.0
    else:
        dia = lognormal_diameters(n, dseed, median=d_median, sigma=sigma)
        d = float(dia.max())
        ratio = d / float(dia.min())
    nLevels = 1 if ratio <= 1.5 else (2 if ratio <= 2.5 else 3)
    # the bed is sized in mean diameters dm (= d when monodisperse)
    dm = d if dia is None else float(np.mean(dia))
    rel3 = 1.0 if dia is None else float(np.mean((dia / dm) ** 3))
    vol_per = (math.pi / 6) / 0.6 * rel3
    R_m = min(15.0, max(5.0, math.sqrt(n * vol_per / (math.pi * bed_height_d))))
    R_d = R_m * dm / d
    H = dm * fill * (n * vol_per / (math.pi * R_m * R_m) + 4)
    inp = make_input(n, d, R_d * d, H, mu, E=E, e=e, cutGap=0.2, nLevels=nLevels,
                     forceCap=1 if liquid > 0 else 0, liquid=liquid, tension=tension,
                     theta=math.radians(theta_deg), rupture_mode=1)
    if dia is not None:
        # level bounds as the default rule draws them, given explicitly so
        # that the top one is the largest diameter itself, not a rounding of it
        dmin = float(dia.min())
        inp.sizeLevelInput = 1
        for l in range(nLevels):
            inp.levelDia[l] = d if l == nLevels - 1 else dmin + (d - dmin) / nLevels * (l + 1)
    # the lists must reach every bridge until it ruptures
    inp.cutGap = max(inp.cutGap, min_cut_gap(inp, dia))
    pos = feed_positions(inp, dia, n, dseed + 1)
    vel = np.zeros_like(pos)
    vel[:, 2] = -v0 / derive(inp, dia)[0].rvunit
    return dict(name=f"c4_{k}", inp=inp, dia=dia, pos=pos, vel=vel,
                params=dict(n=n, mu=mu, liquid=liquid, theta_deg=theta_deg, sigma=sigma,
                            nLevels=nLevels, R_d=R_d))
