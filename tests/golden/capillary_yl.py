"""Pendular liquid bridge between two equal spheres from the Young-Laplace
equation: an independent physical check of the oracle's capillary law.

The oracle's capillary force (kernel.cu:65-81, used at kernel.cu:537-545)
is Willett et al. (2000)'s regression F* = F / (2 pi R gamma) =
exp(f1 - f2 exp(f3 ln S+ + f4 ln^2 S+)), fitted by its authors to numerical
solutions of the Young-Laplace equation.  This module computes those
solutions again from the geometry alone, so the regression's coefficients
can be tested against physics instead of against a retyped table.

Geometry (lengths in units of the sphere radius R = 1): spheres centred at
z = +-(1 + S/2) on the z axis, gap S, bridge volume V, contact angle theta.
The meniscus is a surface of revolution, symmetric about z = 0, with a neck
of radius r0 at z = 0.  Along its meridian, parametrised by arc length s
with psi the angle between the tangent and the z axis,

    dr/ds = sin psi,  dz/ds = cos psi,  dpsi/ds = cos psi / r - P,

where P = dp / gamma is the constant mean-curvature sum (Laplace pressure
of the liquid over surface tension).  r cos psi - (P/2) r^2 is constant
along the meridian (the axial force balance), so the attractive force on
either sphere is F = 2 pi gamma C with C = r0 - (P/2) r0^2 (neck, or
"gorge", method).  The meniscus meets the upper sphere at filling angle
beta, at r_c = sin beta, z_c = S/2 + 1 - cos beta, with psi_c =
pi/2 - beta - theta (contact angle theta measured through the liquid).

The bridges at a gap S form a family parametrised by the neck radius r0.
For given r0 and beta the first integral fixes P; beta is the root at
which the meridian started at the neck reaches r_c at height z_c, and the
volume follows by quadrature.  Along the family the volume has a single
minimum: the bridges with thicker necks are the stable ones, and the
smallest volume is where a bridge of that volume ruptures.  The solution
is checked by virtual work: the neck force equals dE/dS at constant volume,
E = gamma (A_lv - cos theta A_sl), computed from the areas alone.

Running this file writes capillary_yl.json next to it: F* on a grid of
(theta, V* = V/R^3, S+ = (S/2) sqrt(R/V)), with the filling angle.
"""
import json
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp
from scipy.optimize import brentq

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "capillary_yl.json")


def _meridian(r0, P, rc):
    """Meridian from the neck (r0, z = 0) to r = rc: (z, vol, area) there,
    vol and area being those of the half bridge under the meridian; None if
    it turns back first."""
    def rhs(s, y):
        r, z, psi, v, a = y
        return [np.sin(psi), np.cos(psi), np.cos(psi) / r - P,
                np.pi * r * r * np.cos(psi), 2.0 * np.pi * r]

    def reach(s, y):
        return y[0] - rc

    def turn(s, y):
        return y[2] - 0.5 * np.pi

    reach.terminal = True
    reach.direction = 1
    turn.terminal = True
    turn.direction = 1
    sol = solve_ivp(rhs, [0.0, 50.0], [r0, 0.0, 0.0, 0.0, 0.0], events=[reach, turn],
                    rtol=1e-11, atol=1e-14, method="DOP853")
    if len(sol.t_events[0]) == 0:
        return None
    y = sol.y_events[0][0]
    return y[1], y[3], y[4]


def _pressure(r0, theta, beta):
    """P that makes the axial force at the neck, r0 - (P/2) r0^2, equal the
    one at the contact line, sin(beta) sin(beta + theta) - (P/2) sin^2(beta)."""
    rc = np.sin(beta)
    return 2.0 * (r0 - rc * np.sin(beta + theta)) / (r0 * r0 - rc * rc)


def _height_miss(S, theta, r0, beta):
    """Height of the meridian at r = sin(beta) minus the contact line's."""
    m = _meridian(r0, _pressure(r0, theta, beta), np.sin(beta))
    return np.nan if m is None else m[0] - (0.5 * S + 1.0 - np.cos(beta))


def bridge(S, theta, r0, beta_hint=None):
    """The bridge with neck radius r0 at gap S: dict with V, Fstar, beta, P,
    r0, A_lv and A_sl of the whole bridge, or None."""
    b_lo = np.arcsin(min(r0, 0.999)) * 1.0001
    if beta_hint is not None and beta_hint > b_lo:
        grid = np.unique(np.clip(beta_hint * np.array([0.9, 0.97, 1.03, 1.1]), b_lo, 1.4))
        grid = np.concatenate([grid, np.linspace(b_lo, 1.4, 40)])
    else:
        grid = np.linspace(b_lo, 1.4, 40)
    pts = []
    root = None
    for b in grid:
        v = _height_miss(S, theta, r0, b)
        if not np.isfinite(v):
            continue
        # a sign change between this sample and the nearest earlier one on
        # either side brackets the root
        below = [p for p in pts if p[0] < b]
        above = [p for p in pts if p[0] > b]
        for nb in ([max(below)] if below else []) + ([min(above)] if above else []):
            if nb[1] * v < 0.0:
                lo, hi = (nb[0], b) if nb[0] < b else (b, nb[0])
                try:
                    root = brentq(lambda x: _height_miss(S, theta, r0, x), lo, hi,
                                  xtol=1e-15, rtol=4e-15)
                except ValueError:  # the meridian turns back inside the bracket
                    continue
                break
        if root is not None:
            break
        pts.append((b, v))
    if root is None:
        return None
    P = _pressure(r0, theta, root)
    z, vcyl, area = _meridian(r0, P, np.sin(root))
    h = 1.0 - np.cos(root)  # height of the wetted cap of each sphere
    cap = np.pi * h * h * (3.0 - h) / 3.0
    return {"V": 2.0 * (vcyl - cap), "Fstar": r0 - 0.5 * P * r0 * r0, "beta": root, "P": P,
            "r0": r0, "A_lv": 2.0 * area, "A_sl": 2.0 * (2.0 * np.pi * h)}


R0S = np.geomspace(2e-3, 0.95, 56)


def _family(S, theta):
    """V along the family of bridges at gap S, on the R0S grid."""
    out, hint = [], None
    for r0 in R0S:
        r = bridge(S, theta, r0, hint)
        out.append(r)
        if r is not None:
            hint = r["beta"]
    return out


def solve(S, theta, V):
    """Stable bridge of volume V at gap S (R = 1), or None beyond rupture.
    Along the family of bridges at a fixed gap, parametrised by the neck
    radius r0, the volume falls to a minimum (thin, unstable necks) and then
    rises (stable bridges); the stable bridge is where V(r0) crosses V
    upwards."""
    fam = _family(S, theta)
    for k in range(len(R0S) - 1):
        a, b = fam[k], fam[k + 1]
        if a is None or b is None or not (a["V"] < V < b["V"]):
            continue
        hint = [a["beta"]]

        def g(r0):
            r = bridge(S, theta, r0, hint[0])
            if r is None:
                return np.nan
            hint[0] = r["beta"]
            return r["V"] - V

        try:
            r0 = brentq(g, R0S[k], R0S[k + 1], xtol=1e-15, rtol=4e-15)
        except ValueError:  # a member of the family is missing inside the bracket
            continue
        return bridge(S, theta, r0, hint[0])
    return None


def min_volume(S, theta):
    """Smallest volume a bridge at gap S can hold (the fold of V(r0))."""
    from scipy.optimize import minimize_scalar

    fam = _family(S, theta)
    vs = np.array([np.inf if r is None else r["V"] for r in fam])
    k = int(np.argmin(vs))
    lo, hi = R0S[max(k - 1, 0)], R0S[min(k + 1, len(R0S) - 1)]

    def v(lr):
        r = bridge(S, theta, np.exp(lr), fam[k]["beta"])
        return np.inf if r is None else r["V"]

    res = minimize_scalar(v, bounds=(np.log(lo), np.log(hi)), method="bounded",
                          options={"xatol": 1e-9})
    return min(res.fun, vs[k])


def surface_energy(S, theta, V):
    """E / gamma = A_lv - cos(theta) A_sl of the stable bridge, or None."""
    r = solve(S, theta, V)
    if r is None:
        return None
    return r["A_lv"] - np.cos(theta) * r["A_sl"]


def fstar(theta_deg, Vstar, Splus):
    """F / (2 pi R gamma) at Willett's scaled variables, S+ = (S/2) sqrt(R/V)
    with S the full gap: (F*, beta) or None."""
    S = 2.0 * Splus * np.sqrt(Vstar)
    r = solve(S, np.radians(theta_deg), Vstar)
    return None if r is None else (r["Fstar"], r["beta"])


def rupture_gap(theta, V):
    """Largest gap S (R = 1) at which a bridge of volume V exists: the gap
    whose smallest bridge volume is V."""
    lo, hi = 0.2 * V ** (1.0 / 3.0), 2.0 * V ** (1.0 / 3.0)
    return brentq(lambda S: min_volume(S, theta) - V, lo, hi, xtol=1e-9, rtol=1e-9)


RUPTURE_POINTS = [(th, v) for th in (0, 20, 40) for v in (1e-3, 1e-2, 3e-2)]


def _rupture_one(p):
    th, v = p
    return {"theta_deg": th, "Vstar": v, "S_c": float(rupture_gap(np.radians(th), v))}


THETAS = [0, 20, 40, 60]
VSTARS = [1e-5, 3e-5, 1e-4, 3e-4, 1e-3, 3e-3, 1e-2, 3e-2, 1e-1]
SPLUS = [0.005, 0.01, 0.02, 0.05, 0.1, 0.15, 0.2, 0.3, 0.4, 0.5, 0.7, 1.0]


def _one(p):
    th, v, s = p
    try:
        r = fstar(th, v, s)
    except Exception:  # a failed solve is recorded as "no bridge"
        r = None
    return {"theta_deg": th, "Vstar": v, "Splus": s,
            "Fstar": None if r is None else float(r[0]),
            "beta": None if r is None else float(r[1])}


def generate(path=FIXTURE, workers=8):
    from multiprocessing import Pool

    pts = [(th, v, s) for th in THETAS for v in VSTARS for s in SPLUS]
    with Pool(workers) as pool:
        rows = pool.map(_one, pts)
        rupture = pool.map(_rupture_one, RUPTURE_POINTS)
    doc = {
        "what": "Young-Laplace pendular bridge between equal spheres: F* = F/(2 pi R gamma) "
                "at S+ = (S/2) sqrt(R/V), V* = V/R^3; null where no stable bridge was found "
                "(beyond rupture).  rupture: the largest gap S_c/R with a stable bridge.",
        "made_by": "tests/golden/capillary_yl.py",
        "rows": rows,
        "rupture": rupture,
    }
    with open(path, "w") as f:
        f.write("{\n")
        f.write(f' "what": {json.dumps(doc["what"])},\n "made_by": {json.dumps(doc["made_by"])},\n')
        f.write(' "rows": [\n' + ",\n".join(json.dumps(r) for r in rows) + "\n ],\n")
        f.write(' "rupture": [\n' + ",\n".join(json.dumps(r) for r in rupture) + "\n ]\n}\n")
    return doc


if __name__ == "__main__":
    doc = generate(workers=int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    n = sum(1 for r in doc["rows"] if r["Fstar"] is not None)
    print(f"wrote {FIXTURE}: {n} of {len(doc['rows'])} points have a bridge")
