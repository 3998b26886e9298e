"""Material-parameter calibration (host side, no GPU needed).

restitution_to_dmpn: the damping time constant dmpn [s] of the reference's
normal law (kernel.cu:162-165: F_d = -2 dmpn E* sqrt(R* delta) v_n) that
gives a coefficient of restitution e for a head-on collision of two equal
spheres at impact speed v.  The model's e depends weakly on v (as v^(1/5)),
so the speed is part of the calibration.

yield_for_overlap: the yield pressure yldp [Pa] that puts the p-p yield
overlap ydsp_t * R* at a fraction of the diameter (main.cpp:333-334 with
R* = d/4): ydsp_t = (pi sigma/(2 E*))^2, sigma = yldp/2, E* = E/(2(1-nu^2)).
"""
import math

PI_REF = 3.1415926  # heydem.h:15


def yield_for_overlap(E, nu, frac):
    # (pi (yldp/2) / (2 E/(2(1-nu^2))))^2 * d/4 = frac * d
    return 2.0 * math.sqrt(4.0 * frac) * E / (PI_REF * (1.0 - nu * nu))


def collision_restitution(dmpn, d, rho, E, nu, v, steps_per_contact=4000):
    """e of a head-on elastic-regime collision of two spheres of diameter d."""
    r = d / 2.0
    m = rho * PI_REF * d ** 3 / 6.0
    ms = m / 2.0
    Es = (E / (1.0 - nu * nu)) / 2.0
    Rs = r / 2.0
    tc = 2.87 * (ms * ms / (Rs * Es * Es * v)) ** 0.2
    dt = tc / steps_per_contact
    delta, ddot = 0.0, v
    for _ in range(40 * steps_per_contact):
        f = 4.0 / 3.0 * Es * math.sqrt(Rs) * delta ** 1.5 + 2.0 * dmpn * Es * math.sqrt(Rs * delta) * ddot
        ddot -= f / ms * dt
        delta += ddot * dt
        if delta <= 0.0 and ddot < 0.0:
            return -ddot / v
    raise RuntimeError("collision did not end")


def _contact_time(d, rho, E, nu, v):
    r = d / 2.0
    m = rho * PI_REF * d ** 3 / 6.0
    Es = (E / (1.0 - nu * nu)) / 2.0
    return 2.87 * ((m / 2) ** 2 / ((r / 2) * Es * Es * v)) ** 0.2


_TABLE = None


def _restitution_table(steps_per_contact=4000):
    """e as a function of zeta = dmpn / t_c.  With delta in units of the
    elastic maximum overlap and t in units of t_c, the damped Hertz
    collision ms delta'' = -4/3 Es sqrt(Rs) delta^1.5
    - 2 dmpn Es sqrt(Rs delta) delta' depends on dmpn only through dmpn/t_c,
    so one table (integrated exactly as collision_restitution does, for all
    zeta at once) serves every material and speed."""
    global _TABLE
    if _TABLE is None:
        import numpy as np

        d, rho, E, nu, v = 1e-3, 2500.0, 70e9, 0.25, 0.5
        tc = _contact_time(d, rho, E, nu, v)
        zeta = np.linspace(0.0, 0.5, 2501)
        dmpn = zeta * tc
        r = d / 2.0
        ms = rho * PI_REF * d ** 3 / 6.0 / 2.0
        Es = (E / (1.0 - nu * nu)) / 2.0
        Rs = r / 2.0
        dt = tc / steps_per_contact
        delta = np.zeros_like(zeta)
        ddot = np.full_like(zeta, v)
        e = np.full_like(zeta, np.nan)
        live = np.ones(zeta.shape, bool)
        for _ in range(40 * steps_per_contact):
            dl, dd = delta[live], ddot[live]
            f = 4.0 / 3.0 * Es * math.sqrt(Rs) * dl ** 1.5 + 2.0 * dmpn[live] * Es * np.sqrt(Rs * dl) * dd
            dd = dd - f / ms * dt
            dl = dl + dd * dt
            delta[live], ddot[live] = dl, dd
            end = (dl <= 0.0) & (dd < 0.0)
            if end.any():
                idx = np.where(live)[0][end]
                e[idx] = -dd[end] / v
                live[idx] = False
            if not live.any():
                break
        # an overdamped collision never separates: e = 0
        e[np.isnan(e)] = 0.0
        _TABLE = (zeta, e)
    return _TABLE


def restitution_to_dmpn(e, d, rho, E, nu, v=0.5):
    """dmpn giving restitution e (about 0.06 to 1) at impact speed v: table
    inversion, linear between 2501 points of zeta in [0, 0.5]."""
    import numpy as np

    zeta, et = _restitution_table()
    if not (et[-1] < e <= 1.0):
        raise ValueError(f"restitution {e} outside ({et[-1]:.3g}, 1]")
    # et decreases with zeta
    z = float(np.interp(e, et[::-1], zeta[::-1]))
    return z * _contact_time(d, rho, E, nu, v)
