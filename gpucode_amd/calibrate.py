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


def restitution_to_dmpn(e, d, rho, E, nu, v=0.5):
    r = d / 2.0
    m = rho * PI_REF * d ** 3 / 6.0
    Es = (E / (1.0 - nu * nu)) / 2.0
    tc = 2.87 * ((m / 2) ** 2 / ((r / 2) * Es * Es * v)) ** 0.2
    lo, hi = 0.0, tc
    while collision_restitution(hi, d, rho, E, nu, v) > e:
        hi *= 2.0
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        if collision_restitution(mid, d, rho, E, nu, v) > e:
            lo = mid
        else:
            hi = mid
        if hi - lo < 1e-6 * hi:
            break
    return 0.5 * (lo + hi)
