"""Exact solution of the 1-D Riemann problem for an ideal gas (Toro, ch. 4):
Newton iteration on the star pressure, then self-similar sampling. Written
independently of the solver code, as the reference's verification is
(tests/verification/sod_reference.py)."""

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        sq = np.sqrt(a / (p + b))
        return (p - pk) * sq, sq * (1.0 - 0.5 * (p - pk) / (p + b))
    r = p / pk  # rarefaction
    return 2.0 * ck / (g - 1.0) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0), r ** (-(g + 1.0) / (2.0 * g)) / (rho * ck)


def star_state(left, right, g=1.4):
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, g)
        fr, dr = _f(p, rr, pr, cr, g)
        step = (fl + fr + ur - ul) / (dl + dr)
        p_new = max(1e-12, p - step)
        if abs(p_new - p) < 1e-14 * p:
            p = p_new
            break
        p = p_new
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(x, t, x0, left, right, g=1.4):
    """(rho, u, p) at positions x and time t > 0."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    ps, us = star_state(left, right, g)
    s = (np.asarray(x, float) - x0) / t
    rho, u, p = np.empty_like(s), np.empty_like(s), np.empty_like(s)
    for i, si in enumerate(s.ravel()):
        if si <= us:  # left of contact
            if ps > pl:
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                val = (rl, ul, pl) if si <= sl else (rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1)), us, ps)
            else:
                head, cs = ul - cl, cl * (ps / pl) ** ((g - 1) / (2 * g))
                tail = us - cs
                if si <= head:
                    val = (rl, ul, pl)
                elif si >= tail:
                    val = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - si))
                    val = (rl * (c / cl) ** (2 / (g - 1)), 2 / (g + 1) * (cl + (g - 1) / 2 * ul + si), pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                val = (rr, ur, pr) if si >= sr else (rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)), us, ps)
            else:
                head, cs = ur + cr, cr * (ps / pr) ** ((g - 1) / (2 * g))
                tail = us + cs
                if si >= head:
                    val = (rr, ur, pr)
                elif si <= tail:
                    val = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - si))
                    val = (rr * (c / cr) ** (2 / (g - 1)), 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + si), pr * (c / cr) ** (2 * g / (g - 1)))
        rho.flat[i], u.flat[i], p.flat[i] = val
    return rho, u, p


def sod_waves(t, x0=0.5, g=1.4):
    """Rarefaction head, contact and shock positions of the Sod problem."""
    left, right = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    ps, us = star_state(left, right, g)
    cl, cr = np.sqrt(g), np.sqrt(g * 0.1 / 0.125)
    sr = cr * np.sqrt((g + 1) / (2 * g) * ps / 0.1 + (g - 1) / (2 * g))
    return {"head": x0 - cl * t, "contact": x0 + us * t, "shock": x0 + sr * t}
