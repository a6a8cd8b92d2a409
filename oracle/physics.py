"""Pointwise Euler relations, interface fluxes and limiter functions (oracle).

Restates ``src/cfd_demo/euler.py``, ``fluxes.py`` and ``limiters.py`` in NumPy,
keeping the reference's operation order (it fixes the rounding).
"""

from __future__ import annotations

import numpy as np

NUM_VARIABLES = 4  # euler.py:11
ROE_ENTROPY_FIX = 0.1  # fluxes.py:21


def primitive_to_conserved(rho, u, v, p, gamma):
    """``euler.py:14-18``."""
    rho, u, v, p = (np.asarray(x, dtype=np.float64) for x in (rho, u, v, p))
    energy = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack(np.broadcast_arrays(rho, rho * u, rho * v, energy), axis=-1)


def conserved_to_primitive(q, gamma):
    """``euler.py:21-27``."""
    rho = q[..., 0]
    u = q[..., 1] / rho
    v = q[..., 2] / rho
    p = (gamma - 1.0) * (q[..., 3] - 0.5 * (q[..., 1] * u + q[..., 2] * v))
    return rho, u, v, p


def pressure(q, gamma):
    """``euler.py:30-32``."""
    return conserved_to_primitive(q, gamma)[3]


def sound_speed(rho, p, gamma):
    """``euler.py:35-37``."""
    return np.sqrt(gamma * p / rho)


def cartesian_fluxes(q, gamma):
    """``euler.py:40-46``."""
    rho, u, v, p = conserved_to_primitive(q, gamma)
    mx, my, e = q[..., 1], q[..., 2], q[..., 3]
    f = np.stack([mx, mx * u + p, my * u, (e + p) * u], axis=-1)
    g = np.stack([my, mx * v, my * v + p, (e + p) * v], axis=-1)
    return f, g


def normal_flux(q, nx, ny, gamma):
    """``euler.py:49-61``."""
    rho, u, v, p = conserved_to_primitive(q, gamma)
    un = u * nx + v * ny
    return np.stack([rho * un, q[..., 1] * un + p * nx, q[..., 2] * un + p * ny, (q[..., 3] + p) * un], axis=-1)


def _safe_denominator(x):
    """``fluxes.py:24-25``."""
    return np.where(x == 0.0, np.finfo(np.float64).tiny, x)


def rusanov(qm, qp, nx, ny, gamma):
    """``fluxes.py:28-37``."""
    rm, um, vm, pm = conserved_to_primitive(qm, gamma)
    rp, up, vp, pp = conserved_to_primitive(qp, gamma)
    sm = np.abs(um * nx + vm * ny) + sound_speed(rm, pm, gamma)
    sp = np.abs(up * nx + vp * ny) + sound_speed(rp, pp, gamma)
    alpha = np.maximum(sm, sp)[..., None]
    return 0.5 * (normal_flux(qm, nx, ny, gamma) + normal_flux(qp, nx, ny, gamma)) - 0.5 * alpha * (qp - qm)


def _roe_average(qm, qp, gamma):
    """``fluxes.py:40-51``."""
    rm, um, vm, pm = conserved_to_primitive(qm, gamma)
    rp, up, vp, pp = conserved_to_primitive(qp, gamma)
    hm = (qm[..., 3] + pm) / rm
    hp = (qp[..., 3] + pp) / rp
    wm = np.sqrt(rm)
    wp = np.sqrt(rp)
    u = (wm * um + wp * up) / (wm + wp)
    v = (wm * vm + wp * vp) / (wm + wp)
    h = (wm * hm + wp * hp) / (wm + wp)
    c = np.sqrt((gamma - 1.0) * (h - 0.5 * (u * u + v * v)))
    return wm * wp, u, v, h, c


def roe(qm, qp, nx, ny, gamma):
    """Roe flux with Harten's fix on the acoustic waves (``fluxes.py:54-98``)."""
    rm, um, vm, pm = conserved_to_primitive(qm, gamma)
    rp, up, vp, pp = conserved_to_primitive(qp, gamma)
    rho, u, v, h, c = _roe_average(qm, qp, gamma)
    un = u * nx + v * ny
    ut = -u * ny + v * nx
    drho = rp - rm
    dp = pp - pm
    dun = (up * nx + vp * ny) - (um * nx + vm * ny)
    dut = (-up * ny + vp * nx) - (-um * ny + vm * nx)
    c2 = c * c
    a_minus = (dp - rho * c * dun) / (2.0 * c2)
    a_entropy = drho - dp / c2
    a_shear = rho * dut
    a_plus = (dp + rho * c * dun) / (2.0 * c2)
    delta = ROE_ENTROPY_FIX * (np.abs(un) + c)

    def fix(lam):
        mag = np.abs(lam)
        return np.where(mag < delta, (lam ** 2 + delta ** 2) / (2.0 * delta), mag)

    lm, lp, l0 = fix(un - c), fix(un + c), np.abs(un)
    one, zero = np.ones_like(un), np.zeros_like(un)
    r_m = np.stack([one, u - c * nx, v - c * ny, h - c * un], -1)
    r_e = np.stack([one, u, v, 0.5 * (u * u + v * v)], -1)
    r_s = np.stack([zero, -ny * one, nx * one, ut], -1)
    r_p = np.stack([one, u + c * nx, v + c * ny, h + c * un], -1)
    diss = ((lm * a_minus)[..., None] * r_m + (l0 * a_entropy)[..., None] * r_e
            + (l0 * a_shear)[..., None] * r_s + (lp * a_plus)[..., None] * r_p)
    return 0.5 * (normal_flux(qm, nx, ny, gamma) + normal_flux(qp, nx, ny, gamma)) - 0.5 * diss


def hllc(qm, qp, nx, ny, gamma):
    """HLLC with Einfeldt/Roe speeds (``fluxes.py:101-143``)."""
    rm, um, vm, pm = conserved_to_primitive(qm, gamma)
    rp, up, vp, pp = conserved_to_primitive(qp, gamma)
    unm = um * nx + vm * ny
    unp = up * nx + vp * ny
    cm = sound_speed(rm, pm, gamma)
    cp = sound_speed(rp, pp, gamma)
    _, u, v, _, c = _roe_average(qm, qp, gamma)
    un = u * nx + v * ny
    sm = np.minimum(unm - cm, un - c)
    sp = np.maximum(unp + cp, un + c)
    mm = rm * (sm - unm)
    mp = rp * (sp - unp)
    ss = (pp - pm + mm * unm - mp * unp) / _safe_denominator(mm - mp)

    def star(q, rho, uk, vk, p, unk, sk, mk):
        fac = mk / _safe_denominator(sk - ss)
        sh = ss - unk
        return fac[..., None] * np.stack(
            [np.ones_like(rho), uk + sh * nx, vk + sh * ny, q[..., 3] / rho + sh * (ss + p / _safe_denominator(mk))], -1)

    fm = normal_flux(qm, nx, ny, gamma)
    fp = normal_flux(qp, nx, ny, gamma)
    fsm = fm + sm[..., None] * (star(qm, rm, um, vm, pm, unm, sm, mm) - qm)
    fsp = fp + sp[..., None] * (star(qp, rp, up, vp, pp, unp, sp, mp) - qp)
    sm, sp, ss = sm[..., None], sp[..., None], ss[..., None]
    return np.where(sm >= 0.0, fm, np.where(ss >= 0.0, fsm, np.where(sp > 0.0, fsp, fp)))


INTERFACE_FLUXES = {"rusanov": rusanov, "roe": roe, "hllc": hllc}  # fluxes.py:146-150


def get_interface_flux(name):
    """``fluxes.py:153-159``."""
    try:
        return INTERFACE_FLUXES[name]
    except KeyError:
        raise ValueError(f"Unknown interface flux {name!r}; choose one of: {', '.join(sorted(INTERFACE_FLUXES))}.") from None


def minmod(r):
    """``limiters.py:16-18``."""
    return np.maximum(0.0, np.minimum(1.0, r))


def mc(r):
    """``limiters.py:21-23``."""
    return np.maximum(0.0, np.minimum(np.minimum(2.0 * r, 0.5 * (1.0 + r)), 2.0))


def van_leer(r):
    """``limiters.py:26-28``."""
    return (r + np.abs(r)) / (1.0 + np.abs(r))


def superbee(r):
    """``limiters.py:31-33``."""
    return np.maximum(0.0, np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0)))


def unlimited(r):
    """``limiters.py:36-39``."""
    r = np.asarray(r)
    return np.ones_like(r, dtype=np.result_type(r, 1.0))


LIMITER_FUNCTIONS = {"minmod": minmod, "mc": mc, "van_leer": van_leer, "superbee": superbee,
                     "unlimited": unlimited}  # limiters.py:42-48


def get_limiter_function(name):
    """``limiters.py:51-57``."""
    try:
        return LIMITER_FUNCTIONS[name]
    except KeyError:
        raise ValueError(f"Unknown limiter {name!r}; choose one of: {', '.join(sorted(LIMITER_FUNCTIONS))}.") from None
