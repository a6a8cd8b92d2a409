"""Torch float64 CPU restatement of the SSP-RK3 march for autograd (oracle).

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle/``: it gives the
reference gradients the GPU discrete adjoint is checked against. The reference
offers ``jax.grad`` only of its pointwise fluxes and limiters
(``tests/unit/test_physics.py:164-182,224-229``); jax is not available here,
so this module restates the array program of ``oracle/dg.py``,
``oracle/limiting.py`` and ``oracle/solver.py`` (which cite the reference
lines) in torch ops and lets torch.autograd differentiate it. It shares no
code with the HIP adjoint.

The derivative convention is the one the adjoint documents: the time steps
are given, and in every limiter pass the troubled flags and the positivity
scale are computed from the values (with the NumPy oracle) and then held fixed.
Pinned by central finite differences of the NumPy oracle's forward
(``tests/test_oracle_adjoint.py``).
"""

from __future__ import annotations

import numpy as np
import torch

from .limiting import DETECTION_TOLERANCE, NEIGHBOR_FACTOR, troubled_cell_limit, positivity_limit  # noqa: F401
from .mesh import FIXED_STATE, SLIP_WALL, TRANSMISSIVE

TINY = float(np.finfo(np.float64).tiny)


def _t(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype)


# -- physics (euler.py:21-61, fluxes.py:24-143, limiters.py:16-39) ------------------
def prim(q, gamma):
    rho = q[..., 0]
    u = q[..., 1] / rho
    v = q[..., 2] / rho
    p = (gamma - 1.0) * (q[..., 3] - 0.5 * (q[..., 1] * u + q[..., 2] * v))
    return rho, u, v, p


def normal_flux(q, nx, ny, gamma):
    rho, u, v, p = prim(q, gamma)
    un = u * nx + v * ny
    return torch.stack([rho * un, q[..., 1] * un + p * nx, q[..., 2] * un + p * ny, (q[..., 3] + p) * un], -1)


def cartesian_fluxes(q, gamma):
    rho, u, v, p = prim(q, gamma)
    mx, my, e = q[..., 1], q[..., 2], q[..., 3]
    f = torch.stack([mx, mx * u + p, my * u, (e + p) * u], -1)
    g = torch.stack([my, mx * v, my * v + p, (e + p) * v], -1)
    return f, g


def _safe(x):
    return torch.where(x == 0.0, torch.full_like(x, TINY), x)


def rusanov(qm, qp, nx, ny, gamma):
    rm, um, vm, pm = prim(qm, gamma)
    rp, up, vp, pp = prim(qp, gamma)
    sm = torch.abs(um * nx + vm * ny) + torch.sqrt(gamma * pm / rm)
    sp = torch.abs(up * nx + vp * ny) + torch.sqrt(gamma * pp / rp)
    alpha = torch.maximum(sm, sp)[..., None]
    return 0.5 * (normal_flux(qm, nx, ny, gamma) + normal_flux(qp, nx, ny, gamma)) - 0.5 * alpha * (qp - qm)


def _roe_average(qm, qp, gamma):
    rm, um, vm, pm = prim(qm, gamma)
    rp, up, vp, pp = prim(qp, gamma)
    hm = (qm[..., 3] + pm) / rm
    hp = (qp[..., 3] + pp) / rp
    wm = torch.sqrt(rm)
    wp = torch.sqrt(rp)
    u = (wm * um + wp * up) / (wm + wp)
    v = (wm * vm + wp * vp) / (wm + wp)
    h = (wm * hm + wp * hp) / (wm + wp)
    c = torch.sqrt((gamma - 1.0) * (h - 0.5 * (u * u + v * v)))
    return wm * wp, u, v, h, c


def roe(qm, qp, nx, ny, gamma):
    rm, um, vm, pm = prim(qm, gamma)
    rp, up, vp, pp = prim(qp, gamma)
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
    delta = 0.1 * (torch.abs(un) + c)

    def fix(lam):
        mag = torch.abs(lam)
        return torch.where(mag < delta, (lam * lam + delta * delta) / (2.0 * delta), mag)

    lm, lp, l0 = fix(un - c), fix(un + c), torch.abs(un)
    one = torch.ones_like(un)
    nxo, nyo = nx * one, ny * one
    r_m = torch.stack([one, u - c * nx, v - c * ny, h - c * un], -1)
    r_e = torch.stack([one, u, v, 0.5 * (u * u + v * v)], -1)
    r_s = torch.stack([0.0 * one, -nyo, nxo, ut], -1)
    r_p = torch.stack([one, u + c * nx, v + c * ny, h + c * un], -1)
    diss = ((lm * a_minus)[..., None] * r_m + (l0 * a_entropy)[..., None] * r_e
            + (l0 * a_shear)[..., None] * r_s + (lp * a_plus)[..., None] * r_p)
    return 0.5 * (normal_flux(qm, nx, ny, gamma) + normal_flux(qp, nx, ny, gamma)) - 0.5 * diss


def hllc(qm, qp, nx, ny, gamma):
    rm, um, vm, pm = prim(qm, gamma)
    rp, up, vp, pp = prim(qp, gamma)
    unm = um * nx + vm * ny
    unp = up * nx + vp * ny
    cm = torch.sqrt(gamma * pm / rm)
    cp = torch.sqrt(gamma * pp / rp)
    _, u, v, _, c = _roe_average(qm, qp, gamma)
    un = u * nx + v * ny
    sm = torch.minimum(unm - cm, un - c)
    sp = torch.maximum(unp + cp, un + c)
    mm = rm * (sm - unm)
    mp = rp * (sp - unp)
    ss = (pp - pm + mm * unm - mp * unp) / _safe(mm - mp)

    def star(q, rho, uk, vk, p, unk, sk, mk):
        fac = mk / _safe(sk - ss)
        sh = ss - unk
        return fac[..., None] * torch.stack(
            [torch.ones_like(rho), uk + sh * nx, vk + sh * ny, q[..., 3] / rho + sh * (ss + p / _safe(mk))], -1)

    fm = normal_flux(qm, nx, ny, gamma)
    fp = normal_flux(qp, nx, ny, gamma)
    fsm = fm + sm[..., None] * (star(qm, rm, um, vm, pm, unm, sm, mm) - qm)
    fsp = fp + sp[..., None] * (star(qp, rp, up, vp, pp, unp, sp, mp) - qp)
    sm, sp, ss = sm[..., None], sp[..., None], ss[..., None]
    return torch.where(sm >= 0.0, fm, torch.where(ss >= 0.0, fsm, torch.where(sp > 0.0, fsp, fp)))


FLUXES = {"rusanov": rusanov, "roe": roe, "hllc": hllc}


def _zero(x):
    return torch.zeros_like(x)


def phi(name, r):
    z, o = torch.zeros_like(r), torch.ones_like(r)
    if name == "minmod":
        return torch.maximum(z, torch.minimum(o, r))
    if name == "mc":
        return torch.maximum(z, torch.minimum(torch.minimum(2.0 * r, 0.5 * (1.0 + r)), 2.0 * o))
    if name == "van_leer":
        return (r + torch.abs(r)) / (1.0 + torch.abs(r))
    if name == "superbee":
        return torch.maximum(z, torch.maximum(torch.minimum(2.0 * r, o), torch.minimum(r, 2.0 * o)))
    return o


# -- DG right-hand side (dg.py:140-165) ---------------------------------------------
class TorchOperators:
    def __init__(self, ops, lops):
        Np = ops.lift.shape[0]
        self.dr = _t(ops.differentiation[:, :Np])
        self.ds = _t(ops.differentiation[:, Np:])
        self.lift = _t(ops.lift)
        self.face_nodes = _t(ops.face_nodes, torch.int64)
        self.rx, self.ry, self.sx, self.sy = (_t(ops.inverse_metric[i])[None, :, None] for i in range(4))
        self.nx = _t(ops.normal_x)
        self.ny = _t(ops.normal_y)
        self.fscale = _t(ops.face_scale)
        self.ext = _t(ops.exterior_index, torch.int64)
        self.kind = _t(ops.boundary_kind, torch.int64)
        self.bstate = _t(ops.boundary_state)
        # limiter tables (limiting.py:34-48)
        self.avg_w = _t(lops.average_weights)
        self.mid = _t(lops.midpoint_nodes, torch.int64)
        self.p1mid = _t(lops.p1_midpoint)
        self.linb = _t(lops.linear_basis)
        self.nbr = _t(lops.neighbor, torch.int64)
        self.nkind = _t(lops.neighbor_kind, torch.int64)
        self.lnx = _t(lops.normal_x)
        self.lny = _t(lops.normal_y)
        self.fixed = _t(lops.fixed_state)
        self.alpha = _t(lops.alpha)
        self.beta = _t(lops.beta)
        self.partner = _t(lops.partner_face, torch.int64)


def mirror(q, nx, ny):
    mn = q[..., 1] * nx + q[..., 2] * ny
    z = torch.zeros_like(mn)
    return q - torch.stack([z, 2.0 * mn * nx, 2.0 * mn * ny, z], -1)


def dg_rhs(q, T, flux, gamma):
    f, g = cartesian_fluxes(q, gamma)
    ft = T.rx * f + T.ry * g
    gt = T.sx * f + T.sy * g
    div = torch.einsum("ij,jkc->ikc", T.dr, ft) + torch.einsum("ij,jkc->ikc", T.ds, gt)
    qm = q[T.face_nodes]
    qn = q.reshape(-1, 4)[T.ext]
    kind = T.kind[..., None]
    wall = mirror(qm, T.nx, T.ny)
    qp = torch.where(kind == SLIP_WALL, wall,
                     torch.where(kind == FIXED_STATE, T.bstate, torch.where(kind == TRANSMISSIVE, qm, qn)))
    fhat = flux(qm, qp, T.nx, T.ny, gamma)
    fminus = f[T.face_nodes] * T.nx[..., None] + g[T.face_nodes] * T.ny[..., None]
    jump = T.fscale[..., None] * (fminus - fhat)
    return torch.einsum("ij,jkc->ikc", T.lift, jump) - div


# -- limiter with frozen decisions (limiting.py:128-216) -----------------------------
def limit_frozen(a, T, limiter, troubled, theta):
    """apply_limiters' output with the troubled flags (K,) and the positivity
    scale (K,) given."""
    avg = torch.einsum("i,ikc->kc", T.avg_w, a)
    y = a
    if bool(troubled.any()):
        interior = avg[T.nbr]
        wall = mirror(avg[None], T.lnx, T.lny)
        kind = T.nkind[..., None]
        nb = torch.where(kind == SLIP_WALL, wall, torch.where(kind == FIXED_STATE, T.fixed, interior))
        diff = nb - avg[None]
        paired = torch.take_along_dim(diff, T.partner[..., None], dim=0)
        sn = NEIGHBOR_FACTOR * (T.alpha[..., None] * diff + T.beta[..., None] * paired)
        proj = torch.einsum("fi,ikc->fkc", T.p1mid, a) - avg[None]
        nz = proj != 0.0
        ratio = sn / torch.where(nz, proj, torch.ones_like(proj))
        d = torch.where(nz, proj * torch.minimum(torch.ones_like(ratio), phi(limiter, ratio)), _zero(proj))
        zero = _zero(d)
        pos = torch.maximum(d, zero).sum(0)
        neg = torch.maximum(-d, zero).sum(0)
        one = torch.ones_like(pos)
        tp = torch.where(pos > 0.0, torch.minimum(one, neg / torch.where(pos > 0.0, pos, one)), _zero(pos))
        tn = torch.where(neg > 0.0, torch.minimum(one, pos / torch.where(neg > 0.0, neg, one)), _zero(neg))
        d = tp[None] * torch.maximum(d, zero) - tn[None] * torch.maximum(-d, zero)
        linear = avg[None] + torch.einsum("if,fkc->ikc", T.linb, d)
        y = torch.where(troubled[None, :, None], linear, a)
    th = theta[None, :, None]
    return torch.where(th < 1.0, avg[None] + th * (y - avg[None]), y)


def limiter_decisions(a_np, lops, phi_np, gamma, limit_slopes):
    """Troubled flags and positivity scale of one apply_limiters pass (NumPy oracle)."""
    avg = np.einsum("i,ikc->kc", lops.average_weights, a_np)
    q = a_np
    if limit_slopes:
        with np.errstate(all="ignore"):
            q, troubled = troubled_cell_limit(a_np, avg, lops, phi_np)
    else:
        troubled = np.zeros(a_np.shape[1], dtype=bool)
    with np.errstate(all="ignore"):
        _, theta = positivity_limit(q, avg, gamma, return_theta=True)
    return troubled, theta


def march(q0, dts, solver, T):
    """qN from q0 (torch, requires_grad allowed) over the given steps with
    frozen limiter decisions; ``solver`` is an oracle EulerSolver."""
    gamma = solver.config.gamma
    flux = FLUXES[solver.config.flux]
    lim = solver.config.limiter
    lops = solver.limiter_operators

    def limit(a):
        tr, th = limiter_decisions(a.detach().numpy(), lops, solver.phi, gamma, solver.limit_slopes)
        return limit_frozen(a, T, lim, torch.as_tensor(tr), _t(th))

    q = q0
    for dt in dts:
        q1 = limit(q + dt * dg_rhs(q, T, flux, gamma))
        q2 = limit(0.75 * q + 0.25 * (q1 + dt * dg_rhs(q1, T, flux, gamma)))
        q = limit(q / 3.0 + 2.0 / 3.0 * (q2 + dt * dg_rhs(q2, T, flux, gamma)))
    return q


def march_vjp(solver, q0, dts, lam):
    """(dqN/dq0)^T lam and qN (NumPy arrays)."""
    T = TorchOperators(solver.operators, solver.limiter_operators)
    x = _t(q0).requires_grad_(True)
    qn = march(x, [float(d) for d in dts], solver, T)
    (g,) = torch.autograd.grad(qn, x, _t(lam))
    return g.numpy(), qn.detach().numpy()


def rhs_vjp(solver, q, mu):
    T = TorchOperators(solver.operators, solver.limiter_operators)
    x = _t(q).requires_grad_(True)
    r = dg_rhs(x, T, FLUXES[solver.config.flux], solver.config.gamma)
    (g,) = torch.autograd.grad(r, x, _t(mu))
    return g.numpy()


def limit_vjp(solver, q, lam):
    T = TorchOperators(solver.operators, solver.limiter_operators)
    tr, th = limiter_decisions(np.asarray(q), solver.limiter_operators, solver.phi, solver.config.gamma,
                               solver.limit_slopes)
    x = _t(q).requires_grad_(True)
    out = limit_frozen(x, T, solver.config.limiter, torch.as_tensor(tr), _t(th))
    (g,) = torch.autograd.grad(out, x, _t(lam))
    return g.numpy()
