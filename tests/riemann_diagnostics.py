"""Verification diagnostics for Riemann-problem runs (test-side NumPy code).

Restates the reference's ADR 0005 diagnostics (tests/verification/
diagnostics.py:42-245 and sod_reference.py) for the GPU path, vectorised so
that the 256 batched shock tubes can be measured too:

* ``RiemannSolution``: the exact self-similar solution of a 1-D Riemann
  problem for an ideal gas (Toro ch. 4: Newton iteration on the star
  pressure, then the wave fan), with every wave edge as a breakpoint;
* ``SplitQuadrature``: a collapsed Gauss rule on every element, with the
  elements that straddle a breakpoint cut at that vertical line first, so
  the discontinuities of the exact solution fall on sub-triangle edges;
* the degree-4 nodal solution evaluated at the quadrature points through its
  monomial coefficients;
* domain-averaged L1 errors, mesh-column density averages and the shock
  locator (the column face with the largest density drop in a window).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

QUAD_POINTS = 10  # Gauss points per direction of the collapsed rule (diagnostics.py:21)
SPLIT_TOL = 1e-12  # relative slack for a breakpoint to count as inside an element (diagnostics.py:23)


# ---------------------------------------------------------------------------
# Exact solution
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class RiemannSolution:
    """Exact solution for left/right primitive states (rho, u, p) separated
    at x0; ``gamma`` the ratio of specific heats."""

    left: tuple
    right: tuple
    x0: float
    gamma: float = 1.4

    def __post_init__(self):
        g = self.gamma
        (rl, ul, pl), (rr, ur, pr) = self.left, self.right
        cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
        if 2.0 * (cl + cr) / (g - 1.0) <= ur - ul:
            raise ValueError("the data generate vacuum")

        def side(p, rho, pk, ck):
            if p > pk:  # shock branch
                a, b = 2.0 / ((g + 1.0) * rho), (g - 1.0) / (g + 1.0) * pk
                root = np.sqrt(a / (p + b))
                return (p - pk) * root, root * (1.0 - 0.5 * (p - pk) / (p + b))
            ratio = p / pk  # rarefaction branch
            return (2.0 * ck / (g - 1.0) * (ratio ** ((g - 1.0) / (2.0 * g)) - 1.0),
                    ratio ** (-(g + 1.0) / (2.0 * g)) / (rho * ck))

        # Two-rarefaction guess (Toro 4.46), then Newton on f_L + f_R + du = 0.
        z = (g - 1.0) / (2.0 * g)
        p = ((cl + cr - 0.5 * (g - 1.0) * (ur - ul)) / (cl / pl ** z + cr / pr ** z)) ** (1.0 / z)
        for _ in range(200):
            fl, dl = side(p, rl, pl, cl)
            fr, dr = side(p, rr, pr, cr)
            new = max(p - (fl + fr + ur - ul) / (dl + dr), 1e-14 * p)
            done = abs(new - p) <= 1e-15 * (new + p)
            p = new
            if done:
                break
        fl, _ = side(p, rl, pl, cl)
        fr, _ = side(p, rr, pr, cr)
        us = 0.5 * (ul + ur) + 0.5 * (fr - fl)
        object.__setattr__(self, "p_star", p)
        object.__setattr__(self, "u_star", us)
        object.__setattr__(self, "_c", (cl, cr))

    def _speeds(self):
        """(left wave speeds, right wave speeds): one shock speed or the
        rarefaction head and tail speeds on each side, and the contact."""
        g, p, us = self.gamma, self.p_star, self.u_star
        (rl, ul, pl), (rr, ur, pr) = self.left, self.right
        cl, cr = self._c
        if p > pl:
            left = (ul - cl * np.sqrt((g + 1) / (2 * g) * p / pl + (g - 1) / (2 * g)),)
        else:
            left = (ul - cl, us - cl * (p / pl) ** ((g - 1) / (2 * g)))
        if p > pr:
            right = (ur + cr * np.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g)),)
        else:
            right = (us + cr * (p / pr) ** ((g - 1) / (2 * g)), ur + cr)
        return left, right

    def wave_positions(self, t: float) -> dict:
        left, right = self._speeds()
        out = {}
        if len(left) == 1:
            out["left shock"] = self.x0 + left[0] * t
        else:
            out["left head"], out["left tail"] = (self.x0 + s * t for s in left)
        out["contact"] = self.x0 + self.u_star * t
        if len(right) == 1:
            out["right shock"] = self.x0 + right[0] * t
        else:
            out["right tail"], out["right head"] = (self.x0 + s * t for s in right)
        return out

    def sample(self, x, t: float):
        """(rho, u, p) at positions x (any shape) and time t > 0."""
        g, p, us = self.gamma, self.p_star, self.u_star
        (rl, ul, pl), (rr, ur, pr) = self.left, self.right
        cl, cr = self._c
        s = (np.asarray(x, dtype=np.float64) - self.x0) / t
        left, right = self._speeds()
        gm = (g - 1.0) / (g + 1.0)
        rho = np.empty_like(s)
        u = np.empty_like(s)
        pres = np.empty_like(s)
        # left of the contact
        lft = s <= us
        if len(left) == 1:
            rho_star = rl * (p / pl + gm) / (gm * p / pl + 1.0)
            outer = lft & (s <= left[0])
            star = lft & ~outer
            fan = np.zeros_like(lft)
        else:
            rho_star = rl * (p / pl) ** (1.0 / g)
            outer = lft & (s <= left[0])
            star = lft & (s >= left[1])
            fan = lft & ~outer & ~star
        rho[outer], u[outer], pres[outer] = rl, ul, pl
        rho[star], u[star], pres[star] = rho_star, us, p
        c = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * (ul - s[fan]))
        rho[fan] = rl * (c / cl) ** (2.0 / (g - 1.0))
        u[fan] = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * ul + s[fan])
        pres[fan] = pl * (c / cl) ** (2.0 * g / (g - 1.0))
        # right of the contact
        rgt = ~lft
        if len(right) == 1:
            rho_star = rr * (p / pr + gm) / (gm * p / pr + 1.0)
            outer = rgt & (s >= right[0])
            star = rgt & ~outer
            fan = np.zeros_like(rgt)
        else:
            rho_star = rr * (p / pr) ** (1.0 / g)
            outer = rgt & (s >= right[1])
            star = rgt & (s <= right[0])
            fan = rgt & ~outer & ~star
        rho[outer], u[outer], pres[outer] = rr, ur, pr
        rho[star], u[star], pres[star] = rho_star, us, p
        c = 2.0 / (g + 1.0) * (cr - 0.5 * (g - 1.0) * (ur - s[fan]))
        rho[fan] = rr * (c / cr) ** (2.0 / (g - 1.0))
        u[fan] = 2.0 / (g + 1.0) * (-cr + 0.5 * (g - 1.0) * ur + s[fan])
        pres[fan] = pr * (c / cr) ** (2.0 * g / (g - 1.0))
        return rho, u, pres


# ---------------------------------------------------------------------------
# Quadrature
# ---------------------------------------------------------------------------
def collapsed_gauss(n: int = QUAD_POINTS):
    """Reference-triangle rule (vertices (-1,-1), (1,-1), (-1,1)) from the
    Duffy map of an n x n Gauss-Legendre rule: points (r, s) and weights
    normalised to sum to 1."""
    x, w = np.polynomial.legendre.leggauss(n)
    a, b = np.meshgrid(x, x, indexing="ij")
    wa, wb = np.meshgrid(w, w, indexing="ij")
    r = 0.5 * (1.0 + a) * (1.0 - b) - 1.0
    wt = wa * wb * 0.25 * (1.0 - b)
    return r.ravel(), b.ravel(), wt.ravel() / wt.sum()


def _to_physical(r, s, tri):
    """Reference (r, s) (n,) -> physical points (M, n, 2) of triangles (M, 3, 2)."""
    lam = np.stack([-(r + s) / 2.0, (1.0 + r) / 2.0, (1.0 + s) / 2.0], -1)  # (n, 3)
    return np.einsum("nv,mvd->mnd", lam, tri)


def _area(tri):
    e1, e2 = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]
    return 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])


def _clip(poly, line, keep_below):
    """Convex polygon (k, 2) clipped to a x + b y <= c (keep_below) or >= c,
    for line = (a, b, c)."""
    a, b, c = line
    d = c - (a * poly[:, 0] + b * poly[:, 1])
    d = d if keep_below else -d
    inside = d >= 0.0
    if inside.all():
        return poly
    if not inside.any():
        return poly[:0]
    out = []
    k = len(poly)
    for i in range(k):
        j = (i + 1) % k
        if inside[i]:
            out.append(poly[i])
        if inside[i] != inside[j]:
            out.append(poly[i] + d[i] / (d[i] - d[j]) * (poly[j] - poly[i]))
    return np.array(out)


@dataclass(frozen=True, eq=False)
class SplitQuadrature:
    """Quadrature points covering a mesh: x, y, weight (M, n), the parent
    element of each sub-triangle (M,) and the reference coordinates of every
    point in its parent (M, n, 2)."""

    x: np.ndarray
    y: np.ndarray
    weight: np.ndarray
    parent: np.ndarray
    ref: np.ndarray

    @property
    def area(self) -> float:
        return float(self.weight.sum())


def split_quadrature(corners, breakpoints=(), n: int = QUAD_POINTS, lines=()) -> SplitQuadrature:
    """``corners`` (K, 3, 2) element vertices; elements that a vertical
    breakpoint x = b, or a line a x + b y = c of ``lines``, crosses are cut
    along it and the pieces fan-triangulated."""
    corners = np.asarray(corners, dtype=np.float64)
    cuts = [(1.0, 0.0, float(b)) for b in np.sort(np.asarray(breakpoints, dtype=np.float64).ravel())]
    cuts += [tuple(float(v) for v in ln) for ln in lines]
    size = np.maximum(np.ptp(corners[..., 0], 1), np.ptp(corners[..., 1], 1))
    crossed = np.zeros((len(corners), len(cuts)), bool)
    for k, (a, b, c) in enumerate(cuts):
        d = (a * corners[..., 0] + b * corners[..., 1] - c) / np.hypot(a, b)
        slack = SPLIT_TOL * np.maximum(size, SPLIT_TOL)
        crossed[:, k] = (d.min(1) < -slack) & (d.max(1) > slack)
    cut = np.flatnonzero(crossed.any(1))
    keep = np.setdiff1d(np.arange(len(corners)), cut)
    tris, parents = [corners[keep]], [keep]
    for e in cut:
        pieces = [corners[e]]
        for k in np.flatnonzero(crossed[e]):
            pieces = [q for p in pieces for q in (_clip(p, cuts[k], True), _clip(p, cuts[k], False)) if len(q) >= 3]
        for piece in pieces:
            fan = np.stack([np.repeat(piece[:1], len(piece) - 2, 0), piece[1:-1], piece[2:]], 1)
            tris.append(fan)
            parents.append(np.full(len(fan), e))
    tri, parent = np.concatenate(tris), np.concatenate(parents)
    r, s, w = collapsed_gauss(n)
    pts = _to_physical(r, s, tri)
    weight = _area(tri)[:, None] * w[None, :]
    # Reference coordinates in the parent: invert its affine map.
    pc = corners[parent]
    j = np.stack([pc[:, 1] - pc[:, 0], pc[:, 2] - pc[:, 0]], -1)  # (M, 2, 2) columns
    d = pts - pc[:, None, 0]  # (M, n, 2)
    det = (j[:, 0, 0] * j[:, 1, 1] - j[:, 0, 1] * j[:, 1, 0])[:, None]
    lam = np.stack([(j[:, None, 1, 1] * d[..., 0] - j[:, None, 0, 1] * d[..., 1]) / det,
                    (j[:, None, 0, 0] * d[..., 1] - j[:, None, 1, 0] * d[..., 0]) / det], -1)  # barycentric 1, 2
    return SplitQuadrature(pts[..., 0], pts[..., 1], weight, parent, 2.0 * lam - 1.0)


def monomial_powers(degree: int = 4):
    return np.array([(i, j) for i in range(degree + 1) for j in range(degree + 1 - i)])


def _monomials(r, s, pw):
    """r^i s^j for every (i, j) row of pw, by repeated products (shape r.shape + (len(pw),))."""
    deg = int(pw.max())
    rp, sp = [np.ones_like(r)], [np.ones_like(s)]
    for _ in range(deg):
        rp.append(rp[-1] * r)
        sp.append(sp[-1] * s)
    rp, sp = np.stack(rp, -1), np.stack(sp, -1)
    return rp[..., pw[:, 0]] * sp[..., pw[:, 1]]


def evaluate_nodal(state, r_nodes, s_nodes, quad: SplitQuadrature, degree: int = 4):
    """The nodal solution (Np, K, nv) at the quadrature points (M, n, nv)."""
    pw = monomial_powers(degree)
    vand = _monomials(r_nodes, s_nodes, pw)
    coef = np.linalg.solve(vand, state.reshape(state.shape[0], -1)).reshape(state.shape)  # (Np, K, nv)
    basis = _monomials(quad.ref[..., 0], quad.ref[..., 1], pw)  # (M, n, Np)
    return np.matmul(basis, coef.transpose(1, 0, 2)[quad.parent])  # (M, n, nv)


def primitive(q, gamma):
    rho = q[..., 0]
    u, v = q[..., 1] / rho, q[..., 2] / rho
    return rho, u, v, (gamma - 1.0) * (q[..., 3] - 0.5 * (q[..., 1] * u + q[..., 2] * v))


def l1_errors(values: dict, exact: dict, quad: SplitQuadrature) -> dict:
    """Domain-averaged L1 error of each named field."""
    return {k: float(np.sum(quad.weight * np.abs(values[k] - exact[k])) / quad.area) for k in values}


def riemann_l1(solver, state, solution: RiemannSolution, t: float, n: int = QUAD_POINTS) -> dict:
    """L1 errors of density, pressure and both velocities against the exact
    Riemann solution (uniform in y, v = 0), split at its wave edges."""
    m = solver.mesh
    quad = split_quadrature(m.vertices[m.elements], list(solution.wave_positions(t).values()), n)
    num = evaluate_nodal(state, solver.reference.r, solver.reference.s, quad)
    rho, u, v, p = primitive(num, solver.config.gamma)
    er, eu, ep = solution.sample(quad.x, t)
    return l1_errors({"density": rho, "pressure": p, "x velocity": u, "y velocity": v},
                     {"density": er, "pressure": ep, "x velocity": eu, "y velocity": np.zeros_like(eu)}, quad)


def column_density(solver, state, nx: int, n: int = QUAD_POINTS):
    """Density averaged over each of the nx mesh columns of a split-rectangle
    mesh (elements 2(j nx + i) and 2(j nx + i) + 1 lie in column i), and the x
    of the faces between neighbouring columns."""
    m = solver.mesh
    quad = split_quadrature(m.vertices[m.elements], (), n)
    rho = evaluate_nodal(state[..., :1], solver.reference.r, solver.reference.s, quad)[..., 0]
    K = m.num_elements
    integral = np.bincount(quad.parent, weights=(quad.weight * rho).sum(1), minlength=K)
    area = np.bincount(quad.parent, weights=quad.weight.sum(1), minlength=K)
    col = (np.arange(K) // 2) % nx
    avg = np.bincount(col, weights=integral, minlength=nx) / np.bincount(col, weights=area, minlength=nx)
    x0, x1 = m.vertices[:, 0].min(), m.vertices[:, 0].max()
    return avg, x0 + (x1 - x0) * np.arange(1, nx) / nx


def shock_location(avg, face_x, window):
    """The column face inside ``window`` with the largest density drop."""
    inside = (face_x >= window[0]) & (face_x <= window[1])
    if not inside.any():
        raise ValueError(f"no column face lies in {window}")
    drop = np.where(inside, avg[:-1] - avg[1:], -np.inf)
    return float(face_x[int(np.argmax(drop))])


def evaluate_at_points(solver, state, x, y, degree: int = 4):
    """The nodal solution (Np, K, nv) at physical points (n,): each point's
    containing element by its barycentric coordinates (the element with the
    largest smallest coordinate, so points on faces pick one side)."""
    m = solver.mesh
    c = np.asarray(m.vertices)[np.asarray(m.elements)]  # (K, 3, 2)
    j = np.stack([c[:, 1] - c[:, 0], c[:, 2] - c[:, 0]], -1)  # (K, 2, 2)
    inv = np.linalg.inv(j)
    pts = np.stack([np.asarray(x, float), np.asarray(y, float)], -1)  # (n, 2)
    lam12 = np.einsum("kab,nkb->nka", inv, pts[:, None, :] - c[None, :, 0, :])  # (n, K, 2)
    lam = np.concatenate([1.0 - lam12.sum(-1, keepdims=True), lam12], -1)
    e = np.argmax(lam.min(-1), axis=1)
    l12 = lam12[np.arange(len(pts)), e]
    r, s = 2.0 * l12[:, 0] - 1.0, 2.0 * l12[:, 1] - 1.0
    pw = monomial_powers(degree)
    rn, sn = solver.reference.r, solver.reference.s
    vand = rn[:, None] ** pw[None, :, 0] * sn[:, None] ** pw[None, :, 1]
    coef = np.linalg.solve(vand, state.reshape(state.shape[0], -1)).reshape(state.shape)
    basis = r[:, None] ** pw[:, 0] * s[:, None] ** pw[:, 1]  # (n, Np)
    return np.einsum("nb,bnv->nv", basis, coef[:, e, :])
