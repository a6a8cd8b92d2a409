"""Case setups: the reference's Sod tube plus the scale-up workloads.

``sod_shock_tube`` keeps ``src/cfd_demo/cases.py:15-44`` (constants, mesh,
fixed-state ends, slip walls, per-element centroid initialisation). The other
constructors build the four-quadrant Riemann problem, the isentropic vortex,
the seeded random shock tubes, the vortex-plus-bumps fields and the steady
oblique shock over a compression corner that the reference proposes in
docs/adr/0011-oblique-shock-verification.md (on the sheared mesh and
far-field condition of its ADR 0012).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .boundary import ExactState, FixedState, Periodic, SlipWall, Transmissive
from .mesh import rectangle_mesh, sheared_rectangle_mesh
from .physics import primitive_to_conserved
from .solver import DEFAULT_CFL, EulerSolver, SolverConfig

SOD_GAMMA = 1.4
SOD_LEFT = (1.0, 0.0, 0.0, 1.0)
SOD_RIGHT = (0.125, 0.0, 0.0, 0.1)
SOD_DISCONTINUITY = 0.5
SOD_X_RANGE = (0.0, 1.0)
SOD_Y_RANGE = (0.0, 0.1)
SOD_FINAL_TIME = 0.2


@dataclass(frozen=True, eq=False)
class SodShockTube:
    solver: EulerSolver
    initial_state: np.ndarray
    final_time: float


@dataclass(frozen=True, eq=False)
class Case:
    solver: EulerSolver
    initial_state: np.ndarray
    final_time: float


def sod_shock_tube(nx: int = 200, ny: int = 20, *, flux: str, limiter: str, cfl: float = DEFAULT_CFL,
                   transmissive_ends: bool = False) -> SodShockTube:
    """ADR-0005 Sod tube on nx*ny split rectangles: slip walls top and bottom,
    the initial states as fixed exterior states at the ends (or transmissive
    ends with ``transmissive_ends=True``)."""
    mesh = rectangle_mesh(nx, ny, SOD_X_RANGE, SOD_Y_RANGE)
    left = FixedState.from_primitive(*SOD_LEFT, gamma=SOD_GAMMA)
    right = FixedState.from_primitive(*SOD_RIGHT, gamma=SOD_GAMMA)
    ends = (Transmissive(), Transmissive()) if transmissive_ends else (left, right)
    bcs = {"left": ends[0], "right": ends[1], "bottom": SlipWall(), "top": SlipWall()}
    solver = EulerSolver(mesh, bcs, SolverConfig(flux=flux, limiter=limiter, gamma=SOD_GAMMA, cfl=cfl))
    on_left = solver.geometry.centroid[:, 0] < SOD_DISCONTINUITY
    state = np.where(on_left[:, None], np.array(left.state), np.array(right.state))
    return SodShockTube(solver=solver, initial_state=np.broadcast_to(state[None], solver.state_shape).copy(),
                        final_time=SOD_FINAL_TIME)


# Four-quadrant Riemann problem (rho, u, v, p) around (0.5, 0.5).
QUADRANT_STATES = {
    "NE": (1.5, 0.0, 0.0, 1.5),
    "NW": (0.5323, 1.206, 0.0, 0.3),
    "SW": (0.138, 1.206, 1.206, 0.029),
    "SE": (0.5323, 0.0, 1.206, 0.3),
}


def four_quadrant(n: int = 1000, *, flux: str = "hllc", limiter: str = "mc", cfl: float = DEFAULT_CFL,
                  gamma: float = 1.4, final_time: float = 0.3) -> Case:
    """[0,1]^2 on an n x n split-rectangle mesh, transmissive sides, constant
    states per quadrant assigned by element centroid."""
    mesh = rectangle_mesh(n, n)
    bcs = {tag: Transmissive() for tag in ("left", "right", "bottom", "top")}
    solver = EulerSolver(mesh, bcs, SolverConfig(flux=flux, limiter=limiter, gamma=gamma, cfl=cfl))
    cx, cy = solver.geometry.centroid[:, 0], solver.geometry.centroid[:, 1]
    east, north = cx > 0.5, cy > 0.5
    state = np.empty((mesh.num_elements, 4))
    for key, mask in (("NE", east & north), ("NW", ~east & north), ("SW", ~east & ~north), ("SE", east & ~north)):
        state[mask] = primitive_to_conserved(*QUADRANT_STATES[key], gamma)
    return Case(solver, np.broadcast_to(state[None], solver.state_shape).copy(), final_time)


def periodic_square(n: int, length: float):
    """Periodic [0, length]^2 split-rectangle mesh and boundary conditions."""
    mesh = rectangle_mesh(n, n, (0.0, length), (0.0, length))
    bcs = {"left": Periodic("right", (-length, 0.0)), "right": Periodic("left", (length, 0.0)),
           "bottom": Periodic("top", (0.0, -length)), "top": Periodic("bottom", (0.0, length))}
    return mesh, bcs


def isentropic_vortex_state(x, y, t=0.0, *, gamma=1.4, beta=5.0, center=(5.0, 5.0), velocity=(1.0, 1.0),
                            period=10.0):
    """Shu's isentropic vortex advected to time t in a periodic box: (..., 4)."""
    x0 = center[0] + velocity[0] * t
    y0 = center[1] + velocity[1] * t
    dx = (x - x0 + 0.5 * period) % period - 0.5 * period
    dy = (y - y0 + 0.5 * period) % period - 0.5 * period
    r2 = dx * dx + dy * dy
    e = np.exp(0.5 * (1.0 - r2))
    du = -beta / (2.0 * np.pi) * dy * e
    dv = beta / (2.0 * np.pi) * dx * e
    temp = 1.0 - (gamma - 1.0) * beta ** 2 / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = temp ** (1.0 / (gamma - 1.0))
    return primitive_to_conserved(rho, velocity[0] + du, velocity[1] + dv, rho ** gamma, gamma)


def isentropic_vortex(n: int = 512, *, flux: str = "roe", limiter: str = "unlimited", cfl: float = DEFAULT_CFL,
                      gamma: float = 1.4, final_time: float = 2.0) -> Case:
    """Periodic [0,10]^2, n x n split rectangles, nodal initialisation."""
    mesh, bcs = periodic_square(n, 10.0)
    solver = EulerSolver(mesh, bcs, SolverConfig(flux=flux, limiter=limiter, gamma=gamma, cfl=cfl))
    x, y = solver.node_coordinates()
    return Case(solver, isentropic_vortex_state(x, y, gamma=gamma), final_time)


def vortex_bumps_state(x, y, *, n_bumps: int = 8, seed: int = 0, amplitude: float = 0.05, width: float = 0.5,
                       gamma: float = 1.4):
    """The isentropic vortex plus n_bumps seeded density bumps
    rho += amplitude exp(-|x - x_k|^2 / width^2) (periodic distance on
    [0,10]^2; centres uniform in [1, 9]^2 from default_rng(seed)), velocity
    and pressure of the vortex unchanged: (..., 4)."""
    q = isentropic_vortex_state(x, y, gamma=gamma)
    rho = q[..., 0]
    u, v = q[..., 1] / rho, q[..., 2] / rho
    p = rho ** gamma
    centres = np.random.default_rng(seed).uniform(1.0, 9.0, size=(n_bumps, 2))
    for cx, cy in centres:
        dx = (x - cx + 5.0) % 10.0 - 5.0
        dy = (y - cy + 5.0) % 10.0 - 5.0
        rho = rho + amplitude * np.exp(-(dx * dx + dy * dy) / (width * width))
    return primitive_to_conserved(rho, u, v, p, gamma)


def vortex_bumps(n: int = 200, *, n_bumps: int = 8, seed: int = 0, amplitude: float = 0.05, width: float = 0.5,
                 flux: str = "rusanov", limiter: str = "unlimited", cfl: float = 0.1, gamma: float = 1.4,
                 final_time: float = 0.0) -> Case:
    """Periodic [0,10]^2 on n x n split rectangles with
    :func:`vortex_bumps_state`, nodal initialisation."""
    mesh, bcs = periodic_square(n, 10.0)
    solver = EulerSolver(mesh, bcs, SolverConfig(flux=flux, limiter=limiter, gamma=gamma, cfl=cfl))
    x, y = solver.node_coordinates()
    state = vortex_bumps_state(x, y, n_bumps=n_bumps, seed=seed, amplitude=amplitude, width=width, gamma=gamma)
    return Case(solver, state, final_time)


# ---------------------------------------------------------------------------
# Steady oblique shock on an inclined wall (ADR 0011, proposed in the reference)
# ---------------------------------------------------------------------------
def shock_deflection(mach: float, beta, gamma: float = 1.4):
    """Flow deflection of an oblique shock at angle beta (the theta-beta-M
    relation): tan(theta) = 2 cot(beta) (M^2 sin^2 beta - 1) /
    (M^2 (gamma + cos 2 beta) + 2)."""
    m2 = mach * mach
    return np.arctan(2.0 / np.tan(beta) * (m2 * np.sin(beta) ** 2 - 1.0) / (m2 * (gamma + np.cos(2.0 * beta)) + 2.0))


def max_deflection(mach: float, gamma: float = 1.4) -> tuple[float, float]:
    """(theta_max, beta at theta_max): the largest deflection an attached
    shock turns a Mach ``mach`` flow by (golden-section search on beta)."""
    if not mach > 1.0:
        raise ValueError("The inflow Mach number must exceed 1.")
    lo, hi = np.arcsin(1.0 / mach), 0.5 * np.pi
    g = 0.5 * (np.sqrt(5.0) - 1.0)
    for _ in range(200):
        a, b = hi - g * (hi - lo), lo + g * (hi - lo)
        if shock_deflection(mach, a, gamma) < shock_deflection(mach, b, gamma):
            lo = a
        else:
            hi = b
    beta = 0.5 * (lo + hi)
    return float(shock_deflection(mach, beta, gamma)), float(beta)


def weak_shock_angle(mach: float, theta: float, gamma: float = 1.4) -> float:
    """The weak-shock angle beta for a wall deflection 0 < theta <
    theta_max(mach) (bisection on the increasing branch beta in (mu, beta_max))."""
    theta_max, beta_max = max_deflection(mach, gamma)
    if not 0.0 < theta < theta_max:
        raise ValueError(f"The wall angle must lie in (0, {theta_max:.6f}) rad for an attached shock at Mach {mach}.")
    lo, hi = np.arcsin(1.0 / mach), beta_max
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if shock_deflection(mach, mid, gamma) < theta:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


@dataclass(frozen=True, eq=False)
class ObliqueShock:
    """The exact steady solution: uniform upstream state (1, M, 0, 1/gamma)
    (sound speed 1) above the shock line y = x tan(beta) from the corner at
    the origin, and the uniform post-shock state below it, moving parallel to
    the wall (normal-shock relations on M sin(beta))."""

    mach: float
    theta: float
    gamma: float
    beta: float
    upstream: tuple
    post: tuple

    def state(self, x, y):
        """Conserved exact state (..., 4) at physical points."""
        below = np.asarray(y) <= np.asarray(x) * np.tan(self.beta)
        up = primitive_to_conserved(*self.upstream, self.gamma)
        dn = primitive_to_conserved(*self.post, self.gamma)
        return np.where(below[..., None], dn, up)

    def post_fraction(self, corners) -> np.ndarray:
        """Area fraction of each triangle (n, 3, 2) below the shock line."""
        t = np.tan(self.beta)
        out = np.empty(len(corners))
        for i, tri in enumerate(np.asarray(corners, dtype=np.float64)):
            d = tri[:, 1] - tri[:, 0] * t  # <= 0 below the line
            keep = []
            for a in range(3):
                b = (a + 1) % 3
                if d[a] <= 0.0:
                    keep.append(tri[a])
                if (d[a] <= 0.0) != (d[b] <= 0.0):
                    keep.append(tri[a] + d[a] / (d[a] - d[b]) * (tri[b] - tri[a]))
            area = 0.0
            if len(keep) >= 3:
                p = np.array(keep)
                area = 0.5 * abs(np.dot(p[:, 0], np.roll(p[:, 1], -1)) - np.dot(p[:, 1], np.roll(p[:, 0], -1)))
            full = 0.5 * abs((tri[1, 0] - tri[0, 0]) * (tri[2, 1] - tri[0, 1])
                             - (tri[2, 0] - tri[0, 0]) * (tri[1, 1] - tri[0, 1]))
            out[i] = min(1.0, area / full)
        return out

    def average(self, corners) -> np.ndarray:
        """Exact cell averages (n, 4) of the piecewise-constant solution."""
        f = self.post_fraction(corners)[:, None]
        up = primitive_to_conserved(*self.upstream, self.gamma)
        dn = primitive_to_conserved(*self.post, self.gamma)
        return f * np.asarray(dn) + (1.0 - f) * np.asarray(up)


def oblique_shock_solution(mach: float = 2.0, theta: float = np.radians(10.0), gamma: float = 1.4) -> ObliqueShock:
    beta = weak_shock_angle(mach, theta, gamma)
    mn1 = mach * np.sin(beta)
    rho2 = (gamma + 1.0) * mn1 ** 2 / ((gamma - 1.0) * mn1 ** 2 + 2.0)
    p2 = (1.0 + 2.0 * gamma / (gamma + 1.0) * (mn1 ** 2 - 1.0)) / gamma
    un2 = mn1 / rho2  # mass flux through the shock, upstream rho = c = 1
    ut = mach * np.cos(beta)  # tangential velocity is continuous
    u2 = un2 * np.sin(beta) + ut * np.cos(beta)
    v2 = -un2 * np.cos(beta) + ut * np.sin(beta)
    return ObliqueShock(mach, theta, gamma, float(beta), (1.0, mach, 0.0, 1.0 / gamma), (rho2, u2, v2, p2))


def oblique_shock(nx: int = 100, ny: int = 10, *, mach: float = 2.0, theta: float = np.radians(10.0),
                  length: float = 1.0, height: float = 0.5, flux: str = "rusanov", limiter: str = "minmod",
                  cfl: float = DEFAULT_CFL, gamma: float = 1.4, final_time: float = 2.0) -> Case:
    """Supersonic inflow over a compression corner: 0 <= x <= length,
    x tan(theta) <= y <= x tan(theta) + height on the sheared nx x ny mesh;
    FixedState upstream inflow on the left, slip wall below, ExactState far
    field on top and right; upstream state everywhere at t = 0."""
    sol = oblique_shock_solution(mach, theta, gamma)
    mesh = sheared_rectangle_mesh(nx, ny, (0.0, length), (0.0, height), theta)
    far = ExactState(sol.state, sol.average)
    bcs = {"left": FixedState.from_primitive(*sol.upstream, gamma), "bottom": SlipWall(), "top": far, "right": far}
    solver = EulerSolver(mesh, bcs, SolverConfig(flux=flux, limiter=limiter, gamma=gamma, cfl=cfl))
    q0 = np.broadcast_to(primitive_to_conserved(*sol.upstream, gamma), solver.state_shape).copy()
    return Case(solver, q0, final_time)
