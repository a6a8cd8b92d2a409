"""Case setups: the reference's Sod tube plus the scale-up workloads.

``sod_shock_tube`` keeps ``src/cfd_demo/cases.py:15-44`` (constants, mesh,
fixed-state ends, slip walls, per-element centroid initialisation). The other
constructors build the four-quadrant Riemann problem, the isentropic vortex,
the seeded random shock tubes and the vortex-plus-bumps fields.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .boundary import FixedState, Periodic, SlipWall, Transmissive
from .mesh import rectangle_mesh
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
