"""SSP-RK3, CFL step, solver driver and Sod case (oracle restatement).

Restates ``src/cfd_demo/time_integration.py``, ``solver.py`` and ``cases.py``.
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from typing import NamedTuple

import numpy as np

from .dg import build_dg_operators, cell_averages, dg_rhs, domain_integral, exact_boundary_states, \
    resolve_boundary_conditions
from .limiting import apply_limiters, build_limiter_operators
from .mesh import EXACT_STATE, FixedState, SlipWall, Transmissive, build_connectivity, element_geometry, \
    node_coordinates, \
    periodic_connectivity, rectangle_mesh
from .physics import get_interface_flux, get_limiter_function, pressure
from .reference_element import DEGREE, build_reference_triangle

DEFAULT_CFL = 0.5  # time_integration.py:18
MAX_REJECTIONS = 10  # time_integration.py:19


class SolverError(RuntimeError):
    """``solver.py:33-34``."""


class StepReport(NamedTuple):
    """``time_integration.py:22-29``."""

    boundary_flux: np.ndarray
    admissible: np.ndarray
    min_density: np.ndarray
    min_pressure: np.ndarray
    troubled_cells: np.ndarray


def time_step_geometry(geo, degree):
    """``time_integration.py:32-34``."""
    return np.min(2.0 * geo.area[:, None] / ((degree + 1) ** 2 * geo.edge_length), axis=1)


def stable_time_step(q, dtg, gamma, cfl):
    """``time_integration.py:37-41``."""
    rho = q[..., 0]
    u = q[..., 1] / rho
    v = q[..., 2] / rho
    p = (gamma - 1.0) * (q[..., 3] - 0.5 * (q[..., 1] * u + q[..., 2] * v))
    ws = np.sqrt(u * u + v * v) + np.sqrt(gamma * p / rho)
    return cfl * np.min(dtg / np.max(ws, axis=0))


def ssp_rk3_step(q, dt, ops, lops, flux, phi, gamma, limit_slopes, flags_out=None):
    """Shu-Osher SSP-RK3 with limiting after every stage (``time_integration.py:44-84``).

    ``flags_out``, if a list, receives the three stages' troubled flags.
    """

    def limit(u):
        with np.errstate(invalid="ignore", over="ignore"):
            u, avg, tr, fl = apply_limiters(u, lops, phi, gamma, limit_slopes, return_flags=True)
            ok = bool(np.all(np.isfinite(u)) & np.all(avg[:, 0] > 0.0) & np.all(pressure(avg, gamma) > 0.0))
        if flags_out is not None:
            flags_out.append(fl)
        return u, ok, tr

    with np.errstate(invalid="ignore", over="ignore", divide="ignore"):
        r0, b0 = dg_rhs(q, ops, flux, gamma)
        q1, ok1, t1 = limit(q + dt * r0)
        r1, b1 = dg_rhs(q1, ops, flux, gamma)
        q2, ok2, t2 = limit(0.75 * q + 0.25 * (q1 + dt * r1))
        r2, b2 = dg_rhs(q2, ops, flux, gamma)
        q3, ok3, t3 = limit(q / 3.0 + 2.0 / 3.0 * (q2 + dt * r2))
        stages = (q1, q2, q3)
        rep = StepReport(
            boundary_flux=dt * (b0 / 6.0 + b1 / 6.0 + 2.0 * b2 / 3.0),
            admissible=np.bool_(ok1 & ok2 & ok3),
            min_density=np.min(np.stack([np.min(s[..., 0]) for s in stages])),
            min_pressure=np.min(np.stack([np.min(pressure(s, gamma)) for s in stages])),
            troubled_cells=np.max(np.stack([t1, t2, t3])),
        )
    return q3, rep


@dataclass(frozen=True)
class SolverConfig:
    """``solver.py:37-59``."""

    flux: str
    limiter: str
    gamma: float
    cfl: float = DEFAULT_CFL
    max_rejections: int = MAX_REJECTIONS

    def __post_init__(self):
        get_interface_flux(self.flux)
        get_limiter_function(self.limiter)
        if not self.gamma > 1.0:
            raise ValueError("gamma must be greater than one.")
        if not self.cfl > 0.0:
            raise ValueError("cfl must be positive.")
        if self.max_rejections < 0:
            raise ValueError("max_rejections must be nonnegative.")


@dataclass(frozen=True, eq=False)
class RunResult:
    """``solver.py:62-88``."""

    state: np.ndarray
    time: float
    steps: int
    rejected_steps: int
    boundary_flux_integral: np.ndarray
    min_density: float
    min_pressure: float
    max_troubled_cells: int
    wall_time: float


class EulerSolver:
    """``solver.py:91-216``, with the periodic-patch extension."""

    def __init__(self, mesh, boundary_conditions, config):
        self.mesh = mesh
        self.config = config
        self.boundary_conditions = dict(boundary_conditions)
        self.reference = build_reference_triangle(DEGREE)
        conn = build_connectivity(mesh)
        self.connectivity, self.periodic_shift = periodic_connectivity(mesh, conn, self.boundary_conditions)
        shift = self.periodic_shift if np.any(self.periodic_shift) else None
        self.geometry = element_geometry(mesh)
        kind, state = resolve_boundary_conditions(self.connectivity, self.boundary_conditions)
        node_state, ghost = state, state
        if np.any(kind == EXACT_STATE):
            node_state, ghost = exact_boundary_states(mesh, self.connectivity, self.reference, kind, state,
                                                      self.boundary_conditions)
        self.operators = build_dg_operators(mesh, self.connectivity, self.geometry, self.reference, kind, node_state,
                                            shift)
        self.limiter_operators = build_limiter_operators(self.connectivity, self.geometry, self.reference, kind, ghost,
                                                         shift)
        self.dt_geometry = time_step_geometry(self.geometry, self.reference.degree)
        self.flux = get_interface_flux(config.flux)
        self.phi = get_limiter_function(config.limiter)
        self.limit_slopes = config.limiter != "unlimited"

    @property
    def state_shape(self):
        return (self.reference.num_nodes, self.mesh.num_elements, 4)

    def node_coordinates(self):
        return node_coordinates(self.mesh, self.reference.r, self.reference.s)

    def _as_state(self, q):
        q = np.asarray(q, dtype=np.float64)
        if q.shape != self.state_shape:
            raise ValueError(f"Expected a state of shape {self.state_shape}, got {q.shape}.")
        return q

    def rhs(self, q):
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            return dg_rhs(self._as_state(q), self.operators, self.flux, self.config.gamma)

    def limit(self, q, return_flags=False):
        out = apply_limiters(self._as_state(q), self.limiter_operators, self.phi, self.config.gamma,
                             self.limit_slopes, return_flags=True)
        return (out[0], out[3]) if return_flags else out[0]

    def time_step(self, q):
        return float(stable_time_step(self._as_state(q), self.dt_geometry, self.config.gamma, self.config.cfl))

    def step(self, q, dt, flags_out=None):
        return ssp_rk3_step(self._as_state(q), np.float64(dt), self.operators, self.limiter_operators, self.flux,
                            self.phi, self.config.gamma, self.limit_slopes, flags_out)

    def cell_averages(self, q):
        return cell_averages(self._as_state(q), self.operators)

    def integrate(self, q):
        return domain_integral(self._as_state(q), self.operators)

    def run(self, q0, final_time, max_steps=None):
        """``solver.py:167-216`` (``max_steps`` stops early; not in the reference)."""
        t0 = time.perf_counter()
        q = self._as_state(q0)
        rho = q[..., 0]
        with np.errstate(invalid="ignore", divide="ignore"):
            p0 = pressure(q, self.config.gamma)
        if not (np.all(np.isfinite(q)) and np.all(rho > 0.0) and np.all(p0 > 0.0)):
            raise SolverError("The initial state must be finite with positive nodal density and pressure.")
        mind, minp = float(np.min(rho)), float(np.min(p0))
        fint = np.zeros(4)
        t, steps, rej, maxtr = 0.0, 0, 0, 0
        while t < final_time:
            if max_steps is not None and steps >= max_steps:
                break
            dt = self.time_step(q)
            if not np.isfinite(dt) or dt <= 0.0:
                raise SolverError(f"Invalid time step {dt} at t = {t}.")
            remaining = final_time - t
            reaches_end = dt >= remaining
            dt = remaining if reaches_end else dt
            for _ in range(self.config.max_rejections + 1):
                qn, rep = self.step(q, dt)
                if bool(rep.admissible):
                    break
                rej += 1
                reaches_end = False
                dt *= 0.5
            else:
                raise SolverError(f"Step at t = {t} was rejected {self.config.max_rejections + 1} times.")
            q = qn
            t = final_time if reaches_end else t + dt
            steps += 1
            fint += np.asarray(rep.boundary_flux)
            mind = min(mind, float(rep.min_density))
            minp = min(minp, float(rep.min_pressure))
            maxtr = max(maxtr, int(rep.troubled_cells))
        return RunResult(np.asarray(q), t, steps, rej, fint, mind, minp, maxtr, time.perf_counter() - t0)


# Sod shock tube (cases.py:15-44).
SOD_GAMMA = 1.4
SOD_LEFT = (1.0, 0.0, 0.0, 1.0)
SOD_RIGHT = (0.125, 0.0, 0.0, 0.1)
SOD_DISCONTINUITY = 0.5
SOD_X_RANGE = (0.0, 1.0)
SOD_Y_RANGE = (0.0, 0.1)
SOD_FINAL_TIME = 0.2


def sod_shock_tube(nx=200, ny=20, *, flux, limiter, cfl=DEFAULT_CFL, transmissive_ends=False):
    """``cases.py:31-44``: returns (solver, initial_state, final_time).

    ``transmissive_ends`` (extension) replaces the fixed-state ends."""
    mesh = rectangle_mesh(nx, ny, SOD_X_RANGE, SOD_Y_RANGE)
    left = FixedState.from_primitive(*SOD_LEFT, gamma=SOD_GAMMA)
    right = FixedState.from_primitive(*SOD_RIGHT, gamma=SOD_GAMMA)
    ends = (Transmissive(), Transmissive()) if transmissive_ends else (left, right)
    bcs = {"left": ends[0], "right": ends[1], "bottom": SlipWall(), "top": SlipWall()}
    solver = EulerSolver(mesh, bcs, SolverConfig(flux=flux, limiter=limiter, gamma=SOD_GAMMA, cfl=cfl))
    on_left = solver.geometry.centroid[:, 0] < SOD_DISCONTINUITY
    es = np.where(on_left[:, None], np.array(left.state), np.array(right.state))
    return solver, np.broadcast_to(es[None], solver.state_shape).copy(), SOD_FINAL_TIME
