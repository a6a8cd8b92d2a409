"""Euler solver driver: the reference's ``EulerSolver`` surface over libcfdhip.so.

Mirrors ``src/cfd_demo/solver.py`` (``SolverError`` 33-34, ``SolverConfig``
37-59, ``RunResult`` 62-88, ``EulerSolver`` 91-216) and ``StepReport``
(``time_integration.py:22-29``). The four jitted callables the reference binds
at ``solver.py:118-125`` are replaced by C-ABI calls into hand-written gfx950
kernels; ``run`` keeps the state in HBM and syncs once per step.
"""

from __future__ import annotations

import ctypes as C
import time
from collections.abc import Mapping
from dataclasses import dataclass
from types import MappingProxyType
from typing import NamedTuple

import numpy as np

from . import _lib
from .mesh import TriangleMesh, node_coordinates
from .physics import NUM_VARIABLES, get_interface_flux, get_limiter_function
from .reference_element import DEGREE, build_reference_triangle
from .tables import assemble

DEFAULT_CFL = 0.5  # time_integration.py:18
MAX_REJECTIONS = 10  # time_integration.py:19


class SolverError(RuntimeError):
    """Raised when the solver cannot advance an admissible solution."""


class StepReport(NamedTuple):
    """Diagnostics of one SSP-RK3 step (``time_integration.py:22-29``)."""

    boundary_flux: np.ndarray
    admissible: np.ndarray
    min_density: np.ndarray
    min_pressure: np.ndarray
    troubled_cells: np.ndarray


@dataclass(frozen=True)
class SolverConfig:
    """``flux`` and ``limiter`` are required names; ``gamma`` > 1."""

    flux: str
    limiter: str
    gamma: float
    cfl: float = DEFAULT_CFL
    max_rejections: int = MAX_REJECTIONS

    def __post_init__(self) -> None:
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
    """Outcome of :meth:`EulerSolver.run` (fields as ``solver.py:62-88``)."""

    state: np.ndarray
    time: float
    steps: int
    rejected_steps: int
    boundary_flux_integral: np.ndarray
    min_density: float
    min_pressure: float
    max_troubled_cells: int
    wall_time: float


def _report(r: _lib.CfdStepReport) -> StepReport:
    return StepReport(
        boundary_flux=np.array(r.boundary_flux[:], dtype=np.float64),
        admissible=np.bool_(r.admissible != 0),
        min_density=np.float64(r.min_density),
        min_pressure=np.float64(r.min_pressure),
        troubled_cells=np.int32(r.troubled_cells),
    )


class LimiterTables:
    """Handle to a solver's limiter tables, for :func:`apply_limiters`."""

    def __init__(self, solver: "EulerSolver"):
        self.solver = solver


class EulerSolver:
    """Degree-four nodal DG solver for the 2-D Euler equations on the GPU."""

    def __init__(self, mesh: TriangleMesh, boundary_conditions: Mapping, config: SolverConfig) -> None:
        self.mesh = mesh
        self.config = config
        self.boundary_conditions = MappingProxyType(dict(boundary_conditions))
        self.reference = build_reference_triangle(DEGREE)
        # Tables, patched connectivity and geometry from the native assembler.
        self.tables, self.connectivity, self.geometry = assemble(mesh, self.boundary_conditions)
        self.dt_geometry = self.tables.lim[:, 6].copy()  # time_step_geometry (time_integration.py:32-34)
        self.flux = get_interface_flux(config.flux)
        self.phi = get_limiter_function(config.limiter)
        self.limiter_operators = LimiterTables(self)
        self._ctx = None
        self._ctx_by_limiter = {}

    # -- native context -------------------------------------------------------
    def _context(self, limiter_id: int | None = None):
        lid = self.phi.id if limiter_id is None else limiter_id
        ctx = self._ctx_by_limiter.get(lid)
        if ctx is not None:
            return ctx
        _lib.ensure_init(self.reference)
        t = self.tables
        self._keep = (t.geo, t.conn, t.lim, t.bc_states)
        bcs = t.bc_states if t.bc_states.size else np.zeros((1, 4))
        mt = _lib.CfdMeshTables(t.num_elements, _lib.dptr(t.geo), _lib.iptr(t.conn), _lib.dptr(t.lim),
                                _lib.dptr(bcs), int(t.bc_states.shape[0]))
        handle = C.c_void_p()
        _lib.check(_lib.load().cfd_ctx_create(C.byref(mt), self.flux.id, lid, float(self.config.gamma),
                                              float(self.config.cfl), int(self.config.max_rejections),
                                              C.byref(handle)))
        self._ctx_by_limiter[lid] = handle
        return handle

    def close(self) -> None:
        for h in self._ctx_by_limiter.values():
            _lib.load().cfd_ctx_destroy(h)
        self._ctx_by_limiter = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference surface ----------------------------------------------------
    @property
    def state_shape(self) -> tuple[int, int, int]:
        return (self.reference.num_nodes, self.mesh.num_elements, NUM_VARIABLES)

    def node_coordinates(self):
        """Physical node coordinates (x, y), each (Np, K)."""
        return node_coordinates(self.mesh, self.reference.r, self.reference.s)

    def _as_state(self, q) -> np.ndarray:
        q = np.ascontiguousarray(np.asarray(q, dtype=np.float64))
        if q.shape != self.state_shape:
            raise ValueError(f"Expected a state of shape {self.state_shape}, got {q.shape}.")
        return q

    def rhs(self, q):
        """Semi-discrete right-hand side (Np, K, 4) and boundary flux (4,)."""
        q = self._as_state(q)
        out = np.empty_like(q)
        bflux = np.empty(4)
        _lib.check(_lib.load().cfd_rhs(self._context(), _lib.dptr(q), _lib.dptr(out), _lib.dptr(bflux)))
        return out, bflux

    def limit(self, q) -> np.ndarray:
        """Apply troubled-cell and positivity limiting once."""
        return self.limit_with_flags(q)[0]

    def limit_with_flags(self, q, limiter_id: int | None = None):
        """(limited state, averages (K,4), troubled count, troubled flags (K,) bool)."""
        q = self._as_state(q)
        out = np.empty_like(q)
        avg = np.empty((q.shape[1], 4))
        cnt = C.c_int64(0)
        flags = np.empty(q.shape[1], dtype=np.uint8)
        _lib.check(_lib.load().cfd_limit(self._context(limiter_id), _lib.dptr(q), _lib.dptr(out), _lib.dptr(avg),
                                         C.byref(cnt), flags.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out, avg, int(cnt.value), flags.astype(bool)

    def time_step(self, q) -> float:
        """CFL time step of the state ``q``."""
        q = self._as_state(q)
        dt = C.c_double(0.0)
        _lib.check(_lib.load().cfd_time_step(self._context(), _lib.dptr(q), C.byref(dt)))
        return float(dt.value)

    def step(self, q, dt: float):
        """One SSP-RK3 step without rejection handling."""
        return self.step_with_flags(q, dt)[:2]

    def step_with_flags(self, q, dt: float):
        """(new state, StepReport, troubled flags (3, K) bool)."""
        q = self._as_state(q)
        out = np.empty_like(q)
        rep = _lib.CfdStepReport()
        flags = np.empty((3, q.shape[1]), dtype=np.uint8)
        _lib.check(_lib.load().cfd_step(self._context(), _lib.dptr(q), float(dt), _lib.dptr(out), C.byref(rep),
                                        flags.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out, _report(rep), flags.astype(bool)

    def cell_averages(self, q) -> np.ndarray:
        q = self._as_state(q)
        avg = np.empty((q.shape[1], 4))
        _lib.check(_lib.load().cfd_cell_averages(self._context(), _lib.dptr(q), _lib.dptr(avg)))
        return avg

    def integrate(self, q) -> np.ndarray:
        """Domain integral of each conserved variable."""
        q = self._as_state(q)
        out = np.empty(4)
        _lib.check(_lib.load().cfd_integrate(self._context(), _lib.dptr(q), _lib.dptr(out)))
        return out

    def run(self, initial_state, final_time: float, max_steps: int | None = None) -> RunResult:
        """Advance ``initial_state`` from t = 0 to ``final_time`` (``solver.py:167-216``)."""
        start = time.perf_counter()
        q = self._as_state(initial_state)
        out = np.empty_like(q)
        res = _lib.CfdRunResult()
        code = _lib.load().cfd_run(self._context(), _lib.dptr(q), float(final_time),
                                   -1 if max_steps is None else int(max_steps), _lib.dptr(out), C.byref(res))
        if code == _lib.CFD_EINITIAL:
            raise SolverError("The initial state must be finite with positive nodal density and pressure.")
        if code == _lib.CFD_EINVALID_DT:
            raise SolverError(f"Invalid time step {res.fail_dt} at t = {res.fail_time}.")
        if code == _lib.CFD_EREJECTED:
            raise SolverError(f"Step at t = {res.fail_time} was rejected {self.config.max_rejections + 1} times.")
        _lib.check(code)
        return RunResult(
            state=out, time=float(res.time), steps=int(res.steps), rejected_steps=int(res.rejected_steps),
            boundary_flux_integral=np.array(res.boundary_flux_integral[:]), min_density=float(res.min_density),
            min_pressure=float(res.min_pressure), max_troubled_cells=int(res.max_troubled_cells),
            wall_time=time.perf_counter() - start,
        )


def apply_limiters(q, operators: LimiterTables, phi, gamma: float, limit_slopes: bool):
    """``apply_limiters`` (``limiting.py:203-216``) over a solver's tables:
    returns (q, averages, troubled_count). ``phi`` must be a registry limiter
    when ``limit_slopes`` is true; with ``limit_slopes`` false it is unused."""
    solver = operators.solver
    if float(gamma) != float(solver.config.gamma):
        raise ValueError("gamma differs from the solver's configuration.")
    if limit_slopes:
        if not hasattr(phi, "id"):
            raise ValueError("phi must be one of LIMITER_FUNCTIONS on the GPU path.")
        lid = phi.id
    else:
        lid = 4
    out, avg, cnt, _ = solver.limit_with_flags(q, lid)
    return out, avg, np.int32(cnt) if not limit_slopes else np.int64(cnt)
