"""Batched solves: many independent cases advanced together on one GPU.

Each case is an ``EulerSolver`` setup (mesh, boundary conditions, flux,
limiter, CFL, final time, initial state). The meshes are concatenated, each
padded to a multiple of 256 elements, and one batched SSP-RK3 step (one kernel
launch per stage for all cases, replayed from a hipGraph) advances every
active case by one of its own steps with the reference's run-loop semantics
(``src/cfd_demo/solver.py:167-216``). Results per case match ``EulerSolver.run``.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .boundary import SlipWall, Transmissive
from .mesh import rectangle_mesh
from .physics import get_interface_flux, get_limiter_function, primitive_to_conserved
from .solver import EulerSolver, RunResult, SolverConfig, SolverError

PAD = 256


class CfdBatchCase(C.Structure):
    _fields_ = [("first", C.c_int64), ("count", C.c_int64), ("padded", C.c_int64), ("flux_id", C.c_int32),
                ("limiter_id", C.c_int32), ("cfl", C.c_double), ("final_time", C.c_double),
                ("max_rejections", C.c_int32)]


class CfdBatchCaseResult(C.Structure):
    _fields_ = [("time", C.c_double), ("steps", C.c_int64), ("rejected_steps", C.c_int64),
                ("boundary_flux_integral", C.c_double * 4), ("min_density", C.c_double),
                ("min_pressure", C.c_double), ("max_troubled_cells", C.c_int64), ("status", C.c_int32),
                ("fail_time", C.c_double), ("fail_dt", C.c_double)]


_SIGS = {
    "cfd_batch_create": (C.c_int, [C.POINTER(_lib.CfdMeshTables), C.c_int, C.POINTER(CfdBatchCase), C.c_double,
                                   C.POINTER(C.c_void_p)]),
    "cfd_batch_destroy": (C.c_int, [C.c_void_p]),
    "cfd_batch_begin": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "cfd_batch_advance": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_int64)]),
    "cfd_batch_download": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "cfd_batch_results": (C.c_int, [C.c_void_p, C.POINTER(CfdBatchCaseResult)]),
    "cfd_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
}
_lib.SIGNATURES.update(_SIGS)


def _bind():
    lib = _lib.load()
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


@dataclass(frozen=True)
class Case:
    solver: EulerSolver
    initial_state: np.ndarray
    final_time: float


class BatchedSolver:
    """Advance a list of :class:`Case` together on the GPU."""

    def __init__(self, cases: list[Case], gamma: float = 1.4):
        if not cases:
            raise ValueError("at least one case is needed")
        self.cases = list(cases)
        self.gamma = gamma
        # Cases are laid out grouped by flux (so each flux's stage launch spans
        # only its own blocks); results come back in the caller's order.
        self._order = sorted(range(len(self.cases)), key=lambda i: self.cases[i].solver.flux.id)
        geo, conn, lim, states, q_parts, specs = [], [], [], [], [], []
        first = 0
        nstates = 0
        for i in self._order:
            case = self.cases[i]
            s = case.solver
            if s.config.gamma != gamma:
                raise ValueError("all cases of a batch share gamma")
            t = s.tables
            K = t.num_elements
            npad = -(-K // PAD) * PAD
            g = np.empty((npad, 16))
            g[:K] = t.geo
            g[K:] = t.geo[0]
            c = np.empty((npad, 4), dtype=np.int32)
            cc = t.conn.astype(np.int64)
            kind = np.stack([(cc[:, 3] >> (8 * f + 3)) & 7 for f in range(3)], axis=1)
            given = (kind == 2) | (kind == 4)  # fixed-state and exact-state faces index bc_states
            nbr = np.where(given, cc[:, :3] + nstates, cc[:, :3] + first)
            c[:K, :3] = nbr
            c[:K, 3] = t.conn[:, 3]
            pad_ids = first + np.arange(K, npad)
            c[K:, 0] = c[K:, 1] = c[K:, 2] = pad_ids
            c[K:, 3] = (0 | (1 << 6)) | ((1 | (2 << 6)) << 8) | ((2 | (0 << 6)) << 16)  # self faces
            lm = np.empty((npad, 8))
            lm[:K] = t.lim
            lm[K:] = t.lim[0]
            lm[K:, 6] = np.inf
            geo.append(g)
            conn.append(c)
            lim.append(lm)
            states.append(t.bc_states)
            nstates += t.bc_states.shape[0]
            q = np.empty((15, npad, 4))
            q[:, :K] = case.initial_state
            q[:, K:] = case.initial_state[:, :1]
            q_parts.append(q)
            specs.append(CfdBatchCase(first, K, npad, s.flux.id, s.phi.id, float(s.config.cfl), float(case.final_time),
                                      int(s.config.max_rejections)))
            first += npad
        self.K = first
        self.first = [0] * len(specs)
        self.count = [0] * len(specs)
        for pos, i in enumerate(self._order):
            self.first[i], self.count[i] = specs[pos].first, specs[pos].count
        self._geo = np.ascontiguousarray(np.concatenate(geo))
        self._conn = np.ascontiguousarray(np.concatenate(conn))
        self._lim = np.ascontiguousarray(np.concatenate(lim))
        bcs = np.concatenate(states) if nstates else np.zeros((1, 4))
        self._bcs = np.ascontiguousarray(bcs)
        self._q0 = np.ascontiguousarray(np.concatenate(q_parts, axis=1))
        self._specs = (CfdBatchCase * len(specs))(*specs)
        _lib.ensure_init(self.cases[0].solver.reference)
        lib = _bind()
        mt = _lib.CfdMeshTables(self.K, _lib.dptr(self._geo), _lib.iptr(self._conn), _lib.dptr(self._lim),
                                _lib.dptr(self._bcs), int(nstates))
        self._h = C.c_void_p()
        _lib.check(lib.cfd_batch_create(C.byref(mt), len(specs), self._specs, float(gamma), C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            _bind().cfd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def begin(self):
        _lib.check(_bind().cfd_batch_begin(self._h, _lib.dptr(self._q0)))

    def autotune(self, reps: int = 3):
        """Pick the stage-kernel layout by timing; returns (code, ms for codes
        64/128/256 (one lane per element, elements per workgroup), 4 (four
        lanes per element, run-time flux) and 5 (four lanes per element, one
        launch per flux))."""
        ch = C.c_int(0)
        ms = np.zeros(5)
        _lib.check(_bind().cfd_batch_autotune(self._h, reps, C.byref(ch), _lib.dptr(ms)))
        return ch.value, ms

    def advance(self, nsteps: int, use_graph: bool = True) -> int:
        left = C.c_int64(0)
        _lib.check(_bind().cfd_batch_advance(self._h, int(nsteps), 1 if use_graph else 0, C.byref(left)))
        return left.value

    def results(self):
        res = (CfdBatchCaseResult * len(self.cases))()
        _lib.check(_bind().cfd_batch_results(self._h, res))
        out = [None] * len(self.cases)
        for pos, i in enumerate(self._order):
            out[i] = res[pos]
        return out

    def states(self):
        q = np.empty_like(self._q0)
        _lib.check(_bind().cfd_batch_download(self._h, _lib.dptr(q)))
        return [q[:, f:f + n].copy() for f, n in zip(self.first, self.count)]

    def run(self, max_steps: int | None = None, use_graph: bool = True) -> list[RunResult]:
        """Run every case to its final time (or ``max_steps`` batched steps)."""
        import time

        start = time.perf_counter()
        self.begin()
        budget = np.iinfo(np.int64).max if max_steps is None else int(max_steps)
        while budget > 0:
            n = min(budget, 1024)
            if self.advance(n, use_graph) == 0:
                break
            budget -= n
        wall = time.perf_counter() - start
        out = []
        for case, r, q in zip(self.cases, self.results(), self.states()):
            if r.status == 2:
                raise SolverError("The initial state must be finite with positive nodal density and pressure.")
            if r.status == 3:
                raise SolverError(f"Invalid time step {r.fail_dt} at t = {r.fail_time}.")
            if r.status == 4:
                raise SolverError(f"Step at t = {r.fail_time} was rejected "
                                  f"{case.solver.config.max_rejections + 1} times.")
            out.append(RunResult(state=q, time=float(r.time), steps=int(r.steps),
                                 rejected_steps=int(r.rejected_steps),
                                 boundary_flux_integral=np.array(r.boundary_flux_integral[:]),
                                 min_density=float(r.min_density), min_pressure=float(r.min_pressure),
                                 max_troubled_cells=int(r.max_troubled_cells), wall_time=wall))
        return out


FLUX_CYCLE = ("rusanov", "roe", "hllc")
LIMITER_CYCLE = ("minmod", "mc", "superbee")


def random_shock_tube_specs(n_cases: int = 256, seed: int = 0, nx_range=(50, 400)):
    """Seeded shock-tube parameters: Nx ~ U{50..400}, Ny = max(2, Nx // 10),
    rho, p ~ U[0.1, 2], u ~ U[-1, 1], v = 0 on each side, x0 ~ U[0.3, 0.7],
    t_final ~ U[0.1, 0.25]; flux = i mod 3 over (rusanov, roe, hllc), limiter =
    i mod 3 over (minmod, mc, superbee)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_cases):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        left = (rng.uniform(0.1, 2.0), rng.uniform(-1.0, 1.0), 0.0, rng.uniform(0.1, 2.0))
        right = (rng.uniform(0.1, 2.0), rng.uniform(-1.0, 1.0), 0.0, rng.uniform(0.1, 2.0))
        specs.append(dict(nx=nx, ny=max(2, nx // 10), left=left, right=right, x0=float(rng.uniform(0.3, 0.7)),
                          final_time=float(rng.uniform(0.1, 0.25)), flux=FLUX_CYCLE[i % 3],
                          limiter=LIMITER_CYCLE[i % 3]))
    return specs


def shock_tube_case(spec, gamma: float = 1.4, cfl: float = 0.5, max_rejections: int = 10) -> Case:
    """A Riemann problem on [0,1] x [0,0.1]: transmissive ends, slip walls,
    left/right states split at x0 by element centroid (as ``cases.py:41-43``)."""
    mesh = rectangle_mesh(spec["nx"], spec["ny"], (0.0, 1.0), (0.0, 0.1))
    bcs = {"left": Transmissive(), "right": Transmissive(), "bottom": SlipWall(), "top": SlipWall()}
    get_interface_flux(spec["flux"])
    get_limiter_function(spec["limiter"])
    solver = EulerSolver(mesh, bcs, SolverConfig(flux=spec["flux"], limiter=spec["limiter"], gamma=gamma, cfl=cfl,
                                                 max_rejections=max_rejections))
    left = primitive_to_conserved(*spec["left"], gamma)
    right = primitive_to_conserved(*spec["right"], gamma)
    on_left = solver.geometry.centroid[:, 0] < spec["x0"]
    state = np.where(on_left[:, None], left, right)
    return Case(solver, np.broadcast_to(state[None], solver.state_shape).copy(), spec["final_time"])
