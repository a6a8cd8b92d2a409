"""The oracle's RHS and SSP-RK3 step over worker processes (test
infrastructure: the CPU baseline the benchmark times beside the GPU).

The element range is split into contiguous chunks, one per worker process,
forked once after the solver's tables are built. Every worker runs the
oracle's own functions on its chunk -- ``dg_rhs`` (``dg.py:140-165``) with
the exterior traces gathered from the whole shared state, the stage
combinations of ``ssp_rk3_step`` (``time_integration.py:70-74``), and
``apply_limiters`` (``limiting.py:203-216``) with the neighbour averages read
from the shared averages of every element -- so the per-element arithmetic is
the serial oracle's. Barriers separate the phases that read other chunks'
results: a stage's unlimited averages before the limiter, the limited stage
state before the next stage's right-hand side. The per-stage report values
(boundary flux, admissibility, minima, troubled count) are reduced in the
parent, as ``ssp_rk3_step`` reduces them.
"""

from __future__ import annotations

import multiprocessing as mp
import os

import numpy as np

from .dg import DGOperators, dg_rhs
from .limiting import LimiterOperators, apply_limiters
from .physics import pressure
from .solver import StepReport


def _slice_dg(ops: DGOperators, k0: int, k1: int) -> DGOperators:
    c = slice(k0, k1)
    return ops._replace(inverse_metric=ops.inverse_metric[:, c], normal_x=ops.normal_x[:, c],
                        normal_y=ops.normal_y[:, c], face_scale=ops.face_scale[:, c],
                        exterior_index=ops.exterior_index[:, c], boundary_kind=ops.boundary_kind[:, c],
                        boundary_state=ops.boundary_state[:, c], boundary_weight=ops.boundary_weight[:, c],
                        jacobian=ops.jacobian[c])


def _slice_lim(ops: LimiterOperators, k0: int, k1: int) -> LimiterOperators:
    c = slice(k0, k1)
    return ops._replace(neighbor=ops.neighbor[:, c], neighbor_kind=ops.neighbor_kind[:, c],
                        normal_x=ops.normal_x[:, c], normal_y=ops.normal_y[:, c], fixed_state=ops.fixed_state[:, c],
                        alpha=ops.alpha[:, c], beta=ops.beta[:, c], partner_face=ops.partner_face[:, c])


def default_workers() -> int:
    """The cores this process may run on, at most 16."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


class ParallelOracle:
    """``rhs(q)`` and ``step(q, dt)`` of an oracle ``EulerSolver`` on
    ``workers`` processes. Use as a context manager (the workers end on
    ``close``)."""

    def __init__(self, solver, workers: int | None = None):
        self.solver = solver
        self.workers = workers or default_workers()
        shape = solver.state_shape
        self.shape = shape
        K = shape[1]
        ctx = mp.get_context("fork")
        n = int(np.prod(shape))
        # Shared buffers: step input, three stage states, the RHS, averages.
        self._raw = [ctx.RawArray("d", n) for _ in range(5)] + [ctx.RawArray("d", K * 4)]
        self._q0, self._x, self._y, self._z, self._rhs = (np.frombuffer(r, dtype=np.float64).reshape(shape)
                                                          for r in self._raw[:5])
        self._avg = np.frombuffer(self._raw[5], dtype=np.float64).reshape(K, 4)
        bounds = np.linspace(0, K, self.workers + 1).astype(int)
        self._chunks = list(zip(bounds[:-1], bounds[1:]))
        self._barrier = ctx.Barrier(self.workers)
        self._pipes, self._procs = [], []
        for w, (k0, k1) in enumerate(self._chunks):
            parent, child = ctx.Pipe()
            p = ctx.Process(target=self._serve, args=(child, k0, k1), daemon=True)
            p.start()
            child.close()
            self._pipes.append(parent)
            self._procs.append(p)

    # -- worker side ------------------------------------------------------------
    def _serve(self, pipe, k0, k1):
        for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
            os.environ[v] = "1"
        s = self.solver
        ops = _slice_dg(s.operators, k0, k1)
        lops = _slice_lim(s.limiter_operators, k0, k1)
        gamma = s.config.gamma
        c = slice(k0, k1)
        while True:
            cmd, arg = pipe.recv()
            if cmd == "stop":
                return
            with np.errstate(invalid="ignore", over="ignore", divide="ignore"):
                if cmd == "rhs":
                    r, b = dg_rhs(self._x[:, c], ops, s.flux, gamma, q_exterior=self._x)
                    self._rhs[:, c] = r
                    pipe.send(b)
                elif cmd == "step":
                    pipe.send(self._step_chunk(c, arg, ops, lops, gamma))

    def _step_chunk(self, c, dt, ops, lops, gamma):
        s = self.solver
        q = self._q0[:, c]
        out = []
        src = self._q0
        for stage, dst in enumerate((self._x, self._y, self._z)):
            r, b = dg_rhs(src[:, c], ops, s.flux, gamma, q_exterior=src)
            if stage == 0:
                u = q + dt * r
            elif stage == 1:
                u = 0.75 * q + 0.25 * (src[:, c] + dt * r)
            else:
                u = q / 3.0 + 2.0 / 3.0 * (src[:, c] + dt * r)
            self._avg[c] = np.einsum("i,ikc->kc", lops.average_weights, u)
            self._barrier.wait()  # every chunk's averages are written
            u, avg, tr, _ = apply_limiters(u, lops, s.phi, gamma, s.limit_slopes, return_flags=True,
                                           avg_all=self._avg)
            ok = bool(np.all(np.isfinite(u)) & np.all(avg[:, 0] > 0.0) & np.all(pressure(avg, gamma) > 0.0))
            dst[:, c] = u
            out.append((b, ok, float(np.min(u[..., 0])), float(np.min(pressure(u, gamma))), int(tr)))
            self._barrier.wait()  # the limited stage state is complete
            src = dst
        return out

    # -- parent side ------------------------------------------------------------
    def _all(self, cmd, arg=None):
        for p in self._pipes:
            p.send((cmd, arg))
        return [p.recv() for p in self._pipes]

    def rhs(self, q):
        """The oracle's (RHS, boundary flux) of ``q``."""
        self._x[...] = q
        parts = self._all("rhs")
        return self._rhs.copy(), np.sum(parts, axis=0)

    def step(self, q, dt):
        """The oracle's SSP-RK3 step of ``q``: (new state, StepReport)."""
        self._q0[...] = q
        parts = self._all("step", float(dt))
        b = [np.sum([p[s][0] for p in parts], axis=0) for s in range(3)]
        rep = StepReport(
            boundary_flux=dt * (b[0] / 6.0 + b[1] / 6.0 + 2.0 * b[2] / 3.0),
            admissible=np.bool_(all(p[s][1] for p in parts for s in range(3))),
            min_density=np.float64(min(p[s][2] for p in parts for s in range(3))),
            min_pressure=np.float64(min(p[s][3] for p in parts for s in range(3))),
            troubled_cells=np.int64(max(sum(p[s][4] for p in parts) for s in range(3))),
        )
        return self._z.copy(), rep

    def close(self):
        if self._procs:
            for p in self._pipes:
                p.send(("stop", None))
            for p in self._procs:
                p.join()
            self._procs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
