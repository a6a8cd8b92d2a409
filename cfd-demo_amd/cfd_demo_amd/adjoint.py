"""Discrete adjoint of the DG SSP-RK3 march on the GPU, and its torch wrapper.

The reference has no adjoint: its derivatives are ``jax.grad`` of the
pointwise fluxes and limiters (``tests/unit/test_physics.py:164-182,224-229``).
Here the whole march q0 -> qN of ``ssp_rk3_step`` (``time_integration.py:
44-84``) is differentiated with

* the time steps frozen (recorded by the first forward from the CFL rule of
  ``stable_time_step``, ``time_integration.py:37-41``, or given), and
* every limiter pass's troubled-cell flags and positivity scale frozen at
  their forward values (``limiting.py:136-200``).

The backward recomputes each segment of ``checkpoint_every`` steps from its
checkpoint (HBM), keeps the stage states and a tape of the limiter's repaired
elements, and runs hand-written adjoint kernels (transposed Dr/Ds/LIFT,
flux-Jacobian transposes from dual-number evaluation of the same flux code,
the frozen-flag limiter transpose) in reverse.

A batch of B independent fields on one mesh is one context over B copies of
the mesh tables; fields are laid out (B, Np, K, 4) at the API.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .solver import EulerSolver


def rhs_vjp(solver: EulerSolver, q, mu) -> np.ndarray:
    """(dL/dq)^T mu for the DG right-hand side L of ``dg.py:140-165``."""
    q = solver._as_state(q)
    mu = solver._as_state(mu)
    out = np.empty_like(q)
    _lib.check(_lib.load().cfd_rhs_vjp(solver._context(), _lib.dptr(q), _lib.dptr(mu), _lib.dptr(out)))
    return out


def limit_vjp(solver: EulerSolver, q, lam) -> np.ndarray:
    """(dLim/dq)^T lam for ``apply_limiters`` (``limiting.py:203-216``) at q,
    with its troubled flags and positivity scale held fixed."""
    q = solver._as_state(q)
    lam = solver._as_state(lam)
    out = np.empty_like(q)
    _lib.check(_lib.load().cfd_limit_vjp(solver._context(), _lib.dptr(q), _lib.dptr(lam), _lib.dptr(out)))
    return out


def replicate_tables(tables, batch: int):
    """Tables of ``batch`` disjoint copies of one mesh: field b owns elements
    [b K, (b + 1) K); interior neighbours shift by b K, fixed-state indices
    and every other column stay as they are."""
    K = tables.num_elements
    geo = np.tile(tables.geo, (batch, 1))
    lim = np.tile(tables.lim, (batch, 1))
    conn = np.tile(tables.conn, (batch, 1))
    code = tables.conn[:, 3].astype(np.int64)
    kind = np.stack([(code >> (8 * f + 3)) & 7 for f in range(3)], axis=1)
    interior = np.tile(kind == 0, (batch, 1))
    shift = np.repeat(np.arange(batch, dtype=np.int64) * K, K)[:, None]
    nbr = conn[:, :3].astype(np.int64) + np.where(interior, shift, 0)
    if nbr.max(initial=0) >= np.iinfo(np.int32).max:
        raise ValueError("batch too large: element ids are int32")
    conn[:, :3] = nbr.astype(np.int32)
    return np.ascontiguousarray(geo), np.ascontiguousarray(conn), np.ascontiguousarray(lim)


class DiscreteAdjoint:
    """Checkpointed forward and discrete-adjoint backward of ``num_steps``
    SSP-RK3 steps of ``solver`` over a batch of fields.

    ``dts``: the frozen time steps; ``None`` records them from the CFL rule
    (the solver's cfl, min over the batch) in the first :meth:`forward`.
    """

    def __init__(self, solver: EulerSolver, num_steps: int, *, batch: int = 1, dts=None,
                 checkpoint_every: int = 20):
        if batch < 1:
            raise ValueError("batch must be positive")
        self.solver = solver
        self.batch = int(batch)
        self.num_steps = int(num_steps)
        self.checkpoint_every = int(checkpoint_every)
        self.K = solver.mesh.num_elements
        _lib.ensure_init(solver.reference)
        lib = _lib.load()
        t = solver.tables
        geo, conn, lim = replicate_tables(t, self.batch)
        bcs = t.bc_states if t.bc_states.size else np.zeros((1, 4))
        self._keep = (geo, conn, lim, bcs)
        mt = _lib.CfdMeshTables(self.batch * self.K, _lib.dptr(geo), _lib.iptr(conn), _lib.dptr(lim),
                                _lib.dptr(bcs), int(t.bc_states.shape[0]))
        self._ctx = C.c_void_p()
        cfg = solver.config
        _lib.check(lib.cfd_ctx_create(C.byref(mt), solver.flux.id, solver.phi.id, float(cfg.gamma), float(cfg.cfl),
                                      int(cfg.max_rejections), C.byref(self._ctx)))
        self._adj = C.c_void_p()
        d = None
        if dts is not None:
            d = np.ascontiguousarray(dts, dtype=np.float64)
            if d.shape != (self.num_steps,):
                raise ValueError(f"Expected {self.num_steps} time steps, got shape {d.shape}.")
        code = lib.cfd_adjoint_create(self._ctx, self.num_steps, None if d is None else _lib.dptr(d),
                                      self.checkpoint_every, C.byref(self._adj))
        if code != _lib.CFD_OK:
            lib.cfd_ctx_destroy(self._ctx)
            self._ctx = None
            _lib.check(code)

    def close(self):
        lib = _lib.load()
        if getattr(self, "_adj", None):
            lib.cfd_adjoint_destroy(self._adj)
            self._adj = None
        if getattr(self, "_ctx", None):
            lib.cfd_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- layout -----------------------------------------------------------------
    def _field_shape(self):
        return (self.batch,) + self.solver.state_shape

    def _to_internal(self, x):
        """(B, Np, K, 4) or (Np, K, 4) CUDA float64 -> contiguous (Np, B K, 4)."""
        import torch

        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64):
            raise TypeError("expected a CUDA float64 torch tensor")
        if x.dim() == 3 and self.batch == 1:
            x = x.unsqueeze(0)
        if tuple(x.shape) != self._field_shape():
            raise ValueError(f"Expected a state of shape {self._field_shape()}, got {tuple(x.shape)}.")
        Np, K = self.solver.state_shape[:2]
        return x.permute(1, 0, 2, 3).reshape(Np, self.batch * K, 4).contiguous()

    def _to_fields(self, y, like):
        Np, K = self.solver.state_shape[:2]
        out = y.reshape(Np, self.batch, K, 4).permute(1, 0, 2, 3).contiguous()
        return out[0] if like.dim() == 3 else out

    # -- passes -------------------------------------------------------------------
    def forward(self, q0, replay: bool = False):
        """qN from q0 (torch CUDA float64, (B, Np, K, 4) or (Np, K, 4));
        keeps the checkpoints and limiter decisions for :meth:`backward`.

        ``replay=True`` instead marches with the troubled-cell decisions of
        the last plain forward (the frozen map :meth:`backward`
        differentiates) and keeps nothing: for finite-difference checks."""
        import torch

        x = self._to_internal(q0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        fn = _lib.load().cfd_adjoint_forward_replay if replay else _lib.load().cfd_adjoint_forward
        _lib.check(fn(self._adj, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())))
        return self._to_fields(y, q0)

    def backward(self, gn):
        """(dqN/dq0)^T gn for the last forward's q0."""
        import torch

        g = self._to_internal(gn)
        out = torch.empty_like(g)
        torch.cuda.synchronize()
        _lib.check(_lib.load().cfd_adjoint_backward(self._adj, C.c_void_p(g.data_ptr()), C.c_void_p(out.data_ptr())))
        return self._to_fields(out, gn)

    @property
    def time_steps(self) -> np.ndarray:
        d = np.empty(self.num_steps)
        _lib.check(_lib.load().cfd_adjoint_time_steps(self._adj, _lib.dptr(d)))
        return d

    def stats(self) -> dict:
        s = _lib.CfdAdjointStats()
        _lib.check(_lib.load().cfd_adjoint_stats(self._adj, C.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}


def _evolve_function():
    import torch

    class Evolve(torch.autograd.Function):
        """q0 -> qN through :class:`DiscreteAdjoint` (hand-written backward)."""

        @staticmethod
        def forward(ctx, q0, adj):
            ctx.adj = adj
            return adj.forward(q0.detach())

        @staticmethod
        def backward(ctx, g):
            return ctx.adj.backward(g.contiguous()), None

    return Evolve


_EVOLVE = None


def evolve(q0, adj: DiscreteAdjoint):
    """Differentiable march: ``evolve(q0, adj)`` returns qN and backpropagates
    through the discrete adjoint."""
    global _EVOLVE
    if _EVOLVE is None:
        _EVOLVE = _evolve_function()
    return _EVOLVE.apply(q0, adj)


def window_weights(solver: EulerSolver, window=(6.0, 8.0, 6.0, 8.0)) -> np.ndarray:
    """Nodal quadrature weights J_k w_i (Np, K) of the elements whose centroid
    lies inside the window (x0, x1, y0, y1), zero elsewhere."""
    cx, cy = solver.geometry.centroid[:, 0], solver.geometry.centroid[:, 1]
    x0, x1, y0, y1 = window
    inside = (cx > x0) & (cx < x1) & (cy > y0) & (cy < y1)
    w = solver.reference.cell_weights[:, None] * solver.geometry.jacobian[None, :]
    return np.where(inside[None, :], w, 0.0)


def pressure_squared_objective(q, weights, gamma: float):
    """J = sum_i,k W_ik p_ik^2 (torch; q (..., Np, K, 4), weights (Np, K))."""
    p = (gamma - 1.0) * (q[..., 3] - 0.5 * (q[..., 1] * q[..., 1] + q[..., 2] * q[..., 2]) / q[..., 0])
    return (weights * p * p).sum()
