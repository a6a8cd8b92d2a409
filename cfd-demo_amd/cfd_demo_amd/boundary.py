"""Boundary conditions, applied only through the exterior state of the
interface flux (``src/cfd_demo/boundary.py``).

``SlipWall`` and ``FixedState`` are the reference's (kinds 1 and 2).
``Transmissive`` (exterior = interior trace, boundary flux counted) and
``Periodic`` (faces identified by a translation) are extensions the scale-up
workloads need; the reference's tables can express both (a kind-0 boundary
face already points at its own node, ``dg.py:88-91``) but it has no type for
them. ``ExactState`` is the far-field condition the reference proposes in
``docs/adr/0012-sheared-rectangle-mesh-and-far-field-boundaries.md:32-44``.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass

import numpy as np

SLIP_WALL = 1
FIXED_STATE = 2
TRANSMISSIVE = 3
PERIODIC = 4
EXACT_STATE = 5


@dataclass(frozen=True)
class SlipWall:
    """Inviscid wall: mirror the normal momentum (``boundary.py:17-23``)."""

    kind: int = SLIP_WALL


@dataclass(frozen=True)
class FixedState:
    """Prescribed conserved exterior state (``boundary.py:26-43``)."""

    state: tuple[float, float, float, float]
    kind: int = FIXED_STATE

    def __post_init__(self) -> None:
        state = tuple(float(v) for v in self.state)
        if len(state) != 4 or not np.all(np.isfinite(state)):
            raise ValueError("FixedState needs four finite conserved values.")
        object.__setattr__(self, "state", state)

    @classmethod
    def from_primitive(cls, rho: float, u: float, v: float, p: float, gamma: float) -> "FixedState":
        energy = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
        return cls((rho, rho * u, rho * v, energy))


@dataclass(frozen=True)
class Transmissive:
    """Zero-gradient outflow: the exterior state is the interior trace."""

    kind: int = TRANSMISSIVE


@dataclass(frozen=True)
class Periodic:
    """This tag's faces are the faces of ``partner`` translated by ``shift``
    (x_here = x_partner + shift)."""

    partner: str
    shift: tuple[float, float]
    kind: int = PERIODIC


def _collapsed_gauss(n: int):
    """Reference-triangle rule from the Duffy map of an n x n Gauss-Legendre
    rule: barycentric weights (m, 3) of the points and weights summing to 1."""
    x, w = np.polynomial.legendre.leggauss(n)
    a, b = np.meshgrid(x, x, indexing="ij")
    wa, wb = np.meshgrid(w, w, indexing="ij")
    r = (0.5 * (1.0 + a) * (1.0 - b) - 1.0).ravel()
    s = b.ravel()
    wt = (wa * wb * (1.0 - b)).ravel()
    lam = np.stack([-(r + s) / 2.0, (1.0 + r) / 2.0, (1.0 + s) / 2.0], -1)
    return lam, wt / wt.sum()


@dataclass(frozen=True, eq=False)
class ExactState:
    """Far-field condition from a known solution (the reference's proposed
    ADR 0012): ``state(x, y)`` gives the conserved exterior state (..., 4) at
    physical points and is evaluated once per boundary face node when the
    solver is built; it never changes interior values. The limiter's ghost
    cell across such a face (the element reflected across it) takes the
    state's cell average there: ``average(corners)`` for corners (n, 3, 2)
    returns (n, 4) exact averages, or, without it, a 12 x 12-point collapsed
    Gauss rule averages ``state`` over the reflected triangle."""

    state: Callable
    average: Callable | None = None
    kind: int = EXACT_STATE

    def face_states(self, x, y) -> np.ndarray:
        out = np.asarray(self.state(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)),
                         dtype=np.float64)
        if out.shape != np.shape(x) + (4,) or not np.all(np.isfinite(out)):
            raise ValueError("ExactState.state must return finite (..., 4) conserved states.")
        return out

    def ghost_averages(self, corners) -> np.ndarray:
        corners = np.asarray(corners, dtype=np.float64)
        if self.average is not None:
            out = np.asarray(self.average(corners), dtype=np.float64)
        else:
            lam, w = _collapsed_gauss(12)
            pts = np.einsum("qv,nvd->nqd", lam, corners)
            out = np.einsum("q,nqc->nc", w, self.face_states(pts[..., 0], pts[..., 1]))
        if out.shape != (corners.shape[0], 4) or not np.all(np.isfinite(out)):
            raise ValueError("ExactState.average must return finite (n, 4) conserved averages.")
        return out


def reflect_across_faces(corners, normal, midpoint):
    """Each element (K, 3, 2) mirrored across each of its faces: (K, 3, 3, 2),
    [e, f] the element reflected across face f (outward unit normal (K, 3, 2),
    face midpoint (K, 3, 2))."""
    d = corners[:, None, :, :] - midpoint[:, :, None, :]
    off = np.sum(d * normal[:, :, None, :], axis=-1, keepdims=True)
    return corners[:, None, :, :] - 2.0 * off * normal[:, :, None, :]


BoundaryCondition = SlipWall | FixedState | Transmissive | Periodic | ExactState
