"""Boundary conditions, applied only through the exterior state of the
interface flux (``src/cfd_demo/boundary.py``).

``SlipWall`` and ``FixedState`` are the reference's (kinds 1 and 2).
``Transmissive`` (exterior = interior trace, boundary flux counted) and
``Periodic`` (faces identified by a translation) are extensions the scale-up
workloads need; the reference's tables can express both (a kind-0 boundary
face already points at its own node, ``dg.py:88-91``) but it has no type for
them.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SLIP_WALL = 1
FIXED_STATE = 2
TRANSMISSIVE = 3
PERIODIC = 4


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


BoundaryCondition = SlipWall | FixedState | Transmissive | Periodic
