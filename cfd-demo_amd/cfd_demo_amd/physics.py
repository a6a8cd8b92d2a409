"""Flux and limiter registries, and host-side gas relations for case setup.

``INTERFACE_FLUXES`` / ``get_interface_flux`` and ``LIMITER_FUNCTIONS`` /
``get_limiter_function`` keep the reference's names, error messages and call
signatures (``fluxes.py:146-159``, ``limiters.py:42-57``). Calling a registry
entry evaluates the gfx950 device function through the C ABI
(``cfd_flux_pointwise`` / ``cfd_limiter_pointwise``); the entries also carry
the integer id the solver passes to ``cfd_ctx_create``.

The gas relations below (``euler.py:14-37``) are host NumPy used to build
initial states; the hot path evaluates them on the device.
"""

from __future__ import annotations

import numpy as np

from . import _lib

NUM_VARIABLES = 4


def primitive_to_conserved(rho, u, v, p, gamma: float):
    """(..., 4) conserved states from primitives (``euler.py:14-18``)."""
    rho, u, v, p = (np.asarray(x, dtype=np.float64) for x in (rho, u, v, p))
    energy = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack(np.broadcast_arrays(rho, rho * u, rho * v, energy), axis=-1)


def conserved_to_primitive(q, gamma: float):
    """(rho, u, v, p) of conserved states (``euler.py:21-27``)."""
    q = np.asarray(q, dtype=np.float64)
    rho = q[..., 0]
    u = q[..., 1] / rho
    v = q[..., 2] / rho
    return rho, u, v, (gamma - 1.0) * (q[..., 3] - 0.5 * (q[..., 1] * u + q[..., 2] * v))


def pressure(q, gamma: float):
    return conserved_to_primitive(q, gamma)[3]


def sound_speed(rho, p, gamma: float):
    return np.sqrt(gamma * np.asarray(p) / np.asarray(rho))


class InterfaceFlux:
    """A named interface flux evaluated on the GPU: ``flux(qm, qp, nx, ny, gamma)``."""

    def __init__(self, name: str, flux_id: int):
        self.name = name
        self.id = flux_id
        self.__name__ = name

    def __call__(self, q_minus, q_plus, nx, ny, gamma: float):
        qm = np.asarray(q_minus, dtype=np.float64)
        qp = np.asarray(q_plus, dtype=np.float64)
        shape = np.broadcast_shapes(qm.shape[:-1], qp.shape[:-1], np.shape(nx), np.shape(ny))
        qm = np.ascontiguousarray(np.broadcast_to(qm, shape + (4,))).reshape(-1, 4)
        qp = np.ascontiguousarray(np.broadcast_to(qp, shape + (4,))).reshape(-1, 4)
        nxa = np.ascontiguousarray(np.broadcast_to(np.asarray(nx, np.float64), shape)).reshape(-1)
        nya = np.ascontiguousarray(np.broadcast_to(np.asarray(ny, np.float64), shape)).reshape(-1)
        out = np.empty_like(qm)
        _lib.check(_lib.load().cfd_flux_pointwise(self.id, qm.shape[0], _lib.dptr(qm), _lib.dptr(qp),
                                                   _lib.dptr(nxa), _lib.dptr(nya), float(gamma), _lib.dptr(out)))
        return out.reshape(shape + (4,))

    def __repr__(self) -> str:
        return f"<interface flux {self.name!r}>"


class LimiterFunction:
    """A named limiter function phi(r) evaluated on the GPU."""

    def __init__(self, name: str, limiter_id: int):
        self.name = name
        self.id = limiter_id
        self.__name__ = name

    def __call__(self, r):
        arr = np.asarray(r, dtype=np.float64)
        flat = np.ascontiguousarray(arr).reshape(-1)
        out = np.empty_like(flat)
        _lib.check(_lib.load().cfd_limiter_pointwise(self.id, flat.size, _lib.dptr(flat), _lib.dptr(out)))
        return out.reshape(arr.shape)

    def __repr__(self) -> str:
        return f"<limiter {self.name!r}>"


rusanov = InterfaceFlux("rusanov", 0)
roe = InterfaceFlux("roe", 1)
hllc = InterfaceFlux("hllc", 2)
INTERFACE_FLUXES = {"rusanov": rusanov, "roe": roe, "hllc": hllc}

minmod = LimiterFunction("minmod", 0)
mc = LimiterFunction("mc", 1)
van_leer = LimiterFunction("van_leer", 2)
superbee = LimiterFunction("superbee", 3)
unlimited = LimiterFunction("unlimited", 4)
LIMITER_FUNCTIONS = {"minmod": minmod, "mc": mc, "van_leer": van_leer, "superbee": superbee,
                     "unlimited": unlimited}


def get_interface_flux(name: str) -> InterfaceFlux:
    try:
        return INTERFACE_FLUXES[name]
    except KeyError:
        options = ", ".join(sorted(INTERFACE_FLUXES))
        raise ValueError(f"Unknown interface flux {name!r}; choose one of: {options}.") from None


def get_limiter_function(name: str) -> LimiterFunction:
    try:
        return LIMITER_FUNCTIONS[name]
    except KeyError:
        options = ", ".join(sorted(LIMITER_FUNCTIONS))
        raise ValueError(f"Unknown limiter {name!r}; choose one of: {options}.") from None
