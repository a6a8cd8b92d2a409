"""Pins the torch autograd oracle of the discrete adjoint (oracle/adjoint_torch.py).

The reference has no adjoint to compare with (its ``jax.grad`` checks cover
only pointwise fluxes and limiters, tests/unit/test_physics.py:164-182,
224-229). The torch restatement is pinned by central finite differences of
the NumPy oracle's forward (itself pinned to the reference elsewhere), by
agreement of its forward values with the NumPy oracle, and by the closed form
of the frozen positivity scaling.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import adjoint_torch as at
from oracle.dg import dg_rhs
from pairs import make_pair, mixed_bcs, perturbed_vertices, periodic_bcs, random_state, sod_bcs, smooth_state


def _oracle(flux, limiter, bcs, n=3, seed=0):
    v, e, b = perturbed_vertices(n, n, amplitude=0.2, seed=seed)
    _, o = make_pair(v, e, b, bcs, flux=flux, limiter=limiter)
    return o


def _fd(f, x, d, h):
    return (f(x + h * d) - f(x - h * d)) / (2.0 * h)


@pytest.mark.parametrize("flux", ["rusanov", "roe", "hllc"])
@pytest.mark.parametrize("bcs", ["mixed", "sod", "periodic"])
def test_rhs_vjp_matches_finite_differences(flux, bcs):
    spec = {"mixed": mixed_bcs(), "sod": sod_bcs(), "periodic": periodic_bcs()}[bcs]
    o = _oracle(flux, "minmod", spec)
    rng = np.random.default_rng(1)
    q = random_state(o.state_shape, seed=2)
    mu = rng.standard_normal(q.shape)
    d = rng.standard_normal(q.shape)
    g = at.rhs_vjp(o, q, mu)

    def j(x):
        with np.errstate(all="ignore"):
            return float(np.sum(mu * dg_rhs(x, o.operators, o.flux, o.config.gamma)[0]))

    fd = _fd(j, q, d, 1e-6)
    assert abs(np.sum(g * d) - fd) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_march_forward_matches_numpy_oracle(flux):
    o = _oracle(flux, "minmod", periodic_bcs(), n=4)
    x, y = o.node_coordinates()
    q0 = smooth_state(x, y)
    dts = [0.4 * o.time_step(q0)] * 2
    T = at.TorchOperators(o.operators, o.limiter_operators)
    qt = at.march(at._t(q0), dts, o, T).numpy()
    q = q0
    for dt in dts:
        q, _ = o.step(q, dt)
    assert np.max(np.abs(qt - q)) <= 1e-12 * np.max(np.abs(q))


@pytest.mark.parametrize("flux,limiter", [("rusanov", "unlimited"), ("rusanov", "minmod"), ("roe", "mc"),
                                          ("hllc", "minmod"), ("hllc", "van_leer")])
def test_march_vjp_matches_finite_differences(flux, limiter):
    o = _oracle(flux, limiter, periodic_bcs(), n=4, seed=3)
    x, y = o.node_coordinates()
    q0 = smooth_state(x, y, amp=0.3)
    dts = [0.3 * o.time_step(q0)] * 2
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(q0.shape)
    g, _ = at.march_vjp(o, q0, dts, lam)

    def j(qq):
        for dt in dts:
            qq, _ = o.step(qq, dt)
        return float(np.sum(lam * qq))

    for trial in range(3):
        d = rng.standard_normal(q0.shape)
        fd = _fd(j, q0, d, 1e-6)
        assert abs(np.sum(g * d) - fd) <= 1e-6 * max(1.0, abs(fd)), (trial, np.sum(g * d), fd)


def test_frozen_positivity_scale_closed_form():
    # One element whose nodal density dips below zero: the positivity pass
    # scales q - avg by theta < 1; with theta frozen the transpose is
    # theta lam + (1 - theta) w sum_i lam_i.
    o = _oracle("rusanov", "unlimited", sod_bcs(), n=2)
    q = random_state(o.state_shape, seed=5)
    q[3, 1, 0] = -0.05
    q[3, 1, 3] = 0.5
    tr, theta = at.limiter_decisions(q, o.limiter_operators, o.phi, o.config.gamma, False)
    assert theta[1] < 1.0 and np.all(theta[[0, 2, 3, 4, 5, 6, 7]] == 1.0)
    lam = np.random.default_rng(6).standard_normal(q.shape)
    g = at.limit_vjp(o, q, lam)
    w = o.limiter_operators.average_weights
    expect = lam.copy()
    expect[:, 1] = theta[1] * lam[:, 1] + (1.0 - theta[1]) * w[:, None] * lam[:, 1].sum(axis=0)[None]
    assert np.max(np.abs(g - expect)) <= 1e-14 * np.max(np.abs(lam))
