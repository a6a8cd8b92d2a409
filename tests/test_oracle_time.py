"""Pin the oracle's limiting, CFL step and SSP-RK3 run with the reference's
checks (tests/unit/test_limiting_and_time.py) and the ADR-0005 Sod exact
solution (tests/verification/test_sod_shock_tube.py)."""

import numpy as np
import pytest

from oracle import mesh as om
from oracle.limiting import POSITIVITY_FLOOR, apply_limiters
from oracle.physics import conserved_to_primitive, primitive_to_conserved, sound_speed
from oracle.solver import DEFAULT_CFL, EulerSolver, SolverConfig, SolverError, sod_shock_tube, time_step_geometry
from sod_exact import sample, sod_waves
from test_oracle_dg import FREE, GAMMA, fixed_all, perturbed, solver, uniform


def linear_state(s):
    x, y = s.node_coordinates()
    base = primitive_to_conserved(*FREE, GAMMA)
    return base + np.array([0.05, 0.03, -0.02, 0.07]) * (0.3 * x + 0.2 * y)[..., None]


def test_limiting_preserves_averages():
    s = solver(perturbed(6, 5, seed=5), fixed_all())
    q = linear_state(s) * (1.0 + 0.3 * np.random.default_rng(1).standard_normal(s.state_shape))
    assert np.abs(s.cell_averages(s.limit(q)) - s.cell_averages(q)).max() < 1e-12


def test_linear_data_untouched():
    s = solver(om.rectangle_mesh(8, 8), fixed_all())
    q = linear_state(s)
    inner = np.all(s.connectivity.neighbor_element >= 0, axis=1)
    assert np.abs(s.limit(q)[:, inner] - q[:, inner]).max() < 1e-12


def test_oscillation_flattened_in_one_element():
    s = solver(om.rectangle_mesh(8, 8), fixed_all())
    q = uniform(s, FREE)
    tgt = 2 * (4 * 8 + 4)
    ref = s.reference
    osc = 0.2 * np.cos(4 * np.pi * ref.r) * np.cos(3 * np.pi * ref.s)
    osc -= ref.cell_weights @ osc / ref.cell_weights.sum()
    q[:, tgt, 0] += osc
    lim, flags = s.limit(q, return_flags=True)
    assert np.abs(lim[:, tgt] - q[:, tgt]).max() > 1e-6
    others = np.arange(s.mesh.num_elements) != tgt
    assert np.abs(lim[:, others] - q[:, others]).max() < 1e-12
    assert np.allclose(ref.linear_midpoint_basis @ (ref.p1_midpoint @ lim[:, tgt, 0]), lim[:, tgt, 0], atol=1e-12)
    assert flags.sum() == 1 and flags[tgt]


def test_unlimited_skips_slopes():
    s = solver(om.rectangle_mesh(6, 6), fixed_all(), limiter="unlimited")
    q = uniform(s, FREE)
    q[:, 2 * (3 * 6 + 3), 0] += 0.2 * np.cos(4 * np.pi * s.reference.r)
    assert np.abs(s.limit(q) - q).max() < 1e-13
    _, _, cnt = apply_limiters(q, s.limiter_operators, lambda r: r, GAMMA, limit_slopes=False)
    assert int(cnt) == 0


def test_positivity_repair():
    s = solver(om.rectangle_mesh(4, 4), fixed_all(), limiter="unlimited")
    q = uniform(s, (1.0, 0.0, 0.0, 1.0))
    dev = np.linspace(-1.5, 1.5, 15)
    dev -= s.reference.cell_weights @ dev / s.reference.cell_weights.sum()
    q[:, 5, 0] += dev
    q[:, 5, 3] += 1.2 * dev
    assert q[:, 5, 0].min() < 0
    lim = s.limit(q)
    rho, _, _, p = conserved_to_primitive(lim, GAMMA)
    assert rho.min() >= POSITIVITY_FLOOR * (1 - 1e-9) and p.min() >= POSITIVITY_FLOOR * (1 - 1e-9)
    assert np.abs(s.cell_averages(lim) - s.cell_averages(q)).max() < 1e-12
    other = np.arange(s.mesh.num_elements) != 5
    assert np.abs(lim[:, other] - q[:, other]).max() < 1e-13


def test_time_step_formula():
    s = solver(om.rectangle_mesh(5, 4), fixed_all())
    rho, u, v, p = FREE
    exp = DEFAULT_CFL * time_step_geometry(s.geometry, 4).min() / (np.hypot(u, v) + sound_speed(rho, p, GAMMA))
    assert s.time_step(uniform(s, FREE)) == pytest.approx(float(exp), rel=1e-12)
    fine = solver(om.rectangle_mesh(10, 8), fixed_all())
    assert fine.time_step(uniform(fine, FREE)) == pytest.approx(0.5 * s.time_step(uniform(s, FREE)), rel=1e-12)
    assert np.allclose(time_step_geometry(om.element_geometry(om.rectangle_mesh(1, 1)), 4),
                       2.0 * 0.5 / (25.0 * np.sqrt(2.0)))


def sod_bcs():
    left = om.FixedState.from_primitive(1.0, 0.0, 0.0, 1.0, GAMMA)
    right = om.FixedState.from_primitive(0.125, 0.0, 0.0, 0.1, GAMMA)
    return {"left": left, "right": right, "bottom": om.SlipWall(), "top": om.SlipWall()}


def sod_like(nx=20, ny=2, **kw):
    s = EulerSolver(om.rectangle_mesh(nx, ny), sod_bcs(), SolverConfig(flux="rusanov", limiter="minmod", gamma=GAMMA, **kw))
    on_left = s.geometry.centroid[:, 0] < 0.5
    q = np.empty(s.state_shape)
    q[:, on_left] = primitive_to_conserved(1.0, 0.0, 0.0, 1.0, GAMMA)
    q[:, ~on_left] = primitive_to_conserved(0.125, 0.0, 0.0, 0.1, GAMMA)
    return s, q


def test_run_exact_end_time_and_conservation():
    s, q = sod_like()
    r = s.run(q, 0.01)
    assert r.time == 0.01 and r.steps > 1 and r.rejected_steps == 0 and r.state.shape == s.state_shape
    assert np.abs(s.integrate(r.state) - s.integrate(q) + r.boundary_flux_integral).max() < 1e-13


def test_third_order_in_time():
    s = solver(perturbed(5, 4, seed=7), fixed_all(), limiter="unlimited")
    x, y = s.node_coordinates()
    q0 = primitive_to_conserved(*FREE, GAMMA) * (1.0 + 0.02 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y))[..., None]

    def march(n):
        q = q0
        for _ in range(n):
            q, rep = s.step(q, 0.02 / n)
            assert bool(rep.admissible)
        return q

    ref = march(64)
    e4, e8 = (np.abs(march(n) - ref).max() for n in (4, 8))
    assert np.log2(e4 / e8) > 2.5


def test_rejection_paths():
    s, q = sod_like()
    bad = q.copy()
    bad[0, 0, 0] = -1.0
    with pytest.raises(SolverError, match="positive nodal density and pressure"):
        s.run(bad, 0.01)
    s2, q2 = sod_like(4, 2, cfl=1e6, max_rejections=2)
    with pytest.raises(SolverError, match="was rejected 3 times"):
        s2.run(q2, 0.5)


def test_step_report():
    s, q = sod_like()
    _, rep = s.step(q, s.time_step(q))
    assert rep.min_density > 0 and rep.min_pressure > 0 and int(rep.troubled_cells) > 0 and bool(rep.admissible)
    assert rep.boundary_flux.shape == (4,)


def test_sod_against_exact_riemann_solution():
    """ADR-0005 acceptance shape at a coarse 50x2 grid, t = 0.2, CFL 0.5."""
    s, q0, T = sod_shock_tube(50, 2, flux="rusanov", limiter="minmod")
    r = s.run(q0, T)
    assert r.rejected_steps == 0 and r.min_density > 0 and r.min_pressure > 0
    assert np.abs(s.integrate(r.state) - s.integrate(q0) + r.boundary_flux_integral).max() <= 1e-8
    x, _ = s.node_coordinates()
    rho_ex, _, _ = sample(x, T, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    w = s.reference.cell_weights[:, None] * s.geometry.jacobian[None, :]
    l1 = float(np.sum(w * np.abs(r.state[..., 0] - rho_ex)) / 0.1)
    assert l1 < 0.03, l1
    waves = sod_waves(T)
    assert waves["shock"] == pytest.approx(0.85043, abs=1e-4)
    assert waves["contact"] == pytest.approx(0.68549, abs=1e-4)
    assert waves["head"] == pytest.approx(0.26336, abs=1e-4)
