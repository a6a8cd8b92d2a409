"""GPU checks at the full sizes of the benchmark configurations, through
properties that hold at any size (the oracle is too slow there):

* four-quadrant mesh, 1000 x 1000 rectangles = 2,000,000 triangles, HLLC:
  free-stream preservation of the RHS, the discrete conservation identity
  sum_k J_k sum_i w_i RHS_ik + boundary flux = 0, the three stage-kernel
  layouts agreeing on the RHS, and one limited SSP-RK3 step conserving mass,
  momentum and energy to round-off;
* isentropic vortex on 512 x 512 periodic rectangles (524,288 triangles),
  Roe, unlimited: 100 steps against the exactly advected vortex;
* all 256 seeded shock tubes of the batched benchmark: the batch against
  each case's own run over the first steps.
"""

import numpy as np
import pytest

from pairs import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def quadrant():
    from cfd_demo_amd.cases import four_quadrant

    print("case four-quadrant 1000x1000 (2M triangles) setup", flush=True)
    case = four_quadrant(1000, flux="hllc", limiter="mc")
    yield case
    case.solver.close()


def _geometric_scale(s):
    """Largest metric factor of the mesh (|rx| + |ry| + |sx| + |sy| and the
    face scale 0.5 |e| / J): RHS round-off grows with it."""
    g = s.tables.geo
    return float(max(np.abs(g[:, 0:4]).sum(axis=1).max(), np.abs(g[:, 10:13]).max()))


def test_free_stream_rhs_vanishes_at_2M(quadrant):
    from cfd_demo_amd.physics import primitive_to_conserved

    s = quadrant.solver
    q = np.broadcast_to(primitive_to_conserved(1.3, 0.4, -0.25, 0.9, 1.4), s.state_shape).copy()
    print("case free stream rhs 2M", flush=True)
    r, b = s.rhs(q)
    scale = _geometric_scale(s)
    err = float(np.abs(r).max())
    print(f"free stream max |rhs| {err:.3e}, metric scale {scale:.1f}", flush=True)
    # Each RHS entry sums O(30) operator terms of size scale * |F| (|F| < 2
    # here) that cancel exactly in exact arithmetic (measured 2.5e-11 at
    # scale 6000, i.e. about 70 ulps of one term).
    assert err <= 1e-15 * 30 * 2.0 * scale


def test_conservation_identity_at_2M(quadrant):
    s, q0 = quadrant.solver, quadrant.initial_state
    print("case conservation identity rhs 2M", flush=True)
    r, b = s.rhs(q0)
    total = s.integrate(r)
    w = s.reference.cell_weights[:, None] * s.geometry.jacobian[None, :]
    mag = np.array([float(np.sum(w * np.abs(r[..., c]))) for c in range(4)])
    print("integral of rhs", total, "boundary flux", b, "sum |J w rhs|", mag, flush=True)
    assert np.all(np.abs(total + b) <= 1e-12 * mag + 1e-14)


def test_stage_layouts_agree_at_2M(quadrant, monkeypatch):
    s, q0 = quadrant.solver, quadrant.initial_state
    print("case rhs layout quad vs lane 2M", flush=True)
    s.close()
    monkeypatch.setenv("CFDHIP_STAGE_QUAD", "1")
    r_quad, b_quad = s.rhs(q0)
    s.close()
    monkeypatch.setenv("CFDHIP_STAGE_QUAD", "0")
    r_other, b_other = s.rhs(q0)
    s.close()
    print(f"normwise rel diff {rel_err(r_other, r_quad):.3e}", flush=True)
    assert rel_err(r_other, r_quad) <= 1e-12
    assert np.allclose(b_other, b_quad, rtol=1e-12, atol=1e-13)


def test_limited_step_conserves_at_2M(quadrant):
    s, q0 = quadrant.solver, quadrant.initial_state
    print("case one limited step 2M", flush=True)
    dt = s.time_step(q0)
    q1, rep = s.step(q0, dt)
    defect = s.integrate(q1) - s.integrate(q0) + np.asarray(rep.boundary_flux)
    scale = np.abs(s.integrate(np.abs(q0)))
    print(f"dt {dt:.3e} troubled {int(rep.troubled_cells)} defect {defect} min density "
          f"{float(rep.min_density):.3e}", flush=True)
    assert bool(rep.admissible)
    assert int(rep.troubled_cells) > 0  # the quadrant interfaces are troubled
    assert float(rep.min_density) > 0.0 and float(rep.min_pressure) > 0.0
    assert np.all(np.abs(defect) <= 1e-13 * scale)


def test_vortex_512_against_exact_advection():
    from cfd_demo_amd.cases import isentropic_vortex, isentropic_vortex_state

    print("case vortex 512x512 (524288 triangles), 100 steps", flush=True)
    case = isentropic_vortex(512, flux="roe", limiter="unlimited")
    s, q0 = case.solver, case.initial_state
    r = s.run(q0, case.final_time, max_steps=100)
    x, y = s.node_coordinates()
    exact = isentropic_vortex_state(x, y, r.time)
    w = s.reference.cell_weights[:, None] * s.geometry.jacobian[None, :]
    l1 = float(np.sum(w * np.abs(r.state[..., 0] - exact[..., 0])) / 100.0)
    defect = s.integrate(r.state) - s.integrate(q0)
    print(f"steps {r.steps} t {r.time:.4e} L1 density {l1:.3e} defect {defect}", flush=True)
    assert r.steps == 100 and r.rejected_steps == 0
    assert l1 <= 1e-9
    assert np.all(np.abs(defect) <= 1e-12 * np.abs(s.integrate(np.abs(q0))))
    s.close()


def test_batched_256_tubes_match_per_case_runs():
    from cfd_demo_amd.batched import BatchedSolver, random_shock_tube_specs, shock_tube_case

    specs = random_shock_tube_specs(256, seed=0)
    cases = [shock_tube_case(sp) for sp in specs]
    print(f"case 256 batched tubes, {sum(c.solver.mesh.num_elements for c in cases)} triangles, 5 steps", flush=True)
    batch = BatchedSolver(cases)
    batch.begin()
    batch.advance(5)
    results, states = batch.results(), batch.states()
    batch.close()
    worst = 0.0
    for i, (case, rb, qb) in enumerate(zip(cases, results, states)):
        ref = case.solver.run(case.initial_state, case.final_time, max_steps=5)
        assert rb.steps == ref.steps == 5 and rb.rejected_steps == ref.rejected_steps, i
        # Before the final clamp t is a sum of CFL steps, whose wave-speed
        # maxima the run-time-flux kernel may round differently in the last bit.
        assert float(rb.time) == pytest.approx(ref.time, rel=1e-13), i
        worst = max(worst, float(np.abs(qb - ref.state).max() / np.abs(ref.state).max()))
        case.solver.close()
    print(f"worst relative state difference over 256 cases {worst:.3e}", flush=True)
    assert worst <= 1e-13
