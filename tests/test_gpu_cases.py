"""End-to-end GPU runs of the reference's Sod case and the vortex case.

Sod (cases.py:31-44) at the 100 x 10 mesh, Rusanov, minmod, CFL 0.1, t = 0.2:
the reference's verification thresholds (tests/verification/
test_sod_shock_tube.py:199-245; ADR 0005 wave positions), the transmissive
variant against the fixed-state one, and the oracle over the first steps.
Isentropic vortex (periodic, Roe, unlimited): error against the exactly
advected vortex and its convergence order.
"""

from __future__ import annotations

import numpy as np
import pytest

import cfd_demo_amd as gpu
from cfd_demo_amd.cases import isentropic_vortex, isentropic_vortex_state
from oracle import solver as osol
from sod_exact import sample, sod_waves

pytestmark = pytest.mark.gpu


def _l1_density(s, q, rho_ex, area):
    w = s.reference.cell_weights[:, None] * s.geometry.jacobian[None, :]
    return float(np.sum(w * np.abs(q[..., 0] - rho_ex)) / area)


_RUNS = {}


def _sod_run(transmissive):
    if transmissive not in _RUNS:
        case = gpu.sod_shock_tube(100, 10, flux="rusanov", limiter="minmod", cfl=0.1, transmissive_ends=transmissive)
        _RUNS[transmissive] = (case, case.solver.run(case.initial_state, case.final_time))
    return _RUNS[transmissive]


@pytest.mark.parametrize("transmissive", [False, True])
def test_sod_full_run_against_exact_solution(transmissive):
    case, r = _sod_run(transmissive)
    s, q0, T = case.solver, case.initial_state, case.final_time
    print("sod steps", r.steps, "max troubled", r.max_troubled_cells, flush=True)
    assert r.time == T and r.rejected_steps == 0 and r.min_density > 0 and r.min_pressure > 0
    defect = s.integrate(r.state) - s.integrate(q0) + r.boundary_flux_integral
    assert np.abs(defect).max() <= 1e-8
    x, _ = s.node_coordinates()
    rho_ex, _, _ = sample(x, T, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    assert _l1_density(s, r.state, rho_ex, 0.1) <= 0.03
    # Shock: right-most element whose average density is above the midpoint
    # between the post-shock density and the right state.
    avg = s.cell_averages(r.state)[:, 0]
    cx = s.geometry.centroid[:, 0]
    post = sample(np.array([0.8]), T, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))[0][0]
    shock = cx[avg > 0.5 * (post + 0.125)].max()
    assert abs(shock - sod_waves(T)["shock"]) <= 0.02


def test_sod_transmissive_ends_close_to_fixed_ends():
    """No wave reaches the ends by t = 0.2, but q+ = q- does not damp the
    round-off there as the fixed state does; the runs stay close in L1."""
    (case, rf), (_, rt) = _sod_run(False), _sod_run(True)
    s = case.solver
    d = _l1_density(s, rt.state, rf.state[..., 0], 0.1)
    print("L1 |rho_transmissive - rho_fixed|", d, flush=True)
    assert d <= 1e-3


@pytest.mark.parametrize("transmissive", [False, True])
def test_sod_matches_oracle_over_the_first_steps(transmissive):
    case = gpu.sod_shock_tube(100, 10, flux="rusanov", limiter="minmod", cfl=0.1, transmissive_ends=transmissive)
    s, q0 = case.solver, case.initial_state
    so, qo0, _ = osol.sod_shock_tube(100, 10, flux="rusanov", limiter="minmod", cfl=0.1,
                                     transmissive_ends=transmissive)
    assert np.array_equal(q0, qo0)
    T = 0.004
    rg = s.run(q0, T)
    ro = so.run(qo0, T)
    assert rg.steps == ro.steps and rg.time == ro.time == T
    assert rg.max_troubled_cells == ro.max_troubled_cells
    assert np.abs(rg.state[..., 0] - ro.state[..., 0]).max() <= 1e-10
    assert np.allclose(rg.boundary_flux_integral, ro.boundary_flux_integral, rtol=1e-11, atol=1e-15)


def _vortex_error(n, T):
    case = isentropic_vortex(n, flux="roe", limiter="unlimited")
    s = case.solver
    r = s.run(case.initial_state, T)
    x, y = s.node_coordinates()
    exact = isentropic_vortex_state(x, y, T)
    w = s.reference.cell_weights[:, None] * s.geometry.jacobian[None, :]
    return float(np.sum(w * np.abs(r.state[..., 0] - exact[..., 0])) / 100.0), r.steps


def test_isentropic_vortex_converges():
    e1, n1 = _vortex_error(16, 1.0)
    e2, n2 = _vortex_error(32, 1.0)
    order = np.log2(e1 / e2)
    print("vortex L1", e1, e2, "order", order, "steps", n1, n2, flush=True)
    assert e2 < 1e-4 and order > 3.0
