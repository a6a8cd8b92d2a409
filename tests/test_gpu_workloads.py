"""GPU parity with the oracle on each benchmark workload's own data, at sizes
the oracle runs in seconds:

* four-quadrant Riemann problem (24 x 24 split rectangles, HLLC + MC,
  transmissive sides, CFL 0.5), 10 SSP-RK3 steps;
* isentropic vortex (16 x 16 periodic on [0,10]^2, Roe, unlimited), 20 steps;
* vortex plus 8 seeded density bumps (20 x 20 periodic, Rusanov, CFL 0.1),
  20 steps, unlimited and with minmod.

Each runs its own trajectory on both sides (every step's dt from that side's
CFL rule): the dts agree to 1e-15, the troubled flags of every stage are
identical and the states agree to 1e-11 normwise (time_integration.py:44-84,
solver.py:180-204). The run loop's fused CFL step (the next dt computed in
the last stage's limiter pass) is checked against time_step on the same state.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cfd_demo_amd import _lib
from cfd_demo_amd.cases import four_quadrant, isentropic_vortex, vortex_bumps
from oracle import mesh as om
from oracle import solver as osol
from pairs import rel_err

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-11
DT_TOL = 1e-15


def _oracle_for(solver, bc_specs):
    m = solver.mesh
    omesh = om.TriangleMesh(vertices=np.array(m.vertices), elements=np.array(m.elements),
                            boundary_edges={k: np.array(v) for k, v in m.boundary_edges.items()})
    bcs = {}
    for tag, spec in bc_specs.items():
        bcs[tag] = om.Transmissive() if spec[0] == "trans" else om.Periodic(spec[1], spec[2])
    c = solver.config
    return osol.EulerSolver(omesh, bcs, osol.SolverConfig(flux=c.flux, limiter=c.limiter, gamma=c.gamma, cfl=c.cfl,
                                                          max_rejections=c.max_rejections))


TRANS = {t: ("trans",) for t in ("left", "right", "bottom", "top")}
PERIODIC10 = {"left": ("periodic", "right", (-10.0, 0.0)), "right": ("periodic", "left", (10.0, 0.0)),
              "bottom": ("periodic", "top", (0.0, -10.0)), "top": ("periodic", "bottom", (0.0, 10.0))}

WORKLOADS = {
    "four_quadrant_hllc_mc": (lambda: four_quadrant(24, flux="hllc", limiter="mc"), TRANS, 10),
    "vortex_roe_unlimited": (lambda: isentropic_vortex(16, flux="roe", limiter="unlimited"), PERIODIC10, 20),
    "vortex_bumps_rusanov_unlimited": (lambda: vortex_bumps(20, flux="rusanov", limiter="unlimited", cfl=0.1),
                                       PERIODIC10, 20),
    "vortex_bumps_rusanov_minmod": (lambda: vortex_bumps(20, flux="rusanov", limiter="minmod", cfl=0.1),
                                    PERIODIC10, 20),
}


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_workload_steps_match_oracle(name):
    make, specs, nsteps = WORKLOADS[name]
    case = make()
    g = case.solver
    o = _oracle_for(g, specs)
    print(f"case workload {name} K={g.mesh.num_elements} steps={nsteps}", flush=True)
    qg = qo = np.ascontiguousarray(case.initial_state)
    troubled = 0
    for n in range(nsteps):
        dt_g, dt_o = g.time_step(qg), o.time_step(qo)
        assert dt_g == pytest.approx(dt_o, rel=DT_TOL), (n, dt_g, dt_o)
        qg, rep, flags = g.step_with_flags(qg, dt_g)
        fl = []
        qo, rref = o.step(qo, dt_o, flags_out=fl)
        assert np.array_equal(flags, np.stack(fl)), f"step {n}: troubled flags differ"
        err = rel_err(qg, qo)
        assert err <= STATE_TOL, (n, err)
        assert bool(rep.admissible) and bool(rref.admissible)
        assert int(rep.troubled_cells) == int(rref.troubled_cells)
        assert np.allclose(rep.boundary_flux, rref.boundary_flux, rtol=1e-10, atol=1e-14)
        assert rep.min_pressure == pytest.approx(float(rref.min_pressure), rel=1e-12)
        troubled = max(troubled, int(rep.troubled_cells))
    print(f"{name}: last state rel err {err:.2e}, max troubled {troubled}", flush=True)
    if name == "four_quadrant_hllc_mc":
        assert troubled > 0
    g.close()


@pytest.mark.parametrize("name", ["four_quadrant_hllc_mc", "vortex_bumps_rusanov_minmod"])
def test_run_loop_matches_oracle(name):
    """EulerSolver.run (fused CFL step, device-resident state) against the
    oracle's run loop over the same number of steps."""
    make, specs, nsteps = WORKLOADS[name]
    case = make()
    g = case.solver
    o = _oracle_for(g, specs)
    print(f"case run loop {name} steps={nsteps}", flush=True)
    rg = g.run(case.initial_state, 10.0, max_steps=nsteps)
    ro = o.run(case.initial_state, 10.0, max_steps=nsteps)
    assert rg.steps == ro.steps == nsteps
    assert rg.time == pytest.approx(ro.time, rel=1e-14)
    assert rg.max_troubled_cells == ro.max_troubled_cells
    assert rel_err(rg.state, ro.state) <= STATE_TOL
    assert np.allclose(rg.boundary_flux_integral, ro.boundary_flux_integral, rtol=1e-10, atol=1e-14)
    g.close()


@pytest.mark.parametrize("name", ["four_quadrant_hllc_mc", "vortex_roe_unlimited"])
def test_fused_time_step_equals_time_step(name):
    """The dt the resident run loop takes (computed in the previous step's
    last limiter pass) equals time_step() of the same state, and the oracle's."""
    make, specs, _ = WORKLOADS[name]
    case = make()
    g = case.solver
    o = _oracle_for(g, specs)
    lib = _lib.load()
    ctx = g._context()
    _lib.check(lib.cfd_state_upload(ctx, _lib.dptr(np.ascontiguousarray(case.initial_state))))
    dt = C.c_double(0.0)
    rep = _lib.CfdStepReport()
    _lib.check(lib.cfd_state_advance(ctx, 1, C.byref(dt), C.byref(rep)))  # first dt: k_dt of the initial state
    states, fused = [], []
    for _ in range(6):
        q = np.empty(g.state_shape)
        _lib.check(lib.cfd_state_download(ctx, _lib.dptr(q)))
        _lib.check(lib.cfd_state_advance(ctx, 1, C.byref(dt), C.byref(rep)))  # dt fused from the last stage
        states.append(q)
        fused.append(dt.value)
    # time_step() uploads into the context's buffers, so it runs after the
    # resident steps.
    worst = 0.0
    for n, (q, f) in enumerate(zip(states, fused)):
        direct, oracle_dt = g.time_step(q), o.time_step(q)
        worst = max(worst, abs(f - oracle_dt) / oracle_dt)
        print(f"{name} state {n + 1}: fused {f:.17e} time_step {direct:.17e} oracle {oracle_dt:.17e}", flush=True)
        assert f == pytest.approx(direct, rel=DT_TOL)
        assert f == pytest.approx(oracle_dt, rel=DT_TOL)
    print(f"{name}: worst fused-vs-oracle dt rel diff {worst:.2e}", flush=True)
    g.close()
