"""GPU: batched solves (one launch per stage for all cases, hipGraph replay)
reproduce each case's own EulerSolver.run on the GPU, which is parity-tested
against the oracle elsewhere. The step sequence (steps, rejections, end time,
troubled maxima) must agree exactly; states and minima to 1e-13 relative (the
batched kernel dispatches the flux at run time and is compiled separately, so
last-bit differences are allowed)."""

import numpy as np
import pytest

from cfd_demo_amd.batched import BatchedSolver, random_shock_tube_specs, shock_tube_case

pytestmark = pytest.mark.gpu


def _cases(n, seed, final_time=None, nx_range=(50, 90)):
    specs = random_shock_tube_specs(n, seed=seed, nx_range=nx_range)
    for s in specs:
        if final_time is not None:
            s["final_time"] = final_time * (1.0 + 0.5 * (s["nx"] % 3))
    return [shock_tube_case(s) for s in specs]


@pytest.mark.parametrize("use_graph,quad,per_flux", [(True, True, False), (False, True, False), (True, False, False),
                                                     (True, True, True), (False, True, True)])
def test_batched_matches_single_case_runs(use_graph, quad, per_flux, monkeypatch):
    """The stage-kernel layouts (four lanes per element with the run-time
    flux, the default, or one launch per flux; one lane per element) against
    the single-case runs."""
    cases = _cases(6, seed=3, final_time=0.004)
    print(f"case batched use_graph={use_graph} quad={quad} per_flux={per_flux} "
          f"K={[c.solver.mesh.num_elements for c in cases]}", flush=True)
    monkeypatch.setenv("CFDHIP_STAGE_QUAD", "1" if quad else "0")
    monkeypatch.setenv("CFDHIP_STAGE_PER_FLUX", "1" if per_flux else "0")
    batch = BatchedSolver(cases)
    monkeypatch.delenv("CFDHIP_STAGE_QUAD")
    monkeypatch.delenv("CFDHIP_STAGE_PER_FLUX")
    got = batch.run(use_graph=use_graph)
    for case, r in zip(cases, got):
        ref = case.solver.run(case.initial_state, case.final_time)
        assert r.steps == ref.steps and r.rejected_steps == ref.rejected_steps
        assert r.time == ref.time == case.final_time
        assert r.max_troubled_cells == ref.max_troubled_cells
        assert np.abs(r.state - ref.state).max() <= 1e-13 * np.abs(ref.state).max()
        assert r.min_density == pytest.approx(ref.min_density, rel=1e-13)
        assert r.min_pressure == pytest.approx(ref.min_pressure, rel=1e-13)
        assert np.allclose(r.boundary_flux_integral, ref.boundary_flux_integral, rtol=1e-12, atol=1e-15)
    batch.close()


@pytest.mark.parametrize("per_flux", [False, True])
def test_batched_every_flux_and_limiter(per_flux, monkeypatch):
    """All 15 flux x limiter pairs in one batch (van_leer and unlimited
    included, which the seeded benchmark tubes do not cycle through), under
    the run-time-flux and the one-launch-per-flux stage, against the
    single-case runs. Mild tubes (density and pressure ratios about 2): strong
    random tubes can reach the positivity floor, where the CFL step collapses
    and last-bit differences change the step count (see the exact-solution
    test below)."""
    fluxes = ("rusanov", "roe", "hllc")
    limiters = ("minmod", "mc", "van_leer", "superbee", "unlimited")
    specs = []
    for i in range(15):
        nx = 50 + 2 * i
        specs.append(dict(nx=nx, ny=max(2, nx // 10), left=(1.0 + 0.02 * i, 0.1 * (i % 2), 0.0, 1.0),
                          right=(0.5, -0.05 * (i % 3), 0.0, 0.4 + 0.02 * i), x0=0.45 + 0.01 * (i % 5),
                          final_time=0.01, flux=fluxes[i % 3], limiter=limiters[i // 3]))
    cases = [shock_tube_case(sp) for sp in specs]
    print(f"case batched 15 flux x limiter pairs per_flux={per_flux}", flush=True)
    monkeypatch.setenv("CFDHIP_STAGE_PER_FLUX", "1" if per_flux else "0")
    batch = BatchedSolver(cases)
    monkeypatch.delenv("CFDHIP_STAGE_PER_FLUX")
    got = batch.run()
    for sp, case, r in zip(specs, cases, got):
        ref = case.solver.run(case.initial_state, case.final_time)
        what = (sp["flux"], sp["limiter"])
        assert r.steps == ref.steps and r.rejected_steps == ref.rejected_steps, what
        assert r.time == ref.time == case.final_time, what
        assert r.max_troubled_cells == ref.max_troubled_cells, what
        assert np.abs(r.state - ref.state).max() <= 1e-13 * np.abs(ref.state).max(), what
        assert np.allclose(r.boundary_flux_integral, ref.boundary_flux_integral, rtol=1e-12, atol=1e-15), what
    batch.close()


def test_autotuner_and_rejection_case():
    cases = _cases(3, seed=5, final_time=0.003)
    bad_spec = random_shock_tube_specs(1, seed=9, nx_range=(50, 50))[0]
    bad = shock_tube_case(bad_spec, cfl=1e6, max_rejections=2)
    batch = BatchedSolver(cases + [bad])
    batch.begin()
    choice, ms = batch.autotune(reps=2)
    assert choice in (64, 128, 256, 4, 5) and np.all(ms > 0)
    while batch.advance(64) > 0:
        pass
    res = batch.results()
    assert [r.status for r in res[:3]] == [1, 1, 1]
    assert res[3].status == 4 and res[3].rejected_steps == 3
    states = batch.states()
    for case, r, q in zip(cases, res[:3], states[:3]):
        ref = case.solver.run(case.initial_state, case.final_time)
        assert r.steps == ref.steps
        assert np.abs(q - ref.state).max() <= 1e-13 * np.abs(ref.state).max()
    batch.close()


def _wave_bounds(left, right, g=1.4):
    """Slowest and fastest signal speeds of the exact Riemann solution."""
    from sod_exact import star_state

    (rl, ul, pl), (rr, ur, pr) = left, right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    ps, _ = star_state(left, right, g)
    sl = ul - cl * (np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g)) if ps > pl else 1.0)
    sr = ur + cr * (np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g)) if ps > pr else 1.0)
    return sl, sr


# Seed 11, Nx in 100..250: tubes 4 (Roe, MC) and 14 (HLLC, superbee) have a
# contact with a density jump of 4.8 and 7.5 whose low-density side the p = 4
# polynomials overshoot below zero within the first steps; the positivity
# limiter holds those nodes near zero density with finite momentum, |u| + c
# grows past 1e6 and the CFL step collapses (t stays below 0.01 after 1e5
# steps). The oracle marches the same way (dt 1.7e-5 -> 1e-7 in 40 steps on
# tube 14), at CFL 0.5 and at CFL 0.1, so it is the scheme's behaviour; the
# start of it is parity-tested below.
_EXACT_SEED, _STALLING = 11, (4, 14)


def _exact_specs():
    specs = random_shock_tube_specs(16, seed=_EXACT_SEED, nx_range=(100, 250))
    for s in specs:
        (rl, ul, _, pl), (rr, ur, _, pr) = s["left"], s["right"]
        sl, sr = _wave_bounds((rl, ul, pl), (rr, ur, pr))
        t_exit = min(s["x0"] / max(-sl, 1e-12), (1.0 - s["x0"]) / max(sr, 1e-12))
        s["final_time"] = min(s["final_time"], 0.8 * t_exit)
    return specs


def test_batched_random_tubes_against_exact_solutions():
    """14 seeded random shock tubes (the c4 generator, Nx in 100..250, the
    default CFL 0.5) run as one batch to a final time before any wave
    reaches the ends; the L1 density error of each against its exact
    Riemann solution is within 1 % of the exact density's total variation
    (the reference's Sod bound, L1 0.03 at total variation 0.875, is 3.4 %;
    measured 0.2-0.6 %), and mass, momentum and energy are conserved to
    round-off with the boundary fluxes."""
    from sod_exact import sample

    specs = [s for i, s in enumerate(_exact_specs()) if i not in _STALLING]
    cases = [shock_tube_case(s) for s in specs]
    batch = BatchedSolver(cases)
    batch.begin()
    steps = 0
    for _ in range(16):
        left = batch.advance(2048)
        steps += 2048
        print(f"batched steps {steps}, cases running {left}", flush=True)
        if left == 0:
            break
    assert left == 0, [(i, r.time, r.steps) for i, r in enumerate(batch.results()) if r.status == 0]
    results = batch.results()
    states = batch.states()
    worst = 0.0
    for s, case, r, q in zip(specs, cases, results, states):
        assert r.status == 1
        solver = case.solver
        left = (s["left"][0], s["left"][1], s["left"][3])
        right = (s["right"][0], s["right"][1], s["right"][3])
        x, _ = solver.node_coordinates()
        xu, inv = np.unique(x, return_inverse=True)  # the solution depends on x only
        rho_ex = sample(xu, r.time, s["x0"], left, right)[0][inv].reshape(x.shape)
        w = solver.reference.cell_weights[:, None] * solver.geometry.jacobian[None, :]
        l1 = float(np.sum(w * np.abs(q[..., 0] - rho_ex)) / 0.1)
        tv = float(np.abs(np.diff(sample(np.linspace(0.0, 1.0, 100001), r.time, s["x0"], left, right)[0])).sum())
        worst = max(worst, l1 / tv)
        defect = solver.integrate(q) - solver.integrate(case.initial_state) + np.array(r.boundary_flux_integral[:])
        print(f"case nx={s['nx']} {s['flux']} {s['limiter']} t={r.time:.4f} steps={r.steps} L1={l1:.2e} TV={tv:.3f} "
              f"defect={np.abs(defect).max():.1e}", flush=True)
        assert np.abs(defect).max() <= 1e-10, defect
        assert r.time == s["final_time"]
    print("worst L1 density error / total variation", worst, flush=True)
    assert worst <= 0.01
    batch.close()


def test_stalling_tube_start_matches_oracle():
    """Tube 14 of the set above (HLLC, superbee) to t = 4e-4, about 25 steps
    into its density collapse (min density 0.17 -> 0.03): the GPU solver and
    the oracle take the same steps and troubled-cell maxima and agree to
    1e-12 in density (measured 2e-14)."""
    from oracle import solver as osol
    from oracle.mesh import SlipWall as OSlip, Transmissive as OTrans, rectangle_mesh as orect

    s = _exact_specs()[14]
    case = shock_tube_case(s)
    T = 4e-4
    print(f"case stalling tube nx={s['nx']} {s['flux']} {s['limiter']} T={T}", flush=True)
    rg = case.solver.run(case.initial_state, T)
    mesh = orect(s["nx"], s["ny"], (0.0, 1.0), (0.0, 0.1))
    bcs = {"left": OTrans(), "right": OTrans(), "bottom": OSlip(), "top": OSlip()}
    so = osol.EulerSolver(mesh, bcs, osol.SolverConfig(flux=s["flux"], limiter=s["limiter"], gamma=1.4, cfl=0.5))
    assert np.array_equal(so.geometry.centroid, case.solver.geometry.centroid)
    ro = so.run(case.initial_state, T)
    d = np.abs(rg.state[..., 0] - ro.state[..., 0]).max()
    print("steps", rg.steps, ro.steps, "min density", rg.min_density, ro.min_density, "max |drho|", d, flush=True)
    assert rg.steps == ro.steps and rg.time == ro.time == T
    assert rg.max_troubled_cells == ro.max_troubled_cells
    assert ro.min_density < 0.05
    assert d <= 1e-12
