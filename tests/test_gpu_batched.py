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


@pytest.mark.parametrize("use_graph,quad", [(True, True), (False, True), (True, False)])
def test_batched_matches_single_case_runs(use_graph, quad, monkeypatch):
    """Both stage-kernel layouts (four lanes per element, the default, and one
    lane per element) against the single-case runs."""
    cases = _cases(6, seed=3, final_time=0.004)
    print(f"case batched use_graph={use_graph} quad={quad} K={[c.solver.mesh.num_elements for c in cases]}",
          flush=True)
    monkeypatch.setenv("CFDHIP_STAGE_QUAD", "1" if quad else "0")
    batch = BatchedSolver(cases)
    monkeypatch.delenv("CFDHIP_STAGE_QUAD")
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


def test_autotuner_and_rejection_case():
    cases = _cases(3, seed=5, final_time=0.003)
    bad_spec = random_shock_tube_specs(1, seed=9, nx_range=(50, 50))[0]
    bad = shock_tube_case(bad_spec, cfl=1e6, max_rejections=2)
    batch = BatchedSolver(cases + [bad])
    batch.begin()
    choice, ms = batch.autotune(reps=2)
    assert choice in (64, 128, 256, 4) and np.all(ms > 0)
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
