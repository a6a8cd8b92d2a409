"""The Sod case (cases.py:31-44; 100 x 10, Rusanov, minmod, CFL 0.1) run by
EulerSolver.run on the GPU against the oracle's own runs to t = 0.01, 0.02
and 0.05 (tests/golden/sod_100x10_trajectory.npz, written by
tests/golden/make_sod_trajectory.py): identical step counts, end times and
troubled maxima, nodal density within 1e-12, boundary-flux integrals and
minima within 1e-10 relative.
"""

from pathlib import Path

import numpy as np
import pytest

from cfd_demo_amd import sod_shock_tube

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "sod_100x10_trajectory.npz"


def test_sod_trajectory_matches_oracle_runs():
    gold = np.load(FIXTURE)
    case = sod_shock_tube(100, 10, flux="rusanov", limiter="minmod", cfl=0.1)
    s, q0 = case.solver, case.initial_state
    for k, T in enumerate(gold["end_times"]):
        print(f"case sod trajectory T={T}", flush=True)
        r = s.run(q0, float(T))
        diff = float(np.abs(r.state[..., 0] - gold[f"density_{k}"]).max())
        print(f"T={T}: steps {r.steps} (oracle {int(gold[f'steps_{k}'])}), max |d rho| {diff:.2e}", flush=True)
        assert r.steps == int(gold[f"steps_{k}"])
        assert r.time == float(gold[f"time_{k}"])
        assert r.max_troubled_cells == int(gold[f"max_troubled_{k}"])
        assert diff <= 1e-12
        assert np.allclose(r.boundary_flux_integral, gold[f"flux_integral_{k}"], rtol=1e-10, atol=1e-15)
        assert r.min_density == pytest.approx(float(gold[f"min_density_{k}"]), rel=1e-10)
        assert r.min_pressure == pytest.approx(float(gold[f"min_pressure_{k}"]), rel=1e-10)
    s.close()
