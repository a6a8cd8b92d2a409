"""Write the oracle's Sod trajectory fixture (sod_100x10_trajectory.npz).

The reference's Sod case (cases.py:31-44) at 100 x 10 rectangles, Rusanov
flux, minmod limiter, CFL 0.1, run by the oracle's restatement of
EulerSolver.run (solver.py:167-216) from t = 0 to each of t = 0.01, 0.02 and
0.05 (three runs, as three calls of solver.run). Stored per end time: the
nodal density (15, 2000), the step count, the end time, the boundary-flux
integral, the minima and the most troubled cells of a step.

Usage (CPU, about 20 minutes): python tests/golden/make_sod_trajectory.py
"""

from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.solver import sod_shock_tube  # noqa: E402

END_TIMES = (0.01, 0.02, 0.05)
OUT = Path(__file__).resolve().parent / "sod_100x10_trajectory.npz"


def main() -> None:
    solver, q0, _ = sod_shock_tube(100, 10, flux="rusanov", limiter="minmod", cfl=0.1)
    data = {"end_times": np.array(END_TIMES)}
    for k, T in enumerate(END_TIMES):
        t0 = time.perf_counter()
        r = solver.run(q0, T)
        print(f"t = {T}: {r.steps} steps in {time.perf_counter() - t0:.0f} s", flush=True)
        data[f"density_{k}"] = np.ascontiguousarray(r.state[..., 0])
        data[f"steps_{k}"] = np.int64(r.steps)
        data[f"time_{k}"] = np.float64(r.time)
        data[f"flux_integral_{k}"] = np.asarray(r.boundary_flux_integral, dtype=np.float64)
        data[f"min_density_{k}"] = np.float64(r.min_density)
        data[f"min_pressure_{k}"] = np.float64(r.min_pressure)
        data[f"max_troubled_{k}"] = np.int64(r.max_troubled_cells)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
