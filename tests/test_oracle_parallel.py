"""The oracle over worker processes (the CPU baseline's runner) equals the
serial oracle: RHS, limited SSP-RK3 steps and the step report, on a
four-quadrant problem with troubled cells and on a Sod tube with fixed-state
and slip-wall faces."""

import numpy as np
import pytest

from oracle import mesh as om
from oracle.parallel import ParallelOracle
from oracle.physics import primitive_to_conserved
from oracle.solver import EulerSolver, SolverConfig, sod_shock_tube


def _quadrant(n, flux, limiter):
    s = EulerSolver(om.rectangle_mesh(n, n), {t: om.Transmissive() for t in ("left", "right", "bottom", "top")},
                    SolverConfig(flux, limiter, 1.4))
    cx, cy = s.geometry.centroid.T
    st = np.empty((s.mesh.num_elements, 4))
    for (e, nn), prim in {(1, 1): (1.5, 0, 0, 1.5), (0, 1): (0.5323, 1.206, 0, 0.3), (0, 0): (0.138, 1.206, 1.206, 0.029),
                          (1, 0): (0.5323, 0, 1.206, 0.3)}.items():
        st[((cx > .5) == bool(e)) & ((cy > .5) == bool(nn))] = primitive_to_conserved(*prim, 1.4)
    return s, np.broadcast_to(st[None], s.state_shape).copy()


@pytest.mark.parametrize("flux,limiter", [("hllc", "mc"), ("roe", "minmod")])
def test_parallel_steps_equal_serial(flux, limiter):
    s, q = _quadrant(10, flux, limiter)
    with ParallelOracle(s, workers=3) as par:
        r_par, b_par = par.rhs(q)
        r_ser, b_ser = s.rhs(q)
        assert np.array_equal(r_par, r_ser)
        assert np.allclose(b_par, b_ser, rtol=1e-13, atol=1e-15)
        qs = qp = q
        for _ in range(3):
            dt = s.time_step(qs)
            qs, rs = s.step(qs, dt)
            qp, rp = par.step(qp, dt)
            assert np.array_equal(qp, qs)
            assert bool(rp.admissible) == bool(rs.admissible)
            assert int(rp.troubled_cells) == int(rs.troubled_cells)
            assert float(rp.min_density) == float(rs.min_density)
            assert float(rp.min_pressure) == float(rs.min_pressure)
            assert np.allclose(rp.boundary_flux, rs.boundary_flux, rtol=1e-12, atol=1e-16)
        assert int(rs.troubled_cells) > 0


def test_parallel_sod_equals_serial():
    s, q, _ = sod_shock_tube(12, 3, flux="rusanov", limiter="minmod", cfl=0.5)
    with ParallelOracle(s, workers=4) as par:
        for _ in range(2):
            dt = s.time_step(q)
            qs, rs = s.step(q, dt)
            qp, rp = par.step(q, dt)
            assert np.array_equal(qp, qs)
            assert np.allclose(rp.boundary_flux, rs.boundary_flux, rtol=1e-12, atol=1e-16)
            q = qs
