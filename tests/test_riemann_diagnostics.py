"""Pin the test-side verification diagnostics (tests/riemann_diagnostics.py)."""

import numpy as np
import pytest

from riemann_diagnostics import (RiemannSolution, column_density, evaluate_nodal, monomial_powers, shock_location,
                                 split_quadrature)
from sod_exact import sample as scalar_sample


def test_sod_wave_positions_match_adr_0005():
    """ADR 0005 lists head 0.26336, contact 0.68549, shock 0.85043 at t = 0.2."""
    w = RiemannSolution((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.5).wave_positions(0.2)
    assert w["left head"] == pytest.approx(0.26336, abs=1e-5)
    assert w["contact"] == pytest.approx(0.68549, abs=1e-5)
    assert w["right shock"] == pytest.approx(0.85043, abs=1e-5)


def test_vectorised_solution_matches_scalar_solver():
    """Against the separately written per-point solver (tests/sod_exact.py) on
    random data covering shock/rarefaction on either side."""
    rng = np.random.default_rng(4)
    kinds = set()
    for _ in range(40):
        rl, rr = rng.uniform(0.1, 2.0, 2)
        pl, pr = rng.uniform(0.1, 2.0, 2)
        ul, ur = rng.uniform(-1.0, 1.0, 2)
        x0 = rng.uniform(0.3, 0.7)
        sol = RiemannSolution((rl, ul, pl), (rr, ur, pr), x0)
        kinds.add(tuple(sorted(k.split()[1] for k in sol.wave_positions(0.1) if k != "contact")))
        x = np.linspace(0.0, 1.0, 157)
        got = sol.sample(x, 0.17)
        want = scalar_sample(x, 0.17, x0, (rl, ul, pl), (rr, ur, pr))
        for a, b in zip(got, want):
            assert np.allclose(a, b, rtol=1e-10, atol=1e-12)
    assert len(kinds) >= 3  # two shocks, shock + rarefaction, two rarefactions


def test_vacuum_data_rejected():
    with pytest.raises(ValueError, match="vacuum"):
        RiemannSolution((1.0, -5.0, 0.1), (1.0, 5.0, 0.1), 0.5)


def _mesh(nx, ny, seed):
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "cfd-demo_amd"))
    from cfd_demo_amd.mesh import rectangle_mesh

    m = rectangle_mesh(nx, ny, (0.0, 1.0), (0.0, 0.1))
    v = np.array(m.vertices)
    rng = np.random.default_rng(seed)
    inner = (v[:, 0] > 1e-12) & (v[:, 0] < 1 - 1e-12) & (v[:, 1] > 1e-12) & (v[:, 1] < 0.1 - 1e-12)
    v[inner] += 0.2 * rng.uniform(-1, 1, (inner.sum(), 2)) * np.array([1.0 / nx, 0.1 / ny])
    return v[np.asarray(m.elements)]


def test_split_quadrature_is_exact_for_polynomials():
    corners = _mesh(12, 3, 1)
    quad = split_quadrature(corners, [0.1234, 0.5, 0.77777])
    assert quad.area == pytest.approx(0.1, rel=1e-14)
    # x^5 y^3 over [0,1] x [0,0.1]: a degree-8 polynomial, exact for the 10-point rule.
    val = float(np.sum(quad.weight * quad.x ** 5 * quad.y ** 3))
    assert val == pytest.approx((1.0 / 6.0) * (0.1 ** 4 / 4.0), rel=1e-12)
    # every split sub-triangle lies on one side of every breakpoint
    for b in (0.1234, 0.5, 0.77777):
        side = np.sign(quad.x - b)
        assert np.all((side.min(1) >= 0) | (side.max(1) <= 0))


def test_nodal_evaluation_reproduces_degree_four_polynomials():
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "cfd-demo_amd"))
    from cfd_demo_amd.reference_element import build_reference_triangle

    ref = build_reference_triangle(4)
    corners = _mesh(6, 2, 2)
    lam = np.stack([-(ref.r + ref.s) / 2, (1 + ref.r) / 2, (1 + ref.s) / 2], -1)
    xn = np.einsum("nv,kv->nk", lam, corners[..., 0])
    yn = np.einsum("nv,kv->nk", lam, corners[..., 1])
    f = lambda x, y: 1.0 + x - 2.0 * y * x ** 2 + 3.0 * x ** 4 - y ** 3 * x
    state = f(xn, yn)[..., None]
    quad = split_quadrature(corners, [0.3])
    got = evaluate_nodal(state, ref.r, ref.s, quad)[..., 0]
    assert np.allclose(got, f(quad.x, quad.y), rtol=0, atol=1e-12)
    assert len(monomial_powers(4)) == 15


def test_column_average_shock_locator():
    class FakeSolver:
        pass

    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "cfd-demo_amd"))
    from cfd_demo_amd.mesh import rectangle_mesh
    from cfd_demo_amd.reference_element import build_reference_triangle

    s = FakeSolver()
    s.mesh = rectangle_mesh(50, 2, (0.0, 1.0), (0.0, 0.1))
    s.reference = build_reference_triangle(4)
    c = np.asarray(s.mesh.vertices)[np.asarray(s.mesh.elements)]
    cx = c[..., 0].mean(1)
    rho = np.where(cx < 0.84, 0.26, 0.125)
    state = np.broadcast_to(rho[None, :, None], (15, len(cx), 1)).copy()
    avg, face_x = column_density(s, state, 50)
    assert np.allclose(avg[:42], 0.26) and np.allclose(avg[42:], 0.125)
    assert shock_location(avg, face_x, (0.8, 0.9)) == pytest.approx(0.84)
