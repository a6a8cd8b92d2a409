"""Pin the oracle's pointwise physics, fluxes and limiter functions with the
reference's checks (tests/unit/test_physics.py:38-234)."""

import numpy as np
import pytest

from oracle.physics import (INTERFACE_FLUXES, cartesian_fluxes, conserved_to_primitive, get_interface_flux,
                            get_limiter_function, hllc, mc, minmod, normal_flux, pressure, primitive_to_conserved,
                            roe, rusanov, sound_speed, superbee, van_leer)

GAMMA = 1.4
STATES = ((1.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1), (1.3, 0.7, -0.4, 2.1), (0.4, -1.2, 0.9, 0.3))


def cons(p):
    return primitive_to_conserved(*p, GAMMA)


def test_round_trip_and_sound_speed():
    for p in STATES:
        q = cons(p)
        assert np.allclose(conserved_to_primitive(q, GAMMA), p, atol=1e-14)
        assert pressure(q, GAMMA) == pytest.approx(p[3], abs=1e-14)
        assert sound_speed(p[0], p[3], GAMMA) == pytest.approx(np.sqrt(GAMMA * p[3] / p[0]), abs=1e-14)


def test_normal_flux_is_projected_cartesian_flux():
    for p in STATES:
        q = cons(p)
        fx, fy = cartesian_fluxes(q, GAMMA)
        for a in np.linspace(0.0, 2 * np.pi, 7):
            nx, ny = np.cos(a), np.sin(a)
            assert np.allclose(normal_flux(q, nx, ny, GAMMA), fx * nx + fy * ny, atol=1e-13)


def test_cartesian_closed_form():
    rho, u, v, p = 1.3, 0.7, -0.4, 2.1
    q = cons((rho, u, v, p))
    e = q[3]
    fx, fy = cartesian_fluxes(q, GAMMA)
    assert np.allclose(fx, [rho * u, rho * u * u + p, rho * u * v, (e + p) * u], atol=1e-14)
    assert np.allclose(fy, [rho * v, rho * u * v, rho * v * v + p, (e + p) * v], atol=1e-14)


@pytest.mark.parametrize("name", sorted(INTERFACE_FLUXES))
def test_consistency_conservation_rotation(name):
    flux = get_interface_flux(name)
    for p in STATES:
        q = cons(p)
        for nx, ny in ((1.0, 0.0), (0.0, 1.0), (0.6, -0.8)):
            assert np.allclose(flux(q, q, nx, ny, GAMMA), normal_flux(q, nx, ny, GAMMA), atol=1e-12)
    for a in STATES:
        for b in STATES:
            qm, qp = cons(a), cons(b)
            for nx, ny in ((1.0, 0.0), (0.6, -0.8)):
                assert np.allclose(flux(qm, qp, nx, ny, GAMMA), -flux(qp, qm, -nx, -ny, GAMMA), atol=1e-12)
    ang = 0.7
    c, s = np.cos(ang), np.sin(ang)

    def rot(q):
        return np.array([q[0], c * q[1] - s * q[2], s * q[1] + c * q[2], q[3]])

    qm, qp = cons(STATES[2]), cons(STATES[3])
    nx, ny = 0.6, -0.8
    assert np.allclose(rot(flux(qm, qp, nx, ny, GAMMA)), flux(rot(qm), rot(qp), c * nx - s * ny, s * nx + c * ny, GAMMA),
                       atol=1e-12)


def test_contact_and_supersonic():
    qm, qp = cons((1.0, 0.0, 0.0, 1.0)), cons((0.125, 0.0, 0.0, 1.0))
    exact = normal_flux(qm, 1.0, 0.0, GAMMA)
    assert np.allclose(roe(qm, qp, 1.0, 0.0, GAMMA), exact, atol=1e-12)
    assert np.allclose(hllc(qm, qp, 1.0, 0.0, GAMMA), exact, atol=1e-12)
    assert not np.allclose(rusanov(qm, qp, 1.0, 0.0, GAMMA), exact, atol=1e-6)
    rg, lg = cons((1.0, 5.0, 0.0, 1.0)), cons((0.8, 5.0, 0.0, 1.2))
    for f in (roe, hllc):
        assert np.allclose(f(rg, lg, 1.0, 0.0, GAMMA), normal_flux(rg, 1.0, 0.0, GAMMA), atol=1e-12)
        assert np.allclose(f(lg, rg, -1.0, 0.0, GAMMA), normal_flux(rg, -1.0, 0.0, GAMMA), atol=1e-12)


def test_rusanov_definition_and_hllc_expansion():
    qm, qp = cons(STATES[2]), cons(STATES[3])
    nx, ny = 0.6, -0.8
    sp = [abs(conserved_to_primitive(q, GAMMA)[1] * nx + conserved_to_primitive(q, GAMMA)[2] * ny)
          + sound_speed(q[0], pressure(q, GAMMA), GAMMA) for q in (qm, qp)]
    exp = 0.5 * (normal_flux(qm, nx, ny, GAMMA) + normal_flux(qp, nx, ny, GAMMA) - max(sp) * (qp - qm))
    assert np.allclose(rusanov(qm, qp, nx, ny, GAMMA), exp, atol=1e-14)
    out = hllc(cons((1.0, -2.0, 0.0, 0.4)), cons((1.0, 2.0, 0.0, 0.4)), 1.0, 0.0, GAMMA)
    assert np.all(np.isfinite(out)) and out[0] == pytest.approx(0.0, abs=1e-12)


def test_unknown_names():
    with pytest.raises(ValueError, match="Unknown interface flux 'godunov'"):
        get_interface_flux("godunov")
    with pytest.raises(ValueError, match="Unknown limiter 'ospre'"):
        get_limiter_function("ospre")


def test_limiter_tvd_and_closed_forms():
    r = np.linspace(-5.0, 5.0, 401)
    for f in (minmod, mc, van_leer, superbee):
        phi = f(r)
        assert np.all(phi[r <= 0] == 0.0)
        assert f(1.0) == pytest.approx(1.0, abs=1e-14)
        pos = r > 0
        assert np.all(phi[pos] <= np.minimum(2 * r[pos], 2.0) + 1e-12) and np.all(phi[pos] >= 0)
    r = np.array([-1.0, 0.0, 0.25, 0.5, 1.0, 2.0, 4.0])
    assert np.allclose(minmod(r), [0, 0, 0.25, 0.5, 1, 1, 1])
    assert np.allclose(mc(r), [0, 0, 0.5, 0.75, 1, 1.5, 2])
    assert np.allclose(van_leer(r), [0, 0, 0.4, 2 / 3, 1, 4 / 3, 1.6])
    assert np.allclose(superbee(r), [0, 0, 0.5, 1, 1, 2, 2])
    un = get_limiter_function("unlimited")(np.array([-3.0, 0.0, 0.5, 7.0]))
    assert np.allclose(un, 1.0) and un.shape == (4,)
