"""Pin the oracle's mesh and DG operator with the reference's invariants
(tests/unit/test_mesh.py, tests/unit/test_dg_operator.py) and an independent
per-element loop restatement on tiny meshes."""

import numpy as np
import pytest

from oracle import mesh as om
from oracle.dg import dg_rhs, dg_rhs_loop
from oracle.physics import INTERFACE_FLUXES, get_interface_flux, primitive_to_conserved
from oracle.solver import EulerSolver, SolverConfig

GAMMA = 1.4
FREE = (1.2, 0.4, -0.3, 0.9)


def perturbed(nx, ny, amp=0.25, seed=0):
    m = om.rectangle_mesh(nx, ny)
    v = np.array(m.vertices)
    inner = np.all((v > 1e-12) & (v < 1 - 1e-12), axis=1)
    d = np.random.default_rng(seed).uniform(-1, 1, size=v.shape)
    v[inner] += amp * d[inner] / np.array([nx, ny])
    return om.TriangleMesh(vertices=v, elements=m.elements, boundary_edges=m.boundary_edges)


def fixed_all(prim=FREE):
    c = om.FixedState.from_primitive(*prim, GAMMA)
    return {t: c for t in ("left", "right", "bottom", "top")}


def solver(mesh, bcs, flux="rusanov", limiter="minmod"):
    return EulerSolver(mesh, bcs, SolverConfig(flux=flux, limiter=limiter, gamma=GAMMA))


def uniform(s, prim):
    return np.broadcast_to(primitive_to_conserved(*prim, GAMMA), s.state_shape).copy()


def test_mesh_counts_orientation_ordering():
    m = om.rectangle_mesh(4, 3, (0.0, 2.0), (0.0, 0.5))
    assert m.elements.shape == (24, 3) and np.all(om.signed_areas(m) > 0)
    assert om.signed_areas(m).sum() == pytest.approx(1.0)
    g = om.element_geometry(om.rectangle_mesh(3, 2))
    for j in range(2):
        for i in range(3):
            lo, up = 2 * (j * 3 + i), 2 * (j * 3 + i) + 1
            for e in (lo, up):
                assert i / 3 < g.centroid[e, 0] < (i + 1) / 3 and j / 2 < g.centroid[e, 1] < (j + 1) / 2
            assert g.centroid[lo, 0] > g.centroid[up, 0]


def test_mesh_rejections():
    v = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    with pytest.raises(ValueError, match="counterclockwise"):
        om.TriangleMesh(vertices=v, elements=np.array([[0, 2, 1]]), boundary_edges={})
    with pytest.raises(ValueError, match="vertices that do not exist"):
        om.TriangleMesh(vertices=v, elements=np.array([[0, 1, 7]]), boundary_edges={})
    with pytest.raises(ValueError, match="nx and ny must be positive"):
        om.rectangle_mesh(0, 2)


def test_connectivity_symmetric_normals_opposite_and_nodes_paired():
    m = perturbed(5, 4, seed=2)
    c = om.build_connectivity(m)
    g = om.element_geometry(m)
    K = m.num_elements
    for e in range(K):
        for f in range(3):
            nb, nf = c.neighbor_element[e, f], c.neighbor_face[e, f]
            if nb < 0:
                continue
            assert c.neighbor_element[nb, nf] == e and c.neighbor_face[nb, nf] == f
            assert np.allclose(g.normal[e, f], -g.normal[nb, nf], atol=1e-15)
            assert g.edge_length[e, f] == g.edge_length[nb, nf]
    assert np.allclose((g.normal * g.edge_length[..., None]).sum(axis=1), 0.0, atol=1e-15)
    s = solver(m, fixed_all())
    x, y = s.node_coordinates()
    ext = s.operators.exterior_index
    fn = s.operators.face_nodes
    xs, ys = x.reshape(-1), y.reshape(-1)
    own = fn[:, None] * K + np.arange(K)[None, :]
    assert np.allclose(xs[ext], xs[own], atol=1e-12) and np.allclose(ys[ext], ys[own], atol=1e-12)


def test_inverse_metric_closed_form():
    m = perturbed(3, 3, seed=4)
    g = om.element_geometry(m)
    c = m.vertices[m.elements]
    for e in range(m.num_elements):
        jac = 0.5 * np.array([[c[e, 1, 0] - c[e, 0, 0], c[e, 2, 0] - c[e, 0, 0]],
                              [c[e, 1, 1] - c[e, 0, 1], c[e, 2, 1] - c[e, 0, 1]]])
        inv = np.linalg.inv(jac)
        assert np.allclose([[g.rx[e], g.ry[e]], [g.sx[e], g.sy[e]]], inv, rtol=1e-13)


@pytest.mark.parametrize("flux", sorted(INTERFACE_FLUXES))
def test_free_stream_preserved(flux):
    s = solver(perturbed(6, 5), fixed_all(), flux)
    rhs, _ = s.rhs(uniform(s, FREE))
    assert np.abs(rhs).max() < 1e-11


def test_closed_box_and_wall_flux():
    walls = {t: om.SlipWall() for t in ("left", "right", "bottom", "top")}
    s = solver(perturbed(5, 4), walls)
    rhs, b = s.rhs(uniform(s, (1.0, 0.0, 0.0, 1.0)))
    assert np.abs(rhs).max() < 1e-12 and np.abs(b).max() < 1e-12
    s = solver(om.rectangle_mesh(4, 4), walls)
    _, b = s.rhs(uniform(s, (1.0, 0.5, 0.0, 1.0)))
    assert np.abs(b[[0, 2, 3]]).max() < 1e-12
    assert b[1] == pytest.approx(2.0 * (0.5 + np.sqrt(GAMMA)) * 0.5, abs=1e-12)


@pytest.mark.parametrize("flux", sorted(INTERFACE_FLUXES))
def test_discrete_conservation(flux):
    s = solver(perturbed(5, 4, seed=3), fixed_all(), flux)
    prim = np.random.default_rng(11).uniform([0.5, -0.4, -0.4, 0.5], [1.5, 0.4, 0.4, 1.5], size=s.state_shape[:2] + (4,))
    q = primitive_to_conserved(*(prim[..., k] for k in range(4)), GAMMA)
    rhs, b = s.rhs(q)
    assert np.abs(s.integrate(rhs) + b).max() < 1e-12 and np.abs(b).max() > 1e-3


@pytest.mark.parametrize("flux", sorted(INTERFACE_FLUXES))
@pytest.mark.parametrize("bcname", ["fixed", "walls", "trans", "periodic"])
def test_vectorised_rhs_equals_loop_restatement(flux, bcname):
    if bcname == "periodic":
        m = om.rectangle_mesh(3, 2)
        bcs = {"left": om.Periodic("right", (-1.0, 0.0)), "right": om.Periodic("left", (1.0, 0.0)),
               "bottom": om.Periodic("top", (0.0, -1.0)), "top": om.Periodic("bottom", (0.0, 1.0))}
    else:
        m = perturbed(3, 2, seed=1)
        bcs = {"fixed": fixed_all(), "walls": {t: om.SlipWall() for t in ("left", "right", "bottom", "top")},
               "trans": {t: om.Transmissive() for t in ("left", "right", "bottom", "top")}}[bcname]
    s = solver(m, bcs, flux)
    prim = np.random.default_rng(5).uniform([0.5, -0.4, -0.4, 0.5], [1.5, 0.4, 0.4, 1.5], size=s.state_shape[:2] + (4,))
    q = primitive_to_conserved(*(prim[..., k] for k in range(4)), GAMMA)
    a, ba = dg_rhs(q, s.operators, get_interface_flux(flux), GAMMA)
    b, bb = dg_rhs_loop(q, s.operators, get_interface_flux(flux), GAMMA)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    assert np.allclose(ba, bb, rtol=1e-12, atol=1e-14)


def _smooth_prim(x, y):
    tp = 2 * np.pi
    return (1.0 + 0.2 * np.sin(tp * x) * np.sin(tp * y), 0.3 + 0.1 * np.cos(tp * x) * np.sin(tp * y),
            -0.2 + 0.1 * np.sin(tp * x) * np.cos(tp * y), 1.0 + 0.15 * np.sin(tp * (x + y)))


def _exact_rhs(x, y, h=1e-4):
    def flux(x, y, axis):
        rho, u, v, p = _smooth_prim(x, y)
        e = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
        vel = u if axis == 0 else v
        return np.stack([rho * vel, rho * u * vel + (p if axis == 0 else 0.0), rho * v * vel + (p if axis == 1 else 0.0),
                         (e + p) * vel], -1)

    tot = 0.0
    for axis in (0, 1):
        for st, w in ((-2, 1 / 12), (-1, -2 / 3), (1, 2 / 3), (2, -1 / 12)):
            tot = tot + w * flux(x + (st * h if axis == 0 else 0), y + (st * h if axis == 1 else 0), axis) / h
    return -tot


def test_fourth_order_consistency():
    errs = []
    for n in (8, 16, 32):
        s = solver(om.rectangle_mesh(n, n), fixed_all())
        x, y = s.node_coordinates()
        rhs, _ = s.rhs(primitive_to_conserved(*_smooth_prim(x, y), GAMMA))
        inner = np.all(s.connectivity.neighbor_element >= 0, axis=1)
        errs.append(float(np.abs(rhs[:, inner] - _exact_rhs(x, y)[:, inner]).max()))
    assert errs[0] < 5e-2
    orders = [np.log2(a / b) for a, b in zip(errs, errs[1:])]
    assert all(o > 3.5 for o in orders), (errs, orders)


def test_integrate_and_averages_agree():
    s = solver(perturbed(4, 3), fixed_all())
    q = uniform(s, FREE)
    c = primitive_to_conserved(*FREE, GAMMA)
    assert np.allclose(s.cell_averages(q), c) and np.allclose(s.integrate(q), c)
    assert np.allclose((s.cell_averages(q) * s.geometry.area[:, None]).sum(0), s.integrate(q))
