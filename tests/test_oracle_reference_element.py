"""Pin the oracle's reference element (and the product's independent
construction) with the reference's own closed-form checks
(tests/unit/test_reference_element.py:23-131)."""

from math import comb, factorial

import numpy as np
import pytest

from cfd_demo_amd.reference_element import build_reference_triangle as product_triangle
from oracle.reference_element import (DEGREE, build_reference_triangle, equispaced_nodes, gauss_lobatto_nodes,
                                      vandermonde_2d, warp_blend_nodes)

BUILDERS = {"oracle": build_reference_triangle, "product": product_triangle}


@pytest.fixture(scope="module", params=list(BUILDERS))
def ref(request):
    return BUILDERS[request.param]()


def test_counts(ref):
    assert ref.degree == 4 and ref.num_nodes == 15 and ref.num_face_nodes == 5


def test_nodes_in_triangle_with_one_node_per_corner(ref):
    r, s = ref.r, ref.s
    assert np.all(r >= -1 - 1e-12) and np.all(s >= -1 - 1e-12) and np.all(r + s <= 1e-12)
    for corner in ([-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]):
        assert np.isclose(np.hypot(r - corner[0], s - corner[1]), 0.0, atol=1e-12).sum() == 1


def test_face_nodes_are_gll(ref):
    expected = gauss_lobatto_nodes(DEGREE)
    assert np.allclose(expected, [-1.0, -np.sqrt(3 / 7), 0.0, np.sqrt(3 / 7), 1.0])
    r, s = ref.r, ref.s
    for coord in (r[ref.face_nodes[0]], r[ref.face_nodes[1]], s[ref.face_nodes[2]]):
        assert np.allclose(np.sort(coord), expected, atol=1e-14)
    assert np.array_equal(ref.face_nodes, [[0, 1, 2, 3, 4], [4, 8, 11, 13, 14], [0, 5, 9, 12, 14]])
    assert np.array_equal(ref.face_midpoint_nodes, [2, 11, 9])


def test_midpoints(ref):
    mid = np.stack([ref.r[ref.face_midpoint_nodes], ref.s[ref.face_midpoint_nodes]], -1)
    assert np.allclose(mid, [[0.0, -1.0], [0.0, 0.0], [-1.0, 0.0]], atol=1e-14)


def test_differentiation_exact_for_degree_four(ref):
    r, s = ref.r, ref.s
    worst = 0.0
    for i in range(5):
        for j in range(5 - i):
            f = r ** i * s ** j
            dr = i * r ** max(i - 1, 0) * s ** j if i else np.zeros_like(r)
            ds = j * r ** i * s ** max(j - 1, 0) if j else np.zeros_like(r)
            worst = max(worst, np.abs(ref.dr @ f - dr).max(), np.abs(ref.ds @ f - ds).max())
    assert worst < 1e-12


def _monomial_integral(i, j):
    total = 0.0
    for m in range(i + 1):
        for n in range(j + 1):
            coef = comb(i, m) * comb(j, n) * (-1.0) ** (i + j - m - n) * 2.0 ** (m + n)
            total += coef * factorial(m) * factorial(n) / factorial(m + n + 2) * 4.0
    return total


def test_cell_weights_integrate_exactly(ref):
    assert ref.cell_weights.sum() == pytest.approx(2.0, abs=1e-13)
    for i in range(5):
        for j in range(5 - i):
            assert ref.cell_weights @ (ref.r ** i * ref.s ** j) == pytest.approx(_monomial_integral(i, j), abs=1e-12)


def test_lift_reproduces_edge_weights(ref):
    assert np.allclose(ref.cell_weights @ ref.lift, ref.face_weights.reshape(-1), atol=1e-13)
    assert np.allclose(ref.face_weights.sum(axis=1), 2.0)


def test_warp_blend_beats_equispaced():
    assert np.linalg.cond(vandermonde_2d(4, *warp_blend_nodes(4))) < np.linalg.cond(
        vandermonde_2d(4, *equispaced_nodes(4)))


def test_linear_basis_partition_of_unity(ref):
    assert np.allclose(ref.linear_midpoint_basis.sum(axis=1), 1.0)
    assert np.allclose(ref.linear_midpoint_basis[ref.face_midpoint_nodes], np.eye(3), atol=1e-14)


def test_p1_projection_is_a_projection(ref):
    r, s = ref.r, ref.s
    for field in (np.ones_like(r), r, s, 2.0 - 0.5 * r + 3.0 * s):
        assert np.allclose(ref.p1_midpoint @ field, field[ref.face_midpoint_nodes], atol=1e-12)
    assert not np.allclose(ref.p1_midpoint @ (r * r), (r * r)[ref.face_midpoint_nodes])


def test_product_and_oracle_elements_agree():
    p, o = product_triangle(), build_reference_triangle()
    for name, tol in (("r", 1e-15), ("s", 1e-15), ("dr", 1e-13), ("ds", 1e-13), ("lift", 1e-13),
                      ("cell_weights", 1e-15), ("face_weights", 1e-15), ("p1_midpoint", 1e-15)):
        a, b = getattr(p, name), getattr(o, name)
        assert np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max()), name
