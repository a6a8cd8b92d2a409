"""Degree-4 nodal reference triangle (host setup, runs once per process).

Same element as ``src/cfd_demo/reference_element.py:264-316``: warp-and-blend
nodes (Hesthaven & Warburton, alpha_opt = 0.1001 for N = 4), the orthonormal
Koornwinder-Dubiner basis, and the collocated operators Dr, Ds, LIFT, the
cell/face weights, the face-midpoint nodes, the P1 midpoint projection and the
linear midpoint basis. This construction is written for N = 4 only: GLL points
and the equispaced Lagrange warp are closed forms here instead of eigen- and
linear solves, so the tables agree with the reference's to round-off.
"""

from __future__ import annotations

from dataclasses import dataclass
from math import factorial

import numpy as np

DEGREE = 4
NUM_NODES = 15
NUM_FACE_NODES = 5
_ALPHA_OPT = 0.1001  # H&W table 6.1, N = 4 (reference_element.py:23-26)
_GLL4 = np.array([-1.0, -np.sqrt(3.0 / 7.0), 0.0, np.sqrt(3.0 / 7.0), 1.0])
_NODE_TOL = 1e-10


def _jacobi_normalised(x: np.ndarray, a: float, b: float, n: int) -> np.ndarray:
    """Orthonormal Jacobi polynomials P_0..P_n of weight (1-x)^a (1+x)^b, stacked on axis 0."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty((n + 1,) + x.shape)
    h0 = 2.0 ** (a + b + 1) * factorial(int(a)) * factorial(int(b)) / factorial(int(a + b + 1))
    out[0] = 1.0 / np.sqrt(h0)
    if n == 0:
        return out
    h1 = h0 * (a + 1) * (b + 1) / (a + b + 3)
    out[1] = (0.5 * (a + b + 2) * x + 0.5 * (a - b)) / np.sqrt(h1)
    c_prev = 2.0 / (2.0 + a + b) * np.sqrt((a + 1) * (b + 1) / (a + b + 3))
    for k in range(1, n):
        t = 2 * k + a + b
        c_next = 2.0 / (t + 2) * np.sqrt((k + 1) * (k + 1 + a + b) * (k + 1 + a) * (k + 1 + b) / ((t + 1) * (t + 3)))
        shift = -(a * a - b * b) / (t * (t + 2))
        out[k + 1] = ((x - shift) * out[k] - c_prev * out[k - 1]) / c_next
        c_prev = c_next
    return out


def _jacobi_derivative(x, a, b, n):
    """Derivatives of the orthonormal P_0..P_n: d/dx P_k = sqrt(k(k+a+b+1)) P_{k-1}^{(a+1,b+1)}."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros((n + 1,) + x.shape)
    if n >= 1:
        lower = _jacobi_normalised(x, a + 1, b + 1, n - 1)
        for k in range(1, n + 1):
            out[k] = np.sqrt(k * (k + a + b + 1)) * lower[k - 1]
    return out


def _equispaced_warp(r: np.ndarray) -> np.ndarray:
    """GLL-minus-equispaced displacement interpolated at r, divided by 1 - r^2
    on the open interval (H&W ``Warpfactor``), via equispaced Lagrange bases."""
    req = np.linspace(-1.0, 1.0, DEGREE + 1)
    disp = _GLL4 - req
    warp = np.zeros_like(r)
    for k in range(DEGREE + 1):
        basis = np.ones_like(r)
        for m in range(DEGREE + 1):
            if m != k:
                basis = basis * (r - req[m]) / (req[k] - req[m])
        warp = warp + disp[k] * basis
    inside = np.abs(r) < 1.0 - 1.0e-10
    return np.where(inside, warp / np.where(inside, 1.0 - r * r, 1.0), 0.0)


def nodes() -> tuple[np.ndarray, np.ndarray]:
    """Warp-and-blend nodes (r, s), ordered row by row in s, r increasing."""
    l1 = np.array([i / DEGREE for i in range(DEGREE + 1) for _ in range(DEGREE + 1 - i)])
    l3 = np.array([m / DEGREE for i in range(DEGREE + 1) for m in range(DEGREE + 1 - i)])
    l2 = 1.0 - l1 - l3
    x = l3 - l2
    y = (2.0 * l1 - l2 - l3) / np.sqrt(3.0)
    b1 = 4.0 * l2 * l3 * _equispaced_warp(l3 - l2) * (1.0 + (_ALPHA_OPT * l1) ** 2)
    b2 = 4.0 * l1 * l3 * _equispaced_warp(l1 - l3) * (1.0 + (_ALPHA_OPT * l2) ** 2)
    b3 = 4.0 * l1 * l2 * _equispaced_warp(l2 - l1) * (1.0 + (_ALPHA_OPT * l3) ** 2)
    x = x + b1 + np.cos(2 * np.pi / 3) * b2 + np.cos(4 * np.pi / 3) * b3
    y = y + np.sin(2 * np.pi / 3) * b2 + np.sin(4 * np.pi / 3) * b3
    # Equilateral -> reference triangle through barycentric coordinates.
    m1 = (np.sqrt(3.0) * y + 1.0) / 3.0
    m2 = (-3.0 * x - np.sqrt(3.0) * y + 2.0) / 6.0
    m3 = (3.0 * x - np.sqrt(3.0) * y + 2.0) / 6.0
    return -m2 + m3 - m1, -m2 - m3 + m1


def _modes():
    return [(i, j) for i in range(DEGREE + 1) for j in range(DEGREE + 1 - i)]


def _collapsed(r, s):
    top = np.abs(1.0 - s) < 1e-12
    a = np.where(top, -1.0, 2.0 * (1.0 + r) / np.where(top, 1.0, 1.0 - s) - 1.0)
    return a, s


def dubiner(r, s):
    """Orthonormal basis values and (d/dr, d/ds) at (r, s): three (n, 15) arrays."""
    a, b = _collapsed(np.asarray(r, float), np.asarray(s, float))
    pa, dpa = _jacobi_normalised(a, 0, 0, DEGREE), _jacobi_derivative(a, 0, 0, DEGREE)
    v, vr, vs = [], [], []
    for i, j in _modes():
        pb = _jacobi_normalised(b, 2 * i + 1, 0, DEGREE)[j]
        dpb = _jacobi_derivative(b, 2 * i + 1, 0, DEGREE)[j]
        half = 0.5 * (1.0 - b)
        v.append(np.sqrt(2.0) * pa[i] * pb * (1.0 - b) ** i)
        # Chain rule through (a, b); scale 2^(i+1/2) as in H&W GradSimplex2DP.
        ddr = dpa[i] * pb * (half ** (i - 1) if i > 0 else 1.0)
        dds = dpa[i] * pb * 0.5 * (1.0 + a) * (half ** (i - 1) if i > 0 else 1.0)
        tail = dpb * half ** i - (0.5 * i * pb * half ** (i - 1) if i > 0 else 0.0)
        dds = dds + pa[i] * tail
        scale = 2.0 ** (i + 0.5)
        vr.append(scale * ddr)
        vs.append(scale * dds)
    return np.stack(v, -1), np.stack(vr, -1), np.stack(vs, -1)


@dataclass(frozen=True, eq=False)
class ReferenceTriangle:
    """The degree-4 element; field names and shapes follow ``ReferenceTriangle``
    (``reference_element.py:213-261``)."""

    degree: int
    r: np.ndarray
    s: np.ndarray
    vandermonde: np.ndarray
    mass: np.ndarray
    dr: np.ndarray
    ds: np.ndarray
    face_nodes: np.ndarray
    lift: np.ndarray
    cell_weights: np.ndarray
    face_weights: np.ndarray
    face_midpoint_nodes: np.ndarray
    p1_midpoint: np.ndarray
    linear_midpoint_basis: np.ndarray

    @property
    def num_nodes(self) -> int:
        return self.r.size

    @property
    def num_face_nodes(self) -> int:
        return self.face_nodes.shape[1]


def build_reference_triangle(n: int = DEGREE) -> ReferenceTriangle:
    if n != DEGREE:
        raise ValueError("Only the degree-4 element is implemented.")
    r, s = nodes()
    v, vr, vs = dubiner(r, s)
    vinv = np.linalg.inv(v)
    mass = vinv.T @ vinv
    dr, ds = vr @ vinv, vs @ vinv
    on_face = (np.abs(s + 1.0) < _NODE_TOL, np.abs(r + s) < _NODE_TOL, np.abs(r + 1.0) < _NODE_TOL)
    face_nodes = np.stack([np.flatnonzero(m) for m in on_face])
    if face_nodes.shape != (3, NUM_FACE_NODES):
        raise RuntimeError("Reference nodes do not have n + 1 nodes on every face.")
    edge_coord = (r[face_nodes[0]], r[face_nodes[1]], s[face_nodes[2]])
    emat = np.zeros((NUM_NODES, 3 * NUM_FACE_NODES))
    fw = np.zeros((3, NUM_FACE_NODES))
    for f, t in enumerate(edge_coord):
        v1 = _jacobi_normalised(t, 0, 0, DEGREE).T  # (5, 5) Legendre Vandermonde
        m1 = np.linalg.inv(v1 @ v1.T)  # 1-D mass matrix on the edge
        emat[face_nodes[f], f * NUM_FACE_NODES:(f + 1) * NUM_FACE_NODES] = m1
        fw[f] = m1.sum(axis=1)
    lift = v @ (v.T @ emat)
    mid = face_nodes[:, DEGREE // 2]
    linear = [k for k, (i, j) in enumerate(_modes()) if i + j <= 1]
    p1 = v[:, linear] @ vinv[linear, :]
    return ReferenceTriangle(
        degree=DEGREE, r=r, s=s, vandermonde=v, mass=mass, dr=dr, ds=ds, face_nodes=face_nodes, lift=lift,
        cell_weights=mass.sum(axis=0), face_weights=fw, face_midpoint_nodes=mid, p1_midpoint=p1[mid],
        linear_midpoint_basis=np.stack([-s, 1.0 + r + s, -r], axis=-1),
    )
