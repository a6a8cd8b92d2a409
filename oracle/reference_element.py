"""Degree-4 nodal reference triangle (oracle restatement).

Restates ``src/cfd_demo/reference_element.py`` (Hesthaven & Warburton,
*Nodal Discontinuous Galerkin Methods*, 2008, sections 3.1-3.2 and 6.1-6.2):
warp-and-blend nodes, the orthonormal Koornwinder-Dubiner Vandermonde, and the
collocated operators Dr, Ds, LIFT. Reference triangle vertices are (-1,-1),
(1,-1), (-1,1); face f joins vertices f and f+1 (mod 3).
"""

from __future__ import annotations

from dataclasses import dataclass
from math import gamma as _gamma_fn

import numpy as np

DEGREE = 4  # reference_element.py:18
NODE_TOLERANCE = 1e-10  # reference_element.py:19
# Hesthaven & Warburton table 6.1, alpha_opt by degree (reference_element.py:23-26).
WARP_BLEND_ALPHA = (0.0, 0.0, 0.0, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999,
                    1.2832, 1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258)


def jacobi_p(x, a: float, b: float, n: int) -> np.ndarray:
    """Orthonormal Jacobi polynomial P_n^(a,b)(x) by the three-term recurrence.

    Follows ``reference_element.py:29-54`` (H&W ``JacobiP``).
    """
    x = np.asarray(x, dtype=np.float64)
    g0 = 2.0 ** (a + b + 1) / (a + b + 1) * _gamma_fn(a + 1) * _gamma_fn(b + 1) / _gamma_fn(a + b + 1)
    prev = np.full_like(x, 1.0 / np.sqrt(g0))
    if n == 0:
        return prev
    g1 = (a + 1) * (b + 1) / (a + b + 3) * g0
    cur = ((a + b + 2) * x / 2 + (a - b) / 2) / np.sqrt(g1)
    aold = 2 / (2 + a + b) * np.sqrt((a + 1) * (b + 1) / (a + b + 3))
    for i in range(1, n):
        h1 = 2 * i + a + b
        anew = 2 / (h1 + 2) * np.sqrt((i + 1) * (i + 1 + a + b) * (i + 1 + a) * (i + 1 + b) / (h1 + 1) / (h1 + 3))
        bnew = -(a * a - b * b) / h1 / (h1 + 2)
        prev, cur = cur, (-aold * prev + (x - bnew) * cur) / anew
        aold = anew
    return cur


def grad_jacobi_p(x, a: float, b: float, n: int) -> np.ndarray:
    """d/dx of the orthonormal Jacobi polynomial (``reference_element.py:57-62``)."""
    x = np.asarray(x, dtype=np.float64)
    if n == 0:
        return np.zeros_like(x)
    return np.sqrt(n * (n + a + b + 1)) * jacobi_p(x, a + 1, b + 1, n - 1)


def jacobi_gauss(a: float, b: float, n: int):
    """Gauss-Jacobi quadrature, n+1 points, Golub-Welsch (``reference_element.py:65-85``)."""
    if n == 0:
        return np.array([(a - b) / (a + b + 2)]), np.array([2.0])
    h1 = 2.0 * np.arange(n + 1) + a + b
    k = np.arange(1, n + 1, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        diag = -0.5 * (a * a - b * b) / (h1 + 2) / h1
    if a + b < 10 * np.finfo(np.float64).eps:
        diag[0] = 0.0
    off = 2 / (h1[:n] + 2) * np.sqrt(k * (k + a + b) * (k + a) * (k + b) / (h1[:n] + 1) / (h1[:n] + 3))
    jac = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    nodes, vecs = np.linalg.eigh(jac)
    w = vecs[0, :] ** 2 * 2 ** (a + b + 1) / (a + b + 1) * _gamma_fn(a + 1) * _gamma_fn(b + 1) / _gamma_fn(a + b + 1)
    return nodes, w


def gauss_lobatto_nodes(n: int) -> np.ndarray:
    """GLL nodes on [-1, 1] (``reference_element.py:88-93``)."""
    if n == 1:
        return np.array([-1.0, 1.0])
    inner, _ = jacobi_gauss(1.0, 1.0, n - 2)
    return np.concatenate(([-1.0], inner, [1.0]))


def vandermonde_1d(n: int, r) -> np.ndarray:
    """Legendre Vandermonde (``reference_element.py:96-99``)."""
    r = np.asarray(r, dtype=np.float64)
    return np.stack([jacobi_p(r, 0.0, 0.0, j) for j in range(n + 1)], axis=-1)


def _warp_factor(n: int, rout: np.ndarray) -> np.ndarray:
    """H&W ``Warpfactor`` (``reference_element.py:102-113``)."""
    lgl = gauss_lobatto_nodes(n)
    req = np.linspace(-1.0, 1.0, n + 1)
    veq = vandermonde_1d(n, req)
    pmat = np.stack([jacobi_p(rout, 0.0, 0.0, i) for i in range(n + 1)])
    lmat = np.linalg.solve(veq.T, pmat)
    warp = lmat.T @ (lgl - req)
    zerof = (np.abs(rout) < 1.0 - 1.0e-10).astype(np.float64)
    sf = 1.0 - (zerof * rout) ** 2
    return warp / sf + warp * (zerof - 1.0)


def _lattice(n: int):
    """Barycentric lattice in the reference's row order (``reference_element.py:124-132``)."""
    l1 = np.array([i / n for i in range(n + 1) for _ in range(n + 1 - i)])
    l3 = np.array([m / n for i in range(n + 1) for m in range(n + 1 - i)])
    return l1, 1.0 - l1 - l3, l3


def warp_blend_nodes(n: int = DEGREE):
    """H&W ``Nodes2D`` then ``xytors`` (``reference_element.py:116-149``)."""
    alpha = WARP_BLEND_ALPHA[n] if n < len(WARP_BLEND_ALPHA) else 5.0 / 3.0
    l1, l2, l3 = _lattice(n)
    x = -l2 + l3
    y = (-l2 - l3 + 2 * l1) / np.sqrt(3.0)
    w1 = 4 * l2 * l3 * _warp_factor(n, l3 - l2) * (1 + (alpha * l1) ** 2)
    w2 = 4 * l1 * l3 * _warp_factor(n, l1 - l3) * (1 + (alpha * l2) ** 2)
    w3 = 4 * l1 * l2 * _warp_factor(n, l2 - l1) * (1 + (alpha * l3) ** 2)
    x = x + w1 + np.cos(2 * np.pi / 3) * w2 + np.cos(4 * np.pi / 3) * w3
    y = y + np.sin(2 * np.pi / 3) * w2 + np.sin(4 * np.pi / 3) * w3
    b1 = (np.sqrt(3.0) * y + 1.0) / 3.0
    b2 = (-3.0 * x - np.sqrt(3.0) * y + 2.0) / 6.0
    b3 = (3.0 * x - np.sqrt(3.0) * y + 2.0) / 6.0
    return -b2 + b3 - b1, -b2 - b3 + b1


def equispaced_nodes(n: int = DEGREE):
    """Equispaced lattice in the same order (``reference_element.py:152-156``)."""
    l1, _, l3 = _lattice(n)
    return 2 * l3 - 1, 2 * l1 - 1


def _rs_to_ab(r, s):
    """Collapsed coordinates (``reference_element.py:159-166``)."""
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    top = np.abs(1.0 - s) < 1e-12
    a = np.where(top, -1.0, 2.0 * (1.0 + r) / np.where(top, 1.0, 1.0 - s) - 1.0)
    return a, s


def mode_indices(n: int = DEGREE):
    """(i, j) column order of the Vandermonde (``reference_element.py:169-171``)."""
    return [(i, j) for i in range(n + 1) for j in range(n + 1 - i)]


def vandermonde_2d(n: int, r, s) -> np.ndarray:
    """Koornwinder-Dubiner Vandermonde (``reference_element.py:174-175,197-200``)."""
    a, b = _rs_to_ab(r, s)
    cols = [np.sqrt(2.0) * jacobi_p(a, 0.0, 0.0, i) * jacobi_p(b, 2 * i + 1.0, 0.0, j) * (1 - b) ** i
            for i, j in mode_indices(n)]
    return np.stack(cols, axis=-1)


def grad_vandermonde_2d(n: int, r, s):
    """Gradient Vandermonde (``reference_element.py:178-194,203-210``)."""
    a, b = _rs_to_ab(r, s)
    vr, vs = [], []
    for i, j in mode_indices(n):
        fa, dfa = jacobi_p(a, 0.0, 0.0, i), grad_jacobi_p(a, 0.0, 0.0, i)
        gb, dgb = jacobi_p(b, 2 * i + 1.0, 0.0, j), grad_jacobi_p(b, 2 * i + 1.0, 0.0, j)
        hb = 0.5 * (1 - b)
        ddr = dfa * gb
        dds = dfa * (gb * (0.5 * (1 + a)))
        if i > 0:
            ddr = ddr * hb ** (i - 1)
            dds = dds * hb ** (i - 1)
        tmp = dgb * hb ** i
        if i > 0:
            tmp = tmp - 0.5 * i * gb * hb ** (i - 1)
        dds = dds + fa * tmp
        scale = 2.0 ** (i + 0.5)
        vr.append(scale * ddr)
        vs.append(scale * dds)
    return np.stack(vr, axis=-1), np.stack(vs, axis=-1)


@dataclass(frozen=True, eq=False)
class ReferenceTriangle:
    """Fields of ``ReferenceTriangle`` (``reference_element.py:213-261``)."""

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
    """``build_reference_triangle`` (``reference_element.py:264-316``)."""
    r, s = warp_blend_nodes(n)
    v = vandermonde_2d(n, r, s)
    vinv = np.linalg.inv(v)
    mass = vinv.T @ vinv
    vr, vs = grad_vandermonde_2d(n, r, s)
    dr = vr @ vinv
    ds = vs @ vinv
    masks = (np.abs(s + 1) < NODE_TOLERANCE, np.abs(r + s) < NODE_TOLERANCE, np.abs(r + 1) < NODE_TOLERANCE)
    face_nodes = np.stack([np.flatnonzero(m) for m in masks])
    if face_nodes.shape != (3, n + 1):
        raise RuntimeError("Reference nodes do not have n + 1 nodes on every face.")
    coords = (r[face_nodes[0]], r[face_nodes[1]], s[face_nodes[2]])
    nfp = n + 1
    emat = np.zeros((r.size, 3 * nfp))
    fw = np.zeros((3, nfp))
    for f, c in enumerate(coords):
        v1 = vandermonde_1d(n, c)
        m1 = np.linalg.inv(v1 @ v1.T)
        emat[face_nodes[f], f * nfp:(f + 1) * nfp] = m1
        fw[f] = m1.sum(axis=1)
    lift = v @ (v.T @ emat)
    mid = face_nodes[:, n // 2]
    lin = [k for k, (i, j) in enumerate(mode_indices(n)) if i + j <= 1]
    p1 = v[:, lin] @ vinv[lin, :]
    return ReferenceTriangle(
        degree=n, r=r, s=s, vandermonde=v, mass=mass, dr=dr, ds=ds, face_nodes=face_nodes,
        lift=lift, cell_weights=mass.sum(axis=0), face_weights=fw, face_midpoint_nodes=mid,
        p1_midpoint=p1[mid, :], linear_midpoint_basis=np.stack([-s, 1.0 + r + s, -r], axis=-1),
    )
