"""Troubled-cell and positivity limiting (oracle restatement of ``src/cfd_demo/limiting.py``)."""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from .dg import mirror_state
from .mesh import EXACT_STATE, FIXED_STATE, SLIP_WALL
from .physics import pressure

NEIGHBOR_FACTOR = 1.5  # limiting.py:23-24
DETECTION_TOLERANCE = 1e-10  # limiting.py:25-26
DETECTION_COMPONENTS = (0, 3)  # limiting.py:27-28
POSITIVITY_FLOOR = 1e-13  # limiting.py:29-30
BISECTION_STEPS = 60  # limiting.py:30
_COEFFICIENT_TOLERANCE = 1e-12  # limiting.py:31


class LimiterOperators(NamedTuple):
    """``limiting.py:34-48``; face arrays are (3, K)."""

    average_weights: np.ndarray
    midpoint_nodes: np.ndarray
    p1_midpoint: np.ndarray
    linear_basis: np.ndarray
    neighbor: np.ndarray
    neighbor_kind: np.ndarray
    normal_x: np.ndarray
    normal_y: np.ndarray
    fixed_state: np.ndarray
    alpha: np.ndarray
    beta: np.ndarray
    partner_face: np.ndarray


def neighbor_coefficients(conn, geo, shift=None):
    """Cockburn-Shu alpha, beta >= 0 and partner face (``limiting.py:51-94``).

    ``shift`` (K, 3, 2) translates periodic neighbours' centroids (extension).
    """
    cen = geo.centroid
    mid = geo.face_midpoint
    nrm = geo.normal
    nb = conn.neighbor_element
    interior = cen[np.maximum(nb, 0)]
    if shift is not None:
        interior = interior + shift
    off = np.sum((mid - cen[:, None, :]) * nrm, axis=-1, keepdims=True)
    refl = cen[:, None, :] + 2.0 * off * nrm
    ncen = np.where((nb >= 0)[..., None], interior, refl)
    d = ncen - cen[:, None, :]
    tgt = mid - cen[:, None, :]
    K = cen.shape[0]
    alpha = np.zeros((K, 3))
    beta = np.zeros((K, 3))
    partner = -np.ones((K, 3), dtype=np.int64)
    for face in range(3):
        for sh in (1, 2):
            other = (face + sh) % 3
            di, dj, t = d[:, face], d[:, other], tgt[:, face]
            det = di[:, 0] * dj[:, 1] - di[:, 1] * dj[:, 0]
            safe = np.where(det == 0.0, 1.0, det)
            a = (t[:, 0] * dj[:, 1] - t[:, 1] * dj[:, 0]) / safe
            b = (di[:, 0] * t[:, 1] - di[:, 1] * t[:, 0]) / safe
            valid = (det != 0.0) & (a >= -_COEFFICIENT_TOLERANCE) & (b >= -_COEFFICIENT_TOLERANCE)
            ch = valid & (partner[:, face] < 0)
            alpha[ch, face] = np.maximum(a[ch], 0.0)
            beta[ch, face] = np.maximum(b[ch], 0.0)
            partner[ch, face] = other
    if np.any(partner < 0):
        el, fc = np.argwhere(partner < 0)[0]
        raise ValueError(f"No nonnegative Cockburn-Shu neighbor coefficients for face {fc} of element {el}.")
    return alpha, beta, partner


def build_limiter_operators(conn, geo, ref, kind, state, shift=None):
    """``limiting.py:97-121``."""
    alpha, beta, partner = neighbor_coefficients(conn, geo, shift)
    own = np.arange(conn.neighbor_element.shape[0])[:, None]
    nb = np.where(conn.neighbor_element >= 0, conn.neighbor_element, own)
    w = ref.cell_weights / ref.cell_weights.sum()
    return LimiterOperators(
        average_weights=w, midpoint_nodes=np.array(ref.face_midpoint_nodes), p1_midpoint=np.array(ref.p1_midpoint),
        linear_basis=np.array(ref.linear_midpoint_basis), neighbor=np.ascontiguousarray(nb.T),
        neighbor_kind=np.ascontiguousarray(kind.T), normal_x=np.ascontiguousarray(geo.normal[..., 0].T),
        normal_y=np.ascontiguousarray(geo.normal[..., 1].T), fixed_state=np.ascontiguousarray(np.moveaxis(state, 1, 0)),
        alpha=np.ascontiguousarray(alpha.T), beta=np.ascontiguousarray(beta.T),
        partner_face=np.ascontiguousarray(partner.T),
    )


def _minmod(a, b):
    """``limiting.py:124-125``."""
    return np.where(a * b > 0.0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0)


def neighbor_averages(avg, ops, avg_all=None):
    """``limiting.py:128-133``. A TRANSMISSIVE face (kind 3) keeps the own average
    as ghost, which is what ``neighbor`` (= own element) already gives.
    ``avg_all`` (CPU-baseline element chunks): the averages of every element,
    which ``neighbor`` indexes when ``avg`` and ``ops`` cover a chunk."""
    interior = (avg if avg_all is None else avg_all)[ops.neighbor]
    wall = mirror_state(avg[None], ops.normal_x, ops.normal_y)
    kind = ops.neighbor_kind[..., None]
    given = (kind == FIXED_STATE) | (kind == EXACT_STATE)
    return np.where(kind == SLIP_WALL, wall, np.where(given, ops.fixed_state, interior))


def troubled_cell_limit(q, avg, ops, phi, avg_all=None):
    """Cockburn-Shu troubled-cell limiter (``limiting.py:136-163``)."""
    diff = neighbor_averages(avg, ops, avg_all) - avg[None]
    paired = np.take_along_axis(diff, ops.partner_face[..., None], axis=0)
    sn = NEIGHBOR_FACTOR * (ops.alpha[..., None] * diff + ops.beta[..., None] * paired)
    trace = q[ops.midpoint_nodes] - avg[None]
    comps = np.asarray(DETECTION_COMPONENTS)
    excess = np.abs(_minmod(trace, sn) - trace)[..., comps]
    thr = DETECTION_TOLERANCE * np.abs(avg[:, comps])
    troubled = np.any(excess > thr[None], axis=(0, 2))
    proj = np.einsum("fi,ikc->fkc", ops.p1_midpoint, q) - avg[None]
    nz = proj != 0.0
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ratio = sn / np.where(nz, proj, 1.0)
        deltas = np.where(nz, proj * np.minimum(1.0, phi(ratio)), 0.0)
        pos = np.sum(np.maximum(deltas, 0.0), axis=0)
        neg = np.sum(np.maximum(-deltas, 0.0), axis=0)
        tp = np.where(pos > 0.0, np.minimum(1.0, neg / np.where(pos > 0.0, pos, 1.0)), 0.0)
        tn = np.where(neg > 0.0, np.minimum(1.0, pos / np.where(neg > 0.0, neg, 1.0)), 0.0)
    deltas = tp[None] * np.maximum(deltas, 0.0) - tn[None] * np.maximum(-deltas, 0.0)
    linear = avg[None] + np.einsum("if,fkc->ikc", ops.linear_basis, deltas)
    return np.where(troubled[None, :, None], linear, q), troubled


def positivity_limit(q, avg, gamma, return_theta=False):
    """Zhang-Shu scaling (``limiting.py:166-200``), including the global cond.

    ``return_theta`` also returns the overall scale (K,) applied to q - avg
    (1 where nothing changed), the value the adjoint holds fixed."""
    rbar = avg[:, 0]
    pbar = pressure(avg, gamma)
    eps = np.minimum(POSITIVITY_FLOOR, np.minimum(rbar, pbar))
    dev = q - avg[None]
    rmin = np.min(q[..., 0], axis=0)
    need_r = rmin < eps
    span = np.where(need_r, rbar - rmin, 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        theta = np.clip(np.where(need_r, (rbar - eps) / span, 1.0), 0.0, 1.0)
    q = np.where(need_r[None, :, None], avg[None] + theta[None, :, None] * dev, q)
    theta_r = theta
    need_p = pressure(q, gamma) < eps[None]
    if not np.any(need_p):
        return (q, theta_r) if return_theta else q
    dev = q - avg[None]
    lo = np.zeros(q.shape[:2])
    hi = np.ones(q.shape[:2])
    for _ in range(BISECTION_STEPS):
        mid = 0.5 * (lo + hi)
        ok = pressure(avg[None] + mid[..., None] * dev, gamma) >= eps[None]
        lo, hi = np.where(ok, mid, lo), np.where(ok, hi, mid)
    theta = np.min(np.where(need_p, lo, 1.0), axis=0)
    lim = avg[None] + theta[None, :, None] * dev
    out = np.where((theta < 1.0)[None, :, None], lim, q)
    return (out, theta_r * np.where(theta < 1.0, theta, 1.0)) if return_theta else out


def apply_limiters(q, ops, phi, gamma, limit_slopes, return_flags=False, avg_all=None):
    """``limiting.py:203-216``; optionally also returns the troubled flags (K,)."""
    avg = np.einsum("i,ikc->kc", ops.average_weights, q)
    if limit_slopes:
        q, troubled = troubled_cell_limit(q, avg, ops, phi, avg_all)
        count = np.sum(troubled)
    else:
        troubled = np.zeros(q.shape[1], dtype=bool)
        count = np.int32(0)
    with np.errstate(invalid="ignore"):
        q = positivity_limit(q, avg, gamma)
    if return_flags:
        return q, avg, count, troubled
    return q, avg, count
