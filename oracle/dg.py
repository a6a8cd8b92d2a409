"""Strong-form nodal DG operator (oracle restatement of ``src/cfd_demo/dg.py``).

du/dt = -(Dr F~ + Ds G~) + LIFT (Fscale (F(u-) . n - F^)), F~ = rx F + ry G,
G~ = sx F + sy G. Boundary conditions only supply exterior states.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from .mesh import EXACT_STATE, FIXED_STATE, PERIODIC, SLIP_WALL, TRANSMISSIVE, node_coordinates, pair_face_nodes, \
    reflected_element
from .physics import cartesian_fluxes


class DGOperators(NamedTuple):
    """``dg.py:30-49``; face-node arrays are (3 Nfp, K)."""

    differentiation: np.ndarray
    lift: np.ndarray
    face_nodes: np.ndarray
    inverse_metric: np.ndarray
    normal_x: np.ndarray
    normal_y: np.ndarray
    face_scale: np.ndarray
    exterior_index: np.ndarray
    boundary_kind: np.ndarray
    boundary_state: np.ndarray
    boundary_weight: np.ndarray
    cell_weights: np.ndarray
    jacobian: np.ndarray


def resolve_boundary_conditions(conn, boundary_conditions):
    """Per-face kinds (K, 3) and fixed states (K, 3, 4) (``dg.py:52-73``).

    Extension: TRANSMISSIVE and PERIODIC kinds are accepted too (periodic
    faces must already be patched out of ``conn`` by ``periodic_connectivity``).
    """
    tags = set(conn.boundary_tags)
    missing = sorted(tags - set(boundary_conditions))
    if missing:
        raise ValueError(f"Boundary tags without a boundary condition: {missing}.")
    unknown = sorted(set(boundary_conditions) - tags)
    if unknown:
        raise ValueError(f"Boundary conditions for tags that are not in the mesh: {unknown}.")
    kind = np.zeros(conn.face_tag.shape, dtype=np.int32)
    state = np.zeros(conn.face_tag.shape + (4,))
    for ti, tag in enumerate(conn.boundary_tags):
        bc = boundary_conditions[tag]
        mask = conn.face_tag == ti
        if bc.kind == PERIODIC:
            continue
        kind[mask] = bc.kind
        if bc.kind == FIXED_STATE:
            state[mask] = bc.state
        elif bc.kind not in (SLIP_WALL, TRANSMISSIVE, EXACT_STATE):
            raise ValueError(f"Unsupported boundary condition for tag {tag!r}: {bc!r}.")
    return kind, state


def exact_boundary_states(mesh, conn, ref, kind, state, boundary_conditions):
    """Per-face-node exterior states (K, 3, Nfp, 4) and limiter ghost
    averages (K, 3, 4): the fixed states broadcast over their face nodes, and
    on each ExactState face its state function at the physical face nodes and
    its average over the element reflected across the face."""
    nfp = ref.num_face_nodes
    node_state = np.repeat(state[:, :, None, :], nfp, axis=2)
    ghost = state.copy()
    x, y = node_coordinates(mesh, ref.r, ref.s)
    fn = np.asarray(ref.face_nodes).reshape(3, nfp)
    corners = mesh.vertices[mesh.elements]
    for ti, tag in enumerate(conn.boundary_tags):
        bc = boundary_conditions[tag]
        if bc.kind != EXACT_STATE:
            continue
        for e, f in zip(*np.nonzero(conn.face_tag == ti)):
            node_state[e, f] = np.asarray(bc.state(x[fn[f], e], y[fn[f], e]), dtype=np.float64)
            ghost[e, f] = np.asarray(bc.average(reflected_element(corners[e], f)[None]), dtype=np.float64)[0]
    return node_state, ghost


def build_dg_operators(mesh, conn, geo, ref, kind, state, shift=None):
    """Host assembly of the operator tables (``dg.py:76-116``)."""
    K = mesh.num_elements
    nfp = ref.num_face_nodes
    x, y = node_coordinates(mesh, ref.r, ref.s)
    partner = pair_face_nodes(mesh, conn, x, y, ref.face_nodes, shift=shift)
    own = np.arange(K)[:, None]
    nbr = np.where(conn.neighbor_element >= 0, conn.neighbor_element, own)
    ext = partner * K + nbr.T[:, None, :]

    def per_face_node(vals):
        vals = np.moveaxis(vals, 1, 0)[:, None]
        shp = (3, nfp) + vals.shape[2:]
        return np.broadcast_to(vals, shp).reshape((3 * nfp,) + vals.shape[2:])

    fscale = 0.5 * geo.edge_length / geo.jacobian[:, None]
    is_bnd = per_face_node(kind) > 0
    ew = ref.face_weights.reshape(-1)[:, None] * per_face_node(0.5 * geo.edge_length)
    return DGOperators(
        differentiation=np.concatenate([ref.dr, ref.ds], axis=1),
        lift=np.array(ref.lift),
        face_nodes=ref.face_nodes.reshape(-1),
        inverse_metric=np.stack([geo.rx, geo.ry, geo.sx, geo.sy]),
        normal_x=np.ascontiguousarray(per_face_node(geo.normal[..., 0])),
        normal_y=np.ascontiguousarray(per_face_node(geo.normal[..., 1])),
        face_scale=np.ascontiguousarray(per_face_node(fscale)),
        exterior_index=ext.reshape(3 * nfp, K),
        boundary_kind=np.ascontiguousarray(per_face_node(kind)),
        boundary_state=np.ascontiguousarray(per_face_node(state) if state.ndim == 3
                                            else np.moveaxis(state, 0, 2).reshape(3 * nfp, K, 4)),
        boundary_weight=np.where(is_bnd, ew, 0.0),
        cell_weights=np.array(ref.cell_weights),
        jacobian=np.array(geo.jacobian),
    )


def mirror_state(q, nx, ny):
    """Slip-wall mirror (``dg.py:119-126``)."""
    mn = q[..., 1] * nx + q[..., 2] * ny
    z = np.zeros_like(mn)
    return q - np.stack([z, 2.0 * mn * nx, 2.0 * mn * ny, z], axis=-1)


def exterior_states(qm, qn, ops):
    """``dg.py:129-137``; TRANSMISSIVE (extension) takes the interior trace."""
    kind = ops.boundary_kind[..., None]
    wall = mirror_state(qm, ops.normal_x, ops.normal_y)
    given = (kind == FIXED_STATE) | (kind == EXACT_STATE)
    return np.where(kind == SLIP_WALL, wall,
                    np.where(given, ops.boundary_state,
                             np.where(kind == TRANSMISSIVE, qm, qn)))


def dg_rhs(q, ops, flux, gamma, q_exterior=None):
    """Semi-discrete RHS and net outward boundary flux (``dg.py:140-165``).

    ``q_exterior`` (CPU-baseline element chunks): the whole state the
    exterior traces are gathered from when ``q`` and ``ops`` cover a chunk."""
    nv = q.shape[-1]
    f, g = cartesian_fluxes(q, gamma)
    rx, ry, sx, sy = (row[None, :, None] for row in ops.inverse_metric)
    contra = np.concatenate([rx * f + ry * g, sx * f + sy * g], axis=0)
    divergence = np.einsum("ij,jkc->ikc", ops.differentiation, contra)
    nx, ny = ops.normal_x, ops.normal_y
    qm = q[ops.face_nodes]
    qn = (q if q_exterior is None else q_exterior).reshape(-1, nv)[ops.exterior_index]
    qp = exterior_states(qm, qn, ops)
    fhat = flux(qm, qp, nx, ny, gamma)
    fminus = f[ops.face_nodes] * nx[..., None] + g[ops.face_nodes] * ny[..., None]
    jump = ops.face_scale[..., None] * (fminus - fhat)
    surface = np.einsum("ij,jkc->ikc", ops.lift, jump)
    bflux = np.einsum("jk,jkc->c", ops.boundary_weight, fhat)
    return surface - divergence, bflux


def cell_averages(q, ops):
    """``dg.py:168-171``."""
    w = ops.cell_weights / np.sum(ops.cell_weights)
    return np.einsum("i,ikc->kc", w, q)


def domain_integral(q, ops):
    """``dg.py:174-176``."""
    return np.einsum("k,i,ikc->c", ops.jacobian, ops.cell_weights, q)


def dg_rhs_loop(q, ops, flux, gamma):
    """Independent per-element restatement of ``dg_rhs`` for tiny meshes.

    Explicit loops over elements, nodes and face slots, written from the
    operator formula (not from the vectorised code); used only to cross-check
    ``dg_rhs``.
    """
    from .physics import normal_flux

    Np, K, nv = q.shape
    out = np.zeros_like(q)
    bflux = np.zeros(nv)
    Dr, Ds = ops.differentiation[:, :Np], ops.differentiation[:, Np:]
    qflat = q.reshape(-1, nv)
    for k in range(K):
        rx, ry, sx, sy = ops.inverse_metric[:, k]
        Ft = np.zeros((Np, nv))
        Gt = np.zeros((Np, nv))
        for j in range(Np):
            rho, mx, my, e = q[j, k]
            u, v = mx / rho, my / rho
            p = (gamma - 1.0) * (e - 0.5 * (mx * u + my * v))
            F = np.array([mx, mx * u + p, my * u, (e + p) * u])
            G = np.array([my, mx * v, my * v + p, (e + p) * v])
            Ft[j], Gt[j] = rx * F + ry * G, sx * F + sy * G
        for i in range(Np):
            for j in range(Np):
                out[i, k] -= Dr[i, j] * Ft[j] + Ds[i, j] * Gt[j]
        for sl in range(ops.face_nodes.size):
            node = ops.face_nodes[sl]
            nx, ny = ops.normal_x[sl, k], ops.normal_y[sl, k]
            qm = q[node, k]
            kind = ops.boundary_kind[sl, k]
            if kind == SLIP_WALL:
                mn = qm[1] * nx + qm[2] * ny
                qp = qm - np.array([0.0, 2 * mn * nx, 2 * mn * ny, 0.0])
            elif kind == FIXED_STATE:
                qp = ops.boundary_state[sl, k]
            elif kind == TRANSMISSIVE:
                qp = qm
            else:
                qp = qflat[ops.exterior_index[sl, k]]
            fh = flux(qm, qp, nx, ny, gamma)
            jump = ops.face_scale[sl, k] * (normal_flux(qm, nx, ny, gamma) - fh)
            for i in range(Np):
                out[i, k] += ops.lift[i, sl] * jump
            bflux += ops.boundary_weight[sl, k] * fh
    return out, bflux
