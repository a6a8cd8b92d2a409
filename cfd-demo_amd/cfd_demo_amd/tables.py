"""Compact per-element HBM tables for the gfx950 kernels (host setup).

Replaces the reference's per-face-node broadcast tables (``DGOperators``,
``dg.py:30-49``, and ``LimiterOperators``, ``limiting.py:34-48``), which are
~2.9 GB at 2M elements, with 208 bytes per element:

* ``geo``  (K, 16) f64: rx ry sx sy | nx[3] | ny[3] | face_scale[3] | 0.5|e|[3]
* ``conn`` (K, 4) i32: neighbour element per face (or fixed-state index on a
  FIXED face) | one code byte per face: bits 0-1 neighbour face, bit 2 node
  order reversed, bits 3-5 kind, bits 6-7 Cockburn-Shu partner face
* ``lim``  (K, 8) f64: alpha[3] | beta[3] | dt_geometry | jacobian

On an ``ExactState`` face the neighbour field is the first of six
``bc_states`` rows: the exterior state at each of the face's five nodes (own
node order), then the ghost-cell average for the limiter.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .boundary import EXACT_STATE, FIXED_STATE, PERIODIC, SLIP_WALL, TRANSMISSIVE, Periodic, reflect_across_faces
from .mesh import ElementGeometry, MeshConnectivity, node_coordinates

# Boundary kind -> the kind code the kernels read (conn code bits 3-5).
KIND_CODE = {SLIP_WALL: 1, FIXED_STATE: 2, TRANSMISSIVE: 3, EXACT_STATE: 4}
EXACT_ROWS = 6  # bc_states rows per exact-state face: 5 face nodes + the ghost average


@dataclass(frozen=True, eq=False)
class MeshTables:
    geo: np.ndarray
    conn: np.ndarray
    lim: np.ndarray
    bc_states: np.ndarray
    boundary_kind: np.ndarray  # (K, 3) kind per face (0 interior)
    periodic_shift: np.ndarray  # (K, 3, 2)

    @property
    def num_elements(self) -> int:
        return self.geo.shape[0]


def resolve_boundary_conditions(conn: MeshConnectivity, boundary_conditions):
    """Kinds (K, 3) and fixed-state indices (K, 3); same checks and messages
    as ``resolve_boundary_conditions`` (``dg.py:52-73``)."""
    tags = set(conn.boundary_tags)
    missing = sorted(tags - set(boundary_conditions))
    if missing:
        raise ValueError(f"Boundary tags without a boundary condition: {missing}.")
    unknown = sorted(set(boundary_conditions) - tags)
    if unknown:
        raise ValueError(f"Boundary conditions for tags that are not in the mesh: {unknown}.")
    kind = np.zeros(conn.face_tag.shape, dtype=np.int32)
    state_index = np.zeros(conn.face_tag.shape, dtype=np.int32)
    states = []
    for t, tag in enumerate(conn.boundary_tags):
        bc = boundary_conditions[tag]
        mask = conn.face_tag == t
        if bc.kind == PERIODIC:
            continue
        if bc.kind not in KIND_CODE:
            raise ValueError(f"Unsupported boundary condition for tag {tag!r}: {bc!r}.")
        kind[mask] = bc.kind
        if bc.kind == FIXED_STATE:
            state_index[mask] = len(states)
            states.append(bc.state)
    return kind, state_index, np.asarray(states, dtype=np.float64).reshape(-1, 4)


def exact_state_rows(mesh, conn: MeshConnectivity, geometry, boundary_conditions, kind, state_index, states):
    """Append six bc_states rows per ExactState face (its five face-node
    exterior states, evaluated once at the physical face nodes, and the ghost
    average over the element reflected across the face) and point the face's
    state index at the first. Returns (state_index, states)."""
    from .reference_element import DEGREE, build_reference_triangle

    ref = build_reference_triangle(DEGREE)
    x, y = node_coordinates(mesh, ref.r, ref.s)
    fnodes = np.asarray(ref.face_nodes).reshape(3, -1)
    corners = mesh.vertices[mesh.elements]
    rows = [np.asarray(states, dtype=np.float64).reshape(-1, 4)]
    nrow = rows[0].shape[0]
    state_index = state_index.copy()
    for t, tag in enumerate(conn.boundary_tags):
        bc = boundary_conditions.get(tag)
        if bc is None or bc.kind != EXACT_STATE:
            continue
        e, f = np.nonzero(np.asarray(conn.face_tag) == t)
        if e.size == 0:
            continue
        node = fnodes[f]  # (n, 5)
        nodal = bc.face_states(x[node, e[:, None]], y[node, e[:, None]])  # (n, 5, 4)
        ghost = reflect_across_faces(corners[e], geometry.normal[e], geometry.face_midpoint[e])[np.arange(e.size), f]
        block = np.concatenate([nodal, bc.ghost_averages(ghost)[:, None, :]], axis=1)  # (n, 6, 4)
        state_index[e, f] = nrow + EXACT_ROWS * np.arange(e.size)
        rows.append(block.reshape(-1, 4))
        nrow += EXACT_ROWS * e.size
    return state_index, np.concatenate(rows)


def time_step_geometry(geometry, degree: int = 4) -> np.ndarray:
    """min_e 2|K| / ((N+1)^2 |e|) (``time_integration.py:32-34``)."""
    return np.min(2.0 * geometry.area[:, None] / ((degree + 1) ** 2 * geometry.edge_length), axis=1)


def assemble(mesh, boundary_conditions):
    """The compact tables, the connectivity (periodic faces patched to
    interior faces) and the geometry of ``mesh`` under ``boundary_conditions``
    from the native assembler (cfd_build_tables; host code, no GPU). Returns
    (MeshTables, MeshConnectivity, ElementGeometry). Boundary-condition checks
    and messages are those of ``resolve_boundary_conditions`` (``dg.py:52-73``)."""
    from . import _lib

    tags = tuple(sorted(mesh.boundary_edges))
    missing = sorted(set(tags) - set(boundary_conditions))
    if missing:
        raise ValueError(f"Boundary tags without a boundary condition: {missing}.")
    unknown = sorted(set(boundary_conditions) - set(tags))
    if unknown:
        raise ValueError(f"Boundary conditions for tags that are not in the mesh: {unknown}.")
    T = len(tags)
    kind = np.zeros(T, dtype=np.int32)
    state = np.zeros(T, dtype=np.int32)
    partner = -np.ones(T, dtype=np.int32)
    shift = np.zeros((T, 2))
    states = []
    for t, tag in enumerate(tags):
        bc = boundary_conditions[tag]
        if bc.kind == PERIODIC:
            if bc.partner not in tags or not isinstance(boundary_conditions.get(bc.partner), Periodic):
                raise ValueError(f"Periodic tag {tag!r} needs a periodic partner tag {bc.partner!r}.")
            partner[t] = tags.index(bc.partner)
            shift[t] = np.asarray(bc.shift, dtype=np.float64)
        elif bc.kind not in KIND_CODE:
            raise ValueError(f"Unsupported boundary condition for tag {tag!r}: {bc!r}.")
        kind[t] = bc.kind
        if bc.kind == FIXED_STATE:
            state[t] = len(states)
            states.append(bc.state)
    K = mesh.num_elements
    verts = np.ascontiguousarray(mesh.vertices, dtype=np.float64)
    elems = np.ascontiguousarray(mesh.elements, dtype=np.int64)
    edges = [np.ascontiguousarray(mesh.boundary_edges[t], dtype=np.int64).reshape(-1, 2) for t in tags]
    offsets = np.zeros(T + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([len(e) for e in edges])
    tag_edges = np.ascontiguousarray(np.concatenate(edges) if edges else np.zeros((0, 2), np.int64))
    names = "\n".join(tags).encode()
    geo = np.empty((K, 16))
    cn = np.empty((K, 4), dtype=np.int32)
    lim = np.empty((K, 8))
    ne = np.empty((K, 3), dtype=np.int64)
    nf = np.empty((K, 3), dtype=np.int64)
    ftag = np.empty((K, 3), dtype=np.int64)
    cen = np.empty((K, 2))
    nrm = np.empty((K, 3, 2))
    length = np.empty((K, 3))
    midp = np.empty((K, 3, 2))
    pshift = np.empty((K, 3, 2))
    i64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
    i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    inp = _lib.CfdMeshInput(len(verts), dp(verts), K, i64(elems), T, names, i64(offsets), i64(tag_edges),
                            i32(kind), i32(state), i32(partner), dp(shift))
    out = _lib.CfdMeshOutput(dp(geo), i32(cn), dp(lim), i64(ne), i64(nf), i64(ftag), dp(cen), dp(nrm), dp(length),
                             dp(midp), dp(pshift))
    _lib.check(_lib.load().cfd_build_tables(C.byref(inp), C.byref(out)))
    def ro(a):  # freshly written arrays (or views of them): freeze in place, no copy
        a.setflags(write=False)
        return a

    conn = MeshConnectivity(ro(ne), ro(nf), tags, ro(ftag))
    geometry = ElementGeometry(jacobian=ro(lim[:, 7].copy()), rx=ro(geo[:, 0]), ry=ro(geo[:, 1]), sx=ro(geo[:, 2]),
                               sy=ro(geo[:, 3]), centroid=ro(cen), normal=ro(nrm), edge_length=ro(length),
                               face_midpoint=ro(midp))
    face_kind = np.where(ftag >= 0, kind[np.maximum(ftag, 0)], 0).astype(np.int32)
    states = np.asarray(states, dtype=np.float64).reshape(-1, 4)
    if np.any(face_kind == EXACT_STATE):
        index = np.asarray(cn[:, :3], dtype=np.int32)
        index, states = exact_state_rows(mesh, conn, geometry, boundary_conditions, face_kind, index, states)
        exact = face_kind == EXACT_STATE
        cn[:, :3] = np.where(exact, index, cn[:, :3])
    tables = MeshTables(geo=geo, conn=cn, lim=lim, bc_states=np.ascontiguousarray(states), boundary_kind=face_kind,
                        periodic_shift=pshift)
    return tables, conn, geometry
