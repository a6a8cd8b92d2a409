"""Triangle mesh, connectivity, geometry and boundary conditions (oracle restatement).

Restates ``src/cfd_demo/mesh.py`` and ``src/cfd_demo/boundary.py``. Host NumPy,
float64, run once per solver.
"""

from __future__ import annotations

from collections.abc import Mapping
from dataclasses import dataclass
from types import MappingProxyType

import numpy as np

NUM_FACES = 3
SLIP_WALL = 1  # boundary.py:13
FIXED_STATE = 2  # boundary.py:14
# Extension kinds (not in the reference, which only has the two above):
# a transmissive face uses its own trace as exterior state and counts its
# boundary weight; a periodic face is patched into an interior face.
TRANSMISSIVE = 3
PERIODIC = 4
# The far-field condition proposed in docs/adr/0012-sheared-rectangle-mesh-
# and-far-field-boundaries.md:32-44: exterior states from a known solution,
# one per face node, and the limiter ghost average over the reflected element.
EXACT_STATE = 5


@dataclass(frozen=True)
class SlipWall:
    """``boundary.py:17-23``."""

    kind: int = SLIP_WALL


@dataclass(frozen=True)
class FixedState:
    """``boundary.py:26-43``."""

    state: tuple
    kind: int = FIXED_STATE

    def __post_init__(self):
        st = tuple(float(x) for x in self.state)
        if len(st) != 4 or not np.all(np.isfinite(st)):
            raise ValueError("FixedState needs four finite conserved values.")
        object.__setattr__(self, "state", st)

    @classmethod
    def from_primitive(cls, rho, u, v, p, gamma):
        return cls((rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)))


@dataclass(frozen=True)
class Transmissive:
    """Extension: exterior state = interior trace (zero-gradient)."""

    kind: int = TRANSMISSIVE


@dataclass(frozen=True, eq=False)
class ExactState:
    """ADR 0012 (proposed): ``state(x, y)`` -> conserved (..., 4) at the face
    nodes; ``average(corners (n, 3, 2))`` -> (n, 4) ghost-cell averages over
    reflected elements (required here: the oracle takes exact averages only)."""

    state: object
    average: object
    kind: int = EXACT_STATE


def reflected_element(corners, face):
    """Element corners (3, 2) mirrored across the line through its face
    ``face`` (from corner face to corner face + 1): each corner goes to twice
    its projection on the line minus itself."""
    a, b = corners[face], corners[(face + 1) % 3]
    t = (b - a) / np.hypot(*(b - a))
    proj = a + np.outer((corners - a) @ t, t)
    return 2.0 * proj - corners


@dataclass(frozen=True)
class Periodic:
    """Extension: the tag is identified with ``partner`` by translation ``shift``."""

    partner: str
    shift: tuple
    kind: int = PERIODIC


def _ro(a):
    a = np.array(a, copy=True)
    a.setflags(write=False)
    return a


@dataclass(frozen=True, eq=False)
class TriangleMesh:
    """``mesh.py:27-63``."""

    vertices: np.ndarray
    elements: np.ndarray
    boundary_edges: Mapping

    def __post_init__(self):
        v = np.asarray(self.vertices, dtype=np.float64)
        e = np.asarray(self.elements)
        if v.ndim != 2 or v.shape[1] != 2:
            raise ValueError("vertices must have shape (Nv, 2).")
        if e.ndim != 2 or e.shape[1] != 3 or not np.issubdtype(e.dtype, np.integer):
            raise ValueError("elements must be an integer array of shape (K, 3).")
        if e.size and (e.min() < 0 or e.max() >= len(v)):
            raise ValueError("elements reference vertices that do not exist.")
        edges = {str(t): _ro(np.asarray(p, dtype=np.int64).reshape(-1, 2)) for t, p in self.boundary_edges.items()}
        object.__setattr__(self, "vertices", _ro(v))
        object.__setattr__(self, "elements", _ro(e.astype(np.int64)))
        object.__setattr__(self, "boundary_edges", MappingProxyType(edges))
        if np.any(signed_areas(self) <= 0.0):
            raise ValueError("Every element must have counterclockwise vertices and positive area.")

    @property
    def num_elements(self):
        return self.elements.shape[0]


def signed_areas(mesh):
    """``mesh.py:66-69``."""
    p0, p1, p2 = (mesh.vertices[mesh.elements[:, i]] for i in range(3))
    return 0.5 * ((p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1]))


def rectangle_mesh(nx, ny, x_range=(0.0, 1.0), y_range=(0.0, 1.0)):
    """``mesh.py:72-112``: cell (i, j) -> elements 2(j nx + i) (lower-right), +1 (upper-left)."""
    if nx < 1 or ny < 1:
        raise ValueError("nx and ny must be positive.")
    (x0, x1), (y0, y1) = x_range, y_range
    xs = x0 + (x1 - x0) * (np.arange(nx + 1) / nx)
    ys = y0 + (y1 - y0) * (np.arange(ny + 1) / ny)
    xx, yy = np.meshgrid(xs, ys)
    verts = np.stack([xx.ravel(), yy.ravel()], axis=-1)

    def idx(i, j):
        return j * (nx + 1) + i

    i, j = np.meshgrid(np.arange(nx), np.arange(ny))
    i, j = i.ravel(), j.ravel()
    a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
    elems = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], axis=1).reshape(-1, 3)
    cols, rows = np.arange(nx), np.arange(ny)
    bnd = {
        "bottom": np.stack([idx(cols, 0), idx(cols + 1, 0)], -1),
        "top": np.stack([idx(cols, ny), idx(cols + 1, ny)], -1),
        "left": np.stack([idx(0, rows), idx(0, rows + 1)], -1),
        "right": np.stack([idx(nx, rows), idx(nx, rows + 1)], -1),
    }
    return TriangleMesh(vertices=verts, elements=elems, boundary_edges=bnd)


@dataclass(frozen=True, eq=False)
class MeshConnectivity:
    """``mesh.py:115-132``."""

    neighbor_element: np.ndarray
    neighbor_face: np.ndarray
    boundary_tags: tuple
    face_tag: np.ndarray


def build_connectivity(mesh):
    """Face matching by sorted vertex pairs (``mesh.py:135-185``)."""
    K = mesh.num_elements
    first = mesh.elements
    second = np.roll(mesh.elements, -1, axis=1)
    keys = np.sort(np.stack([first, second], -1).reshape(-1, 2), axis=1)
    _, inv, counts = np.unique(keys, axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    if np.any(counts > 2):
        raise ValueError("Nonconforming mesh: a face is shared by more than two elements.")
    order = np.argsort(inv, kind="stable")
    sid = inv[order]
    ne = -np.ones(K * NUM_FACES, dtype=np.int64)
    nf = -np.ones(K * NUM_FACES, dtype=np.int64)
    pr = np.flatnonzero(sid[1:] == sid[:-1])
    left, right = order[pr], order[pr + 1]
    ne[left], nf[left] = right // NUM_FACES, right % NUM_FACES
    ne[right], nf[right] = left // NUM_FACES, left % NUM_FACES
    tags = tuple(sorted(mesh.boundary_edges))
    lookup = {}
    for ti, tag in enumerate(tags):
        for a, b in mesh.boundary_edges[tag]:
            key = (min(a, b), max(a, b))
            if key in lookup:
                raise ValueError(f"Boundary edge {key} has more than one tag.")
            lookup[key] = ti
    ftag = -np.ones(K * NUM_FACES, dtype=np.int64)
    for face in np.flatnonzero(ne < 0):
        key = (int(keys[face, 0]), int(keys[face, 1]))
        if key not in lookup:
            el, loc = divmod(int(face), NUM_FACES)
            raise ValueError(f"Boundary face {loc} of element {el} (vertices {key}) has no tag.")
        ftag[face] = lookup.pop(key)
    if lookup:
        raise ValueError(f"Tagged edges are not boundary faces: {sorted(lookup)[:5]}.")
    shp = (K, NUM_FACES)
    return MeshConnectivity(_ro(ne.reshape(shp)), _ro(nf.reshape(shp)), tags, _ro(ftag.reshape(shp)))


@dataclass(frozen=True, eq=False)
class ElementGeometry:
    """``mesh.py:188-213``."""

    jacobian: np.ndarray
    rx: np.ndarray
    ry: np.ndarray
    sx: np.ndarray
    sy: np.ndarray
    centroid: np.ndarray
    normal: np.ndarray
    edge_length: np.ndarray
    face_midpoint: np.ndarray

    @property
    def area(self):
        return 2.0 * self.jacobian


def element_geometry(mesh):
    """Affine factors from vertex coordinates (``mesh.py:216-243``)."""
    c = mesh.vertices[mesh.elements]
    x1, x2, x3 = c[:, 0, 0], c[:, 1, 0], c[:, 2, 0]
    y1, y2, y3 = c[:, 0, 1], c[:, 1, 1], c[:, 2, 1]
    xr, xs = 0.5 * (x2 - x1), 0.5 * (x3 - x1)
    yr, ys = 0.5 * (y2 - y1), 0.5 * (y3 - y1)
    J = xr * ys - xs * yr
    end = np.roll(c, -1, axis=1)
    dx = end[..., 0] - c[..., 0]
    dy = end[..., 1] - c[..., 1]
    L = np.sqrt(dx * dx + dy * dy)
    n = np.stack([dy / L, -dx / L], -1)
    return ElementGeometry(_ro(J), _ro(ys / J), _ro(-xs / J), _ro(-yr / J), _ro(xr / J),
                           _ro(c.mean(axis=1)), _ro(n), _ro(L), _ro(0.5 * (c + end)))


def node_coordinates(mesh, r, s):
    """Physical node coordinates, each (Np, K) (``mesh.py:246-251``)."""
    c = mesh.vertices[mesh.elements]
    w = np.stack([-(r + s) / 2, (1 + r) / 2, (1 + s) / 2], -1)
    pts = np.einsum("nv,kvd->nkd", w, c)
    return pts[..., 0], pts[..., 1]


def pair_face_nodes(mesh, conn, x, y, face_nodes, relative_tolerance=1e-9, shift=None):
    """Match neighbour face nodes by coordinates (``mesh.py:254-291``).

    Returns partner (3, Nfp, K). ``shift`` (K, 3, 2), when given, is added to
    the neighbour coordinates before matching (used for periodic patches).
    """
    nfaces, nfp = face_nodes.shape
    K = mesh.num_elements
    partner = np.broadcast_to(face_nodes[:, :, None], (nfaces, nfp, K)).copy()
    scale = element_geometry(mesh).edge_length
    for f in range(nfaces):
        els = np.flatnonzero(conn.neighbor_element[:, f] >= 0)
        if els.size == 0:
            continue
        nb = conn.neighbor_element[els, f]
        nbf = conn.neighbor_face[els, f]
        own_nodes = face_nodes[f]
        cand = face_nodes[nbf]
        own = np.stack([x[own_nodes][:, els], y[own_nodes][:, els]], -1)
        oth = np.stack([x[cand.T, nb[None, :]], y[cand.T, nb[None, :]]], -1)
        if shift is not None:
            oth = oth + shift[els, f][None, :, :]
        dist = np.linalg.norm(own[:, None, :, :] - oth[None, :, :, :], axis=-1)
        best = dist.argmin(axis=1)
        bd = np.take_along_axis(dist, best[:, None, :], axis=1)[:, 0, :]
        tol = relative_tolerance * scale[els, f]
        if np.any(bd > tol[None, :]):
            raise ValueError("Neighboring face nodes do not coincide; the mesh is not conforming.")
        partner[f][:, els] = np.take_along_axis(cand.T, best, axis=0)
    return partner


def periodic_connectivity(mesh, conn, boundary_conditions):
    """Extension: patch periodic tag pairs into interior connectivity.

    Returns (connectivity, shift (K, 3, 2)): boundary faces whose tag maps to
    ``Periodic(partner, shift)`` get the element across the translated edge as
    neighbour; ``shift`` is the translation from the neighbour's coordinates
    to this element's. Matching is by translated edge midpoints.
    """
    geo = element_geometry(mesh)
    ne = np.array(conn.neighbor_element)
    nf = np.array(conn.neighbor_face)
    ft = np.array(conn.face_tag)
    shift = np.zeros(ne.shape + (2,))
    tags = conn.boundary_tags
    for ti, tag in enumerate(tags):
        bc = boundary_conditions.get(tag)
        if not isinstance(bc, Periodic):
            continue
        pi = tags.index(bc.partner)
        mine = np.argwhere(np.asarray(conn.face_tag) == ti)
        theirs = np.argwhere(np.asarray(conn.face_tag) == pi)
        mid_t = geo.face_midpoint[theirs[:, 0], theirs[:, 1]] + np.asarray(bc.shift)
        mid_m = geo.face_midpoint[mine[:, 0], mine[:, 1]]
        scale = geo.edge_length[mine[:, 0], mine[:, 1]]
        d = np.linalg.norm(mid_m[:, None, :] - mid_t[None, :, :], axis=-1)
        best = d.argmin(axis=1)
        if np.any(d[np.arange(len(mine)), best] > 1e-9 * scale):
            raise ValueError(f"Periodic tags {tag!r} and {bc.partner!r} do not match.")
        e, f = mine[:, 0], mine[:, 1]
        ne[e, f] = theirs[best, 0]
        nf[e, f] = theirs[best, 1]
        shift[e, f] = np.asarray(bc.shift, dtype=np.float64)
        ft[e, f] = -1
    return MeshConnectivity(_ro(ne), _ro(nf), conn.boundary_tags, _ro(ft)), shift
