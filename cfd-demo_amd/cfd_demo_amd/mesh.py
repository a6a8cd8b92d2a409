"""Conforming triangle meshes, face connectivity and affine geometry (host setup).

Public surface follows ``src/cfd_demo/mesh.py``: ``TriangleMesh`` (27-63),
``signed_areas`` (66-69), ``rectangle_mesh`` (72-112), ``MeshConnectivity`` /
``build_connectivity`` (115-185), ``ElementGeometry`` / ``element_geometry``
(188-243) and ``node_coordinates`` (246-251), with the same validation
messages. Everything is vectorised so that 2M-element meshes assemble in
seconds; the face-node pairing of ``pair_face_nodes`` (254-291) is replaced by
an orientation test on the face end vertices (``face_orientation``).
"""

from __future__ import annotations

from collections.abc import Mapping
from dataclasses import dataclass
from types import MappingProxyType

import numpy as np

NUM_FACES = 3


def _frozen(a) -> np.ndarray:
    a = np.array(a, copy=True)
    a.setflags(write=False)
    return a


@dataclass(frozen=True, eq=False)
class TriangleMesh:
    """Straight-sided conforming mesh: vertices (Nv, 2), CCW elements (K, 3),
    and tagged boundary edges {tag: (n, 2) vertex pairs}."""

    vertices: np.ndarray
    elements: np.ndarray
    boundary_edges: Mapping[str, np.ndarray]

    def __post_init__(self) -> None:
        verts = np.asarray(self.vertices, dtype=np.float64)
        elems = np.asarray(self.elements)
        if verts.ndim != 2 or verts.shape[1] != 2:
            raise ValueError("vertices must have shape (Nv, 2).")
        if elems.ndim != 2 or elems.shape[1] != 3 or not np.issubdtype(elems.dtype, np.integer):
            raise ValueError("elements must be an integer array of shape (K, 3).")
        if elems.size and (elems.min() < 0 or elems.max() >= len(verts)):
            raise ValueError("elements reference vertices that do not exist.")
        tagged = {str(t): _frozen(np.asarray(p, dtype=np.int64).reshape(-1, 2)) for t, p in self.boundary_edges.items()}
        object.__setattr__(self, "vertices", _frozen(verts))
        object.__setattr__(self, "elements", _frozen(elems.astype(np.int64)))
        object.__setattr__(self, "boundary_edges", MappingProxyType(tagged))
        if np.any(signed_areas(self) <= 0.0):
            raise ValueError("Every element must have counterclockwise vertices and positive area.")

    @property
    def num_elements(self) -> int:
        return self.elements.shape[0]


def signed_areas(mesh: TriangleMesh) -> np.ndarray:
    p = mesh.vertices[mesh.elements]
    return 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                  - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))


def rectangle_mesh(nx: int, ny: int, x_range=(0.0, 1.0), y_range=(0.0, 1.0)) -> TriangleMesh:
    """nx*ny rectangles split along the lower-left/upper-right diagonal; cell
    (i, j) holds elements 2(j nx + i) (lower-right) and 2(j nx + i)+1
    (upper-left); tags left/right/bottom/top (``mesh.py:72-112``)."""
    if nx < 1 or ny < 1:
        raise ValueError("nx and ny must be positive.")
    (x0, x1), (y0, y1) = x_range, y_range
    xs = x0 + (x1 - x0) * (np.arange(nx + 1) / nx)
    ys = y0 + (y1 - y0) * (np.arange(ny + 1) / ny)
    verts = np.empty(((nx + 1) * (ny + 1), 2))
    verts[:, 0] = np.tile(xs, ny + 1)
    verts[:, 1] = np.repeat(ys, nx + 1)
    jj, ii = np.divmod(np.arange(nx * ny), nx)
    ll = jj * (nx + 1) + ii  # lower-left vertex of each cell
    elems = np.empty((2 * nx * ny, 3), dtype=np.int64)
    elems[0::2] = np.stack([ll, ll + 1, ll + nx + 2], -1)
    elems[1::2] = np.stack([ll, ll + nx + 2, ll + nx + 1], -1)
    col, row = np.arange(nx), np.arange(ny)
    edges = {
        "bottom": np.stack([col, col + 1], -1),
        "top": np.stack([ny * (nx + 1) + col, ny * (nx + 1) + col + 1], -1),
        "left": np.stack([row * (nx + 1), (row + 1) * (nx + 1)], -1),
        "right": np.stack([row * (nx + 1) + nx, (row + 1) * (nx + 1) + nx], -1),
    }
    return TriangleMesh(vertices=verts, elements=elems, boundary_edges=edges)


def sheared_rectangle_mesh(nx: int, ny: int, x_range=(0.0, 1.0), y_range=(0.0, 1.0),
                           angle: float = 0.0) -> TriangleMesh:
    """``rectangle_mesh`` with its vertices mapped by (x, y) -> (x, y + x
    tan(angle)) (the generator proposed in the reference's
    docs/adr/0012-sheared-rectangle-mesh-and-far-field-boundaries.md:20-31):
    connectivity, the lower-left to upper-right diagonal and the tags are
    unchanged, ``bottom`` becomes the inclined wall. A zero angle reproduces
    ``rectangle_mesh`` exactly."""
    if not abs(angle) < 0.5 * np.pi:
        raise ValueError("The shear angle must lie strictly between -pi/2 and pi/2.")
    m = rectangle_mesh(nx, ny, x_range, y_range)
    v = np.array(m.vertices)
    v[:, 1] = v[:, 1] + v[:, 0] * np.tan(angle)
    return TriangleMesh(vertices=v, elements=np.array(m.elements), boundary_edges=dict(m.boundary_edges))


@dataclass(frozen=True, eq=False)
class MeshConnectivity:
    """neighbor_element / neighbor_face (K, 3), -1 on boundaries; sorted
    boundary_tags; face_tag (K, 3) index into boundary_tags or -1."""

    neighbor_element: np.ndarray
    neighbor_face: np.ndarray
    boundary_tags: tuple
    face_tag: np.ndarray


def _edge_keys(a, b, nv):
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    return lo * np.int64(nv) + hi


def build_connectivity(mesh: TriangleMesh) -> MeshConnectivity:
    """Faces match by their sorted vertex pair; a face shared three times, an
    untagged boundary face or a tagged interior edge is an error."""
    K = mesh.num_elements
    nv = len(mesh.vertices)
    e = mesh.elements
    keys = _edge_keys(e, np.roll(e, -1, axis=1), nv).reshape(-1)
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    if sk.size >= 3 and np.any(sk[2:] == sk[:-2]):
        raise ValueError("Nonconforming mesh: a face is shared by more than two elements.")
    ne = -np.ones(K * NUM_FACES, dtype=np.int64)
    nf = -np.ones(K * NUM_FACES, dtype=np.int64)
    pair = np.flatnonzero(sk[1:] == sk[:-1])
    a, b = order[pair], order[pair + 1]
    ne[a], nf[a] = b // NUM_FACES, b % NUM_FACES
    ne[b], nf[b] = a // NUM_FACES, a % NUM_FACES

    tags = tuple(sorted(mesh.boundary_edges))
    tkeys, tidx = [], []
    for t, tag in enumerate(tags):
        pairs = mesh.boundary_edges[tag]
        tkeys.append(_edge_keys(pairs[:, 0], pairs[:, 1], nv))
        tidx.append(np.full(len(pairs), t, dtype=np.int64))
    tkeys = np.concatenate(tkeys) if tkeys else np.zeros(0, np.int64)
    tidx = np.concatenate(tidx) if tidx else np.zeros(0, np.int64)
    uk, cnt = np.unique(tkeys, return_counts=True)
    if np.any(cnt > 1):
        dup = int(uk[np.argmax(cnt > 1)])
        raise ValueError(f"Boundary edge {(dup // nv, dup % nv)} has more than one tag.")
    torder = np.argsort(tkeys)
    tk_sorted, ti_sorted = tkeys[torder], tidx[torder]
    bfaces = np.flatnonzero(ne < 0)
    ftag = -np.ones(K * NUM_FACES, dtype=np.int64)
    pos = np.searchsorted(tk_sorted, keys[bfaces])
    found = (pos < tk_sorted.size) & (tk_sorted[np.minimum(pos, max(tk_sorted.size - 1, 0))] == keys[bfaces]) \
        if tk_sorted.size else np.zeros(bfaces.size, bool)
    if not np.all(found):
        face = int(bfaces[np.argmin(found)])
        el, loc = divmod(face, NUM_FACES)
        k = int(keys[face])
        raise ValueError(f"Boundary face {loc} of element {el} (vertices {(k // nv, k % nv)}) has no tag.")
    ftag[bfaces] = ti_sorted[pos]
    used = np.zeros(tk_sorted.size, bool)
    used[pos] = True
    if not np.all(used):
        left = sorted((int(k) // nv, int(k) % nv) for k in tk_sorted[~used])
        raise ValueError(f"Tagged edges are not boundary faces: {left[:5]}.")
    shp = (K, NUM_FACES)
    return MeshConnectivity(_frozen(ne.reshape(shp)), _frozen(nf.reshape(shp)), tags, _frozen(ftag.reshape(shp)))


@dataclass(frozen=True, eq=False)
class ElementGeometry:
    """Affine geometry: jacobian (K), rx, ry, sx, sy (K), centroid (K, 2),
    outward unit normal (K, 3, 2), edge_length (K, 3), face_midpoint (K, 3, 2)."""

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
    def area(self) -> np.ndarray:
        return 2.0 * self.jacobian


def element_geometry(mesh: TriangleMesh) -> ElementGeometry:
    """Geometric factors from the vertex coordinates; the arithmetic is the
    reference's (``mesh.py:216-243``) so both sides of a face see bit-equal
    lengths and exactly opposite normals."""
    c = mesh.vertices[mesh.elements]
    xr, xs = 0.5 * (c[:, 1, 0] - c[:, 0, 0]), 0.5 * (c[:, 2, 0] - c[:, 0, 0])
    yr, ys = 0.5 * (c[:, 1, 1] - c[:, 0, 1]), 0.5 * (c[:, 2, 1] - c[:, 0, 1])
    jac = xr * ys - xs * yr
    nxt = np.roll(c, -1, axis=1)
    dx = nxt[..., 0] - c[..., 0]
    dy = nxt[..., 1] - c[..., 1]
    length = np.sqrt(dx * dx + dy * dy)
    normal = np.stack([dy / length, -dx / length], axis=-1)
    return ElementGeometry(
        jacobian=_frozen(jac), rx=_frozen(ys / jac), ry=_frozen(-xs / jac), sx=_frozen(-yr / jac),
        sy=_frozen(xr / jac), centroid=_frozen(c.mean(axis=1)), normal=_frozen(normal),
        edge_length=_frozen(length), face_midpoint=_frozen(0.5 * (c + nxt)),
    )


def node_coordinates(mesh: TriangleMesh, r: np.ndarray, s: np.ndarray):
    """Physical node coordinates (x, y), each (Np, K)."""
    c = mesh.vertices[mesh.elements]
    w = np.stack([-(r + s) / 2, (1 + r) / 2, (1 + s) / 2], axis=-1)
    x = w @ c[:, :, 0].T
    y = w @ c[:, :, 1].T
    return x, y


# Local vertex at which each face's node list (ascending node index) starts.
FACE_START_VERTEX = (0, 1, 0)


def face_orientation(mesh: TriangleMesh, conn: MeshConnectivity, shift=None) -> np.ndarray:
    """(K, 3) bool: True where the neighbour lists the shared face's nodes in
    the opposite direction, so own face node k meets neighbour face node 4-k.

    Equivalent to ``pair_face_nodes`` (``mesh.py:254-291``) on conforming
    straight-sided meshes because the GLL face nodes are symmetric about the
    edge midpoint. ``shift`` (K, 3, 2) translates periodic neighbours.
    """
    K = mesh.num_elements
    start = np.asarray(FACE_START_VERTEX)
    own_start = mesh.vertices[mesh.elements[np.arange(K)[:, None], start[None, :]]]  # (K, 3, 2)
    nb = np.maximum(conn.neighbor_element, 0)
    nbf = np.maximum(conn.neighbor_face, 0)
    nb_start = mesh.vertices[mesh.elements[nb, start[nbf]]]
    if shift is not None:
        nb_start = nb_start + shift
    scale = element_geometry(mesh).edge_length
    same = np.linalg.norm(own_start - nb_start, axis=-1) <= 1e-9 * scale
    interior = conn.neighbor_element >= 0
    # A conforming neighbour either starts at our start vertex or at our end vertex.
    end = np.asarray((1, 2, 2))
    own_end = mesh.vertices[mesh.elements[np.arange(K)[:, None], end[None, :]]]
    flipped = np.linalg.norm(own_end - nb_start, axis=-1) <= 1e-9 * scale
    bad = interior & ~(same | flipped)
    if np.any(bad):
        raise ValueError("Neighboring face nodes do not coincide; the mesh is not conforming.")
    return interior & ~same
