"""Build a GPU solver and the oracle solver for the same mesh, BCs and config."""

from __future__ import annotations

import numpy as np

import cfd_demo_amd as gpu
from cfd_demo_amd import boundary as gb
from cfd_demo_amd.mesh import TriangleMesh, rectangle_mesh
from oracle import mesh as om
from oracle import solver as osol

GAMMA = 1.4
FREE_STREAM = (1.2, 0.4, -0.3, 0.9)


def perturbed_vertices(nx, ny, amplitude=0.25, seed=0, x_range=(0.0, 1.0), y_range=(0.0, 1.0)):
    """Rectangle mesh with seeded interior-vertex jitter (as the reference's
    ``perturbed_mesh`` test helper, tests/unit/test_dg_operator.py:19-27)."""
    m = rectangle_mesh(nx, ny, x_range, y_range)
    v = np.array(m.vertices)
    lo = np.array([x_range[0], y_range[0]])
    hi = np.array([x_range[1], y_range[1]])
    interior = np.all((v > lo + 1e-12) & (v < hi - 1e-12), axis=1)
    rng = np.random.default_rng(seed)
    disp = rng.uniform(-1.0, 1.0, size=v.shape)
    v[interior] += amplitude * disp[interior] * (hi - lo) / np.array([nx, ny])
    return v, np.array(m.elements), {k: np.array(p) for k, p in m.boundary_edges.items()}


def _bc_pair(spec):
    kind = spec[0]
    if kind == "slip":
        return gb.SlipWall(), om.SlipWall()
    if kind == "fixed":
        return gb.FixedState(spec[1]), om.FixedState(spec[1])
    if kind == "trans":
        return gb.Transmissive(), om.Transmissive()
    if kind == "periodic":
        return gb.Periodic(spec[1], spec[2]), om.Periodic(spec[1], spec[2])
    raise ValueError(spec)


def make_pair(vertices, elements, edges, bcs, flux="rusanov", limiter="minmod", cfl=0.5, max_rejections=10):
    gm = TriangleMesh(vertices=vertices, elements=elements, boundary_edges=edges)
    omesh = om.TriangleMesh(vertices=vertices, elements=elements, boundary_edges=edges)
    gbcs, obcs = {}, {}
    for tag, spec in bcs.items():
        gbcs[tag], obcs[tag] = _bc_pair(spec)
    g = gpu.EulerSolver(gm, gbcs, gpu.SolverConfig(flux=flux, limiter=limiter, gamma=GAMMA, cfl=cfl,
                                                   max_rejections=max_rejections))
    o = osol.EulerSolver(omesh, obcs, osol.SolverConfig(flux=flux, limiter=limiter, gamma=GAMMA, cfl=cfl,
                                                         max_rejections=max_rejections))
    return g, o


def fixed_all(primitive=FREE_STREAM):
    st = om.FixedState.from_primitive(*primitive, GAMMA).state
    return {t: ("fixed", st) for t in ("left", "right", "bottom", "top")}


def sod_bcs():
    left = om.FixedState.from_primitive(1.0, 0.0, 0.0, 1.0, GAMMA).state
    right = om.FixedState.from_primitive(0.125, 0.0, 0.0, 0.1, GAMMA).state
    return {"left": ("fixed", left), "right": ("fixed", right), "bottom": ("slip",), "top": ("slip",)}


def periodic_bcs(lx=1.0, ly=1.0):
    return {"left": ("periodic", "right", (-lx, 0.0)), "right": ("periodic", "left", (lx, 0.0)),
            "bottom": ("periodic", "top", (0.0, -ly)), "top": ("periodic", "bottom", (0.0, ly))}


def trans_all():
    return {t: ("trans",) for t in ("left", "right", "bottom", "top")}


def mixed_bcs():
    st = om.FixedState.from_primitive(*FREE_STREAM, GAMMA).state
    return {"left": ("fixed", st), "right": ("trans",), "bottom": ("slip",), "top": ("trans",)}


def random_state(solver_shape, seed, lo=(0.5, -0.4, -0.4, 0.5), hi=(1.5, 0.4, 0.4, 1.5)):
    rng = np.random.default_rng(seed)
    prim = rng.uniform(lo, hi, size=solver_shape[:2] + (4,))
    return gpu_conserved(prim)


def gpu_conserved(prim):
    from oracle.physics import primitive_to_conserved
    return np.ascontiguousarray(primitive_to_conserved(*(prim[..., k] for k in range(4)), GAMMA))


def smooth_state(x, y, amp=0.2):
    two_pi = 2.0 * np.pi
    rho = 1.0 + amp * np.sin(two_pi * x) * np.sin(two_pi * y)
    u = 0.3 + 0.5 * amp * np.cos(two_pi * x) * np.sin(two_pi * y)
    v = -0.2 + 0.5 * amp * np.sin(two_pi * x) * np.cos(two_pi * y)
    p = 1.0 + 0.75 * amp * np.sin(two_pi * (x + y))
    return gpu_conserved(np.stack([rho, u, v, p], -1))


def rel_err(a, b):
    """Normwise relative error per conserved component, max over components."""
    num = np.abs(a - b).reshape(-1, a.shape[-1]).max(axis=0)
    den = np.abs(b).reshape(-1, b.shape[-1]).max(axis=0)
    den = np.where(den > 0, den, 1.0)
    return float(np.max(num / den))
