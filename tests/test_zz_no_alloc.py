"""The allocation-free step (cfd_step_device_static) against cfd_step_device.

Both run the same kernels on the same inputs, so states and reports must be
bit-identical. Two calls in a row on different meshes, the second smaller
(the scratch array keeps the first call's contents), covering the two-launch
limiter (above 4,096 elements) and the one-launch limiter of small meshes;
and the size limit, rejected before any GPU work.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cfd_demo_amd import _lib
from pairs import make_pair, sod_bcs

pytestmark = pytest.mark.gpu


def _sod_pair(nx, ny, flux, limiter):
    from pairs import perturbed_vertices

    v, e, b = perturbed_vertices(nx, ny, amplitude=0.2, seed=5)
    g, _ = make_pair(v, e, b, sod_bcs(), flux=flux, limiter=limiter)
    return g


def _sod_state(g):
    left = g.geometry.centroid[:, 0] < 0.5
    q = np.empty(g.state_shape)
    q[:, left] = [1.0, 0.0, 0.0, 2.5]
    q[:, ~left] = [0.125, 0.0, 0.0, 0.25]
    x, y = g.node_coordinates()
    q[..., 0] *= 1.0 + 0.05 * np.sin(7.0 * x + 3.0 * y)
    return q


def _device_mesh(torch, g):
    t = g.tables
    arrs = {
        "geo": torch.from_numpy(np.ascontiguousarray(t.geo)).cuda(),
        "conn": torch.from_numpy(np.ascontiguousarray(t.conn)).cuda(),
        "ab": torch.from_numpy(np.ascontiguousarray(t.lim[:, :6])).cuda(),
        "dtg": torch.from_numpy(np.ascontiguousarray(t.lim[:, 6])).cuda(),
        "bcs": torch.from_numpy(np.ascontiguousarray(t.bc_states)).cuda(),
    }
    m = _lib.CfdDeviceMesh(t.num_elements, arrs["geo"].data_ptr(), arrs["conn"].data_ptr(), arrs["ab"].data_ptr(),
                           arrs["dtg"].data_ptr(), arrs["bcs"].data_ptr(), int(t.bc_states.shape[0]))
    return m, arrs


def _both(torch, g):
    """(state, report) of cfd_step_device and of cfd_step_device_static."""
    lib = _lib.load()
    q = _sod_state(g)
    dt = g.time_step(q)
    qd = torch.from_numpy(q).cuda()
    out_ref = torch.empty_like(qd)
    out_new = torch.full_like(qd, float("nan"))
    m, keep = _device_mesh(torch, g)
    torch.cuda.synchronize()
    r_ref = _lib.CfdStepReport()
    _lib.check(lib.cfd_step_device(g._context(), C.c_void_p(qd.data_ptr()), dt, C.c_void_p(out_ref.data_ptr()),
                                   C.byref(r_ref)))
    r_new = _lib.CfdStepReport()
    _lib.check(lib.cfd_step_device_static(C.byref(m), g.flux.id, g.phi.id, 1.4, C.c_void_p(qd.data_ptr()), dt,
                                          C.c_void_p(out_new.data_ptr()), C.byref(r_new)))
    torch.cuda.synchronize()
    del keep
    return out_ref.cpu().numpy(), r_ref, out_new.cpu().numpy(), r_new


def _report_tuple(r):
    return (tuple(r.boundary_flux), r.admissible, r.min_density, r.min_pressure, r.troubled_cells)


def test_static_step_matches_context_step_twice():
    torch = pytest.importorskip("torch")
    large = _sod_pair(80, 40, "hllc", "mc")      # 6,400 elements: detection and repair as two launches
    small = _sod_pair(12, 5, "roe", "minmod")    # 120 elements: the one-launch limiter
    for g in (large, small):
        print(f"case no-alloc step K={g.mesh.num_elements} flux={g.config.flux} limiter={g.config.limiter}",
              flush=True)
        q_ref, r_ref, q_new, r_new = _both(torch, g)
        assert np.array_equal(q_new, q_ref)
        assert _report_tuple(r_new) == _report_tuple(r_ref)
        assert r_new.admissible == 1 and r_new.troubled_cells > 0
        g.close()


def test_static_step_rejects_oversized_mesh():
    lib = _lib.load()
    g = _sod_pair(4, 2, "rusanov", "minmod")
    g._context()  # cfd_init on this device
    m = _lib.CfdDeviceMesh(_lib.CFD_STATIC_MAX_ELEMENTS + 1, 1, 1, 1, 1, 0, 0)
    rep = _lib.CfdStepReport()
    code = lib.cfd_step_device_static(C.byref(m), 0, 0, 1.4, C.c_void_p(1), 1e-3, C.c_void_p(1), C.byref(rep))
    assert code == _lib.CFD_EINVAL
    assert "at most 65536" in _lib.last_error()
    g.close()
