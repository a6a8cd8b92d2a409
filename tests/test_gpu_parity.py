"""GPU parity: libcfdhip.so through the C ABI against the oracle, same inputs.

Bars (tolerances written here): per-RHS normwise relative error <= 1e-12 per
conserved component; troubled flags identical; limited states <= 1e-12;
time step equal to 1e-15 relative; step and short runs <= 1e-11.
"""

import numpy as np
import pytest

from pairs import (FREE_STREAM, fixed_all, make_pair, mixed_bcs, perturbed_vertices, periodic_bcs, random_state,
                   rel_err, smooth_state, sod_bcs, trans_all)

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
FLUXES = ["rusanov", "roe", "hllc"]
BCS = {"sod": sod_bcs, "fixed": fixed_all, "trans": trans_all, "mixed": mixed_bcs, "periodic": periodic_bcs}


@pytest.fixture(autouse=True, params=["quad", "lane", "quad-split"])
def stage_layout(request, monkeypatch):
    """Every parity test under each stage-kernel layout: four lanes per
    element (the default), one lane per element; and the default layout with
    the limiter's detection and repair as two launches (what meshes above
    4,096 elements use) instead of the one-pass kernel the small meshes here
    get."""
    monkeypatch.setenv("CFDHIP_STAGE_QUAD", "0" if request.param == "lane" else "1")
    monkeypatch.setenv("CFDHIP_LIMIT_ONEPASS", "0" if request.param == "quad-split" else "1")
    print(f"layout {request.param}", flush=True)
    return request.param


def _pair(bcname, flux="rusanov", limiter="minmod", nx=6, ny=5, seed=3, cfl=0.5):
    v, e, b = perturbed_vertices(nx, ny, seed=seed)
    if bcname == "periodic":
        v, e, b = perturbed_vertices(nx, ny, amplitude=0.0)
    return make_pair(v, e, b, BCS[bcname](), flux=flux, limiter=limiter, cfl=cfl)


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("bcname", list(BCS))
def test_rhs_matches_oracle(flux, bcname):
    g, o = _pair(bcname, flux)
    print(f"case rhs flux={flux} bc={bcname} K={g.mesh.num_elements}", flush=True)
    q = random_state(g.state_shape, 11)
    r_gpu, b_gpu = g.rhs(q)
    r_ref, b_ref = o.rhs(q)
    assert rel_err(r_gpu, r_ref) <= RHS_TOL
    assert np.allclose(b_gpu, b_ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("flux", FLUXES)
def test_rhs_smooth_state(flux):
    g, o = _pair("fixed", flux, nx=8, ny=8)
    x, y = g.node_coordinates()
    q = smooth_state(x, y)
    r_gpu, _ = g.rhs(q)
    r_ref, _ = o.rhs(q)
    assert rel_err(r_gpu, r_ref) <= RHS_TOL


@pytest.mark.parametrize("limiter", ["minmod", "mc", "van_leer", "superbee", "unlimited"])
def test_limit_matches_oracle(limiter):
    g, o = _pair("sod", limiter=limiter)
    rng = np.random.default_rng(1)
    x, y = g.node_coordinates()
    q = smooth_state(x, y) * (1.0 + 0.3 * rng.standard_normal(g.state_shape))
    q_gpu, avg_gpu, cnt, flags = g.limit_with_flags(q)
    q_ref, flags_ref = o.limit(q, return_flags=True)
    assert np.array_equal(flags, flags_ref)
    assert cnt == int(flags_ref.sum())
    assert rel_err(q_gpu, q_ref) <= 1e-12
    assert np.allclose(avg_gpu, o.cell_averages(q), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("name,limit_slopes", [("mc", True), ("superbee", True), ("minmod", False)])
def test_apply_limiters_matches_oracle(name, limit_slopes):
    """The pure-function limiter entry point over a solver's tables, as the
    reference's tests call it (tests/unit/test_limiting_and_time.py:85-87):
    limited state, averages of the input and the troubled count (int32 zero
    without slope limiting) against the oracle's."""
    import cfd_demo_amd as gpu
    from oracle import physics as ophys
    from oracle.limiting import apply_limiters as oracle_apply

    g, o = _pair("sod", limiter=name)
    rng = np.random.default_rng(7)
    x, y = g.node_coordinates()
    q = smooth_state(x, y) * (1.0 + 0.3 * rng.standard_normal(g.state_shape))
    print(f"case apply_limiters phi={name} limit_slopes={limit_slopes}", flush=True)
    q_gpu, avg_gpu, cnt_gpu = gpu.apply_limiters(q, g.limiter_operators, gpu.get_limiter_function(name), 1.4,
                                                 limit_slopes=limit_slopes)
    q_ref, avg_ref, cnt_ref = oracle_apply(q, o.limiter_operators, ophys.get_limiter_function(name), 1.4,
                                           limit_slopes)
    assert rel_err(q_gpu, q_ref) <= 1e-12
    assert np.allclose(avg_gpu, avg_ref, rtol=1e-13, atol=1e-14)
    assert int(cnt_gpu) == int(cnt_ref)
    if not limit_slopes:
        assert cnt_gpu.dtype == np.int32 and int(cnt_gpu) == 0


def test_positivity_repair_matches_oracle():
    g, o = _pair("fixed", limiter="unlimited", nx=4, ny=4)
    q = np.broadcast_to(np.array([1.0, 0.0, 0.0, 2.5]), g.state_shape).copy()
    dev = np.linspace(-1.5, 1.5, 15)
    w = g.reference.cell_weights
    dev -= w @ dev / w.sum()
    q[:, 5, 0] += dev
    q[:, 5, 3] += 1.2 * dev
    q_gpu = g.limit(q)
    q_ref = o.limit(q)
    assert rel_err(q_gpu, q_ref) <= 1e-12
    from oracle.physics import pressure
    assert q_gpu[..., 0].min() >= 1e-13 * (1 - 1e-9)
    assert pressure(q_gpu, 1.4).min() >= 1e-13 * (1 - 1e-9)


@pytest.mark.parametrize("bcname", ["sod", "periodic", "trans"])
def test_time_step_matches_oracle(bcname):
    g, o = _pair(bcname)
    q = random_state(g.state_shape, 5)
    assert g.time_step(q) == pytest.approx(o.time_step(q), rel=1e-15)


@pytest.mark.parametrize("flux,limiter", [("rusanov", "minmod"), ("roe", "mc"), ("hllc", "superbee"),
                                          ("hllc", "van_leer"), ("roe", "unlimited")])
def test_step_matches_oracle(flux, limiter):
    g, o = _pair("sod", flux, limiter, nx=10, ny=3)
    left = g.geometry.centroid[:, 0] < 0.5
    q = np.empty(g.state_shape)
    q[:, left] = [1.0, 0.0, 0.0, 2.5]
    q[:, ~left] = [0.125, 0.0, 0.0, 0.25]
    dt = o.time_step(q)
    q_gpu, rep, flags = g.step_with_flags(q, dt)
    fl_ref = []
    q_ref, rref = o.step(q, dt, flags_out=fl_ref)
    assert rel_err(q_gpu, q_ref) <= 1e-11
    assert np.array_equal(flags, np.stack(fl_ref))
    assert bool(rep.admissible) == bool(rref.admissible)
    assert int(rep.troubled_cells) == int(rref.troubled_cells)
    assert np.allclose(rep.boundary_flux, rref.boundary_flux, rtol=1e-11, atol=1e-15)
    assert rep.min_density == pytest.approx(float(rref.min_density), rel=1e-12)
    assert rep.min_pressure == pytest.approx(float(rref.min_pressure), rel=1e-12)


@pytest.mark.parametrize("flux", FLUXES)
def test_stage_block_switch_matches_oracle(flux):
    """cfd_ctx_set_stage_block on a live context: 64 keeps the layout the
    context was created with, 128 and 256 run one lane per element with that
    many elements per workgroup; each gives the oracle's RHS, boundary flux and
    step (216 elements: partial last workgroup at every size). Other sizes are
    rejected."""
    from cfd_demo_amd import _lib

    g, o = _pair("mixed", flux, "mc", nx=12, ny=9)
    x, y = g.node_coordinates()
    rng = np.random.default_rng(4)
    q = smooth_state(x, y) * (1.0 + 0.1 * rng.standard_normal(g.state_shape))
    r_ref, b_ref = o.rhs(q)
    dt = o.time_step(q)
    q1_ref, rep_ref = o.step(q, dt)
    lib = _lib.load()
    ctx = g._context()
    for sb in (64, 128, 256, 64):
        print(f"case stage block {sb} flux={flux} K={g.mesh.num_elements}", flush=True)
        _lib.check(lib.cfd_ctx_set_stage_block(ctx, sb))
        r, b = g.rhs(q)
        assert rel_err(r, r_ref) <= RHS_TOL
        assert np.allclose(b, b_ref, rtol=1e-12, atol=1e-13)
        q1, rep = g.step(q, dt)
        assert rel_err(q1, q1_ref) <= 1e-11
        assert int(rep.troubled_cells) == int(rep_ref.troubled_cells)
        assert np.allclose(rep.boundary_flux, rep_ref.boundary_flux, rtol=1e-11, atol=1e-15)
        assert rep.min_density == pytest.approx(float(rep_ref.min_density), rel=1e-12)
    with pytest.raises(ValueError, match="stage block"):
        _lib.check(lib.cfd_ctx_set_stage_block(ctx, 100))


def test_run_matches_oracle():
    g, o = _pair("sod", "rusanov", "minmod", nx=20, ny=2)
    left = g.geometry.centroid[:, 0] < 0.5
    q = np.empty(g.state_shape)
    q[:, left] = [1.0, 0.0, 0.0, 2.5]
    q[:, ~left] = [0.125, 0.0, 0.0, 0.25]
    rg = g.run(q, 0.01)
    ro = o.run(q, 0.01)
    assert rg.steps == ro.steps and rg.rejected_steps == ro.rejected_steps
    assert rg.time == ro.time == 0.01
    assert rg.max_troubled_cells == ro.max_troubled_cells
    assert np.abs(rg.state[..., 0] - ro.state[..., 0]).max() <= 1e-10
    defect = g.integrate(rg.state) - g.integrate(q) + rg.boundary_flux_integral
    assert np.abs(defect).max() < 1e-13


@pytest.mark.parametrize("flux", FLUXES)
def test_pointwise_flux_matches_oracle(flux):
    """The interface fluxes (refined hardware reciprocal and square root on
    the GPU) against the oracle's correctly rounded NumPy arithmetic on random
    subsonic and supersonic states and unit normals, to 1e-13 relative."""
    from cfd_demo_amd.physics import get_interface_flux as gpu_flux
    from oracle.physics import get_interface_flux as ref_flux, primitive_to_conserved

    rng = np.random.default_rng(31)
    n = 20000
    def states():
        rho = rng.uniform(0.05, 3.0, n)
        u, v = rng.uniform(-3.0, 3.0, (2, n))
        p = rng.uniform(0.02, 3.0, n)
        return primitive_to_conserved(rho, u, v, p, 1.4)
    qm, qp = states(), states()
    ang = rng.uniform(0.0, 2.0 * np.pi, n)
    nx, ny = np.cos(ang), np.sin(ang)
    got = gpu_flux(flux)(qm, qp, nx, ny, 1.4)
    want = ref_flux(flux)(qm, qp, nx, ny, 1.4)
    # Relative to the size of the terms the flux is assembled from (the
    # central part and the dissipation), so cancellation does not inflate it.
    from oracle.physics import normal_flux, conserved_to_primitive
    speed = lambda q: (np.abs(conserved_to_primitive(q, 1.4)[1] * nx + conserved_to_primitive(q, 1.4)[2] * ny)
                       + np.sqrt(1.4 * conserved_to_primitive(q, 1.4)[3] / q[..., 0]))
    terms = (np.abs(normal_flux(qm, nx, ny, 1.4)) + np.abs(normal_flux(qp, nx, ny, 1.4))
             + np.maximum(speed(qm), speed(qp))[:, None] * np.abs(qp - qm))
    assert np.max(np.abs(got - want) / terms.max(axis=1, keepdims=True)) <= 1e-13


def test_run_error_paths_match_oracle():
    """The reference's run-loop failures (solver.py:171-197) through the C ABI,
    with the oracle's exception class and message: an inadmissible initial
    state, and a step rejected max_rejections + 1 times (CFL 1e6)."""
    import cfd_demo_amd as gpu
    from oracle import solver as osol

    v, e, b = perturbed_vertices(4, 2, seed=3)
    g, o = make_pair(v, e, b, sod_bcs(), cfl=1e6, max_rejections=2)
    left = g.geometry.centroid[:, 0] < 0.5
    q = np.empty(g.state_shape)
    q[:, left] = [1.0, 0.0, 0.0, 2.5]
    q[:, ~left] = [0.125, 0.0, 0.0, 0.25]
    bad = q.copy()
    bad[0, 0, 0] = -1.0
    print("case run error paths", flush=True)
    msgs = []
    for s, err in ((g, gpu.SolverError), (o, osol.SolverError)):
        with pytest.raises(err, match="positive nodal density and pressure") as e1:
            s.run(bad, 0.01)
        with pytest.raises(err, match="was rejected 3 times") as e2:
            s.run(q, 0.5)
        msgs.append((str(e1.value), str(e2.value)))
    assert msgs[0] == msgs[1]


def test_averages_and_domain_integral_match_oracle():
    """cell_averages and integrate (dg.py:168-176) through the C ABI."""
    g, o = _pair("mixed", nx=7, ny=4)
    q = random_state(g.state_shape, 5)
    print("case averages and domain integral", flush=True)
    assert np.allclose(g.cell_averages(q), o.cell_averages(q), rtol=1e-14, atol=1e-15)
    ig, io = g.integrate(q), o.integrate(q)
    assert np.allclose(ig, io, rtol=1e-14, atol=0.0)


def _branch_cases():
    from test_oracle_flux_branches import HLLC_CASES, ROE_CASES, cons

    qm, qp, nx, ny = [], [], [], []
    for left, right, normal, _ in ROE_CASES + HLLC_CASES:
        qm.append(cons(left))
        qp.append(cons(right))
        nx.append(normal[0])
        ny.append(normal[1])
    # A side at zero pressure (sound speed exactly 0) and a side at rest.
    from oracle.physics import primitive_to_conserved
    for a, b in (((1.0, 0.3, 0.0, 0.0), (0.8, 0.1, 0.2, 0.5)), ((0.7, 0.0, 0.0, 0.6), (1.1, 0.0, 0.0, 0.0)),
                 ((1.0, 0.0, 0.0, 1.0), (1.0, 0.0, 0.0, 1.0))):
        for normal in ((1.0, 0.0), (0.6, -0.8)):
            qm.append(primitive_to_conserved(*a, 1.4))
            qp.append(primitive_to_conserved(*b, 1.4))
            nx.append(normal[0])
            ny.append(normal[1])
    return np.array(qm), np.array(qp), np.array(nx), np.array(ny)


@pytest.mark.parametrize("flux", FLUXES)
def test_pointwise_flux_branches_match_oracle(flux, stage_layout):
    """The Roe entropy-fix and HLLC star-branch states that
    tests/test_oracle_flux_branches.py pins independently, and sides with
    zero pressure or at rest (a sound speed or speed of exactly 0), through
    the GPU's interface flux: finite and within 1e-13 of the oracle."""
    if stage_layout != "quad":
        pytest.skip("pointwise fluxes do not depend on the stage layout")
    from cfd_demo_amd.physics import get_interface_flux as gpu_flux
    from oracle.physics import get_interface_flux as ref_flux

    qm, qp, nx, ny = _branch_cases()
    print(f"case pointwise branches flux={flux} n={len(nx)}", flush=True)
    got = gpu_flux(flux)(qm, qp, nx, ny, 1.4)
    want = ref_flux(flux)(qm, qp, nx, ny, 1.4)
    assert np.all(np.isfinite(want)) and np.all(np.isfinite(got))
    scale = np.maximum(np.abs(want).max(axis=1, keepdims=True), 1e-300)
    assert np.all(np.abs(got - want) <= 1e-13 * scale)


def _kinetic_node_state(g):
    """Uniform state with one node whose kinetic energy is ~1e10 times its
    pressure, and whose correctly rounded pressure (pressure(), no
    contraction) lies within 1e-18 below the positivity floor 1e-13."""
    from oracle.physics import pressure

    q = np.broadcast_to(np.array([1.0, 0.0, 0.0, 2.5]), g.state_shape).copy()
    rho, mom = 0.5, 0.03  # ke = 0.5 * mom^2 / rho = 9e-4
    E = 0.5 * mom * mom / rho + 0.999999e-13 / 0.4
    node = np.array([rho, mom, 0.0, E])
    for _ in range(10000):
        p = float(pressure(node, 1.4))
        if 1e-13 - 1e-18 < p < 1e-13:
            break
        node[3] = np.nextafter(node[3], -np.inf if p >= 1e-13 else np.inf)
    assert 1e-13 - 1e-18 < float(pressure(node, 1.4)) < 1e-13
    q[7, 5] = node
    return q


def test_positivity_filter_sends_kinetic_dominated_node():
    """The summary's pre-filter uses the refined hardware reciprocal; a node
    whose correctly rounded pressure is just below 1e-13 while kinetic
    energy dominates must still reach the exact repair: the limiter and a
    dt = 0 step (stage-kernel summary) both match the oracle, which repairs it."""
    from oracle.physics import pressure

    g, o = _pair("fixed", limiter="unlimited", nx=4, ny=4)
    q = _kinetic_node_state(g)
    print(f"case kinetic-dominated node p={float(pressure(q[7, 5], 1.4)):.17e}", flush=True)
    q_ref = o.limit(q)
    repair = np.abs(q_ref[:, 5] - q[:, 5]).max()
    assert repair > 0.0  # the oracle repairs the element (by about 1e-16)
    q_gpu = g.limit(q)
    assert rel_err(q_gpu, q_ref) <= 1e-12
    assert np.abs(q_gpu[:, 5] - q_ref[:, 5]).max() <= 1e-3 * repair  # repaired as the oracle did
    assert float(pressure(q_gpu[7, 5], 1.4)) >= 1e-13 * (1.0 - 1e-9)
    s_gpu, rep = g.step(q, 0.0)
    s_ref, rref = o.step(q, 0.0)
    assert rel_err(s_gpu, s_ref) <= 1e-11
    assert float(rref.min_pressure) >= 1e-13 * (1.0 - 1e-9)
    assert rep.min_pressure == pytest.approx(float(rref.min_pressure), rel=1e-10)
