#!/usr/bin/env python3
"""Benchmark: SSP-RK3 steps of the degree-4 DG Euler solver on one MI355X.

Workload (one "step" = one accepted SSP-RK3 step of the run loop: CFL time
step, three stages of RHS + combination + limiting, and the step report):
four-quadrant Riemann problem on [0,1]^2, 1000x1000 rectangles split into
2,000,000 triangles (120M fp64 DOF), HLLC flux, MC limiter, transmissive
boundaries, CFL 0.5. The state is resident in HBM before the timed region.

value = DOF-stage updates per second = 3 stages x 60 K DOF x steps / time.
roofline = the stage kernel (RHS fused with the stage combination), with
algorithmic bytes (16 + 24 + 24)/3 = 21.33 B per DOF per launch, over its mean
launch duration from HIP events on the solver's stream; also as a fraction of
the streaming ceiling measured in the same run (STREAM copy and add over
2 GiB arrays, weighted by the stages' read/write mix). fp64 = the
same launches' algorithmic flops (about 8.8k per element for HLLC) against
the 78.6 TFLOP/s FP64 peak.
cpu_baseline = the NumPy oracle (a restatement of the reference's JAX path;
the JAX path itself is not available) on all cores of this process's CPU
affinity (at most 16), its element range split over worker processes, on a
bounded sample of the same workload: per-step median of a few steps at the
full 2M-element size. --cpu-plan adds the whole CPU plan: per-RHS and
per-step medians at the Sod, vortex-plus-bumps, vortex and four-quadrant
configurations, the full Sod run, and the one-core step as a labelled extra.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-plan]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cfd-demo_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "p=4 DOF-updates/s (fp64) & % HBM roofline; batched speedup; adjoint/fwd cost"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64, vector and matrix alike
STAGE_BYTES_PER_DOF = (16.0 + 24.0 + 24.0) / 3.0
RHS_BYTES_PER_DOF = 16.0
# Algorithmic flops per element of one HLLC right-hand side (divisions and
# square roots one flop each): volume contractions 2 * (2 * 15 * 15 * 4) =
# 3600, lift 2 * 15 * 15 * 4 = 1800, pointwise fluxes and contravariant
# components 44 * 15 = 660, interface fluxes 15 * (160 + 20) = 2700. A stage
# adds the SSP-RK3 combination (at most 3 flops per DOF).
RHS_FLOPS_PER_ELEMENT_HLLC = 3600 + 1800 + 660 + 15 * (160 + 20)
STAGE_FLOPS_PER_ELEMENT_HLLC = RHS_FLOPS_PER_ELEMENT_HLLC + 3 * 60


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "unknown")
    except OSError:
        return "unknown"


def _oracle_workload(name: str, n: int):
    """(oracle solver, state) of one benchmark configuration at size n."""
    from oracle import mesh as om
    from oracle.physics import primitive_to_conserved
    from oracle.solver import EulerSolver, SolverConfig, sod_shock_tube
    from cfd_demo_amd.cases import QUADRANT_STATES, isentropic_vortex_state, vortex_bumps_state

    if name == "sod":
        s, q, _ = sod_shock_tube(n, max(1, n // 10), flux="rusanov", limiter="minmod", cfl=0.1)
        return s, q
    if name == "quadrant":
        s = EulerSolver(om.rectangle_mesh(n, n), {t: om.Transmissive() for t in ("left", "right", "bottom", "top")},
                        SolverConfig("hllc", "mc", 1.4))
        cx, cy = s.geometry.centroid.T
        st = np.empty((s.mesh.num_elements, 4))
        for key, m in (("NE", (cx > .5) & (cy > .5)), ("NW", (cx < .5) & (cy > .5)),
                       ("SW", (cx < .5) & (cy < .5)), ("SE", (cx > .5) & (cy < .5))):
            st[m] = primitive_to_conserved(*QUADRANT_STATES[key], 1.4)
        return s, np.broadcast_to(st[None], s.state_shape).copy()
    L = 10.0
    bcs = {"left": om.Periodic("right", (-L, 0.0)), "right": om.Periodic("left", (L, 0.0)),
           "bottom": om.Periodic("top", (0.0, -L)), "top": om.Periodic("bottom", (0.0, L))}
    mesh = om.rectangle_mesh(n, n, (0.0, L), (0.0, L))
    if name == "vortex":
        s = EulerSolver(mesh, bcs, SolverConfig("roe", "unlimited", 1.4))
        x, y = s.node_coordinates()
        return s, isentropic_vortex_state(x, y)
    s = EulerSolver(mesh, bcs, SolverConfig("rusanov", "unlimited", 1.4, cfl=0.1))
    x, y = s.node_coordinates()
    return s, vortex_bumps_state(x, y)


def _medians(par, solver, q, rhs_reps: int, step_reps: int) -> dict:
    """Median seconds of par.rhs(q) and par.step(q, dt) after one warm-up each."""
    dt = solver.time_step(q)
    out = {}
    K = solver.mesh.num_elements
    if rhs_reps:
        par.rhs(q)
        ts = []
        for _ in range(rhs_reps):
            t0 = time.perf_counter()
            par.rhs(q)
            ts.append(time.perf_counter() - t0)
            log(f"cpu {K} elements: oracle rhs {ts[-1]:.3f} s")
        out["rhs_median_s"] = float(np.median(ts))
    if step_reps:
        par.step(q, dt)
        ts = []
        for _ in range(step_reps):
            t0 = time.perf_counter()
            par.step(q, dt)
            ts.append(time.perf_counter() - t0)
            log(f"cpu {K} elements: oracle step {ts[-1]:.3f} s")
        out["step_median_s"] = float(np.median(ts))
    out["reps"] = [rhs_reps, step_reps]
    return out


def cpu_baseline(n: int, steps: int) -> dict:
    """The oracle's SSP-RK3 step of the four-quadrant workload at n x n
    rectangles on all cores (worker processes over element chunks)."""
    from oracle.parallel import ParallelOracle, default_workers

    t0 = time.perf_counter()
    s, q = _oracle_workload("quadrant", n)
    setup = time.perf_counter() - t0
    workers = default_workers()
    with ParallelOracle(s, workers) as par:
        med = _medians(par, s, q, 0, steps)
    dof = 60 * s.mesh.num_elements
    return {"value": 3 * dof / med["step_median_s"], "unit": "DOF-stage-updates/s", "cores": workers,
            "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"oracle (NumPy restatement of the reference JAX path) SSP-RK3 step on the four-quadrant "
                      f"HLLC/MC workload at {n}x{n} rectangles ({s.mesh.num_elements} triangles), element range "
                      f"split over {workers} worker processes (bit-identical to the serial oracle); median of "
                      f"{steps} steps after one warm-up step: {med['step_median_s']:.2f} s per step "
                      f"(oracle setup {setup:.0f} s, not timed)"}


def cpu_plan() -> dict:
    """Per-RHS and per-step medians of the oracle on all cores at the Sod
    (100x10), vortex-plus-bumps (200^2), vortex (512^2) and four-quadrant
    (1000^2) configurations (five timed calls each after one warm-up), the
    full Sod run to t = 0.2 at CFL 0.1, and the one-core step of the
    vortex-plus-bumps field as a labelled extra."""
    from oracle.parallel import ParallelOracle, default_workers

    workers = default_workers()
    out = {"cores": workers, "cpu_model": _cpu_model(), "kind": "port", "configs": {}}
    for name, n in (("sod", 100), ("bumps", 200), ("vortex", 512), ("quadrant", 1000)):
        s, q = _oracle_workload(name, n)
        with ParallelOracle(s, workers if s.mesh.num_elements > 20000 else min(workers, 4)) as par:
            med = _medians(par, s, q, 5, 5)
            med["workers"] = par.workers
            if name == "sod":
                t0 = time.perf_counter()
                qq, t, steps = q, 0.0, 0
                while t < 0.2:
                    dt = min(s.time_step(qq), 0.2 - t)
                    qq, _ = par.step(qq, dt)
                    t = 0.2 if dt == 0.2 - t else t + dt
                    steps += 1
                    if steps % 2000 == 0:
                        log(f"cpu sod full run: {steps} steps, t = {t:.4f}, {time.perf_counter() - t0:.0f} s")
                med["full_run_s"] = time.perf_counter() - t0
                med["full_run_steps"] = steps
        K = s.mesh.num_elements
        med["elements"] = K
        med["dof_stage_per_s"] = 3 * 60 * K / med["step_median_s"]
        out["configs"][name] = med
        log(f"cpu plan {name} {K}: {med}")
        if name == "bumps":
            from threadpoolctl import threadpool_limits

            with threadpool_limits(limits=1):
                t0 = time.perf_counter()
                s.step(q, s.time_step(q))
                out["one_core_bumps_step_s"] = time.perf_counter() - t0
    return out


def batched_bench(n_cases: int, steps: int, seed: int = 0) -> dict:
    """Seeded random shock tubes: one batched launch per stage for all cases
    (hipGraph replay) against the plain per-case loop of EulerSolver.run, both
    over the same number of steps per case."""
    from cfd_demo_amd.batched import BatchedSolver, random_shock_tube_specs, shock_tube_case

    cases = [shock_tube_case(s) for s in random_shock_tube_specs(n_cases, seed=seed)]
    dof = sum(60 * c.solver.mesh.num_elements for c in cases)
    batch = BatchedSolver(cases)
    batch.begin()
    sb, ms = batch.autotune(reps=3)
    batch.advance(2)  # warm-up (includes graph capture)
    batch.begin()
    t0 = time.perf_counter()
    left = batch.advance(steps)
    t_batch = time.perf_counter() - t0
    done_steps = [r.steps for r in batch.results()]
    batch.close()
    log(f"batched: {n_cases} cases, {dof} DOF, stage block {sb} (ms {ms.round(3).tolist()}), "
        f"{steps} steps in {t_batch:.3f}s, active left {left}")
    for c in cases:
        c.solver._context()
    t0 = time.perf_counter()
    loop_steps = []
    for c in cases:
        r = c.solver.run(c.initial_state, c.final_time, max_steps=steps)
        loop_steps.append(r.steps)
    t_loop = time.perf_counter() - t0
    for c in cases:
        c.solver.close()
    work = sum(3 * 60 * c.solver.mesh.num_elements * s for c, s in zip(cases, done_steps))
    work_loop = sum(3 * 60 * c.solver.mesh.num_elements * s for c, s in zip(cases, loop_steps))
    log(f"per-case loop: {t_loop:.3f}s")
    return {"cases": n_cases, "seed": seed, "dof": dof, "steps_per_case": steps, "stage_block": sb,
            "batched_dof_stage_per_s": work / t_batch, "loop_dof_stage_per_s": work_loop / t_loop,
            "speedup_vs_loop": (work / t_batch) / (work_loop / t_loop)}


def stream_ceiling(gib: int = 2, reps: int = 10) -> dict:
    """Measured streaming ceilings of this GPU, in the read/write mixes of the
    kernels they bound: STREAM copy (c = a: one read, one write, the
    right-hand side's mix) and STREAM add (c = a + b: two reads, one write,
    the mix of SSP-RK3 stages 2 and 3), each over ``gib`` GiB per array of
    fp64 with torch's elementwise kernels, median of ``reps`` launches timed
    with events on torch's current stream, the one the kernels run on."""
    import torch

    n = gib * (1 << 27)
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def med_ms(fn) -> float:
        for _ in range(3):
            fn()
        ms = []
        for _ in range(reps):
            ev0.record()
            fn()
            ev1.record()
            ev1.synchronize()
            ms.append(ev0.elapsed_time(ev1))
        return float(np.median(ms))

    t_copy = med_ms(lambda: c.copy_(a))
    t_add = med_ms(lambda: torch.add(a, b, out=c))
    del a, b, c
    torch.cuda.empty_cache()
    out = {"copy_gbs": 2.0 * n * 8 / (t_copy * 1e-3) / 1e9, "add_gbs": 3.0 * n * 8 / (t_add * 1e-3) / 1e9,
           "copy_ms": t_copy, "add_ms": t_add, "gib_per_array": gib,
           "kernels": "torch Tensor.copy_ and torch.add(out=) on fp64, device resident"}
    log(f"stream ceilings ({gib} GiB fp64 arrays): copy {out['copy_gbs']:.0f} GB/s, add {out['add_gbs']:.0f} GB/s")
    return out


def setup_bench() -> dict:
    """Host setup time of the native table assembly (cfd_build_tables through
    tables.assemble: connectivity, periodic patch, orientation, geometry,
    Cockburn-Shu coefficients) at 2M and 8M triangles, second call (the
    first pays the OS's first touch of the freshly allocated arrays)."""
    from cfd_demo_amd.boundary import Transmissive
    from cfd_demo_amd.mesh import rectangle_mesh
    from cfd_demo_amd.tables import assemble

    out = {}
    bcs = {t: Transmissive() for t in ("left", "right", "bottom", "top")}
    for n in (1000, 2000):
        t0 = time.perf_counter()
        mesh = rectangle_mesh(n, n)
        t_mesh = time.perf_counter() - t0
        ts = []
        for _ in range(2):
            t0 = time.perf_counter()
            assemble(mesh, bcs)
            ts.append(time.perf_counter() - t0)
        out[f"{2 * n * n}"] = {"rectangle_mesh_s": t_mesh, "assemble_first_s": ts[0], "assemble_s": ts[1]}
        del mesh
    log(f"native table assembly: {out}")
    return out


def batched_full_bench(n_cases: int, seed: int, step_cap: int) -> dict:
    """The batched workload as specified: every seeded tube run to its own
    final time, as one batch (hipGraph replay of one launch per flux per
    stage) and as the plain loop of EulerSolver.run over the cases. A tube
    whose CFL step collapses stops at ``step_cap`` steps on both sides (the
    reference's loop has no cap and would not end)."""
    from cfd_demo_amd.batched import BatchedSolver, random_shock_tube_specs, shock_tube_case

    cases = [shock_tube_case(s) for s in random_shock_tube_specs(n_cases, seed=seed)]
    batch = BatchedSolver(cases)
    batch.begin()
    sb, _ = batch.autotune(reps=3)
    batch.advance(2)  # warm-up (graph capture)
    t0 = time.perf_counter()
    res_b = batch.run(max_steps=step_cap)
    t_batch = time.perf_counter() - t0
    log(f"batched full runs: {t_batch:.1f}s")
    batch.close()
    for c in cases:
        c.solver._context()
    res_l = []
    t0 = time.perf_counter()
    for i, c in enumerate(cases):
        res_l.append(c.solver.run(c.initial_state, c.final_time, max_steps=step_cap))
        if i % 32 == 31:
            log(f"per-case loop: {i + 1} cases in {time.perf_counter() - t0:.1f}s")
    t_loop = time.perf_counter() - t0
    for c in cases:
        c.solver.close()
    steps_b = np.array([r.steps for r in res_b])
    steps_l = np.array([r.steps for r in res_l])
    done_b = np.array([r.time == c.final_time for r, c in zip(res_b, cases)])
    done_l = np.array([r.time == c.final_time for r, c in zip(res_l, cases)])
    work_b = sum(3 * 60 * c.solver.mesh.num_elements * int(n) for c, n in zip(cases, steps_b))
    work_l = sum(3 * 60 * c.solver.mesh.num_elements * int(n) for c, n in zip(cases, steps_l))
    diff = max(float(np.abs(a.state[..., 0] - b.state[..., 0]).max()) for a, b in zip(res_b, res_l))
    log(f"batched full runs: {n_cases} tubes, batch {t_batch:.2f}s, loop {t_loop:.2f}s, speedup "
        f"{t_loop / t_batch:.2f}, done {int(done_b.sum())}/{int(done_l.sum())}, steps {steps_b.min()}-{steps_b.max()}, "
        f"max |d rho| batch vs loop {diff:.2e}")
    return {"cases": n_cases, "seed": seed, "step_cap": step_cap, "stage_block": sb,
            "batch_wall_s": t_batch, "loop_wall_s": t_loop, "speedup_vs_loop": t_loop / t_batch,
            "batched_dof_stage_per_s": work_b / t_batch, "loop_dof_stage_per_s": work_l / t_loop,
            "cases_at_final_time": [int(done_b.sum()), int(done_l.sum())],
            "capped_cases": [int(i) for i in np.flatnonzero(~done_b)],
            "steps_min_max_total": [int(steps_b.min()), int(steps_b.max()), int(steps_b.sum())],
            "same_steps": bool(np.array_equal(steps_b, steps_l)), "max_density_diff_batch_vs_loop": diff}


def adjoint_bench(n: int, batch: int, steps: int, checkpoint_every: int, fd_dirs: int, seed: int = 0) -> dict:
    """Vortex plus 8 seeded Gaussian bumps (field b uses seed + b), periodic
    [0,10]^2 on n x n rectangles, Rusanov, unlimited, CFL 0.1; J = sum over
    the batch of int_[6,8]^2 p^2 by nodal quadrature. Times the checkpointed
    forward (frozen CFL steps) and the backward (segment recomputation plus
    adjoint kernels) through the torch.autograd wrapper's two passes, and
    checks fd_dirs random directions by central differences."""
    import torch

    from cfd_demo_amd.adjoint import DiscreteAdjoint, pressure_squared_objective, window_weights
    from cfd_demo_amd.cases import vortex_bumps, vortex_bumps_state

    case = vortex_bumps(n, flux="rusanov", limiter="unlimited", cfl=0.1)
    s = case.solver
    x, y = s.node_coordinates()
    q0 = torch.from_numpy(np.stack([vortex_bumps_state(x, y, seed=seed + b) for b in range(batch)])).cuda()
    w = torch.from_numpy(window_weights(s)).cuda()
    gamma = s.config.gamma
    run = DiscreteAdjoint(s, steps, batch=batch, checkpoint_every=checkpoint_every)
    t0 = time.perf_counter()
    run.forward(q0)  # records the CFL time steps (one host sync per step)
    t_record = time.perf_counter() - t0
    xg = q0.clone().requires_grad_(True)
    t0 = time.perf_counter()
    qn = run.forward(xg.detach())
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    qn.requires_grad_(True)
    J = pressure_squared_objective(qn, w, gamma)
    (gn,) = torch.autograd.grad(J, qn)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g0 = run.backward(gn)
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    st = run.stats()
    rng = np.random.default_rng(seed + 1000)
    worst = 0.0
    with torch.no_grad():
        for _ in range(fd_dirs):
            d = torch.from_numpy(rng.standard_normal(tuple(q0.shape))).cuda() * q0.abs()
            h = 1e-6
            jp = pressure_squared_objective(run.forward(q0 + h * d), w, gamma).item()
            jm = pressure_squared_objective(run.forward(q0 - h * d), w, gamma).item()
            fd = (jp - jm) / (2 * h)
            worst = max(worst, abs(float((g0 * d).sum()) - fd) / abs(fd))
    run.close()
    dof = 60 * s.mesh.num_elements * batch
    log(f"adjoint: {batch} x {s.mesh.num_elements} elements, {steps} steps, forward {t_fwd:.3f}s, backward "
        f"{t_bwd:.3f}s (recompute {st['recompute_ms'] / 1e3:.3f}s), ratio {t_bwd / t_fwd:.2f}, fd rel {worst:.2e}")
    return {"fields": batch, "elements_per_field": s.mesh.num_elements, "steps": steps,
            "checkpoint_every": checkpoint_every, "forward_s": t_fwd, "backward_s": t_bwd,
            "recompute_s": st["recompute_ms"] / 1e3, "adjoint_over_forward": t_bwd / t_fwd,
            "forward_dof_stage_per_s": 3 * dof * steps / t_fwd, "record_forward_s": t_record,
            "fd_directions": fd_dirs, "fd_max_rel_err": worst, "device_gib": st["device_bytes"] / 2 ** 30}


TRAFFIC_SOURCE = ("not measured in this run: HBM bytes per launch read from profiles/stage_kernel_c2_traffic.json "
                  "(rocprofv3 FETCH_SIZE and WRITE_SIZE on this workload, scaled by a 1 GiB copy kernel)")


def traffic_from_profile(kernel_prefix: str):
    p = ROOT / "profiles" / "stage_kernel_c2_traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return d.get(kernel_prefix)
    except Exception:
        return None


def c2_full_bench(n: int, segments: int = 6) -> dict:
    """The four-quadrant workload run to its final time t = 0.3 through
    ``EulerSolver.run``, as ``segments`` successive calls over 0.3/segments
    each (the equations are autonomous; each call clamps its last step to its
    end time), so every call returns within seconds and reports progress:
    steps, wall time, DOF-stage updates/s, the conservation defect against
    the boundary-flux integral, the minima and the troubled maximum."""
    from cfd_demo_amd.cases import four_quadrant

    case = four_quadrant(n)
    s, q0, T = case.solver, case.initial_state, case.final_time
    s.run(q0, 1e-6)  # context and tables on the device
    q, steps, wall, bflux = q0, 0, 0.0, np.zeros(4)
    rmin, pmin, troubled = np.inf, np.inf, 0
    for k in range(segments):
        t0 = time.perf_counter()
        r = s.run(q, T / segments)
        wall += time.perf_counter() - t0
        q, steps = r.state, steps + r.steps
        bflux += r.boundary_flux_integral
        rmin, pmin, troubled = min(rmin, r.min_density), min(pmin, r.min_pressure), max(troubled, r.max_troubled_cells)
        log(f"c2 full run: segment {k + 1}/{segments}, {steps} steps, {wall:.1f}s")
    K = s.mesh.num_elements
    defect = s.integrate(q) - s.integrate(q0) + bflux
    scale = s.integrate(np.abs(q0))
    out = {"elements": K, "final_time": T, "segments": segments, "steps": steps, "wall_s": wall,
           "ms_per_step": 1e3 * wall / steps, "dof_stage_per_s": 3.0 * 60 * K * steps / wall,
           "conservation_defect_rel": float(np.max(np.abs(defect) / scale)), "min_density": rmin,
           "min_pressure": pmin, "max_troubled_cells": troubled}
    log(f"c2 full run to t = {T}: {steps} steps in {wall:.1f}s ({out['ms_per_step']:.3f} ms/step), "
        f"defect {out['conservation_defect_rel']:.1e}, min rho {rmin:.3e}, min p {pmin:.3e}")
    s.close()
    return out


def cases_bench(c3_n: int) -> dict:
    """Full runs through ``EulerSolver.run``: the reference's Sod case (c1:
    100x10, Rusanov, minmod, CFL 0.1, t = 0.2) and the isentropic vortex (c3:
    periodic [0,10]^2, c3_n^2 rectangles, Roe, unlimited, CFL 0.5, t = 2),
    each with its L1 density error against the exact solution and its
    conservation defect."""
    from cfd_demo_amd.cases import isentropic_vortex, isentropic_vortex_state, sod_shock_tube

    sys.path.insert(0, str(ROOT / "tests"))
    from sod_exact import sample

    def l1(solver, rho, rho_ex, area):
        w = solver.reference.cell_weights[:, None] * solver.geometry.jacobian[None, :]
        return float(np.sum(w * np.abs(rho - rho_ex)) / area)

    out = {}
    case = sod_shock_tube(100, 10, flux="rusanov", limiter="minmod", cfl=0.1)
    s, q0, T = case.solver, case.initial_state, case.final_time
    s.run(q0, 1e-3)  # context and tables on the device
    t0 = time.perf_counter()
    r = s.run(q0, T)
    wall = time.perf_counter() - t0
    x, _ = s.node_coordinates()
    xu, inv = np.unique(x, return_inverse=True)
    rho_ex = sample(xu, T, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))[0][inv].reshape(x.shape)
    defect = s.integrate(r.state) - s.integrate(q0) + r.boundary_flux_integral
    out["c1_sod"] = {"elements": s.mesh.num_elements, "steps": r.steps, "wall_s": wall,
                     "ms_per_step": 1e3 * wall / r.steps, "l1_density_vs_exact": l1(s, r.state[..., 0], rho_ex, 0.1),
                     "conservation_defect": float(np.abs(defect).max())}
    log(f"c1 Sod 100x10: {r.steps} steps in {wall:.2f}s ({1e3 * wall / r.steps:.3f} ms/step), "
        f"L1 {out['c1_sod']['l1_density_vs_exact']:.3e}, defect {out['c1_sod']['conservation_defect']:.1e}")
    s.close()

    case = isentropic_vortex(c3_n, flux="roe", limiter="unlimited")
    s, q0, T = case.solver, case.initial_state, case.final_time
    s.run(q0, 1e-3)
    t0 = time.perf_counter()
    r = s.run(q0, T)
    wall = time.perf_counter() - t0
    x, y = s.node_coordinates()
    exact = isentropic_vortex_state(x, y, T)
    defect = s.integrate(r.state) - s.integrate(q0) + r.boundary_flux_integral
    K = s.mesh.num_elements
    out["c3_vortex"] = {"elements": K, "steps": r.steps, "wall_s": wall, "ms_per_step": 1e3 * wall / r.steps,
                        "dof_stage_per_s": 3.0 * 60 * K * r.steps / wall,
                        "l1_density_vs_exact": l1(s, r.state[..., 0], exact[..., 0], 100.0),
                        "conservation_defect": float(np.abs(defect).max())}
    log(f"c3 vortex {c3_n}^2 ({K} elements): {r.steps} steps in {wall:.2f}s "
        f"({1e3 * wall / r.steps:.3f} ms/step, {out['c3_vortex']['dof_stage_per_s']:.3e} DOF-stage/s), "
        f"L1 {out['c3_vortex']['l1_density_vs_exact']:.3e}, defect {out['c3_vortex']['conservation_defect']:.1e}")
    s.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1000, help="rectangles per side (K = 2 n^2)")
    ap.add_argument("--cpu-n", type=int, default=1000, help="rectangles per side of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=2, help="timed oracle steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-plan", action="store_true",
                    help="also time the whole CPU plan (about 5 minutes): medians at four configurations")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-cases", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=30)
    ap.add_argument("--c2-full", action="store_true",
                    help="also run the four-quadrant workload to its final time t = 0.3 (about 48,000 steps)")
    ap.add_argument("--batched-full", action="store_true",
                    help="run every batched tube to its own final time, batch and per-case loop")
    ap.add_argument("--batched-cap", type=int, default=40000, help="step cap of a tube whose CFL step collapses")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-cases", action="store_true", help="skip the full c1 (Sod) and c3 (vortex) runs")
    ap.add_argument("--c3-n", type=int, default=512)
    ap.add_argument("--adjoint-n", type=int, default=200)
    ap.add_argument("--adjoint-batch", type=int, default=64)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-checkpoint", type=int, default=20)
    ap.add_argument("--adjoint-fd", type=int, default=2)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--cpu-only", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_only:
        out = {"cpu_baseline": cpu_baseline(args.cpu_n, args.cpu_steps)}
        if args.cpu_plan:
            out["cpu_plan"] = cpu_plan()
        print(json.dumps(out), flush=True)
        return 0
    if args.copy_probe:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        log("copy probe: 3 x (1 GiB read + 1 GiB write) through an elementwise kernel")
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported on this build "
              "(every GPU box has one GPU; see the sharding design for more).", file=sys.stderr)
        return 2

    from cfd_demo_amd import _lib
    from cfd_demo_amd.cases import four_quadrant

    t0 = time.perf_counter()
    case = four_quadrant(args.n, flux="hllc", limiter="mc")
    solver = case.solver
    K = solver.mesh.num_elements
    dof = 60 * K
    log(f"setup K={K} in {time.perf_counter() - t0:.1f}s")
    lib = _lib.load()
    ctx = solver._context()
    q0 = np.ascontiguousarray(case.initial_state)
    _lib.check(lib.cfd_state_upload(ctx, _lib.dptr(q0)))
    rep = _lib.CfdStepReport()
    dt = C.c_double(0.0)
    _lib.check(lib.cfd_state_advance(ctx, args.warmup, C.byref(dt), C.byref(rep)))
    _lib.check(lib.cfd_synchronize(ctx))
    log(f"warmup {args.warmup} steps done, dt={dt.value:.3e}")

    t0 = time.perf_counter()
    _lib.check(lib.cfd_state_advance(ctx, args.steps, C.byref(dt), C.byref(rep)))
    _lib.check(lib.cfd_synchronize(ctx))
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = 3.0 * dof * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.4g} DOF-stage/s, "
        f"troubled={rep.troubled_cells}")

    # Kernel durations from HIP events on the solver's stream.
    _lib.check(lib.cfd_timing_enable(ctx, 1))
    _lib.check(lib.cfd_state_advance(ctx, 3, C.byref(dt), C.byref(rep)))
    tim = np.zeros(4)
    _lib.check(lib.cfd_timing_read(ctx, _lib.dptr(tim)))
    _lib.check(lib.cfd_timing_enable(ctx, 0))
    stage_ms = tim[0] / max(tim[1], 1)
    limit_ms = tim[2] / max(tim[3], 1)
    rhs_ms = C.c_double(0.0)
    _lib.check(lib.cfd_time_rhs(ctx, 10, C.byref(rhs_ms)))
    stage_gbs = STAGE_BYTES_PER_DOF * dof / (stage_ms * 1e-3) / 1e9
    rhs_gbs = RHS_BYTES_PER_DOF * dof / (rhs_ms.value * 1e-3) / 1e9
    log(f"stage kernel {stage_ms:.3f} ms ({stage_gbs:.0f} GB/s), limiter {limit_ms:.3f} ms, "
        f"rhs kernel {rhs_ms.value:.3f} ms ({rhs_gbs:.0f} GB/s)")

    # RHS on two inputs, median of 20 single-launch timings after 3 warm-ups:
    # the evolved state, and the same state with a seeded +-1 % multiplicative
    # perturbation of density and pressure (no flat regions with trivial fluxes).
    def rhs_median(reps: int = 20) -> float:
        one = C.c_double(0.0)
        for _ in range(3):
            _lib.check(lib.cfd_time_rhs(ctx, 1, C.byref(one)))
        samples = []
        for _ in range(reps):
            _lib.check(lib.cfd_time_rhs(ctx, 1, C.byref(one)))
            samples.append(one.value)
        return float(np.median(samples))

    rhs_med = rhs_median()
    q_ev = np.empty_like(q0)
    _lib.check(lib.cfd_state_download(ctx, _lib.dptr(q_ev)))
    rng = np.random.default_rng(0)
    gam = solver.config.gamma
    rho = q_ev[..., 0] * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, q_ev.shape[:2]))
    u, v = q_ev[..., 1] / q_ev[..., 0], q_ev[..., 2] / q_ev[..., 0]
    p = (gam - 1.0) * (q_ev[..., 3] - 0.5 * (q_ev[..., 1] * u + q_ev[..., 2] * v))
    p = p * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, q_ev.shape[:2]))
    q_pt = np.ascontiguousarray(np.stack([rho, rho * u, rho * v, p / (gam - 1.0) + 0.5 * rho * (u * u + v * v)], -1))
    _lib.check(lib.cfd_state_upload(ctx, _lib.dptr(q_pt)))
    rhs_med_pt = rhs_median()
    log(f"rhs kernel median of 20: {rhs_med:.3f} ms (evolved state), {rhs_med_pt:.3f} ms (perturbed)")

    traffic = traffic_from_profile("stage_kernel")
    rhs_traffic = traffic_from_profile("rhs_kernel")
    stream = stream_ceiling()
    # The stage launches' read/write mix: stage 1 reads one state and writes
    # one, stages 2 and 3 read two and write one; their byte-weighted ceiling.
    stage_ceiling = (16.0 + 24.0 + 24.0) / (16.0 / stream["copy_gbs"] + 2 * 24.0 / stream["add_gbs"])
    stage_tflops = STAGE_FLOPS_PER_ELEMENT_HLLC * K / (stage_ms * 1e-3) / 1e12
    rhs_tflops = RHS_FLOPS_PER_ELEMENT_HLLC * K / (rhs_ms.value * 1e-3) / 1e12
    log(f"stage kernel {stage_gbs / stage_ceiling:.1%} of the measured streaming ceiling ({stage_ceiling:.0f} GB/s), "
        f"FP64 {stage_tflops:.1f} TFLOP/s ({stage_tflops / FP64_PEAK_TFLOPS:.1%} of {FP64_PEAK_TFLOPS}); "
        f"rhs {rhs_gbs / stream['copy_gbs']:.1%} of copy, {rhs_tflops:.1f} TFLOP/s")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-stage-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"four-quadrant Riemann problem, [0,1]^2, {args.n}x{args.n} rectangles split into {K} "
                        "p=4 triangles (15 nodes), HLLC flux, MC limiter, transmissive sides, SSP-RK3 at CFL 0.5; "
                        "one step = CFL dt + 3 x (RHS + stage combine + limiting) + report",
            "elements": K,
            "dof": dof,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": stage_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": stage_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": TRAFFIC_SOURCE,
            "kernel": "k_stage4<HLLC> (RHS + SSP-RK3 combine + limiter summary, four lanes per element), mean of the 3 stage launches",
            "algorithmic_bytes_per_launch": STAGE_BYTES_PER_DOF * dof,
            "mean_launch_ms": stage_ms,
            "measured_ceiling_gbs": stage_ceiling,
            "frac_of_measured_ceiling": stage_gbs / stage_ceiling,
        },
        "fp64": {
            "achieved": stage_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": stage_tflops / FP64_PEAK_TFLOPS,
            "flops_per_element_stage": STAGE_FLOPS_PER_ELEMENT_HLLC,
            "rhs_achieved": rhs_tflops, "rhs_frac": rhs_tflops / FP64_PEAK_TFLOPS,
            "flops_per_element_rhs": RHS_FLOPS_PER_ELEMENT_HLLC,
        },
        "stream_ceiling": stream,
        "rhs_roofline": {
            "bound": "hbm", "achieved": rhs_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": rhs_gbs / HBM_PEAK_GBS,
            "kernel": "k_stage4<HLLC, RHS only>", "traffic": rhs_traffic, "traffic_source": TRAFFIC_SOURCE,
            "algorithmic_bytes_per_launch": RHS_BYTES_PER_DOF * dof,
            "mean_launch_ms": rhs_ms.value,
            "median_launch_ms": rhs_med,
            "median_launch_ms_perturbed_input": rhs_med_pt,
            "frac_of_measured_ceiling": rhs_gbs / stream["copy_gbs"],
        },
        "limiter_mean_launch_ms": limit_ms,
        "cpu_baseline": None,
    }
    if not args.no_batched:
        line["batched"] = batched_bench(args.batched_cases, args.batched_steps)
    line["setup"] = setup_bench()
    if args.c2_full:
        line["c2_full_run"] = c2_full_bench(args.n)
    if args.batched_full:
        line["batched_full"] = batched_full_bench(args.batched_cases, 0, args.batched_cap)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_bench(args.adjoint_n, args.adjoint_batch, args.adjoint_steps,
                                        args.adjoint_checkpoint, args.adjoint_fd)
    if not args.no_cases:
        line["cases"] = cases_bench(args.c3_n)
    if not args.no_cpu:
        log(f"cpu baseline: oracle at {args.cpu_n}x{args.cpu_n}, {args.cpu_steps} steps")
        # A child process that has not opened the GPU runs the oracle and
        # forks its workers, so none of them holds the device.
        cmd = [sys.executable, "-u", str(Path(__file__).resolve()), "--cpu-only", "--cpu-n", str(args.cpu_n),
               "--cpu-steps", str(args.cpu_steps)] + (["--cpu-plan"] if args.cpu_plan else [])
        import subprocess

        res = subprocess.run(cmd, stdout=subprocess.PIPE, check=True, close_fds=True)
        cpu = json.loads(res.stdout.decode().strip().splitlines()[-1])
        line["cpu_baseline"] = cpu["cpu_baseline"]
        if "cpu_plan" in cpu:
            line["cpu_plan"] = cpu["cpu_plan"]
        log(f"cpu baseline {line['cpu_baseline']['value']:.4g} DOF-stage/s on {line['cpu_baseline']['cores']} cores")
    print(json.dumps(line), flush=True)
    solver.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
