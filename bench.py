#!/usr/bin/env python3
"""Benchmark of the explicit time-domain Maxwell hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: FP32 Yee FDTD on a 256^3 model with a 10-cell CPML on all faces,
eps_r from seeded random ellipsoids in vacuum and a Ricker source (one model
of the batched uneven-size FDTD workload).  A step is one full E+H time step
over the grid.  Prints one JSON line with the whole-job throughput in
Mcell-updates/s, the roofline of the Yee update kernel (HIP events on its
stream, algorithmic bytes per launch) and the CPU baseline.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def gpu_fdtd(n, steps, warmup, precision, kernel, npml=10, seed=11):
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.fdtd import FemSimulation, WAVELET_RICKER
    h = 1e-3
    dt = wl.cfl_dt(h, 0.99)
    er = wl.ellipsoid_model(n, seed)
    src, f0, recs = wl.shot(n, seed)
    sim = FemSimulation((n, n, n), h, dt, wl.EPS0, wl.MU0, 1, precision=precision)
    sim.set_kernel(kernel)
    sim.set_material(er, None)
    sim.set_cpml(npml, 3.0, 1e-6, 1.0, 0.0)
    sim.set_wavelet(WAVELET_RICKER, f0, -1.0, 1.0)
    sim.setSource_z(src)
    for r in recs:
        sim.setObservationPoint([[r], [r], [r]])
    if warmup > 0:
        sim.run(warmup)
    total_ms, kernel_ms = sim.run_timed(steps)
    tr = sim.traces()
    assert np.isfinite(tr).all(), "non-finite traces"
    word = 4 if precision == "f32" else 8
    step_bytes, per_cell = wl.fdtd_cell_bytes(n, npml, word, 1)
    return dict(total_ms=total_ms, kernel_ms=kernel_ms, step_bytes=step_bytes, per_cell=per_cell,
                cells=n ** 3, device_bytes=sim.device_bytes())


def gpu_fem(order, stepper, n, steps, precision="f64", elem="auto"):
    """Structured-hex FEM (kmat operators) on an n^3 fine grid; returns
    DOF-updates/s of the step kernels (HIP events).  elem: the element
    products of orders 2 and 3 ('auto', 'rows' or 'mfma')."""
    from emsolvers_amd.fem import FemSimulation
    from emsolvers_amd.workloads import EPS0, MU0, C0
    kmat = os.path.join(ROOT, "tests", "golden", "kmat")
    h = 1e-3
    dt = h / (C0 * 3 ** 0.5) / 2.0
    sim = FemSimulation(order, (n, n, n), h, dt, EPS0, MU0, 1, 1, stepper=stepper, precision=precision,
                        kmat_dir=kmat)
    sim.set_element_product(elem)
    sim.setSource_z((n // 2, n // 2, n // 2))
    sim.setObservationPoint((n // 2 + order, n // 2, n // 2))
    sim.run(2)
    total_ms, kernel_ms = sim.run_timed(steps)
    ndof = sim.ndof()
    return dict(value=round(ndof * steps / (kernel_ms / 1e3) / 1e6, 1), unit="MDOF-updates/s",
                dofs=ndof, ms_per_step=round(kernel_ms / steps, 4), order=order, stepper=stepper,
                grid=[n] * 3, dtype="fp64" if precision == "f64" else "fp32", element_product=elem)


def gpu_fem1(stepper, n, steps, precision="f64", elem="auto"):
    """First-order stabilized hex FEM (src/fem leapfrog, src/fem_abm) on n^3
    cells with a Hertz dipole; DOF-updates/s of the step kernels."""
    from emsolvers_amd.fem1 import FemSimulation as Fem, AbmSimulation as Abm
    from emsolvers_amd.workloads import EPS0, MU0, C0
    h = 4.0 / n
    dt = 0.3 * h / C0
    s = Fem((n, n, n), h, dt, EPS0, MU0, 1, 0.3, 0, precision=precision) if stepper == "fem" else \
        Abm((n, n, n), h, dt, EPS0, MU0, 1, precision=precision)
    s.set_element_product(elem)
    s.setSourceDipole_x((n // 2, n // 2, n // 2))
    p = (n // 2 + 5, n // 2, n // 2)
    s.setObservationPoint([[p], [p], [p]])
    s.run(2)
    total_ms, kernel_ms = s.run_timed(steps)
    nd = s.ndof()
    return dict(value=round(nd * steps / (kernel_ms / 1e3) / 1e6, 1), unit="MDOF-updates/s", dofs=nd,
                ms_per_step=round(kernel_ms / steps, 4), order=1, stepper=stepper, grid=[n] * 3,
                dtype="fp64" if precision == "f64" else "fp32", element_product=elem)


def gpu_tet(n=40, stepper="symplectic_2", steps=2000, precision="f64", order=1, seed=5, sigma_fraction=0.0,
            power_iters=300, op="auto", solver="auto"):
    """Tetrahedral Nedelec elements (order 1 or 2) on the Kuhn split of an n^3
    unit cube: seeded smoothed log-uniform eps_r (sigma uniform in [0, 0.05]
    S/m on a seeded fraction of the tetrahedra), x-dipole (Gaussian pulse)
    at the centre, 16 random interior receivers, dt = 0.9 x the stepper's
    stable step (kappa / w_max, w_max by power iteration on the same handle);
    DOF-updates/s over the step sequence (HIP events) and the conjugate-
    gradient iterations per solve."""
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.tet import TetSimulation, KAPPA, STEPPERS
    h = 1.0 / n
    er, sg = wl.tet_model(n, seed, sigma_fraction)
    t0 = time.time()
    s = TetSimulation((n, n, n), h, 1e-3 * h / wl.C0, wl.EPS0, wl.MU0, 1, stepper, order, precision)
    s.set_operator(op)
    s.set_solver(solver)
    s.set_material(er, sg)
    wmax = s.max_frequency(power_iters)
    setup_s = time.time() - t0
    dt = 0.9 * KAPPA[STEPPERS[stepper]] / wmax
    s.set_time_step(dt)
    f0 = 2.0 * wl.C0 / (h * n)  # two vacuum wavelengths across the cube
    s.add_dipole([0.5 + 0.1 * h, 0.5 + 0.2 * h, 0.5 + 0.3 * h], [1.0, 0.0, 0.0], f0, 1.2 / f0, 1e-9)
    for p in wl.tet_receivers(n, h, 16, seed + 1):
        s.add_receiver(p)
    ms = s.run_timed(steps)
    last, mx = s.solver_stats()
    nd = s.dof_count()
    tr = s.traces()
    out = dict(value=round(nd * steps / (ms / 1e3) / 1e6, 2), unit="MDOF-updates/s", dofs=nd, tets=6 * n ** 3,
               ms_per_step=round(ms / steps, 4), steps=steps, stepper=stepper, order=order, dt=dt,
               omega_max=wmax, cg_iters_last=last, cg_iters_max=mx, setup_s=round(setup_s, 2),
               finite=bool(np.all(np.isfinite(tr))), dtype="fp64" if precision == "f64" else "fp32",
               sigma_fraction=sigma_fraction, operator=op, solver=solver)
    s.close()
    return out


def gpu_tet_gradient(n=40, steps=200, precision="f64", seed=5):
    """Misfit gradient of a lossless symplectic_4 tetrahedral run (0.46 M
    DOFs at n = 40) by the discrete adjoint with the forward states recovered
    by running the integrator backwards; time per forward+adjoint step, the
    forward step alone, and the reconstruction drift."""
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.tet import TetSimulation, KAPPA, STEPPERS
    h = 1.0 / n
    er, _ = wl.tet_model(n, seed)
    s = TetSimulation((n, n, n), h, 1e-3 * h / wl.C0, wl.EPS0, wl.MU0, 1, "symplectic_4", 1, precision)
    s.set_material(er)
    dt = 0.9 * KAPPA[STEPPERS["symplectic_4"]] / s.max_frequency(300)
    s.set_time_step(dt)
    f0 = 2.0 * wl.C0 / (h * n)
    s.add_dipole([0.5 + 0.1 * h, 0.5 + 0.2 * h, 0.5 + 0.3 * h], [1.0, 0.0, 0.0], f0, 1.2 / f0, 1e-9)
    for p in wl.tet_receivers(n, h, 16, seed + 1):
        s.add_receiver(p)
    fwd_ms = s.run_timed(steps)
    d_obs = 0.9 * s.traces()
    s2 = TetSimulation((n, n, n), h, dt, wl.EPS0, wl.MU0, 1, "symplectic_4", 1, precision)
    s2.set_material(er * 1.02)
    s2.add_dipole([0.5 + 0.1 * h, 0.5 + 0.2 * h, 0.5 + 0.3 * h], [1.0, 0.0, 0.0], f0, 1.2 / f0, 1e-9)
    for p in wl.tet_receivers(n, h, 16, seed + 1):
        s2.add_receiver(p)
    s2.gradient(2, d_obs[:, :3, :])  # warm (tables, buffers)
    t0 = time.time()
    J, ge, gs, drift = s2.gradient(steps, d_obs)
    sec = time.time() - t0
    out = dict(dofs=s2.dof_count(), steps=steps, stepper="symplectic_4", ms_per_step_gradient=round(sec * 1e3 / steps, 3),
               ms_per_step_forward=round(fwd_ms / steps, 3), ratio_to_forward=round(sec * 1e3 / fwd_ms, 2),
               reconstruction_drift=drift, misfit=J, grad_finite=bool(np.all(np.isfinite(ge))),
               dtype="fp64" if precision == "f64" else "fp32 fields, fp64 gradient")
    s.close()
    s2.close()
    return out


def _shot(n, seed, model, npml=10, receivers=(8, 4), z_rec=None, src=None, f0=None, precision="f32"):
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.fdtd import FemSimulation, WAVELET_RICKER
    h = 1e-3
    s = FemSimulation((n, n, n), h, wl.cfl_dt(h, 0.99), wl.EPS0, wl.MU0, 1, precision=precision)
    s.set_cpml(npml, 3.0, 1e-6, 1.0, 0.0)
    rng = np.random.default_rng(seed)
    s.set_wavelet(WAVELET_RICKER, f0 if f0 else rng.uniform(5e9, 15e9), -1.0, 1.0)
    s.set_material(*model)
    if src is None:
        src = tuple(int(v) for v in rng.integers(n // 4, 3 * n // 4, 3))
    s.setSource_z(src)
    ra, rb = receivers
    for a in range(ra):
        for b in range(rb):
            p = (n // 8 + a * (3 * n // 4) // max(ra - 1, 1), n // 8 + b * (3 * n // 4) // max(rb - 1, 1),
                 z_rec if z_rec is not None else n // 2)
            s.setObservationPoint([[p], [p], [p]])
    return s


def gpu_batch_forward(steps, per_model=16, sizes=(96, 128, 192, 256)):
    """Shots x models of uneven size (ellipsoid eps_r, CPML 10, Ricker f0 in
    [5, 15] GHz, 32 receivers each): sequential single-shot runs against one
    batch (autotuned tiling)."""
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.fdtd import Batch
    models = {n: (wl.ellipsoid_model(n, n), None) for n in sizes}
    mem = [_shot(n, 1000 * n + i, models[n]) for n in sizes for i in range(per_model)]
    cells = sum(m.grid[0] * m.grid[1] * m.grid[2] for m in mem)
    for m in mem:
        m.run(1)
    t0 = time.perf_counter()
    for m in mem:
        m.run(steps)
    seq = time.perf_counter() - t0
    b = Batch(mem)
    rows, planes = b.autotune(3)
    b.run(1)
    bat_ms = b.run_timed(steps)
    # the same steps from a HIP graph of 10 steps, captured and instantiated
    # inside the timed call, then replayed
    b.set_graph(10)
    graph_ms = b.run_timed(steps)
    return dict(shots=len(mem), models=list(sizes), steps=steps, cells_per_step=cells,
                sequential_s=round(seq, 4), batched_s=round(bat_ms / 1e3, 4),
                batched_graph_s=round(graph_ms / 1e3, 4),
                sequential_gcells=round(cells * steps / seq / 1e9, 2),
                batched_gcells=round(cells * steps / (bat_ms / 1e3) / 1e9, 2),
                batched_graph_gcells=round(cells * steps / (graph_ms / 1e3) / 1e9, 2),
                ratio=round(seq / (bat_ms / 1e3), 3), ratio_graph=round(seq / (graph_ms / 1e3), 3),
                tiling=dict(rows=rows, planes=planes))


def gpu_batch_gradient(n=192, shots=16, steps=500):
    """Misfit gradient (eps_r and sigma) of `shots` shots on a 4 x 4 source grid
    at z = 20 with 64 receivers on the plane z = n - 20, background eps_r 2 with
    three Gaussian blobs: sequential single-shot gradients against one batched
    gradient, and the batched forward run of the same shots.  Wall-clock,
    second call of each (buffers allocated); default checkpoint interval."""
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.fdtd import Batch
    model = wl.blob_model(n, 7)
    obs = np.zeros((64, steps + 1, 3))
    # 4 x 4 source grid, spacing 30 cells at n = 192 (scaled with n), at z = 20
    # (scaled, at least 12), repeated with a 1-cell shift per further 16 shots
    sp = max(1, round(30 * n / 192))
    zs = max(12, round(20 * n / 192))
    srcs = [(n // 2 - (3 * sp) // 2 + sp * (i % 4) + i // 16, n // 2 - (3 * sp) // 2 + sp * ((i // 4) % 4), zs)
            for i in range(shots)]
    sims = [_shot(n, i, model, receivers=(8, 8), z_rec=n - zs, src=srcs[i], f0=10e9) for i in range(shots)]
    for s in sims:
        s.gradient(steps, obs)
    t0 = time.perf_counter()
    for s in sims:
        s.gradient(steps, obs)
    seq = time.perf_counter() - t0
    del sims
    mem = [_shot(n, i, model, receivers=(8, 8), z_rec=n - zs, src=srcs[i], f0=10e9) for i in range(shots)]
    b = Batch(mem)
    b.gradient(steps, [obs] * shots)
    t0 = time.perf_counter()
    b.gradient(steps, [obs] * shots)
    bat = time.perf_counter() - t0
    fwd_ms = b.run_timed(steps)
    return dict(grid=[n] * 3, shots=shots, steps=steps, sequential_s=round(seq, 3), batched_s=round(bat, 3),
                ratio=round(seq / bat, 3), ms_per_shot_step_batched=round(bat * 1e3 / shots / steps, 4),
                forward_ms_per_shot_step=round(fwd_ms / shots / steps, 4),
                gradient_to_forward=round(bat * 1e3 / fwd_ms, 3), dtype="fp32 fields, fp64 gradient")


def _trace_errors(tr, ref, dt):
    """Relative L2 of FP32 traces against FP64 ones, and the largest TF
    envelope / phase misfit (Kristekova et al., emsolvers_amd.tfmisfit) over
    the components carrying at least 1e-2 of the peak."""
    from emsolvers_amd import tfmisfit
    fin = np.isfinite(ref)
    rel = float(np.linalg.norm((tr - ref)[fin]) / np.linalg.norm(ref[fin]))
    nslot = ref.shape[1]
    fmin, fmax = 2.0 / (nslot * dt), 0.25 / dt
    floor = 1e-2 * np.abs(ref[fin]).max()
    em = pm = 0.0
    for o in range(ref.shape[0]):
        for c in range(ref.shape[2]):
            r = ref[o, :, c]
            if not np.all(np.isfinite(r)) or not np.abs(r).max() > floor:
                continue
            m = tfmisfit.misfits(tr[o, :, c], r, dt, fmin, fmax, nf=40)
            em, pm = max(em, m["EM"]), max(pm, m["PM"])
    return dict(rel_l2=rel, tf_em=em, tf_pm=pm)


def fp32_errors():
    """FP32 against FP64 on the same inputs for each benchmark configuration
    (the FP64 GPU path equals the FP64 oracle: bit for bit for the FDTD and
    hex FEM, within 1e-10 for tetrahedra)."""
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.fdtd import FemSimulation as Fdtd, WAVELET_RICKER
    from emsolvers_amd.tet import TetSimulation, KAPPA, STEPPERS
    from emsolvers_amd.fem import FemSimulation as Hex
    out = {}

    # config 1: 64^3 FDTD, sphere eps_r 4 (r = 16), CPML 10, Ricker 10 GHz at
    # Ez (16, 32, 32), 8 receivers on y = z = 32, 1000 steps
    def cfg1(prec):
        n, h = 64, 1e-3
        s = Fdtd((n, n, n), h, wl.cfl_dt(h), wl.EPS0, wl.MU0, 1, precision=prec)
        s.set_material(wl.sphere_model(n, 16.0, 4.0, (32, 32, 32)), None)
        s.set_cpml(10, 3.0, 1e-6, 1.0, 0.0)
        s.set_wavelet(WAVELET_RICKER, 10e9, -1.0, 1.0)
        s.setSource_z((16, 32, 32))
        for r in range(8):
            p = (10 + int(round((r + 0.5) * 44 / 8)), 32, 32)
            s.setObservationPoint([[p], [p], [p]])
        s.run(1000)
        return s.traces()
    out["config1_fdtd_64_sphere_cpml"] = _trace_errors(cfg1("f32"), cfg1("f64"), wl.cfl_dt(1e-3))

    # config 2: tetrahedra p = 1 on 40^3 cubes, symplectic_2 and ABM
    def tet(n, order, stepper, prec, steps, sigma_fraction, dt=None):
        h = 1.0 / n
        er, sg = wl.tet_model(n, 5, sigma_fraction)
        s = TetSimulation((n, n, n), h, 1e-3 * h / wl.C0, wl.EPS0, wl.MU0, 1, stepper, order, prec)
        s.set_material(er, sg)
        if dt is None:
            dt = 0.9 * KAPPA[STEPPERS[stepper]] / s.max_frequency(40)
        s.set_time_step(dt)
        f0 = 2.0 * wl.C0 / (h * n)
        s.add_dipole([0.5 + 0.1 * h, 0.5 + 0.2 * h, 0.5 + 0.3 * h], [1.0, 0.0, 0.0], f0, 1.2 / f0, 1e-9)
        for p in wl.tet_receivers(n, h, 16, 6):
            s.add_receiver(p)
        s.run(steps)
        tr = s.traces()
        s.close()
        return tr, dt
    for stepper in ("symplectic_2", "abm"):
        ref, dt = tet(40, 1, stepper, "f64", 600, 0.0)
        tr, _ = tet(40, 1, stepper, "f32", 600, 0.0, dt)
        out["config2_tet_p1_40_" + stepper] = _trace_errors(tr, ref, dt)

    # config 3: hex p = 2 (64^3 elements) and tetrahedra p = 2 (64^3 cubes, sigma
    # on 20 % of the tetrahedra), symplectic_4
    def hex2(prec):
        kmat = os.path.join(ROOT, "tests", "golden", "kmat")
        n, h = 128, 1e-3
        dt = h / (wl.C0 * 3 ** 0.5) / 2.0
        s = Hex(2, (n, n, n), h, dt, wl.EPS0, wl.MU0, 1, 1, stepper="symplectic_4", precision=prec, kmat_dir=kmat)
        s.setSource_z((n // 2, n // 2, n // 2))
        for d in (4, 12, 24):
            s.setObservationPoint((n // 2 + d, n // 2, n // 2))
        s.run(600)
        return s.traces()[..., 1:], dt
    ref, dt = hex2("f64")
    out["config3_hex_p2_symplectic_4"] = _trace_errors(hex2("f32")[0], ref, dt)
    # (the pulse peaks near step 230: 300 steps cover its arrival)
    ref, dt = tet(64, 2, "symplectic_4", "f64", 300, 0.2)
    tr, _ = tet(64, 2, "symplectic_4", "f32", 300, 0.2, dt)
    out["config3_tet_p2_symplectic_4"] = _trace_errors(tr, ref, dt)

    # config 4: one shot per model size of the batched workload (each batched
    # shot equals its single run bit for bit)
    for n in (96, 256):
        er = wl.ellipsoid_model(n, n)
        tr = {}
        for prec in ("f32", "f64"):
            s = _shot(n, 1000 * n, (er, None), precision=prec)
            s.run(300)
            tr[prec] = s.traces()
        out["config4_fdtd_%d_ellipsoids_shot" % n] = _trace_errors(tr["f32"], tr["f64"], wl.cfl_dt(1e-3))

    # config 5: misfit gradients with FP32 fields (FP64 accumulation) against
    # FP64 fields; FDTD 192^3 blob model, one shot, 600 steps (the wave reaches
    # the receiver plane after about 270); tetrahedral lossless symplectic_4
    # on 40^3 cubes, 100 steps
    n = 192
    model = wl.blob_model(n, 7)
    g = {}
    for prec in ("f32", "f64"):
        s = _shot(n, 0, model, receivers=(8, 8), z_rec=n - 20, src=(n // 2 - 45, n // 2 - 45, 20), f0=10e9,
                  precision=prec)
        J, ge, gs = s.gradient(600, np.zeros((64, 601, 3)))
        g[prec] = (np.asarray(ge), np.asarray(gs))
    out["config5_fdtd_192_gradient"] = dict(
        rel_l2_eps=float(np.linalg.norm(g["f32"][0] - g["f64"][0]) / np.linalg.norm(g["f64"][0])),
        rel_l2_sigma=float(np.linalg.norm(g["f32"][1] - g["f64"][1]) / np.linalg.norm(g["f64"][1])))
    gt = {}
    dt_t = None
    for prec in ("f64", "f32"):
        h = 1.0 / 40
        er, _ = wl.tet_model(40, 5, 0.0)
        s = TetSimulation((40, 40, 40), h, 1e-3 * h / wl.C0, wl.EPS0, wl.MU0, 1, "symplectic_4", 1, prec)
        s.set_material(er, np.zeros_like(er))
        if dt_t is None:  # the FP64 handle's stable step, reused for FP32
            dt_t = 0.9 * KAPPA[STEPPERS["symplectic_4"]] / s.max_frequency(40)
        s.set_time_step(dt_t)
        f0 = 2.0 * wl.C0 / (h * 40)
        s.add_dipole([0.5 + 0.1 * h, 0.5 + 0.2 * h, 0.5 + 0.3 * h], [1.0, 0.0, 0.0], f0, 1.2 / f0, 1e-9)
        for p in wl.tet_receivers(40, h, 16, 6):
            s.add_receiver(p)
        J, ge, gs, drift = s.gradient(100, np.zeros((16, 101, 3)))
        gt[prec] = (np.asarray(ge), np.asarray(gs))
        s.close()
    out["config5_tet_p1_40_gradient"] = dict(
        rel_l2_eps=float(np.linalg.norm(gt["f32"][0] - gt["f64"][0]) / np.linalg.norm(gt["f64"][0])),
        rel_l2_sigma=float(np.linalg.norm(gt["f32"][1] - gt["f64"][1]) / np.linalg.norm(gt["f64"][1])))
    return out


def pmc_traffic(path):
    """HBM bytes per step from a committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a device copy of
    known size measured in the same runs)."""
    if not os.path.exists(path):
        return None
    kv = {}
    kern = []
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "kernel":
            kern.append((float(t[3]), float(t[5])))
        elif t[0] in ("copy_bytes", "copy_fetch_kib", "copy_write_kib", "steps_per_group"):
            kv[t[0]] = float(t[1])
    fs = kv["copy_bytes"] / (kv["copy_fetch_kib"] * 1024.0)
    ws = kv["copy_bytes"] / (kv["copy_write_kib"] * 1024.0)
    # the kernels listed make up one launch group of steps_per_group steps
    return sum(f * 1024.0 * fs + w * 1024.0 * ws for f, w in kern) / kv.get("steps_per_group", 1.0)


def cpu_reference_sample(n=256, steps=None, budget_s=12.0):
    """Reference FDTD (src/fdtd, compiled serially from its sources into
    oracle/_ref) on an n^3 vacuum/PEC grid: the closest workload it supports
    (no CPML, no per-cell material, FP64)."""
    from oracle import ref
    if not ref.available():
        return None
    h = 1e-3
    eps0, mu0 = 8.854187817e-12, 1.25663706212e-6
    dt = 0.99 * h / ((1 / (eps0 * mu0) ** 0.5) * 3 ** 0.5)
    # calibrate on a few steps, then size the sample to the budget
    probe = 3
    _, _, s = ref.run_reference((n, n, n), h, dt, eps0, mu0, 1, probe, [(2, (n // 2, n // 2, n // 2))],
                                [[[(n // 2 + 5, n // 2, n // 2)], [], []]], want_output=False)
    if steps is None:
        steps = max(probe, int(budget_s / max(s / probe, 1e-6)))
    _, _, s = ref.run_reference((n, n, n), h, dt, eps0, mu0, 1, steps, [(2, (n // 2, n // 2, n // 2))],
                                [[[(n // 2 + 5, n // 2, n // 2)], [], []]], want_output=False)
    return dict(value=n ** 3 * steps / s / 1e6, unit="Mcell-updates/s", cores=1, kind="reference",
                sample="%d^3 vacuum/PEC FP64, %d steps, %.1f s (src/fdtd built with g++ -O3, serial)"
                       % (n, steps, s))


def host_info():
    """Logical CPUs usable by this process (affinity), physical cores (lscpu)
    and the CPU model."""
    info = {"nproc": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()}
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        cpus = [l.split(",") for l in out.splitlines() if l and not l.startswith("#")]
        allowed = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else None
        cores = {(c[2], c[1]) for c in cpus if allowed is None or int(c[0]) in allowed}
        info["physical_cores"] = len(cores)
        model = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for l in model.splitlines():
            if l.startswith("Model name:"):
                info["model"] = l.split(":", 1)[1].strip()
    except Exception as e:  # reported, not fatal
        info["lscpu_error"] = str(e)
    return info


def cpu_port_sample(n=256, budget_s=10.0):
    """FP64 CPU oracle on the same model (CPML + ellipsoids), timed for at
    least budget_s each on all physical cores of this process's CPU share (capped by
    OMP_NUM_THREADS when set) and on 1 thread."""
    from oracle import fdtd as of
    from oracle.fdtd import OracleFDTD, WAVELET_RICKER
    from emsolvers_amd import workloads as wl
    info = host_info()
    allc = info.get("physical_cores") or info["nproc"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        allc = min(allc, int(os.environ["OMP_NUM_THREADS"]))
    h = 1e-3
    o = OracleFDTD((n, n, n), h, wl.cfl_dt(h), wl.EPS0, wl.MU0, wavelet=WAVELET_RICKER, f0=10e9, npml=10)
    o.set_material(wl.ellipsoid_model(n, 11), None)
    o.setSource_z((n // 2, n // 2, n // 2))
    legs = []
    for threads in (allc, 1):
        of.lib().ems_oracle_set_threads(int(threads))
        o.run(1)  # warm-up (first touch, thread start)
        t0 = time.perf_counter()
        o.run(2)  # calibration
        per = (time.perf_counter() - t0) / 2
        # chunks of about a quarter of the budget until at least budget_s has passed
        chunk = max(1, int(0.25 * budget_s / max(per, 1e-6)))
        steps = 0
        t0 = time.perf_counter()
        while True:
            o.run(chunk)
            steps += chunk
            s = time.perf_counter() - t0
            if s >= budget_s:
                break
        legs.append(dict(value=round(n ** 3 * steps / s / 1e6, 2), cores=int(threads), steps=steps,
                         seconds=round(s, 2)))
    best = legs[0]
    return dict(value=best["value"], unit="Mcell-updates/s", cores=best["cores"], kind="port",
                sample="%d^3 CPML+ellipsoids FP64 oracle, %d steps, %.1f s on %d threads; 1 thread: %.1f Mcell/s"
                       " (%d steps, %.1f s)" % (n, best["steps"], best["seconds"], best["cores"], legs[1]["value"],
                                                legs[1]["steps"], legs[1]["seconds"]),
                one_thread=legs[1], host=info)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", default="f32")
    ap.add_argument("--kernel", type=int, default=0, help="0 auto, 1 split, 2 fused, 3 z-march, 4 two-step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-fem", action="store_true", help="skip the FEM secondary figures")
    ap.add_argument("--tet-steps", type=int, default=200, help="steps of the tetrahedral FEM figures")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched forward / gradient figures")
    ap.add_argument("--gradient64", action="store_true",
                    help="also time the 64-shot batched gradient against 64 sequential gradients")
    ap.add_argument("--fem-profile", action="store_true",
                    help="run only short FEM legs (tet configs 2 and 3, fem_v2 and fem at 128^3), for rocprofv3")
    ap.add_argument("--fem-elem-profile", action="store_true",
                    help="run only the hex FEM order 2 and 3 legs with both element products, for rocprofv3")
    ap.add_argument("--gradient-profile", action="store_true",
                    help="run only a short FDTD gradient leg (192^3, 4 shots, 60 steps), for rocprofv3")
    ap.add_argument("--fp32-errors", action="store_true",
                    help="FP32 against FP64 traces and gradients per benchmark configuration, then exit")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="first run a 1 GiB device copy (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU runs would shard the grid into slabs with halo exchange, not built)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: multi-rank launch is not supported", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibrate:
        import torch
        x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    if args.fp32_errors:
        print(json.dumps({"metric": "FP32 error against FP64", "configs": fp32_errors()}), flush=True)
        return
    if args.fem_profile:
        legs = [gpu_tet(40, "symplectic_2", 50, "f64"),
                gpu_tet(64, "symplectic_4", 2, "f32", order=2, sigma_fraction=0.2, power_iters=10),
                gpu_fem(1, "fem_v2", 128, 20), gpu_fem(2, "symplectic_4", 128, 10), gpu_fem1("fem", 128, 20),
                gpu_fem1("fem_abm", 128, 10)]
        print(json.dumps({"metric": "FEM DOF-updates/s", "legs": legs}), flush=True)
        return
    if args.gradient_profile:
        print(json.dumps({"metric": "FDTD gradient", "legs": [gpu_batch_gradient(shots=4, steps=60)]}), flush=True)
        return
    if args.fem_elem_profile:
        legs = [gpu_fem(o, st, n, k, p, e) for (o, st, n, k) in ((3, "fem_v2", 96, 20), (2, "symplectic_4", 128, 10))
                for p in ("f64", "f32") for e in ("rows", "mfma")]
        print(json.dumps({"metric": "FEM DOF-updates/s", "legs": legs}), flush=True)
        return
    t_start = time.perf_counter()

    def note(leg):
        # one progress line per finished leg, on stderr (stdout carries the JSON line)
        print("bench.py: %s done at %.1f s" % (leg, time.perf_counter() - t_start), file=sys.stderr, flush=True)

    r = gpu_fdtd(args.n, args.steps, args.warmup, args.precision, args.kernel)
    note("fdtd")
    cells = r["cells"]
    ms_per_step = r["total_ms"] / args.steps
    value = cells * args.steps / (r["total_ms"] / 1e3) / 1e6
    k_avg_ms = r["kernel_ms"] / args.steps
    achieved = r["step_bytes"] / (k_avg_ms / 1e3) / 1e9
    out = {
        "metric": "FDTD Mcell-updates/s",
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "f32" else "fp64",
        "data": "synthetic (seeded ellipsoid eps_r model, Ricker source)",
        "config": {
            "workload": "fdtd_yee_%d^3_%s_cpml10_ellipsoids" % (args.n, args.precision),
            "grid": [args.n] * 3,
            "cpml_cells": 10,
            "material_words": 1,
            "kernel": {0: "auto (two-step sweep: two time steps per launch group, CPML-free interior plus "
                          "CPML shell slabs on side streams)", 1: "split", 2: "fused", 3: "z-march",
                       4: "two-step"}[args.kernel],
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic(os.path.join(ROOT, "profiles", "fdtd_yee_256_f32_cpml10_tb2_kernel_stats_pmc.txt"))
            if (args.n == 256 and args.precision == "f32" and args.kernel in (0, 4)) else None,
            "kernel_avg_ms": round(k_avg_ms, 5),
            "algorithmic_bytes_per_step": int(r["step_bytes"]),
            "bytes_per_cell": round(r["per_cell"], 3),
        },
        "device_bytes": r["device_bytes"],
    }
    if not args.no_fem:
        out["fem"] = [gpu_fem(2, "symplectic_4", 128, 20), gpu_fem(2, "symplectic_4", 128, 20, "f32"),
                      gpu_fem(3, "fem_v2", 96, 20), gpu_fem(3, "fem_v2", 96, 20, "f64", "rows"),
                      gpu_fem(3, "fem_v2", 96, 20, "f32"), gpu_fem(3, "symplectic_4", 96, 10),
                      gpu_fem(1, "fem_v2", 128, 50),
                      gpu_fem(1, "symplectic_2", 128, 50, "f32"), gpu_fem1("fem", 128, 50),
                      gpu_fem1("fem_abm", 128, 30)]
        note("hex fem")
    if not args.no_fem:
        out["tet"] = [gpu_tet(40, "symplectic_2", args.tet_steps), gpu_tet(40, "abm", args.tet_steps),
                      gpu_tet(40, "symplectic_2", args.tet_steps, "f32"),
                      gpu_tet(64, "symplectic_4", 5, "f32", order=2, sigma_fraction=0.2, power_iters=40),
                      gpu_tet(64, "symplectic_4", 5, "f32", order=2, sigma_fraction=0.2, power_iters=40,
                              op="ebe_mfma"),
                      gpu_tet(64, "symplectic_4", 3, "f64", order=2, sigma_fraction=0.2, power_iters=40)]
        note("tet")
    if not args.no_fem:
        out["tet_gradient"] = [gpu_tet_gradient(40, 100, "f64"), gpu_tet_gradient(40, 100, "f32")]
        note("tet gradient")
    if not args.no_batch:
        out["batched_forward"] = gpu_batch_forward(20)
        note("batched forward")
        out["batched_gradient"] = gpu_batch_gradient()
        note("batched gradient")
        out["batched_gradient_small"] = gpu_batch_gradient(n=64, shots=64, steps=200)
        note("small-grid batched gradient")
        if args.gradient64:
            out["batched_gradient_64"] = gpu_batch_gradient(shots=64, steps=100)
            note("64-shot gradient")
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_reference_sample(args.n)
        except Exception as e:  # reported, not fatal
            out["cpu_baseline"] = {"error": str(e)}
        note("cpu reference")
        try:
            out["cpu_port"] = cpu_port_sample(args.n)
        except Exception as e:
            out["cpu_port"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
