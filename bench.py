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


def gpu_fem(order, stepper, n, steps, precision="f64"):
    """Structured-hex FEM (kmat operators) on an n^3 fine grid; returns
    DOF-updates/s of the step kernels (HIP events)."""
    from emsolvers_amd.fem import FemSimulation
    from emsolvers_amd.workloads import EPS0, MU0, C0
    kmat = os.path.join(ROOT, "tests", "golden", "kmat")
    h = 1e-3
    dt = h / (C0 * 3 ** 0.5) / 2.0
    sim = FemSimulation(order, (n, n, n), h, dt, EPS0, MU0, 1, 1, stepper=stepper, precision=precision,
                        kmat_dir=kmat)
    sim.setSource_z((n // 2, n // 2, n // 2))
    sim.setObservationPoint((n // 2 + order, n // 2, n // 2))
    sim.run(2)
    total_ms, kernel_ms = sim.run_timed(steps)
    ndof = sim.ndof()
    return dict(value=round(ndof * steps / (kernel_ms / 1e3) / 1e6, 1), unit="MDOF-updates/s",
                dofs=ndof, ms_per_step=round(kernel_ms / steps, 4), order=order, stepper=stepper,
                grid=[n] * 3, dtype="fp64" if precision == "f64" else "fp32")


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
        else:
            kv[t[0]] = float(t[1])
    fs = kv["copy_bytes"] / (kv["copy_fetch_kib"] * 1024.0)
    ws = kv["copy_bytes"] / (kv["copy_write_kib"] * 1024.0)
    return sum(f * 1024.0 * fs + w * 1024.0 * ws for f, w in kern)


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


def cpu_port_sample(n=256, steps=4, threads=None):
    """FP64 CPU oracle on the same model (CPML + ellipsoids), OpenMP threads."""
    from oracle.fdtd import OracleFDTD, WAVELET_RICKER
    from emsolvers_amd import workloads as wl
    threads = threads or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    h = 1e-3
    o = OracleFDTD((n, n, n), h, wl.cfl_dt(h), wl.EPS0, wl.MU0, wavelet=WAVELET_RICKER, f0=10e9,
                   npml=10)
    o.set_material(wl.ellipsoid_model(n, 11), None)
    o.setSource_z((n // 2, n // 2, n // 2))
    o.run(1)
    t0 = time.perf_counter()
    o.run(steps)
    s = time.perf_counter() - t0
    return dict(value=n ** 3 * steps / s / 1e6, unit="Mcell-updates/s", cores=threads, kind="port",
                sample="%d^3 CPML+ellipsoids FP64 oracle, %d steps, %.1f s" % (n, steps, s))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", default="f32")
    ap.add_argument("--kernel", type=int, default=0, help="0 auto, 1 split, 2 fused")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-fem", action="store_true", help="skip the FEM secondary figures")
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
    r = gpu_fdtd(args.n, args.steps, args.warmup, args.precision, args.kernel)
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
            "kernel": {0: "auto (split H/E sweeps when a CPML is present, fused otherwise)", 1: "split", 2: "fused"}[args.kernel],
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic(os.path.join(ROOT, "profiles", "fdtd_yee_256_f32_cpml10_pmc.txt"))
            if (args.n == 256 and args.precision == "f32" and args.kernel in (0, 1)) else None,
            "kernel_avg_ms": round(k_avg_ms, 5),
            "algorithmic_bytes_per_step": int(r["step_bytes"]),
            "bytes_per_cell": round(r["per_cell"], 3),
        },
        "device_bytes": r["device_bytes"],
    }
    if not args.no_fem:
        out["fem"] = [gpu_fem(2, "symplectic_4", 128, 20), gpu_fem(1, "fem_v2", 128, 50),
                      gpu_fem(1, "symplectic_2", 128, 50, "f32")]
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_reference_sample(args.n)
        except Exception as e:  # reported, not fatal
            out["cpu_baseline"] = {"error": str(e)}
        try:
            out["cpu_port"] = cpu_port_sample(args.n)
        except Exception as e:
            out["cpu_port"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
