"""The benchmark workloads at their full sizes on the GPU, with the step
counts stated in each test (reduced where the full count would take minutes).

FDTD and hex FEM paths equal their FP64 oracles bit for bit at small sizes
(test_fdtd_gpu, test_fem_gpu); the tetrahedral FP64 path equals its oracle
within 1e-10 (test_tet_gpu).  Here, at full size: batched members against
their single runs (bit for bit), FP32 against FP64 on the same inputs, the
operator and solver variants against each other, and the gradients' FP32
error.  Bars, each stated in its test: FP32 traces against FP64 by relative
L2 and by the TF envelope / phase misfits (Kristekova et al.) of the
components carrying at least 1 % of the peak.
"""
import numpy as np
import pytest

from tests.conftest import rel_l2, tf_misfits

pytestmark = pytest.mark.gpu


def fdtd_shot(n, seed, model, precision, npml=10, src=None, f0=None, receivers=(8, 4), z_rec=None):
    """An FDTD shot of the batched workloads: n^3 cells of 1 mm, dt = 0.99 x
    CFL, CPML npml cells, Ricker (f0 seeded in [5, 15] GHz unless given),
    z source (seeded in the inner 50 % unless given), a planar grid of
    receivers[0] x receivers[1] points at z = z_rec (default n/2)."""
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


def test_config4_batched_four_model_sizes():
    """64 shots over four models (96^3, 128^3, 192^3, 256^3 cells, 16 shots
    each; CPML 10, seeded ellipsoids eps_r in [2, 8], Ricker f0 in [5, 15]
    GHz, source in the inner 50 %, 32 receivers on an 8 x 4 grid), 300 steps
    (of 3000).  FP32: one batch with autotuned tiles and HIP-graph replay,
    every member bit for bit its single run.  FP64: the first 4 shots of each
    size batched, each bit for bit its single run.  FP32 against FP64 on
    those 16 shots: relative L2 below 1e-5 on every shot, TF misfits within
    1e-4 on the first shot of each size."""
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.fdtd import Batch
    sizes, per, steps = (96, 128, 192, 256), 16, 300
    models = {n: (wl.ellipsoid_model(n, n), None) for n in sizes}
    keys = [(n, 1000 * n + i) for n in sizes for i in range(per)]
    mem = [fdtd_shot(n, seed, models[n], "f32") for n, seed in keys]
    b = Batch(mem)
    rows, planes = b.autotune(3)  # on scratch fields: the members stay at rest
    b.set_graph(10)
    b.run(steps)
    print("config 4 fp32 batch: tiling rows %d planes %d" % (rows, planes), flush=True)
    tr32 = {}
    for (n, seed), m in zip(keys, mem):
        s = fdtd_shot(n, seed, models[n], "f32")
        s.run(steps)
        assert np.array_equal(m.traces(), s.traces()), (n, seed)
        assert np.array_equal(m.field(2), s.field(2)), (n, seed)
        if seed % 1000 < 4:
            tr32[(n, seed)] = m.traces()
        s.close()
    del b, mem
    sub = [k for k in keys if k[1] % 1000 < 4]
    mem64 = [fdtd_shot(n, seed, models[n], "f64") for n, seed in sub]
    b64 = Batch(mem64)
    b64.run(steps)
    dt = wl.cfl_dt(1e-3, 0.99)
    for (n, seed), m in zip(sub, mem64):
        s = fdtd_shot(n, seed, models[n], "f64")
        s.run(steps)
        ref = m.traces()
        assert np.array_equal(ref, s.traces()), (n, seed)
        s.close()
        err = rel_l2(tr32[(n, seed)], ref)
        msg = "config 4 n=%d seed=%d: fp32 rel L2 %.2e" % (n, seed, err)
        if seed % 1000 == 0:
            em, pm = tf_misfits(tr32[(n, seed)], ref, dt)
            msg += ", TF EM %.2e PM %.2e" % (em, pm)
            assert em <= 1e-4 and pm <= 1e-4
        print(msg, flush=True)
        assert err < 1e-5


def test_config5_fdtd_gradient_16_shots():
    """192^3 cells, 16 shots on a 4 x 4 source grid (spacing 30) at z = 20,
    64 receivers on the plane z = 172; observed data from the true model
    (background eps_r 2 plus 3 seeded Gaussian blobs: d eps_r 0.5, width 12
    cells, d sigma 0.01 S/m), gradient at the background model; 500 steps (of
    2000).  FP32 fields with FP64 sums against FP64: per-shot misfits within
    1e-4 relative, gradients (eps_r and sigma) within 1e-3 relative L2."""
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.fdtd import Batch
    n, shots, steps = 192, 16, 500
    true = wl.blob_model(n, 7)
    nm = (n + 1) ** 3
    background = (np.full(nm, 2.0), np.zeros(nm))
    srcs = [(n // 2 - 45 + 30 * (i % 4), n // 2 - 45 + 30 * (i // 4), 20) for i in range(shots)]

    def shots_on(model, prec):
        return [fdtd_shot(n, i, model, prec, src=srcs[i], f0=10e9, receivers=(8, 8), z_rec=n - 20)
                for i in range(shots)]
    obs_mem = shots_on(true, "f64")
    Batch(obs_mem).run(steps)
    d_obs = [m.traces() for m in obs_mem]
    del obs_mem
    out = {}
    for prec in ("f64", "f32"):
        mem = shots_on(background, prec)
        J, mis, ge, gs = Batch(mem).gradient(steps, d_obs)
        out[prec] = (np.asarray(mis), ge, gs)
        del mem
    mis64, ge64, gs64 = out["f64"]
    mis32, ge32, gs32 = out["f32"]
    assert np.all(np.isfinite(ge32)) and np.all(np.isfinite(gs32))
    e_mis = float(np.max(np.abs(mis32 - mis64) / np.abs(mis64)))
    e_eps, e_sig = rel_l2(ge32, ge64), rel_l2(gs32, gs64)
    print("config 5 fdtd 192^3 x 16 shots: misfit %.4e, fp32 misfit rel %.2e, grad eps rel L2 %.2e, sigma %.2e"
          % (mis64.sum(), e_mis, e_eps, e_sig), flush=True)
    assert e_mis < 1e-4
    assert e_eps < 1e-3 and e_sig < 1e-3


def _tet(n, order, stepper, precision, sigma_fraction, seed=5, op="auto", solver="auto", dt=None,
         delay=1.2, receivers=None):
    """The tetrahedral workloads: Kuhn split of an n^3 unit cube, smoothed
    log-uniform eps_r, sigma in [0, 0.05] S/m on a seeded fraction of the
    tetrahedra, x dipole (Gaussian pulse, f0 = two vacuum wavelengths across
    the cube, peak at delay / f0) at the centre, dt = 0.9 x the stable step
    of the FP64 handle unless given."""
    from emsolvers_amd import workloads as wl
    from emsolvers_amd.tet import TetSimulation, KAPPA, STEPPERS
    h = 1.0 / n
    er, sg = wl.tet_model(n, seed, sigma_fraction)
    s = TetSimulation((n, n, n), h, 1e-3 * h / wl.C0, wl.EPS0, wl.MU0, 1, stepper, order, precision)
    s.set_operator(op)
    s.set_solver(solver)
    s.set_material(er, sg if sg is not None else np.zeros_like(er))
    if dt is None:
        dt = 0.9 * KAPPA[STEPPERS[stepper]] / s.max_frequency(40)
    s.set_time_step(dt)
    f0 = 2.0 * wl.C0 / (h * n)
    s.add_dipole([0.5 + 0.1 * h, 0.5 + 0.2 * h, 0.5 + 0.3 * h], [1.0, 0.0, 0.0], f0, delay / f0, 1e-9)
    pts = wl.tet_receivers(n, h, 16, seed + 1) if receivers is None else receivers
    for p in pts:
        s.add_receiver(p)
    return s, dt


def test_config2_tet_full_2000_steps():
    """First-order tetrahedra on 40^3 cubes (0.46 M DOFs), PEC, lossless,
    2000 steps of symplectic_2 and of ABM.  FP64 with the default mass solver
    (Chebyshev-accelerated Jacobi at order 1) against FP64 conjugate
    gradients: within 1e-10 relative L2 (both solve to 1e-14).  FP32
    against FP64: relative L2 below 1e-4, TF misfits within 1e-3 (2000 steps
    of a closed, lossless cavity: the rounding error of the echoes grows
    linearly with time)."""
    for stepper in ("symplectic_2", "abm"):
        s, dt = _tet(40, 1, stepper, "f64", 0.0)
        s.run(2000)
        ref = s.traces()
        s.close()
        c, _ = _tet(40, 1, stepper, "f64", 0.0, solver="cg", dt=dt)
        c.run(2000)
        e_cg = rel_l2(c.traces(), ref)
        c.close()
        f, _ = _tet(40, 1, stepper, "f32", 0.0, dt=dt)
        f.run(2000)
        tr = f.traces()
        f.close()
        assert np.all(np.isfinite(ref)) and np.all(np.isfinite(tr))
        err = rel_l2(tr, ref)
        em, pm = tf_misfits(tr, ref, dt)
        print("config 2 %s, 2000 steps: Chebyshev vs CG %.2e, fp32 rel L2 %.2e, TF EM %.2e PM %.2e"
              % (stepper, e_cg, err, em, pm), flush=True)
        assert e_cg < 1e-10
        assert err < 1e-4 and em <= 1e-3 and pm <= 1e-3


def test_config3_tet_p2_64_sigma20():
    """Second-order tetrahedra on 64^3 cubes (10.1 M DOFs), sigma on 20 % of
    the tetrahedra, symplectic_4; the pulse peaks at step ~100 (delay 0.45 /
    f0) and 8 receivers sit 2-9 cells from the dipole; 150 steps (of 5000).
    FP32 against FP64: relative L2 below 1e-5, TF misfits within 1e-4; the
    per-element dense-product (MFMA) mass operator against the gather one:
    within 1e-5 relative L2 in FP32 and 1e-10 in FP64."""
    n = 64
    h = 1.0 / n
    c = np.array([0.5 + 0.1 * h, 0.5 + 0.2 * h, 0.5 + 0.3 * h])
    rng = np.random.default_rng(9)
    rec = [c + h * r * v / np.linalg.norm(v) for r, v in zip(np.linspace(2, 9, 8), rng.normal(size=(8, 3)))]
    s, dt = _tet(n, 2, "symplectic_4", "f64", 0.2, delay=0.45, receivers=rec)
    s.run(150)
    ref = s.traces()
    s.close()
    m64, _ = _tet(n, 2, "symplectic_4", "f64", 0.2, delay=0.45, receivers=rec, dt=dt, op="ebe_mfma")
    m64.run(150)
    e_m64 = rel_l2(m64.traces(), ref)
    m64.close()
    out = {}
    for op in ("ebe", "ebe_mfma"):
        f, _ = _tet(n, 2, "symplectic_4", "f32", 0.2, delay=0.45, receivers=rec, dt=dt, op=op)
        f.run(150)
        out[op] = f.traces()
        f.close()
    err = rel_l2(out["ebe"], ref)
    em, pm = tf_misfits(out["ebe"], ref, dt)
    e_m32 = rel_l2(out["ebe_mfma"], out["ebe"])
    print("config 3 tet p2 64^3: peak |E| %.3e, fp32 rel L2 %.2e, TF EM %.2e PM %.2e, mfma vs gather fp32 %.2e"
          " fp64 %.2e" % (np.abs(ref).max(), err, em, pm, e_m32, e_m64), flush=True)
    assert np.abs(ref).max() > 0
    assert err < 1e-5 and em <= 1e-4 and pm <= 1e-4
    assert e_m32 < 1e-5 and e_m64 < 1e-10


def test_config5_tet_adjoint_fp32_against_fp64():
    """Misfit gradient of the lossless first-order symplectic_4 run on 40^3
    cubes (0.46 M DOFs), forward states recovered by running the integrator
    backwards, 200 steps; observed data 0.9 x the traces of the model.  FP32
    fields (FP64 gradient sums) against FP64: gradient within 1e-3 relative
    L2; reconstruction drift (largest recovered-state error relative to the
    state) below 1e-10 in FP64 and 1e-2 in FP32 (measured 7e-12 and 2e-3:
    the FP32 rounding of 200 forward and 200 backward steps, which moves the
    gradient by 3e-4)."""
    out = {}
    dt = None
    for prec in ("f64", "f32"):
        s, dt = _tet(40, 1, "symplectic_4", prec, 0.0, dt=dt)
        s.run(200)
        d_obs = 0.9 * s.traces() if prec == "f64" else d_obs
        s.close()
        s, _ = _tet(40, 1, "symplectic_4", prec, 0.0, dt=dt)
        J, ge, gs, drift = s.gradient(200, d_obs)
        out[prec] = (J, ge, gs, drift)
        s.close()
    J64, ge64, gs64, d64 = out["f64"]
    J32, ge32, gs32, d32 = out["f32"]
    e_eps = rel_l2(ge32, ge64)
    e_sig = rel_l2(gs32, gs64)
    print("config 5 tet adjoint 40^3: J %.4e (fp32 %.4e), grad eps rel L2 %.2e sigma %.2e, drift fp64 %.2e fp32 %.2e"
          % (J64, J32, e_eps, e_sig, d64, d32), flush=True)
    assert np.all(np.isfinite(ge32))
    assert e_eps < 1e-3 and e_sig < 1e-3
    assert d64 < 1e-10 and d32 < 1e-2
