"""GPU misfit gradient (discrete adjoint of the Yee step with sqrt(N)
checkpointing) against the FP64 adjoint oracle and finite differences.

Bars: FP64 J, traces and gradients bit-exact with the oracle (same operation
order); results independent of the checkpoint interval (recomputation is
exact); the FP32 gradient within 1e-3 of central finite differences of the
FP64 misfit along random directions.
"""
import numpy as np
import pytest

from oracle.fdtd import OracleFDTD, EPS0, MU0, C0, WAVELET_RICKER

pytestmark = pytest.mark.gpu

CASES = {
    "cpml-lossy": dict(grid=(9, 8, 10), npml=3, kmax=2.0, amax=0.05, lossy=True, eps=True, tf=1, steps=60),
    "eps-only-tf2": dict(grid=(10, 7, 9), npml=0, kmax=1.0, amax=0.0, lossy=False, eps=True, tf=2, steps=50),
    "uniform-cpml": dict(grid=(8, 9, 8), npml=2, kmax=1.0, amax=0.0, lossy=False, eps=False, tf=1, steps=40),
}


def _dt(h):
    return 0.9 * h / (C0 * np.sqrt(3))


def _material(case, seed=3):
    nx, ny, nz = case["grid"]
    nm = (nx + 1) * (ny + 1) * (nz + 1)
    rng = np.random.default_rng(seed)
    er = 1 + rng.uniform(0, 3, nm) if case["eps"] else None
    sg = rng.uniform(0, 2.0, nm) if case["lossy"] else None
    return er, sg


def _setup(s, case, er, sg):
    if er is not None or sg is not None:
        s.set_material(er, sg)
    nx, ny, nz = case["grid"]
    s.setSource_z((nx // 2, ny // 2, nz // 2))
    s.setSource_x((nx // 3, ny // 2 + 1, nz // 2 - 1))
    s.setObservationPoint([[(2, 3, 3)], [(nx - 3, ny - 3, nz - 4)], [(nx - 4, 2, nz - 3), (nx - 4, 3, nz - 3)]])
    s.setObservationPoint([[], [], [(3, ny - 2, 3)]])


def _oracle(case, er, sg):
    o = OracleFDTD(case["grid"], 1e-3, _dt(1e-3), EPS0, MU0, case["tf"], wavelet=WAVELET_RICKER, f0=60e9,
                   t0=25e-12, npml=case["npml"], pml_kappa_max=case["kmax"], pml_alpha_max=case["amax"])
    _setup(o, case, er, sg)
    return o


def _gpu(case, er, sg, precision="f64"):
    from emsolvers_amd.fdtd import FemSimulation, WAVELET_RICKER as WR
    s = FemSimulation(case["grid"], 1e-3, _dt(1e-3), EPS0, MU0, case["tf"], precision=precision)
    if case["npml"]:
        s.set_cpml(case["npml"], 3.0, 1e-6, case["kmax"], case["amax"])
    s.set_wavelet(WR, 60e9, 25e-12, 1.0)
    _setup(s, case, er, sg)
    return s


def _observed(case, er, sg):
    er_t = (np.ones_like(sg) if er is None and sg is not None else er)
    if er_t is None:
        nx, ny, nz = case["grid"]
        er_t = np.ones((nx + 1) * (ny + 1) * (nz + 1))
    er_t = er_t.copy()
    er_t[::7] += 0.5
    o = _oracle(case, er_t, sg)
    o.run(case["steps"])
    return o.traces()


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_gradient_bitwise_with_oracle(name):
    case = CASES[name]
    print("case", name, flush=True)
    er, sg = _material(case)
    dobs = _observed(case, er, sg)
    o = _oracle(case, er, sg)
    Jo, geo, gso = o.gradient(case["steps"], dobs)
    s = _gpu(case, er, sg)
    Jg, geg, gsg = s.gradient(case["steps"], dobs)
    assert Jo > 0
    assert Jg == Jo
    np.testing.assert_array_equal(s.traces(), o.traces())
    assert np.abs(geo).max() > 0
    assert np.array_equal(geg, geo)
    assert np.array_equal(gsg, gso)


def test_checkpoint_interval_does_not_change_results():
    case = CASES["cpml-lossy"]
    er, sg = _material(case)
    dobs = _observed(case, er, sg)
    ref = None
    for k in (0, 1, 7, 60, 1000):
        s = _gpu(case, er, sg)
        s.set_checkpoint_interval(k)
        out = s.gradient(case["steps"], dobs)
        if ref is None:
            ref = out
        else:
            assert out[0] == ref[0]
            assert np.array_equal(out[1], ref[1]) and np.array_equal(out[2], ref[2])
    # the handle holds the final forward state afterwards
    p = _gpu(case, er, sg)
    p.run(case["steps"])
    for w in range(6):
        assert np.array_equal(s.field(w), p.field(w))
    assert s.current_time == p.current_time


def test_memory_caps_do_not_change_results():
    """Memory caps that leave one-level schedules of several intervals and
    two-level ones (every segment rerun twice): each plan within its cap, and
    misfit, gradients and final state bit for bit those of the uncapped run.
    A cap that nothing fits is an error."""
    case = CASES["cpml-lossy"]
    er, sg = _material(case)
    dobs = _observed(case, er, sg)
    N = case["steps"]
    s = _gpu(case, er, sg)
    k1, ks, full = s.gradient_plan(N)
    assert (k1, ks) == (N, N)  # a small grid: every step's snapshot fits
    ref = s.gradient(N, dobs)
    kinds = set()
    for frac in (0.6, 0.45, 0.4, 0.35, 0.3):
        c = _gpu(case, er, sg)
        cap = int(full * frac)
        c.set_gradient_memory(cap)
        try:
            k1, ks, b = c.gradient_plan(N)
        except Exception:
            continue  # nothing fits this cap
        print("cap %.2f: level-1 interval %d, segment %d, %d bytes" % (frac, k1, ks, b), flush=True)
        assert b <= cap and ks < N
        kinds.add(k1 < N)
        out = c.gradient(N, dobs)
        assert out[0] == ref[0]
        assert np.array_equal(out[1], ref[1]) and np.array_equal(out[2], ref[2])
        for w in range(6):
            assert np.array_equal(c.field(w), s.field(w))
    assert kinds == {False, True}, kinds
    c = _gpu(case, er, sg)
    c.set_gradient_memory(1)
    with pytest.raises(Exception):
        c.gradient(N, dobs)


@pytest.mark.parametrize("name", ["cpml-lossy", "eps-only-tf2"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_gradient_two_step_kernel_bitwise(name, precision):
    """With the two-step kernel the gradient's forward stretches without
    snapshots advance in pairs of steps (the same bits as single steps):
    misfit, gradients, traces and final state equal those of the split
    sweeps, for the default schedule and for capped one- and two-level ones."""
    case = CASES[name]
    er, sg = _material(case)
    dobs = _observed(case, er, sg)
    N = case["steps"]
    print("two-step gradient", name, precision, flush=True)
    full = _gpu(case, er, sg, precision).gradient_plan(N)[2]
    for frac in (None, 0.6, 0.35):
        outs = []
        for k in (1, 4):
            s = _gpu(case, er, sg, precision)
            s.set_kernel(k)
            if frac is not None:
                s.set_gradient_memory(int(full * frac))
                try:
                    s.gradient_plan(N)
                except Exception:
                    break  # nothing fits this cap
            outs.append((s.gradient(N, dobs), s.traces(), [s.field(w) for w in range(6)]))
        if len(outs) < 2:
            continue
        (a, ta, fa), (b, tb, fb) = outs
        assert a[0] == b[0], frac
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), frac
        np.testing.assert_array_equal(ta, tb)
        for w in range(6):
            assert np.array_equal(fa[w], fb[w]), (frac, w)


def test_fp32_gradient_against_finite_differences():
    case = CASES["cpml-lossy"]
    er, sg = _material(case)
    dobs = _observed(case, er, sg)
    s32 = _gpu(case, er, sg, "f32")
    _, ge, gs = s32.gradient(case["steps"], dobs)
    rng = np.random.default_rng(11)
    d = 1e-4
    for which, g in ((0, ge), (1, gs)):
        v = rng.standard_normal(er.size)

        def J(delta):
            e = er + delta * v if which == 0 else er
            sgm = sg + delta * v if which == 1 else sg
            return _gpu(case, e, sgm).gradient(case["steps"], dobs)[0]
        fd = (J(d) - J(-d)) / (2 * d)
        err = abs(g @ v - fd) / abs(fd)
        print("fp32 gradient vs FD (%s): %.2e" % ("eps_r" if which == 0 else "sigma", err), flush=True)
        assert err < 1e-3


def test_torch_autograd_binding():
    """torch.autograd.Function over the C ABI: backward returns the adjoint
    gradient, and gradcheck (FP64, random projections) agrees with finite
    differences of the forward."""
    import torch
    from emsolvers_amd.autograd import fdtd_misfit
    case = dict(CASES["cpml-lossy"], grid=(5, 5, 6), npml=1, steps=25)
    er, sg = _material(case)
    dobs = _observed(case, er, sg)
    s = _gpu(case, er, sg)
    e = torch.tensor(er, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(sg, dtype=torch.float64, requires_grad=True)
    J = fdtd_misfit(e, g, s, case["steps"], dobs)
    J.backward()
    J2, ge, gs = _gpu(case, er, sg).gradient(case["steps"], dobs)
    assert J.item() == J2
    assert np.array_equal(e.grad.numpy(), ge) and np.array_equal(g.grad.numpy(), gs)
    f = lambda a, b: fdtd_misfit(a, b, s, case["steps"], dobs)
    assert torch.autograd.gradcheck(f, (e.detach().requires_grad_(), g.detach().requires_grad_()), eps=1e-6,
                                    atol=1e-9, rtol=1e-4, fast_mode=True)


@pytest.mark.parametrize("group", [0, 1, 2])
def test_batched_gradient_equals_sum_of_single_shots(group):
    """Several shots on one model in one batch (one group of 3, or groups of
    1 or 2 shots): per-shot misfits and traces bit-exact with single runs and
    with the oracle; the summed gradient equal to the sum of the single-shot
    gradients up to the summation order (1e-12), and to the sum of the
    oracle's single-shot gradients likewise."""
    from emsolvers_amd.fdtd import Batch
    case = CASES["cpml-lossy"]
    er, sg = _material(case)
    nx, ny, nz = case["grid"]
    srcs = [(nx // 2, ny // 2, nz // 2), (2, 3, 4), (nx - 3, ny - 2, 3)]

    def shot(k, kind):
        s = _gpu(case, er, sg) if kind == "gpu" else _oracle(case, er, sg)
        s.setSource_y(srcs[k])
        return s
    obs = []
    for k in range(3):
        o = shot(k, "oracle")
        er_t = er.copy()
        er_t[::5] += 0.3
        o.set_material(er_t, sg)
        o.run(case["steps"])
        obs.append(o.traces())
    singles = [shot(k, "gpu").gradient(case["steps"], obs[k]) for k in range(3)]
    mem = [shot(k, "gpu") for k in range(3)]
    b = Batch(mem)
    b.set_gradient_group(group)
    J, mis, ge, gs = b.gradient(case["steps"], obs)
    for k in range(3):
        assert mis[k] == singles[k][0]
    assert J == singles[0][0] + singles[1][0] + singles[2][0]
    se = singles[0][1] + singles[1][1] + singles[2][1]
    ss = singles[0][2] + singles[1][2] + singles[2][2]
    assert np.abs(ge - se).max() <= 1e-12 * np.abs(se).max()
    assert np.abs(gs - ss).max() <= 1e-12 * np.abs(ss).max()
    # the oracle's single-shot misfits and gradients
    oe, osg = 0.0, 0.0
    for k in range(3):
        Jo, geo, gso = shot(k, "oracle").gradient(case["steps"], obs[k])
        assert mis[k] == Jo
        oe, osg = oe + geo, osg + gso
    assert np.abs(ge - oe).max() <= 1e-12 * np.abs(oe).max()
    assert np.abs(gs - osg).max() <= 1e-12 * np.abs(osg).max()
    ref = shot(1, "gpu")
    ref.run(case["steps"])
    np.testing.assert_array_equal(mem[1].traces(), ref.traces())
    assert np.array_equal(mem[1].field(2), ref.field(2))
