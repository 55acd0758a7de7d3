"""GPU misfit gradient of the tetrahedral solver (discrete adjoint with the
forward states recovered by running the symplectic integrator backwards for
lossless models, by checkpoints and reruns for lossy ones) against the
oracle's discrete adjoint with stored states (oracle/tet.py misfit_gradient,
itself pinned by central finite differences).

Bar: FP64 misfit within 1e-12 and gradients within 1e-8 relative L2 of the
oracle, reconstruction drift below 1e-9; FP32 gradients within 1e-3."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = {
    "p1-s2": dict(order=1, stepper="symplectic_2", n=3, dt=0.02, steps=40, tf=2),
    "p1-s4": dict(order=1, stepper="symplectic_4", n=3, dt=0.02, steps=40, tf=1),
    "p2-s4": dict(order=2, stepper="symplectic_4", n=2, dt=0.01, steps=30, tf=3),
    # lossy: checkpoints (ceil(sqrt(steps)) per segment) and reruns
    "p1-s2-lossy": dict(order=1, stepper="symplectic_2", n=3, dt=0.02, steps=40, tf=2, lossy=True),
    "p1-s4-lossy": dict(order=1, stepper="symplectic_4", n=3, dt=0.02, steps=37, tf=1, lossy=True),
    "p2-s4-lossy": dict(order=2, stepper="symplectic_4", n=2, dt=0.01, steps=30, tf=3, lossy=True),
}
ST = {"symplectic_2": 0, "symplectic_4": 1, "abm": 2}


def setup(case, kind, precision="f64", seed=3, op="auto"):
    n = case["n"]
    rng = np.random.default_rng(seed)
    er = rng.uniform(1.0, 3.0, 6 * n ** 3)
    sg = rng.uniform(0.0, 0.8, 6 * n ** 3) if case.get("lossy") else None
    pos_d = ([0.41, 0.57, 0.36], [0.3, 0.5, 0.8], 3.0, 0.3, 1.0)
    recs = [[0.62, 0.31, 0.55], [0.27, 0.71, 0.66], [0.55, 0.49, 0.23]]
    if kind == "oracle":
        from oracle import tet as ot
        s = ot.OracleTet((n, n, n), 1.0 / n, case["dt"], 1.0, 1.0, tf=case["tf"], stepper=ST[case["stepper"]],
                         eps_r=er, sigma=sg, order=case["order"])
    else:
        from emsolvers_amd.tet import TetSimulation
        s = TetSimulation((n, n, n), 1.0 / n, case["dt"], 1.0, 1.0, case["tf"], case["stepper"], case["order"],
                          precision)
        s.set_operator(op)
        s.set_material(er, sg)
    s.add_dipole(*pos_d)
    for r in recs:
        s.add_receiver(r)
    nslot = case["steps"] // case["tf"] + 1
    d = np.zeros((len(recs), nslot, 3))
    d[:, 1:, :] = 0.05 * np.random.default_rng(seed + 1).standard_normal((len(recs), nslot - 1, 3))
    return s, d, er


_ref = {}


def reference(name):
    if name not in _ref:
        from oracle import tet as ot
        case = CASES[name]
        o, d, er = setup(case, "oracle")
        _ref[name] = ot.misfit_gradient(o, case["steps"], d)[:3]
    return _ref[name]


@pytest.mark.parametrize("op", ["csr", "ebe"])
@pytest.mark.parametrize("name", list(CASES))
def test_fp64_gradient_matches_oracle(name, op):
    case = CASES[name]
    print("tet gradient fp64", name, op, flush=True)
    J_o, ge_o, gs_o = reference(name)
    s, d, _ = setup(case, "gpu", op=op)
    J, ge, gs, drift = s.gradient(case["steps"], d)
    print("misfit %.6e drift %.2e" % (J, drift))
    assert J == pytest.approx(J_o, rel=1e-12)
    assert rel_l2(ge, ge_o) < 1e-8
    assert rel_l2(gs, gs_o) < 1e-8
    assert drift < 1e-9
    assert np.abs(gs).max() > 0
    # the handle is back at rest: a second evaluation repeats the first
    J2, ge2, _, _ = s.gradient(case["steps"], d)
    assert J2 == J and np.array_equal(ge2, ge)


@pytest.mark.parametrize("name", ["p1-s4", "p2-s4", "p1-s4-lossy", "p2-s4-lossy"])
def test_fp32_gradient(name):
    case = CASES[name]
    J_o, ge_o, gs_o = reference(name)
    s, d, _ = setup(case, "gpu", "f32")
    J, ge, gs, drift = s.gradient(case["steps"], d)
    print("fp32 %s: misfit rel %.2e, grad rel %.2e, drift %.2e" % (name, abs(J - J_o) / J_o, rel_l2(ge, ge_o), drift))
    assert abs(J - J_o) <= 1e-4 * J_o
    assert rel_l2(ge, ge_o) < 1e-3
    assert rel_l2(gs, gs_o) < 1e-3


def test_lossy_gradient_against_finite_differences():
    """The checkpointed lossy gradient (FP64) along random directions of
    eps_r and sigma against central differences of the GPU misfit."""
    case = CASES["p1-s2-lossy"]
    s, d, er = setup(case, "gpu")
    rng = np.random.default_rng(4)
    sg = rng.uniform(0.0, 0.8, er.size)
    s.set_material(er, sg)
    J, ge, gs, _ = s.gradient(case["steps"], d)
    for k in range(2):
        de, ds = rng.standard_normal(er.size), rng.standard_normal(er.size)
        h = 1e-5

        def Jat(t):
            s.set_material(er + t * de, sg + t * ds)
            return s.gradient(case["steps"], d)[0]
        fd = (Jat(h) - Jat(-h)) / (2 * h)
        an = float(ge @ de + gs @ ds)
        print("lossy FD %d: %.8e analytic %.8e" % (k, fd, an), flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an)
    s.set_material(er, sg)


def test_unsupported_models_rejected():
    from emsolvers_amd import _lib
    from emsolvers_amd.tet import TetSimulation
    a = TetSimulation((2, 2, 2), 0.5, 0.01, 1.0, 1.0, 1, "abm")
    a.add_receiver([0.3, 0.4, 0.6])
    with pytest.raises(_lib.EmsError) as ei:
        a.gradient(4, np.zeros((1, 5, 3)))
    assert ei.value.status == 5


def test_autograd_binding():
    import torch
    from emsolvers_amd.autograd import tet_misfit
    case = CASES["p1-s2"]
    s, d, er = setup(case, "gpu")
    J_o, ge_o, _ = reference("p1-s2")
    e = torch.tensor(er, dtype=torch.float64, requires_grad=True)
    J = tet_misfit(e, s, case["steps"], torch.tensor(d))
    J.backward()
    assert float(J) == pytest.approx(J_o, rel=1e-12)
    assert rel_l2(e.grad.numpy(), ge_o) < 1e-8
