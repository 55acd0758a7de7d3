"""Pins of the FP64 discrete-adjoint oracle of the Yee step (oracle/fdtd_oracle.c,
adjoint section): the transpose identity <F x, y> = <x, F^T y> of one full
step (CPML and lossy material included), and the misfit gradient against
central finite differences of the forward oracle along random directions."""
import numpy as np
import pytest

from oracle.fdtd import OracleFDTD, EPS0, MU0, C0, WAVELET_RICKER

NX, NY, NZ = 9, 8, 10
H = 1e-3
DT = 0.9 * H / (C0 * np.sqrt(3))
NM = (NX + 1) * (NY + 1) * (NZ + 1)


def make(npml=3, kmax=2.0, amax=0.05):
    return OracleFDTD((NX, NY, NZ), H, DT, EPS0, MU0, 1, wavelet=WAVELET_RICKER, f0=60e9, t0=25e-12, npml=npml,
                      pml_kappa_max=kmax, pml_alpha_max=amax)


def e_interior(comp):
    """mask of the E entries the update touches (the ghost ring stays 0)"""
    if comp == 0:
        m = np.zeros((NZ + 2, NY + 2, NX + 1), bool)
        m[1:NZ + 1, 1:NY + 1, :] = True
    elif comp == 1:
        m = np.zeros((NZ + 2, NY + 1, NX + 2), bool)
        m[1:NZ + 1, :, 1:NX + 1] = True
    else:
        m = np.zeros((NZ + 1, NY + 2, NX + 2), bool)
        m[:, 1:NY + 1, 1:NX + 1] = True
    return m.ravel()


@pytest.mark.parametrize("npml,lossy", [(3, True), (0, False), (2, False)])
def test_adjoint_step_is_transpose(npml, lossy):
    rng = np.random.default_rng(1 + npml)
    o = make(npml)
    o.set_material(1 + rng.uniform(0, 3, NM), rng.uniform(0, 2.0, NM) if lossy else None)
    nst = 6 + (12 if npml else 0)
    x, y = {}, {}
    for w in range(6):
        v = rng.standard_normal(o.field(w).size)
        if w < 3:
            v *= e_interior(w)
        x[w] = v
        o.set_field(w, v)
    for p in range(12 if npml else 0):
        v = rng.standard_normal(o.uni_size())
        x[6 + p] = v
        o.set_psi(p, v)
    o.bare_step()
    Fx = {w: o.field(w) for w in range(6)}
    for p in range(12 if npml else 0):
        Fx[6 + p] = o.psi(p)
    for w in range(nst):
        v = rng.standard_normal(o.adjoint(w).size)
        if w < 3:
            v *= e_interior(w)
        y[w] = v
        o.set_adjoint(w, v)
    o.adjoint_bare_step()
    lhs = sum(Fx[w] @ y[w] for w in range(nst))
    rhs = sum(x[w] @ o.adjoint(w) for w in range(nst))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def _setup(o, er, sg):
    o.set_material(er, sg)
    o.setSource_z((4, 4, 5))
    o.setSource_x((3, 5, 4))
    o.setObservationPoint([[(2, 3, 3)], [(6, 5, 6)], [(5, 2, 7), (5, 3, 7)]])
    o.setObservationPoint([[], [], [(3, 6, 3)]])


def _model(seed):
    rng = np.random.default_rng(seed)
    return 1 + rng.uniform(0, 3, NM), rng.uniform(0, 2.0, NM)


def _observed(er, sg, nsteps):
    er_t = er.copy()
    er_t[::7] += 0.5
    o = make()
    _setup(o, er_t, sg * 1.1)
    o.run(nsteps)
    return o.traces()


def test_gradient_matches_finite_differences():
    nsteps = 60
    er, sg = _model(3)
    dobs = _observed(er, sg, nsteps)

    def J(e, s):
        o = make()
        _setup(o, e, s)
        return o.gradient(nsteps, dobs)

    J0, ge, gs = J(er, sg)
    assert J0 > 0
    rng = np.random.default_rng(4)
    d = 1e-5
    for g, which in ((ge, 0), (gs, 1)):
        v = rng.standard_normal(NM)
        if which == 0:
            fd = (J(er + d * v, sg)[0] - J(er - d * v, sg)[0]) / (2 * d)
        else:
            fd = (J(er, sg + d * v)[0] - J(er, sg - d * v)[0]) / (2 * d)
        assert abs(fd - g @ v) <= 1e-7 * abs(g @ v)


def test_gradient_zero_at_true_model_and_traces_match_run():
    nsteps = 30
    er, sg = _model(5)
    o = make()
    _setup(o, er, sg)
    o.run(nsteps)
    tr = o.traces()
    o2 = make()
    _setup(o2, er, sg)
    J, ge, gs = o2.gradient(nsteps, tr)
    assert J == 0.0
    assert not ge.any() and not gs.any()
    np.testing.assert_array_equal(o2.traces()[:, 1:], tr[:, 1:])
