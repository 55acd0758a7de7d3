"""Pins of the FP64 FDTD oracle (CPU only).

- Reference behaviour: bit-exact agreement with golden vectors written by the
  compiled reference (src/fdtd/fdtd_simulation.cpp, tests/golden/).
- Extensions (no reference counterpart): discrete energy identities of the
  Yee scheme (lossless and lossy), CPML reflection against a larger PEC
  domain, lossy decay against the continuous solution, wavelet properties.
"""
import numpy as np
import pytest

from oracle import fdtd as of
from tests.conftest import rel_l2
from tests.fdtd_cases import EPS0, MU0, C0, cfl_dt, apply_sources_observations

GOLDEN_CASES = ["fdtd_small", "fdtd_rect", "fdtd_main50", "fdtd_vacuum64"]
NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


def run_oracle_from_golden(cfg):
    o = of.OracleFDTD(cfg["grid"], cfg["h"], cfg["dt"], cfg["eps"], cfg["mu"], tf=cfg["tf"])
    apply_sources_observations(o, cfg)
    o.run(cfg["nsteps"])
    return o


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_matches_reference_golden(golden, name):
    g = golden(name)
    o = run_oracle_from_golden(g["config"])
    assert np.array_equal(o.traces(), g["traces"], equal_nan=True)
    for w, nm in enumerate(NAMES):
        f = o.field(w)
        st = np.array([f.sum(), (f * f).sum(), np.abs(f).max()])
        assert np.array_equal(st, g[nm + "_stats"]), nm
        if nm in g:
            assert np.array_equal(f, g[nm]), nm


def test_golden_traces_are_nontrivial(golden):
    for name in GOLDEN_CASES:
        tr = golden(name)["traces"]
        assert np.isfinite(tr).all()
        assert np.abs(tr).max() > 0, name


def test_material_identity_reproduces_vacuum(golden):
    """eps_r = 1, sigma = 0 arrays must give the reference bits."""
    g = golden("fdtd_rect")
    cfg = g["config"]
    nx, ny, nz = cfg["grid"]
    o = of.OracleFDTD(cfg["grid"], cfg["h"], cfg["dt"], cfg["eps"], cfg["mu"], tf=cfg["tf"])
    apply_sources_observations(o, cfg)
    n = (nx + 1) * (ny + 1) * (nz + 1)
    o.set_material(np.ones(n), np.zeros(n))
    o.run(cfg["nsteps"])
    assert np.array_equal(o.traces(), g["traces"])
    assert np.array_equal(o.field("ez"), g["ez"])


# ---------------------------------------------------------------- energy
def eps_ref_layout(grid, eps_bg, eps_r):
    """Permittivity per E sample in the reference E layouts (ghosts -> eps_bg)."""
    nx, ny, nz = grid
    er = np.ones((nz + 1, ny + 1, nx + 1)) if eps_r is None else eps_r.reshape(nz + 1, ny + 1, nx + 1)
    out = []
    for dims in ((nx + 1, ny + 2, nz + 2), (nx + 2, ny + 1, nz + 2), (nx + 2, ny + 2, nz + 1)):
        a = np.ones((dims[2], dims[1], dims[0]))
        kz, jy, ix = min(dims[2], nz + 1), min(dims[1], ny + 1), min(dims[0], nx + 1)
        a[:kz, :jy, :ix] = er[:kz, :jy, :ix]
        out.append(eps_bg * a.ravel())
    return out


def sigma_ref_layout(grid, sigma):
    nx, ny, nz = grid
    sg = sigma.reshape(nz + 1, ny + 1, nx + 1)
    out = []
    for dims in ((nx + 1, ny + 2, nz + 2), (nx + 2, ny + 1, nz + 2), (nx + 2, ny + 2, nz + 1)):
        a = np.zeros((dims[2], dims[1], dims[0]))
        kz, jy, ix = min(dims[2], nz + 1), min(dims[1], ny + 1), min(dims[0], nx + 1)
        a[:kz, :jy, :ix] = sg[:kz, :jy, :ix]
        out.append(a.ravel())
    return out


def random_interior_state(o, grid, rng):
    """Random E on the updated samples (ghosts zero) and random H."""
    nx, ny, nz = grid
    ex = np.zeros((nz + 2, ny + 2, nx + 1)); ex[1:nz + 1, 1:ny + 1, :] = rng.standard_normal((nz, ny, nx + 1))
    ey = np.zeros((nz + 2, ny + 1, nx + 2)); ey[1:nz + 1, :, 1:nx + 1] = rng.standard_normal((nz, ny + 1, nx))
    ez = np.zeros((nz + 1, ny + 2, nx + 2)); ez[:, 1:ny + 1, 1:nx + 1] = rng.standard_normal((nz + 1, ny, nx))
    o.set_field("ex", ex)
    o.set_field("ey", ey)
    o.set_field("ez", ez)
    for w in (3, 4, 5):
        o.set_field(w, 1e-3 * rng.standard_normal(o.field(w).size))


@pytest.mark.parametrize("with_material", [False, True])
def test_lossless_discrete_energy_is_conserved(with_material):
    """W^n = sum eps E^n.E^n + mu H^{n-1/2}.H^{n+1/2} is invariant for the Yee
    scheme exactly when the H-curl is the transpose of the E-curl; a wrong
    sign, index or operand breaks it."""
    grid = (9, 11, 13)
    h = 1e-3
    dt = cfl_dt(h, 0.9)
    rng = np.random.default_rng(1)
    o = of.OracleFDTD(grid, h, dt, EPS0, MU0)
    er = None
    if with_material:
        er = 1.0 + 5.0 * rng.random((grid[0] + 1) * (grid[1] + 1) * (grid[2] + 1))
        o.set_material(er, None)
    epsE = eps_ref_layout(grid, EPS0, er)
    random_interior_state(o, grid, rng)
    Wp = None
    h_prev = [o.field(w) for w in (3, 4, 5)]
    e_now = [o.field(w) for w in (0, 1, 2)]
    Ws = []
    for n in range(40):
        o.bare_step()
        h_new = [o.field(w) for w in (3, 4, 5)]
        W = sum((epsE[c] * e_now[c] ** 2).sum() for c in range(3)) + MU0 * sum(
            (h_prev[c] * h_new[c]).sum() for c in range(3))
        Ws.append(W)
        h_prev = h_new
        e_now = [o.field(w) for w in (0, 1, 2)]
    Ws = np.array(Ws)
    assert np.abs(Ws - Ws[0]).max() / abs(Ws[0]) < 1e-12


def test_lossy_discrete_energy_identity():
    """W^{n+1} - W^n = -(dt/2) sum sigma (E^{n+1} + E^n)^2 for the Ca/Cb
    update; pins the relation of Ca and Cb to sigma and eps."""
    grid = (8, 10, 12)
    h = 1e-3
    dt = cfl_dt(h, 0.9)
    rng = np.random.default_rng(2)
    o = of.OracleFDTD(grid, h, dt, EPS0, MU0)
    n = (grid[0] + 1) * (grid[1] + 1) * (grid[2] + 1)
    er = 1.0 + 3.0 * rng.random(n)
    sg = 0.5 * rng.random(n)  # sigma*dt/(2 eps) up to ~0.05
    o.set_material(er, sg)
    epsE = eps_ref_layout(grid, EPS0, er)
    sgE = sigma_ref_layout(grid, sg)
    random_interior_state(o, grid, rng)
    e_prev = [o.field(w) for w in (0, 1, 2)]
    h_m = [o.field(w) for w in (3, 4, 5)]
    o.bare_step()
    e_now = [o.field(w) for w in (0, 1, 2)]
    h_p = [o.field(w) for w in (3, 4, 5)]
    W_prev = None
    worst = 0.0
    for step in range(25):
        o.bare_step()
        e_next = [o.field(w) for w in (0, 1, 2)]
        h_pp = [o.field(w) for w in (3, 4, 5)]
        W0 = sum((epsE[c] * e_prev[c] ** 2).sum() for c in range(3)) + MU0 * sum((h_m[c] * h_p[c]).sum() for c in range(3))
        W1 = sum((epsE[c] * e_now[c] ** 2).sum() for c in range(3)) + MU0 * sum((h_p[c] * h_pp[c]).sum() for c in range(3))
        D = 0.5 * dt * sum((sgE[c] * (e_now[c] + e_prev[c]) ** 2).sum() for c in range(3))
        worst = max(worst, abs(W1 - W0 + D) / abs(W0))
        assert D > 0
        e_prev, e_now = e_now, e_next
        h_m, h_p = h_p, h_pp
    assert worst < 1e-12


def test_lossy_decay_matches_continuous_solution():
    """Uniform field in a uniform conductor: E(t) = E0 exp(-sigma t / eps).
    The centre sample is unaffected by the walls for the first steps."""
    grid = (41, 41, 41)
    h = 1e-3
    dt = cfl_dt(h, 0.9)
    sigma, er = 0.5, 2.0  # sigma*dt/(2 eps) ~ 0.024 per step
    eps = EPS0 * er
    n = 42 ** 3
    o = of.OracleFDTD(grid, h, dt, EPS0, MU0)
    o.set_material(np.full(n, er), np.full(n, sigma))
    nx, ny, nz = grid
    ex = np.zeros((nz + 2, ny + 2, nx + 1)); ex[1:nz + 1, 1:ny + 1, :] = 1.0
    o.set_field("ex", ex)
    steps = 20
    for _ in range(steps):
        o.bare_step()
    c = o.field("ex").reshape(nz + 2, ny + 2, nx + 1)[21, 21, 20]
    exact = np.exp(-sigma * steps * dt / eps)
    assert abs(c - exact) / exact < 1e-3
    assert abs(c - exact) / exact > 0  # discrete, not identical


# ------------------------------------------------------------------ CPML
def _cpml_setup(n, npml, src, rec, steps, f0):
    h = 1e-3
    dt = cfl_dt(h, 0.99)
    o = of.OracleFDTD((n, n, n), h, dt, EPS0, MU0, wavelet=of.WAVELET_RICKER, f0=f0, t0=2 ** 0.5 / f0,
                      npml=npml, pml_order=3.0, pml_r0=1e-6, pml_kappa_max=1.0, pml_alpha_max=0.0)
    o.setSource_z(src)
    o.setObservationPoint([[rec], [rec], [rec]])
    o.run(steps)
    return o.traces()[0]


@pytest.mark.slow
def test_cpml_absorbs_against_large_pec_domain():
    """A CPML-bounded 32^3 box against a 100^3 PEC box whose wall echoes do
    not reach the receiver within the run: the difference is the CPML
    reflection (and is large without the CPML)."""
    f0 = 15e9
    steps = 160
    off = (6, 0, 0)
    big = _cpml_setup(100, 0, (50, 50, 50), (56, 50, 50), steps, f0)
    small = _cpml_setup(32, 8, (16, 16, 16), (16 + off[0], 16, 16), steps, f0)
    bare = _cpml_setup(32, 0, (16, 16, 16), (16 + off[0], 16, 16), steps, f0)
    ref = big[:, 2]
    err_pml = np.abs(small[:, 2] - ref).max() / np.abs(ref).max()
    err_pec = np.abs(bare[:, 2] - ref).max() / np.abs(ref).max()
    assert err_pml < 2e-2, err_pml
    assert err_pec > 10 * err_pml, (err_pec, err_pml)


# --------------------------------------------------------------- wavelets
def test_ricker_properties():
    f0 = 10e9
    t0 = 2 ** 0.5 / f0
    assert of.ricker(t0, f0, t0) == 1.0
    tz = 1.0 / (np.pi * f0 * 2 ** 0.5)
    assert abs(of.ricker(t0 + tz, f0, t0)) < 1e-15
    assert abs(of.ricker(t0 - tz, f0, t0)) < 1e-15
    # amplitude spectrum peaks at f0
    dt = 1e-13
    t = np.arange(200000) * dt
    r = np.array([of.ricker(x, f0, t0) for x in t[:4000]])
    spec = np.abs(np.fft.rfft(r, n=200000))
    f = np.fft.rfftfreq(200000, dt)
    assert abs(f[np.argmax(spec)] - f0) / f0 < 1e-3


def test_gauss_properties():
    f0 = 5e9
    t0 = 3e-10
    assert of.gauss(t0, f0, t0) == 1.0
    assert abs(of.gauss(t0 + 1.0 / (np.pi * f0), f0, t0) - np.exp(-1.0)) < 1e-15


def test_hertz_zero_crossing_and_symmetry():
    """Derivative of a Gaussian centred on 1/f0: zero at t = 1/f0, odd about it."""
    h = 1e-3
    f0 = 1e9
    assert of.hertz(1.0 / f0, h) == 0.0
    for d in (1e-11, 1e-10, 3e-10):
        assert abs(of.hertz(1.0 / f0 + d, h) + of.hertz(1.0 / f0 - d, h)) < 1e-9 * abs(of.hertz(1.0 / f0 + d, h))


def test_check_params_reference_cases():
    # src/fdtd/main.cpp:12-27 parameters are valid
    eps = 8.854187817e-12
    mu = 1.25663706212e-6
    c = 1.0 / (eps * mu) ** 0.5
    assert of.check_params([1, 1, 1], [500, 500, 500], 3e-9, 3e-9 / 780, 1.0 / 500, c)
    assert not of.check_params([1, 1, 1], [500, 500, 500], 3e-9, 1e-9, 1.0 / 500, c)
    assert not of.check_params([0, 1, 1], [500, 500, 500], 3e-9, 1e-12, 1.0 / 500, c)
    assert not of.check_params([1, 1, 1], [500, 0, 500], 3e-9, 1e-12, 1.0 / 500, c)
    assert not of.check_params([1, 1, 1], [500, 500, 500], -1, 1e-12, 1.0 / 500, c)
    # 1-D test only: a 3-D-unstable step passes (reference quirk)
    h = 1.0 / 500
    assert of.check_params([1, 1, 1], [500, 500, 500], 3e-9, 0.9 * h / c, h, c)
