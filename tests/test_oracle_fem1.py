"""Pins of the FP64 oracle of the first-order stabilized hex solvers
(oracle/fem1_oracle.c, oracle/fem1.py): src/fem (leapfrog) and src/fem_abm.

The reference needs Eigen, which is not in this image, so it is not compiled;
the restatement is pinned by what the mathematics fixes:
  - the element tables (fixture extracted from the reference sources) annihilate
    discrete gradients (A G = 0, G^T C = 0) and B is symmetric positive definite,
    so K = c C_s B_s^-1 A_s has the gradient null space on both sides;
  - the assembled leapfrog operator has the sampled PEC-cavity TE110 mode as an
    exact eigenvector, and its frequency converges at second order to
    c pi sqrt(2) / L;
  - the leapfrog discrete energy |E^{n+1}-E^n|^2 + E^{n+1}.K E^n is conserved;
  - time convergence orders (leapfrog 2, fem_abm 1: the reference's predictor
    is not PEC-projected, so the boundary lowers its order);
  - fem_abm's E side is the adjoint of its H side, and a source-free cavity
    does not grow;
  - the two source functions are tied by sf_fem = -mu d/dt sf_abm;
  - the dipole pattern's symmetry and support.
"""
import numpy as np
import pytest

from oracle import fem1

EPS0, MU0 = 8.854187817e-12, 1.25663706212e-6
C0 = 1.0 / np.sqrt(EPS0 * MU0)


def elem_gradient():
    """12 x 8 edge-node incidence of one hex in the reference's local order."""
    G = np.zeros((12, 8))

    def node(a, b, c):
        return a + 2 * b + 4 * c
    for b in range(2):
        for c in range(2):
            G[b + 2 * c, node(1, b, c)] += 1
            G[b + 2 * c, node(0, b, c)] -= 1
    for a in range(2):
        for c in range(2):
            G[4 + 2 * a + c, node(a, 1, c)] += 1
            G[4 + 2 * a + c, node(a, 0, c)] -= 1
    for a in range(2):
        for b in range(2):
            G[8 + a + 2 * b, node(a, b, 1)] += 1
            G[8 + a + 2 * b, node(a, b, 0)] -= 1
    return G


def arr(t):
    return np.array([[float(x) for x in r] for r in t])


def test_tables_annihilate_gradients():
    T = fem1.tables()
    G = elem_gradient()
    for key in ("fem_ofem0", "fem_ofem1"):
        t = T[key]
        for n in ("Amat", "_Amat"):
            assert np.abs(arr(t[n]) @ G).max() == 0, (key, n)
        for n in ("Cmat", "_Cmat"):
            assert np.abs(G.T @ arr(t[n])).max() == 0, (key, n)
        B = arr(t["Bmat"])
        assert np.array_equal(B, B.T)
        assert np.linalg.eigvalsh(B).min() > 0
    for n in ("Amat", "_Amat"):
        assert np.abs(arr(T["fem_abm"][n]) @ G).max() == 0
    # the conventional tables of fem and fem_abm are the same
    assert T["fem_abm"]["Amat"] == T["fem_ofem0"]["Amat"]
    assert T["fem_abm"]["_Amat"] == T["fem_ofem0"]["_Amat"]


@pytest.mark.parametrize("ofem", [0, 1])
def test_element_matrix_nullspace_and_symmetry(ofem):
    h = 0.01
    dt = 0.5 * h / C0 / np.sqrt(3)
    K = fem1.element_matrix(fem1.FEM, dt, h, EPS0, MU0, 0.3, ofem)
    G = elem_gradient()
    scale = np.abs(K).max()
    assert np.abs(K @ G).max() <= 1e-14 * scale
    assert np.abs(G.T @ K).max() <= 1e-14 * scale
    assert np.abs(K - K.T).max() <= 1e-14 * scale
    # positive semidefinite with the 7-dimensional gradient null space
    ev = np.sort(np.linalg.eigvalsh(0.5 * (K + K.T)))
    assert ev[0] > -1e-13 * scale
    assert (ev < 1e-12 * scale).sum() == 7
    # exact rational evaluation agrees with a float64 solve
    tab = fem1.tables()["fem_ofem1" if ofem else "fem_ofem0"]
    As, Bs, Cs = (np.array(fem1.stabilized(tab, n, 0.3)) for n in ("Amat", "Bmat", "Cmat"))
    c = 0.5 * dt * dt / MU0 / EPS0 / h / h
    Kf = c * Cs @ np.linalg.solve(Bs, As)
    assert np.abs(Kf - K).max() <= 1e-12 * scale


def pec_mask(N):
    ex = np.zeros((N + 1, N + 1, N), bool)
    ex[:, [0, N], :] = True
    ex[[0, N], :, :] = True
    ey = np.zeros((N + 1, N, N + 1), bool)
    ey[:, :, [0, N]] = True
    ey[[0, N], :, :] = True
    ez = np.zeros((N, N + 1, N + 1), bool)
    ez[:, :, [0, N]] = True
    ez[:, [0, N], :] = True
    return np.concatenate([ex.ravel(), ey.ravel(), ez.ravel()])


def te110(N, L=1.0):
    h = L / N
    x = np.arange(N + 1) * h
    ez = np.sin(np.pi * x / L)[None, None, :] * np.sin(np.pi * x / L)[None, :, None] * np.ones((N, 1, 1))
    n = N * (N + 1) * (N + 1)
    return np.concatenate([np.zeros(2 * n), ez.ravel()])


@pytest.mark.parametrize("ofem", [0, 1])
def test_fem_cavity_mode_eigenvector_and_frequency(ofem):
    L = 1.0
    we = C0 * np.pi * np.sqrt(2) / L
    err = []
    for N in (8, 16):
        h = L / N
        dt = 0.5 * h / C0 / np.sqrt(3)
        o = fem1.OracleFem1((N, N, N), h, dt, EPS0, MU0, 1, fem1.FEM, 0.3, ofem)
        v = te110(N, L)
        Kv = o.apply(v)
        Kv[pec_mask(N)] = 0
        lam = v @ Kv / (v @ v)
        assert np.linalg.norm(Kv - lam * v) <= 1e-12 * np.linalg.norm(Kv)
        err.append(abs(np.arccos(1 - lam / 2) / dt / we - 1))
    assert err[0] < 0.025
    assert 3.6 < err[0] / err[1] < 4.4


def test_fem_leapfrog_energy_conserved():
    N = 6
    h = 0.05
    dt = 0.5 * h / C0 / np.sqrt(3)
    o = fem1.OracleFem1((N, N, N), h, dt, EPS0, MU0, 1, fem1.FEM, 0.3, 0)
    rng = np.random.default_rng(5)
    m = pec_mask(N)
    sizes = [o.size(q) for q in range(3)]
    x0 = rng.standard_normal(sum(sizes))
    x0[m] = 0
    x1 = x0 + 0.01 * rng.standard_normal(sum(sizes))
    x1[m] = 0
    split = lambda v: np.split(v, np.cumsum(sizes)[:-1])
    for q, (a, b) in enumerate(zip(split(x0), split(x1))):
        o.set_field(q, 0, a)
        o.set_field(q, 1, b)

    def energy():
        a = np.concatenate([o.field(q, 0) for q in range(3)])
        b = np.concatenate([o.field(q, 1) for q in range(3)])
        Kb = o.apply(b)
        Kb[m] = 0
        return (b - a) @ (b - a) + a @ Kb
    w0 = energy()
    for _ in range(10):
        o.run(20)
        assert abs(energy() - w0) <= 1e-11 * abs(w0)


def _run_dipole(stepper, stab, ns, T=1.6e-9, N=10, h=0.05):
    o = fem1.OracleFem1((N, N, N), h, T / ns, EPS0, MU0, 1, stepper, stab)
    o.setSourceDipole_x((4, 4, 4))
    o.run(ns)
    return np.concatenate([o.field(q, 1) for q in range(3)])


@pytest.mark.parametrize("stepper,stab,lo,hi", [(fem1.FEM, 0.3, 3.6, 4.5), (fem1.ABM, 0.05, 1.7, 2.6)])
def test_time_convergence_order(stepper, stab, lo, hi):
    ref = _run_dipole(stepper, stab, 1600)
    e = [np.linalg.norm(_run_dipole(stepper, stab, ns) - ref) / np.linalg.norm(ref) for ns in (100, 200)]
    assert lo < e[0] / e[1] < hi, e


def test_abm_e_side_is_adjoint_of_h_side():
    N = 5
    o = fem1.OracleFem1((N, 4, 6), 0.1, 1e-12, EPS0, MU0, 1, fem1.ABM, 0.05)
    n = sum(o.size(q) for q in range(3))
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    lhs = o.apply(x, transpose=True) @ y
    rhs = x @ o.apply(y, transpose=False)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1)
    # the H side annihilates discrete gradients of nodal fields
    phi = rng.standard_normal((7, 5, 6))  # nodes [k][j][i] of a 5 x 4 x 6 mesh
    gx = (phi[:, :, 1:] - phi[:, :, :-1]).ravel()
    gy = (phi[:, 1:, :] - phi[:, :-1, :]).ravel()
    gz = (phi[1:, :, :] - phi[:-1, :, :]).ravel()
    g = np.concatenate([gx, gy, gz])
    assert np.abs(o.apply(g, transpose=False)).max() <= 1e-13 * np.abs(g).max()


def test_abm_cavity_does_not_grow():
    N = 8
    L = 1.0
    h = L / N
    dt = 0.2 * h / C0 / np.sqrt(3)
    o = fem1.OracleFem1((N, N, N), h, dt, EPS0, MU0, 1, fem1.ABM, 0.05)
    v = te110(N, L)
    ez = v[-o.size(2):]
    o.set_field(2, 0, ez)
    o.set_field(2, 1, ez)
    c = N // 2
    o.setObservationPoint([[], [], [(c, c, c)]])
    period = 2 * np.pi / (C0 * np.pi * np.sqrt(2) / L)
    o.run(int(3 * period / dt))
    tr = o.traces()[0, 1:, 2]
    assert np.abs(tr).max() <= 1.0 + 1e-12
    assert np.abs(tr[-len(tr) // 3:]).max() > 0.5  # still oscillating


def test_source_functions():
    mu, h = MU0, 0.02
    ts = np.linspace(0.1e-9, 2.5e-9, 25)
    d = 1e-15
    for t in ts:
        fd = (fem1.fem_source(fem1.ABM, t + d, mu, h) - fem1.fem_source(fem1.ABM, t - d, mu, h)) / (2 * d)
        ff = fem1.fem_source(fem1.FEM, t, mu, h)
        assert abs(ff + mu * fd) <= 1e-6 * max(abs(ff), 1e-3 * mu / h ** 3 * 4e18)
    # sf_abm is odd about t0 = 1/f0 and vanishes there
    assert fem1.fem_source(fem1.ABM, 1e-9, mu, h) == 0.0
    for t in (0.3e-9, 0.7e-9):
        a = fem1.fem_source(fem1.ABM, 1e-9 - t, mu, h)
        b = fem1.fem_source(fem1.ABM, 1e-9 + t, mu, h)
        assert abs(a + b) <= 1e-12 * abs(a)


def _one_step_pattern(stepper, ofem=0, stab=0.3):
    N = 8
    h = 0.05
    o = fem1.OracleFem1((N, N, N), h, 1e-11, EPS0, MU0, 1, stepper, stab, ofem)
    o.setSourceDipole_x((3, 2, 2))
    o.run(1)
    return [o.field(q, 1) for q in range(3)], N


def test_dipole_pattern_support_and_symmetry():
    (ex, ey, ez), N = _one_step_pattern(fem1.FEM)
    ex = ex.reshape(N + 1, N + 1, N)  # [k][j][i]
    nz = np.argwhere(ex != 0)
    # 9 x-edges at i = 3, j, k in 2..4, symmetric about (j, k) = (3, 3)
    assert len(nz) == 9 and set(nz[:, 2]) == {3}
    blk = ex[2:5, 2:5, 3]
    assert np.array_equal(blk, blk[::-1, :]) and np.array_equal(blk, blk[:, ::-1])
    assert np.array_equal(blk, blk.T)
    ab = np.abs(blk)
    assert ab[1, 1] == ab.max() and ab[0, 0] == ab.min() > 0 and np.all(np.sign(blk) == np.sign(blk[1, 1]))
    assert blk[1, 1] == 36.0 * blk[0, 0] and blk[0, 1] == 6.0 * blk[0, 0]  # 9/2, 3/4, 1/8 of one value
    # stabilization terms: 8 y- and 8 z-edges, antisymmetric along the
    # dipole's transverse direction
    assert (ey != 0).sum() == 8 and (ez != 0).sum() == 8
    assert abs(ey.sum()) <= 1e-12 * np.abs(ey).max()
    assert abs(ez.sum()) <= 1e-12 * np.abs(ez).max()
    (ex1, ey1, ez1), _ = _one_step_pattern(fem1.FEM, ofem=1)
    assert (ex1 != 0).sum() == 1 and (ey1 != 0).sum() == 0 and (ez1 != 0).sum() == 0
    assert ex1[ex1 != 0][0] == 64.0 * blk[0, 0]  # 8 temp vs temp / 8


def test_receivers_average_and_shape():
    N = 6
    o = fem1.OracleFem1((N, N, N), 0.05, 1e-11, EPS0, MU0, 2, fem1.FEM, 0.3, 0)
    o.setObservationPoint([[(2, 2, 2), (3, 2, 2)], [(1, 1, 1)], [(2, 3, 1), (2, 3, 2), (2, 3, 3)]])
    o.setSourceDipole_x((2, 1, 1))
    o.run(9)
    tr = o.traces()
    assert tr.shape == (1, 9 // 2 + 1, 3)
    assert np.all(tr[0, 0] == 0)
    ex = o.field(0, 1)
    ey = o.field(1, 1)
    ez = o.field(2, 1)
    fx = lambda i, j, k: i + j * N + k * N * (N + 1)
    fy = lambda i, j, k: i + j * (N + 1) + k * (N + 1) * N
    fz = lambda i, j, k: i + j * (N + 1) + k * (N + 1) * (N + 1)
    # the last recorded step (8) is two steps before the end, so re-run a
    # fresh instance to step 8 and compare its fields with row 4
    o2 = fem1.OracleFem1((N, N, N), 0.05, 1e-11, EPS0, MU0, 2, fem1.FEM, 0.3, 0)
    o2.setSourceDipole_x((2, 1, 1))
    o2.run(8)
    ex, ey, ez = (o2.field(q, 1) for q in range(3))
    assert tr[0, 4, 0] == (0.0 + ex[fx(2, 2, 2)] + ex[fx(3, 2, 2)]) / 2
    assert tr[0, 4, 1] == ey[fy(1, 1, 1)] / 1
    assert tr[0, 4, 2] == ((0.0 + ez[fz(2, 3, 1)] + ez[fz(2, 3, 2)]) + ez[fz(2, 3, 3)]) / 3
