"""Pins of the FP64 FEM oracle (fem_v2 / symplectic_2 / symplectic_4), CPU only.

The reference FEM solvers need Eigen and MPI, which this image lacks, so they
cannot be compiled here; the oracle is pinned by the reference-held element
operators (kmat/*.csv, copied as data fixtures into tests/golden/kmat) and by
what the mathematics fixes:
  - kmat: exact symmetry, positive semi-definiteness, the discrete-gradient
    nullspace for p = 1 and the nullities of p = 2, 3;
  - the assembled operator is symmetric (dot-product test);
  - the leapfrog scheme conserves |E^{n+1}-E^n|^2 + E^{n+1}.K E^n;
  - the lowest PEC-cavity eigenfrequency converges to c*pi*sqrt(2)/L;
  - symplectic_2 / symplectic_4 reach orders 2 / 4 on an eigenmode and are
    time reversible;
  - the receivers average the four surrounding x / y edges.
"""
import numpy as np
import pytest

from oracle import fem as om

EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6
C0 = 1.0 / (EPS0 * MU0) ** 0.5


# ----------------------------------------------------------------- kmat
@pytest.mark.parametrize("name,order", [("1", 1), ("lamped1", 1), ("2", 2), ("3", 3)])
def test_kmat_symmetric_psd(name, order):
    K = om.load_kmat(om.kmat_path(order, 0 if name.startswith("lamped") else 1))
    n = 3 * order * (order + 1) ** 2
    assert K.shape == (n, n)
    assert np.array_equal(K, K.T)
    w = np.linalg.eigvalsh(K)
    assert w.min() > -1e-10 * w.max()


def p1_gradient():
    """Discrete gradient of the 8 nodal hats in the fem_v2 local edge order."""
    G = np.zeros((12, 8))
    node = lambda x, y, z: x + 2 * y + 4 * z
    for m in range(2):
        for n in range(2):
            r = m + 2 * n  # x-edge at (y=m, z=n)
            G[r, node(1, m, n)] += 1
            G[r, node(0, m, n)] -= 1
    for l in range(2):
        for n in range(2):
            r = 4 + l + 2 * n  # y-edge at (x=l, z=n)
            G[r, node(l, 1, n)] += 1
            G[r, node(l, 0, n)] -= 1
    for l in range(2):
        for m in range(2):
            r = 8 + l + 2 * m  # z-edge at (x=l, y=m)
            G[r, node(l, m, 1)] += 1
            G[r, node(l, m, 0)] -= 1
    return G


@pytest.mark.parametrize("name", ["1", "lamped1"])
def test_kmat_p1_annihilates_gradients(name):
    K = om.load_kmat(om.kmat_path(1, 0 if name == "lamped1" else 1))
    assert np.abs(K @ p1_gradient()).max() == 0.0
    assert np.linalg.matrix_rank(K) == 5


@pytest.mark.parametrize("order,nullity", [(2, 10), (3, 29)])
def test_kmat_higher_order_nullity(order, nullity):
    K = om.load_kmat(om.kmat_path(order))
    assert K.shape[0] - np.linalg.matrix_rank(K, tol=1e-9 * np.abs(K).max()) == nullity


def test_parse_value_forms():
    assert om.parse_value('"7/4"') == 1.75
    assert om.parse_value(' "-3/8" \r') == -0.375
    assert om.parse_value("2.5") == 2.5
    assert om.parse_value('""') == 0.0


# ------------------------------------------------------- assembled operator
def dense_operator(o):
    n = o.ndof()
    A = np.zeros((n, n))
    e = np.zeros(n)
    for i in range(n):
        e[i] = 1.0
        A[:, i] = o.apply(e)
        e[i] = 0.0
    return A


@pytest.mark.parametrize("order,grid", [(1, (3, 4, 5)), (2, (4, 6, 4)), (3, (6, 6, 6))])
def test_assembled_operator_symmetric(order, grid):
    o = om.OracleFEMv2(order, grid, 1e-2, 1e-12, EPS0, MU0, stepper=om.SYMPLECTIC2)
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(o.ndof()), rng.standard_normal(o.ndof())
    a, b = x @ o.apply(y), y @ o.apply(x)
    assert abs(a - b) <= 1e-12 * (abs(a) + abs(b))


@pytest.mark.parametrize("order,grid", [(1, (3, 4, 5)), (2, (4, 6, 4)), (3, (6, 6, 9))])
def test_separable_fields_in_global_nullspace(order, grid):
    """Every kmat operator annihilates element fields whose x-edge values
    depend only on the x sub-index (and likewise for y, z): curl-free fields
    (f(x), g(y), h(z)).  On the assembled mesh, random f, g, h of the fine
    index must give K E = 0 exactly up to rounding; a wrong local ordering
    or global connectivity mixes the axes and breaks it."""
    o = om.OracleFEMv2(order, grid, 1e-2, 1e-12, EPS0, MU0, stepper=om.SYMPLECTIC2)
    Nx, Ny, Nz = grid
    rng = np.random.default_rng(order)
    f, g, hh = rng.standard_normal(Nx), rng.standard_normal(Ny), rng.standard_normal(Nz)
    ex = np.broadcast_to(f[None, None, :], (Nz + 1, Ny + 1, Nx)).ravel()
    ey = np.broadcast_to(g[None, :, None], (Nz + 1, Ny, Nx + 1)).ravel()
    ez = np.broadcast_to(hh[:, None, None], (Nz, Ny + 1, Nx + 1)).ravel()
    x = np.concatenate([ex, ey, ez])
    y = o.apply(x)
    ref = o.apply(rng.standard_normal(x.size))
    assert np.abs(y).max() < 1e-12 * np.abs(ref).max()
    # a mixed field (depends on two indices) is not annihilated
    ex2 = (rng.standard_normal((Nz + 1, Ny + 1, 1)) * np.ones((1, 1, Nx))).ravel()
    y2 = o.apply(np.concatenate([ex2, np.zeros(ey.size + ez.size)]))
    assert np.abs(y2).max() > 1e-3 * np.abs(ref).max()


def test_leapfrog_discrete_energy_conserved():
    grid = (4, 5, 6)
    h = 1e-2
    dt = 0.3 * h / C0
    o = om.OracleFEMv2(1, grid, h, dt, EPS0, MU0, stepper=om.LEAPFROG)
    mask = om.boundary_mask(grid)
    rng = np.random.default_rng(4)
    e0 = rng.standard_normal(o.ndof()); e0[mask] = 0
    e1 = e0 + 0.1 * rng.standard_normal(o.ndof()); e1[mask] = 0
    o.set_state(0, e0)
    o.set_state(1, e1)
    Ws = []
    for _ in range(60):
        prev = o.get_state(1)
        o.run(1)
        new = o.get_state(1)
        # the step used K scaled by dt^2: W = |new-prev|^2 + new.K prev
        Ws.append(np.dot(new - prev, new - prev) + new @ o.apply(prev))
    Ws = np.array(Ws)
    assert np.abs(Ws - Ws[0]).max() < 1e-11 * abs(Ws[0])


def interior_eigs(order, grid, h):
    """Eigenvalues of the symplectic-scaled K restricted to interior DOFs."""
    o = om.OracleFEMv2(order, grid, h, 1e-12, EPS0, MU0, stepper=om.SYMPLECTIC2)
    A = dense_operator(o)
    keep = ~om.boundary_mask(grid)
    return np.linalg.eigvalsh(A[np.ix_(keep, keep)])


@pytest.mark.parametrize("order,n", [(1, 4), (1, 8)])
def test_cavity_lowest_mode(order, n):
    """Cube of side L = n h: lowest PEC mode omega = c pi sqrt(2) / L (x3).
    The symplectic operator satisfies m U'' = -K U with m = 8 / p^3.  (The
    p = 2, 3 reference operators annihilate only 10 / 29 local vectors
    instead of all element gradients, so their low spectrum holds spurious
    gradient-like modes; they are pinned by the nullspace test below.)"""
    h = 1e-2
    lam = interior_eigs(order, (n, n, n), h)
    m = 8.0 / order ** 3
    nz = lam[lam > 1e-6 * lam.max()]
    w = np.sqrt(nz[:3] / m)
    exact = C0 * np.pi * np.sqrt(2.0) / (n * h)
    err = np.abs(w - exact) / exact
    assert err.max() < (0.06 if n == 4 else 0.02), err
    assert np.abs(w - w[0]).max() < 1e-8 * w[0]  # three-fold degenerate


def test_cavity_convergence_second_order():
    h = 1e-2
    errs = []
    for n in (4, 8):
        lam = interior_eigs(1, (n, n, n), h * 4 / n)
        w = np.sqrt(lam[lam > 1e-6 * lam.max()][0] / 8.0)
        exact = C0 * np.pi * np.sqrt(2.0) / (4 * h)
        errs.append(abs(w - exact) / exact)
    assert errs[1] < errs[0] / 3.0, errs


# ------------------------------------------------------------ symplectic
def mode_run(stepper, order, grid, h, dt, T):
    o = om.OracleFEMv2(order, grid, h, dt, EPS0, MU0, stepper=stepper)
    A = dense_operator(o)
    keep = ~om.boundary_mask(grid)
    lam, vec = np.linalg.eigh(A[np.ix_(keep, keep)])
    i = np.argmax(lam > 1e-6 * lam.max())
    v = np.zeros(o.ndof())
    v[keep] = vec[:, i]
    w = np.sqrt(lam[i] / (8.0 / order ** 3))
    o.set_state(0, v)
    o.set_state(1, np.zeros(o.ndof()))
    nsteps = int(round(T / dt))
    o.run(nsteps)
    return np.linalg.norm(o.get_state(0) - np.cos(w * nsteps * dt) * v), w


@pytest.mark.parametrize("stepper,rate", [(om.SYMPLECTIC2, 4.0), (om.SYMPLECTIC4, 16.0)])
def test_symplectic_order_of_accuracy(stepper, rate):
    grid, h = (4, 4, 4), 1e-2
    _, w = mode_run(stepper, 1, grid, h, 1e-13, 1e-13)
    T = 2.5 * np.pi / w  # cos has maximal slope there: error linear in phase
    dt = T / 50
    e1, _ = mode_run(stepper, 1, grid, h, dt, T)
    e2, _ = mode_run(stepper, 1, grid, h, dt / 2, T)
    assert rate * 0.7 < e1 / e2 < rate * 1.4, e1 / e2


def test_symplectic_stage_is_time_reversible():
    grid, h = (4, 4, 6), 1e-2
    dt = 0.2 * h / C0
    o = om.OracleFEMv2(2, grid, h, dt, EPS0, MU0, stepper=om.SYMPLECTIC2)
    rng = np.random.default_rng(5)
    mask = om.boundary_mask(grid)
    u = rng.standard_normal(o.ndof()); u[mask] = 0
    o.set_state(0, u)
    o.set_state(1, np.zeros(o.ndof()))
    for _ in range(30):
        o.stage(dt)
    for _ in range(30):
        o.stage(-dt)
    assert np.linalg.norm(o.get_state(0) - u) < 1e-12 * np.linalg.norm(u)
    assert np.linalg.norm(o.get_state(1)) < 1e-12 * np.linalg.norm(u) * C0 / h


# -------------------------------------------------------------- receivers
def test_receiver_averages_surrounding_edges():
    """Ex = y + z, Ey = x + 2 z, Ez = 3 x.  The x receiver averages the x
    edges (i-1|i, j, k|k+1) and the y receiver the y edges (i, j-1|j, k|k+1),
    so the averages are j + k + 1/2 and i + 2 k + 1; Ez is one edge."""
    grid = (6, 6, 6)
    Nx, Ny, Nz = grid
    o = om.OracleFEMv2(1, grid, 1e-2, 1e-14, EPS0, MU0, stepper=om.SYMPLECTIC2)
    z, y, x = np.meshgrid(np.arange(Nz + 1), np.arange(Ny + 1), np.arange(Nx), indexing="ij")
    ex = (y + z).astype(float)
    z, y, x = np.meshgrid(np.arange(Nz + 1), np.arange(Ny), np.arange(Nx + 1), indexing="ij")
    ey = (x + 2 * z).astype(float)
    z, y, x = np.meshgrid(np.arange(Nz), np.arange(Ny + 1), np.arange(Nx + 1), indexing="ij")
    ez = (3 * x).astype(float)
    o.set(0, 0, ex); o.set(1, 0, ey); o.set(2, 0, ez)
    i, j, k = 3, 2, 4
    o.setObservationPoint((i, j, k))
    o.run(1)  # symplectic records U before the step
    t = o.traces()[0, 0]
    assert t[0] == 0.0
    assert t[1] == j + k + 0.5
    assert t[2] == i + 2 * k + 1
    assert t[3] == 3 * i


def test_fem_source_properties():
    """Second derivative of exp(-zeta (t - 1/f)^2), times 1e-3."""
    f = 1e9
    zeta = 2 * np.pi ** 2 * f ** 2
    assert abs(om.fem_source(1 / f) + 2 * zeta * 1e-3) < 1e-12 * 2 * zeta * 1e-3
    tz = 1 / f + 1 / np.sqrt(2 * zeta)
    assert abs(om.fem_source(tz)) < 1e-9 * 2 * zeta * 1e-3
    t = np.linspace(0, 2 / f, 20001)
    s = np.array([om.fem_source(v) for v in t])
    # integral of a second derivative of a (numerically) compact pulse ~ 0
    assert abs(np.trapezoid(s, t)) < 1e-6 * np.abs(s).max() * (2 / f)
