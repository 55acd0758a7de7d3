"""Checks of the FEM oracles against what the mathematics fixes: helpers of
tests/test_oracle_fem_pins.py and tests/test_oracle_mutants.py (CPU only).

Each function measures one quantity of an oracle run whose correct value is
known in closed form; the tests assert it, and the mutant test asserts that a
deliberately wrong oracle fails the same assertion.

  cavity_eigen_error   fem_v2 family, orders 1-3: the PEC-cube TM110 triplet
                       of m U'' = -K U, against (pi c sqrt 2 / L)^2; pins the
                       operator scale 4/((h p)^2 mu eps) and the mass 8/p^3
                       (src/fem_v2/fem_simulation.cpp:476,
                       src/symplectic_2/fem_simulation.cpp:473, :588).
  charge_errors        the discrete Gauss law of a point source: for any
                       vector g with interior support and g^T K = 0 (a
                       discrete gradient), g^T E(t) follows the source alone,
                       in closed form; pins the source amplitudes, signs and
                       evaluation times of every stepper
                       (src/fem_v2/fem_simulation.cpp:651,
                       src/symplectic_2/fem_simulation.cpp:619,
                       src/fem/fem_simulation.cpp:619-633,
                       src/fem_abm/fem_simulation.cpp:735-758).
  abm_step_weights     src/fem_abm's predictor and corrector weights read
                       off one step of linear probe states
                       (src/fem_abm/fem_simulation.cpp:577-612).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6
C0 = 1.0 / np.sqrt(EPS0 * MU0)
F0 = 1.0e9  # centre frequency of the reference source functions
ZETA = 2.0 * np.pi ** 2 * F0 ** 2


def pulse(t):
    """g(t) = 1e-3 exp(-zeta (t - 1/f)^2): the dipole moment whose second
    derivative is the fem_v2-family source (dl = 1e-3 m)."""
    return 1e-3 * np.exp(-ZETA * (t - 1.0 / F0) ** 2)


def pulse_rate(t):
    return -2.0 * ZETA * (t - 1.0 / F0) * pulse(t)


# ------------------------------------------------------------------ hex FEM
def assemble_femv2(o):
    """The oracle's global operator as a sparse matrix, assembled from its
    scaled element matrix and the fem_v2 element-to-edge map: local x-edge
    l + m p + n p (p+1) of element (I, J, K) is the fine x-edge
    (pI + l, pJ + m, pK + n) (src/fem_v2/fem_simulation.cpp:196-240); y and
    z edges likewise with the direction index last in its axis."""
    from oracle import fem as om
    p = o.order
    Nx, Ny, Nz = o.grid
    n = o.n
    D = n // 3
    Ks = np.ctypeslib.as_array(om.lib().ems_oracle_femv2_matrix(o._h), shape=(n * n,)).reshape(n, n).copy()
    nX, nY = Nx * (Ny + 1) * (Nz + 1), (Nx + 1) * Ny * (Nz + 1)
    l = np.arange(D)
    I, J, K = np.meshgrid(np.arange(Nx // p), np.arange(Ny // p), np.arange(Nz // p), indexing="ij")
    I, J, K = (a.ravel()[:, None] * p for a in (I, J, K))
    a, b, c = l % p, (l // p) % (p + 1), l // (p * (p + 1))
    gx = (I + a) + (J + b) * Nx + (K + c) * Nx * (Ny + 1)
    a, b, c = l % (p + 1), (l // (p + 1)) % p, l // (p * (p + 1))
    gy = nX + (I + a) + (J + b) * (Nx + 1) + (K + c) * (Nx + 1) * Ny
    a, b, c = l % (p + 1), (l // (p + 1)) % (p + 1), l // ((p + 1) ** 2)
    gz = nX + nY + (I + a) + (J + b) * (Nx + 1) + (K + c) * (Nx + 1) * (Ny + 1)
    G = np.concatenate([gx, gy, gz], axis=1)
    rows = np.repeat(G, n, axis=1).ravel()
    cols = np.tile(G, (1, n)).ravel()
    vals = np.tile(Ks.ravel(), G.shape[0])
    N = o.ndof()
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def tm_modes(N, h):
    """Sampled TM110, TM101, TM011 modes of the PEC cube of side N h (E
    tangential on the edges of the fine grid), as columns [Ex | Ey | Ez]."""
    L = N * h
    s = np.sin(np.pi * np.arange(N + 1) * h / L)
    nx = N * (N + 1) * (N + 1)
    z = np.zeros(nx)
    ez = (s[None, None, :] * s[None, :, None]) * np.ones((N, 1, 1))
    ey = (s[:, None, None] * s[None, None, :]) * np.ones((1, N, 1))
    ex = (s[:, None, None] * s[None, :, None]) * np.ones((1, 1, N))
    return np.stack([np.concatenate([z, z, ez.ravel()]), np.concatenate([z, ey.ravel(), z]),
                     np.concatenate([ex.ravel(), z, z])], axis=1)


def cavity_eigen_error(order, N, h=1e-2, shift=1.0):
    """Signed relative error of the PEC-cube TM110 eigenvalue of the
    symplectic operator, m U'' = -K U with m = 8/p^3, on an N^3 fine grid:
    the largest deviation from the exact omega^2 = 2 (pi c / L)^2 among the
    three eigenvalues (searched near shift * omega^2) whose eigenvectors
    overlap most with the sampled TM modes."""
    from oracle import fem as om
    o = om.OracleFEMv2(order, (N, N, N), h, 1e-13, EPS0, MU0, stepper=om.SYMPLECTIC2)
    A = assemble_femv2(o)
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    y = o.apply(x)
    assert np.abs(A @ x - y).max() <= 1e-12 * np.abs(y).max(), "assembly differs from the oracle's operator"
    keep = ~om.boundary_mask((N, N, N))
    exact = 2.0 * (np.pi * C0 / (N * h)) ** 2
    m = 8.0 / order ** 3
    B = (A[keep][:, keep] / (m * exact)).tocsc()
    lu = sla.splu((B - shift * sp.identity(B.shape[0], format="csc")).tocsc())
    op = sla.LinearOperator(B.shape, matvec=lu.solve, dtype=np.float64)
    mu, vec = sla.eigsh(op, k=min(30, B.shape[0] - 2), which="LM", tol=1e-12)
    lam = shift + 1.0 / mu
    Q, _ = np.linalg.qr(tm_modes(N, h)[keep])
    ov = np.linalg.norm(Q.T @ vec, axis=0)
    pick = np.argsort(-ov)[:3]
    assert ov[pick].min() > 0.9, "no eigenvector close to the TM110 triplet"
    d = lam[pick] - 1.0
    return float(d[np.argmax(np.abs(d))])


def femv2_kernel_vector(o, src):
    """A vector g with support on interior edges, g^T K = 0 and g[src] != 0:
    the projection of the source edge onto the null space of the interior
    block of K (a discrete gradient; for order 1 the gradient of the hat of
    the source edge's lower node)."""
    from oracle import fem as om
    keep = ~om.boundary_mask(o.grid)
    Ai = assemble_femv2(o)[keep][:, keep].toarray()
    w, V = np.linalg.eigh(Ai)
    N0 = V[:, np.abs(w) < 1e-9 * np.abs(w).max()]
    idx = np.flatnonzero(keep)
    e = np.zeros(idx.size)
    e[np.searchsorted(idx, src)] = 1.0
    gi = N0 @ (N0.T @ e)
    g = np.zeros(o.ndof())
    g[idx] = gi
    return g


def femv2_charge_errors(stepper, order, steps_list, T=2.0e-9):
    """Discrete Gauss law of the fem_v2 / symplectic z source.  With U'' =
    -(K/m) U - s(t)/(eps h^3) e_src and s = g'' of the dipole moment g(t)
    (the source amplitude dt^2/(eps h^3), resp. 8/(eps (hp)^3) divided by the
    mass 8/p^3), every g with g^T K = 0 gives g.U(t) = -g_src (g(t) - g(0) -
    g'(0) t)/(eps h^3) for all t, whatever the waves do.  Returns, per step
    count, max_n |g.U_n - exact(t_n)| / max |exact| over a run to T."""
    from oracle import fem as om
    p = order
    # the source edge starts at a node inside an element (order 2: the
    # element centre, whose node gradient is in the kernel)
    N = {1: 4, 2: 6, 3: 9}[p]
    c = {1: 2, 2: 3, 3: 4}[p]
    h = 0.3  # coarse cells: large stable steps, the identity does not see K
    eps = EPS0
    o0 = om.OracleFEMv2(p, (N, N, N), h, T / steps_list[0], eps, MU0, stepper=stepper)
    nX, nY = N * (N + 1) ** 2, (N + 1) * N * (N + 1)
    src = nX + nY + c + c * (N + 1) + c * (N + 1) ** 2
    g = femv2_kernel_vector(o0, src)
    errs = []
    for ns in steps_list:
        dt = T / ns
        o = om.OracleFEMv2(p, (N, N, N), h, dt, eps, MU0, stepper=stepper)
        o.setSource_z((c, c, c))
        slot = 1 if stepper == om.LEAPFROG else 0  # E^n / U
        got, t = [], []
        for k in range(ns):
            o.run(1)
            got.append(g @ o.get_state(slot))
            t.append((k + 1) * dt)
        t = np.array(t)
        exact = -g[src] * (pulse(t) - pulse(0.0) - pulse_rate(0.0) * t) / (eps * h ** 3)
        got = np.array(got)
        assert np.all(np.isfinite(got))
        errs.append(float(np.abs(got - exact).max() / np.abs(exact).max()))
    return errs


# ---------------------------------------------------- first-order stabilized
def node_gradient(grid, node):
    """Incidence vector of the six edges at an interior node (the discrete
    gradient of its hat), [Ex | Ey | Ez] in the reference layouts; +1 on the
    three edges ending at the node, -1 on the three starting there."""
    nx, ny, nz = grid
    i, j, k = node
    n0 = nx * (ny + 1) * (nz + 1)
    n1 = (nx + 1) * ny * (nz + 1)
    g = np.zeros(n0 + n1 + (nx + 1) * (ny + 1) * nz)
    g[i + j * nx + k * nx * (ny + 1)] -= 1
    g[i - 1 + j * nx + k * nx * (ny + 1)] += 1
    g[n0 + i + j * (nx + 1) + k * (nx + 1) * ny] -= 1
    g[n0 + i + (j - 1) * (nx + 1) + k * (nx + 1) * ny] += 1
    g[n0 + n1 + i + j * (nx + 1) + k * (nx + 1) * (ny + 1)] -= 1
    g[n0 + n1 + i + j * (nx + 1) + (k - 1) * (nx + 1) * (ny + 1)] += 1
    return g


def fem1_charge_errors(stepper, steps_list, T=2.0e-9):
    """Discrete Gauss law of the src/fem and src/fem_abm z point source:
    g^T K = 0 (src/fem, G^T C = 0) and A g = 0 (src/fem_abm E update A^T H),
    so g.E follows the source alone.  src/fem: E'' = sf_fem/(mu eps h^3) at
    the source with sf_fem = -mu g''(t)/h^3, i.e. g.E = -g_src (g(t) - g(0) -
    g'(0) t)/(eps h^6); src/fem_abm: E' = -sf_abm/(eps h^3) with sf_abm =
    g'(t)/h^3, i.e. g.E = -g_src (g(t) - g(0))/(eps h^6).  Returns the
    relative error per step count (leapfrog: order 2; AB2/AM3: order 3)."""
    from oracle import fem1
    N = 6
    h = 0.06 if stepper == fem1.FEM else 0.2
    c = N // 2
    g = node_gradient((N, N, N), (c, c, c))
    src = 2 * N * (N + 1) ** 2 + c + c * (N + 1) + c * (N + 1) ** 2
    errs = []
    for ns in steps_list:
        dt = T / ns
        o = fem1.OracleFem1((N, N, N), h, dt, EPS0, MU0, 1, stepper, 0.3 if stepper == fem1.FEM else 0.05)
        o.setSource_z((c, c, c))
        got, t = [], []
        for k in range(ns):
            o.run(1)
            got.append(g @ np.concatenate([o.field(q, 1) for q in range(3)]))
            t.append((k + 1) * dt)
        t = np.array(t)
        if stepper == fem1.FEM:
            exact = -g[src] * (pulse(t) - pulse(0.0) - pulse_rate(0.0) * t) / (EPS0 * h ** 6)
        else:
            exact = -g[src] * (pulse(t) - pulse(0.0)) / (EPS0 * h ** 6)
        got = np.array(got)
        assert np.all(np.isfinite(got))
        errs.append(float(np.abs(got - exact).max() / np.abs(exact).max()))
    return errs


def adams_weights(nodes):
    """Weights w_j with int_0^1 q(t) dt = sum_j w_j q(nodes_j) for every
    polynomial q of degree < len(nodes): Adams-Bashforth (nodes 0, -1) and
    Adams-Moulton (nodes 1, 0, -1) in units of the step."""
    nodes = np.asarray(nodes, dtype=np.float64)
    V = np.vander(nodes, increasing=True).T  # V[k, j] = nodes_j^k
    rhs = 1.0 / np.arange(1, len(nodes) + 1)
    return np.linalg.solve(V, rhs)


def abm_step_weights():
    """Reads src/fem_abm's time weights off single steps of linear probes,
    separately for the E and the H update.  With E' = a A^T H and H' = -b A E
    (a = 1/(4 eps h), b = 1/(4 mu h)), a PECE step whose E update has
    predictor weights (P0e at t_n, P1e at t_n-1) and corrector weights (C0e at
    t_n+1, C1e at t_n, C2e at t_n-1), and H update (P0h, P1h, C0h, C1h, C2h),
    maps the probe states (every other history entry zero) to
      H_n = v:    E_n+1 = dt a (C0e + C1e) A^T v,  H_n+1 - v = -dt^2 a b C0h P0e A A^T v
      H_n-1 = v:  E_n+1 = dt a C2e A^T v,          H_n+1 = -dt^2 a b C0h P1e A A^T v
      E_n = w:    E_n+1 - w = -dt^2 a b C0e P0h A^T A w,  H_n+1 = -dt b (C0h + C1h) A w
      E_n-1 = w:  E_n+1 = -dt^2 a b C0e P1h A^T A w,      H_n+1 = -dt b C2h A w
    (E compared on interior edges: the step zeroes tangential E on the
    faces).  With consistent predictors (P0 + P1 = 1) these give all ten
    weights; returns (P0e, P1e, C0e, C1e, C2e, P0h, P1h, C0h, C1h, C2h)."""
    from oracle import fem1
    from tests.test_oracle_fem1 import pec_mask
    N = 5
    h = 0.05
    dt = 0.1 * h / C0
    a, b = 1.0 / (4.0 * EPS0 * h), 1.0 / (4.0 * MU0 * h)
    rng = np.random.default_rng(3)
    o = fem1.OracleFem1((N, N, N), h, dt, EPS0, MU0, 1, fem1.ABM, 0.05)
    sizes = [o.size(q) for q in range(3)]
    n = sum(sizes)
    interior = ~pec_mask(N)
    v = rng.standard_normal(n)
    w = rng.standard_normal(n)
    w[~interior] = 0.0
    ATv = o.apply(v, transpose=True)
    AATv = o.apply(ATv, transpose=False)
    Aw = o.apply(w, transpose=False)
    ATAw = o.apply(Aw, transpose=True)
    split = lambda x: np.split(x, np.cumsum(sizes)[:-1])
    zero = np.zeros(n)

    def step(En, Em, Hn, Hm):
        o2 = fem1.OracleFem1((N, N, N), h, dt, EPS0, MU0, 1, fem1.ABM, 0.05)
        for base, cur, prev in ((0, En, Em), (3, Hn, Hm)):
            for q, (x1, x0) in enumerate(zip(split(cur), split(prev))):
                o2.set_field(base + q, 1, x1)
                o2.set_field(base + q, 0, x0)
        o2.run(1)
        return (np.concatenate([o2.field(q, 1) for q in range(3)]),
                np.concatenate([o2.field(3 + q, 1) for q in range(3)]))

    def coef(y, basis, mask):
        sel = mask & (np.abs(basis) > 1e-3 * np.abs(basis).max())
        r = y[sel] / basis[sel]
        assert np.abs(r - r.mean()).max() <= 1e-9 * max(abs(r.mean()), 1e-300), "probe response is not a multiple"
        return float(r.mean())

    full = np.ones(n, bool)
    E1, H1 = step(zero, zero, v, zero)
    E2, H2 = step(zero, zero, zero, v)
    E3, H3 = step(w, zero, zero, zero)
    E4, H4 = step(zero, w, zero, zero)
    c01e = coef(E1, dt * a * ATv, interior)
    c2e = coef(E2, dt * a * ATv, interior)
    c0p0h = coef(w - E3, dt * dt * a * b * ATAw, interior)
    c0p1h = coef(-E4, dt * dt * a * b * ATAw, interior)
    c0p0e = coef(v - H1, dt * dt * a * b * AATv, full)
    c0p1e = coef(-H2, dt * dt * a * b * AATv, full)
    c01h = coef(-H3, dt * b * Aw, full)
    c2h = coef(-H4, dt * b * Aw, full)
    c0e, c0h = c0p0h + c0p1h, c0p0e + c0p1e
    return (c0p0e / c0h, c0p1e / c0h, c0e, c01e - c0e, c2e,
            c0p0h / c0e, c0p1h / c0e, c0h, c01h - c0h, c2h)
