"""Pins of the tetrahedral edge-element oracle (oracle/tet.py) by what the
mathematics fixes: the discrete de Rham null space, the PEC cube cavity
spectrum, closed-form element matrices, exact interpolation of constant
fields and the convergence orders of the steppers on a discrete eigenmode."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import tet as ot


def interior_pencil(n, eps=1.0, mu=1.0):
    m = ot.Mesh(n, n, n, 1.0 / n)
    S, Me, _ = ot.assemble(m, eps, mu)
    keep = ~m.boundary()
    return m, S[keep][:, keep].toarray(), Me[keep][:, keep].toarray(), keep


def test_edge_counts():
    m = ot.Mesh(2, 3, 4, 0.5)
    # 7 families: x, y, z edges + 3 face diagonals + 1 body diagonal per cube
    nx, ny, nz = 2, 3, 4
    exp = (nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
           + nx * ny * (nz + 1) + nx * (ny + 1) * nz + (nx + 1) * ny * nz + nx * ny * nz)
    assert m.ne == exp
    E = m.tet_edges()
    assert E.shape == (6 * nx * ny * nz, 6)
    assert E.min() == 0 and E.max() == m.ne - 1
    # every edge belongs to some tetrahedron
    assert np.unique(E).size == m.ne


def test_kuhn_split_fills_the_cube():
    vols = [ot.bary_grad(ot.tet_vertices(t).astype(float))[1] for t in range(6)]
    assert sum(vols) == pytest.approx(1.0, abs=1e-15)
    assert all(v == pytest.approx(1 / 6) for v in vols)


@pytest.mark.parametrize("t", range(6))
def test_element_matrices_closed_form(t):
    """Quadrature of the Whitney functions equals the closed forms
    M_lm = sum of g.g V (1 + d_ij)/20 moments, S_lm = 4 V (g_a x g_b).(g_c x g_d)."""
    h = 0.37
    M, S = ot.element_matrices(t, h)
    G, V = ot.bary_grad(ot.tet_vertices(t) * h)
    I2 = lambda i, j: V * (2.0 if i == j else 1.0) / 20.0
    Mc = np.zeros((6, 6))
    Sc = np.zeros((6, 6))
    for l, (a, b) in enumerate(ot.LOC_EDGE):
        for q, (c, d) in enumerate(ot.LOC_EDGE):
            Mc[l, q] = (G[b] @ G[d]) * I2(a, c) - (G[b] @ G[c]) * I2(a, d) - (G[a] @ G[d]) * I2(b, c) \
                + (G[a] @ G[c]) * I2(b, d)
            Sc[l, q] = 4 * V * np.cross(G[a], G[b]) @ np.cross(G[c], G[d])
    np.testing.assert_allclose(M, Mc, rtol=1e-13, atol=1e-15 * h)
    np.testing.assert_allclose(S, Sc, rtol=1e-13, atol=1e-13 / h)
    # symmetric positive definite mass, positive semi-definite curl-curl of rank 3
    assert np.all(np.linalg.eigvalsh(M) > 0)
    ev = np.linalg.eigvalsh(S)
    assert np.sum(ev > 1e-10 * ev.max()) == 3


def test_gradients_are_curl_free():
    """S G = 0: the incidence matrix maps vertex potentials into ker S."""
    m = ot.Mesh(3, 2, 4, 0.5)
    S, Me, _ = ot.assemble(m, 1.0, 1.0)
    G = m.vertex_incidence()
    # without PEC rows (full curl-curl)
    E = m.tet_edges()
    Sfull = np.zeros((m.ne, m.ne))
    for tid, e in enumerate(E):
        Sfull[np.ix_(e, e)] += ot.element_matrices(tid % 6, m.h)[1]
    assert np.abs(Sfull @ G.toarray()).max() < 1e-12 * np.abs(Sfull).max()


@pytest.mark.slow
def test_pec_cavity_null_space_and_lowest_modes():
    """dim ker S = interior vertices; the lowest cavity modes of the unit cube
    are (pi^2) (1+1+0) (x3) and (pi^2) (1+1+1) (x2), converging as h^2."""
    errs = []
    for n in (3, 4, 6):
        m, S, M, keep = interior_pencil(n)
        w = sla.eigh(S, M, eigvals_only=True)
        tol = 1e-8 * w.max()
        nz = np.sum(np.abs(w) < tol)
        assert nz == (n - 1) ** 3
        w = np.sort(w[np.abs(w) >= tol])
        lam110, lam111 = 2 * np.pi ** 2, 3 * np.pi ** 2
        assert np.all(np.abs(w[:3] / lam110 - 1) < 0.3)
        assert np.all(np.abs(w[3:5] / lam111 - 1) < 0.3)
        # the degenerate triplet stays within the discretisation error of one another
        errs.append(np.abs(w[:3] / lam110 - 1).max())
    # second-order convergence of the (1,1,0) triplet: 3 -> 6 quarters the error
    assert errs[2] < errs[0] / 3.0
    assert errs[2] < 0.05


def test_receivers_reproduce_constant_field():
    m = ot.Mesh(3, 3, 2, 0.4)
    P, D = m.edge_geometry()
    E0 = np.array([0.3, -1.2, 0.7])
    U = D @ E0  # edge integrals of a constant field
    rng = np.random.default_rng(3)
    for _ in range(20):
        x = rng.uniform(0, 1, 3) * m.n * m.h
        e, W = m.basis_at(x)
        np.testing.assert_allclose(U[e] @ W, E0, rtol=0, atol=1e-13)


def test_whitney_tangential_moments():
    """The edge integral of w_l along edge k is d_lk (Whitney dual basis)."""
    h = 0.5
    for t in range(6):
        X = ot.tet_vertices(t) * h
        G, _ = ot.bary_grad(X)
        for l, (a, b) in enumerate(ot.LOC_EDGE):
            for k, (c, d) in enumerate(ot.LOC_EDGE):
                # w_l is linear along edge (c, d): Simpson is exact
                tvec = X[d] - X[c]
                val = 0.0
                for s, ws in ((0.0, 1 / 6), (0.5, 4 / 6), (1.0, 1 / 6)):
                    lam = np.zeros(4)
                    lam[c], lam[d] = 1 - s, s
                    val += ws * (lam[a] * G[b] - lam[b] * G[a]) @ tvec
                assert val == pytest.approx(1.0 if l == k else 0.0, abs=1e-13)


def _mode(n=3):
    m, S, M, keep = interior_pencil(n)
    w, X = sla.eigh(S, M)
    tol = 1e-8 * w.max()
    q = np.argmax(w > tol)
    x = np.zeros(m.ne)
    x[keep] = X[:, q]
    return w[q], x


@pytest.mark.parametrize("stepper,order", [(ot.SYMPLECTIC2, 2), (ot.SYMPLECTIC4, 4), (ot.ABM, 2)])
@pytest.mark.parametrize("sigma", [0.0, 0.8])
def test_stepper_order_on_eigenmode(stepper, order, sigma):
    """U(0) = x (a discrete mode S x = lam M x), V(0) = 0: U(t) = u(t) x with
    u'' + (sigma/eps) u' + lam u = 0."""
    lam, x = _mode(3)
    g = 0.5 * sigma
    wd = np.sqrt(lam - g * g)
    T = 1.0
    exact = np.exp(-g * T) * (np.cos(wd * T) + g / wd * np.sin(wd * T))
    errs = []
    # ABM: the startup error (F_-1 = F_0) and the corrector's error change sign
    # near dt = 0.03, so its asymptotic rate is read at smaller steps
    for dt in ((0.005, 0.0025, 0.00125) if stepper == ot.ABM else (0.04, 0.02, 0.01)):
        o = ot.OracleTet((3, 3, 3), 1.0 / 3, dt, 1.0, 1.0, stepper=stepper,
                         sigma=np.full(6 * 27, sigma) if sigma else None)
        o.set_field(0, x)
        for _ in range(int(round(T / dt))):
            o.step()
        errs.append(np.linalg.norm(o.U - exact * x) / np.linalg.norm(x))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > order - 0.3), (errs, rates)


def test_lossless_symplectic_energy_bounded():
    """Lossless symplectic steps keep the discrete energy within O(dt^2)."""
    o = ot.OracleTet((3, 3, 3), 1.0 / 3, 0.02, 1.0, 1.0, stepper=ot.SYMPLECTIC2)
    o.add_dipole([0.43, 0.52, 0.61], [0.0, 0.0, 1.0], f0=3.0, t0=0.4)
    for _ in range(60):
        o.step()
    en = []
    for _ in range(200):
        o.step()
        en.append(0.5 * o.V @ (o.Me @ o.V) + 0.5 * o.U @ (o.S @ o.U))
    en = np.array(en)
    assert en.min() > 0
    assert (en.max() - en.min()) / en.mean() < 0.05


def test_dipole_load_is_point_evaluation():
    """g . U = p'(t) (E(r0) . d) for the interpolant U of a constant field."""
    o = ot.OracleTet((3, 3, 3), 0.5, 0.01, 1.0, 1.0)
    d = np.array([0.2, 0.5, -0.4])
    o.add_dipole([0.61, 0.77, 0.83], d, f0=2.0, t0=0.5, amp=1.5)
    _, D = o.mesh.edge_geometry()
    E0 = np.array([1.1, -0.3, 0.4])
    U = D @ E0
    t = 0.37
    assert o.load(t) @ U == pytest.approx(ot.pulse_rate(t, 2.0, 0.5, 1.5) * (E0 @ d), rel=1e-12)
