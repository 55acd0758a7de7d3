"""Pins of the tetrahedral edge-element oracle (oracle/tet.py) by what the
mathematics fixes: the discrete de Rham null space, the PEC cube cavity
spectrum, closed-form element matrices, exact interpolation of constant
fields and the convergence orders of the steppers on a discrete eigenmode."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

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


# ------------------------------------------------------------------ order 2
def test_order2_counts_match_the_64_cube_workload():
    m = ot.Mesh(64, 64, 64, 1 / 64, order=2)
    assert m.ndof == 10084736 and m.nt == 1572864
    m1 = ot.Mesh(40, 40, 40, 1 / 40)
    assert m1.ne == 462520 and m1.nt == 384000


def test_order2_face_numbering():
    m = ot.Mesh(2, 3, 2, 0.5, order=2)
    F = m.tet_faces()
    assert F.min() == 0 and F.max() == m.nf - 1
    # every interior face is shared by exactly two tetrahedra, boundary faces by one
    cnt = np.bincount(F.ravel(), minlength=m.nf)
    bf = m.dof_boundary()[2 * m.ne::2]
    assert np.all(cnt[bf] == 1) and np.all(cnt[~bf] == 2)
    # Euler: V - E + F - T = 1 for a ball
    nv = np.prod(m.n + 1)
    assert nv - m.ne + m.nf - m.nt == 1


@pytest.mark.parametrize("t", range(6))
def test_order2_element_matrices(t):
    M, S = ot.element_matrices(t, 0.3, order=2)
    assert np.allclose(M, M.T, rtol=0, atol=1e-15) and np.allclose(S, S.T, rtol=0, atol=1e-13)
    assert np.all(np.linalg.eigvalsh(M) > 0)
    ev = np.linalg.eigvalsh(S)
    assert np.sum(ev > 1e-9 * ev.max()) == 11  # 20 - dim grad P2 (9)
    # the leading 6 x 6 block is the Whitney element
    M1, S1 = ot.element_matrices(t, 0.3, order=1)
    np.testing.assert_allclose(M[:6, :6], M1, rtol=1e-13)
    np.testing.assert_allclose(S[:6, :6], S1, rtol=1e-13)


@pytest.mark.parametrize("order", [1, 2])
def test_library_element_matrices_equal_quadrature(order):
    """The library's exact monomial integration against the oracle's
    quadrature of the hand-written basis and curls (host-only call)."""
    from emsolvers_amd.tet import element_matrices as lib_em
    for t in range(6):
        for h in (1.0, 0.013):
            Ml, Sl = lib_em(order, t, h)
            Mo, So = ot.element_matrices(t, h, order)
            np.testing.assert_allclose(Ml, Mo, rtol=1e-12, atol=1e-14 * np.abs(Mo).max())
            np.testing.assert_allclose(Sl, So, rtol=1e-12, atol=1e-12 * np.abs(So).max())


def test_order2_null_space():
    """ker S (PEC) = grad of P2 Lagrange functions vanishing on the boundary:
    dimension = interior vertices + interior edges."""
    n = 2
    m = ot.Mesh(n, n, n, 1.0 / n, order=2)
    S, Me, _ = ot.assemble(m, 1.0, 1.0)
    keep = ~m.dof_boundary()
    Sk = S[keep][:, keep].toarray()
    Mk = Me[keep][:, keep].toarray()
    w = sla.eigh(Sk, Mk, eigvals_only=True)
    nz = np.sum(np.abs(w) < 1e-8 * w.max())
    interior_edges = np.sum(~m.boundary())
    assert nz == (n - 1) ** 3 + interior_edges


@pytest.mark.slow
def test_order2_cavity_fourth_order():
    errs = []
    for n in (2, 3):
        m = ot.Mesh(n, n, n, 1.0 / n, order=2)
        S, Me, _ = ot.assemble(m, 1.0, 1.0)
        keep = ~m.dof_boundary()
        w = sla.eigh(S[keep][:, keep].toarray(), Me[keep][:, keep].toarray(), eigvals_only=True)
        w = np.sort(w[w > 1e-8 * w.max()])
        lam110 = 2 * np.pi ** 2
        errs.append(np.abs(w[:3] / lam110 - 1).max())
        assert np.all(np.abs(w[3:5] / (3 * np.pi ** 2) - 1) < 0.1)
    rate = np.log(errs[0] / errs[1]) / np.log(3 / 2)
    assert rate > 3.5, (errs, rate)
    assert errs[1] < 5e-3


def test_order2_projection_reproduces_linear_fields():
    """L2 projection (full system, no PEC) of E(x) = A x + c onto the order-2
    space is exact; the receivers' evaluation then returns E(r)."""
    m = ot.Mesh(2, 2, 1, 0.5, order=2)
    Edofs = m.tet_dofs()
    rng = np.random.default_rng(5)
    A = rng.standard_normal((3, 3))
    c0 = rng.standard_normal(3)
    field = lambda x: A @ x + c0
    M = np.zeros((m.ndof, m.ndof))
    b = np.zeros(m.ndof)
    for tid, e in enumerate(Edofs):
        t = tid % 6
        cube = tid // 6
        ci = np.array([cube % m.n[0], (cube // m.n[0]) % m.n[1], cube // (m.n[0] * m.n[1])])
        X = (ot.tet_vertices(t) + ci) * m.h
        G, V = ot.bary_grad(X)
        for lam, wq in zip(ot.QPTS, ot.QW):
            W, _ = ot.basis_values(lam, G, 2)
            x = lam @ X
            M[np.ix_(e, e)] += wq * V * (W @ W.T)
            b[e] += wq * V * (W @ field(x))
    u = np.linalg.solve(M, b)
    for _ in range(10):
        x = rng.uniform(0.05, 0.95, 3) * m.n * m.h
        dofs, W = m.basis_at(x)
        np.testing.assert_allclose(u[dofs] @ W, field(x), rtol=0, atol=1e-11)


@pytest.mark.parametrize("order,grid", [(1, (3, 2, 4)), (2, (2, 3, 2)), (2, (3, 3, 3))])
def test_library_assembly_equals_oracle(order, grid):
    """The library's row-wise assembly (host-only call) against the oracle's
    scipy assembly, with per-tetrahedron eps_r and sigma."""
    from emsolvers_amd.tet import assemble as lib_assemble
    rng = np.random.default_rng(2)
    nt = 6 * int(np.prod(grid))
    er, sg = rng.uniform(1, 4, nt), rng.uniform(0, 0.5, nt)
    h, eps, mu = 0.31, 2.0, 0.7
    m = ot.Mesh(*grid, h, order=order)
    So, Mo, Go = ot.assemble(m, eps, mu, er, sg)
    Sl, Ml, Gl = lib_assemble(grid, h, eps, mu, order, er, sg)
    for A, B in ((Sl, So), (Ml, Mo), (Gl, Go)):
        D = (A - B).toarray()
        assert np.abs(D).max() <= 1e-12 * np.abs(B.toarray()).max()
    # the pattern: every nonzero of the oracle is a stored entry of the library
    Mo.eliminate_zeros()
    pat_l = sp.csr_matrix((np.ones(Ml.nnz), Ml.indices, Ml.indptr), shape=Ml.shape)
    pat_o = (abs(Mo) > 0).astype(float)
    assert (pat_o - pat_o.multiply(pat_l)).nnz == 0
    assert np.all(np.diff(Ml.indptr) >= 1)
    # columns sorted within each row
    for r in range(0, Ml.shape[0], 7):
        c = Ml.indices[Ml.indptr[r]:Ml.indptr[r + 1]]
        assert np.all(np.diff(c) > 0)


# ------------------------------------------------------------------ adjoint
def _grad_case(order, stepper, lossy, n=2, steps=30):
    rng = np.random.default_rng(order * 10 + stepper)
    nt = 6 * n ** 3
    er = rng.uniform(1.0, 3.0, nt)
    sg = rng.uniform(0.0, 0.3, nt) if lossy else None
    dt = 0.02 if order == 1 else 0.01
    def make(e, s):
        o = ot.OracleTet((n, n, n), 1.0 / n, dt, 1.0, 1.0, tf=2, stepper=stepper, eps_r=e, sigma=s, order=order)
        o.add_dipole([0.41, 0.57, 0.36], [0.3, 0.5, 0.8], f0=3.0, t0=0.3)
        o.add_receiver([0.62, 0.31, 0.55])
        o.add_receiver([0.27, 0.71, 0.66])
        return o
    o = make(er, sg)
    d = np.zeros((2, steps // 2 + 1, 3))
    d[:, 1:, :] = 0.01 * rng.standard_normal((2, steps // 2, 3))
    return make, er, sg, d, steps


@pytest.mark.parametrize("order,stepper,lossy", [(1, ot.SYMPLECTIC2, False), (1, ot.SYMPLECTIC4, True),
                                                 (2, ot.SYMPLECTIC4, False)])
def test_adjoint_gradient_matches_central_differences(order, stepper, lossy):
    make, er, sg, d, steps = _grad_case(order, stepper, lossy)
    J, ge, gs, _ = ot.misfit_gradient(make(er, sg), steps, d)
    rng = np.random.default_rng(4)
    sg0 = sg if sg is not None else np.zeros_like(er)
    for which in ("eps", "sigma"):
        v = rng.standard_normal(er.size)
        h = 1e-4 if which == "eps" else 1e-4
        def Jat(a):
            e, s = (er + a * v, sg0) if which == "eps" else (er, sg0 + a * v)
            return ot.misfit_gradient(make(e, s), steps, d)[0]
        fd = (Jat(h) - Jat(-h)) / (2 * h)
        g = (ge if which == "eps" else gs) @ v
        assert abs(fd - g) <= 1e-6 * max(abs(fd), abs(g)) + 1e-14, (which, fd, g)


def test_whitney_stiffness_is_a_product_of_element_curls():
    """First order: curl w_ab = 2 grad l_a x grad l_b is constant on a
    tetrahedron, so the quadrature curl-curl matrix equals V C^T C (the
    factorized form the device applies, curl then transposed curl).  Pinned
    for all six Kuhn types against the oracle's quadrature matrices, and C
    annihilates the gradients of the vertex functions (the null space)."""
    from oracle import tet as ot
    h = 0.37
    for t in range(6):
        G, V = ot.bary_grad(ot.tet_vertices(t) * h)
        C = np.stack([2.0 * np.cross(G[a], G[b]) for a, b in ot.LOC_EDGE], axis=1)  # 3 x 6
        M, S = ot.element_matrices(t, h, 1)
        np.testing.assert_allclose(V * C.T @ C, S, rtol=1e-12, atol=1e-12 * np.abs(S).max())
        # gradient of vertex function v as edge coefficients: (l_b - l_a)(grad) -> +-1
        for v in range(4):
            g = np.array([(1.0 if b == v else 0.0) - (1.0 if a == v else 0.0) for a, b in ot.LOC_EDGE])
            assert np.abs(C @ g).max() < 1e-12 * np.abs(C).max()
