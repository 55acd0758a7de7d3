"""FP64 oracle of the tetrahedral Nedelec edge-element solver (numpy / scipy).

TEST INFRASTRUCTURE ONLY.  The solver has no reference counterpart; it
applies the time steppers of src/symplectic_2/fem_simulation.cpp:557-672
(symplectic stages), src/symplectic_4 (the x1, x2, x1 composition) and
src/fem_abm/fem_simulation.cpp:278-833 (AB2 predictor, AM3 corrector) to the
edge-element semi-discretisation on tetrahedra

    M_eps U'' + M_sig U' + S U = -g(t),  g_i = p'(t) (w_i(r0) . d).

This restatement is independent of the library's host code: the element
matrices come from quadrature of the basis functions and of their curls
written out by hand (the library integrates barycentric monomials exactly),
barycentric coordinates from a 4 x 4 inverse, assembly from scipy.sparse and
the mass solves from a sparse LU factorisation (the library uses conjugate
gradients on the device).  Order 1: the 6 Whitney functions; order 2 (first
kind, degree 2, 20 functions): the Whitney functions, the edge gradients
grad(l_a l_b) and per face (a, b, c) the functions l_a w_bc, l_b w_ac.

Pins (tests/test_oracle_tet.py):
  * null space: S G = 0 for the vertex-to-edge incidence G (a discrete
    gradient is curl-free), and dim ker S = number of interior vertices;
  * PEC cube cavity: the lowest nonzero eigenvalues of S x = w^2 M x converge
    at second order in h to (pi / L)^2 (l^2 + m^2 + n^2) with the analytic
    multiplicities (3 for (1,1,0), 2 for (1,1,1));
  * element matrices equal the closed forms used by the library (Whitney
    mass V (1 + d_ij) / 20 moments, curl-curl 4 V (g_a x g_b).(g_c x g_d));
  * interpolation: receivers reproduce a constant field exactly;
  * steppers: on a discrete eigenmode the exact solution is cos(w t)
    (damped for uniform sigma): symplectic_2 and ABM converge at second
    order, symplectic_4 at fourth;
  * order 2: element curl-curl rank 11 (its kernel grad P2 has dimension
    9), dim ker S = interior vertices + interior edges, the cavity modes
    converge at fourth order in h, and L2 projection reproduces linear fields.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

SYMPLECTIC2, SYMPLECTIC4, ABM = 0, 1, 2

# edge families (directions) and their order
FAM = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]])
# Kuhn tetrahedron t of a cube: the path 0 -> e_p0 -> e_p0 + e_p1 -> (1,1,1)
PERM = np.array([[0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]])
LOC_EDGE = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]

# face families: (edge family of q - p, edge family of r - p) of a face p < q < r
FFAM = [(0, 3), (0, 4), (0, 6), (1, 3), (1, 5), (1, 6), (2, 4), (2, 5), (2, 6), (3, 6), (4, 6), (5, 6)]
LOC_TRI = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]


def _collapsed_rule(n):
    """Collapsed-coordinate Gauss rule on the reference tetrahedron: barycentric
    points (npts x 4) and weights summing to 1 (fraction of the volume); exact
    for polynomials of degree <= 2n - 3."""
    x, w = np.polynomial.legendre.leggauss(n)
    x, w = 0.5 * (x + 1), 0.5 * w
    pts, wts = [], []
    for a, wa in zip(x, w):
        for b, wb in zip(x, w):
            for c, wc in zip(x, w):
                X, Y, Z = a, (1 - a) * b, (1 - a) * (1 - b) * c
                pts.append([1 - X - Y - Z, X, Y, Z])
                wts.append(6.0 * wa * wb * wc * (1 - a) ** 2 * (1 - b))
    return np.array(pts), np.array(wts)


QPTS, QW = _collapsed_rule(5)


def tet_vertices(t):
    """Vertex offsets (4 x 3, in units of h) of Kuhn tetrahedron t."""
    v = np.zeros((4, 3), dtype=int)
    v[1, PERM[t][0]] = 1
    v[2, PERM[t][0]] = 1
    v[2, PERM[t][1]] = 1
    v[3] = 1
    return v


def bary_grad(X):
    """Barycentric gradients (4 x 3) and volume of the tetrahedron with vertices X (4 x 3)."""
    A = np.hstack([np.ones((4, 1)), X])  # lambda_i(x) = C[0, i] + C[1:, i] . x
    C = np.linalg.inv(A)
    return C[1:, :].T.copy(), abs(np.linalg.det(A)) / 6.0


def basis_values(lam, G, order=1):
    """(nl x 3) basis values and curls at barycentric point lam."""
    W, C = [], []
    for a, b in LOC_EDGE:
        W.append(lam[a] * G[b] - lam[b] * G[a])
        C.append(2.0 * np.cross(G[a], G[b]))
    if order == 2:
        for a, b in LOC_EDGE:
            W.append(lam[a] * G[b] + lam[b] * G[a])
            C.append(np.zeros(3))
        for a, b, c in LOC_TRI:
            # l_a w_bc: curl = grad l_a x w_bc + l_a curl w_bc
            wbc = lam[b] * G[c] - lam[c] * G[b]
            W.append(lam[a] * wbc)
            C.append(np.cross(G[a], wbc) + 2.0 * lam[a] * np.cross(G[b], G[c]))
            wac = lam[a] * G[c] - lam[c] * G[a]
            W.append(lam[b] * wac)
            C.append(np.cross(G[b], wac) + 2.0 * lam[b] * np.cross(G[a], G[c]))
    return np.array(W), np.array(C)


def element_matrices(t, h, order=1):
    """(M, S) by quadrature: M_lm = int w_l . w_m, S_lm = int curl w_l . curl w_m."""
    X = tet_vertices(t) * float(h)
    G, V = bary_grad(X)
    nl = 6 if order == 1 else 20
    M = np.zeros((nl, nl))
    S = np.zeros((nl, nl))
    for lam, wq in zip(QPTS, QW):
        W, C = basis_values(lam, G, order)
        M += wq * V * (W @ W.T)
        S += wq * V * (C @ C.T)
    return M, S


class Mesh:
    """Kuhn split of an nx x ny x nz grid of cubes of edge h; edge numbering of
    family f at base vertex (I, J, K): off_f + I + cx (J + cy K)."""

    def __init__(self, nx, ny, nz, h, order=1):
        self.n = np.array([nx, ny, nz])
        self.h = float(h)
        self.order = int(order)
        self.cnt = [(nx + 1 - d[0]) * (ny + 1 - d[1]) * (nz + 1 - d[2]) for d in FAM]
        self.off = np.concatenate([[0], np.cumsum(self.cnt)])
        self.ne = int(self.off[-1])
        self.fcnt = [(nx + 1 - FAM[f2][0]) * (ny + 1 - FAM[f2][1]) * (nz + 1 - FAM[f2][2]) for _, f2 in FFAM]
        self.foff = np.concatenate([[0], np.cumsum(self.fcnt)])
        self.nf = int(self.foff[-1])
        self.nt = 6 * nx * ny * nz
        self.nl = 6 if order == 1 else 20
        self.ndof = self.ne if order == 1 else 2 * self.ne + 2 * self.nf

    def fam_of(self, d):
        for f in range(7):
            if tuple(FAM[f]) == tuple(d):
                return f
        raise ValueError(d)

    def edge(self, f, I, J, K):
        cx = self.n[0] + 1 - FAM[f][0]
        cy = self.n[1] + 1 - FAM[f][1]
        return self.off[f] + I + cx * (J + cy * K)

    def face(self, q, I, J, K):
        d = FAM[FFAM[q][1]]
        cx = self.n[0] + 1 - d[0]
        cy = self.n[1] + 1 - d[1]
        return self.foff[q] + I + cx * (J + cy * K)

    def tet_faces(self):
        """[nt][4] global faces in local face order."""
        nx, ny, nz = self.n
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        out = np.zeros((nx * ny * nz, 6, 4), dtype=np.int64)
        for t in range(6):
            v = tet_vertices(t)
            for q, (a, b, c) in enumerate(LOC_TRI):
                ff = FFAM.index((self.fam_of(v[b] - v[a]), self.fam_of(v[c] - v[a])))
                out[:, t, q] = self.face(ff, i + v[a][0], j + v[a][1], k + v[a][2])
        return out.reshape(-1, 4)

    def tet_dofs(self):
        """[nt][nl] global dofs in the local basis order."""
        E = self.tet_edges()
        if self.order == 1:
            return E
        F = self.tet_faces()
        Fd = np.stack([2 * self.ne + 2 * F, 2 * self.ne + 2 * F + 1], axis=2).reshape(-1, 8)
        return np.hstack([E, self.ne + E, Fd])

    def dof_boundary(self):
        be = self.boundary()
        if self.order == 1:
            return be
        # a face is on the boundary when its three vertices lie in one outer plane
        nx, ny, nz = self.n
        bf = np.zeros(self.nf, dtype=bool)
        for q, (f1, f2) in enumerate(FFAM):
            d1, d2 = FAM[f1], FAM[f2]
            cx, cy, cz = self.n[0] + 1 - d2[0], self.n[1] + 1 - d2[1], self.n[2] + 1 - d2[2]
            K, J, I = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
            P = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
            flag = np.zeros(P.shape[0], dtype=bool)
            for a in range(3):
                if d1[a] == 0 and d2[a] == 0:
                    flag |= (P[:, a] == 0) | (P[:, a] == self.n[a])
            bf[self.foff[q]:self.foff[q + 1]] = flag
        return np.concatenate([be, be, np.repeat(bf, 2)])

    def tet_edges(self):
        """[nt][6] global edges, tetrahedron id 6 (i + nx (j + ny k)) + t."""
        nx, ny, nz = self.n
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        out = np.zeros((nx * ny * nz, 6, 6), dtype=np.int64)
        for t in range(6):
            v = tet_vertices(t)
            for l, (a, b) in enumerate(LOC_EDGE):
                f = self.fam_of(v[b] - v[a])
                out[:, t, l] = self.edge(f, i + v[a][0], j + v[a][1], k + v[a][2])
        return out.reshape(-1, 6)

    def edge_geometry(self):
        """start point and vector (physical) of every edge."""
        P = np.zeros((self.ne, 3))
        D = np.zeros((self.ne, 3))
        for f in range(7):
            cx = self.n[0] + 1 - FAM[f][0]
            cy = self.n[1] + 1 - FAM[f][1]
            cz = self.n[2] + 1 - FAM[f][2]
            K, J, I = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
            sl = slice(self.off[f], self.off[f + 1])
            P[sl] = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1) * self.h
            D[sl] = FAM[f] * self.h
        return P, D

    def boundary(self):
        """edges lying in an outer face (tangential E = 0 there)."""
        P, D = self.edge_geometry()
        L = self.n * self.h
        out = np.zeros(self.ne, dtype=bool)
        for a in range(3):
            flat = D[:, a] == 0
            out |= flat & ((np.abs(P[:, a]) < 1e-9 * self.h) | (np.abs(P[:, a] - L[a]) < 1e-9 * self.h))
        return out

    def vertex_incidence(self):
        """G (ne x nv): edge = end vertex minus start vertex, vertex (I,J,K) = I + (nx+1)(J + (ny+1) K)."""
        P, D = self.edge_geometry()
        n1 = self.n + 1
        a = np.rint(P / self.h).astype(int)
        b = np.rint((P + D) / self.h).astype(int)
        va = a[:, 0] + n1[0] * (a[:, 1] + n1[1] * a[:, 2])
        vb = b[:, 0] + n1[0] * (b[:, 1] + n1[1] * b[:, 2])
        r = np.arange(self.ne)
        nv = int(np.prod(n1))
        return sp.csr_matrix((np.r_[np.ones(self.ne), -np.ones(self.ne)], (np.r_[r, r], np.r_[vb, va])),
                             shape=(self.ne, nv))

    def locate(self, x):
        """(cube (i,j,k), tet type, barycentric lambda) of point x."""
        u = np.asarray(x, dtype=float) / self.h
        c = np.minimum(np.floor(u).astype(int), self.n - 1)
        if np.any(u < 0) or np.any(u > self.n):
            raise ValueError("point outside the mesh")
        for t in range(6):
            X = (tet_vertices(t) + c) * self.h
            A = np.hstack([np.ones((4, 1)), X])
            lam = np.linalg.solve(A.T, np.r_[1.0, x])
            if np.all(lam >= -1e-12):
                return c, t, lam
        raise RuntimeError("no tetrahedron contains the point")

    def basis_at(self, x):
        """(global dofs [nl], values [nl][3]) of the basis functions at point x."""
        c, t, lam = self.locate(x)
        X = (tet_vertices(t) + c) * self.h
        G, _ = bary_grad(X)
        i, j, k = c
        tid = 6 * (i + self.n[0] * (j + self.n[1] * k)) + t
        e = self.tet_dofs()[tid]
        W, _ = basis_values(lam, G, self.order)
        return e, W


def assemble(mesh, eps, mu, eps_r=None, sigma=None):
    """S, M_eps, M_sig (csr) with PEC: boundary rows/columns removed, identity
    rows in the mass-type matrices, zero rows in S."""
    E = mesh.tet_dofs()
    nt, nl = E.shape
    er = np.ones(nt) if eps_r is None else np.asarray(eps_r, dtype=float)
    sg = np.zeros(nt) if sigma is None else np.asarray(sigma, dtype=float)
    Ms_loc = [element_matrices(t, mesh.h, mesh.order) for t in range(6)]
    typ = np.arange(nt) % 6
    Ml = np.stack([Ms_loc[t][0] for t in typ])  # [nt][nl][nl]
    Sl = np.stack([Ms_loc[t][1] for t in typ])
    rows = np.repeat(E, nl, axis=1).ravel()
    cols = np.tile(E, (1, nl)).ravel()
    bnd = mesh.dof_boundary()
    keep = ~bnd[rows] & ~bnd[cols]
    r, c = rows[keep], cols[keep]
    n = mesh.ndof

    def build(vals, diag):
        A = sp.coo_matrix((vals.ravel()[keep], (r, c)), shape=(n, n)).tocsr()
        return (A + sp.diags(np.where(bnd, diag, 0.0))).tocsr()

    S = build(Sl / mu, 0.0)
    Me = build(eps * er[:, None, None] * Ml, 1.0)
    Ms = build(sg[:, None, None] * Ml, 0.0)
    return S, Me, Ms


def pulse_rate(t, f0, t0, amp):
    """p'(t) of p(t) = amp exp(-(pi f0 (t - t0))^2)."""
    a = np.pi * f0 * (t - t0)
    return amp * (-2.0 * np.pi * f0 * a) * np.exp(-a * a)


class OracleTet:
    def __init__(self, grid, h, dt, eps, mu, tf=1, stepper=SYMPLECTIC2, eps_r=None, sigma=None, order=1):
        self.mesh = Mesh(*grid, h, order=order)
        self.dt, self.tf, self.stepper = float(dt), int(tf), int(stepper)
        self.S, self.Me, self.Ms = assemble(self.mesh, eps, mu, eps_r, sigma)
        self.Me_eps0 = float(eps)
        self.bnd = self.mesh.dof_boundary()
        n = self.mesh.ndof
        self.U = np.zeros(n)
        self.V = np.zeros(n)
        self.t = 0.0
        self.steps = 0
        self.dips = []  # (dofs, weights, f0, t0, amp)
        self.recs = []  # (dofs, values)
        self.samples = []  # per receiver list of slots
        self._lu = {}
        self._hist = None

    def add_dipole(self, pos, direction, f0, t0, amp=1.0):
        e, W = self.mesh.basis_at(pos)
        w = W @ np.asarray(direction, dtype=float)
        w[self.bnd[e]] = 0.0
        self.dips.append((e, w, float(f0), float(t0), float(amp)))

    def add_receiver(self, pos):
        self.recs.append(self.mesh.basis_at(pos))

    def load(self, t):
        g = np.zeros(self.mesh.ndof)
        for e, w, f0, t0, amp in self.dips:
            np.add.at(g, e, pulse_rate(t, f0, t0, amp) * w)
        return g

    def receive(self, U=None):
        U = self.U if U is None else U
        return np.array([U[e] @ W for e, W in self.recs]).reshape(-1, 3)

    def _solve(self, D, b):
        key = round(D, 300) if D else 0.0
        if key not in self._lu:
            self._lu[key] = spla.splu((self.Me + 0.5 * D * self.Ms).tocsc())
        return self._lu[key].solve(b)

    def _stage(self, D, t_mid):
        self.U = self.U + 0.5 * D * self.V
        b = (self.Me @ self.V - 0.5 * D * (self.Ms @ self.V)) - D * (self.S @ self.U + self.load(t_mid))
        self.V = self._solve(D, b)
        self.U = self.U + 0.5 * D * self.V

    def _accel(self, U, V, t):
        return self._solve(0.0, -(self.S @ U) - self.Ms @ V - self.load(t))

    def step(self):
        dt, t = self.dt, self.t
        if self.stepper == SYMPLECTIC2:
            self._stage(dt, t + 0.5 * dt)
        elif self.stepper == SYMPLECTIC4:
            x1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
            x2 = 1.0 - 2.0 * x1
            o = 0.0
            for x in (x1, x2, x1):
                self._stage(x * dt, t + o + 0.5 * x * dt)
                o += x * dt
        else:
            if self._hist is None:
                a0 = self._accel(self.U, self.V, t)
                self._hist = (self.V.copy(), a0, a0.copy())  # (V_n-1, a_n-1, a_n)
            Vm, am, an = self._hist
            Up = self.U + 0.5 * dt * (3 * self.V - Vm)
            Vp = self.V + 0.5 * dt * (3 * an - am)
            ap = self._accel(Up, Vp, t + dt)
            U1 = self.U + dt / 12.0 * (5 * Vp + 8 * self.V - Vm)
            V1 = self.V + dt / 12.0 * (5 * ap + 8 * an - am)
            a1 = self._accel(U1, V1, t + dt)
            self._hist = (self.V, an, a1)
            self.U, self.V = U1, V1
        self.t = t + dt
        self.steps += 1

    def run(self, nsteps):
        """Receiver traces [nobs][nslot][3] of this run (slot 0 = the state at entry)."""
        out = [self.receive()]
        for s in range(1, nsteps + 1):
            self.step()
            if s % self.tf == 0:
                out.append(self.receive())
        return np.stack(out, axis=1) if self.recs else np.zeros((0, len(out), 3))

    def set_field(self, which, values):
        if which == 0:
            self.U = np.array(values, dtype=float)
        else:
            self.V = np.array(values, dtype=float)
        self._hist = None


# ------------------------------------------------------------------ adjoint
def stage_schedule(stepper, dt):
    """(D, offset) of the symplectic stages of one step."""
    if stepper == SYMPLECTIC2:
        return [(dt, 0.0)]
    x1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
    x2 = 1.0 - 2.0 * x1
    return [(x1 * dt, 0.0), (x2 * dt, x1 * dt), (x1 * dt, (x1 + x2) * dt)]


def misfit_gradient(o, nsteps, d_obs):
    """Misfit J = 1/2 sum over slots 1..nsteps/tf of |trace - d_obs|^2 of a
    run from rest at t = 0, and dJ/deps_r, dJ/dsigma per tetrahedron, by the
    discrete adjoint of the symplectic stages with stored forward states:
        stage  U1 = U + D/2 V;  A V' = B V - D (S U1 + g);  U' = U1 + D/2 V'
               (A = Me + D/2 Ms, B = Me - D/2 Ms)
        adjoint mu = A^-1 (lV' + D/2 lU');  lU = lU' - D S mu;  lV = B mu + D/2 lU
        dJ/deps_r(T) += eps0 mu_T' M_T (V - V')_T,
        dJ/dsigma(T) += -D/2 mu_T' M_T (V + V')_T   (interior dofs only)."""
    if o.stepper == ABM:
        raise ValueError("the adjoint covers the symplectic steppers")
    m = o.mesh
    n = m.ndof
    o.U = np.zeros(n)
    o.V = np.zeros(n)
    o.t = 0.0
    sched = stage_schedule(o.stepper, o.dt)
    states = []  # per stage: (U_in, V_in, V_out, D)
    traces = [o.receive()]
    for s in range(1, nsteps + 1):
        t = o.t
        for D, off in sched:
            U_in, V_in = o.U.copy(), o.V.copy()
            o._stage(D, t + off + 0.5 * D)
            states.append((U_in, V_in, o.V.copy(), D))
        o.t = t + o.dt
        if s % o.tf == 0:
            traces.append(o.receive())
    tr = np.stack(traces, axis=1)
    res = tr - d_obs
    res[:, 0, :] = 0.0
    J = 0.5 * float(np.sum(res[:, 1:, :] ** 2))
    E = m.tet_dofs()
    typ = np.arange(m.nt) % 6
    Mloc = [element_matrices(t, m.h, m.order)[0] for t in range(6)]
    inner = (~o.bnd).astype(float)
    eps0 = o.Me_eps0
    ge = np.zeros(m.nt)
    gs = np.zeros(m.nt)
    lU = np.zeros(n)
    lV = np.zeros(n)
    q = len(states)
    for s in range(nsteps, 0, -1):
        if s % o.tf == 0:
            r = res[:, s // o.tf, :]
            for (dofs, W), rr in zip(o.recs, r):
                np.add.at(lU, dofs, W @ rr)
        for _ in sched:
            q -= 1
            U_in, V_in, V_out, D = states[q]
            mu = o._solve(D, lV + 0.5 * D * lU)
            mu_i = mu * inner
            dm = (V_in - V_out) * inner
            sm = (V_in + V_out) * inner
            for t in range(6):
                sel = typ == t
                Mu = mu_i[E[sel]]
                ge[sel] += eps0 * np.einsum("ti,ij,tj->t", Mu, Mloc[t], dm[E[sel]])
                gs[sel] += -0.5 * D * np.einsum("ti,ij,tj->t", Mu, Mloc[t], sm[E[sel]])
            lU1 = lU - D * (o.S @ mu)
            lV = (o.Me @ mu - 0.5 * D * (o.Ms @ mu)) + 0.5 * D * lU1
            lU = lU1
    return J, ge, gs, tr
