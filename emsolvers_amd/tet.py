"""Tetrahedral Nedelec (first kind, order 1 and 2) edge elements over the C ABI.

No reference counterpart: the structured-hex edge-FEM steppers of
src/symplectic_2, src/symplectic_4 (symplectic stages) and src/fem_abm
(AB2 / AM3) on the Kuhn 6-tetrahedron split of a cube grid, with the
consistent mass matrix solved by conjugate gradients on the GPU.  The
method names follow the reference solvers (run, observation traces), the
source is an electric dipole at an arbitrary point.
"""
import ctypes

import numpy as np

from . import _lib

SYMPLECTIC2, SYMPLECTIC4, ABM = 0, 1, 2
STEPPERS = {"symplectic_2": SYMPLECTIC2, "symplectic_4": SYMPLECTIC4, "abm": ABM}
# stability limits w_max dt < kappa of each stepper on y'' = -w^2 y
KAPPA = {SYMPLECTIC2: 2.0, SYMPLECTIC4: 1.5735, ABM: 1.2005}
# operators: assembled CSR, element-by-element (auto: element-by-element)
OPERATORS = {"auto": 0, "csr": 1, "ebe": 2, "ebe_mfma": 3, "ebe_rows": 4}


def element_matrices(order, tet_type, h):
    """(M, S) element matrices of Kuhn tetrahedron `tet_type` (host computation)."""
    nl = 6 if order == 1 else 20
    M = np.zeros((nl, nl))
    S = np.zeros((nl, nl))
    _lib.check(_lib.load().ems_tet_element_matrices(int(order), int(tet_type), float(h), _lib.dptr(M),
                                                    _lib.dptr(S)))
    return M, S


def assemble(grid, h, eps, mu, order=1, eps_r=None, sigma=None):
    """The library's assembled (S, M_eps, M_sig) as scipy CSR matrices (host computation)."""
    import scipy.sparse as sp
    g = [int(v) for v in grid]
    d = _lib.TetDesc(g[0], g[1], g[2], float(h), 1.0, float(eps), float(mu), 1, 0, int(order), 0)
    er = None if eps_r is None else np.ascontiguousarray(eps_r, dtype=np.float64)
    sg = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
    L = _lib.load()
    nnz = ctypes.c_size_t()
    args = (_lib.dptr(er) if er is not None else None, _lib.dptr(sg) if sg is not None else None)
    _lib.check(L.ems_tet_assemble(ctypes.byref(d), *args, None, None, None, None, None, 0, ctypes.byref(nnz)))
    n = nnz.value
    ne = sum((g[0] + 1 - a) * (g[1] + 1 - b) * (g[2] + 1 - c)
             for a, b, c in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)))
    nf = sum((g[0] + 1 - a) * (g[1] + 1 - b) * (g[2] + 1 - c)
             for a, b, c in ((1, 1, 0),) * 2 + ((1, 0, 1),) * 2 + ((0, 1, 1),) * 2 + ((1, 1, 1),) * 6)
    nd = ne if order == 1 else 2 * ne + 2 * nf
    rp = np.zeros(nd + 1, dtype=np.int32)
    col = np.zeros(n, dtype=np.int32)
    S, Me, Ms = np.zeros(n), np.zeros(n), np.zeros(n)
    _lib.check(L.ems_tet_assemble(ctypes.byref(d), *args, _lib.iptr(rp), _lib.iptr(col), _lib.dptr(S),
                                  _lib.dptr(Me), _lib.dptr(Ms), n, ctypes.byref(nnz)))
    mk = lambda v: sp.csr_matrix((v, col, rp), shape=(nd, nd))
    return mk(S), mk(Me), mk(Ms)


class TetSimulation:
    """grid = (nx, ny, nz) cubes of edge h, 6 tetrahedra per cube, PEC outer faces."""

    def __init__(self, grid, h, dt, eps, mu, tf=1, stepper="symplectic_2", order=1, precision="f64"):
        L = _lib.load()
        self.grid = tuple(int(g) for g in grid)
        st = STEPPERS[stepper] if isinstance(stepper, str) else int(stepper)
        prec = 1 if precision in ("f32", "fp32", "float32") else 0
        d = _lib.TetDesc(self.grid[0], self.grid[1], self.grid[2], float(h), float(dt), float(eps), float(mu),
                         int(tf), st, int(order), prec)
        hnd = ctypes.c_void_p()
        _lib.check(L.ems_tet_create(ctypes.byref(d), ctypes.byref(hnd)))
        self._h = hnd
        self.dt, self.tf, self.stepper = float(dt), int(tf), st

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ems_tet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_tets(self):
        return 6 * self.grid[0] * self.grid[1] * self.grid[2]

    def set_material(self, eps_r=None, sigma=None):
        """Per-tetrahedron relative permittivity / conductivity (6 nx ny nz values)."""
        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
            if a.size != self.num_tets:
                raise ValueError("expected %d values per tetrahedron" % self.num_tets)
            return a
        e, s = arr(eps_r), arr(sigma)
        _lib.check(_lib.load().ems_tet_set_material(self._h, _lib.dptr(e) if e is not None else None,
                                                    _lib.dptr(s) if s is not None else None))

    def add_dipole(self, pos, direction, f0, t0, amp=1.0):
        p = np.asarray(pos, dtype=np.float64)
        d = np.asarray(direction, dtype=np.float64)
        _lib.check(_lib.load().ems_tet_add_dipole(self._h, _lib.dptr(p), _lib.dptr(d), float(f0), float(t0),
                                                  float(amp)))

    def add_receiver(self, pos):
        p = np.asarray(pos, dtype=np.float64)
        _lib.check(_lib.load().ems_tet_add_receiver(self._h, _lib.dptr(p)))

    def set_operator(self, op):
        """'auto', 'csr', 'ebe', 'ebe_mfma' or 'ebe_rows' (at rest)."""
        _lib.check(_lib.load().ems_tet_set_operator(self._h, OPERATORS[op] if isinstance(op, str) else int(op)))

    def set_time_step(self, dt):
        """New dt before the first step (no reassembly)."""
        _lib.check(_lib.load().ems_tet_set_time_step(self._h, float(dt)))
        self.dt = float(dt)

    def run(self, num_steps):
        _lib.check(_lib.load().ems_tet_run(self._h, int(num_steps)))

    def run_timed(self, num_steps):
        ms = ctypes.c_double()
        _lib.check(_lib.load().ems_tet_run_timed(self._h, int(num_steps), ctypes.byref(ms)))
        return ms.value

    def gradient(self, num_steps, d_obs):
        """(J, dJ/deps_r, dJ/dsigma, drift) of a run from rest (symplectic
        steppers; lossless models by running the integrator backwards,
        lossy ones by checkpoints, drift 0); d_obs [nobs][num_steps/tf + 1][3]."""
        d = np.ascontiguousarray(np.asarray(d_obs, dtype=np.float64))
        J, dr = ctypes.c_double(), ctypes.c_double()
        ge = np.zeros(self.num_tets)
        gs = np.zeros(self.num_tets)
        _lib.check(_lib.load().ems_tet_gradient(self._h, int(num_steps), _lib.dptr(d), d.size, ctypes.byref(J),
                                                _lib.dptr(ge), _lib.dptr(gs), ctypes.byref(dr)))
        return J.value, ge, gs, dr.value

    def traces(self):
        """[nobs][nslot][3] receiver samples, slot q at step q tf."""
        no, ns = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().ems_tet_trace_shape(self._h, ctypes.byref(no), ctypes.byref(ns)))
        out = np.zeros((no.value, ns.value, 3))
        _lib.check(_lib.load().ems_tet_get_traces(self._h, _lib.dptr(out), out.size))
        return out

    def dof_count(self):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().ems_tet_dof_count(self._h, ctypes.byref(n)))
        return n.value

    def get_field(self, which=0):
        out = np.zeros(self.dof_count())
        _lib.check(_lib.load().ems_tet_get_field(self._h, int(which), _lib.dptr(out), out.size))
        return out

    def set_field(self, which, values):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
        _lib.check(_lib.load().ems_tet_set_field(self._h, int(which), _lib.dptr(v), v.size))

    def time(self):
        t = ctypes.c_double()
        _lib.check(_lib.load().ems_tet_get_time(self._h, ctypes.byref(t)))
        return t.value

    def max_frequency(self, iters=200):
        """w_max (rad/s) of M_eps^-1 S by power iteration (a lower bound)."""
        w2 = ctypes.c_double()
        _lib.check(_lib.load().ems_tet_max_frequency(self._h, int(iters), ctypes.byref(w2)))
        return float(np.sqrt(w2.value))

    def set_solver(self, solver):
        """'auto' (Chebyshev-accelerated Jacobi, CG fallback) or 'cg'."""
        _lib.check(_lib.load().ems_tet_set_solver(self._h, {"auto": 0, "cg": 1}[solver] if isinstance(solver, str)
                                                  else int(solver)))

    def solver_stats(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().ems_tet_solver_stats(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value
