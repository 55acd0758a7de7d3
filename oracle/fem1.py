"""ctypes binding of the FP64 oracle of the first-order stabilized hex solvers
(oracle/fem1_oracle.c) and the exact element matrix.

TEST INFRASTRUCTURE ONLY.  Mirrors the reference classes ``FemSimulation`` of
src/fem (leapfrog, stabilized K = C_s B_s^-1 A_s) and src/fem_abm
(Adams-Bashforth-Moulton on co-located E/H).

The element matrix follows src/fem/fem_simulation.cpp:125-434 and
src/fem_abm/fem_simulation.cpp:143-246: the stabilized tables
X_s = X + s * _X are formed entrywise in double (as the reference does), then
  fem:      K = c * C_s * B_s^-1 * A_s,  c = 0.5*dt*dt/mu/eps/h/h (left to right)
  fem_abm:  K = A_s
For fem the product is evaluated exactly in rationals and rounded once
(the reference evaluates it with Eigen's blocked double GEMM and an LU
inverse; that rounding is not reproducible without Eigen, so K agrees with the
reference to within its own rounding, O(1e-16) relative).
"""
import ctypes
import json
import os
from fractions import Fraction

import numpy as np

from . import fdtd as _fdtd

FEM, ABM = 0, 1
TABLES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden",
                      "fem1_tables.json")


class Fem1Cfg(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("h", ctypes.c_double), ("dt", ctypes.c_double), ("eps", ctypes.c_double),
                ("mu", ctypes.c_double), ("tf", ctypes.c_int), ("stepper", ctypes.c_int),
                ("stab", ctypes.c_double), ("use_ofem", ctypes.c_int)]


_sig_done = False


def lib():
    global _sig_done
    L = _fdtd.lib()
    if not _sig_done:
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        i = ctypes.c_int
        L.ems_oracle_fem1_create.restype = P
        L.ems_oracle_fem1_create.argtypes = [ctypes.POINTER(Fem1Cfg), dp]
        L.ems_oracle_fem1_destroy.argtypes = [P]
        L.ems_oracle_fem1_add_source.argtypes = [P, i, i, i, i]
        L.ems_oracle_fem1_add_dipole_x.argtypes = [P, i, i, i]
        L.ems_oracle_fem1_add_observation.argtypes = [P, ip, ip]
        L.ems_oracle_fem1_run.argtypes = [P, i]
        L.ems_oracle_fem1_apply.argtypes = [P, dp, dp, i, i]
        L.ems_oracle_fem1_size.argtypes = [P, i]
        L.ems_oracle_fem1_size.restype = ctypes.c_long
        L.ems_oracle_fem1_nslot.argtypes = [P]
        L.ems_oracle_fem1_nobs.argtypes = [P]
        L.ems_oracle_fem1_time.argtypes = [P]
        L.ems_oracle_fem1_time.restype = ctypes.c_double
        L.ems_oracle_fem1_get_traces.argtypes = [P, dp]
        L.ems_oracle_fem1_get_field.argtypes = [P, i, i, dp]
        L.ems_oracle_fem1_set_field.argtypes = [P, i, i, dp]
        L.ems_oracle_fem1_source.restype = ctypes.c_double
        L.ems_oracle_fem1_source.argtypes = [i, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _sig_done = True
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def tables():
    """{'fem_ofem0': {'Amat': 12x12 Fractions, ...}, 'fem_ofem1': ..., 'fem_abm': ...}"""
    js = json.load(open(TABLES))
    return {k: {n: [[Fraction(x) for x in row] for row in t] for n, t in v.items()} for k, v in js.items()}


def stabilized(tab, name, s):
    """X_s = X + s * _X, entrywise in double (Amat_Eigen + s * _Amat_Eigen)."""
    a, b = tab[name], tab["_" + name]
    return [[float(a[r][c]) + s * float(b[r][c]) for c in range(12)] for r in range(12)]


def _solve_exact(B, A):
    """B^-1 A over the rationals (Gauss-Jordan, exact)."""
    n = len(B)
    M = [[Fraction(B[r][c]) for c in range(n)] + [Fraction(A[r][c]) for c in range(n)] for r in range(n)]
    for col in range(n):
        piv = next(r for r in range(col, n) if M[r][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        pv = M[col][col]
        M[col] = [x / pv for x in M[col]]
        for r in range(n):
            if r != col and M[r][col] != 0:
                f = M[r][col]
                M[r] = [x - f * y for x, y in zip(M[r], M[col])]
    return [row[n:] for row in M]


def element_matrix(stepper, dt, h, eps, mu, stab, use_ofem=0, exact=False):
    """12x12 element matrix as used by the step (row-major float64 array);
    exact=True returns the rational matrix before the final rounding."""
    T = tables()
    if stepper == ABM:
        return np.array(stabilized(T["fem_abm"], "Amat", stab), dtype=np.float64)
    tab = T["fem_ofem1" if use_ofem else "fem_ofem0"]
    As, Bs, Cs = (stabilized(tab, n, stab) for n in ("Amat", "Bmat", "Cmat"))
    X = _solve_exact(Bs, As)
    M = [[sum(Fraction(Cs[r][k]) * X[k][c] for k in range(12)) for c in range(12)] for r in range(12)]
    c = 0.5 * dt * dt / mu / eps / h / h
    K = [[Fraction(c) * M[r][q] for q in range(12)] for r in range(12)]
    if exact:
        return K
    return np.array([[float(x) for x in row] for row in K], dtype=np.float64)


def fem_source(stepper, t, mu, h):
    return lib().ems_oracle_fem1_source(stepper, t, mu, h)


class OracleFem1:
    """FP64 oracle of src/fem (stepper=FEM) and src/fem_abm (stepper=ABM)."""

    def __init__(self, grid, h, dt, eps, mu, tf, stepper=FEM, stab=0.3, use_ofem=0, K=None):
        self.grid = tuple(int(g) for g in grid)
        self.stepper = stepper
        if K is None:
            K = element_matrix(stepper, dt, h, eps, mu, stab, use_ofem)
        self.K = np.ascontiguousarray(K, dtype=np.float64).reshape(12, 12)
        cfg = Fem1Cfg(*self.grid, h, dt, eps, mu, tf, stepper, stab, use_ofem)
        self._L = lib()
        self._h = self._L.ems_oracle_fem1_create(ctypes.byref(cfg), _dp(self.K))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.ems_oracle_fem1_destroy(h)
            self._h = None

    def setSource(self, comp, pos):
        self._L.ems_oracle_fem1_add_source(self._h, comp, *[int(v) for v in pos])

    def setSource_x(self, pos):
        self.setSource(0, pos)

    def setSource_y(self, pos):
        self.setSource(1, pos)

    def setSource_z(self, pos):
        self.setSource(2, pos)

    def setSourceDipole_x(self, pos):
        self._L.ems_oracle_fem1_add_dipole_x(self._h, *[int(v) for v in pos])

    def setObservationPoint(self, lists):
        counts = (ctypes.c_int * 3)(*[len(l) for l in lists])
        flat = [int(v) for l in lists for p in l for v in p] or [0]
        pts = (ctypes.c_int * len(flat))(*flat)
        self._L.ems_oracle_fem1_add_observation(self._h, counts, pts)

    def run(self, n):
        self._L.ems_oracle_fem1_run(self._h, int(n))

    def size(self, comp):
        return self._L.ems_oracle_fem1_size(self._h, comp % 3)

    def traces(self):
        no, ns = self._L.ems_oracle_fem1_nobs(self._h), self._L.ems_oracle_fem1_nslot(self._h)
        out = np.zeros(no * ns * 3 + 1)
        self._L.ems_oracle_fem1_get_traces(self._h, _dp(out))
        return out[:no * ns * 3].reshape(no, ns, 3)

    def field(self, f, slot=1):
        out = np.zeros(self.size(f))
        self._L.ems_oracle_fem1_get_field(self._h, f, slot, _dp(out))
        return out

    def set_field(self, f, slot, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.size(f)
        self._L.ems_oracle_fem1_set_field(self._h, f, slot, _dp(v))

    def apply(self, x, transpose=False, subtract=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self._L.ems_oracle_fem1_apply(self._h, _dp(x), _dp(y), int(transpose), int(subtract))
        return y

    @property
    def time(self):
        return self._L.ems_oracle_fem1_time(self._h)
