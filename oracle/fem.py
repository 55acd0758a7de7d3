"""ctypes binding of the FP64 FEM oracle (oracle/fem_oracle.c) and a kmat
CSV reader.

TEST INFRASTRUCTURE ONLY.  Mirrors the reference class ``FemSimulation`` of
src/fem_v2 (and src/symplectic_2, src/symplectic_4) with one rank owning the
whole domain.
"""
import ctypes
import os

import numpy as np

from . import fdtd as _fdtd

LEAPFROG, SYMPLECTIC2, SYMPLECTIC4 = 0, 1, 2


class FemCfg(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("h", ctypes.c_double), ("dt", ctypes.c_double), ("eps", ctypes.c_double),
                ("mu", ctypes.c_double), ("tf", ctypes.c_int), ("stepper", ctypes.c_int)]


_sig_done = False


def lib():
    global _sig_done
    L = _fdtd.lib()
    if not _sig_done:
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.ems_oracle_femv2_create.restype = P
        L.ems_oracle_femv2_create.argtypes = [ctypes.POINTER(FemCfg), dp]
        L.ems_oracle_femv2_destroy.argtypes = [P]
        L.ems_oracle_femv2_add_source_z.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ems_oracle_femv2_add_observation.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ems_oracle_femv2_run.argtypes = [P, ctypes.c_int]
        L.ems_oracle_femv2_stage.argtypes = [P, ctypes.c_double, ctypes.c_double]
        L.ems_oracle_femv2_nslot.argtypes = [P]
        L.ems_oracle_femv2_nobs.argtypes = [P]
        L.ems_oracle_femv2_time.argtypes = [P]
        L.ems_oracle_femv2_time.restype = ctypes.c_double
        L.ems_oracle_femv2_get_traces.argtypes = [P, dp]
        L.ems_oracle_femv2_size.argtypes = [P, ctypes.c_int]
        L.ems_oracle_femv2_size.restype = ctypes.c_size_t
        L.ems_oracle_femv2_get.argtypes = [P, ctypes.c_int, ctypes.c_int, dp]
        L.ems_oracle_femv2_set.argtypes = [P, ctypes.c_int, ctypes.c_int, dp]
        L.ems_oracle_femv2_apply.argtypes = [P, dp, dp]
        L.ems_oracle_femv2_matrix.restype = dp
        L.ems_oracle_femv2_matrix.argtypes = [P]
        L.ems_oracle_fem_source.restype = ctypes.c_double
        L.ems_oracle_fem_source.argtypes = [ctypes.c_double]
        _sig_done = True
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def parse_value(cell):
    """parseValue of src/fem_v2/fem_simulation.cpp:1477-1493: trim blanks and
    quotes, "a/b" -> a / b, else a plain number."""
    t = cell.strip(" \t\r\n\"")
    if not t:
        return 0.0
    if "/" in t:
        num, den = t.split("/", 1)
        return float(num) / float(den)
    return float(t)


def load_kmat(path):
    """loadMatrixFromCSV (src/fem_v2/fem_simulation.cpp:1495-1524)."""
    rows = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            row = [parse_value(c) for c in line.split(",")] if line else []
            if row:
                rows.append(row)
    return np.array(rows, dtype=np.float64)


def kmat_path(order, use_ofem=1, root=None):
    root = root or os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                                "golden", "kmat")
    name = ("%d.csv" % order) if use_ofem else ("lamped%d.csv" % order)
    return os.path.join(root, name)


class OracleFEMv2:
    """fem_v2 / symplectic_2 / symplectic_4 on one rank (dims = 1 1 1)."""

    def __init__(self, order, grid, h, dt, eps, mu, tf=1, stepper=LEAPFROG, use_ofem=1, kraw=None):
        if kraw is None:
            kraw = load_kmat(kmat_path(order, use_ofem))
        kraw = np.ascontiguousarray(kraw, dtype=np.float64)
        self.order = order
        self.grid = tuple(int(g) for g in grid)
        self.cfg = FemCfg(order, self.grid[0], self.grid[1], self.grid[2], h, dt, eps, mu, tf, stepper)
        self._kraw = kraw
        self._h = lib().ems_oracle_femv2_create(ctypes.byref(self.cfg), _dp(kraw))
        if not self._h:
            raise ValueError("oracle: invalid FEM configuration")
        self.n = kraw.shape[0]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ems_oracle_femv2_destroy(self._h)
            self._h = None

    def setSource_z(self, pos):
        lib().ems_oracle_femv2_add_source_z(self._h, *[int(p) for p in pos])

    def setObservationPoint(self, pos):
        lib().ems_oracle_femv2_add_observation(self._h, *[int(p) for p in pos])

    def run(self, num_steps):
        lib().ems_oracle_femv2_run(self._h, int(num_steps))

    def stage(self, deltat, offset=0.0):
        lib().ems_oracle_femv2_stage(self._h, float(deltat), float(offset))

    @property
    def current_time(self):
        return lib().ems_oracle_femv2_time(self._h)

    def traces(self):
        nobs = lib().ems_oracle_femv2_nobs(self._h)
        nslot = lib().ems_oracle_femv2_nslot(self._h)
        out = np.zeros((nobs, nslot, 4))
        if out.size:
            lib().ems_oracle_femv2_get_traces(self._h, _dp(out))
        return out

    def size(self, comp):
        return lib().ems_oracle_femv2_size(self._h, comp)

    def get(self, comp, slot):
        out = np.empty(self.size(comp))
        lib().ems_oracle_femv2_get(self._h, comp, slot, _dp(out))
        return out

    def set(self, comp, slot, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.size(comp))
        lib().ems_oracle_femv2_set(self._h, comp, slot, _dp(v))

    def get_state(self, slot):
        return np.concatenate([self.get(c, slot) for c in range(3)])

    def set_state(self, slot, v):
        n0, n1 = self.size(0), self.size(1)
        self.set(0, slot, v[:n0])
        self.set(1, slot, v[n0:n0 + n1])
        self.set(2, slot, v[n0 + n1:])

    def apply(self, x):
        """Assembled y = K x (scaled element matrices, all elements)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib().ems_oracle_femv2_apply(self._h, _dp(x), _dp(y))
        return y

    def ndof(self):
        return self.size(0) + self.size(1) + self.size(2)


def fem_source(t):
    return lib().ems_oracle_fem_source(t)


def boundary_mask(grid):
    """True for the DOFs the PEC condition zeroes (tangential on the faces),
    ordered [Ex | Ey | Ez] in the reference layouts."""
    Nx, Ny, Nz = grid
    mx = np.zeros((Nz + 1, Ny + 1, Nx), bool)
    mx[:, 0, :] = mx[:, Ny, :] = True
    mx[0, :, :] = mx[Nz, :, :] = True
    my = np.zeros((Nz + 1, Ny, Nx + 1), bool)
    my[:, :, 0] = my[:, :, Nx] = True
    my[0, :, :] = my[Nz, :, :] = True
    mz = np.zeros((Nz, Ny + 1, Nx + 1), bool)
    mz[:, :, 0] = mz[:, :, Nx] = True
    mz[:, 0, :] = mz[:, Ny, :] = True
    return np.concatenate([mx.ravel(), my.ravel(), mz.ravel()])
