"""Python mirror of the reference structured-hex FEM classes over the C ABI.

``FemSimulation`` keeps the constructor and method names of the reference
class in src/fem_v2/fem_simulation.hpp:10-60 (and the identical interface of
src/symplectic_2, src/symplectic_4).  The MPI topology arguments of the
reference are accepted and ignored: one GPU owns the whole domain.  The
element matrix is read from kmat/<order>.csv (use_ofem = 1) or
kmat/lamped<order>.csv (use_ofem = 0) relative to kmat_dir (default: the
current directory, as the reference).
"""
import ctypes
import os

import numpy as np

from . import _lib

LEAPFROG, SYMPLECTIC2, SYMPLECTIC4 = 0, 1, 2
STEPPERS = {"fem_v2": LEAPFROG, "leapfrog": LEAPFROG, "symplectic_2": SYMPLECTIC2,
            "symplectic_4": SYMPLECTIC4}


def load_kmat(path):
    """loadMatrixFromCSV (src/fem_v2/fem_simulation.cpp:1495-1524)."""
    L = _lib.load()
    rows = ctypes.c_int()
    _lib.check(L.ems_fem_load_kmat(path.encode(), None, 0, ctypes.byref(rows)))
    n = rows.value
    out = np.empty((n, n), dtype=np.float64)
    _lib.check(L.ems_fem_load_kmat(path.encode(), _lib.dptr(out), out.size, None))
    return out


def kmat_file(order, use_ofem, kmat_dir=None):
    name = ("%d.csv" % order) if use_ofem else ("lamped%d.csv" % order)
    return os.path.join(kmat_dir if kmat_dir is not None else "kmat", name)


class FemSimulation:
    """Order-p hex edge FEM on one GPU (reference: src/fem_v2/fem_simulation.hpp:14-24)."""

    def __init__(self, order, grid_size, domain_size, time_step, permittivity, permeability,
                 time_frequency, use_ofem=1, *mpi_args, stepper="fem_v2", precision="f64",
                 kmat=None, kmat_dir=None):
        L = _lib.load()
        if isinstance(stepper, str):
            stepper = STEPPERS[stepper]
        if kmat is None:
            kmat = load_kmat(kmat_file(order, use_ofem, kmat_dir))
        kmat = np.ascontiguousarray(kmat, dtype=np.float64)
        prec = 1 if precision in ("f32", "fp32", "float32") else 0
        self.grid = tuple(int(g) for g in grid_size)
        self.order = int(order)
        d = _lib.FemDesc(int(order), self.grid[0], self.grid[1], self.grid[2], float(domain_size),
                         float(time_step), float(permittivity), float(permeability), int(time_frequency),
                         int(stepper), prec)
        h = ctypes.c_void_p()
        _lib.check(L.ems_fem_create(ctypes.byref(d), _lib.dptr(kmat), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ems_fem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setSource_x(self, position):
        self._src(0, position)

    def setSource_y(self, position):
        self._src(1, position)

    def setSource_z(self, position):
        self._src(2, position)

    def _src(self, c, p):
        _lib.check(_lib.load().ems_fem_add_source(self._h, c, int(p[0]), int(p[1]), int(p[2])))

    def setObservationPoint(self, position):
        _lib.check(_lib.load().ems_fem_add_observation(self._h, *[int(v) for v in position]))

    def run(self, num_steps):
        _lib.check(_lib.load().ems_fem_run(self._h, int(num_steps)))

    def run_timed(self, num_steps):
        tot, ker = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.load().ems_fem_run_timed(self._h, int(num_steps), ctypes.byref(tot), ctypes.byref(ker)))
        return tot.value, ker.value

    def saveResults(self, num_steps, filename):
        _lib.check(_lib.load().ems_fem_save_results(self._h, int(num_steps), filename.encode()))

    def traces(self):
        L = _lib.load()
        nobs, nslot = ctypes.c_int(), ctypes.c_int()
        _lib.check(L.ems_fem_trace_shape(self._h, ctypes.byref(nobs), ctypes.byref(nslot)))
        out = np.zeros((nobs.value, nslot.value, 4))
        _lib.check(L.ems_fem_get_traces(self._h, _lib.dptr(out), out.size))
        return out

    def size(self, comp):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().ems_fem_field_size(self._h, comp, ctypes.byref(n)))
        return n.value

    def get(self, comp, slot):
        out = np.empty(self.size(comp))
        _lib.check(_lib.load().ems_fem_get_field(self._h, comp, slot, _lib.dptr(out), out.size))
        return out

    def set(self, comp, slot, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _lib.check(_lib.load().ems_fem_set_field(self._h, comp, slot, _lib.dptr(v), v.size))

    def get_state(self, slot):
        return np.concatenate([self.get(c, slot) for c in range(3)])

    def set_state(self, slot, v):
        n0, n1 = self.size(0), self.size(1)
        self.set(0, slot, v[:n0])
        self.set(1, slot, v[n0:n0 + n1])
        self.set(2, slot, v[n0 + n1:])

    def set_tiling(self, rows, planes=8):
        """First-order kernels: rows 4, 8 or 16 = one z-marching pass over all
        components (64 x rows workgroups, `planes` planes each; 16 for the
        FP32 stencil only, else 8), 0 = one gather kernel per component.  Results do not depend on it, except
        that rows 0 always takes the element row dots (not the stencil)."""
        _lib.check(_lib.load().ems_fem_set_tiling(self._h, int(rows), int(planes)))

    def set_element_product(self, kind):
        """Orders 2 and 3: 'rows' (one thread per element, the reference's
        sums), 'mfma' (matrix cores; equal to rounding) or 'auto' (the
        default: mfma at order 3, rows at order 2).  Order 1: 'stencil' (one
        33-point stencil per component in the z-march; equal to rounding;
        also 'auto') or 'rows' (the element row dots in the reference's
        order, bit for bit)."""
        k = {"auto": 0, "rows": 1, "mfma": 2, "stencil": 3}[kind] if isinstance(kind, str) else int(kind)
        _lib.check(_lib.load().ems_fem_set_element_product(self._h, k))

    def ndof(self):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().ems_fem_dof_count(self._h, ctypes.byref(n)))
        return n.value

    @property
    def current_time(self):
        t = ctypes.c_double()
        _lib.check(_lib.load().ems_fem_get_time(self._h, ctypes.byref(t)))
        return t.value
