"""Python mirror of the reference classes of src/fem and src/fem_abm over the C ABI.

``FemSimulation`` keeps the constructor and method names of
src/fem/fem_simulation.hpp:10-40 (grid, h, dt, eps, mu, tf, stabilization
factor, use_ofem); ``AbmSimulation`` those of src/fem_abm/fem_simulation.hpp:
10-40 (no stabilization argument: the reference hard-codes 0.05).
"""
import ctypes

import numpy as np

from . import _lib

LEAPFROG, ABM = 0, 1
ABM_STABILIZATION = 0.05  # src/fem_abm/fem_simulation.cpp:29


def _desc(grid, h, dt, eps, mu, tf, stepper, stab, use_ofem, precision):
    prec = 1 if precision in ("f32", "fp32", "float32") else 0
    g = [int(v) for v in grid]
    return _lib.Fem1Desc(g[0], g[1], g[2], float(h), float(dt), float(eps), float(mu), int(tf), int(stepper),
                         float(stab), int(use_ofem), prec)


def element_matrix(stepper, grid, h, dt, eps, mu, stab, use_ofem=0):
    """The 12 x 12 element matrix the step uses (host computation)."""
    out = np.zeros(144)
    d = _desc(grid, h, dt, eps, mu, 1, stepper, stab, use_ofem, "f64")
    _lib.check(_lib.load().ems_fem1_element_matrix(ctypes.byref(d), _lib.dptr(out)))
    return out.reshape(12, 12)


class _Fem1:
    def __init__(self, grid_size, domain_size, time_step, permittivity, permeability, time_frequency,
                 stepper, stab, use_ofem, precision):
        L = _lib.load()
        self.grid = tuple(int(g) for g in grid_size)
        self.stepper = stepper
        d = _desc(self.grid, domain_size, time_step, permittivity, permeability, time_frequency, stepper, stab,
                  use_ofem, precision)
        h = ctypes.c_void_p()
        _lib.check(L.ems_fem1_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ems_fem1_destroy(self._h)
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
        _lib.check(_lib.load().ems_fem1_add_source(self._h, c, int(p[0]), int(p[1]), int(p[2])))

    def setSourceDipole_x(self, position):
        p = [int(v) for v in position]
        _lib.check(_lib.load().ems_fem1_add_dipole_x(self._h, p[0], p[1], p[2]))

    def setObservationPoint(self, lists):
        counts = (ctypes.c_int * 3)(*[len(l) for l in lists])
        flat = [int(v) for l in lists for p in l for v in p] or [0]
        pts = (ctypes.c_int * len(flat))(*flat)
        _lib.check(_lib.load().ems_fem1_add_observation(self._h, counts, pts))

    def run(self, num_steps):
        _lib.check(_lib.load().ems_fem1_run(self._h, int(num_steps)))

    def run_timed(self, num_steps):
        tot, ker = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.load().ems_fem1_run_timed(self._h, int(num_steps), ctypes.byref(tot), ctypes.byref(ker)))
        return tot.value, ker.value

    def traces(self):
        L = _lib.load()
        nobs, nslot = ctypes.c_int(), ctypes.c_int()
        _lib.check(L.ems_fem1_trace_shape(self._h, ctypes.byref(nobs), ctypes.byref(nslot)))
        out = np.zeros((nobs.value, nslot.value, 3))
        _lib.check(L.ems_fem1_get_traces(self._h, _lib.dptr(out), out.size))
        return out

    def size(self, field):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().ems_fem1_field_size(self._h, field, ctypes.byref(n)))
        return n.value

    def field(self, f, slot=1):
        out = np.empty(self.size(f))
        _lib.check(_lib.load().ems_fem1_get_field(self._h, f, slot, _lib.dptr(out), out.size))
        return out

    def set_field(self, f, slot, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _lib.check(_lib.load().ems_fem1_set_field(self._h, f, slot, _lib.dptr(v), v.size))

    def matrix(self):
        out = np.zeros(144)
        _lib.check(_lib.load().ems_fem1_matrix(self._h, _lib.dptr(out)))
        return out.reshape(12, 12)

    def set_element_product(self, kind):
        """'stencil' (the default, also 'auto': one 33-point stencil per
        component on interior edges; equal to rounding) or 'rows' (the
        element row dots in the reference's order, bit for bit)."""
        k = {"auto": 0, "rows": 1, "stencil": 3}[kind] if isinstance(kind, str) else int(kind)
        _lib.check(_lib.load().ems_fem1_set_element_product(self._h, k))

    def ndof(self):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().ems_fem1_dof_count(self._h, ctypes.byref(n)))
        return n.value

    @property
    def current_time(self):
        t = ctypes.c_double()
        _lib.check(_lib.load().ems_fem1_get_time(self._h, ctypes.byref(t)))
        return t.value


class FemSimulation(_Fem1):
    """src/fem: FemSimulation(grid, h, dt, eps, mu, tf, stabilization_factor, use_ofem)."""

    def __init__(self, grid_size, domain_size, time_step, permittivity, permeability, time_frequency,
                 stabilization_factor, use_ofem, precision="f64"):
        super().__init__(grid_size, domain_size, time_step, permittivity, permeability, time_frequency, LEAPFROG,
                         stabilization_factor, use_ofem, precision)

    def saveResults(self, num_steps, filename):
        _lib.check(_lib.load().ems_fem1_save_results(self._h, int(num_steps), filename.encode()))


class AbmSimulation(_Fem1):
    """src/fem_abm: FemSimulation(grid, h, dt, eps, mu, tf)."""

    def __init__(self, grid_size, domain_size, time_step, permittivity, permeability, time_frequency,
                 precision="f64", stabilization_factor=ABM_STABILIZATION):
        super().__init__(grid_size, domain_size, time_step, permittivity, permeability, time_frequency, ABM,
                         stabilization_factor, 0, precision)

    def saveResults(self, num_steps, directory=""):
        _lib.check(_lib.load().ems_fem1_save_results(self._h, int(num_steps), directory.encode()))
