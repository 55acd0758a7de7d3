"""ctypes binding of the FP64 FDTD oracle (oracle/fdtd_oracle.c).

TEST INFRASTRUCTURE ONLY: used as the checker by tests/, smoke() and the
cpu_baseline leg of bench.py.  Mirrors the reference class
``FemSimulation`` of src/fdtd/fdtd_simulation.hpp:9-80.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# EMS_ORACLE_LIB selects another build of the same sources (the mutation
# tests load deliberately altered copies through it)
_LIB_PATH = os.environ.get("EMS_ORACLE_LIB") or os.path.join(_HERE, "_build", "libems_oracle.so")
_lib = None

WAVELET_HERTZ, WAVELET_RICKER, WAVELET_GAUSS = 0, 1, 2


class OracleCfg(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("h", ctypes.c_double), ("dt", ctypes.c_double),
        ("eps", ctypes.c_double), ("mu", ctypes.c_double),
        ("tf", ctypes.c_int),
        ("wavelet", ctypes.c_int),
        ("f0", ctypes.c_double), ("t0", ctypes.c_double), ("amp", ctypes.c_double),
        ("npml", ctypes.c_int),
        ("pml_order", ctypes.c_double), ("pml_r0", ctypes.c_double),
        ("pml_kappa_max", ctypes.c_double), ("pml_alpha_max", ctypes.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.ems_oracle_fdtd_create.restype = P
        L.ems_oracle_fdtd_create.argtypes = [ctypes.POINTER(OracleCfg)]
        L.ems_oracle_fdtd_destroy.argtypes = [P]
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ems_oracle_fdtd_set_material.argtypes = [P, dp, dp]
        L.ems_oracle_fdtd_bind_source_material.argtypes = [P, dp, dp]
        L.ems_oracle_fdtd_add_source.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ems_oracle_fdtd_add_observation.argtypes = [P, ip, ip]
        L.ems_oracle_fdtd_run.argtypes = [P, ctypes.c_int]
        L.ems_oracle_fdtd_nslot.argtypes = [P]
        L.ems_oracle_fdtd_nobs.argtypes = [P]
        L.ems_oracle_set_threads.argtypes = [ctypes.c_int]
        L.ems_oracle_get_threads.restype = ctypes.c_int
        L.ems_oracle_fdtd_time.argtypes = [P]
        L.ems_oracle_fdtd_time.restype = ctypes.c_double
        L.ems_oracle_fdtd_field_size.argtypes = [P, ctypes.c_int]
        L.ems_oracle_fdtd_field_size.restype = ctypes.c_size_t
        L.ems_oracle_fdtd_get_field.argtypes = [P, ctypes.c_int, dp]
        L.ems_oracle_fdtd_set_field.argtypes = [P, ctypes.c_int, dp]
        L.ems_oracle_fdtd_get_traces.argtypes = [P, dp]
        L.ems_oracle_fdtd_bare_step.argtypes = [P]
        L.ems_oracle_hertz.restype = ctypes.c_double
        L.ems_oracle_hertz.argtypes = [ctypes.c_double] * 3
        L.ems_oracle_ricker.restype = ctypes.c_double
        L.ems_oracle_ricker.argtypes = [ctypes.c_double] * 3
        L.ems_oracle_gauss.restype = ctypes.c_double
        L.ems_oracle_gauss.argtypes = [ctypes.c_double] * 3
        L.ems_oracle_fdtd_adjoint_bare_step.argtypes = [P]
        L.ems_oracle_fdtd_get_adjoint.argtypes = [P, ctypes.c_int, dp]
        L.ems_oracle_fdtd_set_adjoint.argtypes = [P, ctypes.c_int, dp]
        L.ems_oracle_fdtd_get_psi.argtypes = [P, ctypes.c_int, dp]
        L.ems_oracle_fdtd_set_psi.argtypes = [P, ctypes.c_int, dp]
        L.ems_oracle_fdtd_uni_size.argtypes = [P]
        L.ems_oracle_fdtd_uni_size.restype = ctypes.c_size_t
        L.ems_oracle_fdtd_gradient.argtypes = [P, ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.ems_oracle_check_params.restype = ctypes.c_int
        L.ems_oracle_check_params.argtypes = [dp, ip, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


FIELD_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


class OracleFDTD:
    """FP64 CPU FDTD with the reference's constructor and method names."""

    def __init__(self, grid, h, dt, eps, mu, tf=1, wavelet=WAVELET_HERTZ, f0=1.0e9, t0=None,
                 amp=1.0, npml=0, pml_order=3.0, pml_r0=1e-6, pml_kappa_max=1.0,
                 pml_alpha_max=0.0):
        if t0 is None:
            t0 = 2 ** 0.5 / f0
        self.grid = tuple(int(g) for g in grid)
        self.cfg = OracleCfg(self.grid[0], self.grid[1], self.grid[2], h, dt, eps, mu, int(tf),
                             int(wavelet), f0, t0, amp, int(npml), pml_order, pml_r0,
                             pml_kappa_max, pml_alpha_max)
        self._h = lib().ems_oracle_fdtd_create(ctypes.byref(self.cfg))
        if not self._h:
            raise ValueError("oracle: invalid FDTD configuration")
        self._mat = (None, None)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ems_oracle_fdtd_destroy(self._h)
            self._h = None

    # reference: setSource_{x,y,z} (src/fdtd/fdtd_simulation.hpp:20-22)
    def setSource(self, comp, pos):
        lib().ems_oracle_fdtd_add_source(self._h, int(comp), *[int(p) for p in pos])
        if self._mat[0] is not None or self._mat[1] is not None:
            self._bind_src()

    def setSource_x(self, pos):
        self.setSource(0, pos)

    def setSource_y(self, pos):
        self.setSource(1, pos)

    def setSource_z(self, pos):
        self.setSource(2, pos)

    # reference: setObservationPoint (src/fdtd/fdtd_simulation.hpp:25)
    def setObservationPoint(self, lists):
        counts = np.array([len(l) for l in lists], dtype=np.int32)
        pts = np.array([p for l in lists for p in l] or [[0, 0, 0]], dtype=np.int32).reshape(-1, 3)
        pts = np.ascontiguousarray(pts)
        lib().ems_oracle_fdtd_add_observation(self._h, _ip(counts), _ip(pts))

    def set_material(self, eps_r=None, sigma=None):
        nx, ny, nz = self.grid
        n = (nx + 1) * (ny + 1) * (nz + 1)
        er = None if eps_r is None else np.ascontiguousarray(eps_r, dtype=np.float64).reshape(n)
        sg = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64).reshape(n)
        self._mat = (er, sg)
        lib().ems_oracle_fdtd_set_material(self._h, None if er is None else _dp(er),
                                           None if sg is None else _dp(sg))
        self._bind_src()

    def _bind_src(self):
        er, sg = self._mat
        lib().ems_oracle_fdtd_bind_source_material(self._h, None if er is None else _dp(er),
                                                   None if sg is None else _dp(sg))

    def run(self, num_steps):
        if lib().ems_oracle_fdtd_run(self._h, int(num_steps)) != 0:
            raise MemoryError("oracle run failed")

    def bare_step(self):
        lib().ems_oracle_fdtd_bare_step(self._h)

    @property
    def current_time(self):
        return lib().ems_oracle_fdtd_time(self._h)

    def traces(self):
        nobs = lib().ems_oracle_fdtd_nobs(self._h)
        nslot = lib().ems_oracle_fdtd_nslot(self._h)
        out = np.zeros((nobs, nslot, 3), dtype=np.float64)
        if out.size:
            lib().ems_oracle_fdtd_get_traces(self._h, _dp(out))
        return out

    def field(self, which):
        if isinstance(which, str):
            which = FIELD_NAMES.index(which)
        n = lib().ems_oracle_fdtd_field_size(self._h, which)
        out = np.empty(n, dtype=np.float64)
        lib().ems_oracle_fdtd_get_field(self._h, which, _dp(out))
        return out

    def set_field(self, which, values):
        if isinstance(which, str):
            which = FIELD_NAMES.index(which)
        n = lib().ems_oracle_fdtd_field_size(self._h, which)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(n)
        lib().ems_oracle_fdtd_set_field(self._h, which, _dp(v))

    def fields(self):
        return [self.field(w) for w in range(6)]

    # ---- discrete adjoint (no reference counterpart)
    def uni_size(self):
        return lib().ems_oracle_fdtd_uni_size(self._h)

    def adjoint_bare_step(self):
        lib().ems_oracle_fdtd_adjoint_bare_step(self._h)

    def adjoint(self, which):
        n = self.uni_size() if which >= 6 else lib().ems_oracle_fdtd_field_size(self._h, which)
        out = np.empty(n, dtype=np.float64)
        lib().ems_oracle_fdtd_get_adjoint(self._h, which, _dp(out))
        return out

    def set_adjoint(self, which, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        lib().ems_oracle_fdtd_set_adjoint(self._h, which, _dp(v))

    def psi(self, slot):
        out = np.empty(self.uni_size(), dtype=np.float64)
        if lib().ems_oracle_fdtd_get_psi(self._h, slot, _dp(out)) != 0:
            raise ValueError("no CPML")
        return out

    def set_psi(self, slot, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        lib().ems_oracle_fdtd_set_psi(self._h, slot, _dp(v))

    def gradient(self, nsteps, d_obs):
        """Misfit J and its gradient with respect to eps_r and sigma (material
        grid) of one run of nsteps from a zero state; d_obs [nobs][nsteps/tf+1][3]."""
        nx, ny, nz = self.grid
        nm = (nx + 1) * (ny + 1) * (nz + 1)
        er, sg = self._mat
        d = np.ascontiguousarray(d_obs, dtype=np.float64)
        ge, gs = np.zeros(nm), np.zeros(nm)
        J = ctypes.c_double()
        lib().ems_oracle_fdtd_gradient(self._h, int(nsteps), None if er is None else _dp(er),
                                       None if sg is None else _dp(sg), _dp(d),
                                       ctypes.cast(ctypes.byref(J), ctypes.POINTER(ctypes.c_double)),
                                       _dp(ge), _dp(gs))
        return J.value, ge, gs


def hertz(t, h, f0=1.0e9):
    return lib().ems_oracle_hertz(t, h, f0)


def ricker(t, f0, t0):
    return lib().ems_oracle_ricker(t, f0, t0)


def gauss(t, f0, t0):
    return lib().ems_oracle_gauss(t, f0, t0)


def check_params(domain_sizes, grid_num, duration, time_step, domain_size, c):
    ds = np.ascontiguousarray(domain_sizes, dtype=np.float64)
    gn = np.ascontiguousarray(grid_num, dtype=np.int32)
    return bool(lib().ems_oracle_check_params(_dp(ds), _ip(gn), duration, time_step, domain_size, c))


# Reference physical constants (src/fdtd/main.cpp:24-27)
EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6
C0 = 1.0 / (EPS0 * MU0) ** 0.5
