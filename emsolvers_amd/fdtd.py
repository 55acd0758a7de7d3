"""Python mirror of the reference FDTD class over the C ABI.

``FemSimulation`` keeps the names, argument meanings and error behaviour of
the reference class ``FemSimulation`` in src/fdtd/fdtd_simulation.hpp:9-80
(yes, the reference names its FDTD class that way).  All work runs in
libems.so on the GPU.
"""
import ctypes

import numpy as np

from . import _lib

F64, F32 = 0, 1
WAVELET_HERTZ, WAVELET_RICKER, WAVELET_GAUSS = 0, 1, 2
KERNEL_AUTO, KERNEL_SPLIT, KERNEL_FUSED, KERNEL_ZMARCH, KERNEL_TB2 = 0, 1, 2, 3, 4
FIELD_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")

# Reference physical constants (src/fdtd/main.cpp:24-27)
EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6
C0 = 1.0 / (EPS0 * MU0) ** 0.5


def _prec(p):
    if p in (F64, "f64", "fp64", "float64", np.float64):
        return F64
    if p in (F32, "f32", "fp32", "float32", np.float32):
        return F32
    raise ValueError("precision must be f64 or f32")


class FemSimulation:
    """Yee FDTD on one GPU (reference: src/fdtd/fdtd_simulation.hpp:13-28)."""

    def __init__(self, grid_size, domain_size, time_step, permittivity, permeability,
                 time_frequency, precision="f64"):
        L = _lib.load()
        nx, ny, nz = (int(g) for g in grid_size)
        self.grid = (nx, ny, nz)
        self.domain_size = float(domain_size)
        self.time_step = float(time_step)
        self.time_frequency = int(time_frequency)
        self.precision = _prec(precision)
        d = _lib.FdtdDesc(nx, ny, nz, float(domain_size), float(time_step), float(permittivity),
                          float(permeability), int(time_frequency), self.precision)
        h = ctypes.c_void_p()
        _lib.check(L.ems_fdtd_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ems_fdtd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- reference API -----------------------------------------------------
    def setSource_x(self, position):
        self._source(0, position)

    def setSource_y(self, position):
        self._source(1, position)

    def setSource_z(self, position):
        self._source(2, position)

    def _source(self, comp, p):
        _lib.check(_lib.load().ems_fdtd_add_source(self._h, comp, int(p[0]), int(p[1]), int(p[2])))

    def setObservationPoint(self, position):
        """position: three lists of (i,j,k) points for Ex, Ey and Ez."""
        counts = np.array([len(l) for l in position], dtype=np.int32)
        pts = np.array([tuple(p) for l in position for p in l] or [(0, 0, 0)], dtype=np.int32)
        pts = np.ascontiguousarray(pts.reshape(-1, 3))
        _lib.check(_lib.load().ems_fdtd_add_observation(self._h, _lib.iptr(counts), _lib.iptr(pts)))

    def run(self, num_steps):
        _lib.check(_lib.load().ems_fdtd_run(self._h, int(num_steps)))

    def run_static(self, num_steps):
        """ems_fdtd_run_static: the same run with its staging in the
        library's static device array (no device allocation)."""
        _lib.check(_lib.load().ems_fdtd_run_static(self._h, int(num_steps)))

    def saveResults(self, num_steps, c, directory=""):
        _lib.check(_lib.load().ems_fdtd_save_results(self._h, int(num_steps), float(c),
                                                     directory.encode()))

    # --- extensions ----------------------------------------------------------
    def set_material(self, eps_r=None, sigma=None):
        """Per-cell eps_r / sigma on the (nx+1)(ny+1)(nz+1) material grid."""
        nx, ny, nz = self.grid
        n = (nx + 1) * (ny + 1) * (nz + 1)
        er = None if eps_r is None else np.ascontiguousarray(eps_r, dtype=np.float64).reshape(n)
        sg = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64).reshape(n)
        _lib.check(_lib.load().ems_fdtd_set_material(
            self._h, None if er is None else _lib.dptr(er), None if sg is None else _lib.dptr(sg)))

    def set_cpml(self, ncells, order=3.0, r0=1e-6, kappa_max=1.0, alpha_max=0.0):
        _lib.check(_lib.load().ems_fdtd_set_cpml(self._h, int(ncells), float(order), float(r0),
                                                 float(kappa_max), float(alpha_max)))

    def set_wavelet(self, kind, f0, t0=-1.0, amp=1.0):
        _lib.check(_lib.load().ems_fdtd_set_wavelet(self._h, int(kind), float(f0), float(t0),
                                                    float(amp)))

    def set_kernel(self, variant):
        _lib.check(_lib.load().ems_fdtd_set_kernel(self._h, int(variant)))

    def set_tiling(self, rows, planes, flags=0):
        """Fused-kernel workgroup rows (8, 16), planes per workgroup (or per
        split-kernel thread) and flags (bit 0 prefetch, bit 1 non-temporal
        stores, bit 2 split sweeps with a CPML as interior + CPML-shell launches)."""
        _lib.check(_lib.load().ems_fdtd_set_tiling(self._h, int(rows), int(planes), int(flags)))

    def set_zmarch(self, rows=4, planes=8, waves=1):
        """Tiling of the z-march sweep (kernel 3): strip rows (2, 4), planes
        per chunk, wavefronts per workgroup (1, 4)."""
        _lib.check(_lib.load().ems_fdtd_set_zmarch(self._h, int(rows), int(planes), int(waves)))

    def set_tb2(self, cols=0, rows=0, planes=0):
        """Tiling of the two-step sweep (kernel 4): thread tile cols x rows
        (cols * rows <= 1024; stored core (cols-3) x (rows-3)), planes per
        workgroup; 0 = default."""
        _lib.check(_lib.load().ems_fdtd_set_tb2(self._h, int(cols), int(rows), int(planes)))

    def autotune_tb2(self, trial_steps=10):
        """Times candidate two-step tilings on scratch fields (the state is
        untouched), keeps the fastest and returns it as (cols, rows, planes)."""
        c, r, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().ems_fdtd_autotune_tb2(self._h, int(trial_steps), ctypes.byref(c), ctypes.byref(r),
                                                     ctypes.byref(p)))
        return c.value, r.value, p.value

    def run_timed(self, num_steps):
        """Returns (total_ms, kernel_ms) measured with HIP events."""
        tot, ker = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.load().ems_fdtd_run_timed(self._h, int(num_steps), ctypes.byref(tot),
                                                  ctypes.byref(ker)))
        return tot.value, ker.value

    def traces(self):
        """saved_electric_field_ as an array [nobs][nslot][3]."""
        nobs, nslot = ctypes.c_int(), ctypes.c_int()
        L = _lib.load()
        _lib.check(L.ems_fdtd_trace_shape(self._h, ctypes.byref(nobs), ctypes.byref(nslot)))
        out = np.zeros((nobs.value, nslot.value, 3), dtype=np.float64)
        _lib.check(L.ems_fdtd_get_traces(self._h, _lib.dptr(out), out.size))
        return out

    def field_size(self, which):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().ems_fdtd_field_size(self._h, self._which(which), ctypes.byref(n)))
        return n.value

    def field(self, which):
        w = self._which(which)
        out = np.empty(self.field_size(w), dtype=np.float64)
        _lib.check(_lib.load().ems_fdtd_get_field(self._h, w, _lib.dptr(out), out.size))
        return out

    def set_field(self, which, values):
        w = self._which(which)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _lib.check(_lib.load().ems_fdtd_set_field(self._h, w, _lib.dptr(v), v.size))

    def fields(self):
        return [self.field(w) for w in range(6)]

    @property
    def current_time(self):
        t = ctypes.c_double()
        _lib.check(_lib.load().ems_fdtd_get_time(self._h, ctypes.byref(t)))
        return t.value

    def set_checkpoint_interval(self, interval):
        """One-level checkpoint interval of gradient() (0: chosen per call
        within the memory cap; results do not depend on it)."""
        _lib.check(_lib.load().ems_fdtd_set_checkpoint_interval(self._h, int(interval)))

    def set_gradient_memory(self, nbytes):
        """Cap on the snapshot and checkpoint bytes of gradient() (0: 3/4 of
        the free device memory); the schedule with the fewest recomputed
        steps within it is taken (results do not depend on it)."""
        _lib.check(_lib.load().ems_fdtd_set_gradient_memory(self._h, int(nbytes)))

    def gradient_plan(self, num_steps):
        """(level-1 interval, snapshot segment length, bytes) of the schedule
        gradient(num_steps) would take now (level-1 interval = num_steps: one
        level)."""
        k1, ks, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        _lib.check(_lib.load().ems_fdtd_gradient_plan(self._h, int(num_steps), ctypes.byref(k1), ctypes.byref(ks),
                                                      ctypes.byref(b)))
        return k1.value, ks.value, b.value

    def gradient(self, num_steps, d_obs):
        """Misfit J = 1/2 sum (trace - d_obs)^2 of a run of num_steps from a
        zero state, and dJ/deps_r, dJ/dsigma on the material grid (FP64).
        d_obs: [nobs][num_steps/tf + 1][3]."""
        nx, ny, nz = self.grid
        nm = (nx + 1) * (ny + 1) * (nz + 1)
        d = np.ascontiguousarray(d_obs, dtype=np.float64)
        ge, gs = np.zeros(nm), np.zeros(nm)
        J = ctypes.c_double()
        _lib.check(_lib.load().ems_fdtd_gradient(self._h, int(num_steps), _lib.dptr(d), d.size,
                                                 ctypes.cast(ctypes.byref(J), _lib._DP), _lib.dptr(ge),
                                                 _lib.dptr(gs)))
        return J.value, ge, gs

    def device_bytes(self):
        b = ctypes.c_size_t()
        _lib.check(_lib.load().ems_fdtd_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    @staticmethod
    def _which(w):
        return FIELD_NAMES.index(w) if isinstance(w, str) else int(w)


class Batch:
    """Many FemSimulation handles advanced together (one launch per sweep for
    all members); members keep their own sources, receivers and material."""

    def __init__(self, members=()):
        h = ctypes.c_void_p()
        _lib.check(_lib.load().ems_batch_create(ctypes.byref(h)))
        self._h = h
        self.members = []
        for m in members:
            self.add(m)

    def add(self, sim):
        _lib.check(_lib.load().ems_batch_add(self._h, sim._h))
        self.members.append(sim)

    def set_tiling(self, rows, planes):
        _lib.check(_lib.load().ems_batch_set_tiling(self._h, int(rows), int(planes)))

    def autotune(self, trial_steps=5):
        r, p = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().ems_batch_autotune(self._h, int(trial_steps), ctypes.byref(r), ctypes.byref(p)))
        return r.value, p.value

    def set_graph(self, steps_per_graph):
        _lib.check(_lib.load().ems_batch_set_graph(self._h, int(steps_per_graph)))

    def run(self, num_steps):
        _lib.check(_lib.load().ems_batch_run(self._h, int(num_steps)))

    def run_timed(self, num_steps):
        t = ctypes.c_double()
        _lib.check(_lib.load().ems_batch_run_timed(self._h, int(num_steps), ctypes.byref(t)))
        return t.value

    def set_gradient_group(self, shots):
        """Shots per group of gradient() (0: the most whose snapshots of every
        step fit in half of the device memory); the summation order, and so
        the gradient's last bits, depend on it."""
        _lib.check(_lib.load().ems_batch_set_gradient_group(self._h, int(shots)))

    def gradient(self, num_steps, d_obs):
        """Shots on one model: (sum of misfits, per-member misfits, dJ/deps_r,
        dJ/dsigma of the sum); d_obs[q] as for FemSimulation.gradient."""
        nx, ny, nz = self.members[0].grid
        nm = (nx + 1) * (ny + 1) * (nz + 1)
        ds = [np.ascontiguousarray(d, dtype=np.float64) for d in d_obs]
        if len(ds) != len(self.members):
            raise ValueError("one d_obs per member")
        ptrs = (_lib._DP * len(ds))(*[_lib.dptr(d) for d in ds])
        mis = np.zeros(len(ds))
        ge, gs = np.zeros(nm), np.zeros(nm)
        J = ctypes.c_double()
        _lib.check(_lib.load().ems_batch_gradient(self._h, int(num_steps), ptrs, _lib.dptr(mis),
                                                  ctypes.cast(ctypes.byref(J), _lib._DP), _lib.dptr(ge),
                                                  _lib.dptr(gs)))
        return J.value, mis, ge, gs

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ems_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_params(domain_sizes, grid_num, duration, time_step, domain_size, c):
    """Reference check_params (src/fdtd/fdtd_simulation.cpp:333-366)."""
    ds = np.ascontiguousarray(domain_sizes, dtype=np.float64)
    gn = np.ascontiguousarray(grid_num, dtype=np.int32)
    return bool(_lib.load().ems_fdtd_check_params(_lib.dptr(ds), _lib.iptr(gn), float(duration),
                                                  float(time_step), float(domain_size), float(c)))
