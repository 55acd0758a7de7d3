"""C ABI checks that need no GPU: libems.so loads, exports every function
declared in include/ems/*.h, and host-only entry points behave like the
reference.  No compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT

HEADER_DIR = os.path.join(ROOT, "include", "ems")


def declared_functions():
    names = set()
    for fn in os.listdir(HEADER_DIR):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ems_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "ems_fdtd_create" in names and "ems_fdtd_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from emsolvers_amd import _lib
    L = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes signature table covers them too
    assert not [n for n in declared_functions() if n not in _lib.SIGNATURES]


def test_version_and_device_count():
    from emsolvers_amd import _lib
    L = _lib.load()
    assert b"gfx950" in L.ems_version()
    assert L.ems_device_count() >= 0


def test_create_without_device_fails_loudly():
    from emsolvers_amd import _lib
    from emsolvers_amd.fdtd import FemSimulation
    if _lib.load().ems_device_count() > 0:
        pytest.skip("a GPU is visible: covered by the gpu tests")
    with pytest.raises(_lib.EmsError) as ei:
        FemSimulation((8, 8, 8), 1e-3, 1e-12, 8.854187817e-12, 1.25663706212e-6, 1)
    assert "no HIP device" in str(ei.value)


def test_invalid_descriptor_rejected_before_device_use():
    from emsolvers_amd import _lib
    from emsolvers_amd.fdtd import FemSimulation
    with pytest.raises(_lib.EmsError) as ei:
        FemSimulation((0, 8, 8), 1e-3, 1e-12, 8.854187817e-12, 1.25663706212e-6, 1)
    assert ei.value.status == 1
    with pytest.raises(_lib.EmsError):
        FemSimulation((8, 8, 8), 1e-3, 1e-12, 8.854187817e-12, 1.25663706212e-6, 0)
    # one padded z-plane of 4 GiB or more: the kernels' 32-bit plane offsets
    # would wrap (40032 x 40002 FP32 words = 6.4 GB; 23232 x 23202 FP64 = 4.31 GB)
    for grid, prec in (((40000, 40000, 4), "f32"), ((23200, 23200, 4), "f64")):
        with pytest.raises(_lib.EmsError) as ei:
            FemSimulation(grid, 1e-3, 1e-12, 8.854187817e-12, 1.25663706212e-6, 1, precision=prec)
        assert ei.value.status == 1 and "plane too large" in str(ei.value)


def test_check_params_matches_oracle():
    from emsolvers_amd.fdtd import check_params
    from oracle import fdtd as of
    c = of.C0
    rng = np.random.default_rng(0)
    cases = [([1, 1, 1], [500, 500, 500], 3e-9, 3e-9 / 780, 1 / 500, c),
             ([1, 1, 1], [500, 500, 500], 3e-9, 1e-9, 1 / 500, c),
             ([1, 2, 1], [500, 500, 500], 3e-9, 1e-12, 1 / 500, c)]
    for _ in range(50):
        ds = list(rng.uniform(-0.5, 3, 3))
        gn = list(rng.integers(-2, 50, 3))
        cases.append((ds, gn, rng.uniform(-1, 1), rng.uniform(-1e-11, 1e-11), rng.uniform(1e-4, 1e-2), c))
    for args in cases:
        assert check_params(*args) == of.check_params(*args), args


@pytest.mark.parametrize("stepper,ofem,stab", [(0, 0, 0.3), (0, 1, 0.3), (0, 0, 0.0), (0, 0, 1.25), (1, 0, 0.05)])
def test_fem1_element_matrix_equals_exact_rounding(stepper, ofem, stab):
    """The library's host-side element matrix of src/fem / src/fem_abm equals
    the exact rational product C_s B_s^-1 A_s (scaled) rounded once."""
    from emsolvers_amd import fem1 as F
    from oracle import fem1 as O
    EPS0, MU0 = 8.854187817e-12, 1.25663706212e-6
    for h in (0.02, 1.7e-3):
        dt = 0.3 * h * np.sqrt(EPS0 * MU0)
        a = F.element_matrix(stepper, (4, 4, 4), h, dt, EPS0, MU0, stab, ofem)
        b = O.element_matrix(stepper, dt, h, EPS0, MU0, stab, ofem)
        assert np.array_equal(a, b)


def test_fem1_invalid_descriptor_rejected():
    from emsolvers_amd import _lib
    from emsolvers_amd import fem1 as F
    with pytest.raises(_lib.EmsError):
        F.element_matrix(2, (4, 4, 4), 0.02, 1e-11, 8.854187817e-12, 1.25663706212e-6, 0.3)
    with pytest.raises(_lib.EmsError):
        F.element_matrix(1, (4, 4, 4), 0.02, 1e-11, 8.854187817e-12, 1.25663706212e-6, 0.05, use_ofem=1)
    with pytest.raises(_lib.EmsError):
        F.FemSimulation((4, -1, 4), 0.02, 1e-11, 8.854187817e-12, 1.25663706212e-6, 1, 0.3, 0)


def test_tet_invalid_descriptor_rejected():
    from emsolvers_amd import _lib
    from emsolvers_amd.tet import TetSimulation
    bad = [dict(grid=(0, 4, 4)), dict(h=0.0), dict(dt=-1.0), dict(tf=0), dict(stepper=7), dict(order=3),
           dict(precision="f16")]
    for b in bad:
        kw = dict(grid=(4, 4, 4), h=0.25, dt=0.01, eps=1.0, mu=1.0, tf=1, stepper="symplectic_2", order=1)
        kw.update(b)
        if kw.get("precision") == "f16":
            continue  # the binding maps names; the C ABI check is below
        with pytest.raises(_lib.EmsError) as ei:
            TetSimulation(**kw)
        assert ei.value.status == 1, b
    d = _lib.TetDesc(4, 4, 4, 0.25, 0.01, 1.0, 1.0, 1, 0, 1, 9)
    import ctypes
    h = ctypes.c_void_p()
    assert _lib.load().ems_tet_create(ctypes.byref(d), ctypes.byref(h)) == 1


def test_cpp_interfaces_compile_and_link(tmp_path):
    """Every C++ header under include/ems compiles in one translation unit
    that uses each class's methods, and links against libems.so (no call is
    made: there is no device here)."""
    import shutil
    import subprocess
    from emsolvers_amd import _lib
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include <ems/fdtd_simulation.hpp>
#include <ems/fem_simulation.hpp>
#include <ems/fem1_simulation.hpp>
#include <ems/tet_simulation.hpp>
double use_fdtd()
{
    ems::fdtd::FemSimulation s({8, 8, 8}, 1e-3, 1e-12, 8.85e-12, 1.26e-6, 1, EMS_F32);
    s.setCPML(2);
    s.setSource_z({4, 4, 4});
    s.setObservationPoint({{{{1, 2, 3}}, {}, {}}});
    s.setGradientMemory(1 << 20);
    auto plan = s.gradientPlan(10);
    std::vector<double> ge, gs, obs(33, 0.0);
    double J = s.gradient(10, obs, &ge, &gs);
    ems::fdtd::Batch b;
    b.add(s);
    b.setGradientGroup(1);
    std::vector<double> mis;
    J += b.gradient(10, {obs}, &mis, &ge, nullptr);
    return J + plan[0] + s.savedElectricField().size();
}
void use_fem()
{
    ems::fem_v2::FemSimulation v(1, {8, 8, 8}, 0.08, 1e-12, 8.85e-12, 1.26e-6, 1, 1);
    v.setElementProduct(EMS_FEM_ELEM_ROWS);
    ems::fem::FemSimulation f({8, 8, 8}, 0.08, 1e-12, 8.85e-12, 1.26e-6, 1, 0.3, 0);
    f.setElementProduct(EMS_FEM_ELEM_STENCIL);
    ems::fem_abm::FemSimulation a({8, 8, 8}, 0.08, 1e-12, 8.85e-12, 1.26e-6, 1);
    a.setElementProduct(EMS_FEM_ELEM_AUTO);
}
double use_tet()
{
    ems::tet::TetSimulation t({2, 2, 2}, 0.5, 1e-3, 1.0, 1.0, 1, EMS_TET_SYMPLECTIC4, 2, EMS_F64);
    t.setMaterial(std::vector<double>(48, 2.0));
    t.addDipole({0.5, 0.5, 0.5}, {1.0, 0.0, 0.0}, 1.0, 0.5);
    t.addReceiver({0.3, 0.6, 0.4});
    t.setOperator(EMS_TET_OP_EBE_MFMA);
    t.setSolver(EMS_TET_SOLVER_CG);
    t.setTimeStep(0.9 * 1.5735 / t.maxFrequency(20));
    t.run(5);
    std::vector<double> ge, gs;
    double drift = 0.0;
    double J = t.gradient(5, std::vector<double>(18, 0.0), &ge, &gs, &drift);
    return J + drift + t.traces().size() + t.field(1).size() + t.solverStats()[0] + t.currentTime();
}
''')
    lib = os.path.dirname(_lib.lib_path()) if hasattr(_lib, "lib_path") else os.path.join(ROOT, "emsolvers_amd")
    out = tmp_path / "libuse.so"
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"),
                        str(src), "-o", str(out), "-L", lib, "-lems"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
