"""GPU parity of the first-order stabilized hex solvers (src/fem leapfrog,
src/fem_abm Adams-Bashforth-Moulton) against the FP64 oracle.

Bar: FP64 bit-exact with the element row dots ('rows': one thread per edge
gathers its elements in the serial reference's order, each element row
summed from 0 as the reference does; the element matrix is the same bits on
both sides, pinned separately against its exact rational value); 1e-10
relative L2 with the merged stencil (the default); FP32 receiver traces
within 1e-5 relative L2 of the FP64 oracle.
"""
import os

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6

CASES = {
    # src/fem: dt = duration / num_steps of src/fem/main.cpp (0.3 h / c at h = 0.02)
    "fem-ofem0": dict(stepper=0, stab=0.3, ofem=0, grid=(7, 6, 8), h=0.02, dt=2e-11, tf=1, steps=60,
                      dip=[(2, 1, 2)], pts=[(0, (3, 2, 2)), (1, (1, 2, 3)), (2, (2, 2, 2)), (0, (3, 2, 2))]),
    "fem-ofem1": dict(stepper=0, stab=0.3, ofem=1, grid=(6, 8, 6), h=0.02, dt=2e-11, tf=2, steps=50,
                      dip=[(2, 3, 1), (1, 1, 1)], pts=[]),
    "fem-nostab": dict(stepper=0, stab=0.0, ofem=0, grid=(5, 5, 5), h=0.02, dt=2e-11, tf=1, steps=40,
                       dip=[(1, 1, 1)], pts=[(2, (2, 2, 2))]),
    "abm": dict(stepper=1, stab=0.05, ofem=0, grid=(6, 6, 7), h=0.02, dt=1e-11, tf=1, steps=60,
                dip=[(2, 2, 2)], pts=[(0, (2, 3, 3)), (2, (3, 3, 3))]),
    "abm-tf3": dict(stepper=1, stab=0.05, ofem=0, grid=(8, 5, 6), h=0.02, dt=1e-11, tf=3, steps=50,
                    dip=[(3, 1, 2)], pts=[]),
}

OBS = [
    [[(3, 2, 3), (3, 3, 3)], [(2, 2, 3), (2, 3, 3), (3, 2, 3), (3, 3, 3)], [(2, 3, 2), (2, 3, 3)]],
    [[(1, 1, 1)], [(1, 1, 1)], [(1, 1, 1)]],
]


def build(kind, case, precision="f64", elem="rows"):
    st, g = case["stepper"], case["grid"]
    if kind == "oracle":
        from oracle.fem1 import OracleFem1
        s = OracleFem1(g, case["h"], case["dt"], EPS0, MU0, case["tf"], st, case["stab"], case["ofem"])
    else:
        from emsolvers_amd.fem1 import FemSimulation, AbmSimulation
        if st == 0:
            s = FemSimulation(g, case["h"], case["dt"], EPS0, MU0, case["tf"], case["stab"], case["ofem"],
                              precision=precision)
        else:
            s = AbmSimulation(g, case["h"], case["dt"], EPS0, MU0, case["tf"], precision=precision,
                              stabilization_factor=case["stab"])
        s.set_element_product(elem)
    for comp, p in case["pts"]:
        [s.setSource_x, s.setSource_y, s.setSource_z][comp](p)
    for d in case["dip"]:
        s.setSourceDipole_x(d)
    for o in OBS:
        s.setObservationPoint(o)
    return s


def nfields(case):
    return 6 if case["stepper"] == 1 else 3


@pytest.fixture(scope="module")
def oracle_runs():
    cache = {}

    def get(name):
        if name not in cache:
            c = CASES[name]
            o = build("oracle", c)
            o.run(c["steps"])
            cache[name] = o
        return cache[name]
    return get


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_bitwise_with_oracle(oracle_runs, name):
    case = CASES[name]
    print("case", name, flush=True)
    o = oracle_runs(name)
    g = build("gpu", case)
    assert np.array_equal(g.matrix(), o.K)
    g.run(case["steps"])
    to, tg = o.traces(), g.traces()
    assert tg.shape == to.shape
    assert np.abs(to).max() > 0
    assert np.array_equal(tg, to)
    for f in range(nfields(case)):
        for slot in (0, 1):
            assert np.array_equal(g.field(f, slot), o.field(f, slot)), (f, slot)
    assert g.current_time == o.time


@pytest.mark.parametrize("name", list(CASES))
def test_stencil_against_oracle(oracle_runs, name):
    """The merged stencil (the library default): FP64 traces and final fields
    within 1e-10 relative L2 of the oracle, FP32 traces within 1e-5."""
    case = CASES[name]
    print("stencil case", name, flush=True)
    o = oracle_runs(name)
    g = build("gpu", case, "f64", "stencil")
    g.run(case["steps"])
    err = rel_l2(g.traces(), o.traces())
    ferr = max(rel_l2(g.field(f, slot), o.field(f, slot)) for f in range(nfields(case)) for slot in (0, 1))
    f32 = build("gpu", case, "f32", "auto")
    f32.run(case["steps"])
    e32 = rel_l2(f32.traces(), o.traces())
    print("stencil %s: fp64 traces %.2e fields %.2e; fp32 %.2e" % (name, err, ferr, e32), flush=True)
    assert np.abs(o.traces()).max() > 0
    assert err <= 1e-10 and ferr <= 1e-10
    assert e32 <= 1e-5


@pytest.mark.parametrize("name", ["fem-ofem0", "abm"])
def test_fp32_traces_close_to_fp64_oracle(oracle_runs, name):
    case = CASES[name]
    o = oracle_runs(name)
    g = build("gpu", case, "f32")
    g.run(case["steps"])
    err = rel_l2(g.traces(), o.traces())
    print("fp32 trace rel L2 %s: %.3e" % (name, err), flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("name", ["fem-ofem0", "abm"])
def test_two_runs_and_state_round_trip(name):
    """A second run() continues the state and accumulates into the resized
    trace rows as the reference does; set_field then run continues bitwise."""
    from oracle.fem1 import OracleFem1  # noqa: F401
    case = CASES[name]
    o = build("oracle", case)
    g = build("gpu", case)
    for n in (17, 9):
        o.run(n)
        g.run(n)
        assert np.array_equal(g.traces(), o.traces())
    rng = np.random.default_rng(7)
    for f in range(nfields(case)):
        for slot in (0, 1):
            v = rng.standard_normal(g.size(f)) * 1e-3
            g.set_field(f, slot, v)
            o.set_field(f, slot, v)
            assert np.array_equal(g.field(f, slot), v)
    o.run(11)
    g.run(11)
    for f in range(nfields(case)):
        assert np.array_equal(g.field(f, 1), o.field(f, 1))


def test_save_results_formats(tmp_path):
    case = CASES["fem-ofem0"]
    g = build("gpu", case)
    g.run(10)
    path = str(tmp_path / "fem_0.3_1.csv")
    g.saveResults(10, path)
    lines = open(path).read().splitlines()
    assert lines[0] == "time,Ex,Ey,Ez" and len(lines) == 12
    assert lines[1].startswith("0,0,0,0")
    tr = g.traces()
    vals = [float(x) for x in lines[-1].split(",")]
    assert vals[0] == pytest.approx(10 * case["dt"], rel=1e-5)
    np.testing.assert_allclose(vals[1:], tr[-1, -1], rtol=1e-5)  # default 6 significant digits
    a = build("gpu", CASES["abm"])
    a.run(4)
    a.saveResults(4, str(tmp_path))
    names = sorted(os.listdir(tmp_path))
    assert "results_0_0.000000_0.020000_4.csv" in names and "results_1_0.000000_0.020000_4.csv" in names


def test_invalid_arguments():
    from emsolvers_amd._lib import EmsError
    from emsolvers_amd.fem1 import FemSimulation, AbmSimulation
    with pytest.raises(EmsError):
        FemSimulation((0, 4, 4), 0.02, 1e-11, EPS0, MU0, 1, 0.3, 0)
    with pytest.raises(EmsError):
        FemSimulation((4, 4, 4), 0.02, 1e-11, EPS0, MU0, 1, 0.3, 2)
    s = FemSimulation((4, 4, 4), 0.02, 1e-11, EPS0, MU0, 1, 0.3, 0)
    with pytest.raises(EmsError):
        s.setSourceDipole_x((0, 3, 0))  # pattern reaches j + 2 = 5 > 4
    with pytest.raises(EmsError):
        s.setSource_x((4, 0, 0))  # x-edges have i < nx
    with pytest.raises(EmsError):
        s.setObservationPoint([[], [], [(0, 0, 4)]])  # z-edges have k < nz
    with pytest.raises(EmsError):
        s.field(3)  # no H in src/fem
    with pytest.raises(EmsError):
        s.set_element_product(2)  # no matrix-core path at first order
    a = AbmSimulation((4, 4, 4), 0.02, 1e-11, EPS0, MU0, 1)
    assert a.field(3).shape == (4 * 5 * 5,)
