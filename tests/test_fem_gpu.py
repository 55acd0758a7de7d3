"""GPU parity of the structured-hex FEM (fem_v2 leapfrog, symplectic_2,
symplectic_4; orders 1-3) against the FP64 oracle on the same inputs.

Bar: FP64 bit-exact with the element products in the reference's order
('rows': the owner-gather kernels visit each DOF's elements in the serial
reference's order and sum every element row from 0 in the same order);
1e-10 relative L2 for the order-1 stencil and the order-2/3 matrix-core
products; FP32 traces within 1e-5 relative L2 of the FP64 oracle.
"""
import numpy as np
import pytest

from tests.conftest import rel_l2, tf_misfits

pytestmark = pytest.mark.gpu

EPS0 = 8.854187817e-12
MU0 = 1.25663706212e-6
C0 = 1.0 / (EPS0 * MU0) ** 0.5

# reference time step: h / (c sqrt(3)) / 2 (src/fem_v2/main.cpp:101)
CASES = {
    (1, 0): dict(grid=(8, 6, 10), steps=70, src=(4, 3, 5), obs=[(4, 3, 6), (2, 2, 2)], tf=1),
    (1, 1): dict(grid=(8, 6, 10), steps=70, src=(4, 3, 5), obs=[(5, 3, 5)], tf=2),
    (1, 2): dict(grid=(6, 8, 6), steps=50, src=(3, 4, 3), obs=[(3, 5, 3)], tf=1),
    (2, 0): dict(grid=(8, 8, 6), steps=50, src=(4, 4, 3), obs=[(4, 5, 3), (3, 3, 2)], tf=1),
    (2, 1): dict(grid=(8, 6, 8), steps=50, src=(4, 3, 4), obs=[(5, 3, 4)], tf=3),
    (2, 2): dict(grid=(6, 6, 8), steps=40, src=(3, 3, 4), obs=[(3, 3, 5)], tf=1),
    (3, 0): dict(grid=(9, 6, 6), steps=40, src=(4, 3, 3), obs=[(5, 3, 3)], tf=1),
    (3, 1): dict(grid=(6, 9, 6), steps=40, src=(3, 4, 3), obs=[(3, 5, 3)], tf=1),
    (3, 2): dict(grid=(6, 6, 9), steps=30, src=(3, 3, 4), obs=[(3, 3, 5)], tf=2),
}


def build(cls, order, stepper, case, precision=None, h=1e-2):
    dt = h / (C0 * np.sqrt(3.0)) / 2.0
    if cls == "oracle":
        from oracle.fem import OracleFEMv2
        s = OracleFEMv2(order, case["grid"], h, dt, EPS0, MU0, tf=case["tf"], stepper=stepper)
    else:
        from emsolvers_amd.fem import FemSimulation
        from oracle.fem import load_kmat, kmat_path
        s = FemSimulation(order, case["grid"], h, dt, EPS0, MU0, case["tf"], 1, stepper=stepper,
                          precision=precision, kmat=load_kmat(kmat_path(order)))
    s.setSource_z(case["src"])
    for o in case["obs"]:
        s.setObservationPoint(o)
    return s


@pytest.fixture(scope="module")
def oracle_cache():
    cache = {}

    def get(key):
        if key not in cache:
            order, stepper = key
            case = CASES[key]
            o = build("oracle", order, stepper, case)
            o.run(case["steps"])
            cache[key] = o
        return cache[key]
    return get


@pytest.mark.parametrize("rows,planes", [(0, 8), (4, 1), (4, 3), (8, 5), (8, 8), (16, 3)])
@pytest.mark.parametrize("stepper", [0, 1, 2])
def test_first_order_tilings_bitwise(oracle_cache, stepper, rows, planes):
    """First order: the z-marching all-component kernel (64 x rows workgroups,
    any number of planes per workgroup; 16 rows falls back to 8 for the row
    dots) and the per-component gather kernels (rows = 0) give the oracle's
    bits."""
    key = (1, stepper)
    case = CASES[key]
    print("tiling", key, rows, planes, flush=True)
    o = oracle_cache(key)
    g = build("gpu", 1, stepper, case, "f64")
    g.set_tiling(rows, planes)
    g.set_element_product("rows")  # the reference's sums (the default at order 1 is the stencil)
    g.run(case["steps"])
    assert np.array_equal(g.traces(), o.traces())
    for c in range(3):
        for slot in (0, 1):
            assert np.array_equal(g.get(c, slot), o.get(c, slot)), (c, slot)


@pytest.mark.parametrize("stepper", [0, 2])
def test_first_order_wide_grid_bitwise(stepper):
    """A grid wider than one 64-column tile and taller than one row block,
    with the source and receiver near tile edges."""
    case = dict(grid=(70, 11, 9), steps=25, src=(64, 8, 4), obs=[(63, 8, 4), (65, 3, 4)], tf=1)
    print("wide grid", stepper, flush=True)
    o = build("oracle", 1, stepper, case)
    o.run(case["steps"])
    for rows in (4, 8):
        g = build("gpu", 1, stepper, case, "f64")
        g.set_tiling(rows, 4)
        g.set_element_product("rows")
        g.run(case["steps"])
        assert np.array_equal(g.traces(), o.traces())
        for c in range(3):
            assert np.array_equal(g.get(c, 1), o.get(c, 1)), c


@pytest.mark.parametrize("key", list(CASES), ids=lambda k: "p%d-s%d" % k)
def test_fp64_bitwise_with_oracle(oracle_cache, key):
    order, stepper = key
    case = CASES[key]
    print("case order", order, "stepper", stepper, flush=True)
    o = oracle_cache(key)
    g = build("gpu", order, stepper, case, "f64")
    g.set_element_product("rows")  # the reference's sums (the defaults at orders 1 and 3 are not)
    g.run(case["steps"])
    to, tg = o.traces(), g.traces()
    assert to.shape == tg.shape
    assert np.abs(to[..., 1:]).max() > 0
    assert np.array_equal(tg, to)
    for c in range(3):
        for slot in (0, 1):
            assert np.array_equal(g.get(c, slot), o.get(c, slot)), (c, slot)
    assert g.current_time == o.current_time


@pytest.mark.parametrize("key", [k for k in CASES if k[0] >= 2], ids=lambda k: "p%d-s%d" % k)
def test_mfma_element_products_against_oracle(oracle_cache, key):
    """Orders 2 and 3 with the element products on matrix cores: FP64 traces
    and final fields within 1e-10 relative L2 of the oracle (the MFMA sums
    run in another order than the reference's), FP32 traces within 1e-5 of
    the FP64 oracle and within 1e-5 of the FP32 per-element-thread run."""
    order, stepper = key
    case = CASES[key]
    print("case order", order, "stepper", stepper, flush=True)
    o = oracle_cache(key)
    g = build("gpu", order, stepper, case, "f64")
    g.set_element_product("mfma")
    g.run(case["steps"])
    err = rel_l2(g.traces()[..., 1:], o.traces()[..., 1:])
    ferr = max(rel_l2(g.get(c, slot), o.get(c, slot)) for c in range(3) for slot in (0, 1))
    f32 = build("gpu", order, stepper, case, "f32")
    f32.set_element_product("mfma")
    f32.run(case["steps"])
    r32 = build("gpu", order, stepper, case, "f32")
    r32.set_element_product("rows")
    r32.run(case["steps"])
    e32 = rel_l2(f32.traces()[..., 1:], o.traces()[..., 1:])
    e32r = rel_l2(f32.traces()[..., 1:], r32.traces()[..., 1:])
    print("mfma p%d s%d: fp64 traces %.2e fields %.2e; fp32 vs oracle %.2e vs rows %.2e" % (order, stepper, err, ferr,
                                                                                         e32, e32r), flush=True)
    assert err <= 1e-10 and ferr <= 1e-10
    assert e32 <= 1e-5 and e32r <= 1e-5


@pytest.mark.parametrize("rows,planes", [(4, 3), (8, 5), (16, 3)])
@pytest.mark.parametrize("stepper", [0, 1, 2])
def test_first_order_stencil_against_oracle(oracle_cache, stepper, rows, planes):
    """First order with each component's 4 element rows merged into one
    33-point stencil: FP64 traces and fields within 1e-10 relative L2 of the
    oracle (the stencil sums in another order), FP32 traces within 1e-5 of
    the FP64 oracle; also on a grid wider than one tile."""
    key = (1, stepper)
    case = CASES[key]
    print("stencil", key, rows, planes, flush=True)
    o = oracle_cache(key)
    g = build("gpu", 1, stepper, case, "f64")
    g.set_tiling(rows, planes)
    g.set_element_product("stencil")
    g.run(case["steps"])
    err = rel_l2(g.traces()[..., 1:], o.traces()[..., 1:])
    ferr = max(rel_l2(g.get(c, slot), o.get(c, slot)) for c in range(3) for slot in (0, 1))
    f32 = build("gpu", 1, stepper, case, "f32")
    f32.set_tiling(rows, planes)
    f32.set_element_product("stencil")
    f32.run(case["steps"])
    e32 = rel_l2(f32.traces()[..., 1:], o.traces()[..., 1:])
    print("stencil s%d: fp64 traces %.2e fields %.2e; fp32 %.2e" % (stepper, err, ferr, e32), flush=True)
    assert np.abs(o.traces()[..., 1:]).max() > 0
    assert err <= 1e-10 and ferr <= 1e-10
    assert e32 <= 1e-5
    if rows >= 8:
        wide = dict(grid=(70, 11, 9), steps=25, src=(64, 8, 4), obs=[(63, 8, 4), (65, 3, 4)], tf=1)
        ow = build("oracle", 1, stepper, wide)
        ow.run(wide["steps"])
        gw = build("gpu", 1, stepper, wide, "f64")
        gw.set_tiling(rows, 4)
        gw.set_element_product("stencil")
        gw.run(wide["steps"])
        assert rel_l2(gw.traces()[..., 1:], ow.traces()[..., 1:]) <= 1e-10
        for c in range(3):
            assert rel_l2(gw.get(c, 1), ow.get(c, 1)) <= 1e-10, c
        fw = build("gpu", 1, stepper, wide, "f32")  # 16 rows run in FP32 (one partial row block here)
        fw.set_tiling(rows, 4)
        fw.run(wide["steps"])
        assert rel_l2(fw.traces()[..., 1:], ow.traces()[..., 1:]) <= 1e-5


def test_stencil_is_first_order_only():
    from emsolvers_amd import _lib
    g = build("gpu", 2, 0, CASES[(2, 0)], "f64")
    with pytest.raises(_lib.EmsError):
        g.set_element_product("stencil")


@pytest.mark.parametrize("key", list(CASES), ids=lambda k: "p%d-s%d" % k)
def test_fp32_error_against_fp64_oracle(oracle_cache, key):
    order, stepper = key
    case = CASES[key]
    print("case order", order, "stepper", stepper, flush=True)
    o = oracle_cache(key)
    g = build("gpu", order, stepper, case, "f32")
    g.run(case["steps"])
    err = rel_l2(g.traces()[..., 1:], o.traces()[..., 1:])
    print("fp32 trace rel L2 p%d s%d: %.3e" % (order, stepper, err))
    assert err <= 1e-5
    dt = g.traces()[0, 1, 0] - g.traces()[0, 0, 0]
    em, pm = tf_misfits(g.traces()[..., 1:], o.traces()[..., 1:], dt)
    print("fp32 TF envelope / phase misfit p%d s%d: EM %.2e PM %.2e" % (order, stepper, em, pm))
    assert em <= 1e-4 and pm <= 1e-4
    assert np.array_equal(g.traces()[..., 0], o.traces()[..., 0])


@pytest.mark.parametrize("stepper", [0, 1, 2])
def test_state_continuation_from_random_state(stepper):
    """Both sides start from the same random interior state, with no source."""
    from oracle.fem import OracleFEMv2, boundary_mask
    from emsolvers_amd.fem import FemSimulation
    from oracle.fem import load_kmat, kmat_path
    grid, h = (6, 8, 6), 1e-2
    dt = h / (C0 * np.sqrt(3.0)) / 2.0
    o = OracleFEMv2(2, grid, h, dt, EPS0, MU0, stepper=stepper)
    g = FemSimulation(2, grid, h, dt, EPS0, MU0, 1, 1, stepper=stepper, kmat=load_kmat(kmat_path(2)))
    rng = np.random.default_rng(stepper)
    mask = boundary_mask(grid)
    for slot in (0, 1):
        v = rng.standard_normal(o.ndof())
        v[mask] = 0
        o.set_state(slot, v)
        g.set_state(slot, v)
    o.run(25)
    g.run(25)
    for slot in (0, 1):
        assert np.array_equal(g.get_state(slot), o.get_state(slot))


def test_kmat_loader_matches_python_parser():
    from emsolvers_amd.fem import load_kmat as lib_load
    from oracle.fem import load_kmat, kmat_path
    for order in (1, 2, 3):
        assert np.array_equal(lib_load(kmat_path(order)), load_kmat(kmat_path(order)))
    assert np.array_equal(lib_load(kmat_path(1, 0)), load_kmat(kmat_path(1, 0)))


def test_fem_error_paths(tmp_path):
    from emsolvers_amd import _lib
    from emsolvers_amd.fem import FemSimulation, load_kmat
    from oracle.fem import load_kmat as ol, kmat_path
    K = ol(kmat_path(2))
    with pytest.raises(_lib.EmsError):
        load_kmat(str(tmp_path / "missing.csv"))
    with pytest.raises(_lib.EmsError):  # not a multiple of the order
        FemSimulation(2, (7, 8, 8), 1e-2, 1e-12, EPS0, MU0, 1, 1, kmat=K)
    with pytest.raises(_lib.EmsError):  # one element along z
        FemSimulation(2, (8, 8, 2), 1e-2, 1e-12, EPS0, MU0, 1, 1, kmat=K)
    s = FemSimulation(2, (8, 8, 8), 1e-2, 1e-12, EPS0, MU0, 1, 1, kmat=K)
    with pytest.raises(_lib.EmsError):
        s.setObservationPoint((0, 0, 0))  # reads outside the Ex array
    with pytest.raises(_lib.EmsError):
        s.setSource_z((0, 0, 8))


def test_save_results_format(tmp_path, oracle_cache):
    from emsolvers_amd.fem import FemSimulation
    case = CASES[(1, 0)]
    g = build("gpu", 1, 0, case, "f64")
    g.run(case["steps"])
    fn = str(tmp_path / "out.csv")
    g.saveResults(case["steps"], fn)
    lines = open(fn).read().splitlines()
    assert lines[0] == "time,Ex,Ey,Ez"
    assert len(lines) == 1 + g.traces().shape[1]
    assert "e" in lines[5].split(",")[1]  # scientific, 15 digits
    assert len(lines[5].split(",")[3].split("e")[0].replace("-", "").replace(".", "")) == 16
