"""Closed-form pins of the hex FEM oracles' constants (CPU only).

The reference FEM solvers cannot be compiled here (Eigen and MPI are absent),
so the oracles' constants are pinned by what the mathematics fixes (helpers
in tests/fem_pins.py):
  - the PEC-cube TM110 eigenvalue of m U'' = -K U converges to
    2 (pi c / L)^2 for orders 1, 2, 3 (operator scale 4/((h p)^2 mu eps) and
    mass 8/p^3);
  - the discrete Gauss law of a point source (g^T K = 0 for a discrete
    gradient g) fixes every stepper's source amplitude and sign exactly, and
    the convergence order of g.E(t) to its closed form fixes the source
    evaluation times (orders 2 for leapfrog and symplectic_2, 4 for
    symplectic_4, 3 for the AB2/AM3 pair);
  - src/fem_abm's predictor and corrector weights, read off linear probes,
    equal the Adams-Bashforth-2 and Adams-Moulton-3 weights obtained by
    integrating the Lagrange interpolants.
tests/test_oracle_mutants.py rebuilds the oracle with each of these constants
changed and shows that the corresponding check fails.
"""
import numpy as np
import pytest

from tests import fem_pins as fp


@pytest.mark.parametrize("order,grids", [(1, (4, 8)), (3, (9, 12))])
def test_cavity_eigenvalue_converges(order, grids):
    """Orders 1 and 3: omega^2 of the TM110 triplet converges to the exact
    value (order 1 at second order)."""
    e = [abs(fp.cavity_eigen_error(order, n)) for n in grids]
    assert e[1] < (0.025 if order == 1 else 0.005), e
    assert e[1] < e[0] / (3.0 if order == 1 else 1.2), e


def test_cavity_eigenvalue_order_2():
    """Order 2: the reference's kmat/2.csv operator (element nullity 10, not
    the 26 gradients of a consistent order-2 element) converges to omega^2
    about 7.8 % above the exact value (1.034, 1.058, 1.067, 1.071, 1.073 at
    N = 8, 12, 16, 20, 24, measured on this oracle).  A wrong scale or mass
    (a factor p^2 = 4 or p^3 = 8) is far outside the band asserted here."""
    e = [fp.cavity_eigen_error(2, n, shift=1.05) for n in (8, 12)]
    assert 0.02 < e[0] < e[1] < 0.09, e


@pytest.mark.parametrize("stepper,order,rate", [
    ("leapfrog", 1, 4.0), ("leapfrog", 2, 4.0), ("leapfrog", 3, 4.0),
    ("symplectic_2", 1, 4.0), ("symplectic_2", 2, 4.0), ("symplectic_2", 3, 4.0),
    ("symplectic_4", 1, 16.0), ("symplectic_4", 3, 16.0)])
def test_point_source_gauss_law(stepper, order, rate):
    from oracle import fem as om
    st = {"leapfrog": om.LEAPFROG, "symplectic_2": om.SYMPLECTIC2, "symplectic_4": om.SYMPLECTIC4}[stepper]
    e = fp.femv2_charge_errors(st, order, (40, 80))
    assert e[0] < 0.1, e  # amplitude and sign: a wrong factor gives an O(1) error
    assert 0.75 * rate < e[0] / e[1] < 1.3 * rate, e


@pytest.mark.parametrize("stepper,rate", [("fem", 4.0), ("fem_abm", 8.0)])
def test_fem1_point_source_gauss_law(stepper, rate):
    from oracle import fem1
    e = fp.fem1_charge_errors(fem1.FEM if stepper == "fem" else fem1.ABM, (40, 80))
    assert e[0] < 0.1, e
    assert 0.75 * rate < e[0] / e[1] < 1.3 * rate, e


def test_abm_weights_are_adams_weights():
    got = np.array(fp.abm_step_weights())
    want = np.concatenate([fp.adams_weights([0, -1]), fp.adams_weights([1, 0, -1])])
    np.testing.assert_allclose(got[5:], want, rtol=1e-9, atol=1e-12)  # H update
    got = got[:5]  # E update
    np.testing.assert_allclose(want, [1.5, -0.5, 5 / 12, 8 / 12, -1 / 12], rtol=1e-14)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_leapfrog_operator_is_dt2_over_mass_times_symplectic(order):
    """fem_v2 scales the element matrix by 4 dt^2/((h p)^2 mu eps (8/p^3))
    (src/fem_v2/fem_simulation.cpp:476), symplectic by 4/((h p)^2 mu eps)
    (src/symplectic_2/fem_simulation.cpp:473): the leapfrog operator is
    dt^2 / m times the one the cavity test pins."""
    from oracle import fem as om
    dt = 3e-12
    N = 2 * order
    mats = []
    for st in (om.LEAPFROG, om.SYMPLECTIC2):
        o = om.OracleFEMv2(order, (N, N, N), 1e-2, dt, fp.EPS0, fp.MU0, stepper=st)
        mats.append(np.ctypeslib.as_array(om.lib().ems_oracle_femv2_matrix(o._h), shape=(o.n * o.n,)).copy())
    np.testing.assert_allclose(mats[0], mats[1] * dt * dt / (8.0 / order ** 3), rtol=1e-14, atol=0)
