"""Hex edge-element operators generated from the basis (no shipped tables).

The order-p hex FEM of src/fem_v2 reads its reference-element operator from
kmat/<p>.csv (use_ofem = 1) or kmat/lamped<p>.csv (use_ofem = 0)
(src/fem_v2/fem_simulation.cpp:458-490) and scales it by
4 dt^2 / ((h p)^2 mu eps (8 / p^3)).  At order 1 both tables follow from the
lowest-order Nedelec (first kind) element on the reference cube [-1, 1]^3:

    N_(c, a, b) = (1 + s_a t1)(1 + s_b t2) / 4  e_c,   s_a = 2a - 1,

for component c and (t1, t2) its two transverse coordinates in ascending
axis order, local index 4c + a + 2b, so that the tangential value is 1 on
the element's own edge and 0 on the others.  The operator is the curl-curl
Gram matrix K_ij = integral of curl N_i . curl N_j over the cube, each 1-D
factor of the integrand evaluated with the symmetric two-point rule at
+-sqrt(r) (weights 1):

    r = 1/3  (Gauss: exact)                 -> kmat/lamped1.csv
    r = 3/4  (the modified integration rule) -> kmat/1.csv

Both are generated here in exact rational arithmetic and written in the
tables' "a/b" format; they equal the shipped files byte for byte.

Orders 2 and 3 are not generated: the shipped kmat/2.csv and 3.csv factor
exactly into 1-D tensor products (the x-x block is A (x) (S (x) M + M (x) S)
with A acting along the edge), but their mass-like factor M has a positive
corner entry, which no nodal Lagrange basis on the element's fine nodes
integrated with positive weights produces; the tables stay the input for
those orders (tests/test_kmat.py checks both facts).
"""
from fractions import Fraction

import numpy as np

# two-point rule parameter r (nodes at +-sqrt(r)) per use_ofem
RULE = {0: Fraction(1, 3), 1: Fraction(3, 4)}


def _mul(p, q):
    out = {}
    for ea, ca in p.items():
        for eb, cb in q.items():
            e = (ea[0] + eb[0], ea[1] + eb[1], ea[2] + eb[2])
            out[e] = out.get(e, 0) + ca * cb
    return {e: c for e, c in out.items() if c != 0}


def _add(p, q, sign=1):
    out = dict(p)
    for e, c in q.items():
        out[e] = out.get(e, 0) + sign * c
    return {e: c for e, c in out.items() if c != 0}


def _diff(p, axis):
    out = {}
    for e, c in p.items():
        if e[axis] > 0:
            f = list(e)
            f[axis] -= 1
            out[tuple(f)] = out.get(tuple(f), 0) + c * e[axis]
    return out


def _half_linear(axis, s):
    """(1 + s x_axis) / 2 as a polynomial {exponents: coefficient}."""
    e = [0, 0, 0]
    e[axis] = 1
    return {(0, 0, 0): Fraction(1, 2), tuple(e): Fraction(s, 2)}


def _basis():
    out = []
    for c in range(3):
        t1, t2 = [a for a in range(3) if a != c]
        for b in range(2):
            for a in range(2):
                f = _mul(_half_linear(t1, 2 * a - 1), _half_linear(t2, 2 * b - 1))
                v = [{}, {}, {}]
                v[c] = f
                out.append(v)
    return out


def _curl(v):
    return [_add(_diff(v[2], 1), _diff(v[1], 2), -1),
            _add(_diff(v[0], 2), _diff(v[2], 0), -1),
            _add(_diff(v[1], 0), _diff(v[0], 1), -1)]


def _rule_1d(k, r):
    """Two-point rule at +-sqrt(r), weights 1, applied to x^k on [-1, 1]."""
    return Fraction(0) if k % 2 else 2 * r ** (k // 2)


def _integrate(p, r):
    return sum(c * _rule_1d(e[0], r) * _rule_1d(e[1], r) * _rule_1d(e[2], r) for e, c in p.items())


def hex_order1(use_ofem):
    """12 x 12 reference-element operator of order 1 as exact fractions:
    kmat/1.csv (use_ofem = 1) or kmat/lamped1.csv (use_ofem = 0)."""
    r = RULE[1 if use_ofem else 0]
    curls = [_curl(v) for v in _basis()]
    n = len(curls)
    k = [[Fraction(0)] * n for _ in range(n)]
    for i in range(n):
        for j in range(i, n):
            g = {}
            for c in range(3):
                g = _add(g, _mul(curls[i][c], curls[j][c]))
            k[i][j] = k[j][i] = _integrate(g, r)
    return k


def csv_text(matrix):
    """The kmat table format: rows of quoted "a/b" entries ("a" for integers)."""
    def fmt(x):
        x = Fraction(x)
        return '"%d"' % x.numerator if x.denominator == 1 else '"%d/%d"' % (x.numerator, x.denominator)
    return "".join(",".join(fmt(x) for x in row) + "\n" for row in matrix)


def element_operator(order, use_ofem):
    """Unscaled reference-element operator as float64 (what the FEM reads from
    kmat/), generated for order 1; orders 2 and 3 come from their tables."""
    if order != 1:
        raise NotImplementedError("orders 2 and 3: read kmat/<p>.csv (see the module docstring)")
    return np.array([[float(x) for x in row] for row in hex_order1(use_ofem)], dtype=np.float64)
