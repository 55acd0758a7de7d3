// Hierarchical Nedelec (first kind) bases on a tetrahedron and their exact
// element matrices.  Host code of the tetrahedral edge-element solver.
//
// Every basis function is a sum of terms c * lambda^a * grad(lambda_g) in
// the barycentric coordinates lambda_0..3 of the tetrahedron (local vertices
// in increasing global order, so that shared edges and faces see the same
// local vertex order from both sides):
//   order 1 (6):  w_ab = l_a grad l_b - l_b grad l_a          (Whitney, edges)
//   order 2 (20): w_ab (6), g_ab = grad(l_a l_b) (6), and per face (a, b, c)
//                 f1 = l_a w_bc, f2 = l_b w_ac (8)
// The order-2 set spans the first-kind Nedelec space of degree 2 (complete
// linear plus the quadratic fields with x . p(x) = 0), dimension 20; its
// curl-free part is grad P2 (dimension 9).  Integrals of barycentric
// monomials are exact:  int_T l^a = 6 V a0! a1! a2! a3! / (|a| + 3)!.
#pragma once

#include <cmath>

#include "tet_mesh.hpp"

namespace ems {
namespace tet {

struct BTerm {
    double c;
    int a[4];  // exponents of lambda_0..3
    int g;     // gradient index
};
struct BFun {
    int n;
    BTerm t[2];
};

inline BTerm term(double c, int i, int j, int g)
{
    // c * lambda_i * lambda_j (j = -1: linear) * grad lambda_g
    BTerm t{c, {0, 0, 0, 0}, g};
    t.a[i] += 1;
    if (j >= 0) t.a[j] += 1;
    return t;
}

// basis of the given order in local dof order: order 1: 6 edge (Whitney);
// order 2: 6 Whitney, 6 edge gradients, 4 faces x 2
inline int basis(int order, BFun *f)
{
    int n = 0;
    for (int l = 0; l < 6; ++l) {
        const int a = kLocEdge[l][0], b = kLocEdge[l][1];
        f[n++] = BFun{2, {term(1.0, a, -1, b), term(-1.0, b, -1, a)}};
    }
    if (order == 1) return n;
    for (int l = 0; l < 6; ++l) {
        const int a = kLocEdge[l][0], b = kLocEdge[l][1];
        f[n++] = BFun{2, {term(1.0, a, -1, b), term(1.0, b, -1, a)}};
    }
    for (int q = 0; q < 4; ++q) {
        const int a = kLocTri[q][0], b = kLocTri[q][1], c = kLocTri[q][2];
        f[n++] = BFun{2, {term(1.0, a, b, c), term(-1.0, a, c, b)}};  // l_a w_bc
        f[n++] = BFun{2, {term(1.0, b, a, c), term(-1.0, b, c, a)}};  // l_b w_ac
    }
    return n;
}

inline int basis_size(int order) { return order == 1 ? 6 : 20; }

// int over the tet of l^a (volume V)
inline double mono_int(const int a[4], double V)
{
    static const double fact[16] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880, 3628800, 39916800,
                                    479001600, 6227020800.0, 87178291200.0, 1307674368000.0};
    const int s = a[0] + a[1] + a[2] + a[3];
    return 6.0 * V * fact[a[0]] * fact[a[1]] * fact[a[2]] * fact[a[3]] / fact[s + 3];
}

inline double dot3(const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; }
inline void cross3(const double *x, const double *y, double *z)
{
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
}

// Element mass M_lm = int w_l . w_m and curl-curl S_lm = int curl w_l . curl w_m
// (nl x nl, row-major) of tetrahedron type t scaled by h.
inline int element_matrices(int order, int t, double h, double *M, double *S)
{
    BFun f[20];
    const int nl = basis(order, f);
    double G[4][3], V;
    bary(t, h, G, V);
    // curl of c l^a grad l_g = sum_i c a_i l^(a - e_i) grad l_i x grad l_g
    struct CTerm {
        double c;
        int a[4];
        double v[3];
    };
    CTerm ct[20][8];
    int nc[20];
    for (int l = 0; l < nl; ++l) {
        nc[l] = 0;
        for (int s = 0; s < f[l].n; ++s) {
            const BTerm &b = f[l].t[s];
            for (int i = 0; i < 4; ++i) {
                if (!b.a[i]) continue;
                CTerm c{b.c * b.a[i], {b.a[0], b.a[1], b.a[2], b.a[3]}, {0, 0, 0}};
                c.a[i] -= 1;
                cross3(G[i], G[b.g], c.v);
                ct[l][nc[l]++] = c;
            }
        }
    }
    for (int l = 0; l < nl; ++l)
        for (int m = 0; m < nl; ++m) {
            double mm = 0.0, ss = 0.0;
            for (int s = 0; s < f[l].n; ++s)
                for (int r = 0; r < f[m].n; ++r) {
                    const BTerm &x = f[l].t[s], &y = f[m].t[r];
                    const int a[4] = {x.a[0] + y.a[0], x.a[1] + y.a[1], x.a[2] + y.a[2], x.a[3] + y.a[3]};
                    mm += x.c * y.c * dot3(G[x.g], G[y.g]) * mono_int(a, V);
                }
            for (int s = 0; s < nc[l]; ++s)
                for (int r = 0; r < nc[m]; ++r) {
                    const CTerm &x = ct[l][s], &y = ct[m][r];
                    const int a[4] = {x.a[0] + y.a[0], x.a[1] + y.a[1], x.a[2] + y.a[2], x.a[3] + y.a[3]};
                    ss += x.c * y.c * dot3(x.v, y.v) * mono_int(a, V);
                }
            M[l * nl + m] = mm;
            S[l * nl + m] = ss;
        }
    return nl;
}

// First order: the curls of the 6 Whitney functions of type t are constant,
// curl w_ab = 2 grad l_a x grad l_b; C[k][l] is component k of the curl of
// local function l, so that the element curl-curl matrix is S = V C^T C.
inline void whitney_curls(int t, double h, double C[3][6], double &V)
{
    double G[4][3];
    bary(t, h, G, V);
    for (int l = 0; l < 6; ++l) {
        double c[3];
        cross3(G[kLocEdge[l][0]], G[kLocEdge[l][1]], c);
        for (int k = 0; k < 3; ++k) C[k][l] = 2.0 * c[k];
    }
}

// values of the nl basis functions at barycentric point lam
inline int basis_at(int order, int t, double h, const double lam[4], double (*w)[3])
{
    BFun f[20];
    const int nl = basis(order, f);
    double G[4][3], V;
    bary(t, h, G, V);
    for (int l = 0; l < nl; ++l) {
        w[l][0] = w[l][1] = w[l][2] = 0.0;
        for (int s = 0; s < f[l].n; ++s) {
            const BTerm &b = f[l].t[s];
            double p = b.c;
            for (int i = 0; i < 4; ++i)
                for (int e = 0; e < b.a[i]; ++e) p *= lam[i];
            for (int c = 0; c < 3; ++c) w[l][c] += p * G[b.g][c];
        }
    }
    return nl;
}

}  // namespace tet
}  // namespace ems
