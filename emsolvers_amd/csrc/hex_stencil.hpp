// Merged first-order hex edge-element stencil, shared by the fem_v2 family
// (k_fem1z) and the stabilized src/fem, src/fem_abm kernels.
//
// An edge DOF of component C whose two transverse coordinates are both
// strictly inside the grid lies in exactly 4 elements, and with one element
// operator for all elements the sum of its 4 element row dots is a fixed
// 33-point stencil: 9 entries of array C in the plane across the edge and 12
// of each other array.  Term t of component C (a1 < a2 the axes transverse
// to C, which are also the other components):
//   t <  9: array C at offsets t % 3 - 1 along a1, t / 3 - 1 along a2;
//   t >= 9: array q = a1 (t < 21) or a2, u = (t - 9) % 12: u & 1 along C,
//           (u >> 1 & 1) - 1 along q, (u >> 2) - 1 along the remaining axis.
// Offsets are relative to the DOF's own (X, Y, Z) in each array's own index
// space (x fastest).
#pragma once

#include <hip/hip_runtime.h>

namespace ems {

struct Z1Term {
    int q, d[3];
};

constexpr int kZ1Terms = 33;

__host__ __device__ constexpr Z1Term z1_term(int C, int t)
{
    const int a1 = C == 0 ? 1 : 0, a2 = C == 2 ? 1 : 2;
    Z1Term r{C, {0, 0, 0}};
    if (t < 9) {
        r.d[a1] = t % 3 - 1;
        r.d[a2] = t / 3 - 1;
        return r;
    }
    const int q = t < 21 ? a1 : a2, third = q == a1 ? a2 : a1;
    const int u = (t - 9) % 12;
    r.q = q;
    r.d[C] = u & 1;
    r.d[q] = ((u >> 1) & 1) - 1;
    r.d[third] = (u >> 2) - 1;
    return r;
}

// Host: the coefficients c[3][33] from a 12 x 12 element operator M
// (row-major).  Element (i, j, k) holds the edge of array q at (i, j, k) plus
// offsets a, b along q's transverse axes (ascending) under local index
// 4 q + a + 2 b, or 4 q + 2 a + b for the y edges when swap_y (the src/fem
// numbering, fem_simulation.cpp:98-121); the DOF's own row in an element is
// its local index there.  Each element row entry is added to the term its
// (array, offset) maps to, in element order (lower elements along a1, then
// a2), in FP64.  Returns false if an entry maps to no term.
inline bool z1_stencil_coefs(const double *M, bool swap_y, double *c)
{
    auto local = [swap_y](int q, int a, int b) { return 4 * q + ((swap_y && q == 1) ? 2 * a + b : a + 2 * b); };
    for (int t = 0; t < 3 * kZ1Terms; ++t) c[t] = 0.0;
    for (int C = 0; C < 3; ++C) {
        const int a1 = C == 0 ? 1 : 0, a2 = C == 2 ? 1 : 2;
        for (int code = 0; code < 4; ++code) {
            int e[3] = {0, 0, 0};
            if (code & 1) e[a1] = -1;  // the element below along a1: the DOF at offset 1 there
            if (code & 2) e[a2] = -1;
            const int r = local(C, code & 1, (code >> 1) & 1);
            for (int q = 0; q < 3; ++q) {
                const int b1 = q == 0 ? 1 : 0, b2 = q == 2 ? 1 : 2;
                for (int b = 0; b < 2; ++b)
                    for (int a = 0; a < 2; ++a) {
                        int d[3] = {e[0], e[1], e[2]};
                        d[b1] += a;
                        d[b2] += b;
                        int t = 0;
                        for (; t < kZ1Terms; ++t) {
                            const Z1Term z = z1_term(C, t);
                            if (z.q == q && z.d[0] == d[0] && z.d[1] == d[1] && z.d[2] == d[2]) break;
                        }
                        if (t == kZ1Terms) return false;
                        c[C * kZ1Terms + t] += M[r * 12 + local(q, a, b)];
                    }
            }
        }
    }
    return true;
}

}  // namespace ems
