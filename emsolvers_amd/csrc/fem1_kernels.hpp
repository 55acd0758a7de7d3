// First-order stabilized hex edge-element kernels for gfx950: the steppers
// of src/fem (leapfrog with K = c C_s B_s^-1 A_s) and src/fem_abm
// (Adams-Bashforth-2 predictor / Adams-Moulton-3 corrector on co-located E
// and H).
//
// Layouts are the reference's (src/fem/fem_simulation.cpp:66-72), x fastest:
//   x-edges Nx(Ny+1)(Nz+1), y-edges (Nx+1)Ny(Nz+1), z-edges (Nx+1)(Ny+1)Nz.
// Element (i,j,k) owns the edges (src/fem/fem_simulation.cpp:98-121)
//   x: (i, j+b, k+c) local b + 2c      y: (i+a, j, k+c) local 4 + 2a + c
//   z: (i+a, j+b, k) local 8 + a + 2b
// The reference scatters each element's K x_e with atomics.  Here one thread
// owns one edge and gathers the (at most 4) element contributions in the
// order the reference's serial element loop (k, j, i) visits them, each row
// dot product summed from 0 over m in order.  With contraction off the FP64
// results equal a serial run of the reference bit for bit.  These gathers are
// the EMS_FEM_ELEM_ROWS path; by default the interior edges take the merged
// stencil of hex_stencil.hpp (fem1.hip: the fem leapfrog on k_fem1z, the
// fem_abm passes on k_fem1z_sum), equal to the reference to rounding.
#pragma once

#include <hip/hip_runtime.h>

#include "hex_stencil.hpp"
#include "lane_addr.hpp"

namespace ems {
namespace fem1 {

struct Hex1 {
    int Nx, Ny, Nz;
    long n[3];  // entries per component
};

template <typename T> struct F3 {
    T *c[3];
};
template <typename T> struct CF3 {
    const T *c[3];
};

// Row r of the 12 x 12 matrix M dotted with the 12 edge values of element
// (i, j, k), m = 0..11 in order, summed from 0.  UNI: the callers' j and k
// are wave-uniform (only i varies over the lanes), and in FP32 the addresses
// are taken as uniform bases plus i through lane_at (11 % faster at 128^3);
// in FP64 that form measured 35 % slower and the plain 64-bit addresses are
// kept.  Callers whose j or k vary over a wave pass UNI = false.
template <typename T, bool UNI = true>
__device__ __forceinline__ T elem_dot(const Hex1 &g, const T *__restrict__ M, int r, CF3<T> x, int i, int j, int k)
{
    const long Nx = g.Nx, Ny = g.Ny;
    const long sx = Nx * (Ny + 1);
    const long sy = (Nx + 1) * Ny;
    const T *__restrict__ kr = M + r * 12;
    const T *__restrict__ X = x.c[0];
    const T *__restrict__ Y = x.c[1];
    const T *__restrict__ Z = x.c[2];
    T t = 0;
    if constexpr (sizeof(T) == 4 && UNI) {
        const long bx = Nx * (j + (Ny + 1) * (long)k);
        const long by = (Nx + 1) * (j + Ny * (long)k);
        const long bz = (Nx + 1) * (j + (Ny + 1) * (long)k);
        const unsigned io = (unsigned)i * sizeof(T);
        t += kr[0] * lane_at(X + bx, io);
        t += kr[1] * lane_at(X + (bx + Nx), io);
        t += kr[2] * lane_at(X + (bx + sx), io);
        t += kr[3] * lane_at(X + (bx + sx + Nx), io);
        t += kr[4] * lane_at(Y + by, io);
        t += kr[5] * lane_at(Y + (by + sy), io);
        t += kr[6] * lane_at(Y + (by + 1), io);
        t += kr[7] * lane_at(Y + (by + 1 + sy), io);
        t += kr[8] * lane_at(Z + bz, io);
        t += kr[9] * lane_at(Z + (bz + 1), io);
        t += kr[10] * lane_at(Z + (bz + Nx + 1), io);
        t += kr[11] * lane_at(Z + (bz + Nx + 2), io);
        return t;
    }
    const long bx = i + Nx * (j + (Ny + 1) * (long)k);
    const long by = i + (Nx + 1) * (j + Ny * (long)k);
    const long bz = i + (Nx + 1) * (j + (Ny + 1) * (long)k);
    t += kr[0] * X[bx];
    t += kr[1] * X[bx + Nx];
    t += kr[2] * X[bx + sx];
    t += kr[3] * X[bx + sx + Nx];
    t += kr[4] * Y[by];
    t += kr[5] * Y[by + sy];
    t += kr[6] * Y[by + 1];
    t += kr[7] * Y[by + 1 + sy];
    t += kr[8] * Z[bz];
    t += kr[9] * Z[bz + 1];
    t += kr[10] * Z[bz + Nx + 1];
    t += kr[11] * Z[bz + Nx + 2];
    return t;
}

__device__ __forceinline__ bool pec1(const Hex1 &g, int comp, int X, int Y, int Z)
{
    if (comp == 0) return Y == 0 || Y == g.Ny || Z == 0 || Z == g.Nz;
    if (comp == 1) return X == 0 || X == g.Nx || Z == 0 || Z == g.Nz;
    return X == 0 || X == g.Nx || Y == 0 || Y == g.Ny;
}

// The (up to 4) elements of edge (COMP, X, Y, Z) in the serial visiting
// order, with the edge's local row.  For every component that order is
// (1,1), (0,1), (1,0), (0,0) in the two transverse element offsets, i.e.
// local rows 3, 2, 1, 0 (x), 3, 1, 2, 0 (y), 3, 2, 1, 0 (z) of the block.
// f(i, j, k, r) is called for each of them in that order (no candidate
// array: an array filled with a data-dependent count lives in scratch memory).
template <int COMP, typename F>
__device__ __forceinline__ void for_cands(const Hex1 &g, int X, int Y, int Z, F &&f)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int u = (q & 1) ? 0 : 1;  // inner offset: 1 then 0
        const int v = (q & 2) ? 0 : 1;  // outer offset: 1 then 0
        int i, j, k, r;
        if (COMP == 0) {  // u = b (y), v = c (z)
            i = X;
            j = Y - u;
            k = Z - v;
            r = u + 2 * v;
        } else if (COMP == 1) {  // u = a (x), v = c (z)
            i = X - u;
            j = Y;
            k = Z - v;
            r = 4 + 2 * u + v;
        } else {  // u = a (x), v = b (y)
            i = X - u;
            j = Y - v;
            k = Z;
            r = 8 + u + 2 * v;
        }
        if (i >= 0 && i < g.Nx && j >= 0 && j < g.Ny && k >= 0 && k < g.Nz) f(i, j, k, r);
    }
}

// Launch: workgroup 64 x 4, grid (ceil(dx/64), ceil(dy/4), dz) per component.
template <int COMP> __device__ __forceinline__ bool coords(const Hex1 &g, int &X, int &Y, int &Z, long &idx)
{
    const int dx = g.Nx + (COMP != 0), dy = g.Ny + (COMP != 1);
    X = blockIdx.x * 64 + threadIdx.x;
    Y = blockIdx.y * 4 + threadIdx.y;
    Z = blockIdx.z;
    if (X >= dx || Y >= dy) return false;
    idx = X + (long)dx * (Y + (long)dy * Z);
    return true;
}

// src/fem (fem_simulation.cpp:466-613, :743-805):
//   E2 = (0 - sum_e K_e E1) + (2 E1 - E0);  PEC-tangential edges -> 0
template <typename T, int COMP>
__global__ __launch_bounds__(256) void k_fem1_leap(Hex1 g, const T *__restrict__ K, CF3<T> e0, CF3<T> e1, F3<T> e2)
{
    int X, Y, Z;
    long idx;
    if (!coords<COMP>(g, X, Y, Z, idx)) return;
    if (pec1(g, COMP, X, Y, Z)) {
        e2.c[COMP][idx] = (T)0;
        return;
    }
    T acc = 0;
    for_cands<COMP>(g, X, Y, Z, [&](int i, int j, int k, int r) { acc -= elem_dot<T>(g, K, r, e1, i, j, k); });
    e2.c[COMP][idx] = acc + ((T)2 * e1.c[COMP][idx] - e0.c[COMP][idx]);
}

// src/fem_abm scratch fields (fem_simulation.cpp:294-315, :331-344, :576-618):
//   predictor  E3 = (3 H1 - H0) dt/8/eps/h,           H3 = -(3 E1 - E0) dt/8/mu/h
//   corrector  E3 = (5 Hp + 8 H1 - H0) dt/48/eps/h,   H3 = -(5 Ep + 8 E1 - E0) dt/48/mu/h
// 1-D launch over each component (blockIdx.y = component).
template <typename T, bool CORR>
__global__ __launch_bounds__(256) void k_fem1_scratch(Hex1 g, CF3<T> e0, CF3<T> e1, CF3<T> ep, CF3<T> h0,
                                                      CF3<T> h1, CF3<T> hp, F3<T> e3, F3<T> h3, T dt, T eps,
                                                      T mu, T h)
{
    const int q = blockIdx.y;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= g.n[q]) return;
    if (CORR) {
        e3.c[q][i] = ((T)5 * hp.c[q][i] + (T)8 * h1.c[q][i] - h0.c[q][i]) * dt / (T)48 / eps / h;
        h3.c[q][i] = -((T)5 * ep.c[q][i] + (T)8 * e1.c[q][i] - e0.c[q][i]) * dt / (T)48 / mu / h;
    } else {
        e3.c[q][i] = ((T)3 * h1.c[q][i] - h0.c[q][i]) * dt / (T)8 / eps / h;
        h3.c[q][i] = -((T)3 * e1.c[q][i] - e0.c[q][i]) * dt / (T)8 / mu / h;
    }
}

// src/fem_abm element passes (fem_simulation.cpp:319-369, :521-558, :619-710)
// followed by the "+= slot 1" pass:
//   Eo = (0 + sum_e K_e^T E3) + E1,   Ho = (0 + sum_e K_e H3) + H1
// Kt is K transposed (row-major), so both sides read rows.  PEC: when set,
// PEC-tangential E edges are written as 0 (the corrector; the predictor
// leaves them, as the reference applies the boundary condition only at the
// end of the step).
template <typename T, int COMP, bool PEC, bool UNI>
__device__ __forceinline__ void abm_edge(const Hex1 &g, const T *__restrict__ K, const T *__restrict__ Kt,
                                         CF3<T> e3, CF3<T> h3, CF3<T> e1, CF3<T> h1, F3<T> eo, F3<T> ho, int X,
                                         int Y, int Z, long idx)
{
    T ah = 0;
    for_cands<COMP>(g, X, Y, Z, [&](int i, int j, int k, int r) { ah += elem_dot<T, UNI>(g, K, r, h3, i, j, k); });
    ho.c[COMP][idx] = ah + h1.c[COMP][idx];
    if (PEC && pec1(g, COMP, X, Y, Z)) {
        eo.c[COMP][idx] = (T)0;
        return;
    }
    T ae = 0;
    for_cands<COMP>(g, X, Y, Z, [&](int i, int j, int k, int r) { ae += elem_dot<T, UNI>(g, Kt, r, e3, i, j, k); });
    eo.c[COMP][idx] = ae + e1.c[COMP][idx];
}

template <typename T, int COMP, bool PEC>
__global__ __launch_bounds__(256) void k_fem1_abm(Hex1 g, const T *__restrict__ K, const T *__restrict__ Kt,
                                                  CF3<T> e3, CF3<T> h3, CF3<T> e1, CF3<T> h1, F3<T> eo, F3<T> ho)
{
    int X, Y, Z;
    long idx;
    if (!coords<COMP>(g, X, Y, Z, idx)) return;
    abm_edge<T, COMP, PEC, true>(g, K, Kt, e3, h3, e1, h1, eo, ho, X, Y, Z, idx);
}

// Edges of component COMP on the grid's faces, i.e. with a transverse
// coordinate at 0 or N (the merged stencil covers the others), as two pairs
// of sheets: A, a2 = 0 or N_a2 with a1 in [0, N_a1]; B, a1 = 0 or N_a1 with
// a2 in [1, N_a2 - 1]; COMP runs over [0, N_COMP) on both.  Within a sheet
// the lower-numbered free axis is the fastest, so that x runs along the
// lanes wherever it is free.  Entry i of the N_COMP (2 (N_a1 + 1) +
// 2 (N_a2 - 1)) face edges.
template <int COMP>
__device__ __forceinline__ bool face_edge(const Hex1 &g, long i, int &X, int &Y, int &Z, long &idx)
{
    const int N[3] = {g.Nx, g.Ny, g.Nz};
    constexpr int a1 = COMP == 0 ? 1 : 0, a2 = COMP == 2 ? 1 : 2;
    const long nA = (long)N[COMP] * (N[a1] + 1), nB = (long)N[COMP] * (N[a2] - 1);
    int c[3];
    if (i < 2 * nA) {
        const int side = i >= nA;
        const long j = i - side * nA;
        constexpr int u = COMP < a1 ? COMP : a1, v = COMP < a1 ? a1 : COMP;
        const int nu = u == COMP ? N[COMP] : N[a1] + 1;
        c[u] = (int)(j % nu);
        c[v] = (int)(j / nu);
        c[a2] = side ? N[a2] : 0;
    } else if (i < 2 * nA + 2 * nB) {
        const long k = i - 2 * nA;
        const int side = k >= nB;
        const long j = k - side * nB;
        constexpr int u = COMP < a2 ? COMP : a2, v = COMP < a2 ? a2 : COMP;
        const int nu = u == COMP ? N[COMP] : N[a2] - 1;
        c[u] = (int)(j % nu);
        c[v] = (int)(j / nu);
        c[a2] += 1;
        c[a1] = side ? N[a1] : 0;
    } else {
        return false;
    }
    X = c[0];
    Y = c[1];
    Z = c[2];
    const int dx = g.Nx + (COMP != 0), dy = g.Ny + (COMP != 1);
    idx = X + (long)dx * (Y + (long)dy * Z);
    return true;
}

// The element passes above on the face edges only (blockIdx.y = component),
// next to the z-marching stencil pass of the interior edges; j and k vary
// over a wave here, so the FP32 row dots take plain addresses.
template <typename T, bool PEC>
__global__ __launch_bounds__(256) void k_fem1_abm_faces(Hex1 g, const T *__restrict__ K, const T *__restrict__ Kt,
                                                        CF3<T> e3, CF3<T> h3, CF3<T> e1, CF3<T> h1, F3<T> eo,
                                                        F3<T> ho)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int X, Y, Z;
    long idx;
    if (blockIdx.y == 0) {
        if (face_edge<0>(g, i, X, Y, Z, idx)) abm_edge<T, 0, PEC, false>(g, K, Kt, e3, h3, e1, h1, eo, ho, X, Y, Z, idx);
    } else if (blockIdx.y == 1) {
        if (face_edge<1>(g, i, X, Y, Z, idx)) abm_edge<T, 1, PEC, false>(g, K, Kt, e3, h3, e1, h1, eo, ho, X, Y, Z, idx);
    } else {
        if (face_edge<2>(g, i, X, Y, Z, idx)) abm_edge<T, 2, PEC, false>(g, K, Kt, e3, h3, e1, h1, eo, ho, X, Y, Z, idx);
    }
}

// One source update in the reference's order: field[comp][idx] (+|-)= value,
// value = point value of the step, or coef * dipole value of the step.
struct SrcOp {
    long idx;
    double coef;
    int comp;
    int kind;      // 0: point source value, 1: coef * dipole value
    int subtract;  // 1: -=
    int pec;       // target is a PEC-tangential edge
};

// Sources (thread 0, serial, in the reference's order), then the receivers
// (src/fem/fem_simulation.cpp:858-893): for each observation set and
// component, v = (v + sum of the listed edges) / count, into the
// device-resident trace row `row` (row < 0: no record).
template <typename T>
__global__ void k_fem1_post(F3<T> e, int nops, const SrcOp *__restrict__ ops, const double *__restrict__ val,
                            int skip_pec, int nobs, const int *__restrict__ ocnt, const long *__restrict__ oidx,
                            const long *__restrict__ ooff, double *__restrict__ tr, int nslot, int row)
{
    if (threadIdx.x == 0) {
        const double vp = val[0], vd = val[1];
        for (int s = 0; s < nops; ++s) {
            const SrcOp o = ops[s];
            if (skip_pec && o.pec) continue;
            const double v = o.kind == 0 ? vp : o.coef * vd;
            T *f = e.c[o.comp] + o.idx;
            *f = o.subtract ? *f - (T)v : *f + (T)v;
        }
    }
    __syncthreads();
    if (row < 0) return;
    for (int o = threadIdx.x; o < nobs; o += blockDim.x) {
        for (int q = 0; q < 3; ++q) {
            double *v = tr + ((long)o * nslot + row) * 3 + q;
            double acc = *v;
            const int n = ocnt[3 * o + q];
            const long *ix = oidx + ooff[3 * o + q];
            for (int a = 0; a < n; ++a) acc += (double)e.c[q][ix[a]];
            acc /= (double)n;
            *v = acc;
        }
    }
}

}  // namespace fem1
}  // namespace ems
