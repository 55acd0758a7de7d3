// Structured-hex edge-element FEM kernels (fem_v2 / symplectic_2 / symplectic_4
// family) for gfx950.
//
// Fields keep the reference layouts (src/fem_v2/fem_simulation.cpp:169-171),
// x fastest:  Ex[Z][Y][X] Nx(Ny+1)(Nz+1),  Ey (Nx+1)Ny(Nz+1),  Ez (Nx+1)(Ny+1)Nz
// on the fine grid of Nx x Ny x Nz sub-edges, p^3 sub-cells per element.
//
// The element-by-element product of the reference (fem_kernels.hpp:13-175)
// scatters K x_e into the output with atomics.  Here every thread owns one
// DOF and gathers: it finds the (at most 4) elements that contain its edge,
// orders them by the position at which the reference's serial loop visits
// them (outer-boundary list first, src/fem_v2/fem_simulation.cpp:359-395,
// then the inner elements k, j, i), and accumulates each element's K-row dot
// product (summed from 0 over m, as the reference) in that order.  No
// atomics, deterministic, and with contraction off the FP64 result equals a
// serial run of the reference bit for bit.  The element geometry is
// arithmetic (no connectivity arrays are read): the local-to-base offset of
// each element DOF comes from a small table.
#pragma once

#include <hip/hip_runtime.h>

namespace ems {
namespace fem {

struct FemGrid {
    int p, Nx, Ny, Nz;
    int nex, ney, nez;
    long NX, NY, NZ;  // entries per component
    long OE;          // outer elements: rank of the first inner element
    int D, n;         // DOFs per component per element, 3*D
};

// Position of element (i, j, k) in the reference's serial visiting order.
__device__ __forceinline__ long elem_rank(const FemGrid &g, int i, int j, int k)
{
    const long a = g.nex, b = g.ney, c = g.nez;
    if (k == 0) return i + j * a;
    if (k == c - 1) return a * b + i + j * a;
    if (i == 0) return 2 * a * b + j + (k - 1) * b;
    if (i == a - 1) return 2 * a * b + b * (c - 2) + j + (k - 1) * b;
    if (j == 0) return 2 * a * b + 2 * b * (c - 2) + (k - 1) + (i - 1) * (c - 2);
    if (j == b - 1) return 2 * a * b + 2 * b * (c - 2) + (c - 2) * (a - 2) + (k - 1) + (i - 1) * (c - 2);
    return g.OE + (i - 1) + (long)(j - 1) * (a - 2) + (long)(k - 1) * (a - 2) * (b - 2);
}

struct Cand {
    long rank;
    int i, j, k;  // element
    int r;        // local row of the DOF in that element
};

// Elements containing DOF (comp, X, Y, Z), sorted by rank.  Returns count.
template <int P>
__device__ __forceinline__ int dof_elements(const FemGrid &g, int comp, int X, int Y, int Z, Cand *out)
{
    constexpr int p = P;
    // per axis: up to two (element index, sub-index) choices
    int ei[3][2], li[3][2], cnt[3];
    const int C[3] = {X, Y, Z};
    const int ne[3] = {g.nex, g.ney, g.nez};
    for (int a = 0; a < 3; ++a) {
        const int v = C[a];
        if (a == comp || v % p != 0) {  // along the edge, or strictly inside
            ei[a][0] = v / p;
            li[a][0] = v % p;
            cnt[a] = 1;
        } else {
            cnt[a] = 0;
            if (v / p < ne[a]) {
                ei[a][cnt[a]] = v / p;
                li[a][cnt[a]] = 0;
                cnt[a]++;
            }
            if (v / p - 1 >= 0) {
                ei[a][cnt[a]] = v / p - 1;
                li[a][cnt[a]] = p;
                cnt[a]++;
            }
        }
    }
    // all candidates inner elements (1 <= index <= ne-2 on every axis): the
    // serial order of inner elements is k, then j, then i ascending, so the
    // candidates come out sorted without ranking them
    bool inner = true;
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < cnt[a]; ++q) inner = inner && ei[a][q] >= 1 && ei[a][q] <= ne[a] - 2;
    int nc = 0;
    if (inner) {
        for (int c = cnt[2] - 1; c >= 0; --c)      // ascending k (the +1 choice is stored first)
            for (int b = cnt[1] - 1; b >= 0; --b)  // ascending j
                for (int a = cnt[0] - 1; a >= 0; --a) {
                    Cand q;
                    q.i = ei[0][a];
                    q.j = ei[1][b];
                    q.k = ei[2][c];
                    const int l = li[0][a], m = li[1][b], n = li[2][c];
                    if (comp == 0) q.r = l + m * p + n * p * (p + 1);
                    else if (comp == 1) q.r = g.D + l + m * (p + 1) + n * p * (p + 1);
                    else q.r = 2 * g.D + l + m * (p + 1) + n * (p + 1) * (p + 1);
                    q.rank = g.OE;  // inner (only compared against OE)
                    out[nc++] = q;
                }
        return nc;
    }
    for (int a = 0; a < cnt[0]; ++a)
        for (int b = 0; b < cnt[1]; ++b)
            for (int c = 0; c < cnt[2]; ++c) {
                Cand q;
                q.i = ei[0][a];
                q.j = ei[1][b];
                q.k = ei[2][c];
                const int l = li[0][a], m = li[1][b], n = li[2][c];
                if (comp == 0) q.r = l + m * p + n * p * (p + 1);
                else if (comp == 1) q.r = g.D + l + m * (p + 1) + n * p * (p + 1);
                else q.r = 2 * g.D + l + m * (p + 1) + n * (p + 1) * (p + 1);
                q.rank = elem_rank(g, q.i, q.j, q.k);
                // insertion by rank
                int s = nc++;
                while (s > 0 && out[s - 1].rank > q.rank) {
                    out[s] = out[s - 1];
                    --s;
                }
                out[s] = q;
            }
    return nc;
}

__device__ __forceinline__ void elem_bases(const FemGrid &g, int i, int j, int k, long &bx, long &by, long &bz)
{
    const long p = g.p, Nx = g.Nx, Ny = g.Ny;
    bx = p * i + p * j * Nx + p * k * Nx * (Ny + 1);
    by = p * i + p * j * (Nx + 1) + p * k * (Nx + 1) * Ny;
    bz = p * i + p * j * (Nx + 1) + p * k * (Nx + 1) * (Ny + 1);
}

__device__ __forceinline__ bool pec_dof(const FemGrid &g, int comp, int X, int Y, int Z)
{
    if (comp == 0) return Y == 0 || Y == g.Ny || Z == 0 || Z == g.Nz;
    if (comp == 1) return X == 0 || X == g.Nx || Z == 0 || Z == g.Nz;
    return X == 0 || X == g.Nx || Y == 0 || Y == g.Ny;
}

// Extent of component COMP's array (x, y, z) and its coordinates from a 3-D
// launch: workgroup 64 x 4, grid (ceil(nx/64), ceil(ny/4), nz).  No integer
// division on the device.
template <int COMP> __device__ __forceinline__ void comp_dims(const FemGrid &g, int &dx, int &dy, int &dz)
{
    dx = g.Nx + (COMP != 0);
    dy = g.Ny + (COMP != 1);
    dz = g.Nz + (COMP != 2);
}

template <typename T> struct Comp3 {
    T *x, *y, *z;
};
template <typename T> struct CComp3 {
    const T *x, *y, *z;
};

// Value of gathered DOF: plain (leapfrog) or drifted U + ((dt/2) P)/m (symplectic)
template <typename T, bool DRIFT>
__device__ __forceinline__ T gval(const T *__restrict__ u, const T *__restrict__ pm, long q, T hd, T mm)
{
    if (DRIFT) return u[q] + hd * pm[q] / mm;
    return u[q];
}

// dot(K row r, x_e) for element (i, j, k), summed from 0 over m in order.
template <typename T, int P, bool DRIFT>
__device__ __forceinline__ T row_dot(const FemGrid &g, const T *__restrict__ K, const int *__restrict__ off, int r,
                                     int i, int j, int k, CComp3<T> u, CComp3<T> pm, T hd, T mm)
{
    constexpr int D = P * (P + 1) * (P + 1);
    constexpr int N = 3 * D;
    long bx, by, bz;
    elem_bases(g, i, j, k, bx, by, bz);
    const T *__restrict__ kr = K + (long)r * N;
    T temp = 0;
    if constexpr (P == 1) {
        // first order: the 12 offsets in the local order of the table
        // (x: +Y, +Z; y: +X, +Z; z: +X, +Y), without the table lookups
        const long Nx = g.Nx, Ny = g.Ny, sx = Nx * (Ny + 1), sy = (Nx + 1) * Ny;
        temp += kr[0] * gval<T, DRIFT>(u.x, pm.x, bx, hd, mm);
        temp += kr[1] * gval<T, DRIFT>(u.x, pm.x, bx + Nx, hd, mm);
        temp += kr[2] * gval<T, DRIFT>(u.x, pm.x, bx + sx, hd, mm);
        temp += kr[3] * gval<T, DRIFT>(u.x, pm.x, bx + sx + Nx, hd, mm);
        temp += kr[4] * gval<T, DRIFT>(u.y, pm.y, by, hd, mm);
        temp += kr[5] * gval<T, DRIFT>(u.y, pm.y, by + 1, hd, mm);
        temp += kr[6] * gval<T, DRIFT>(u.y, pm.y, by + sy, hd, mm);
        temp += kr[7] * gval<T, DRIFT>(u.y, pm.y, by + sy + 1, hd, mm);
        temp += kr[8] * gval<T, DRIFT>(u.z, pm.z, bz, hd, mm);
        temp += kr[9] * gval<T, DRIFT>(u.z, pm.z, bz + 1, hd, mm);
        temp += kr[10] * gval<T, DRIFT>(u.z, pm.z, bz + Nx + 1, hd, mm);
        temp += kr[11] * gval<T, DRIFT>(u.z, pm.z, bz + Nx + 2, hd, mm);
        return temp;
    }
#pragma unroll 4
    for (int m = 0; m < D; ++m) temp += kr[m] * gval<T, DRIFT>(u.x, pm.x, bx + off[m], hd, mm);
#pragma unroll 4
    for (int m = D; m < 2 * D; ++m) temp += kr[m] * gval<T, DRIFT>(u.y, pm.y, by + off[m], hd, mm);
#pragma unroll 4
    for (int m = 2 * D; m < N; ++m) temp += kr[m] * gval<T, DRIFT>(u.z, pm.z, bz + off[m], hd, mm);
    return temp;
}

// ---------------------------------------------------------------------------
// fem_v2 leapfrog step (src/fem_v2/fem_simulation.cpp:560-655 + BC):
//   E2 = ((0 - sum_e K_e E1) + (2 E1 - E0)) - src,  PEC-tangential DOFs -> 0
// ---------------------------------------------------------------------------
template <typename T, int P, int COMP>
__global__ __launch_bounds__(256) void k_fem_leapfrog(FemGrid g, const T *__restrict__ K, const int *__restrict__ off,
                                                      CComp3<T> e0, CComp3<T> e1, Comp3<T> e2, int nsrc,
                                                      const long *__restrict__ src_idx, const double *__restrict__ src_val)
{
    int dx, dy, dz;
    comp_dims<COMP>(g, dx, dy, dz);
    const int X = blockIdx.x * 64 + threadIdx.x;
    const int Y = blockIdx.y * 4 + threadIdx.y;
    const int Z = blockIdx.z;
    if (X >= dx || Y >= dy) return;
    const long idx = X + (long)dx * (Y + (long)dy * Z);
    T *out = COMP == 0 ? e2.x : COMP == 1 ? e2.y : e2.z;
    if (pec_dof(g, COMP, X, Y, Z)) {
        out[idx] = (T)0;
        return;
    }
    Cand c[4];
    const int nc = dof_elements<P>(g, COMP, X, Y, Z, c);
    T acc = 0;
    for (int q = 0; q < nc; ++q)
        acc -= row_dot<T, P, false>(g, K, off, c[q].r, c[q].i, c[q].j, c[q].k, e1, e1, (T)0, (T)1);
    const T *a1 = COMP == 0 ? e1.x : COMP == 1 ? e1.y : e1.z;
    const T *a0 = COMP == 0 ? e0.x : COMP == 1 ? e0.y : e0.z;
    T v = acc + ((T)2 * a1[idx] - a0[idx]);
    if (COMP == 2)
        for (int s = 0; s < nsrc; ++s)
            if (src_idx[s] == idx) v -= (T)src_val[0];
    out[idx] = v;
}

// ---------------------------------------------------------------------------
// Symplectic drift-kick-drift stage (src/symplectic_2/fem_simulation.cpp:557-672):
//   U' = U + ((dt/2) P)/m ;  tmp = sum_e K_e U' (+ src between the outer and
//   inner elements, as the reference adds it) ;  P' = P - dt*tmp ;
//   U'' = U' + ((dt/2) P')/m ;  PEC-tangential U, P -> 0.
// Reads (U, P) from one buffer pair and writes the other; neighbours' drift
// is recomputed in the gather, so one pass per stage.
// ---------------------------------------------------------------------------
template <typename T, int P, int COMP>
__global__ __launch_bounds__(256) void k_fem_symplectic(FemGrid g, const T *__restrict__ K,
                                                        const int *__restrict__ off, CComp3<T> u, CComp3<T> pm,
                                                        Comp3<T> u_out, Comp3<T> p_out, T deltat, T hd, T mm,
                                                        int nsrc, const long *__restrict__ src_idx,
                                                        const double *__restrict__ src_val)
{
    int dx, dy, dz;
    comp_dims<COMP>(g, dx, dy, dz);
    const int X = blockIdx.x * 64 + threadIdx.x;
    const int Y = blockIdx.y * 4 + threadIdx.y;
    const int Z = blockIdx.z;
    if (X >= dx || Y >= dy) return;
    const long idx = X + (long)dx * (Y + (long)dy * Z);
    T *uo = COMP == 0 ? u_out.x : COMP == 1 ? u_out.y : u_out.z;
    T *po = COMP == 0 ? p_out.x : COMP == 1 ? p_out.y : p_out.z;
    if (pec_dof(g, COMP, X, Y, Z)) {
        uo[idx] = (T)0;
        po[idx] = (T)0;
        return;
    }
    Cand c[4];
    const int nc = dof_elements<P>(g, COMP, X, Y, Z, c);
    int nsrc_here = 0;
    if (COMP == 2)
        for (int s = 0; s < nsrc; ++s) nsrc_here += src_idx[s] == idx;
    const T sv = (T)src_val[0];
    T tmp = 0;
    bool src_done = nsrc_here == 0;
    for (int q = 0; q < nc; ++q) {
        if (!src_done && c[q].rank >= g.OE) {
            for (int s = 0; s < nsrc_here; ++s) tmp += sv;
            src_done = true;
        }
        tmp += row_dot<T, P, true>(g, K, off, c[q].r, c[q].i, c[q].j, c[q].k, u, pm, hd, mm);
    }
    if (!src_done)
        for (int s = 0; s < nsrc_here; ++s) tmp += sv;
    const T *ua = COMP == 0 ? u.x : COMP == 1 ? u.y : u.z;
    const T *pa = COMP == 0 ? pm.x : COMP == 1 ? pm.y : pm.z;
    const T u1 = ua[idx] + hd * pa[idx] / mm;
    const T p1 = pa[idx] - deltat * tmp;
    uo[idx] = u1 + hd * p1 / mm;
    po[idx] = p1;
}

// ---------------------------------------------------------------------------
// First order with the fields staged in LDS.  A 64 x 4 workgroup of
// component COMP's DOFs at plane Z first loads, for each field array q, the
// box of values its elements touch (element range along axis a:
// [T0 - (a != COMP), T1]; array q adds +1 along the axes a != q), and the
// 12 x 12 element matrix; symplectic stages store the drifted value
// u + ((dt/2) P)/m, computed as in the global-memory kernels.  The per-DOF
// work is then that of the kernels above (same candidate order, same row
// dot order), with every operand read from LDS: the same bits.
// ---------------------------------------------------------------------------
template <int COMP, int Q, int A> __device__ __host__ constexpr int t1_ext()
{
    return (A == 0 ? 64 : A == 1 ? 4 : 1) + (A != COMP ? 1 : 0) + (A != Q ? 1 : 0);
}
template <int COMP, int Q> __device__ __host__ constexpr int t1_size()
{
    return t1_ext<COMP, Q, 0>() * t1_ext<COMP, Q, 1>() * t1_ext<COMP, Q, 2>();
}
template <int COMP> __device__ __host__ constexpr int t1_total()
{
    return t1_size<COMP, 0>() + t1_size<COMP, 1>() + t1_size<COMP, 2>();
}

// load array q's box (values or drifted values; outside the array: 0)
template <typename T, int COMP, int Q, bool DRIFT>
__device__ __forceinline__ void t1_load(const FemGrid &g, T *dst, const T *__restrict__ u, const T *__restrict__ pm,
                                        T hd, T mm, int ox, int oy, int oz)
{
    constexpr int ex = t1_ext<COMP, Q, 0>(), ey = t1_ext<COMP, Q, 1>(), n = t1_size<COMP, Q>();
    const int dx = g.Nx + (Q != 0), dy = g.Ny + (Q != 1), dz = g.Nz + (Q != 2);
    const int tid = threadIdx.y * 64 + threadIdx.x;
    for (int q = tid; q < n; q += 256) {
        const int lx = q % ex, ly = (q / ex) % ey, lz = q / (ex * ey);
        const int X = ox + lx, Y = oy + ly, Z = oz + lz;
        T v = 0;
        if (X >= 0 && Y >= 0 && Z >= 0 && X < dx && Y < dy && Z < dz)
            v = gval<T, DRIFT>(u, pm, X + (long)dx * (Y + (long)dy * Z), hd, mm);
        dst[q] = v;
    }
}

template <typename T, int COMP, bool DRIFT>
__device__ __forceinline__ void t1_stage(const FemGrid &g, const T *__restrict__ K, T *sK, T *sE, CComp3<T> u,
                                         CComp3<T> pm, T hd, T mm, int ox, int oy, int oz)
{
    const int tid = threadIdx.y * 64 + threadIdx.x;
    if (tid < 144) sK[tid] = K[tid];
    t1_load<T, COMP, 0, DRIFT>(g, sE, u.x, pm.x, hd, mm, ox, oy, oz);
    t1_load<T, COMP, 1, DRIFT>(g, sE + t1_size<COMP, 0>(), u.y, pm.y, hd, mm, ox, oy, oz);
    t1_load<T, COMP, 2, DRIFT>(g, sE + t1_size<COMP, 0>() + t1_size<COMP, 1>(), u.z, pm.z, hd, mm, ox, oy, oz);
}

// dot(K row r, x_e) of element (i, j, k) from the staged box (origin ox, oy, oz)
template <typename T, int COMP>
__device__ __forceinline__ T t1_row_dot(const T *sK, const T *sE, int r, int i, int j, int k, int ox, int oy, int oz)
{
    constexpr int e0x = t1_ext<COMP, 0, 0>(), e0y = t1_ext<COMP, 0, 1>();
    constexpr int e1x = t1_ext<COMP, 1, 0>(), e1y = t1_ext<COMP, 1, 1>();
    constexpr int e2x = t1_ext<COMP, 2, 0>(), e2y = t1_ext<COMP, 2, 1>();
    const T *sx = sE, *sy = sE + t1_size<COMP, 0>(), *sz = sy + t1_size<COMP, 1>();
    const int li = i - ox, lj = j - oy, lk = k - oz;
    const int bx = li + e0x * (lj + e0y * lk), by = li + e1x * (lj + e1y * lk), bz = li + e2x * (lj + e2y * lk);
    const T *kr = sK + r * 12;
    T temp = 0;
    temp += kr[0] * sx[bx];
    temp += kr[1] * sx[bx + e0x];
    temp += kr[2] * sx[bx + e0x * e0y];
    temp += kr[3] * sx[bx + e0x * e0y + e0x];
    temp += kr[4] * sy[by];
    temp += kr[5] * sy[by + 1];
    temp += kr[6] * sy[by + e1x * e1y];
    temp += kr[7] * sy[by + e1x * e1y + 1];
    temp += kr[8] * sz[bz];
    temp += kr[9] * sz[bz + 1];
    temp += kr[10] * sz[bz + e2x];
    temp += kr[11] * sz[bz + e2x + 1];
    return temp;
}

template <typename T, int COMP>
__global__ __launch_bounds__(256) void k_fem_leapfrog1(FemGrid g, const T *__restrict__ K, CComp3<T> e0, CComp3<T> e1,
                                                       Comp3<T> e2, int nsrc, const long *__restrict__ src_idx,
                                                       const double *__restrict__ src_val)
{
    __shared__ T sK[144];
    __shared__ T sE[t1_total<COMP>()];
    const int X0 = blockIdx.x * 64, Y0 = blockIdx.y * 4, Z = blockIdx.z;
    const int ox = X0 - (COMP != 0), oy = Y0 - (COMP != 1), oz = Z - (COMP != 2);
    t1_stage<T, COMP, false>(g, K, sK, sE, e1, e1, (T)0, (T)1, ox, oy, oz);
    __syncthreads();
    int dx, dy, dz;
    comp_dims<COMP>(g, dx, dy, dz);
    const int X = X0 + threadIdx.x, Y = Y0 + threadIdx.y;
    if (X >= dx || Y >= dy) return;
    const long idx = X + (long)dx * (Y + (long)dy * Z);
    T *out = COMP == 0 ? e2.x : COMP == 1 ? e2.y : e2.z;
    if (pec_dof(g, COMP, X, Y, Z)) {
        out[idx] = (T)0;
        return;
    }
    Cand c[4];
    const int nc = dof_elements<1>(g, COMP, X, Y, Z, c);
    T acc = 0;
    for (int q = 0; q < nc; ++q) acc -= t1_row_dot<T, COMP>(sK, sE, c[q].r, c[q].i, c[q].j, c[q].k, ox, oy, oz);
    const T *a1 = COMP == 0 ? e1.x : COMP == 1 ? e1.y : e1.z;
    const T *a0 = COMP == 0 ? e0.x : COMP == 1 ? e0.y : e0.z;
    T v = acc + ((T)2 * a1[idx] - a0[idx]);
    if (COMP == 2)
        for (int s = 0; s < nsrc; ++s)
            if (src_idx[s] == idx) v -= (T)src_val[0];
    out[idx] = v;
}

template <typename T, int COMP>
__global__ __launch_bounds__(256) void k_fem_symplectic1(FemGrid g, const T *__restrict__ K, CComp3<T> u,
                                                         CComp3<T> pm, Comp3<T> u_out, Comp3<T> p_out, T deltat, T hd,
                                                         T mm, int nsrc, const long *__restrict__ src_idx,
                                                         const double *__restrict__ src_val)
{
    __shared__ T sK[144];
    __shared__ T sE[t1_total<COMP>()];
    const int X0 = blockIdx.x * 64, Y0 = blockIdx.y * 4, Z = blockIdx.z;
    const int ox = X0 - (COMP != 0), oy = Y0 - (COMP != 1), oz = Z - (COMP != 2);
    t1_stage<T, COMP, true>(g, K, sK, sE, u, pm, hd, mm, ox, oy, oz);
    __syncthreads();
    int dx, dy, dz;
    comp_dims<COMP>(g, dx, dy, dz);
    const int X = X0 + threadIdx.x, Y = Y0 + threadIdx.y;
    if (X >= dx || Y >= dy) return;
    const long idx = X + (long)dx * (Y + (long)dy * Z);
    T *uo = COMP == 0 ? u_out.x : COMP == 1 ? u_out.y : u_out.z;
    T *po = COMP == 0 ? p_out.x : COMP == 1 ? p_out.y : p_out.z;
    if (pec_dof(g, COMP, X, Y, Z)) {
        uo[idx] = (T)0;
        po[idx] = (T)0;
        return;
    }
    Cand c[4];
    const int nc = dof_elements<1>(g, COMP, X, Y, Z, c);
    int nsrc_here = 0;
    if (COMP == 2)
        for (int s = 0; s < nsrc; ++s) nsrc_here += src_idx[s] == idx;
    const T sv = (T)src_val[0];
    T tmp = 0;
    bool src_done = nsrc_here == 0;
    for (int q = 0; q < nc; ++q) {
        if (!src_done && c[q].rank >= g.OE) {
            for (int s = 0; s < nsrc_here; ++s) tmp += sv;
            src_done = true;
        }
        tmp += t1_row_dot<T, COMP>(sK, sE, c[q].r, c[q].i, c[q].j, c[q].k, ox, oy, oz);
    }
    if (!src_done)
        for (int s = 0; s < nsrc_here; ++s) tmp += sv;
    const T *ua = COMP == 0 ? u.x : COMP == 1 ? u.y : u.z;
    const T *pa = COMP == 0 ? pm.x : COMP == 1 ? pm.y : pm.z;
    const T u1 = ua[idx] + hd * pa[idx] / mm;
    const T p1 = pa[idx] - deltat * tmp;
    uo[idx] = u1 + hd * p1 / mm;
    po[idx] = p1;
}

// Receivers (src/fem_v2/fem_simulation.cpp:1275-1310): the four x edges and
// four y edges around the point and one z edge, raw, into rec[obs][9].
template <typename T>
__global__ void k_fem_record(CComp3<T> f, int nobs, const long *__restrict__ obs9, double *rec)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nobs) return;
    const long *q = obs9 + 9 * o;
    for (int a = 0; a < 4; ++a) rec[9 * o + a] = (double)f.x[q[a]];
    for (int a = 4; a < 8; ++a) rec[9 * o + a] = (double)f.y[q[a]];
    rec[9 * o + 8] = (double)f.z[q[8]];
}

}  // namespace fem
}  // namespace ems
