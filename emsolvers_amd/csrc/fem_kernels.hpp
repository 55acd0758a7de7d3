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
// each element DOF comes from a small table.  At first order the default
// instead sums each DOF's element rows into one merged stencil
// (hex_stencil.hpp), and at order 3 the element products run on matrix
// cores: both equal the reference to rounding.
#pragma once

#include <hip/hip_runtime.h>

#include "hex_stencil.hpp"
#include "mfma.hpp"

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

// Elements of a DOF, held in registers (a candidate array filled with a
// data-dependent count would live in scratch memory).  Along each of the DOF's
// two transverse axes the DOF lies in the element at v / p, or, when v is on
// an element boundary (v % p == 0), also in the element v / p - 1 (local
// index p).  Code a | b << 1 picks the choice on the first (a) and second (b)
// transverse axis (x edge: y, z; y edge: x, z; z edge: x, y); 1 = the lower
// element.  Returns false when that element does not exist.
template <int P>
__device__ __forceinline__ bool elem_of(const FemGrid &g, int comp, int X, int Y, int Z, int code, int &i, int &j,
                                        int &k, int &r)
{
    constexpr int p = P;
    const int C[3] = {X, Y, Z};
    const int ne[3] = {g.nex, g.ney, g.nez};
    int e[3], l[3];
    bool ok = true;
    int t = 0;  // transverse axis counter
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int v = C[a];
        if (a == comp) {
            e[a] = v / p;
            l[a] = v % p;
            continue;
        }
        const int lower = (code >> t) & 1;
        ++t;
        if (v % p != 0) {
            e[a] = v / p;
            l[a] = v % p;
            ok = ok && lower == 0;
        } else if (lower == 0) {
            e[a] = v / p;
            l[a] = 0;
            ok = ok && e[a] < ne[a];
        } else {
            e[a] = v / p - 1;
            l[a] = p;
            ok = ok && e[a] >= 0;
        }
    }
    i = e[0];
    j = e[1];
    k = e[2];
    if (comp == 0) r = l[0] + l[1] * p + l[2] * p * (p + 1);
    else if (comp == 1) r = g.D + l[0] + l[1] * (p + 1) + l[2] * p * (p + 1);
    else r = 2 * g.D + l[0] + l[1] * (p + 1) + l[2] * (p + 1) * (p + 1);
    return ok;
}

constexpr long kNoRank = 0x7fffffffffffffffL;

// The (up to 4) element codes of a DOF in the reference's serial order
// (ascending element rank), missing elements last with rank kNoRank.  Inner
// elements are visited k, j, i ascending, i.e. codes 3, 2, 1, 0; otherwise
// the ranks are sorted by a 5-comparator network, in registers.
template <int P>
__device__ __forceinline__ void order4(const FemGrid &g, int comp, int X, int Y, int Z, int (&cd)[4], long (&rk)[4])
{
    bool inner = true;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int i, j, k, r;
        cd[q] = 3 - q;
        if (elem_of<P>(g, comp, X, Y, Z, 3 - q, i, j, k, r)) {
            inner = inner && i >= 1 && j >= 1 && k >= 1 && i <= g.nex - 2 && j <= g.ney - 2 && k <= g.nez - 2;
            rk[q] = elem_rank(g, i, j, k);
        } else {
            rk[q] = kNoRank;
        }
    }
    if (inner) return;
    auto cswap = [&](int a, int b) {
        if (rk[b] < rk[a]) {
            const long t = rk[a];
            rk[a] = rk[b];
            rk[b] = t;
            const int c = cd[a];
            cd[a] = cd[b];
            cd[b] = c;
        }
    };
    cswap(0, 1);
    cswap(2, 3);
    cswap(0, 2);
    cswap(1, 3);
    cswap(1, 2);
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

// ---------------------------------------------------------------------------
// Orders 2 and 3, element by element.  Phase 1 (k_fem_elem): one thread per element
// loads its n field values once and stores all n row dots
//   y_e[r] = sum_m K[r][m] x_e[m]   (summed from 0 over m in order)
// at tmp[r * nelem + e]; the matrix is read at addresses uniform over the
// wavefront.  Phase 2: the per-DOF kernels below add each element's row dot
// read from tmp: the same bits as a per-DOF gather, which would load every
// element's n values once per DOF of it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ long elem_slot(const FemGrid &g, int r, int i, int j, int k)
{
    const long nelem = (long)g.nex * g.ney * g.nez;
    return r * nelem + i + (long)g.nex * (j + (long)g.ney * k);
}

// Workgroup size of k_fem_elem: at order 3 the 144 values of an element take
// up to 288 registers, which one wave per workgroup can hold.
template <int P> constexpr int kElemWG = P == 3 ? 64 : 256;

template <typename T, int P>
__global__ __launch_bounds__(kElemWG<P>) void k_fem_elem(FemGrid g, const T *__restrict__ K,
                                                         const int *__restrict__ off, CComp3<T> x, T *__restrict__ tmp)
{
    constexpr int D = P * (P + 1) * (P + 1), N = 3 * D;
    const long nelem = (long)g.nex * g.ney * g.nez;
    const long e = (long)blockIdx.x * kElemWG<P> + threadIdx.x;
    if (e >= nelem) return;
    const int i = (int)(e % g.nex), j = (int)((e / g.nex) % g.ney), k = (int)(e / ((long)g.nex * g.ney));
    long bx, by, bz;
    elem_bases(g, i, j, k, bx, by, bz);
    T v[N];
#pragma unroll
    for (int m = 0; m < D; ++m) v[m] = x.x[bx + off[m]];
#pragma unroll
    for (int m = D; m < 2 * D; ++m) v[m] = x.y[by + off[m]];
#pragma unroll
    for (int m = 2 * D; m < N; ++m) v[m] = x.z[bz + off[m]];
    // rows rolled (one copy of the n-term dot), columns unrolled: v stays in
    // registers
#pragma unroll 1
    for (int r = 0; r < N; ++r) {
        const T *__restrict__ kr = K + r * N;
        T t = 0;
#pragma unroll
        for (int m = 0; m < N; ++m) t += kr[m] * v[m];
        tmp[r * nelem + e] = t;
    }
}

// Orders 2 and 3 on matrix cores: the row dots of k_fem_elem (tmp[r nelem +
// e]) as the product K X of the n x n element matrix and the n x 16 block of
// 16 elements' local vectors, one block per wavefront, in 16 x 16 x 4 MFMA
// steps (row tiles q of 16 rows, k steps of 4 columns).  The sums run in the
// instructions' order with fused multiply-adds, not in k_fem_elem's order, so
// the row dots agree with it to rounding, not bit for bit.  Kp holds K in
// the A-operand layout, Kp[(q KS + ks) 64 + lane] = K[16 q + lane % 16][4 ks +
// lane / 16] (zero outside n x n); workgroups of 4 wavefronts.
template <int P> struct ElemTiles {
    static constexpr int D = P * (P + 1) * (P + 1), N = 3 * D, MT = (N + 15) / 16, KS = (N + 3) / 4;
};

template <typename T, int P>
__global__ __launch_bounds__(256) void k_fem_elem_mfma(FemGrid g, const T *__restrict__ Kp,
                                                       const int *__restrict__ off, CComp3<T> x, T *__restrict__ tmp)
{
    constexpr int D = ElemTiles<P>::D, N = ElemTiles<P>::N, MT = ElemTiles<P>::MT, KS = ElemTiles<P>::KS;
    typedef typename MfmaAcc<T>::V V;
    const long nelem = (long)g.nex * g.ney * g.nez;
    const int lane = threadIdx.x & 63;
    const long base = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;  // first element of this wavefront
    if (base >= nelem) return;  // uniform per wavefront; no barrier below
    const int n = lane & 15, kq = lane >> 4;
    const long e = base + n;
    const bool valid = e < nelem;
    long bx = 0, by = 0, bz = 0;
    if (valid) {
        const int i = (int)(e % g.nex), j = (int)((e / g.nex) % g.ney), k = (int)(e / ((long)g.nex * g.ney));
        elem_bases(g, i, j, k, bx, by, bz);
    }
    V acc[MT];
#pragma unroll
    for (int q = 0; q < MT; ++q) acc[q] = V{0, 0, 0, 0};
#pragma unroll 4
    for (int ks = 0; ks < KS; ++ks) {
        const int m = 4 * ks + kq;
        T bv = 0;
        if (valid && m < N) bv = m < D ? x.x[bx + off[m]] : m < 2 * D ? x.y[by + off[m]] : x.z[bz + off[m]];
        const T *kp = Kp + (long)ks * 64 + lane;
#pragma unroll
        for (int q = 0; q < MT; ++q) acc[q] = MfmaAcc<T>::step(kp[(long)q * KS * 64], bv, acc[q]);
    }
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < MT; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * q + MfmaAcc<T>::row(lane, v);
            if (r < N) tmp[r * nelem + e] = acc[q][v];
        }
}

// ---------------------------------------------------------------------------
// fem_v2 leapfrog step (src/fem_v2/fem_simulation.cpp:560-655 + BC):
//   E2 = ((0 - sum_e K_e E1) + (2 E1 - E0)) - src,  PEC-tangential DOFs -> 0
// ---------------------------------------------------------------------------
// Orders 2 and 3: the row dots come from etmp (k_fem_elem).
template <typename T, int P, int COMP>
__global__ __launch_bounds__(256) void k_fem_leapfrog(FemGrid g, CComp3<T> e0, CComp3<T> e1, Comp3<T> e2, int nsrc,
                                                      const long *__restrict__ src_idx, const double *__restrict__ src_val,
                                                      const T *__restrict__ etmp)
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
    T acc = 0;
    {  // one stored value per element: candidates in registers
        int cd[4];
        long rk[4];
        order4<P>(g, COMP, X, Y, Z, cd, rk);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int i, j, k, r;
            if (rk[q] == kNoRank || !elem_of<P>(g, COMP, X, Y, Z, cd[q], i, j, k, r)) continue;
            acc -= etmp[elem_slot(g, r, i, j, k)];
        }
    }
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
// Drifted field u + ((dt/2) P)/m of every DOF (the value the symplectic
// gathers read), once per stage instead of once per gathered term: the same
// expression, the same bits.  Grid (ceil(max count / 256), 3 components).
template <typename T>
__global__ __launch_bounds__(256) void k_fem_drift(FemGrid g, CComp3<T> u, CComp3<T> pm, Comp3<T> d, T hd, T mm)
{
    const int c = blockIdx.y;
    const long n = c == 0 ? g.NX : c == 1 ? g.NY : g.NZ;
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const T *uc = c == 0 ? u.x : c == 1 ? u.y : u.z;
    const T *pc = c == 0 ? pm.x : c == 1 ? pm.y : pm.z;
    T *dc = c == 0 ? d.x : c == 1 ? d.y : d.z;
    dc[i] = gval<T, true>(uc, pc, i, hd, mm);
}

// Orders 2 and 3: etmp holds the row dots of the drifted field of
// k_fem_drift (k_fem_elem); u, pm: the DOF's own values.
template <typename T, int P, int COMP>
__global__ __launch_bounds__(256) void k_fem_symplectic(FemGrid g, CComp3<T> u, CComp3<T> pm, Comp3<T> u_out,
                                                        Comp3<T> p_out, T deltat, T hd, T mm, int nsrc,
                                                        const long *__restrict__ src_idx,
                                                        const double *__restrict__ src_val,
                                                        const T *__restrict__ etmp)
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
    int nsrc_here = 0;
    if (COMP == 2)
        for (int s = 0; s < nsrc; ++s) nsrc_here += src_idx[s] == idx;
    const T sv = (T)src_val[0];
    T tmp = 0;
    bool src_done = nsrc_here == 0;
    {  // one stored value per element: candidates in registers
        int cd[4];
        long rk[4];
        order4<P>(g, COMP, X, Y, Z, cd, rk);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int i, j, k, r;
            if (rk[q] == kNoRank || !elem_of<P>(g, COMP, X, Y, Z, cd[q], i, j, k, r)) continue;
            if (!src_done && rk[q] >= g.OE) {
                for (int s = 0; s < nsrc_here; ++s) tmp += sv;
                src_done = true;
            }
            tmp += etmp[elem_slot(g, r, i, j, k)];
        }
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
    int cd[4];
    long rk[4];
    order4<1>(g, COMP, X, Y, Z, cd, rk);
    T acc = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int i, j, k, r;
        if (!elem_of<1>(g, COMP, X, Y, Z, cd[q], i, j, k, r)) continue;
        acc -= t1_row_dot<T, COMP>(sK, sE, r, i, j, k, ox, oy, oz);
    }
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
    int cd[4];
    long rk[4];
    order4<1>(g, COMP, X, Y, Z, cd, rk);
    int nsrc_here = 0;
    if (COMP == 2)
        for (int s = 0; s < nsrc; ++s) nsrc_here += src_idx[s] == idx;
    const T sv = (T)src_val[0];
    T tmp = 0;
    bool src_done = nsrc_here == 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int i, j, k, r;
        if (!elem_of<1>(g, COMP, X, Y, Z, cd[q], i, j, k, r)) continue;
        if (!src_done && rk[q] >= g.OE) {
            for (int s = 0; s < nsrc_here; ++s) tmp += sv;
            src_done = true;
        }
        tmp += t1_row_dot<T, COMP>(sK, sE, r, i, j, k, ox, oy, oz);
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
// First order, all three components in one z-marching pass.
//
// A workgroup of 64 x BY threads owns the points X0 <= X < X0 + 64,
// Y0 <= Y < Y0 + BY of a chunk of planes and updates, at each plane Z, the
// x, y and z edge DOFs based at (X, Y, Z).  Every DOF on the faces X = Nx,
// Y = Ny, Z = Nz is PEC-tangential (zero): the threads at X = Nx-1,
// Y = Ny-1, Z = Nz-1 write those, so the thread grid covers Nx x Ny x Nz.
// The fields (drifted in symplectic stages, as gval) are staged in LDS as a
// rolling window of 4 plane slots, each holding the x, y and z arrays on the
// 66 x (BY+2) box (X0-1 .. X0+64, Y0-1 .. Y0+BY) of one plane; plane P lives
// in slot P & 3.  Plane Z+2 is loaded into registers while plane Z is
// computed from planes Z-1 .. Z+1 and stored after it: one barrier per plane.
// Candidate order and row dots are those of the kernels above: same bits.
// ---------------------------------------------------------------------------
template <int BY> struct Z1Win {
    static constexpr int EX = 66, EY = BY + 2;
    static constexpr int PLANE = EX * EY;  // one array of one plane
    static constexpr int SLOT = 3 * PLANE;  // x, y, z arrays of one plane
    static constexpr int NT = 64 * BY;
    static constexpr int NPF = (SLOT + NT - 1) / NT;  // entries per thread per plane
};

template <typename T, int BY>
__device__ __forceinline__ T z1_row_dot(const T *sK, const T *win, int r, int i, int j, int k, int ox, int oy)
{
    using W = Z1Win<BY>;
    const T *p0 = win + (k & 3) * W::SLOT, *p1 = win + ((k + 1) & 3) * W::SLOT;
    const int b = (i - ox) + W::EX * (j - oy);
    const T *kr = sK + r * 12;
    T temp = 0;
    temp += kr[0] * p0[b];
    temp += kr[1] * p0[b + W::EX];
    temp += kr[2] * p1[b];
    temp += kr[3] * p1[b + W::EX];
    temp += kr[4] * p0[W::PLANE + b];
    temp += kr[5] * p0[W::PLANE + b + 1];
    temp += kr[6] * p1[W::PLANE + b];
    temp += kr[7] * p1[W::PLANE + b + 1];
    temp += kr[8] * p0[2 * W::PLANE + b];
    temp += kr[9] * p0[2 * W::PLANE + b + 1];
    temp += kr[10] * p0[2 * W::PLANE + b + W::EX];
    temp += kr[11] * p0[2 * W::PLANE + b + W::EX + 1];
    return temp;
}

// Merged first-order stencil (hex_stencil.hpp): off the PEC faces every DOF
// lies in exactly 4 elements, and the sum of their row dots is one 33-point
// stencil per component; the host sums the scaled element rows into its
// coefficients once.
// Stencil sum of component C at the window position lo of plane Z: one
// partial sum per array, coefficients at uniform addresses (scalar loads).
template <typename T, int BY, int C>
__device__ __forceinline__ T z1_stencil(const T *__restrict__ cf, const T *win, int Z, int lo)
{
    using W = Z1Win<BY>;
    const T *pz[3] = {win + ((Z + 3) & 3) * W::SLOT + lo, win + (Z & 3) * W::SLOT + lo,
                      win + ((Z + 1) & 3) * W::SLOT + lo};
    T s[3] = {0, 0, 0};
#pragma unroll
    for (int t = 0; t < kZ1Terms; ++t) {
        const Z1Term m = z1_term(C, t);
        s[m.q] += cf[t] * pz[m.d[2] + 1][m.q * W::PLANE + m.d[0] + W::EX * m.d[1]];
    }
    return (s[0] + s[1]) + s[2];
}

// The step of one DOF from the merged stencil: the update of z1_dof below,
// with the sources added after the operator sum.
template <typename T, int BY, bool DRIFT>
__device__ __forceinline__ void z1_dof_stn(const T *__restrict__ cf, const T *win, int comp, int Z, int lo, long idx,
                                           T own, const T *__restrict__ other, T *o1, T *o2, T deltat, T hd, T mm,
                                           int nsrc, const long *__restrict__ src_idx, T sv)
{
    const T sum = comp == 0   ? z1_stencil<T, BY, 0>(cf, win, Z, lo)
                  : comp == 1 ? z1_stencil<T, BY, 1>(cf + kZ1Terms, win, Z, lo)
                              : z1_stencil<T, BY, 2>(cf + 2 * kZ1Terms, win, Z, lo);
    int nsrc_here = 0;
    if (comp == 2)
        for (int s = 0; s < nsrc; ++s) nsrc_here += src_idx[s] == idx;
    if (!DRIFT) {
        T v = (0 - sum) + ((T)2 * own - other[idx]);
        for (int s = 0; s < nsrc_here; ++s) v -= sv;
        o1[idx] = v;
    } else {
        T tmp = sum;
        for (int s = 0; s < nsrc_here; ++s) tmp += sv;
        const T p1 = other[idx] - deltat * tmp;
        o1[idx] = own + hd * p1 / mm;
        o2[idx] = p1;
    }
}

// Global index of window entry e of plane P, or -1 outside the array.
template <int BY>
__device__ __forceinline__ long z1_entry(const FemGrid &g, int e, int ox, int oy, int P, int &q)
{
    using W = Z1Win<BY>;
    q = e / W::PLANE;
    const int rem = e - q * W::PLANE;
    const int ly = rem / W::EX, lx = rem - ly * W::EX;
    const int X = ox + lx, Y = oy + ly;
    const int dx = g.Nx + (q != 0), dy = g.Ny + (q != 1), dz = g.Nz + (q != 2);
    if (X < 0 || Y < 0 || P < 0 || X >= dx || Y >= dy || P >= dz) return -1;
    return X + (long)dx * (Y + (long)dy * P);
}

template <typename T, int BY, bool DRIFT> struct Z1Regs {
    T v[Z1Win<BY>::NPF];
    T m[DRIFT ? Z1Win<BY>::NPF : 1];
};

template <typename T, int BY, bool DRIFT>
__device__ __forceinline__ void z1_fetch(const FemGrid &g, CComp3<T> u, CComp3<T> pm, int ox, int oy, int P,
                                         Z1Regs<T, BY, DRIFT> &r)
{
    using W = Z1Win<BY>;
    const int tid = threadIdx.y * 64 + threadIdx.x;
#pragma unroll
    for (int t = 0; t < W::NPF; ++t) {
        const int e = tid + t * W::NT;
        r.v[t] = 0;
        if (DRIFT) r.m[t] = 0;
        if (e < W::SLOT) {
            int q;
            const long a = z1_entry<BY>(g, e, ox, oy, P, q);
            if (a >= 0) {
                const T *uq = q == 0 ? u.x : q == 1 ? u.y : u.z;
                r.v[t] = uq[a];
                if (DRIFT) {
                    const T *mq = q == 0 ? pm.x : q == 1 ? pm.y : pm.z;
                    r.m[t] = mq[a];
                }
            }
        }
    }
}

template <typename T, int BY, bool DRIFT>
__device__ __forceinline__ void z1_store(T *win, int P, const Z1Regs<T, BY, DRIFT> &r, T hd, T mm)
{
    using W = Z1Win<BY>;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    T *slot = win + (P & 3) * W::SLOT;
#pragma unroll
    for (int t = 0; t < W::NPF; ++t) {
        const int e = tid + t * W::NT;
        if (e < W::SLOT) slot[e] = DRIFT ? r.v[t] + hd * r.m[t] / mm : r.v[t];
    }
}

// The step of one DOF from its candidates: leapfrog (DRIFT = false)
//   out = ((0 - sum K_e E1) + (2 E1 - E0)) - src
// or symplectic stage (DRIFT = true), as k_fem_leapfrog / k_fem_symplectic.
template <typename T, int BY, bool DRIFT>
__device__ __forceinline__ void z1_dof(const FemGrid &g, const T *sK, const T *win, int comp, int X, int Y, int Z,
                                       int ox, int oy, long idx, T own, const T *__restrict__ other, T *o1, T *o2,
                                       T deltat, T hd, T mm, int nsrc, const long *__restrict__ src_idx, T sv)
{
    int cd[4];
    long rk[4];
    order4<1>(g, comp, X, Y, Z, cd, rk);
    if (!DRIFT) {
        T acc = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int i, j, k, r;
            if (!elem_of<1>(g, comp, X, Y, Z, cd[q], i, j, k, r)) continue;
            acc -= z1_row_dot<T, BY>(sK, win, r, i, j, k, ox, oy);
        }
        T v = acc + ((T)2 * own - other[idx]);
        if (comp == 2)
            for (int s = 0; s < nsrc; ++s)
                if (src_idx[s] == idx) v -= sv;
        o1[idx] = v;
    } else {
        int nsrc_here = 0;
        if (comp == 2)
            for (int s = 0; s < nsrc; ++s) nsrc_here += src_idx[s] == idx;
        T tmp = 0;
        bool src_done = nsrc_here == 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int i, j, k, r;
            if (!elem_of<1>(g, comp, X, Y, Z, cd[q], i, j, k, r)) continue;
            if (!src_done && rk[q] >= g.OE) {
                for (int s = 0; s < nsrc_here; ++s) tmp += sv;
                src_done = true;
            }
            tmp += z1_row_dot<T, BY>(sK, win, r, i, j, k, ox, oy);
        }
        if (!src_done)
            for (int s = 0; s < nsrc_here; ++s) tmp += sv;
        // own = u + ((dt/2) P)/m, the drifted value staged in the window
        const T p1 = other[idx] - deltat * tmp;
        o1[idx] = own + hd * p1 / mm;
        o2[idx] = p1;
    }
}

template <typename T, bool DRIFT>
__device__ __forceinline__ void z1_zero(T *o1, T *o2, long idx)
{
    o1[idx] = (T)0;
    if (DRIFT) o2[idx] = (T)0;
}

// Leapfrog: u = E1, other = E0, out1 = E2.  Symplectic stage: u = U, pm = P,
// other = P (the DOF's own momentum), out1 = U', out2 = P'.  STN: the merged
// stencil (coefficients stc[3][33]) instead of the 4 element row dots.
template <typename T, int BY, bool DRIFT, bool STN>
__global__ __launch_bounds__(64 * BY) void k_fem1z(FemGrid g, const T *__restrict__ K, const T *__restrict__ stc,
                                                    CComp3<T> u, CComp3<T> pm, CComp3<T> other, Comp3<T> out1,
                                                    Comp3<T> out2, int lz, T deltat, T hd, T mm, int nsrc,
                                                    const long *__restrict__ src_idx,
                                                    const double *__restrict__ src_val)
{
    using W = Z1Win<BY>;
    __shared__ T sK[STN ? 1 : 144];
    __shared__ T win[4 * W::SLOT];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    if (!STN && tid < 144) sK[tid] = K[tid];
    const int X0 = blockIdx.x * 64, Y0 = blockIdx.y * BY;
    const int ox = X0 - 1, oy = Y0 - 1;
    const int Z0 = blockIdx.z * lz, Z1 = min(Z0 + lz, g.Nz);
    const int X = X0 + threadIdx.x, Y = Y0 + threadIdx.y;
    const bool act = X < g.Nx && Y < g.Ny;
    const T sv = nsrc > 0 ? (T)src_val[0] : (T)0;
    // warm-up: planes Z0-1 .. Z0+1 (symplectic stages stage the drifted
    // value; a separate drift pass measured slower at first order)
    for (int P = Z0 - 1; P <= Z0 + 1; ++P) {
        Z1Regs<T, BY, DRIFT> r;
        z1_fetch<T, BY, DRIFT>(g, u, pm, ox, oy, P, r);
        z1_store<T, BY, DRIFT>(win, P, r, hd, mm);
    }
    __syncthreads();
    const long sx1 = g.Nx, sx2 = (long)g.Nx * (g.Ny + 1);             // x array strides
    const long sy1 = g.Nx + 1, sy2 = (long)(g.Nx + 1) * g.Ny;         // y array
    const long sz1 = g.Nx + 1, sz2 = (long)(g.Nx + 1) * (g.Ny + 1);   // z array
    const int lo = ((int)threadIdx.x + 1) + W::EX * ((int)threadIdx.y + 1);  // own (X, Y) in the window
    for (int Z = Z0; Z < Z1; ++Z) {
        Z1Regs<T, BY, DRIFT> r;
        z1_fetch<T, BY, DRIFT>(g, u, pm, ox, oy, Z + 2, r);
        if (act) {
            const T *cur = win + (Z & 3) * W::SLOT;
            // the x, y, z edges based at (X, Y, Z), one code path (register
            // pressure); PEC on the low faces here, on X = Nx, Y = Ny, Z = Nz below
#pragma unroll 1
            for (int comp = 0; comp < 3; ++comp) {
                const long i1 = comp == 0 ? sx1 : comp == 1 ? sy1 : sz1;
                const long i2 = comp == 0 ? sx2 : comp == 1 ? sy2 : sz2;
                const long idx = X + i1 * Y + i2 * Z;
                T *a1 = comp == 0 ? out1.x : comp == 1 ? out1.y : out1.z;
                T *a2 = comp == 0 ? out2.x : comp == 1 ? out2.y : out2.z;
                const T *ot = comp == 0 ? other.x : comp == 1 ? other.y : other.z;
                const bool pec = comp == 0 ? (Y == 0 || Z == 0) : comp == 1 ? (X == 0 || Z == 0) : (X == 0 || Y == 0);
                if (pec) z1_zero<T, DRIFT>(a1, a2, idx);
                else if (STN)
                    z1_dof_stn<T, BY, DRIFT>(stc, win, comp, Z, lo, idx, cur[comp * W::PLANE + lo], ot, a1, a2,
                                             deltat, hd, mm, nsrc, src_idx, sv);
                else z1_dof<T, BY, DRIFT>(g, sK, win, comp, X, Y, Z, ox, oy, idx, cur[comp * W::PLANE + lo], ot, a1,
                                          a2, deltat, hd, mm, nsrc, src_idx, sv);
            }
            // the PEC faces X = Nx, Y = Ny, Z = Nz
            const bool ex = X == g.Nx - 1, ey = Y == g.Ny - 1, ez = Z == g.Nz - 1;
            if (ey) {
                z1_zero<T, DRIFT>(out1.x, out2.x, X + sx1 * g.Ny + sx2 * Z);
                z1_zero<T, DRIFT>(out1.z, out2.z, X + sz1 * g.Ny + sz2 * Z);
            }
            if (ex) {
                z1_zero<T, DRIFT>(out1.y, out2.y, g.Nx + sy1 * Y + sy2 * Z);
                z1_zero<T, DRIFT>(out1.z, out2.z, g.Nx + sz1 * Y + sz2 * Z);
            }
            if (ez) {
                z1_zero<T, DRIFT>(out1.x, out2.x, X + sx1 * Y + sx2 * g.Nz);
                z1_zero<T, DRIFT>(out1.y, out2.y, X + sy1 * Y + sy2 * g.Nz);
            }
            if (ey && ez) z1_zero<T, DRIFT>(out1.x, out2.x, X + sx1 * g.Ny + sx2 * g.Nz);
            if (ex && ez) z1_zero<T, DRIFT>(out1.y, out2.y, g.Nx + sy1 * Y + sy2 * g.Nz);
            if (ex && ey) z1_zero<T, DRIFT>(out1.z, out2.z, g.Nx + sz1 * g.Ny + sz2 * Z);
        }
        z1_store<T, BY, DRIFT>(win, Z + 2, r, hd, mm);
        __syncthreads();
    }
}

// Sum pass on the interior edges (both transverse coordinates strictly
// inside): out = (merged stencil of x, coefficients stc[3][33]) + base, the
// src/fem_abm element passes with their "+= slot 1"
// (src/fem_abm/fem_simulation.cpp:319-369, :619-710).  The window and plane
// march of k_fem1z over x; the face edges are another launch's.
template <typename T, int BY>
__global__ __launch_bounds__(64 * BY) void k_fem1z_sum(FemGrid g, const T *__restrict__ stc, CComp3<T> x,
                                                       CComp3<T> base, Comp3<T> out, int lz)
{
    using W = Z1Win<BY>;
    __shared__ T win[4 * W::SLOT];
    const int X0 = blockIdx.x * 64, Y0 = blockIdx.y * BY;
    const int ox = X0 - 1, oy = Y0 - 1;
    const int Z0 = blockIdx.z * lz, Z1 = min(Z0 + lz, g.Nz);
    const int X = X0 + threadIdx.x, Y = Y0 + threadIdx.y;
    const bool act = X < g.Nx && Y < g.Ny;
    for (int P = Z0 - 1; P <= Z0 + 1; ++P) {
        Z1Regs<T, BY, false> r;
        z1_fetch<T, BY, false>(g, x, x, ox, oy, P, r);
        z1_store<T, BY, false>(win, P, r, (T)0, (T)1);
    }
    __syncthreads();
    const long sx1 = g.Nx, sx2 = (long)g.Nx * (g.Ny + 1);
    const long sy1 = g.Nx + 1, sy2 = (long)(g.Nx + 1) * g.Ny;
    const long sz1 = g.Nx + 1, sz2 = (long)(g.Nx + 1) * (g.Ny + 1);
    const int lo = ((int)threadIdx.x + 1) + W::EX * ((int)threadIdx.y + 1);
    for (int Z = Z0; Z < Z1; ++Z) {
        Z1Regs<T, BY, false> r;
        z1_fetch<T, BY, false>(g, x, x, ox, oy, Z + 2, r);
        if (act) {
            if (Y > 0 && Z > 0) {
                const long i = X + sx1 * Y + sx2 * Z;
                out.x[i] = z1_stencil<T, BY, 0>(stc, win, Z, lo) + base.x[i];
            }
            if (X > 0 && Z > 0) {
                const long i = X + sy1 * Y + sy2 * Z;
                out.y[i] = z1_stencil<T, BY, 1>(stc + kZ1Terms, win, Z, lo) + base.y[i];
            }
            if (X > 0 && Y > 0) {
                const long i = X + sz1 * Y + sz2 * Z;
                out.z[i] = z1_stencil<T, BY, 2>(stc + 2 * kZ1Terms, win, Z, lo) + base.z[i];
            }
        }
        z1_store<T, BY, false>(win, Z + 2, r, (T)0, (T)1);
        __syncthreads();
    }
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
