// First-order Nedelec (Whitney) edge elements on the Kuhn tetrahedral split
// of a hex grid (no reference counterpart; nearest hooks: the order-p hex
// operator of src/fem_v2/fem_kernels.hpp:13-175 and setSourceDipole_x of
// src/fem/fem_simulation.cpp:454).
//
//   M_eps U'' + M_sig U' + S U = -g(t),  g_i = p'(t) (w_i(r0) . d)  (dipole d at r0)
//   M_eps = sum_T eps0 eps_r(T) M_T,  M_sig = sum_T sig(T) M_T,  S = sum_T S_T / mu
// with PEC (tangential E = 0) on the outer faces.  The consistent mass is
// solved, not lumped: a Jacobi-preconditioned conjugate gradient on the
// device (reductions in a fixed order, convergence tested on the device).
// Steppers (U = E, V = E'):
//   symplectic_2 / _4 (the stage of src/symplectic_2/fem_simulation.cpp:557-672
//   with the mass solve, semi-implicit in sigma):
//      U += (D/2) V;  (M_eps + D/2 M_sig) V' = (M_eps - D/2 M_sig) V - D (S U + g(t + o + D/2));
//      U += (D/2) V'
//     symplectic_4: stages (x1 dt, 0), (x2 dt, x1 dt), (x1 dt, (x1+x2) dt)
//   ABM (the AB2 predictor / AM3 corrector of src/fem_abm/fem_simulation.cpp:
//   278-833 on y = (U, V), y' = (V, a), M_eps a = -S U - M_sig V - g):
//      y_p = y_n + dt/2 (3 F_n - F_{n-1});  y_{n+1} = y_n + dt/12 (5 F(y_p) + 8 F_n - F_{n-1})
// The matrices share one CSR pattern, assembled on the host in a fixed
// order (entries summed in (row, col, tet) order).
#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <thread>
#include <vector>

#include "common.hpp"
#include "lane_addr.hpp"
#include "mfma.hpp"
#include "tet_basis.hpp"
#include "tet_mesh.hpp"

using namespace ems;
using namespace ems::tet;

namespace {

template <typename T> struct Csr {
    int n;
    const int *rp, *col;
};

// Sparse rows are processed by LPR consecutive lanes (LPR = 4 or 8, from the
// longest row): lane l sums entries rp[r] + l, + 2 LPR, ... in order, and the
// lanes combine by a fixed butterfly, so every product is deterministic.
template <int LPR, typename T> __device__ __forceinline__ T lanes_sum(T v)
{
#pragma unroll
    for (int m = 1; m < LPR; m <<= 1) v += __shfl_xor(v, m, LPR);
    return v;
}

template <int LPR, typename T>
__device__ __forceinline__ T row_dot(const Csr<T> &A, const T *__restrict__ val, const T *__restrict__ x, int r,
                                     int lane)
{
    T s = 0;
    for (int q = A.rp[r] + lane; q < A.rp[r + 1]; q += LPR) s += val[q] * x[A.col[q]];
    return lanes_sum<LPR>(s);
}

// ---------------------------------------------------------------------------
// Element-by-element operator (no assembled matrix).  Phase 1, one thread per
// tetrahedron, waves of one tetrahedron type (so the type's nl x nl element
// matrices are read at uniform addresses): gather the tetrahedron's nl dofs
// (boundary dofs as 0), form the local product(s) and store them type-major,
// tmp[(t nl + l) ncube + cube], coalesced over cubes.  Phase 2, one thread per
// dof: sum the contributions of the tetrahedra containing its entity in the
// fixed order of the row type's incidence list; boundary rows are identity
// rows of the mass-type operators.  Nothing is accumulated with atomics.
// ---------------------------------------------------------------------------
struct EbeMesh {
    int nx, ny, nz, nl;
    long ne, nf, nd, ncube;
    long off[8];       // first edge of each edge family
    long foff[13];     // first face of each face family
    int ecx[7], ecy[7];  // edge family extents (index = off + I + cx (J + cy K))
    int fcx[12], fcy[12];
    int tedge[6][6];   // per tet type and local edge: family | base offset bits << 4
    int tface[6][4];   // per tet type and local face: face family | base offset bits << 4
    int inc_start[39];  // per row type: first entry of its incidence list
};

__device__ __forceinline__ long ebe_at(const long *off, const int *cx, const int *cy, int code, int ci, int cj, int ck)
{
    const int f = code & 15;
    return off[f] + (ci + ((code >> 4) & 1)) + (long)cx[f] * ((cj + ((code >> 5) & 1)) + (long)cy[f] * (ck + ((code >> 6) & 1)));
}

// global dof of local dof l of tetrahedron type t in cube (ci, cj, ck)
__device__ __forceinline__ long ebe_dof(const EbeMesh &m, int t, int l, int ci, int cj, int ck)
{
    if (l < 6) return ebe_at(m.off, m.ecx, m.ecy, m.tedge[t][l], ci, cj, ck);
    if (l < 12) return m.ne + ebe_at(m.off, m.ecx, m.ecy, m.tedge[t][l - 6], ci, cj, ck);
    const int q = (l - 12) >> 1;
    return 2 * m.ne + 2 * ebe_at(m.foff, m.fcx, m.fcy, m.tface[t][q], ci, cj, ck) + ((l - 12) & 1);
}

// (I, J, K) of index r = I + cx (J + cy K), in 32-bit unsigned arithmetic
// (ems_tet_create keeps every cube and family index below 2^31; 64-bit
// division expands to a long instruction sequence per lane)
__device__ __forceinline__ void cube_ijk(unsigned r, unsigned cx, unsigned cy, int &I, int &J, int &K)
{
    const unsigned q = r / cx;
    I = (int)(r - q * cx);
    J = (int)(q % cy);
    K = (int)(q / cy);
}

// row type and base vertex of a dof (see row_types on the host)
__device__ __forceinline__ int ebe_row(const EbeMesh &m, long r, int &I, int &J, int &K)
{
    if (r < 2 * m.ne) {
        const int kind = r >= m.ne;
        const long e = kind ? r - m.ne : r;
        int f = 0;
#pragma unroll
        for (int q = 1; q < 7; ++q) f += e >= m.off[q];
        cube_ijk((unsigned)(e - m.off[f]), (unsigned)m.ecx[f], (unsigned)m.ecy[f], I, J, K);
        return f + 7 * kind;
    }
    const long fd = r - 2 * m.ne;
    const long fid = fd >> 1;
    int q = 0;
#pragma unroll
    for (int w = 1; w < 12; ++w) q += fid >= m.foff[w];
    cube_ijk((unsigned)(fid - m.foff[q]), (unsigned)m.fcx[q], (unsigned)m.fcy[q], I, J, K);
    return 14 + 2 * q + (int)(fd & 1);
}

// phase 2: sum over the tetrahedra containing dof r, in incidence order.
// All loads of the row are issued before the sum (at most kMaxInc entries;
// entries outside the mesh contribute +0).  When the wave holds one row type
// (all but the waves at family boundaries) the incidence entries are
// wave-uniform and each tmp address is a uniform base plus the lane's cube
// offset (lane_at; 6-7 % faster at order 1); otherwise per-lane addresses.
// The cube offset in bytes stays below 2^32 (the mesh-size limit of
// ems_tet_create keeps ncube under 2.9e8).
constexpr int kMaxInc = 8;

template <typename T>
__device__ __forceinline__ T ebe_gather(const EbeMesh &m, const int *__restrict__ inc, const T *__restrict__ tmp, long r)
{
    int I, J, K;
    const int rt = ebe_row(m, r, I, J, K);
    const int q0 = m.inc_start[rt], cnt = m.inc_start[rt + 1] - q0;
    T v[kMaxInc];
    if (__all(rt == __builtin_amdgcn_readfirstlane(rt))) {
        // one row type on the wave: the incidence entries are uniform, and
        // each tmp address is a uniform base plus the lane's cube offset
        const unsigned co = (unsigned)(I + m.nx * (J + m.ny * K)) * (unsigned)sizeof(T);
        const int rtu = __builtin_amdgcn_readfirstlane(rt);
        const int q0u = m.inc_start[rtu], cntu = m.inc_start[rtu + 1] - q0u;
#pragma unroll
        for (int q = 0; q < kMaxInc; ++q) {
            v[q] = 0;
            if (q < cntu) {
                const int c = inc[q0u + q];
                const int di = (c & 3) - 1, dj = ((c >> 2) & 3) - 1, dk = ((c >> 4) & 3) - 1;
                const int ci = I + di, cj = J + dj, ck = K + dk;
                const int t = (c >> 6) & 7, L = (c >> 9) & 31;
                const long sh = (long)(t * m.nl + L) * m.ncube + di + (long)m.nx * (dj + (long)m.ny * dk);
                if (ci >= 0 && cj >= 0 && ck >= 0 && ci < m.nx && cj < m.ny && ck < m.nz)
                    v[q] = lane_at(tmp + sh, co);
            }
        }
        T s = 0;
#pragma unroll
        for (int q = 0; q < kMaxInc; ++q) s += v[q];
        return s;
    }
#pragma unroll
    for (int q = 0; q < kMaxInc; ++q) {
        v[q] = 0;
        if (q < cnt) {
            const int c = inc[q0 + q];
            const int ci = I + (c & 3) - 1, cj = J + ((c >> 2) & 3) - 1, ck = K + ((c >> 4) & 3) - 1;
            if (ci >= 0 && cj >= 0 && ck >= 0 && ci < m.nx && cj < m.ny && ck < m.nz) {
                const int t = (c >> 6) & 7, L = (c >> 9) & 31;
                v[q] = tmp[(long)(t * m.nl + L) * m.ncube + ci + (long)m.nx * (cj + (long)m.ny * ck)];
            }
        }
    }
    T s = 0;
#pragma unroll
    for (int q = 0; q < kMaxInc; ++q) s += v[q];
    return s;
}

// Phase 1 modes: MASS y = (eps0 eps_r + hh sigma) M x1;  STIFF y = S x1;
// FORCE y = -S x2 - sigma M x1 (S already divided by mu; x1 = V, x2 = U).
// First order applies S = (V/mu) C^T C as the element curl C x followed by
// (V/mu) C^T (Ct: [type][C | (V/mu) C], 3 x 6 each): a curl-free field then
// gives a curl that is zero up to the rounding of C x alone, so the rounding
// of S x no longer leaks ~ (V/mu) |C|^2 |x| into the curl-free null space,
// which no oscillation removes (in FP32 it made the fields drift).
enum { EBE_MASS = 0, EBE_STIFF = 1, EBE_FORCE = 3 };

// cmt (MASS, optional): the coefficient (T)(e0 eps_r + hh sigma) per
// tetrahedron, type-major, computed by k_mass_coef with the same expression.
template <typename T, int NL, int MODE>
__global__ __launch_bounds__(256) void k_ebe_local(EbeMesh m, const T *__restrict__ Mt, const T *__restrict__ St,
                                                   const double *__restrict__ er, const double *__restrict__ sg,
                                                   const T *__restrict__ Ct, double e0, double hh,
                                                   const unsigned *__restrict__ tmask, const T *__restrict__ x1,
                                                   const T *__restrict__ x2, T *__restrict__ tmp, const int *done,
                                                   const T *__restrict__ cmt)
{
    if (done && *done) return;
    const long cube = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (cube >= m.ncube) return;
    int ci, cj, ck;
    cube_ijk((unsigned)cube, (unsigned)m.nx, (unsigned)m.ny, ci, cj, ck);
    const long tid = 6 * cube + t;
    const bool pre = MODE == EBE_MASS && cmt;
    const double epsr = (!pre && er) ? er[tid] : 1.0, sig = (!pre && sg) ? sg[tid] : 0.0;
    const unsigned bm = tmask[(long)t * m.ncube + cube];  // bit l: local dof l is a boundary dof
    constexpr bool USE2 = MODE == EBE_FORCE;
    T a[NL], b[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const long d = ebe_dof(m, t, l, ci, cj, ck);
        const bool in = !((bm >> l) & 1u);
        a[l] = in ? x1[d] : (T)0;
        b[l] = (USE2 && in) ? x2[d] : (T)0;
    }
    const T *M = Mt + t * NL * NL, *S = St + t * NL * NL;
    // first order: the element curl of the stiffness operand (x1 in STIFF,
    // x2 in FORCE)
    T cu[3] = {0, 0, 0};
    if constexpr (NL == 6 && MODE != EBE_MASS) {
        const T *C = Ct + t * 36;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            T v = 0;
#pragma unroll
            for (int q = 0; q < 6; ++q) v += C[k * 6 + q] * (MODE == EBE_STIFF ? a[q] : b[q]);
            cu[k] = v;
        }
    }
    T cm;
    if (MODE == EBE_MASS) cm = pre ? cmt[(long)t * m.ncube + cube] : (T)(e0 * epsr + hh * sig);
    else cm = (T)sig;
#pragma unroll 4
    for (int l = 0; l < NL; ++l) {
        T y;
        if (MODE == EBE_STIFF && NL == 6) {
            const T *Cs = Ct + t * 36 + 18;
            y = (Cs[l] * cu[0] + Cs[6 + l] * cu[1]) + Cs[12 + l] * cu[2];
        } else if (MODE == EBE_STIFF) {
            T s = 0;
#pragma unroll
            for (int q = 0; q < NL; ++q) s += S[l * NL + q] * a[q];
            y = s;
        } else if (MODE == EBE_MASS) {
            T s = 0;
#pragma unroll
            for (int q = 0; q < NL; ++q) s += M[l * NL + q] * a[q];
            y = cm * s;
        } else {
            T sm = 0, ss = 0;
            if constexpr (NL == 6) {
                const T *Cs = Ct + t * 36 + 18;
                ss = (Cs[l] * cu[0] + Cs[6 + l] * cu[1]) + Cs[12 + l] * cu[2];
#pragma unroll
                for (int q = 0; q < NL; ++q) sm += M[l * NL + q] * a[q];
            } else {
#pragma unroll
                for (int q = 0; q < NL; ++q) {
                    sm += M[l * NL + q] * a[q];
                    ss += S[l * NL + q] * b[q];
                }
            }
            y = -ss - cm * sm;
        }
        tmp[(long)(t * NL + l) * m.ncube + cube] = y;
    }
}

// Phase 1 of the mass operator with matrix cores: a wave owns 16
// tetrahedra of one type in consecutive cubes and forms Y = M_t X (NL x 16)
// with 16x16x4 MFMA tiles (rows padded to 16, K in steps of 4), then scales
// column n by the tetrahedron's eps0 eps_r + hh sigma.  Operand layouts of
// the 16x16x4 builtins: A[i][k] in lane (k 16 + i), B[k][j] in lane (k 16 + j);
// result element v of lane l is D[4 (l / 16) + v][l % 16] (f32) or
// D[l / 16 + 4 v][l % 16] (f64).

// Mass coefficient (T)(e0 eps_r + hh sigma) of every tetrahedron, type-major
// (the expression of k_ebe_local, so the same bits).
template <typename T>
__global__ void k_mass_coef(long ncube, const double *__restrict__ er, const double *__restrict__ sg, double e0,
                            double hh, T *__restrict__ out)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 6 * ncube) return;
    const long t = q / ncube, cube = q - t * ncube, tid = 6 * cube + t;
    const double epsr = er ? er[tid] : 1.0, sig = sg ? sg[tid] : 0.0;
    out[q] = (T)(e0 * epsr + hh * sig);
}

template <typename T, int NL>
__global__ __launch_bounds__(256) void k_ebe_mass_mfma(EbeMesh m, const T *__restrict__ Mt,
                                                       const double *__restrict__ er, const double *__restrict__ sg,
                                                       double e0, double hh, const unsigned *__restrict__ tmask,
                                                       const T *__restrict__ x, T *__restrict__ tmp, const int *done)
{
    if (done && *done) return;
    constexpr int MT = (NL + 15) / 16, KS = (NL + 3) / 4;
    const int t = blockIdx.y, lane = threadIdx.x & 63;
    const long base = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;  // first cube of this wave
    if (base >= m.ncube) return;  // uniform per wave; no barrier below
    const int n = lane & 15, kq = lane >> 4;
    const long cube = base + n;
    const bool valid = cube < m.ncube;
    int ci = 0, cj = 0, ck = 0;
    if (valid) cube_ijk((unsigned)cube, (unsigned)m.nx, (unsigned)m.ny, ci, cj, ck);
    const T *M = Mt + t * NL * NL;
    typedef typename MfmaAcc<T>::V V;
    V acc[MT];
#pragma unroll
    for (int q = 0; q < MT; ++q) acc[q] = V{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const int k = 4 * ks + kq;
        T bv = 0;
        if (valid && k < NL) {
            const long d = ebe_dof(m, t, k, ci, cj, ck);
            if (!((tmask[(long)t * m.ncube + cube] >> k) & 1u)) bv = x[d];
        }
#pragma unroll
        for (int q = 0; q < MT; ++q) {
            const int i = 16 * q + n;
            const T av = (i < NL && k < NL) ? M[i * NL + k] : (T)0;
            acc[q] = MfmaAcc<T>::step(av, bv, acc[q]);
        }
    }
    if (!valid) return;
    const long tid = 6 * cube + t;
    const T c = (T)(e0 * (er ? er[tid] : 1.0) + hh * (sg ? sg[tid] : 0.0));
#pragma unroll
    for (int q = 0; q < MT; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 16 * q + MfmaAcc<T>::row(lane, v);
            if (row < NL) tmp[(long)(t * NL + row) * m.ncube + cube] = c * acc[q][v];
        }
}

// Operator policies of the solver kernels: apply(x, r, lane) returns (A x)_r
// for lane 0 of the row's lanes.
template <int LPR, typename T> struct CsrOp {
    static constexpr int lanes = LPR;
    Csr<T> A;
    const T *val;
    __device__ __forceinline__ int n() const { return A.n; }
    __device__ __forceinline__ T apply(const T *__restrict__ x, int r, int lane) const
    {
        return row_dot<LPR>(A, val, x, r, lane);
    }
};

template <typename T> struct EbeOp {
    static constexpr int lanes = 1;
    EbeMesh m;
    const int *inc;
    const T *tmp;
    const unsigned char *bnd;
    int ident;  // boundary rows: 1 identity (mass-type), 0 zero
    __device__ __forceinline__ int n() const { return (int)m.nd; }
    __device__ __forceinline__ T apply(const T *__restrict__ x, int r, int) const
    {
        if (bnd[r]) return ident ? x[r] : (T)0;
        return ebe_gather(m, inc, tmp, r);
    }
};

// The mass-type operator (eps0 eps_r + hh sigma) M in one pass per row, with
// no per-tetrahedron buffer: row r forms, for each tetrahedron of its
// incidence list (fixed order), that tetrahedron's local product row L,
// cm * sum_l M_t[L][l] x_l (boundary dofs as 0), and sums them.  The same
// expressions in the same order as phase 1 + phase 2 of the element-by-element
// operator, so the same bits; the tetrahedron's dofs are gathered once per
// row instead of once per tetrahedron (cache hits), and the tmp round trip
// through HBM disappears.  Boundary rows are identity rows.
template <typename T, int NL> struct EbeRowOp {
    static constexpr int lanes = 1;
    EbeMesh m;
    const int *inc;
    const T *Mt;             // [type][NL][NL]
    const unsigned *tmask;   // [type][cube]: bit l = local dof l on the boundary
    const T *cmt;            // [type][cube]: eps0 eps_r + hh sigma
    const unsigned char *bnd;
    __device__ __forceinline__ int n() const { return (int)m.nd; }
    __device__ __forceinline__ T apply(const T *__restrict__ x, int r, int) const
    {
        if (bnd[r]) return x[r];
        int I, J, K;
        const int rt = ebe_row(m, r, I, J, K);
        const int q0 = m.inc_start[rt], cnt = m.inc_start[rt + 1] - q0;
        T s = 0;
#pragma unroll
        for (int q = 0; q < kMaxInc; ++q) {
            T v = 0;
            if (q < cnt) {
                const int c = inc[q0 + q];
                const int ci = I + (c & 3) - 1, cj = J + ((c >> 2) & 3) - 1, ck = K + ((c >> 4) & 3) - 1;
                if (ci >= 0 && cj >= 0 && ck >= 0 && ci < m.nx && cj < m.ny && ck < m.nz) {
                    const int t = (c >> 6) & 7, L = (c >> 9) & 31;
                    const long cube = ci + (long)m.nx * (cj + (long)m.ny * ck);
                    const unsigned bm = tmask[(long)t * m.ncube + cube];
                    const T *Mr = Mt + (t * NL + L) * NL;
                    T acc = 0;
#pragma unroll
                    for (int l = 0; l < NL; ++l) {
                        const T a = ((bm >> l) & 1u) ? (T)0 : x[ebe_dof(m, t, l, ci, cj, ck)];
                        acc += Mr[l] * a;
                    }
                    v = cmt[(long)t * m.ncube + cube] * acc;
                }
            }
            s += v;
        }
        return s;
    }
};

// y = A x (phase 2 / CSR row products)
template <typename OP, typename T>
__global__ __launch_bounds__(256) void k_apply(OP op, const T *__restrict__ x, T *__restrict__ y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = t / OP::lanes, lane = t % OP::lanes;
    if (r >= op.n()) return;
    const T v = op.apply(x, r, lane);
    if (lane == 0) y[r] = v;
}

// b = phase-2 gather with boundary rows b = V (the stage right-hand side) or
// 0 (the ABM force)
template <typename T>
__global__ __launch_bounds__(256) void k_ebe_rhs(EbeMesh m, const int *__restrict__ inc, const T *__restrict__ tmp,
                                                 const unsigned char *__restrict__ bnd, const T *__restrict__ V,
                                                 int keep_v, T *__restrict__ b)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m.nd) return;
    b[r] = bnd[r] ? (keep_v ? V[r] : (T)0) : ebe_gather(m, inc, tmp, r);
}

// Jacobi diagonal of the mass-type operator (eps0 eps_r + hh sigma) M: dinv
template <typename T>
__global__ __launch_bounds__(256) void k_ebe_diag(EbeMesh m, const int *__restrict__ inc, const T *__restrict__ Mt,
                                                  const double *__restrict__ er, const double *__restrict__ sg,
                                                  double e0, double hh, const unsigned char *__restrict__ bnd,
                                                  T *__restrict__ dinv)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m.nd) return;
    if (bnd[r]) {
        dinv[r] = 1;
        return;
    }
    int I, J, K;
    const int rt = ebe_row(m, r, I, J, K);
    double s = 0.0;
    for (int q = m.inc_start[rt]; q < m.inc_start[rt + 1]; ++q) {
        const int c = inc[q];
        const int ci = I + (c & 3) - 1, cj = J + ((c >> 2) & 3) - 1, ck = K + ((c >> 4) & 3) - 1;
        if (ci < 0 || cj < 0 || ck < 0 || ci >= m.nx || cj >= m.ny || ck >= m.nz) continue;
        const int t = (c >> 6) & 7, L = (c >> 9) & 31;
        const long tid = 6 * (ci + (long)m.nx * (cj + (long)m.ny * ck)) + t;
        const double co = e0 * (er ? er[tid] : 1.0) + hh * (sg ? sg[tid] : 0.0);
        s += co * (double)Mt[(t * m.nl + L) * m.nl + L];
    }
    dinv[r] = (T)(1.0 / s);
}

// y = A x
template <int LPR, typename T>
__global__ __launch_bounds__(256) void k_spmv(Csr<T> A, const T *__restrict__ val, const T *__restrict__ x,
                                              T *__restrict__ y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = t / LPR, lane = t % LPR;
    if (r >= A.n) return;
    const T v = row_dot<LPR>(A, val, x, r, lane);
    if (lane == 0) y[r] = v;
}

// b = -S U - Ms V (the force of the first-order system)
template <int LPR, typename T, bool LOSSY>
__global__ __launch_bounds__(256) void k_force(Csr<T> A, const T *__restrict__ ms, const T *__restrict__ st,
                                               const T *__restrict__ U, const T *__restrict__ V, T *__restrict__ b)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = t / LPR, lane = t % LPR;
    if (r >= A.n) return;
    T c = 0, s = 0;
    for (int q = A.rp[r] + lane; q < A.rp[r + 1]; q += LPR) {
        const int j = A.col[q];
        s += st[q] * U[j];
        if (LOSSY) c += ms[q] * V[j];
    }
    s = lanes_sum<LPR>(s);
    if (LOSSY) c = lanes_sum<LPR>(c);
    if (lane == 0) b[r] = LOSSY ? -s - c : -s;
}

// b[dof] += f[dipole] * w (dipole loads, nl weights per dipole), one thread
// in dipole order so that loads sharing a dof add up in a fixed order
template <typename T>
__global__ void k_load(T *b, int nw, int nl, const int *__restrict__ dof, const double *__restrict__ w,
                       const double *__restrict__ f)
{
    if (threadIdx.x || blockIdx.x) return;
    for (int q = 0; q < nw; ++q) b[dof[q]] = b[dof[q]] + (T)(f[q / nl] * w[q]);
}

// ---------------------------------------------------------------------------
// Preconditioned conjugate gradients, two kernels per iteration (the
// direction recurrence Ap <- A z + beta Ap lets the product and the new
// direction share one pass).  Dot products: per-block partials (rows in a
// fixed order, then a fixed LDS tree); every block re-reduces the partials
// it needs in the same fixed order, so all blocks see the same scalars and
// no scalar kernels or grid barriers are needed (a last-block reduction
// would need a device-scope fence per block, which costs more here).  The
// block count nb grows with the rows (cg_blocks); partial buffers of nb
// entries each:
//   P[0], P[1]  (r, z) of even / odd updates,  P[2] (p, Ap),  P[3] (b, M^-1 b)
// The solve stops when (r, z) <= tol^2 (b, M^-1 b); b = 0 gives x = 0.
// Block 0 records done and the iteration count for the host.
// ---------------------------------------------------------------------------
constexpr int kRB = 256;         // threads per block
constexpr int kCgMaxNB = 16384;  // most blocks of the CG kernels (partials per dot)
constexpr int kCgInlineNB = 512; // up to this many blocks every block re-reduces the partials

// CG blocks for n rows: about 1024 rows per block, 512 to kCgMaxNB.  Above
// kCgInlineNB the partials are reduced once per dot by a one-block kernel
// (the same fixed order), and the consumers read the scalar.
inline int cg_blocks(long n) { return (int)std::min<long>(kCgMaxNB, std::max<long>(512, (n + 1023) / 1024)); }

struct CgState {
    int done, iters;
    double scal[4];  // reduced dots (slots as the partial buffers) when nb > kCgInlineNB
};

// scalar of partial slot `slot`: re-reduced in place or read after k_cg_finalize
__device__ __forceinline__ double block_total(const double *__restrict__ P, int nb);
__device__ __forceinline__ double cg_dot(const double *__restrict__ P, const CgState *st, int slot, int nb)
{
    return nb > kCgInlineNB ? st->scal[slot] : block_total(P + slot * nb, nb);
}

// sum of nb partials, the same fixed order in every block; all threads get it
__device__ __forceinline__ double block_total(const double *__restrict__ P, int nb)
{
    __shared__ double sh[kRB];
    double s = 0.0;
    for (int q = threadIdx.x; q < nb; q += kRB) s += P[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kRB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double out = sh[0];
    __syncthreads();
    return out;
}

// per-block partial of the per-thread sums v
__device__ __forceinline__ void block_partial(double v, double *P)
{
    __shared__ double sh[kRB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kRB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) P[blockIdx.x] = sh[0];
    __syncthreads();
}

__device__ __forceinline__ bool cg_converged(double rz, double ref, double tol2) { return !(rz > tol2 * ref); }

// r = b - A x, z = M^-1 r; partials (r, z) -> P0 and (b, M^-1 b) -> P3
template <typename OP, typename T>
__global__ __launch_bounds__(kRB) void k_cg_init(OP op, const T *__restrict__ dinv, const T *__restrict__ b,
                                                 const T *__restrict__ x, T *r, T *z, double *P, CgState *st, int nb)
{
    constexpr int LPR = OP::lanes;
    double rz = 0.0, bb = 0.0;
    const int stride = gridDim.x * kRB / LPR;
    for (int i = (blockIdx.x * kRB + threadIdx.x) / LPR; i < op.n(); i += stride) {
        const int lane = threadIdx.x % LPR;
        const T ax = op.apply(x, i, lane);
        if (lane == 0) {
            const T ri = b[i] - ax, zi = dinv[i] * ri;
            r[i] = ri;
            z[i] = zi;
            rz += (double)ri * (double)zi;
            bb += (double)b[i] * (double)(dinv[i] * b[i]);
        }
    }
    block_partial(rz, P);
    block_partial(bb, P + 3 * nb);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->done = 0;
        st->iters = 0;
    }
}

// p = z, Ap = A z; partial (p, Ap) -> P2 (or done when (r, z) is small
// already; x = 0 when b = 0)
template <typename OP, typename T>
__global__ __launch_bounds__(kRB) void k_cg_first(OP op, const T *__restrict__ z, T *p, T *ap, T *x, double *P,
                                                  CgState *st, double tol2, int nb)
{
    constexpr int LPR = OP::lanes;
    const double rz = cg_dot(P, st, 0, nb), ref = cg_dot(P, st, 3, nb);
    if (!(ref > 0.0)) {
        for (int i = blockIdx.x * kRB + threadIdx.x; i < op.n(); i += gridDim.x * kRB) x[i] = 0;
        if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
        return;
    }
    if (cg_converged(rz, ref, tol2)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
        return;
    }
    double pap = 0.0;
    const int stride = gridDim.x * kRB / LPR;
    for (int i = (blockIdx.x * kRB + threadIdx.x) / LPR; i < op.n(); i += stride) {
        const int lane = threadIdx.x % LPR;
        const T q = op.apply(z, i, lane);
        if (lane == 0) {
            const T zi = z[i];
            p[i] = zi;
            ap[i] = q;
            pap += (double)zi * (double)q;
        }
    }
    block_partial(pap, P + 2 * nb);
}

// alpha = (r, z) / (p, Ap); x += alpha p, r -= alpha Ap, z = M^-1 r;
// partial (r, z) -> P[par ^ 1]
template <typename T>
__global__ __launch_bounds__(kRB) void k_cg_update(int n, T *x, T *r, T *z, const T *__restrict__ p,
                                                   const T *__restrict__ ap, const T *__restrict__ dinv, double *P,
                                                   int par, const CgState *st, int nb)
{
    if (st->done) return;
    const double rz = cg_dot(P, st, par, nb), pap = cg_dot(P, st, 2, nb);
    const T a = (T)(rz / pap);
    double s = 0.0;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += gridDim.x * kRB) {
        x[i] = x[i] + a * p[i];
        const T ri = r[i] - a * ap[i], zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        s += (double)ri * (double)zi;
    }
    block_partial(s, P + (par ^ 1) * nb);
}

// converged?  else beta = (r, z)_new / (r, z)_old; p = z + beta p,
// Ap = A z + beta Ap; partial (p, Ap) -> P2
template <typename OP, typename T>
__global__ __launch_bounds__(kRB) void k_cg_direction(OP op, const T *__restrict__ z, T *p, T *ap, double *P, int par,
                                                      CgState *st, int it, double tol2, int nb)
{
    constexpr int LPR = OP::lanes;
    if (st->done) return;
    const double rz_old = cg_dot(P, st, par, nb), rz = cg_dot(P, st, par ^ 1, nb), ref = cg_dot(P, st, 3, nb);
    if (cg_converged(rz, ref, tol2) || !(rz_old > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->done = 1;
            st->iters = it + 1;
        }
        return;
    }
    const T beta = (T)(rz / rz_old);
    double pap = 0.0;
    const int stride = gridDim.x * kRB / LPR;
    for (int i = (blockIdx.x * kRB + threadIdx.x) / LPR; i < op.n(); i += stride) {
        const int lane = threadIdx.x % LPR;
        const T q = op.apply(z, i, lane);
        if (lane == 0) {
            const T pi = z[i] + beta * p[i], api = q + beta * ap[i];
            p[i] = pi;
            ap[i] = api;
            pap += (double)pi * (double)api;
        }
    }
    block_partial(pap, P + 2 * nb);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->iters = it + 1;
}

// ---------------------------------------------------------------------------
// Chebyshev acceleration of Jacobi for the mass solves (Saad, Iterative
// Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1): with the spectrum
// of D^-1 A inside [lo, hi], theta = (hi + lo)/2, delta = (hi - lo)/2,
// sigma = theta/delta, rho_0 = 1/sigma, d_0 = D^-1 r_0 / theta and, for k >= 1,
// rho_k = 1/(2 sigma - rho_k-1), d_k = rho_k rho_k-1 d_k-1 + (2 rho_k/delta)
// D^-1 r_k, x_k+1 = x_k + d_k.  The coefficients are known in advance: no
// inner products (no reductions, no host polls) between the first and the
// last residual; the iteration count follows from the Chebyshev bound
// 2 q^k, q = (sqrt(kappa) - 1)/(sqrt(kappa) + 1), kappa = hi/lo.
// ---------------------------------------------------------------------------
// r = b - A x (row i), z = D^-1 r, d = c1 d + c2 z, xn = x + d  (first: d = c2 z)
template <typename OP, typename T>
__global__ __launch_bounds__(kRB) void k_cheb_step(OP op, const T *__restrict__ b, const T *__restrict__ dinv,
                                                   const T *__restrict__ x, T *__restrict__ d, T *__restrict__ xn,
                                                   T c1, T c2, int first)
{
    constexpr int LPR = OP::lanes;
    const int stride = gridDim.x * kRB / LPR;
    for (int i = (blockIdx.x * kRB + threadIdx.x) / LPR; i < op.n(); i += stride) {
        const int lane = threadIdx.x % LPR;
        const T ax = op.apply(x, i, lane);
        if (lane == 0) {
            const T zi = dinv[i] * (b[i] - ax);
            const T di = first ? c2 * zi : c1 * d[i] + c2 * zi;
            d[i] = di;
            xn[i] = x[i] + di;
        }
    }
}

// per-block partials of sum a_i b_i (w: 0) or sum a_i b_i / dinv_i (w: 1)
template <typename T>
__global__ __launch_bounds__(kRB) void k_dot(int n, const T *__restrict__ a, const T *__restrict__ b,
                                             const T *__restrict__ dinv, int w, double *P)
{
    double s = 0.0;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += gridDim.x * kRB) {
        const double v = (double)a[i] * (double)b[i];
        s += w ? v / (double)dinv[i] : v;
    }
    block_partial(s, P);
}

// Lanczos vector update of D^-1 A in the D inner product:
// w = D^-1 y - alpha v - beta vp
template <typename T>
__global__ void k_lanczos_w(int n, const T *__restrict__ dinv, const T *__restrict__ y, const T *__restrict__ v,
                            const T *__restrict__ vp, double alpha, double beta, T *__restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    w[i] = (T)((double)dinv[i] * (double)y[i] - alpha * (double)v[i] - beta * (double)vp[i]);
}

// v = x / s
template <typename T> __global__ void k_scale_to(int n, const T *__restrict__ x, double s, T *__restrict__ v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (T)((double)x[i] / s);
}

// deterministic start vector in [-1/2, 1/2] (integer hash of the index)
template <typename T> __global__ void k_hash_fill(int n, T *__restrict__ v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned h = (unsigned)i * 2654435761u + 0x9e3779b9u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    v[i] = (T)((double)(h & 0xffffff) / 16777216.0 - 0.5);
}

// one block: scal[slot] = fixed-order sum of the slot's nb partials
// One block of kFB threads: each thread sums its strided partials of both
// slots in one pass (fixed order), then a fixed LDS tree.
constexpr int kFB = 1024;
__global__ __launch_bounds__(kFB) void k_cg_finalize(const double *__restrict__ P, CgState *st, int s0, int s1,
                                                     int nb)
{
    if (st->done) return;
    __shared__ double sa[kFB], sb[kFB];
    const int t = threadIdx.x;
    const double *p0 = P + s0 * nb, *p1 = P + (s1 >= 0 ? s1 : s0) * nb;
    double a = 0.0, b = 0.0;
    for (int q = t; q < nb; q += kFB) {
        a += p0[q];
        if (s1 >= 0) b += p1[q];
    }
    sa[t] = a;
    sb[t] = b;
    __syncthreads();
    for (int w = kFB / 2; w > 0; w >>= 1) {
        if (t < w) {
            sa[t] += sa[t + w];
            sb[t] += sb[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        st->scal[s0] = sa[0];
        if (s1 >= 0) st->scal[s1] = sb[0];
    }
}

// A = Me + hh Ms over the stored entries
template <typename T>
__global__ void k_stage_matrix(long nnz, const T *__restrict__ me, const T *__restrict__ ms, T hh, T *__restrict__ a)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nnz) a[q] = me[q] + hh * ms[q];
}

// dinv[r] = 1 / A[r][r] (1 where a row stores no diagonal)
template <typename T> __global__ void k_diag_inv(Csr<T> A, const T *__restrict__ val, T *__restrict__ dinv)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.n) return;
    T d = 1;
    for (int q = A.rp[r]; q < A.rp[r + 1]; ++q)
        if (A.col[q] == r) d = (T)1 / val[q];
    dinv[r] = d;
}

// y += a x
template <typename T> __global__ void k_axpy(int n, T a, const T *__restrict__ x, T *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + a * x[i];
}

// y += a x with Kahan compensation c (the low-order part y has lost so
// far): the stored state does not accumulate the rounding of its updates.
template <typename T> __global__ void k_axpy_comp(int n, T a, const T *__restrict__ x, T *y, T *c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T yv = a * x[i] - c[i];
    const T t = y[i] + yv;
    c[i] = (t - y[i]) - yv;
    y[i] = t;
}

// x = y + r (y - yp): linear extrapolation of the stage solutions in time
template <typename T>
__global__ void k_extrap(int n, const T *__restrict__ y, const T *__restrict__ yp, T r, T *__restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = y[i] + r * (y[i] - yp[i]);
}

// y = c0 + k (a1 x1 + a2 x2 + a3 x3), in that order
template <typename T>
__global__ void k_comb(int n, const T *__restrict__ c0, T k, T a1, const T *__restrict__ x1, T a2,
                       const T *__restrict__ x2, T a3, const T *__restrict__ x3, T *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T s = a1 * x1[i] + a2 * x2[i];
    if (x3) s = s + a3 * x3[i];
    y[i] = c0[i] + k * s;
}

// Misfit gradient per tetrahedron (one thread each, fixed order):
//   ge[T] += eps0 mu_T' M_T (Vin - Vout)_T,  gs[T] += -D/2 mu_T' M_T (Vin + Vout)_T
// over the interior dofs of T (boundary dofs are masked).
template <typename T, int NL>
__global__ __launch_bounds__(128) void k_tet_grad(long nt, const int *__restrict__ t2d, const double *__restrict__ mloc,
                           const unsigned char *__restrict__ bnd, const T *__restrict__ mu,
                           const T *__restrict__ vin, const T *__restrict__ vout, double eps0, double mhalf,
                           double *__restrict__ ge, double *__restrict__ gs)
{
    // NL (6 or 20) known at compile time: the local vectors stay in registers
    constexpr int nl = NL;
    const long tt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tt >= nt) return;
    const int *d = t2d + tt * nl;
    const double *M = mloc + (tt % 6) * nl * nl;
    double dm[nl], sm[nl];
#pragma unroll
    for (int l = 0; l < nl; ++l) {
        const int q = d[l];
        const bool in = !bnd[q];
        dm[l] = in ? (double)vin[q] - (double)vout[q] : 0.0;
        sm[l] = in ? (double)vin[q] + (double)vout[q] : 0.0;
    }
    double a = 0.0, b = 0.0;
    // rows rolled, columns unrolled: dm, sm are indexed only by the constant k
#pragma unroll 1
    for (int l = 0; l < nl; ++l) {
        const int q = d[l];
        const double ml = !bnd[q] ? (double)mu[q] : 0.0;
        double ra = 0.0, rb = 0.0;
#pragma unroll
        for (int k = 0; k < nl; ++k) {
            ra += M[l * nl + k] * dm[k];
            rb += M[l * nl + k] * sm[k];
        }
        a += ml * ra;
        b += ml * rb;
    }
    ge[tt] += eps0 * a;
    gs[tt] += mhalf * b;
}

// lU[dof] += sum_c w_c res_c over the receivers' basis functions (one thread,
// receiver order)
template <typename T>
__global__ void k_tet_inject(int nobs, int nl, const int *__restrict__ dof, const double *__restrict__ w,
                             const double *__restrict__ res, T *lu)
{
    if (threadIdx.x || blockIdx.x) return;
    for (int o = 0; o < nobs; ++o)
        for (int l = 0; l < nl; ++l) {
            const int q = dof[nl * o + l];
            const double *wl = w + (nl * o + l) * 3;
            const double v = (wl[0] * res[3 * o] + wl[1] * res[3 * o + 1]) + wl[2] * res[3 * o + 2];
            lu[q] = lu[q] + (T)v;
        }
}

// receivers: E(r) = sum_l U[dof_l] w_l(r) over the nl basis functions of the
// enclosing tetrahedron, 3 components, into rec[obs][3]
template <typename T>
__global__ void k_tet_record(const T *__restrict__ U, int nobs, int nl, const int *__restrict__ dof,
                             const double *__restrict__ w, double *rec)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nobs) return;
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int l = 0; l < nl; ++l) s += (double)U[dof[nl * o + l]] * w[(nl * o + l) * 3 + c];
        rec[3 * o + c] = s;
    }
}

}  // namespace

struct ems_tet {
    ems_tet_desc d{};
    Mesh m;
    size_t esz = 8;
    std::vector<double> eps_r, sigma;  // per tet (empty: 1 / 0)
    std::vector<int> rp, col;
    std::vector<char> bnd;
    DevBuf d_rp, d_col, d_S, d_Me, d_Ms;
    std::vector<double> deltas;  // stage steps with an assembled A_+
    DevBuf d_A[3], d_dinv[3];
    const void *a_val[3] = {}, *a_dinv[3] = {};  // solve matrix per stage step (d_A or d_Me)
    // element-by-element: per-tetrahedron mass coefficient eps0 eps_r + hh sigma
    // (type-major, [t * ncube + cube]) for hh = 0 and per stage step
    DevBuf d_cm_me, d_cm[3];
    const void *a_cm[3] = {};
    DevBuf d_dinv_me;  // Jacobi preconditioner of M_eps alone
    DevBuf U, V, b, r, z, p, ap;        // state and CG scratch
    DevBuf Vm, am, an, Up, Vp, Ap;      // ABM: history (V_n-1, a_n-1, a_n) and predictor
    // symplectic stages solve for y = (V' - V) / D, an acceleration at the
    // stage midpoint: Y, Yp keep the last two (at times ty, typ; have_prev
    // counts the valid ones) for the next solve's initial guess, X is the
    // solve's buffer
    DevBuf Y, Yp, X;
    double ty = 0.0, typ = 0.0;
    int have_prev = 0;
    // FP32 symplectic stages: Kahan compensation of the U and V updates,
    // valid for the current state (cleared when the state is set)
    DevBuf Uc, Vc;
    bool comp_valid = false, comp_off = false;
    DevBuf d_partial, d_cg;
    int lpr = 4;  // lanes per sparse row (4, or 8 for rows longer than 32)
    int op_kind = EMS_TET_OP_AUTO;
    bool ebe = false;   // element-by-element operator (no assembled matrices)
    bool mfma = false;  // element products of the mass operator on matrix cores
    bool rows = false;  // mass operator fused per row (EbeRowOp), no tmp buffer
    EbeMesh em{};
    DevBuf d_inc, d_Mt, d_St, d_Ct, d_er, d_sg, d_bnd, d_tmp, d_tmask;
    int nl = 6;   // basis functions per tetrahedron
    long nd = 0;  // dofs
    bool assembled = false, lossy = false;
    // dipoles: per dipole 6 dofs and the load weights w_l(r0).d
    std::vector<int> dip_dof;
    std::vector<double> dip_w, dip_par;  // par: f0, t0, amp per dipole
    DevBuf d_dip_dof, d_dip_w, d_dip_f;
    std::vector<int> obs_dof;
    std::vector<double> obs_w;
    DevBuf d_obs_dof, d_obs_w, d_rec;
    std::vector<double> traces;  // [nobs][nslot][3]
    int nslot = 0;
    long steps = 0;  // steps taken
    double t_now = 0.0;
    bool abm_started = false;
    int cg_last = 0, cg_max = 0;
    // mass solver: EMS_TET_SOLVER_AUTO (Chebyshev-accelerated Jacobi with
    // spectral bounds from Lanczos, CG where that fails) or _CG.  Bounds per
    // solve matrix: slot 3 = M_eps, 0..2 = the stage steps' matrices.
    int solver = EMS_TET_SOLVER_AUTO;
    double cheb_lo[4] = {}, cheb_hi[4] = {};
    int cheb_state[4] = {};  // 0 unknown, 1 bounds valid, -1 use CG
    int cheb_steps = 0;      // Lanczos steps of the last bounds
    DevBuf cd, cx, cx0;      // Chebyshev direction, ping-pong iterate, initial guess
    // after a run or gradient failed part way the device state is undefined:
    // bit 0 / bit 1 = E / dE/dt must be set again before the next run
    int need_state = 0;
    hipStream_t stream = nullptr;
    // misfit gradient (allocated on first use)
    DevBuf g_t2d, g_mloc, g_bnd, g_lU, g_lV, g_mu, g_y, g_vout, g_tmp, g_ge, g_gs, g_res, g_f;
    DevBuf g_ck, g_st, g_fwd_f;  // checkpointed gradient: (U, V) checkpoints, stage states, forward load factors
};

namespace {

double pulse_rate(double t, double f0, double t0, double amp)
{
    // p(t) = amp exp(-(pi f0 (t - t0))^2),  returns p'(t)
    const double a = M_PI * f0 * (t - t0);
    return amp * (-2.0 * M_PI * f0 * a) * std::exp(-a * a);
}

// ---------------------------------------------------------------------------
// Degrees of freedom.  Order 1: one per edge (the Whitney function), global
// index = edge.  Order 2: Whitney dofs [0, ne), edge-gradient dofs [ne, 2 ne)
// (ne + edge), face dofs [2 ne, 2 ne + 2 nf) (2 ne + 2 face + kappa).  Local
// dofs of a tetrahedron follow the basis order of tet_basis.hpp.
// ---------------------------------------------------------------------------
long dof_count(const Mesh &m, int order) { return order == 1 ? m.ne : 2 * m.ne + 2 * m.nf; }

void tet_dofs(const Mesh &m, int order, int i, int j, int k, int t, long *d)
{
    long e[6];
    m.tet_edges(i, j, k, t, e);
    for (int l = 0; l < 6; ++l) d[l] = e[l];
    if (order == 1) return;
    long fc[4];
    m.tet_faces(i, j, k, t, fc);
    for (int l = 0; l < 6; ++l) d[6 + l] = m.ne + e[l];
    for (int q = 0; q < 4; ++q) {
        d[12 + 2 * q] = 2 * m.ne + 2 * fc[q];
        d[13 + 2 * q] = 2 * m.ne + 2 * fc[q] + 1;
    }
}

bool dof_boundary(const Mesh &m, long d)
{
    if (d < 2 * m.ne) return m.boundary(d % m.ne);
    return m.face_boundary((d - 2 * m.ne) / 2);
}

// ---------------------------------------------------------------------------
// Assembly, one CSR row per dof, without a global sort.  A dof belongs to a
// row type (edge family and kind, or face family and kind); the tetrahedra
// containing its entity sit at fixed cube offsets from the entity's base
// vertex, so each row type has a fixed list of column slots (sorted in
// global index order) and every (tetrahedron, column) contribution lands in
// a precomputed slot.  A row sums its contributions in the fixed incidence
// order; slots that no in-range tetrahedron touches and boundary columns are
// dropped.  PEC: boundary rows are identity rows of the mass-type matrices
// and zero rows of S.
// ---------------------------------------------------------------------------
struct SlotKey {
    int cls, fam, rk, rj, ri, kap;  // cls 0 Whitney edge, 1 gradient edge, 2 face
    bool operator<(const SlotKey &o) const
    {
        if (cls != o.cls) return cls < o.cls;
        if (fam != o.fam) return fam < o.fam;
        if (rk != o.rk) return rk < o.rk;
        if (rj != o.rj) return rj < o.rj;
        if (ri != o.ri) return ri < o.ri;
        return kap < o.kap;
    }
    bool operator==(const SlotKey &o) const
    {
        return cls == o.cls && fam == o.fam && rk == o.rk && rj == o.rj && ri == o.ri && kap == o.kap;
    }
};

struct RowType {
    struct Inc {
        int d[3], t, L;  // cube offset from the row's base vertex, tet type, row's local dof
    };
    std::vector<Inc> inc;
    std::vector<SlotKey> slots;
    std::vector<int> slot_of;  // [inc][nl]
};

// the entity (class, family, base offset, kappa) of local dof m of tet type t
SlotKey local_key(int t, int m)
{
    int v[4][3];
    Mesh::verts(t, v);
    SlotKey k{};
    if (m < 12) {
        const int l = m % 6, a = kLocEdge[l][0], b = kLocEdge[l][1];
        k.cls = m / 6;
        k.fam = Mesh::family_of(v[b][0] - v[a][0], v[b][1] - v[a][1], v[b][2] - v[a][2]);
        k.ri = v[a][0];
        k.rj = v[a][1];
        k.rk = v[a][2];
        k.kap = 0;
    } else {
        const int q = (m - 12) / 2, a = kLocTri[q][0], b = kLocTri[q][1], c = kLocTri[q][2];
        const int f1 = Mesh::family_of(v[b][0] - v[a][0], v[b][1] - v[a][1], v[b][2] - v[a][2]);
        const int f2 = Mesh::family_of(v[c][0] - v[a][0], v[c][1] - v[a][1], v[c][2] - v[a][2]);
        k.cls = 2;
        k.fam = Mesh::face_family(f1, f2);
        k.ri = v[a][0];
        k.rj = v[a][1];
        k.rk = v[a][2];
        k.kap = (m - 12) % 2;
    }
    return k;
}

// row types: edges rt = f + 7 kappa (kappa < order), faces rt = 14 + 2 q + kappa
std::vector<RowType> row_types(int order)
{
    const int nl = basis_size(order);
    std::vector<RowType> rts(order == 1 ? 7 : 38);
    for (int t = 0; t < 6; ++t)
        for (int m = 0; m < nl; ++m) {
            const SlotKey k = local_key(t, m);
            const int rt = k.cls < 2 ? k.fam + 7 * k.cls : 14 + 2 * k.fam + k.kap;
            rts[rt].inc.push_back(RowType::Inc{{-k.ri, -k.rj, -k.rk}, t, m});
        }
    for (RowType &R : rts) {
        for (const RowType::Inc &ic : R.inc)
            for (int m = 0; m < nl; ++m) {
                SlotKey k = local_key(ic.t, m);
                k.ri += ic.d[0];
                k.rj += ic.d[1];
                k.rk += ic.d[2];
                R.slots.push_back(k);
            }
        std::sort(R.slots.begin(), R.slots.end());
        R.slots.erase(std::unique(R.slots.begin(), R.slots.end()), R.slots.end());
        for (const RowType::Inc &ic : R.inc)
            for (int m = 0; m < nl; ++m) {
                SlotKey k = local_key(ic.t, m);
                k.ri += ic.d[0];
                k.rj += ic.d[1];
                k.rk += ic.d[2];
                R.slot_of.push_back((int)(std::lower_bound(R.slots.begin(), R.slots.end(), k) - R.slots.begin()));
            }
    }
    return rts;
}

// row type and base vertex of a global dof
void dof_entity(const Mesh &m, long d, int &rt, int &I, int &J, int &K)
{
    if (d < 2 * m.ne) {
        int f;
        m.edge_coords(d % m.ne, f, I, J, K);
        rt = f + 7 * (int)(d / m.ne);
    } else {
        int q;
        m.face_coords((d - 2 * m.ne) / 2, q, I, J, K);
        rt = 14 + 2 * q + (int)((d - 2 * m.ne) % 2);
    }
}

long slot_column(const Mesh &m, const SlotKey &k, int I, int J, int K)
{
    if (k.cls < 2) return k.cls * m.ne + m.edge(k.fam, I + k.ri, J + k.rj, K + k.rk);
    return 2 * m.ne + 2 * m.face(k.fam, I + k.ri, J + k.rj, K + k.rk) + k.kap;
}

ems_status assemble(ems_tet *s, std::vector<double> &S, std::vector<double> &Me, std::vector<double> &Ms)
{
    const Mesh &m = s->m;
    const int order = s->d.order, nl = basis_size(order);
    const long nd = dof_count(m, order);
    s->bnd.assign(nd, 0);
    for (long d = 0; d < nd; ++d) s->bnd[d] = dof_boundary(m, d) ? 1 : 0;
    std::vector<double> Mt(6 * nl * nl), St(6 * nl * nl);
    for (int t = 0; t < 6; ++t) element_matrices(order, t, m.h, &Mt[t * nl * nl], &St[t * nl * nl]);
    const std::vector<RowType> rts = row_types(order);
    const double mu = s->d.mu, e0 = s->d.eps;
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto for_rows = [&](auto &&body) {
        std::vector<std::thread> th;
        for (int w = 0; w < nth; ++w)
            th.emplace_back([&, w]() {
                const long r0 = nd * w / nth, r1 = nd * (w + 1) / nth;
                std::vector<double> vs, vm, vg;
                std::vector<char> touched;
                for (long r = r0; r < r1; ++r) body(r, vs, vm, vg, touched);
            });
        for (std::thread &t : th) t.join();
    };
    // one row: slot values and touched flags; returns the row type
    auto row_values = [&](long r, bool values, std::vector<double> &vs, std::vector<double> &vm,
                          std::vector<double> &vg, std::vector<char> &touched, int &I, int &J, int &K) -> int {
        int rt;
        dof_entity(m, r, rt, I, J, K);
        const RowType &R = rts[rt];
        const size_t ns = R.slots.size();
        touched.assign(ns, 0);
        if (values) {
            vs.assign(ns, 0.0);
            vm.assign(ns, 0.0);
            vg.assign(ns, 0.0);
        }
        for (size_t ic = 0; ic < R.inc.size(); ++ic) {
            const RowType::Inc &c = R.inc[ic];
            const int ci = I + c.d[0], cj = J + c.d[1], ck = K + c.d[2];
            if (ci < 0 || cj < 0 || ck < 0 || ci >= m.nx || cj >= m.ny || ck >= m.nz) continue;
            const long tid = 6 * (ci + (long)m.nx * (cj + (long)m.ny * ck)) + c.t;
            const double er = s->eps_r.empty() ? 1.0 : s->eps_r[tid];
            const double sg = s->sigma.empty() ? 0.0 : s->sigma[tid];
            const double *Mrow = &Mt[(c.t * nl + c.L) * nl], *Srow = &St[(c.t * nl + c.L) * nl];
            const int *so = &R.slot_of[ic * nl];
            for (int q = 0; q < nl; ++q) {
                touched[so[q]] = 1;
                if (values) {
                    vs[so[q]] += Srow[q] / mu;
                    vm[so[q]] += (e0 * er) * Mrow[q];
                    vg[so[q]] += sg * Mrow[q];
                }
            }
        }
        return rt;
    };
    // pass 1: row lengths
    std::vector<int> len(nd);
    for_rows([&](long r, std::vector<double> &vs, std::vector<double> &vm, std::vector<double> &vg,
                 std::vector<char> &touched) {
        if (s->bnd[r]) {
            len[r] = 1;
            return;
        }
        int I, J, K;
        const RowType &R = rts[row_values(r, false, vs, vm, vg, touched, I, J, K)];
        int c = 0;
        for (size_t q = 0; q < R.slots.size(); ++q)
            if (touched[q] && !s->bnd[slot_column(m, R.slots[q], I, J, K)]) ++c;
        len[r] = c;
    });
    s->rp.assign(nd + 1, 0);
    long total = 0;
    for (long r = 0; r < nd; ++r) {
        s->rp[r] = (int)total;
        total += len[r];
        if (total > INT_MAX) return invalid("tetrahedral mesh too large: more than 2^31 matrix entries");
    }
    s->rp[nd] = (int)total;
    s->col.assign(total, 0);
    S.assign(total, 0.0);
    Me.assign(total, 0.0);
    Ms.assign(total, 0.0);
    // pass 2: columns and values
    for_rows([&](long r, std::vector<double> &vs, std::vector<double> &vm, std::vector<double> &vg,
                 std::vector<char> &touched) {
        long o = s->rp[r];
        if (s->bnd[r]) {
            s->col[o] = (int)r;
            S[o] = 0.0;
            Me[o] = 1.0;
            Ms[o] = 0.0;
            return;
        }
        int I, J, K;
        const RowType &R = rts[row_values(r, true, vs, vm, vg, touched, I, J, K)];
        for (size_t q = 0; q < R.slots.size(); ++q) {
            if (!touched[q]) continue;
            const long c = slot_column(m, R.slots[q], I, J, K);
            if (s->bnd[c]) continue;
            s->col[o] = (int)c;
            S[o] = vs[q];
            Me[o] = vm[q];
            Ms[o] = vg[q];
            ++o;
        }
    });
    return EMS_OK;
}

template <typename T> ems_status upload(DevBuf &d, const std::vector<double> &v)
{
    std::vector<T> t(v.begin(), v.end());
    EMS_TRY(d.reserve(sizeof(T) * std::max<size_t>(t.size(), 1)));
    if (!t.empty()) EMS_HIP(hipMemcpy(d.p, t.data(), sizeof(T) * t.size(), hipMemcpyHostToDevice));
    return EMS_OK;
}

std::vector<double> stage_deltas(const ems_tet *s)
{
    const double dt = s->d.dt;
    if (s->d.stepper == EMS_TET_SYMPLECTIC4) {
        const double x1 = 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0)), x2 = 1.0 - 2.0 * x1;
        return {x1 * dt, x2 * dt};
    }
    if (s->d.stepper == EMS_TET_ABM) return {0.0};
    return {dt};
}

// Per stage step D: the solve matrix A_+ = Me + D/2 Ms and its Jacobi
// preconditioner, built on the device from the stored Me and Ms; lossless
// and ABM solves use Me itself.
template <typename T> ems_status stage_matrices(ems_tet *s)
{
    const long nd = s->nd, nnz = (long)s->col.size();
    const Csr<T> A{(int)nd, s->d_rp.as<int>(), s->d_col.as<int>()};
    s->deltas = stage_deltas(s);
    for (int q = 0; q < 4; ++q) s->cheb_state[q] = 0;  // new matrices: new spectral bounds
    if (s->ebe) {
        // Jacobi diagonals of (eps0 eps_r + hh sigma) M per stage step (hh = 0: M_eps)
        for (size_t q = 0; s->lossy && q < s->deltas.size(); ++q)
            for (size_t t = 0; t < s->sigma.size(); ++t) {
                const double er = s->eps_r.empty() ? 1.0 : s->eps_r[t];
                if (!(s->d.eps * er + 0.5 * s->deltas[q] * s->sigma[t] > 0.0))
                    return invalid("conductivity too large for the symplectic_4 stages: eps + x2 dt sigma / 2 <= 0");
            }
        const int g = (int)((nd + 255) / 256);
        const double *er = s->eps_r.empty() ? nullptr : s->d_er.as<double>();
        const double *sg = s->sigma.empty() ? nullptr : s->d_sg.as<double>();
        EMS_TRY(s->d_dinv_me.reserve(sizeof(T) * nd));
        k_ebe_diag<T><<<g, 256, 0, s->stream>>>(s->em, s->d_inc.as<int>(), s->d_Mt.as<T>(), er, sg, s->d.eps, 0.0,
                                                s->d_bnd.as<unsigned char>(), s->d_dinv_me.as<T>());
        const long nt = 6 * s->em.ncube;
        const int gt = (int)((nt + 255) / 256);
        EMS_TRY(s->d_cm_me.reserve(sizeof(T) * nt));
        k_mass_coef<T><<<gt, 256, 0, s->stream>>>(s->em.ncube, er, sg, s->d.eps, 0.0, s->d_cm_me.as<T>());
        for (size_t q = 0; q < s->deltas.size(); ++q) {
            if (!s->lossy || s->deltas[q] == 0.0) {
                s->a_dinv[q] = s->d_dinv_me.p;
                s->a_cm[q] = s->d_cm_me.p;
                continue;
            }
            EMS_TRY(s->d_dinv[q].reserve(sizeof(T) * nd));
            k_ebe_diag<T><<<g, 256, 0, s->stream>>>(s->em, s->d_inc.as<int>(), s->d_Mt.as<T>(), er, sg, s->d.eps,
                                                    0.5 * s->deltas[q], s->d_bnd.as<unsigned char>(),
                                                    s->d_dinv[q].as<T>());
            s->a_dinv[q] = s->d_dinv[q].p;
            EMS_TRY(s->d_cm[q].reserve(sizeof(T) * nt));
            k_mass_coef<T><<<gt, 256, 0, s->stream>>>(s->em.ncube, er, sg, s->d.eps, 0.5 * s->deltas[q],
                                                      s->d_cm[q].as<T>());
            s->a_cm[q] = s->d_cm[q].p;
        }
        EMS_HIP(hipGetLastError());
        EMS_HIP(hipStreamSynchronize(s->stream));
        return EMS_OK;
    }
    // Me + D/2 Ms is positive definite when every tet has eps0 eps_r + D/2 sigma > 0
    // (the symplectic_4 middle stage has D < 0)
    for (size_t q = 0; s->lossy && q < s->deltas.size(); ++q)
        for (size_t t = 0; t < s->sigma.size(); ++t) {
            const double er = s->eps_r.empty() ? 1.0 : s->eps_r[t];
            if (!(s->d.eps * er + 0.5 * s->deltas[q] * s->sigma[t] > 0.0))
                return invalid("conductivity too large for the symplectic_4 stages: eps + x2 dt sigma / 2 <= 0");
        }
    EMS_TRY(s->d_dinv_me.reserve(sizeof(T) * nd));
    k_diag_inv<T><<<(int)((nd + 255) / 256), 256, 0, s->stream>>>(A, s->d_Me.as<T>(), s->d_dinv_me.as<T>());
    for (size_t q = 0; q < s->deltas.size(); ++q) {
        if (!s->lossy || s->deltas[q] == 0.0) {
            s->a_val[q] = s->d_Me.p;
            s->a_dinv[q] = s->d_dinv_me.p;
            continue;
        }
        EMS_TRY(s->d_A[q].reserve(sizeof(T) * std::max<long>(nnz, 1)));
        EMS_TRY(s->d_dinv[q].reserve(sizeof(T) * nd));
        k_stage_matrix<T><<<(int)((nnz + 255) / 256), 256, 0, s->stream>>>(
            nnz, s->d_Me.as<T>(), s->d_Ms.as<T>(), (T)(0.5 * s->deltas[q]), s->d_A[q].as<T>());
        k_diag_inv<T><<<(int)((nd + 255) / 256), 256, 0, s->stream>>>(A, s->d_A[q].as<T>(), s->d_dinv[q].as<T>());
        s->a_val[q] = s->d_A[q].p;
        s->a_dinv[q] = s->d_dinv[q].p;
    }
    EMS_HIP(hipGetLastError());
    EMS_HIP(hipStreamSynchronize(s->stream));
    return EMS_OK;
}

// Element-by-element setup: boundary flags, mesh tables, incidence lists,
// element matrices (S divided by mu), per-tetrahedron material.
template <typename T> ems_status setup_ebe(ems_tet *s)
{
    const Mesh &m = s->m;
    const int order = s->d.order, nl = s->nl;
    const long nd = s->nd;
    s->bnd.assign(nd, 0);
    for (long d = 0; d < nd; ++d) s->bnd[d] = dof_boundary(m, d) ? 1 : 0;
    EbeMesh &e = s->em;
    e = EbeMesh{};
    e.nx = m.nx;
    e.ny = m.ny;
    e.nz = m.nz;
    e.nl = nl;
    e.ne = m.ne;
    e.nf = m.nf;
    e.nd = nd;
    e.ncube = (long)m.nx * m.ny * m.nz;
    for (int f = 0; f < 8; ++f) e.off[f] = m.off[f];
    for (int q = 0; q < 13; ++q) e.foff[q] = m.foff[q];
    for (int f = 0; f < 7; ++f) {
        e.ecx[f] = m.nx + 1 - kFam[f][0];
        e.ecy[f] = m.ny + 1 - kFam[f][1];
    }
    for (int q = 0; q < 12; ++q) {
        e.fcx[q] = m.nx + 1 - kFam[kFFam[q][1]][0];
        e.fcy[q] = m.ny + 1 - kFam[kFFam[q][1]][1];
    }
    for (int t = 0; t < 6; ++t) {
        for (int l = 0; l < 6; ++l) {
            const SlotKey k = local_key(t, l);
            e.tedge[t][l] = k.fam | (k.ri << 4) | (k.rj << 5) | (k.rk << 6);
        }
        for (int q = 0; q < 4; ++q) {
            const SlotKey k = local_key(t, 12 + 2 * q);
            e.tface[t][q] = k.fam | (k.ri << 4) | (k.rj << 5) | (k.rk << 6);
        }
    }
    const std::vector<RowType> rts = row_types(order);
    std::vector<int> inc;
    for (size_t rt = 0; rt < 39; ++rt) {
        e.inc_start[rt] = (int)inc.size();
        if (rt >= rts.size()) continue;
        for (const RowType::Inc &c : rts[rt].inc)
            inc.push_back((c.d[0] + 1) | ((c.d[1] + 1) << 2) | ((c.d[2] + 1) << 4) | (c.t << 6) | (c.L << 9));
    }
    std::vector<double> Mt(6 * nl * nl), St(6 * nl * nl);
    for (int t = 0; t < 6; ++t) element_matrices(order, t, m.h, &Mt[t * nl * nl], &St[t * nl * nl]);
    for (double &v : St) v /= s->d.mu;
    {  // first order: element curls C and (V/mu) C per type (S = (V/mu) C^T C)
        std::vector<double> ct(6 * 36, 0.0);
        if (order == 1)
            for (int t = 0; t < 6; ++t) {
                double C[3][6], V;
                whitney_curls(t, m.h, C, V);
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 6; ++l) {
                        ct[t * 36 + k * 6 + l] = C[k][l];
                        ct[t * 36 + 18 + k * 6 + l] = C[k][l] * (V / s->d.mu);
                    }
            }
        EMS_TRY(upload<T>(s->d_Ct, ct));
    }
    EMS_TRY(s->d_inc.reserve(sizeof(int) * inc.size()));
    EMS_HIP(hipMemcpy(s->d_inc.p, inc.data(), sizeof(int) * inc.size(), hipMemcpyHostToDevice));
    EMS_TRY(upload<T>(s->d_Mt, Mt));
    EMS_TRY(upload<T>(s->d_St, St));
    std::vector<unsigned char> b(s->bnd.begin(), s->bnd.end());
    EMS_TRY(s->d_bnd.reserve(b.size()));
    EMS_HIP(hipMemcpy(s->d_bnd.p, b.data(), b.size(), hipMemcpyHostToDevice));
    // per tetrahedron: bit l set when local dof l is a boundary dof
    {
        // boundary mask per tetrahedron, type-major ([t * ncube + cube], as tmp)
        std::vector<unsigned> tm((size_t)m.nt, 0u);
        const size_t ncube = (size_t)m.nx * m.ny * m.nz;
        long dd[20];
        size_t cube = 0;
        for (int k = 0; k < m.nz; ++k)
            for (int j = 0; j < m.ny; ++j)
                for (int i = 0; i < m.nx; ++i, ++cube)
                    for (int t = 0; t < 6; ++t) {
                        tet_dofs(m, order, i, j, k, t, dd);
                        unsigned w = 0;
                        for (int l = 0; l < nl; ++l) w |= (s->bnd[dd[l]] ? 1u : 0u) << l;
                        tm[(size_t)t * ncube + cube] = w;
                    }
        EMS_TRY(s->d_tmask.reserve(sizeof(unsigned) * tm.size()));
        EMS_HIP(hipMemcpy(s->d_tmask.p, tm.data(), sizeof(unsigned) * tm.size(), hipMemcpyHostToDevice));
    }
    for (size_t rt = 0; rt + 1 < 39; ++rt)
        if (e.inc_start[rt + 1] - e.inc_start[rt] > kMaxInc) return invalid("internal: incidence list too long");
    if (!s->eps_r.empty()) {
        EMS_TRY(s->d_er.reserve(sizeof(double) * s->eps_r.size()));
        EMS_HIP(hipMemcpy(s->d_er.p, s->eps_r.data(), sizeof(double) * s->eps_r.size(), hipMemcpyHostToDevice));
    }
    s->lossy = false;
    for (double v : s->sigma)
        if (v != 0.0) s->lossy = true;
    if (!s->sigma.empty()) {
        EMS_TRY(s->d_sg.reserve(sizeof(double) * s->sigma.size()));
        EMS_HIP(hipMemcpy(s->d_sg.p, s->sigma.data(), sizeof(double) * s->sigma.size(), hipMemcpyHostToDevice));
    }
    EMS_TRY(s->d_tmp.reserve(sizeof(T) * 6 * nl * (size_t)e.ncube));
    s->col.clear();
    return EMS_OK;
}

// State and solver vectors.  U and V hold the state: they are zeroed when
// first allocated only, so that a field set at rest survives the reassembly
// that a later set_material / set_operator triggers.
template <typename T> ems_status alloc_vectors(ems_tet *s)
{
    const size_t vb = sizeof(T) * s->nd;
    for (DevBuf *v : {&s->U, &s->V}) {
        const bool fresh = v->p == nullptr || v->bytes < vb;
        EMS_TRY(v->reserve(vb));
        if (fresh) EMS_HIP(hipMemset(v->p, 0, vb));
    }
    for (DevBuf *v : {&s->b, &s->r, &s->z, &s->p, &s->ap, &s->Vm, &s->am, &s->an, &s->Up, &s->Vp, &s->Ap}) {
        EMS_TRY(v->reserve(vb));
        EMS_HIP(hipMemset(v->p, 0, vb));
    }
    EMS_TRY(s->d_partial.reserve(sizeof(double) * 4 * kCgMaxNB));
    EMS_TRY(s->d_cg.reserve(sizeof(CgState)));
    return EMS_OK;
}

template <typename T> ems_status setup(ems_tet *s)
{
    if (s->assembled) return EMS_OK;
    s->ebe = s->op_kind != EMS_TET_OP_CSR;
    s->mfma = s->op_kind == EMS_TET_OP_EBE_MFMA;
    s->rows = s->op_kind == EMS_TET_OP_EBE_ROWS;
    const long nd = s->nd;
    if (s->ebe) {
        EMS_TRY(setup_ebe<T>(s));
        EMS_TRY(stage_matrices<T>(s));
        EMS_TRY(alloc_vectors<T>(s));
        s->assembled = true;
        return EMS_OK;
    }
    std::vector<double> S, Me, Ms;
    EMS_TRY(assemble(s, S, Me, Ms));
    s->lossy = false;
    for (double v : Ms)
        if (v != 0.0) s->lossy = true;
    EMS_TRY(s->d_rp.reserve(sizeof(int) * s->rp.size()));
    EMS_TRY(s->d_col.reserve(sizeof(int) * std::max<size_t>(s->col.size(), 1)));
    EMS_HIP(hipMemcpy(s->d_rp.p, s->rp.data(), sizeof(int) * s->rp.size(), hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(s->d_col.p, s->col.data(), sizeof(int) * s->col.size(), hipMemcpyHostToDevice));
    EMS_TRY(upload<T>(s->d_S, S));
    S = std::vector<double>();
    EMS_TRY(upload<T>(s->d_Me, Me));
    Me = std::vector<double>();
    if (s->lossy) EMS_TRY(upload<T>(s->d_Ms, Ms));
    Ms = std::vector<double>();
    EMS_TRY(stage_matrices<T>(s));
    EMS_TRY(alloc_vectors<T>(s));
    int longest = 0;
    for (long r = 0; r < nd; ++r) longest = std::max(longest, s->rp[r + 1] - s->rp[r]);
    s->lpr = longest > 32 ? 8 : 4;
    s->assembled = true;
    return EMS_OK;
}

void swap_dev(DevBuf &a, DevBuf &b)
{
    std::swap(a.p, b.p);
    std::swap(a.bytes, b.bytes);
}

// calls f(std::integral_constant<int, LPR>) for the handle's lanes per row
template <typename F> void with_lpr(int lpr, F &&f)
{
    if (lpr == 8) f(std::integral_constant<int, 8>{});
    else f(std::integral_constant<int, 4>{});
}

// blocks of 256 threads covering n rows at LPR lanes each
inline int row_blocks(long n, int lpr) { return (int)((n * lpr + 255) / 256); }

template <typename F> void with_nl(int nl, F &&f)
{
    if (nl == 20) f(std::integral_constant<int, 20>{});
    else f(std::integral_constant<int, 6>{});
}

// phase 1 of the element-by-element operator
template <typename T, int MODE>
void ebe_local(ems_tet *s, double hh, const T *x1, const T *x2, const int *done, const T *cmt = nullptr)
{
    const dim3 grid((unsigned)((s->em.ncube + 255) / 256), 6);
    const double *er = s->eps_r.empty() ? nullptr : s->d_er.as<double>();
    const double *sg = s->sigma.empty() ? nullptr : s->d_sg.as<double>();
    if (MODE == EBE_MASS && s->mfma) {
        const dim3 gm((unsigned)((s->em.ncube + 63) / 64), 6);
        with_nl(s->nl, [&](auto N) {
            constexpr int NL = decltype(N)::value;
            k_ebe_mass_mfma<T, NL><<<gm, 256, 0, s->stream>>>(s->em, s->d_Mt.as<T>(), er, sg, s->d.eps, hh,
                                                             s->d_tmask.as<unsigned>(), x1, s->d_tmp.as<T>(),
                                                             done);
        });
        return;
    }
    with_nl(s->nl, [&](auto N) {
        constexpr int NL = decltype(N)::value;
        k_ebe_local<T, NL, MODE><<<grid, 256, 0, s->stream>>>(s->em, s->d_Mt.as<T>(), s->d_St.as<T>(), er, sg,
                                                             s->d_Ct.as<T>(), s->d.eps, hh, s->d_tmask.as<unsigned>(),
                                                             x1, x2,
                                                             s->d_tmp.as<T>(), done, cmt);
    });
}

template <typename T> EbeOp<T> ebe_op(ems_tet *s, int ident)
{
    return EbeOp<T>{s->em, s->d_inc.as<int>(), s->d_tmp.as<T>(), s->d_bnd.as<unsigned char>(), ident};
}

// Element-by-element mass operator of solve matrix `which` (M_eps for
// which < 0) applied to x: f(op) launches the kernel that consumes op.  The
// two-phase form runs phase 1 on x first; the per-row form needs nothing.
template <typename T, typename F> void with_mass_op(ems_tet *s, int which, const T *x, const int *done, F &&f)
{
    const double hh = (which < 0 || !s->lossy) ? 0.0 : 0.5 * s->deltas[which];
    const T *cmt = (const T *)(which < 0 ? s->d_cm_me.p : s->a_cm[which]);
    if (s->rows) {
        with_nl(s->nl, [&](auto N) {
            constexpr int NL = decltype(N)::value;
            f(EbeRowOp<T, NL>{s->em, s->d_inc.as<int>(), s->d_Mt.as<T>(), s->d_tmask.as<unsigned>(), cmt,
                              s->d_bnd.as<unsigned char>()});
        });
        return;
    }
    ebe_local<T, EBE_MASS>(s, hh, x, nullptr, done, cmt);
    f(ebe_op<T>(s, 1));
}

// y = S x (which 0) or y = M_eps x (which 1)
template <typename T> ems_status apply_op(ems_tet *s, int which, const T *x, T *y)
{
    const int n = (int)s->nd;
    if (s->ebe) {
        if (which == 0) ebe_local<T, EBE_STIFF>(s, 0.0, x, nullptr, nullptr);
        else ebe_local<T, EBE_MASS>(s, 0.0, x, nullptr, nullptr);
        k_apply<EbeOp<T>, T><<<(n + 255) / 256, 256, 0, s->stream>>>(ebe_op<T>(s, which), x, y);
    } else {
        const Csr<T> A{n, s->d_rp.as<int>(), s->d_col.as<int>()};
        const T *val = which == 0 ? s->d_S.as<T>() : s->d_Me.as<T>();
        with_lpr(s->lpr, [&](auto L) {
            constexpr int LPR = decltype(L)::value;
            k_spmv<LPR, T><<<row_blocks(n, LPR), 256, 0, s->stream>>>(A, val, x, y);
        });
    }
    EMS_HIP(hipGetLastError());
    return EMS_OK;
}

// x <- A^-1 b (initial guess x) with A = M_eps (which < 0) or the solve
// matrix M_eps + D/2 M_sig of stage step `which`
template <typename T> ems_status cg(ems_tet *s, int which, DevBuf &x)
{
    const int n = (int)s->nd;
    // FP32: the recursively updated residual keeps converging below the
    // rounding of the matrix entries, so 1e-7 is reachable
    const double tol = s->esz == 4 ? 1e-7 : 1e-14;
    const double tol2 = tol * tol;
    CgState *st = s->d_cg.as<CgState>();
    double *P = s->d_partial.as<double>();
    T *r = s->r.as<T>(), *z = s->z.as<T>(), *p = s->p.as<T>(), *ap = s->ap.as<T>(), *xx = x.as<T>();
    const T *b = s->b.as<T>();
    const T *dinv = (const T *)(which < 0 ? s->d_dinv_me.p : s->a_dinv[which]);
    const int maxit = 512, check = 16, nb = cg_blocks(n);
    const bool fin = nb > kCgInlineNB;
    if (s->ebe) {
        with_mass_op<T>(s, which, xx, nullptr, [&](auto op) {
            k_cg_init<decltype(op), T><<<nb, kRB, 0, s->stream>>>(op, dinv, b, xx, r, z, P, st, nb);
        });
        if (fin) k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, 0, 3, nb);
        with_mass_op<T>(s, which, z, nullptr, [&](auto op) {
            k_cg_first<decltype(op), T><<<nb, kRB, 0, s->stream>>>(op, z, p, ap, xx, P, st, tol2, nb);
        });
        if (fin) k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, 2, -1, nb);
        EMS_HIP(hipGetLastError());
        for (int it = 0; it < maxit; ++it) {
            const int par = it & 1;
            k_cg_update<T><<<nb, kRB, 0, s->stream>>>(n, xx, r, z, p, ap, dinv, P, par, st, nb);
            if (fin) k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, par ^ 1, -1, nb);
            with_mass_op<T>(s, which, z, &st->done, [&](auto op) {
                k_cg_direction<decltype(op), T><<<nb, kRB, 0, s->stream>>>(op, z, p, ap, P, par, st, it, tol2, nb);
            });
            if (fin) k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, 2, -1, nb);
            if ((it + 1) % check == 0) {
                CgState h;
                EMS_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s->stream));
                EMS_HIP(hipStreamSynchronize(s->stream));
                if (h.done) {
                    s->cg_last = h.iters;
                    s->cg_max = std::max(s->cg_max, h.iters);
                    return EMS_OK;
                }
            }
        }
        set_error("tetrahedral mass solve did not converge");
        return EMS_ERR_STATE;
    }
    const Csr<T> A{n, s->d_rp.as<int>(), s->d_col.as<int>()};
    const T *val = (const T *)(which < 0 ? s->d_Me.p : s->a_val[which]);
    with_lpr(s->lpr, [&](auto L) {
        constexpr int LPR = decltype(L)::value;
        const CsrOp<LPR, T> op{A, val};
        k_cg_init<CsrOp<LPR, T>, T><<<nb, kRB, 0, s->stream>>>(op, dinv, b, xx, r, z, P, st, nb);
        if (fin) k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, 0, 3, nb);
        k_cg_first<CsrOp<LPR, T>, T><<<nb, kRB, 0, s->stream>>>(op, z, p, ap, xx, P, st, tol2, nb);
        if (fin) k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, 2, -1, nb);
    });
    EMS_HIP(hipGetLastError());
    for (int it = 0; it < maxit; ++it) {
        const int par = it & 1;
        k_cg_update<T><<<nb, kRB, 0, s->stream>>>(n, xx, r, z, p, ap, dinv, P, par, st, nb);
        if (fin) k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, par ^ 1, -1, nb);
        with_lpr(s->lpr, [&](auto L) {
            constexpr int LPR = decltype(L)::value;
            const CsrOp<LPR, T> op{A, val};
            k_cg_direction<CsrOp<LPR, T>, T><<<nb, kRB, 0, s->stream>>>(op, z, p, ap, P, par, st, it, tol2, nb);
        });
        if (fin) k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, 2, -1, nb);
        if ((it + 1) % check == 0) {
            CgState h;
            EMS_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s->stream));
            EMS_HIP(hipStreamSynchronize(s->stream));
            if (h.done) {
                s->cg_last = h.iters;
                s->cg_max = std::max(s->cg_max, h.iters);
                return EMS_OK;
            }
        }
    }
    set_error("tetrahedral mass solve did not converge");
    return EMS_ERR_STATE;
}

// y = A x with A = M_eps (which < 0) or the solve matrix of stage step `which`
template <typename T> void apply_solve(ems_tet *s, int which, const T *x, T *y)
{
    const int n = (int)s->nd;
    if (s->ebe) {
        with_mass_op<T>(s, which, x, nullptr, [&](auto op) {
            k_apply<decltype(op), T><<<(n + 255) / 256, 256, 0, s->stream>>>(op, x, y);
        });
    } else {
        const Csr<T> A{n, s->d_rp.as<int>(), s->d_col.as<int>()};
        const T *val = (const T *)(which < 0 ? s->d_Me.p : s->a_val[which]);
        with_lpr(s->lpr, [&](auto L) {
            constexpr int LPR = decltype(L)::value;
            k_spmv<LPR, T><<<row_blocks(n, LPR), 256, 0, s->stream>>>(A, val, x, y);
        });
    }
}

// sum a_i b_i (w 0) or sum a_i b_i / dinv_i (w 1), block partials summed on
// the host in a fixed order
template <typename T> ems_status dot_host(ems_tet *s, const T *a, const T *b, const T *dinv, int w, double &out)
{
    const int n = (int)s->nd, nb = cg_blocks(n);
    double *P = s->d_partial.as<double>();
    k_dot<T><<<nb, kRB, 0, s->stream>>>(n, a, b, dinv, w, P);
    EMS_HIP(hipGetLastError());
    std::vector<double> h((size_t)nb);
    EMS_HIP(hipMemcpyAsync(h.data(), P, sizeof(double) * nb, hipMemcpyDeviceToHost, s->stream));
    EMS_HIP(hipStreamSynchronize(s->stream));
    double t = 0.0;
    for (double v : h) t += v;
    out = t;
    return EMS_OK;
}

// Eigenvalues of the symmetric tridiagonal matrix (diagonal a, off-diagonal
// b[i] between rows i-1 and i): number below x by the Sturm sequence, and the
// k-th smallest by bisection inside the Gershgorin interval.
int sturm_count(const std::vector<double> &a, const std::vector<double> &b, double x)
{
    int c = 0;
    double q = 1.0;
    for (size_t i = 0; i < a.size(); ++i) {
        q = a[i] - x - (i > 0 ? b[i] * b[i] / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0.0) ++c;
    }
    return c;
}

double tridiag_eig(const std::vector<double> &a, const std::vector<double> &b, int k)
{
    double lo = 1e300, hi = -1e300;
    for (size_t i = 0; i < a.size(); ++i) {
        const double r = (i > 0 ? std::fabs(b[i]) : 0.0) + (i + 1 < a.size() ? std::fabs(b[i + 1]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (sturm_count(a, b, mid) > k) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// Spectral bounds of D^-1 A by Lanczos in the D inner product from a fixed
// pseudo-random start, until both extreme Ritz values change by under 1e-4
// over 8 steps (at most 240): the extreme Ritz values widened by 5 % (Ritz
// values lie inside the spectrum).  Uses the CG scratch vectors.
template <typename T> ems_status cheb_bounds(ems_tet *s, int which, double &lo, double &hi)
{
    const int n = (int)s->nd, g = (n + 255) / 256;
    const T *dinv = (const T *)(which < 0 ? s->d_dinv_me.p : s->a_dinv[which]);
    T *bufA = s->p.as<T>(), *bufB = s->ap.as<T>(), *w = s->r.as<T>(), *y = s->z.as<T>();
    k_hash_fill<T><<<g, 256, 0, s->stream>>>(n, w);
    double nn = 0.0;
    EMS_TRY(dot_host(s, w, w, dinv, 1, nn));
    if (!(nn > 0.0)) return not_supported("Lanczos start vector vanished");
    T *vp = bufA, *v = bufB;
    k_scale_to<T><<<g, 256, 0, s->stream>>>(n, w, std::sqrt(nn), v);
    k_scale_to<T><<<g, 256, 0, s->stream>>>(n, w, HUGE_VAL, vp);  // 0
    std::vector<double> al, be;
    double beta = 0.0, last_lo = 0.0, last_hi = 0.0;
    for (int j = 0; j < 240; ++j) {
        apply_solve<T>(s, which, v, y);
        double alpha = 0.0;
        EMS_TRY(dot_host(s, y, v, dinv, 0, alpha));
        k_lanczos_w<T><<<g, 256, 0, s->stream>>>(n, dinv, y, v, vp, alpha, beta, w);
        double ww = 0.0;
        EMS_TRY(dot_host(s, w, w, dinv, 1, ww));
        al.push_back(alpha);
        be.push_back(beta);
        beta = std::sqrt(std::max(ww, 0.0));
        if (!(beta > 1e-10 * std::fabs(alpha))) break;
        k_scale_to<T><<<g, 256, 0, s->stream>>>(n, w, beta, vp);  // the next v, in the old vp
        std::swap(v, vp);
        // every 8 steps: stop once both extreme Ritz values are settled
        if ((j + 1) % 8 == 0) {
            const double a = tridiag_eig(al, be, 0), b = tridiag_eig(al, be, (int)al.size() - 1);
            const bool settled = std::fabs(a - last_lo) <= 1e-4 * a && std::fabs(b - last_hi) <= 1e-4 * b;
            last_lo = a;
            last_hi = b;
            if (settled) break;
        }
    }
    EMS_HIP(hipGetLastError());
    const int m = (int)al.size();
    lo = 0.95 * tridiag_eig(al, be, 0);
    hi = 1.05 * tridiag_eig(al, be, m - 1);
    s->cheb_steps = m;
    return EMS_OK;
}

// x <- A^-1 b (initial guess x) by Chebyshev-accelerated Jacobi: one residual
// (with its inner products) before and one after a fixed number of
// dot-free iterations; if the final residual misses the tolerance (bounds too
// narrow), the matrix falls back to CG from the saved initial guess.
template <typename T> ems_status solve(ems_tet *s, int which, DevBuf &x)
{
    // FP32: the true residual b - A x of a fixed-coefficient iteration
    // stagnates near the FP32 rounding of A x, above the 1e-7 tolerance that
    // CG's recursively updated residual reaches; CG there
    // Order 2 (Jacobi condition number ~130): about 10 % more iterations
    // than CG and slower per solve (measured); CG there too.
    if (s->solver == EMS_TET_SOLVER_CG || s->esz == 4 || s->nl != 6) return cg<T>(s, which, x);
    const int slot = which < 0 ? 3 : which;
    if (s->cheb_state[slot] == 0) {
        double lo = 0.0, hi = 0.0;
        EMS_TRY(cheb_bounds<T>(s, which, lo, hi));
        if (lo > 0.0 && hi > lo && std::isfinite(hi)) {
            s->cheb_lo[slot] = lo;
            s->cheb_hi[slot] = hi;
            s->cheb_state[slot] = 1;
        } else {
            s->cheb_state[slot] = -1;
        }
    }
    if (s->cheb_state[slot] < 0) return cg<T>(s, which, x);
    const int n = (int)s->nd, nb = cg_blocks(n);
    const double tol = s->esz == 4 ? 1e-7 : 1e-14;
    CgState *st = s->d_cg.as<CgState>();
    double *P = s->d_partial.as<double>();
    T *r = s->r.as<T>(), *z = s->z.as<T>();
    const T *b = s->b.as<T>();
    const T *dinv = (const T *)(which < 0 ? s->d_dinv_me.p : s->a_dinv[which]);
    const size_t vb = sizeof(T) * n;
    EMS_TRY(s->cd.reserve(vb));
    EMS_TRY(s->cx.reserve(vb));
    EMS_TRY(s->cx0.reserve(vb));
    EMS_HIP(hipMemcpyAsync(s->cx0.p, x.p, vb, hipMemcpyDeviceToDevice, s->stream));
    // r = b - A xx, z = D^-1 r; (r, z) and (b, D^-1 b) to the host
    auto residual = [&](const T *xx, double &rz, double &bb) -> ems_status {
        if (s->ebe) {
            with_mass_op<T>(s, which, xx, nullptr, [&](auto op) {
                k_cg_init<decltype(op), T><<<nb, kRB, 0, s->stream>>>(op, dinv, b, xx, r, z, P, st, nb);
            });
        } else {
            const Csr<T> A{n, s->d_rp.as<int>(), s->d_col.as<int>()};
            const T *val = (const T *)(which < 0 ? s->d_Me.p : s->a_val[which]);
            with_lpr(s->lpr, [&](auto L) {
                constexpr int LPR = decltype(L)::value;
                k_cg_init<CsrOp<LPR, T>, T><<<nb, kRB, 0, s->stream>>>(CsrOp<LPR, T>{A, val}, dinv, b, xx, r, z, P, st,
                                                                      nb);
            });
        }
        k_cg_finalize<<<1, kFB, 0, s->stream>>>(P, st, 0, 3, nb);
        EMS_HIP(hipGetLastError());
        CgState h;
        EMS_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s->stream));
        EMS_HIP(hipStreamSynchronize(s->stream));
        rz = h.scal[0];
        bb = h.scal[3];
        return EMS_OK;
    };
    double rz = 0.0, bb = 0.0;
    EMS_TRY(residual(x.as<T>(), rz, bb));
    if (!(bb > 0.0)) {  // b = 0: x = 0
        k_scale_to<T><<<(n + 255) / 256, 256, 0, s->stream>>>(n, b, HUGE_VAL, x.as<T>());
        s->cg_last = 0;
        return EMS_OK;
    }
    if (!(rz > tol * tol * bb)) {
        s->cg_last = 0;
        return EMS_OK;
    }
    const double lo = s->cheb_lo[slot], hi = s->cheb_hi[slot];
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    const double sk = std::sqrt(hi / lo), q = (sk - 1.0) / (sk + 1.0);
    const double need = 0.5 * tol * std::sqrt(bb / rz);
    const int K = std::max(1, std::min(4000, (int)std::ceil(std::log(need) / std::log(q)) + 2));
    T *cur = x.as<T>(), *nxt = s->cx.as<T>(), *d = s->cd.as<T>();
    double rho_prev = 1.0 / sigma;
    // no reductions in the steps: a full grid (latency is hidden by waves,
    // not by the grid-stride loop of the CG kernels)
    const int gs = (int)std::min<long>(((long)n + kRB - 1) / kRB, 1L << 20);
    for (int k = 0; k < K; ++k) {
        double c1 = 0.0, c2 = 1.0 / theta;
        if (k > 0) {
            const double rho = 1.0 / (2.0 * sigma - rho_prev);
            c1 = rho * rho_prev;
            c2 = 2.0 * rho / delta;
            rho_prev = rho;
        }
        if (s->ebe) {
            with_mass_op<T>(s, which, cur, nullptr, [&](auto op) {
                k_cheb_step<decltype(op), T><<<gs, kRB, 0, s->stream>>>(op, b, dinv, cur, d, nxt, (T)c1, (T)c2, k == 0);
            });
        } else {
            const Csr<T> A{n, s->d_rp.as<int>(), s->d_col.as<int>()};
            const T *val = (const T *)(which < 0 ? s->d_Me.p : s->a_val[which]);
            with_lpr(s->lpr, [&](auto L) {
                constexpr int LPR = decltype(L)::value;
                const int gl = (int)std::min<long>(((long)n * LPR + kRB - 1) / kRB, 1L << 20);
                k_cheb_step<CsrOp<LPR, T>, T><<<gl, kRB, 0, s->stream>>>(CsrOp<LPR, T>{A, val}, b, dinv, cur, d, nxt,
                                                                         (T)c1, (T)c2, k == 0);
            });
        }
        std::swap(cur, nxt);
    }
    EMS_HIP(hipGetLastError());
    if (cur != x.as<T>()) EMS_HIP(hipMemcpyAsync(x.p, cur, vb, hipMemcpyDeviceToDevice, s->stream));
    double rz2 = 0.0;
    EMS_TRY(residual(x.as<T>(), rz2, bb));
    if (!(rz2 > tol * tol * bb)) {
        s->cg_last = K;
        s->cg_max = std::max(s->cg_max, K);
        return EMS_OK;
    }
    // bounds too narrow for this matrix: CG from the initial guess from now on
    s->cheb_state[slot] = -1;
    EMS_HIP(hipMemcpyAsync(x.p, s->cx0.p, vb, hipMemcpyDeviceToDevice, s->stream));
    return cg<T>(s, which, x);
}

template <typename T> void load(ems_tet *s, DevBuf &b, const double *f)
{
    const int nw = (int)s->dip_dof.size();
    if (nw) k_load<T><<<1, 64, 0, s->stream>>>(b.as<T>(), nw, s->nl, s->d_dip_dof.as<int>(), s->d_dip_w.as<double>(), f);
}

// load scale per dipole weight at time t: -p'(t) (the force of stages and ABM)
void dipole_factors(const ems_tet *s, double t, double scale, std::vector<double> &out)
{
    const int nd = (int)s->dip_par.size() / 3;
    for (int q = 0; q < nd; ++q) out.push_back(scale * pulse_rate(t, s->dip_par[3 * q], s->dip_par[3 * q + 1], s->dip_par[3 * q + 2]));
}

// b = -S U - Ms V + load (the force of the first-order system)
template <typename T> void force(ems_tet *s, const T *U, const T *V, const double *fdev)
{
    const int n = (int)s->nd, gr = row_blocks(n, s->lpr);
    const Csr<T> A{n, s->d_rp.as<int>(), s->d_col.as<int>()};
    if (s->ebe) {
        ebe_local<T, EBE_FORCE>(s, 0.0, V, U, nullptr);
        k_ebe_rhs<T><<<(n + 255) / 256, 256, 0, s->stream>>>(s->em, s->d_inc.as<int>(), s->d_tmp.as<T>(),
                                                             s->d_bnd.as<unsigned char>(), V, 0, s->b.as<T>());
    } else with_lpr(s->lpr, [&](auto L) {
        constexpr int LPR = decltype(L)::value;
        if (s->lossy)
            k_force<LPR, T, true><<<gr, 256, 0, s->stream>>>(A, s->d_Ms.as<T>(), s->d_S.as<T>(), U, V, s->b.as<T>());
        else
            k_force<LPR, T, false><<<gr, 256, 0, s->stream>>>(A, nullptr, s->d_S.as<T>(), U, V, s->b.as<T>());
    });
    load<T>(s, s->b, fdev);
}

// a = Me^-1 (-S U - Ms V + load) into dst (initial guess: dst)
template <typename T> ems_status accel(ems_tet *s, const T *U, const T *V, DevBuf &dst, const double *fdev)
{
    force<T>(s, U, V, fdev);
    return solve<T>(s, 0, dst);
}

// One symplectic stage of step D, midpoint tm in units of dt (load factors -p'
// at the midpoint):
//   U += D/2 V;  (M_eps + D/2 M_sigma) V' = (M_eps - D/2 M_sigma) V - D (S U - load);
//   U += D/2 V'
// solved for the increment: (M_eps + D/2 M_sigma) y = -S U - M_sigma V + load,
// V' = V + D y.  The solve's tolerance is then relative to the increment
// rather than to V (in FP32 the error of solving for V' accumulated over the
// steps).  Initial guess: the last two stage solutions extrapolated linearly
// to tm (second order) when the span is moderate, else the last one.
template <typename T> ems_status stage(ems_tet *s, int q, double D, double tm, const double *fdev)
{
    const int n = (int)s->nd, g = (n + 255) / 256;
    const bool comp = sizeof(T) == 4 && !s->comp_off;
    if (comp && !s->comp_valid) {
        EMS_TRY(s->Uc.reserve(sizeof(T) * n));
        EMS_TRY(s->Vc.reserve(sizeof(T) * n));
        EMS_HIP(hipMemsetAsync(s->Uc.p, 0, sizeof(T) * n, s->stream));
        EMS_HIP(hipMemsetAsync(s->Vc.p, 0, sizeof(T) * n, s->stream));
        s->comp_valid = true;
    }
    auto update = [&](T a, const T *x, DevBuf &y, DevBuf &c) {
        if (comp) k_axpy_comp<T><<<g, 256, 0, s->stream>>>(n, a, x, y.as<T>(), c.as<T>());
        else k_axpy<T><<<g, 256, 0, s->stream>>>(n, a, x, y.as<T>());
    };
    update((T)(0.5 * D), s->V.as<T>(), s->U, s->Uc);
    force<T>(s, s->U.as<T>(), s->V.as<T>(), fdev);
    if (!s->X.p) {
        EMS_TRY(s->Y.reserve(sizeof(T) * n));
        EMS_TRY(s->Yp.reserve(sizeof(T) * n));
        EMS_TRY(s->X.reserve(sizeof(T) * n));
        s->have_prev = 0;
    }
    // extrapolate only over moderate spans: symplectic_4's stage midpoints
    // (0.675, 0.5, 0.325 of a step) would give a ratio of -7.7 into the next
    // step, which amplifies the solutions' differences instead of predicting
    const double r = s->have_prev >= 2 && s->ty != s->typ ? (tm - s->ty) / (s->ty - s->typ) : 0.0;
    if (s->have_prev >= 2 && std::fabs(r) <= 1.5)
        k_extrap<T><<<g, 256, 0, s->stream>>>(n, s->Y.as<T>(), s->Yp.as<T>(), (T)r, s->X.as<T>());
    else if (s->have_prev >= 1)
        EMS_HIP(hipMemcpyAsync(s->X.p, s->Y.p, sizeof(T) * n, hipMemcpyDeviceToDevice, s->stream));
    else
        EMS_HIP(hipMemsetAsync(s->X.p, 0, sizeof(T) * n, s->stream));
    EMS_TRY(solve<T>(s, q, s->X));
    update((T)D, s->X.as<T>(), s->V, s->Vc);
    update((T)(0.5 * D), s->V.as<T>(), s->U, s->Uc);
    // history: Yp <- Y <- X
    swap_dev(s->Yp, s->Y);
    swap_dev(s->Y, s->X);
    s->typ = s->ty;
    s->ty = tm;
    s->have_prev = std::min(s->have_prev + 1, 2);
    return EMS_OK;
}

// Largest eigenvalue of M_eps^-1 S by power iteration from a fixed start
// vector (interior edges 1 + (e mod 7) / 7); the Rayleigh quotient
// x'Sx / x'M_eps x of the last iterate, a lower bound of the true value.
template <typename T> ems_status max_frequency(ems_tet *s, int iters, double *omega2)
{
    EMS_TRY(setup<T>(s));
    const int n = (int)s->nd;
    std::vector<T> x(n), sx(n), mx(n);
    for (int e = 0; e < n; ++e) x[e] = s->bnd[e] ? (T)0 : (T)(1.0 + (e % 7) / 7.0);
    DevBuf dx, dy;
    EMS_TRY(dx.reserve(sizeof(T) * n));
    EMS_TRY(dy.reserve(sizeof(T) * n));
    double lam = 0.0;
    for (int it = 0; it <= iters; ++it) {
        EMS_HIP(hipMemcpy(dx.p, x.data(), sizeof(T) * n, hipMemcpyHostToDevice));
        EMS_TRY(apply_op<T>(s, 0, dx.as<T>(), s->b.as<T>()));
        EMS_TRY(apply_op<T>(s, 1, dx.as<T>(), dy.as<T>()));
        EMS_HIP(hipMemcpy(sx.data(), s->b.p, sizeof(T) * n, hipMemcpyDeviceToHost));
        EMS_HIP(hipMemcpy(mx.data(), dy.p, sizeof(T) * n, hipMemcpyDeviceToHost));
        double xs = 0.0, xm = 0.0;
        for (int e = 0; e < n; ++e) {
            xs += (double)x[e] * (double)sx[e];
            xm += (double)x[e] * (double)mx[e];
        }
        if (!(xm > 0.0)) return invalid("mass matrix is not positive on the start vector");
        lam = xs / xm;
        if (it == iters) break;
        // x <- M_eps^-1 S x, normalised (b holds S x)
        EMS_HIP(hipMemset(dy.p, 0, sizeof(T) * n));
        EMS_TRY(solve<T>(s, -1, dy));
        EMS_HIP(hipMemcpy(x.data(), dy.p, sizeof(T) * n, hipMemcpyDeviceToHost));
        double nn = 0.0;
        for (int e = 0; e < n; ++e) nn += (double)x[e] * (double)x[e];
        if (!(nn > 0.0)) break;
        const double sc = 1.0 / std::sqrt(nn);
        for (int e = 0; e < n; ++e) x[e] = (T)(x[e] * sc);
    }
    *omega2 = lam;
    return EMS_OK;
}

template <typename T> ems_status run_impl(ems_tet *s, int N, double *total_ms)
{
    EMS_TRY(setup<T>(s));
    const int tf = s->d.tf, nobs = (int)s->obs_dof.size() / s->nl;
    const long total = s->steps + N;
    if (total / tf + 1 > 2147483647L / 3 / std::max(nobs, 1)) return invalid("too many receiver samples");
    const int nslot = (int)(total / tf + 1);
    if (s->steps == 0) {
        // a run from rest defines every slot: none of an earlier (longer)
        // run or gradient may survive in the tail
        s->traces.assign((size_t)nobs * nslot * 3, 0.0);
        s->nslot = nslot;
    } else if (nslot > s->nslot) {
        std::vector<double> t((size_t)nobs * nslot * 3, 0.0);
        for (int o = 0; o < nobs; ++o)
            for (int q = 0; q < s->nslot; ++q)
                for (int c = 0; c < 3; ++c) t[((size_t)o * nslot + q) * 3 + c] = s->traces[((size_t)o * s->nslot + q) * 3 + c];
        s->traces.swap(t);
        s->nslot = nslot;
    }
    if (N <= 0) return EMS_OK;
    const int n = (int)s->nd, g = (n + 255) / 256;
    const double dt = s->d.dt;
    const int nd = (int)s->dip_par.size() / 3;
    // per-call tables: load factors per stage / evaluation, in order
    std::vector<double> f;
    double t = s->t_now;
    const bool abm = s->d.stepper == EMS_TET_ABM;
    const bool s4 = s->d.stepper == EMS_TET_SYMPLECTIC4;
    const double x1 = 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0)), x2 = 1.0 - 2.0 * x1;
    if (abm && !s->abm_started) dipole_factors(s, t, -1.0, f);  // a_0
    for (int st = 0; st < N; ++st) {
        if (abm) {
            dipole_factors(s, t + dt, -1.0, f);  // F(y_p)
            dipole_factors(s, t + dt, -1.0, f);  // F(y_n+1)
        } else if (s4) {
            const double Dq[3] = {x1 * dt, x2 * dt, x1 * dt}, Oq[3] = {0.0, x1 * dt, (x1 + x2) * dt};
            for (int q = 0; q < 3; ++q) dipole_factors(s, t + Oq[q] + 0.5 * Dq[q], -1.0, f);
        } else {
            dipole_factors(s, t + 0.5 * dt, -1.0, f);
        }
        t += dt;
    }
    if (f.empty()) f.push_back(0.0);
    DevBuf &df = s->d_dip_f;
    EMS_TRY(df.reserve(sizeof(double) * f.size()));
    EMS_HIP(hipMemcpy(df.p, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice));
    // receiver samples of this call: steps s0+1..s0+N that are multiples of tf
    const long s0 = s->steps;
    const int nrec = (int)(total / tf - s0 / tf);
    EMS_TRY(s->d_rec.reserve(sizeof(double) * std::max(1, nobs) * 3 * (size_t)(nrec + 1)));
    const double *fp = df.as<double>();
    size_t fo = 0;
    auto fnext = [&]() { const double *r = fp + fo * (size_t)std::max(nd, 1); fo += 1; return r; };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (total_ms) {
        EMS_HIP(hipEventCreate(&e0));
        EMS_HIP(hipEventCreate(&e1));
        EMS_HIP(hipEventRecord(e0, s->stream));
    }
    if (nobs && s0 == 0)  // slot 0: the initial state, stored after this call's samples
        k_tet_record<T><<<(nobs + 63) / 64, 64, 0, s->stream>>>(s->U.as<T>(), nobs, s->nl, s->d_obs_dof.as<int>(),
                                                                 s->d_obs_w.as<double>(),
                                                                 s->d_rec.as<double>() + (size_t)nrec * nobs * 3);
    if (abm && !s->abm_started) {
        // a_0 from the initial state; the AB2 history starts as (V_0, a_0)
        EMS_TRY(accel<T>(s, s->U.as<T>(), s->V.as<T>(), s->an, fnext()));
        EMS_HIP(hipMemcpyAsync(s->am.p, s->an.p, sizeof(T) * n, hipMemcpyDeviceToDevice, s->stream));
        EMS_HIP(hipMemcpyAsync(s->Vm.p, s->V.p, sizeof(T) * n, hipMemcpyDeviceToDevice, s->stream));
        s->abm_started = true;
    }
    for (int st = 1; st <= N; ++st) {
        if (abm) {
            const T h2 = (T)(0.5 * dt), h12 = (T)(dt / 12.0);
            // predictor y_p = y_n + dt/2 (3 F_n - F_n-1)
            k_comb<T><<<g, 256, 0, s->stream>>>(n, s->U.as<T>(), h2, (T)3, s->V.as<T>(), (T)-1, s->Vm.as<T>(),
                                                (T)0, nullptr, s->Up.as<T>());
            k_comb<T><<<g, 256, 0, s->stream>>>(n, s->V.as<T>(), h2, (T)3, s->an.as<T>(), (T)-1, s->am.as<T>(),
                                                (T)0, nullptr, s->Vp.as<T>());
            EMS_HIP(hipMemcpyAsync(s->Ap.p, s->an.p, sizeof(T) * n, hipMemcpyDeviceToDevice, s->stream));
            EMS_TRY(accel<T>(s, s->Up.as<T>(), s->Vp.as<T>(), s->Ap, fnext()));
            // corrector y_n+1 = y_n + dt/12 (5 F(y_p) + 8 F_n - F_n-1)
            k_comb<T><<<g, 256, 0, s->stream>>>(n, s->U.as<T>(), h12, (T)5, s->Vp.as<T>(), (T)8, s->V.as<T>(),
                                                (T)-1, s->Vm.as<T>(), s->Up.as<T>());
            k_comb<T><<<g, 256, 0, s->stream>>>(n, s->V.as<T>(), h12, (T)5, s->Ap.as<T>(), (T)8, s->an.as<T>(),
                                                (T)-1, s->am.as<T>(), s->Vp.as<T>());
            // history: (V_n-1, a_n-1) <- (V_n, a_n); state <- corrected; a_n+1 (guess a_n)
            swap_dev(s->Vm, s->V);
            swap_dev(s->V, s->Vp);
            swap_dev(s->U, s->Up);
            swap_dev(s->am, s->an);
            EMS_HIP(hipMemcpyAsync(s->an.p, s->am.p, sizeof(T) * n, hipMemcpyDeviceToDevice, s->stream));
            EMS_TRY(accel<T>(s, s->U.as<T>(), s->V.as<T>(), s->an, fnext()));
        } else if (s4) {
            // stage midpoints in units of dt from the absolute step index, so
            // that split runs extrapolate exactly as one run
            const double tn = (double)(s0 + st - 1);
            EMS_TRY(stage<T>(s, 0, x1 * dt, tn + 0.5 * x1, fnext()));
            EMS_TRY(stage<T>(s, 1, x2 * dt, tn + x1 + 0.5 * x2, fnext()));
            EMS_TRY(stage<T>(s, 0, x1 * dt, tn + (x1 + x2) + 0.5 * x1, fnext()));
        } else {
            EMS_TRY(stage<T>(s, 0, dt, (double)(s0 + st - 1) + 0.5, fnext()));
        }
        if (nobs && (s0 + st) % tf == 0)
            k_tet_record<T><<<(nobs + 63) / 64, 64, 0, s->stream>>>(
                s->U.as<T>(), nobs, s->nl, s->d_obs_dof.as<int>(), s->d_obs_w.as<double>(),
                s->d_rec.as<double>() + (size_t)((s0 + st) / tf - s0 / tf - 1) * nobs * 3);
        EMS_HIP(hipGetLastError());
    }
    (void)g;
    if (total_ms) {
        EMS_HIP(hipEventRecord(e1, s->stream));
        EMS_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        EMS_HIP(hipEventElapsedTime(&ms, e0, e1));
        *total_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    EMS_HIP(hipStreamSynchronize(s->stream));
    s->t_now = t;
    s->steps = total;
    if (nobs && (nrec || s0 == 0)) {
        std::vector<double> raw((size_t)(nrec + 1) * nobs * 3);
        EMS_HIP(hipMemcpy(raw.data(), s->d_rec.p, sizeof(double) * raw.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < nrec; ++r)
            for (int o = 0; o < nobs; ++o)
                for (int c = 0; c < 3; ++c)
                    s->traces[((size_t)o * s->nslot + s0 / tf + r + 1) * 3 + c] = raw[((size_t)r * nobs + o) * 3 + c];
        for (int o = 0; s0 == 0 && o < nobs; ++o)
            for (int c = 0; c < 3; ++c) s->traces[(size_t)o * s->nslot * 3 + c] = raw[((size_t)nrec * nobs + o) * 3 + c];
    }
    return EMS_OK;
}

// Misfit gradient of a run from rest at t = 0 (symplectic steppers, lossless
// model): the forward run records the receivers; the backward pass recovers
// each stage's input state by running the stage backwards,
//   U1 = U' - D/2 V';  V = V' + D M_eps^-1 (S U1 + g(t_s));  U = U1 - D/2 V,
// and applies the discrete adjoint of the stage,
//   mu = M_eps^-1 (lV' + D/2 lU');  lU = lU' - D S mu;  lV = M_eps mu + D/2 lU,
// accumulating per tetrahedron eps0 mu' M_T (V - V') (eps_r) and
// -D/2 mu' M_T (V + V') (sigma, at sigma = 0).  `drift` = the recovered
// initial state's largest entry relative to the final state's.
// Tables of the gradient kernels (once per handle): tet -> dofs, element
// masses, boundary mask.
template <typename T> ems_status gradient_tables(ems_tet *s)
{
    if (s->g_t2d.p) return EMS_OK;
    const long nt = s->m.nt;
    const int nl = s->nl;
    std::vector<int> t2d((size_t)nt * nl);
    long tid = 0, dd[20];
    for (int k = 0; k < s->m.nz; ++k)
        for (int j = 0; j < s->m.ny; ++j)
            for (int i = 0; i < s->m.nx; ++i)
                for (int t = 0; t < 6; ++t, ++tid) {
                    tet_dofs(s->m, s->d.order, i, j, k, t, dd);
                    for (int l = 0; l < nl; ++l) t2d[tid * nl + l] = (int)dd[l];
                }
    std::vector<double> mloc(6 * nl * nl), sl(nl * nl);
    for (int t = 0; t < 6; ++t) element_matrices(s->d.order, t, s->m.h, &mloc[t * nl * nl], sl.data());
    std::vector<unsigned char> bnd(s->bnd.begin(), s->bnd.end());
    EMS_TRY(s->g_t2d.reserve(sizeof(int) * t2d.size()));
    EMS_TRY(s->g_mloc.reserve(sizeof(double) * mloc.size()));
    EMS_TRY(s->g_bnd.reserve(bnd.size()));
    EMS_HIP(hipMemcpy(s->g_t2d.p, t2d.data(), sizeof(int) * t2d.size(), hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(s->g_mloc.p, mloc.data(), sizeof(double) * mloc.size(), hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(s->g_bnd.p, bnd.data(), bnd.size(), hipMemcpyHostToDevice));
    return EMS_OK;
}

template <typename T>
ems_status gradient_ckpt(ems_tet *s, int N, const double *d_obs, double *J, double *ge, double *gs, double *drift);

template <typename T>
ems_status gradient_impl(ems_tet *s, int N, const double *d_obs, double *J, double *ge, double *gs, double *drift)
{
    if (s->d.stepper == EMS_TET_ABM) return not_supported("the tetrahedral gradient needs a symplectic stepper");
    for (double v : s->sigma)
        if (v != 0.0) return gradient_ckpt<T>(s, N, d_obs, J, ge, gs, drift);  // lossy: no time reversal
    EMS_TRY(setup<T>(s));
    const long nd = s->nd, nt = s->m.nt;
    const int nl = s->nl, nobs = (int)s->obs_dof.size() / nl, tf = s->d.tf;
    const int n = (int)nd, g = (n + 255) / 256;
    EMS_TRY(gradient_tables<T>(s));
    const size_t vb = sizeof(T) * nd;
    for (DevBuf *v : {&s->g_lU, &s->g_lV, &s->g_mu, &s->g_y, &s->g_vout, &s->g_tmp}) {
        EMS_TRY(v->reserve(vb));
        EMS_HIP(hipMemset(v->p, 0, vb));
    }
    EMS_TRY(s->g_ge.reserve(sizeof(double) * nt));
    EMS_TRY(s->g_gs.reserve(sizeof(double) * nt));
    EMS_HIP(hipMemset(s->g_ge.p, 0, sizeof(double) * nt));
    EMS_HIP(hipMemset(s->g_gs.p, 0, sizeof(double) * nt));
    // forward from rest
    EMS_HIP(hipMemset(s->U.p, 0, vb));
    EMS_HIP(hipMemset(s->V.p, 0, vb));
    s->t_now = 0.0;
    s->steps = 0;
    s->nslot = 0;
    s->traces.assign(0, 0.0);
    s->have_prev = 0;
    s->comp_valid = false;
    s->comp_off = true;  // the reversal below undoes plain updates
    const ems_status fwd = run_impl<T>(s, N, nullptr);
    s->comp_off = false;
    EMS_TRY(fwd);
    const int nslot = N / tf + 1;
    std::vector<double> res((size_t)nslot * nobs * 3, 0.0);  // [slot][obs][3]
    double jj = 0.0;
    for (int o = 0; o < nobs; ++o)
        for (int q = 1; q < nslot; ++q)
            for (int c = 0; c < 3; ++c) {
                const size_t k = ((size_t)o * nslot + q) * 3 + c;
                const double r = s->traces[k] - d_obs[k];
                res[((size_t)q * nobs + o) * 3 + c] = r;
                jj += 0.5 * r * r;
            }
    *J = jj;
    EMS_TRY(s->g_res.reserve(sizeof(double) * std::max<size_t>(res.size(), 1)));
    EMS_HIP(hipMemcpy(s->g_res.p, res.data(), sizeof(double) * res.size(), hipMemcpyHostToDevice));
    // stage schedule and the forward step times (same accumulation as run)
    std::vector<double> Dq, Oq;
    const double dt = s->d.dt;
    if (s->d.stepper == EMS_TET_SYMPLECTIC4) {
        const double x1 = 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0)), x2 = 1.0 - 2.0 * x1;
        Dq = {x1 * dt, x2 * dt, x1 * dt};
        Oq = {0.0, x1 * dt, (x1 + x2) * dt};
    } else {
        Dq = {dt};
        Oq = {0.0};
    }
    const int ns = (int)Dq.size(), ndip = (int)s->dip_par.size() / 3;
    std::vector<double> tstep(N + 1, 0.0);
    for (int k = 1; k <= N; ++k) tstep[k] = tstep[k - 1] + dt;
    // load factors +p'(t_s) of the backward stages, in backward order
    std::vector<double> f;
    for (int st = N; st >= 1; --st)
        for (int q = ns - 1; q >= 0; --q) dipole_factors(s, tstep[st - 1] + Oq[q] + 0.5 * Dq[q], 1.0, f);
    if (f.empty()) f.push_back(0.0);
    EMS_TRY(s->g_f.reserve(sizeof(double) * f.size()));
    EMS_HIP(hipMemcpy(s->g_f.p, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice));
    // scale of the final state for the drift
    auto inf_norm = [&](DevBuf &b) -> double {
        std::vector<T> h(nd);
        if (hipMemcpy(h.data(), b.p, vb, hipMemcpyDeviceToHost) != hipSuccess) return -1.0;
        double mx = 0.0;
        for (T v : h) mx = std::max(mx, std::fabs((double)v));
        return mx;
    };
    EMS_HIP(hipStreamSynchronize(s->stream));
    const double u_end = inf_norm(s->U), v_end = inf_norm(s->V);
    T *U = s->U.as<T>(), *V = s->V.as<T>(), *lU = s->g_lU.as<T>(), *lV = s->g_lV.as<T>();
    T *tmp = s->g_tmp.as<T>(), *vout = s->g_vout.as<T>();
    size_t fo = 0;
    const int tb = (int)((nt + 127) / 128);
    for (int st = N; st >= 1; --st) {
        if (nobs && st % tf == 0)
            k_tet_inject<T><<<1, 64, 0, s->stream>>>(nobs, nl, s->d_obs_dof.as<int>(), s->d_obs_w.as<double>(),
                                                     s->g_res.as<double>() + (size_t)(st / tf) * nobs * 3, lU);
        for (int q = ns - 1; q >= 0; --q, ++fo) {
            const double D = Dq[q];
            // recover the stage input
            EMS_HIP(hipMemcpyAsync(vout, V, vb, hipMemcpyDeviceToDevice, s->stream));
            k_axpy<T><<<g, 256, 0, s->stream>>>(n, (T)(-0.5 * D), V, U);  // U1
            EMS_TRY(apply_op<T>(s, 0, U, s->b.as<T>()));
            load<T>(s, s->b, s->g_f.as<double>() + fo * (size_t)std::max(ndip, 1));
            EMS_TRY(solve<T>(s, -1, s->g_y));
            k_axpy<T><<<g, 256, 0, s->stream>>>(n, (T)D, s->g_y.as<T>(), V);        // V_in
            k_axpy<T><<<g, 256, 0, s->stream>>>(n, (T)(-0.5 * D), V, U);            // U_in
            // adjoint of the stage
            k_comb<T><<<g, 256, 0, s->stream>>>(n, lV, (T)(0.5 * D), (T)1, lU, (T)0, lU, (T)0, nullptr,
                                                s->b.as<T>());
            EMS_TRY(solve<T>(s, -1, s->g_mu));
            if (nl == 6)
                k_tet_grad<T, 6><<<tb, 128, 0, s->stream>>>(nt, s->g_t2d.as<int>(), s->g_mloc.as<double>(),
                                                            s->g_bnd.as<unsigned char>(), s->g_mu.as<T>(), V, vout,
                                                            s->d.eps, -0.5 * D, s->g_ge.as<double>(),
                                                            s->g_gs.as<double>());
            else
                k_tet_grad<T, 20><<<tb, 128, 0, s->stream>>>(nt, s->g_t2d.as<int>(), s->g_mloc.as<double>(),
                                                             s->g_bnd.as<unsigned char>(), s->g_mu.as<T>(), V, vout,
                                                             s->d.eps, -0.5 * D, s->g_ge.as<double>(),
                                                             s->g_gs.as<double>());
            EMS_TRY(apply_op<T>(s, 0, s->g_mu.as<T>(), tmp));
            k_axpy<T><<<g, 256, 0, s->stream>>>(n, (T)(-D), tmp, lU);  // lU1
            EMS_TRY(apply_op<T>(s, 1, s->g_mu.as<T>(), tmp));
            k_comb<T><<<g, 256, 0, s->stream>>>(n, tmp, (T)(0.5 * D), (T)1, lU, (T)0, lU, (T)0, nullptr, lV);
            EMS_HIP(hipGetLastError());
        }
    }
    EMS_HIP(hipStreamSynchronize(s->stream));
    const double u0 = inf_norm(s->U), v0 = inf_norm(s->V);
    if (drift) *drift = std::max(u_end > 0 ? u0 / u_end : u0, v_end > 0 ? v0 / v_end : v0);
    if (ge) EMS_HIP(hipMemcpy(ge, s->g_ge.p, sizeof(double) * nt, hipMemcpyDeviceToHost));
    if (gs) EMS_HIP(hipMemcpy(gs, s->g_gs.p, sizeof(double) * nt, hipMemcpyDeviceToHost));
    // the handle returns to rest at t = 0 and keeps the run's traces
    EMS_HIP(hipMemset(s->U.p, 0, vb));
    EMS_HIP(hipMemset(s->V.p, 0, vb));
    s->t_now = 0.0;
    s->steps = 0;
    s->have_prev = 0;
    s->comp_valid = false;
    return EMS_OK;
}

// y = (eps0 eps_r + hh sigma) M x, PEC rows identity (the stage matrices
// A_+- = M_eps +- D/2 M_sig for hh = +-D/2).
template <typename T> ems_status apply_mass(ems_tet *s, double hh, const T *x, T *y, T *work)
{
    const int n = (int)s->nd;
    if (s->ebe) {
        ebe_local<T, EBE_MASS>(s, hh, x, nullptr, nullptr);
        k_apply<EbeOp<T>, T><<<(n + 255) / 256, 256, 0, s->stream>>>(ebe_op<T>(s, 1), x, y);
    } else {
        const Csr<T> A{n, s->d_rp.as<int>(), s->d_col.as<int>()};
        with_lpr(s->lpr, [&](auto L) {
            constexpr int LPR = decltype(L)::value;
            k_spmv<LPR, T><<<row_blocks(n, LPR), 256, 0, s->stream>>>(A, s->d_Me.as<T>(), x, y);
            k_spmv<LPR, T><<<row_blocks(n, LPR), 256, 0, s->stream>>>(A, s->d_Ms.as<T>(), x, work);
        });
        k_axpy<T><<<(n + 255) / 256, 256, 0, s->stream>>>(n, (T)hh, work, y);
    }
    EMS_HIP(hipGetLastError());
    return EMS_OK;
}

// Misfit gradient with checkpoints (symplectic steppers; the path for lossy
// models, whose integrator cannot run backwards).  The forward run keeps
// (U, V) every K = ceil(sqrt(N)) steps.  Backwards over the segments, each is
// rerun from its checkpoint keeping every stage's input (U, V) and output V,
// then the discrete adjoint of the stages runs over it:
//   mu = A_+^-1 (lV' + D/2 lU');  lU = lU' - D S mu;  lV = A_- mu + D/2 lU,
//   A_+- = M_eps +- D/2 M_sig,
// accumulating eps0 mu' M_T (V_in - V_out) (eps_r) and -D/2 mu' M_T (V_in +
// V_out) (sigma) per tetrahedron.  *drift = 0 (nothing is reconstructed).
template <typename T>
ems_status gradient_ckpt(ems_tet *s, int N, const double *d_obs, double *J, double *ge, double *gs, double *drift)
{
    EMS_TRY(setup<T>(s));
    EMS_TRY(gradient_tables<T>(s));
    const long nd = s->nd, nt = s->m.nt;
    const int nl = s->nl, nobs = (int)s->obs_dof.size() / nl, tf = s->d.tf;
    const int n = (int)nd, g = (n + 255) / 256;
    const size_t vb = sizeof(T) * nd;
    for (DevBuf *v : {&s->g_lU, &s->g_lV, &s->g_mu, &s->g_y, &s->g_vout, &s->g_tmp}) {
        EMS_TRY(v->reserve(vb));
        EMS_HIP(hipMemset(v->p, 0, vb));
    }
    EMS_TRY(s->g_ge.reserve(sizeof(double) * nt));
    EMS_TRY(s->g_gs.reserve(sizeof(double) * nt));
    EMS_HIP(hipMemset(s->g_ge.p, 0, sizeof(double) * nt));
    EMS_HIP(hipMemset(s->g_gs.p, 0, sizeof(double) * nt));
    // stage schedule: step D, offset, stage-matrix index
    const double dt = s->d.dt;
    std::vector<double> Dq, Oq;
    std::vector<int> Wq;
    if (s->d.stepper == EMS_TET_SYMPLECTIC4) {
        const double x1 = 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0)), x2 = 1.0 - 2.0 * x1;
        Dq = {x1 * dt, x2 * dt, x1 * dt};
        Oq = {0.0, x1 * dt, (x1 + x2) * dt};
        Wq = {0, 1, 0};
    } else {
        Dq = {dt};
        Oq = {0.0};
        Wq = {0};
    }
    const int ns = (int)Dq.size(), ndip = (int)s->dip_par.size() / 3;
    const int K = std::max(1, (int)std::ceil(std::sqrt((double)N)));
    const int nseg = (N + K - 1) / K;
    EMS_TRY(s->g_ck.reserve(2 * vb * (size_t)nseg));
    EMS_TRY(s->g_st.reserve(3 * vb * (size_t)K * ns));
    char *ck = static_cast<char *>(s->g_ck.p), *stq = static_cast<char *>(s->g_st.p);
    auto cpy = [&](void *dst, const void *src) { return hipMemcpyAsync(dst, src, vb, hipMemcpyDeviceToDevice, s->stream); };
    // forward from rest, a checkpoint at every segment start (plain updates
    // throughout: the reruns then repeat the forward's arithmetic)
    EMS_HIP(hipMemset(s->U.p, 0, vb));
    EMS_HIP(hipMemset(s->V.p, 0, vb));
    s->t_now = 0.0;
    s->steps = 0;
    s->nslot = 0;
    s->traces.assign(0, 0.0);
    s->have_prev = 0;
    s->comp_valid = false;
    s->comp_off = true;
    struct PlainUpdates {  // restored on every exit
        ems_tet *s;
        ~PlainUpdates() { s->comp_off = false; }
    } plain{s};
    ems_status fwd = EMS_OK;
    for (int q = 0; q < nseg && fwd == EMS_OK; ++q) {
        EMS_HIP(cpy(ck + 2 * vb * q, s->U.p));
        EMS_HIP(cpy(ck + 2 * vb * q + vb, s->V.p));
        fwd = run_impl<T>(s, std::min(K, N - q * K), nullptr);
    }
    EMS_TRY(fwd);
    const int nslot = N / tf + 1;
    std::vector<double> res((size_t)nslot * nobs * 3, 0.0);  // [slot][obs][3]
    double jj = 0.0;
    for (int o = 0; o < nobs; ++o)
        for (int q = 1; q < nslot; ++q)
            for (int c = 0; c < 3; ++c) {
                const size_t k = ((size_t)o * nslot + q) * 3 + c;
                const double r = s->traces[k] - d_obs[k];
                res[((size_t)q * nobs + o) * 3 + c] = r;
                jj += 0.5 * r * r;
            }
    *J = jj;
    EMS_TRY(s->g_res.reserve(sizeof(double) * std::max<size_t>(res.size(), 1)));
    EMS_HIP(hipMemcpy(s->g_res.p, res.data(), sizeof(double) * res.size(), hipMemcpyHostToDevice));
    // load factors of every forward stage (the run's times and signs)
    std::vector<double> f;
    {
        double t = 0.0;
        for (int st = 0; st < N; ++st) {
            for (int q = 0; q < ns; ++q) dipole_factors(s, t + Oq[q] + 0.5 * Dq[q], -1.0, f);
            t += dt;
        }
    }
    if (f.empty()) f.push_back(0.0);
    EMS_TRY(s->g_fwd_f.reserve(sizeof(double) * f.size()));
    EMS_HIP(hipMemcpy(s->g_fwd_f.p, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice));
    const double *fw = s->g_fwd_f.as<double>();
    T *lU = s->g_lU.as<T>(), *lV = s->g_lV.as<T>(), *tmp = s->g_tmp.as<T>(), *work = s->g_y.as<T>();
    const int tb = (int)((nt + 127) / 128);
    ems_status st_err = EMS_OK;
    for (int q = nseg - 1; q >= 0 && st_err == EMS_OK; --q) {
        const int a = q * K, b = std::min(N, a + K);
        // rerun the segment, keeping each stage's (U_in, V_in, V_out)
        EMS_HIP(cpy(s->U.p, ck + 2 * vb * q));
        EMS_HIP(cpy(s->V.p, ck + 2 * vb * q + vb));
        s->have_prev = 0;
        for (int st = a; st < b && st_err == EMS_OK; ++st)
            for (int k = 0; k < ns && st_err == EMS_OK; ++k) {
                char *slot = stq + 3 * vb * ((size_t)(st - a) * ns + k);
                EMS_HIP(cpy(slot, s->U.p));
                EMS_HIP(cpy(slot + vb, s->V.p));
                const double tm = (double)st + (Oq[k] + 0.5 * Dq[k]) / dt;
                st_err = stage<T>(s, Wq[k], Dq[k], tm, fw + ((size_t)st * ns + k) * (size_t)std::max(ndip, 1));
                if (st_err == EMS_OK) EMS_HIP(cpy(slot + 2 * vb, s->V.p));
            }
        // adjoint steps b .. a+1
        for (int st = b; st > a && st_err == EMS_OK; --st) {
            if (nobs && st % tf == 0)
                k_tet_inject<T><<<1, 64, 0, s->stream>>>(nobs, nl, s->d_obs_dof.as<int>(), s->d_obs_w.as<double>(),
                                                         s->g_res.as<double>() + (size_t)(st / tf) * nobs * 3, lU);
            for (int k = ns - 1; k >= 0 && st_err == EMS_OK; --k) {
                const double D = Dq[k];
                const char *slot = stq + 3 * vb * ((size_t)(st - 1 - a) * ns + k);
                const T *Vin = reinterpret_cast<const T *>(slot + vb), *Vout = reinterpret_cast<const T *>(slot + 2 * vb);
                k_comb<T><<<g, 256, 0, s->stream>>>(n, lV, (T)(0.5 * D), (T)1, lU, (T)0, lU, (T)0, nullptr,
                                                    s->b.as<T>());
                st_err = solve<T>(s, Wq[k], s->g_mu);
                if (st_err != EMS_OK) break;
                if (nl == 6)
                    k_tet_grad<T, 6><<<tb, 128, 0, s->stream>>>(nt, s->g_t2d.as<int>(), s->g_mloc.as<double>(),
                                                                s->g_bnd.as<unsigned char>(), s->g_mu.as<T>(), Vin,
                                                                Vout, s->d.eps, -0.5 * D, s->g_ge.as<double>(),
                                                                s->g_gs.as<double>());
                else
                    k_tet_grad<T, 20><<<tb, 128, 0, s->stream>>>(nt, s->g_t2d.as<int>(), s->g_mloc.as<double>(),
                                                                 s->g_bnd.as<unsigned char>(), s->g_mu.as<T>(), Vin,
                                                                 Vout, s->d.eps, -0.5 * D, s->g_ge.as<double>(),
                                                                 s->g_gs.as<double>());
                EMS_TRY(apply_op<T>(s, 0, s->g_mu.as<T>(), tmp));
                k_axpy<T><<<g, 256, 0, s->stream>>>(n, (T)(-D), tmp, lU);  // lU1
                EMS_TRY(apply_mass<T>(s, -0.5 * D, s->g_mu.as<T>(), tmp, work));
                k_comb<T><<<g, 256, 0, s->stream>>>(n, tmp, (T)(0.5 * D), (T)1, lU, (T)0, lU, (T)0, nullptr, lV);
                EMS_HIP(hipGetLastError());
            }
        }
    }
    EMS_TRY(st_err);
    EMS_HIP(hipStreamSynchronize(s->stream));
    if (drift) *drift = 0.0;
    if (ge) EMS_HIP(hipMemcpy(ge, s->g_ge.p, sizeof(double) * nt, hipMemcpyDeviceToHost));
    if (gs) EMS_HIP(hipMemcpy(gs, s->g_gs.p, sizeof(double) * nt, hipMemcpyDeviceToHost));
    // the handle returns to rest at t = 0 and keeps the run's traces
    EMS_HIP(hipMemset(s->U.p, 0, vb));
    EMS_HIP(hipMemset(s->V.p, 0, vb));
    s->t_now = 0.0;
    s->steps = 0;
    s->have_prev = 0;
    s->comp_valid = false;
    return EMS_OK;
}

}  // namespace

extern "C" {

ems_status ems_tet_create(const ems_tet_desc *desc, ems_tet **out)
{
    clear_error();
    if (!desc || !out) return invalid("ems_tet_create: null argument");
    *out = nullptr;
    if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1) return invalid("grid sizes must be >= 1");
    if (desc->order != 1 && desc->order != 2) return invalid("tetrahedral element order must be 1 or 2");
    if (!(desc->h > 0) || !(desc->dt > 0) || !(desc->eps > 0) || !(desc->mu > 0))
        return invalid("h, dt, eps and mu must be positive");
    if (desc->tf <= 0) return invalid("time_frequency must be >= 1");
    if (desc->stepper < EMS_TET_SYMPLECTIC2 || desc->stepper > EMS_TET_ABM) return invalid("unknown stepper");
    if (desc->precision != EMS_F64 && desc->precision != EMS_F32) return invalid("unknown precision");
    if ((double)desc->nx * desc->ny * desc->nz * (desc->order == 1 ? 7.0 : 38.0) > 2.0e9)
        return invalid("mesh too large for 32-bit indices");
    if (ems_device_count() <= 0) {
        set_error("no HIP device available");
        return EMS_ERR_HIP;
    }
    ems_tet *s = new ems_tet();
    s->d = *desc;
    s->esz = desc->precision == EMS_F32 ? 4 : 8;
    s->m.init(desc->nx, desc->ny, desc->nz, desc->h);
    s->nl = basis_size(desc->order);
    s->nd = dof_count(s->m, desc->order);
    *out = s;
    return EMS_OK;
}

void ems_tet_destroy(ems_tet *s) { delete s; }

ems_status ems_tet_set_material(ems_tet *s, const double *eps_r, const double *sigma)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (s->steps > 0) return invalid("the material can only change at rest (before the first step)");
    const size_t nt = (size_t)s->m.nt;
    for (size_t q = 0; eps_r && q < nt; ++q)
        if (!(eps_r[q] > 0)) return invalid("eps_r must be positive");
    for (size_t q = 0; sigma && q < nt; ++q)
        if (!(sigma[q] >= 0)) return invalid("sigma must be non-negative");
    if (eps_r) s->eps_r.assign(eps_r, eps_r + nt);
    else s->eps_r.clear();
    if (sigma) s->sigma.assign(sigma, sigma + nt);
    else s->sigma.clear();
    s->assembled = false;  // reassembled by the next run (the state is at rest)
    return EMS_OK;
}

ems_status ems_tet_add_dipole(ems_tet *s, const double pos[3], const double dir[3], double f0, double t0, double amp)
{
    clear_error();
    if (!s || !pos || !dir) return invalid("null argument");
    if (!(f0 > 0)) return invalid("f0 must be positive");
    int i, j, k;
    double lam[4], w[20][3];
    const int t = locate(s->m, pos, i, j, k, lam);
    if (t < 0) return invalid("dipole outside the mesh");
    basis_at(s->d.order, t, s->m.h, lam, w);
    long e[20];
    tet_dofs(s->m, s->d.order, i, j, k, t, e);
    for (int l = 0; l < s->nl; ++l) {
        s->dip_dof.push_back((int)e[l]);
        const double v = w[l][0] * dir[0] + w[l][1] * dir[1] + w[l][2] * dir[2];
        s->dip_w.push_back(dof_boundary(s->m, e[l]) ? 0.0 : v);
    }
    s->dip_par.push_back(f0);
    s->dip_par.push_back(t0);
    s->dip_par.push_back(amp);
    // per dipole its own factor: weights of dipole q scale factor q
    const int nd = (int)s->dip_dof.size();
    EMS_TRY(s->d_dip_dof.reserve(sizeof(int) * nd));
    EMS_TRY(s->d_dip_w.reserve(sizeof(double) * nd));
    EMS_HIP(hipMemcpy(s->d_dip_dof.p, s->dip_dof.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(s->d_dip_w.p, s->dip_w.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
    return EMS_OK;
}

ems_status ems_tet_add_receiver(ems_tet *s, const double pos[3])
{
    clear_error();
    if (!s || !pos) return invalid("null argument");
    int i, j, k;
    double lam[4], w[20][3];
    const int t = locate(s->m, pos, i, j, k, lam);
    if (t < 0) return invalid("receiver outside the mesh");
    basis_at(s->d.order, t, s->m.h, lam, w);
    long e[20];
    tet_dofs(s->m, s->d.order, i, j, k, t, e);
    for (int l = 0; l < s->nl; ++l) {
        s->obs_dof.push_back((int)e[l]);
        for (int c = 0; c < 3; ++c) s->obs_w.push_back(w[l][c]);
    }
    const int no = (int)s->obs_dof.size() / s->nl;
    EMS_TRY(s->d_obs_dof.reserve(sizeof(int) * s->obs_dof.size()));
    EMS_TRY(s->d_obs_w.reserve(sizeof(double) * s->obs_w.size()));
    EMS_HIP(hipMemcpy(s->d_obs_dof.p, s->obs_dof.data(), sizeof(int) * s->obs_dof.size(), hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(s->d_obs_w.p, s->obs_w.data(), sizeof(double) * s->obs_w.size(), hipMemcpyHostToDevice));
    std::vector<double> t2((size_t)no * s->nslot * 3, 0.0);
    std::copy(s->traces.begin(), s->traces.end(), t2.begin());
    s->traces.swap(t2);
    return EMS_OK;
}

namespace {

// A run that fails after its first device step leaves U, V and the solver
// history part way through a step: mark the state as undefined (argument
// errors are raised before any device work and leave it intact).
ems_status run_checked(ems_tet *s, int num_steps, double *total_ms)
{
    if (s->need_state) {
        set_error("the state is undefined after a failed run: set E and dE/dt (ems_tet_set_field) first");
        return EMS_ERR_STATE;
    }
    const ems_status st =
        s->esz == 4 ? run_impl<float>(s, num_steps, total_ms) : run_impl<double>(s, num_steps, total_ms);
    if (st != EMS_OK && st != EMS_ERR_INVALID_ARG) s->need_state = 3;
    return st;
}

}  // namespace

ems_status ems_tet_run(ems_tet *s, int num_steps)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return run_checked(s, num_steps, nullptr);
}

ems_status ems_tet_run_timed(ems_tet *s, int num_steps, double *total_ms)
{
    clear_error();
    if (!s || !total_ms) return invalid("null argument");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return run_checked(s, num_steps, total_ms);
}

ems_status ems_tet_set_operator(ems_tet *s, int op)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (op < EMS_TET_OP_AUTO || op > EMS_TET_OP_EBE_ROWS) return invalid("unknown operator");
    if (s->steps > 0) return invalid("the operator can only change at rest (before the first step)");
    s->op_kind = op;
    s->assembled = false;
    return EMS_OK;
}

ems_status ems_tet_set_time_step(ems_tet *s, double dt)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (!(dt > 0)) return invalid("dt must be positive");
    if (s->steps > 0) return invalid("the time step can only change before the first step");
    const double old = s->d.dt;
    s->d.dt = dt;
    if (!s->assembled) return EMS_OK;
    const ems_status st = s->esz == 4 ? stage_matrices<float>(s) : stage_matrices<double>(s);
    if (st != EMS_OK) {
        s->d.dt = old;
        const ems_status back = s->esz == 4 ? stage_matrices<float>(s) : stage_matrices<double>(s);
        (void)back;
    }
    return st;
}

ems_status ems_tet_gradient(ems_tet *s, int num_steps, const double *d_obs, size_t count, double *misfit,
                            double *grad_eps_r, double *grad_sigma, double *drift)
{
    clear_error();
    if (!s || !d_obs || !misfit) return invalid("null argument");
    if (num_steps < 1) return invalid("num_steps must be >= 1");
    const int nobs = (int)s->obs_dof.size() / s->nl;
    if (count != (size_t)nobs * (num_steps / s->d.tf + 1) * 3) return invalid("d_obs has the wrong size");
    // the gradient starts from rest and ends at rest: it defines the state
    const ems_status st = s->esz == 4
                              ? gradient_impl<float>(s, num_steps, d_obs, misfit, grad_eps_r, grad_sigma, drift)
                              : gradient_impl<double>(s, num_steps, d_obs, misfit, grad_eps_r, grad_sigma, drift);
    if (st == EMS_OK) s->need_state = 0;
    else if (st != EMS_ERR_INVALID_ARG && st != EMS_ERR_NOT_SUPPORTED) s->need_state = 3;
    return st;
}

ems_status ems_tet_trace_shape(const ems_tet *s, int *nobs, int *nslot)
{
    clear_error();
    if (!s || !nobs || !nslot) return invalid("null argument");
    *nobs = (int)s->obs_dof.size() / s->nl;
    *nslot = s->nslot;
    return EMS_OK;
}

ems_status ems_tet_get_traces(ems_tet *s, double *out, size_t count)
{
    clear_error();
    if (!s || (!out && count)) return invalid("null argument");
    if (count != s->traces.size()) return invalid("trace buffer has the wrong size");
    std::copy(s->traces.begin(), s->traces.end(), out);
    return EMS_OK;
}

ems_status ems_tet_dof_count(const ems_tet *s, size_t *n)
{
    clear_error();
    if (!s || !n) return invalid("null argument");
    *n = (size_t)s->nd;
    return EMS_OK;
}

// which: 0 = E (U), 1 = dE/dt (V)
ems_status ems_tet_get_field(ems_tet *s, int which, double *out, size_t count)
{
    clear_error();
    if (!s || !out) return invalid("null argument");
    if (which < 0 || which > 1) return invalid("field must be 0 (E) or 1 (dE/dt)");
    if (count != (size_t)s->nd) return invalid("field buffer has the wrong size");
    DevBuf &b = which == 0 ? s->U : s->V;
    if (!b.p || b.bytes < count * s->esz) {  // never allocated: the state is at rest
        std::fill(out, out + count, 0.0);
        return EMS_OK;
    }
    EMS_HIP(hipStreamSynchronize(s->stream));
    if (s->esz == 8) {
        EMS_HIP(hipMemcpy(out, b.p, count * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t(count);
        EMS_HIP(hipMemcpy(t.data(), b.p, count * 4, hipMemcpyDeviceToHost));
        std::copy(t.begin(), t.end(), out);
    }
    return EMS_OK;
}

ems_status ems_tet_set_field(ems_tet *s, int which, const double *in, size_t count)
{
    clear_error();
    if (!s || !in) return invalid("null argument");
    if (which < 0 || which > 1) return invalid("field must be 0 (E) or 1 (dE/dt)");
    if (count != (size_t)s->nd) return invalid("field buffer has the wrong size");
    ems_status st = s->esz == 4 ? setup<float>(s) : setup<double>(s);
    if (st != EMS_OK) return st;
    DevBuf &b = which == 0 ? s->U : s->V;
    if (s->esz == 8) {
        EMS_HIP(hipMemcpy(b.p, in, count * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(in, in + count);
        EMS_HIP(hipMemcpy(b.p, t.data(), count * 4, hipMemcpyHostToDevice));
    }
    s->abm_started = false;
    s->have_prev = 0;  // a new state: no stage solution to extrapolate
    s->comp_valid = false;
    s->need_state &= ~(1 << which);
    return EMS_OK;
}

ems_status ems_tet_get_time(const ems_tet *s, double *t)
{
    clear_error();
    if (!s || !t) return invalid("null argument");
    *t = s->t_now;
    return EMS_OK;
}

ems_status ems_tet_assemble(const ems_tet_desc *desc, const double *eps_r, const double *sigma, int *rp, int *col,
                            double *S, double *Me, double *Ms, size_t capacity, size_t *nnz)
{
    clear_error();
    if (!desc || !nnz) return invalid("null argument");
    if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1) return invalid("grid sizes must be >= 1");
    if (desc->order != 1 && desc->order != 2) return invalid("tetrahedral element order must be 1 or 2");
    if (!(desc->h > 0) || !(desc->eps > 0) || !(desc->mu > 0)) return invalid("h, eps and mu must be positive");
    ems_tet s;
    s.d = *desc;
    s.m.init(desc->nx, desc->ny, desc->nz, desc->h);
    s.nl = basis_size(desc->order);
    s.nd = dof_count(s.m, desc->order);
    if (eps_r) s.eps_r.assign(eps_r, eps_r + s.m.nt);
    if (sigma) s.sigma.assign(sigma, sigma + s.m.nt);
    std::vector<double> vS, vM, vG;
    EMS_TRY(assemble(&s, vS, vM, vG));
    *nnz = s.col.size();
    if (capacity == 0) return EMS_OK;
    if (capacity < s.col.size() || !rp || !col || !S || !Me || !Ms) return invalid("output arrays too small");
    std::copy(s.rp.begin(), s.rp.end(), rp);
    std::copy(s.col.begin(), s.col.end(), col);
    std::copy(vS.begin(), vS.end(), S);
    std::copy(vM.begin(), vM.end(), Me);
    std::copy(vG.begin(), vG.end(), Ms);
    return EMS_OK;
}

ems_status ems_tet_element_matrices(int order, int type, double h, double *M, double *S)
{
    clear_error();
    if (order != 1 && order != 2) return invalid("order must be 1 or 2");
    if (type < 0 || type > 5) return invalid("tetrahedron type must be in [0, 5]");
    if (!(h > 0)) return invalid("h must be positive");
    if (!M || !S) return invalid("null argument");
    element_matrices(order, type, h, M, S);
    return EMS_OK;
}

ems_status ems_tet_max_frequency(ems_tet *s, int iters, double *omega2)
{
    clear_error();
    if (!s || !omega2) return invalid("null argument");
    if (iters < 1 || iters > 100000) return invalid("iters must be in [1, 100000]");
    return s->esz == 4 ? max_frequency<float>(s, iters, omega2) : max_frequency<double>(s, iters, omega2);
}

// conjugate-gradient iterations of the last mass solve and the maximum so far
ems_status ems_tet_set_solver(ems_tet *s, int solver)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (solver != EMS_TET_SOLVER_AUTO && solver != EMS_TET_SOLVER_CG) return invalid("unknown mass solver");
    s->solver = solver;
    return EMS_OK;
}

ems_status ems_tet_solver_stats(const ems_tet *s, int *last_iters, int *max_iters)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (last_iters) *last_iters = s->cg_last;
    if (max_iters) *max_iters = s->cg_max;
    return EMS_OK;
}

}  // extern "C"
