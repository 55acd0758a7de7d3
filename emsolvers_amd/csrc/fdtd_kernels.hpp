// Yee FDTD device kernels for gfx950.
//
// Device layout ("unified grid"): all six components share one padded index
// space (I, J, K) in [0, nx+1] x [0, ny+1] x [0, nz+1],
//     u(I, J, K) = off + I + pitch*J + slab*K,   slab = pitch*(ny+2),
// with pitch a multiple of 32 elements, so that every row starts on a 128-byte
// (FP32) / 256-byte (FP64) boundary.  Reference arrays map into it as
//     E*[k][j][i] -> (i, j, k)
//     Hx[k][j][i] -> (i+1, j, k)    Hy[k][j][i] -> (i, j+1, k)    Hz[k][j][i] -> (i, j, k+1)
// (reference layouts: src/fdtd/fdtd_simulation.cpp:55-88), which turns the
// reference update loops (:119-220) into the textbook staggered form
//     H(u) -= ch * (E(u + d) - E(u) ...),   E(u) += ce * (H(u) - H(u - d) ...)
// over one box I, J, K in [0, n]:
//     Hx needs I >= 1, Hy needs J >= 1, Hz needs K >= 1,
//     Ex needs J, K >= 1, Ey needs I, K >= 1, Ez needs I, J >= 1.
// Entries outside a component's range are zero and never written (the
// reference's PEC ghost ring).
//
// Brackets are evaluated as ((a - b) - c) + d outside the CPML, exactly as
// the reference, and the library is compiled with -ffp-contract=off, so the
// FP64 kernels reproduce the reference bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "lane_addr.hpp"

namespace ems {
namespace fdtd {

struct GridView {
    int nx, ny, nz;
    long pitch, slab, off;
};

template <typename T> struct Fields {
    T *ex, *ey, *ez, *hx, *hy, *hz;
};

// CPML: per-axis profile tables indexed by I, J or K (H-type positions at
// I*h, E-type positions at (I - 1/2)*h) and compact psi slabs, 2P+1 planes
// thick per axis: planes I <= P and I >= n - P + 1.
// CPML grading at one position: psi <- b psi + a t, term = t k + psi (k = 1/kappa)
template <typename T> struct alignas(4 * sizeof(T)) Prof {
    T b, a, k, pad;
};

template <typename T> struct PmlView {
    int P, W;
    const Prof<T> *hp[3], *ep[3];  // H-type (I*h) and E-type ((I-1/2)*h) positions
    // psi_h[axis][0/1], psi_e[axis][0/1]:
    //   x: hy_x hz_x | ey_x ez_x    y: hx_y hz_y | ex_y ez_y    z: hx_z hy_z | ex_z ey_z
    T *psi_h[3][2];
    T *psi_e[3][2];
};

__device__ __forceinline__ bool in_pml(int P, int n, int I) { return P > 0 && (I <= P || I >= n - P + 1); }

// Index into a psi slab.  32-bit: every slab has fewer than 2^31 entries
// (ems_fdtd_set_cpml checks it), and the narrower indices keep the CPML
// instantiations of the split sweeps at 6 waves per SIMD in FP32.
__device__ __forceinline__ int slab_x(const GridView &g, int P, int W, int I, int J, int K)
{
    int l = I <= P ? I : I - g.nx + 2 * P;
    return l + W * (J + (g.ny + 2) * K);
}
__device__ __forceinline__ int slab_y(const GridView &g, int P, int W, int I, int J, int K)
{
    int l = J <= P ? J : J - g.ny + 2 * P;
    return I + (g.nx + 2) * (l + W * K);
}
__device__ __forceinline__ int slab_z(const GridView &g, int P, int W, int I, int J, int K)
{
    int l = K <= P ? K : K - g.nz + 2 * P;
    return I + (g.nx + 2) * (J + (g.ny + 2) * l);
}

// psi <- b*psi + a*t ; returns t/kappa + psi  (Roden & Gedney recursion)
template <typename T> __device__ __forceinline__ T cpml_term(T *psi, long q, const Prof<T> pr, T t)
{
    T p = pr.b * psi[q] + pr.a * t;
    psi[q] = p;
    return t * pr.k + p;
}

// The same recursion with psi already in a register (psi_load); stores the
// new psi at q.  Bit for bit cpml_term.
template <typename T> __device__ __forceinline__ T cpml_reg(T psi_old, T *psi, int q, const Prof<T> pr, T t)
{
    T p = pr.b * psi_old + pr.a * t;
    psi[q] = p;
    return t * pr.k + p;
}

// The split sweeps load psi at its plane: prefetching the next plane's psi
// with its fields costs registers, and at 256^3 the wave per SIMD it takes
// is worth more than the latency it hides (FP32 and FP64).

// psi of one cell for the split sweeps, in the order the updates use them:
//   [0] y of component x   [1] z of component x   [2] z of component y
//   [3] x of component y   [4] x of component z   [5] y of component z
// (ps[axis][0/1] as in PmlView: x: y z | y: x z | z: x y components)
template <typename T>
__device__ __forceinline__ void psi_load(const GridView &g, const PmlView<T> &pm, T *const (&ps)[3][2], int I, int J,
                                         int K, bool px, bool py, bool pz, T (&v)[6])
{
    if (py) {
        const int q = slab_y(g, pm.P, pm.W, I, J, K);
        v[0] = ps[1][0][q];
        v[5] = ps[1][1][q];
    }
    if (pz) {
        const int q = slab_z(g, pm.P, pm.W, I, J, K);
        v[1] = ps[2][0][q];
        v[2] = ps[2][1][q];
    }
    if (px) {
        const int q = slab_x(g, pm.P, pm.W, I, J, K);
        v[3] = ps[0][0][q];
        v[4] = ps[0][1][q];
    }
}

// E <- update(E, bracket) for the three material modes
template <typename T, int MAT>
__device__ __forceinline__ T e_update(T e, T br, T ce, const T *__restrict__ ca,
                                      const T *__restrict__ cb, long u)
{
    if (MAT == 0) return e + ce * br;
    if (MAT == 1) return e + cb[u] * br;
    return ca[u] * e + cb[u] * br;
}

// ---------------------------------------------------------------------------
// Split kernels: an H sweep then an E sweep, in place.  Each thread owns one
// (I, J) column and marches up a chunk of lz planes in K, carrying the
// K-neighbour values in registers and loading the next plane before the
// current one's stores; I and J neighbours are served by the caches (same
// wavefront / workgroup / XCD).  Workgroup 64 x BY (BY rows, one wavefront
// per row); a 1-D grid of nbx workgroups per tile (row block, K chunk), in
// the order of split_tile.
// ---------------------------------------------------------------------------
// Column of a split-kernel thread.  With a CPML the 2P+1 columns of the two
// x-slabs (I <= P, I >= nx-P+1) are packed into the first npw wavefronts of
// every row, so that no wavefront mixes CPML and interior lanes (the CPML
// branch would otherwise run for most rows).  The interior wavefronts that
// follow cover I = 64w .. 64w+63, aligned with the rows' 128-byte lines, with
// the x-slab lanes idle: a wavefront starting at I = P+1 would straddle a line
// at each end, and the neighbouring wavefront sits in another workgroup,
// usually on another XCD with an L2 of its own (measured at 256^3 FP32
// with a 10-cell CPML: 0.41 -> 0.38 ms per step).
// Returns -1 for idle lanes.
__device__ __forceinline__ int split_column(const GridView &g, int P, int npw, int bx, int tx)
{
    if (npw == 0) return bx * 64 + tx;
    if (bx < npw) {
        const int L = bx * 64 + tx;
        if (L > 2 * P) return -1;
        return L <= P ? L : g.nx - 2 * P + L;
    }
    const int I = (bx - npw) * 64 + tx;
    return I > P && I <= g.nx - P ? I : -1;
}

// Wavefronts per row of the split sweeps: npw x-slab wavefronts plus the
// aligned interior ones (P > 0), or all columns 0..nx (P = 0).
__host__ __device__ inline int split_row_waves(const GridView &g, int P)
{
    if (P <= 0) return (g.nx + 1 + 63) / 64;
    return (2 * P + 1 + 63) / 64 + (g.nx - P + 1 + 63) / 64;
}

// The split sweeps address every field as (uniform plane base) + (lane's
// offset in the plane, under 4 GiB: ems_fdtd_create checks it), through
// lane_at.  This took the FP32 CPML sweeps from 73-82 VGPRs (6 waves per
// SIMD) to 51-56 (8).

// H sweep over planes [K0, K1) of the column (I, J).  K0 and K1 are
// wave-uniform; I and J may differ between lanes.
template <typename T, bool PML>
__device__ __forceinline__ void h_column(const GridView &g, T ch, const Fields<T> &f, const PmlView<T> &pm, int I,
                                         int J, int K0, int K1)
{
    const T *__restrict__ ex = f.ex;
    const T *__restrict__ ey = f.ey;
    const T *__restrict__ ez = f.ez;
    T *__restrict__ hx = f.hx;
    T *__restrict__ hy = f.hy;
    T *__restrict__ hz = f.hz;
    const long P = g.pitch, S = g.slab;
    long ub = g.off + S * K0;                              // wave-uniform part of the cell index
    const unsigned ui = (unsigned)(I + P * J) * sizeof(T);  // the lane's part, in bytes
    bool px = false, py = false;
    if (PML) {
        px = in_pml(pm.P, g.nx, I);
        py = in_pml(pm.P, g.ny, J);
    }
    T ex_k = lane_at(ex + ub, ui), ey_k = lane_at(ey + ub, ui);
    // values of the current plane; the next plane is loaded before the stores
    T ex_k1 = lane_at(ex + (ub + S), ui), ey_k1 = lane_at(ey + (ub + S), ui);
    T ez_k = lane_at(ez + ub, ui), ez_j1 = lane_at(ez + (ub + P), ui), ez_i1 = lane_at(ez + (ub + 1), ui);
    T ey_i1 = lane_at(ey + (ub + 1), ui), ex_j1 = lane_at(ex + (ub + P), ui);
    T hxo = lane_at(hx + ub, ui), hyo = lane_at(hy + ub, ui), hzo = lane_at(hz + ub, ui);
    // CPML: the column's x and y gradings once (psi is private to its cell)
    Prof<T> prx{}, pry{}, prz{};
    T ps[6] = {};
    bool pz = false;
    if (PML) {
        if (px) prx = pm.hp[0][I];
        if (py) pry = pm.hp[1][J];
    }
    for (int K = K0; K < K1; ++K, ub += S) {
        T n_ex_k1 = 0, n_ey_k1 = 0, n_ez_k = 0, n_ez_j1 = 0, n_ez_i1 = 0, n_ey_i1 = 0, n_ex_j1 = 0;
        T n_hx = 0, n_hy = 0, n_hz = 0;
        if (K + 1 < K1) {
            const long vb = ub + S;
            n_ex_k1 = lane_at(ex + (vb + S), ui);
            n_ey_k1 = lane_at(ey + (vb + S), ui);
            n_ez_k = lane_at(ez + vb, ui);
            n_ez_j1 = lane_at(ez + (vb + P), ui);
            n_ez_i1 = lane_at(ez + (vb + 1), ui);
            n_ey_i1 = lane_at(ey + (vb + 1), ui);
            n_ex_j1 = lane_at(ex + (vb + P), ui);
            n_hx = lane_at(hx + vb, ui);
            n_hy = lane_at(hy + vb, ui);
            n_hz = lane_at(hz + vb, ui);
        }
        if (PML) {
            pz = in_pml(pm.P, g.nz, K);
            if (pz) prz = pm.hp[2][K];
            psi_load(g, pm, pm.psi_h, I, J, K, px, py, pz, ps);
        }
        int qx = 0, qy = 0, qz = 0;
        if (PML) {
            if (px) qx = slab_x(g, pm.P, pm.W, I, J, K);
            if (py) qy = slab_y(g, pm.P, pm.W, I, J, K);
            if (pz) qz = slab_z(g, pm.P, pm.W, I, J, K);
        }
        if (I >= 1) {  // Hx: (Ez(J+1) - Ez(J)) - Ey(K+1) + Ey(K)
            const T a = ez_j1, b = ez_k, c = ey_k1, d = ey_k;
            T br;
            if (PML && (py || pz)) {
                T t1 = a - b, t2 = c - d;
                if (py) t1 = cpml_reg(ps[0], pm.psi_h[1][0], qy, pry, t1);
                if (pz) t2 = cpml_reg(ps[1], pm.psi_h[2][0], qz, prz, t2);
                br = t1 - t2;
            } else {
                br = a - b - c + d;
            }
            lane_at(hx + ub, ui) = hxo - ch * br;
        }
        if (J >= 1) {  // Hy: (Ex(K+1) - Ex(K)) - Ez(I+1) + Ez(I)
            const T a = ex_k1, b = ex_k, c = ez_i1, d = ez_k;
            T br;
            if (PML && (pz || px)) {
                T t1 = a - b, t2 = c - d;
                if (pz) t1 = cpml_reg(ps[2], pm.psi_h[2][1], qz, prz, t1);
                if (px) t2 = cpml_reg(ps[3], pm.psi_h[0][0], qx, prx, t2);
                br = t1 - t2;
            } else {
                br = a - b - c + d;
            }
            lane_at(hy + ub, ui) = hyo - ch * br;
        }
        if (K >= 1) {  // Hz: (Ey(I+1) - Ey(I)) - Ex(J+1) + Ex(J)
            const T a = ey_i1, b = ey_k, c = ex_j1, d = ex_k;
            T br;
            if (PML && (px || py)) {
                T t1 = a - b, t2 = c - d;
                if (px) t1 = cpml_reg(ps[4], pm.psi_h[0][1], qx, prx, t1);
                if (py) t2 = cpml_reg(ps[5], pm.psi_h[1][1], qy, pry, t2);
                br = t1 - t2;
            } else {
                br = a - b - c + d;
            }
            lane_at(hz + ub, ui) = hzo - ch * br;
        }
        ex_k = ex_k1;
        ey_k = ey_k1;
        ex_k1 = n_ex_k1;
        ey_k1 = n_ey_k1;
        ez_k = n_ez_k;
        ez_j1 = n_ez_j1;
        ez_i1 = n_ez_i1;
        ey_i1 = n_ey_i1;
        ex_j1 = n_ex_j1;
        hxo = n_hx;
        hyo = n_hy;
        hzo = n_hz;
    }
}

// E sweep over planes [K0, K1) of the column (I, J), as the H sweep.
template <typename T, int MAT, bool PML>
__device__ __forceinline__ void e_column(const GridView &g, T ce, const Fields<T> &f, const T *__restrict__ ca,
                                         const T *__restrict__ cb, const PmlView<T> &pm, int I, int J, int K0, int K1)
{
    T *__restrict__ ex = f.ex;
    T *__restrict__ ey = f.ey;
    T *__restrict__ ez = f.ez;
    const T *__restrict__ hx = f.hx;
    const T *__restrict__ hy = f.hy;
    const T *__restrict__ hz = f.hz;
    const long P = g.pitch, S = g.slab;
    long ub = g.off + S * K0;                                 // wave-uniform part of the cell index
    const unsigned ui = (unsigned)(I + P * J) * sizeof(T);     // the lane's part, in bytes
    const unsigned uim = ui - (I >= 1 ? (unsigned)sizeof(T) : 0u);       // I-1 (clamped)
    const unsigned ujm = ui - (J >= 1 ? (unsigned)(P * sizeof(T)) : 0u);  // J-1 (clamped)
    bool px = false, py = false;
    if (PML) {
        px = in_pml(pm.P, g.nx, I);
        py = in_pml(pm.P, g.ny, J);
    }
    // H at K-1 (zero below the grid: never used there)
    T hx_km1 = K0 >= 1 ? lane_at(hx + (ub - S), ui) : (T)0, hy_km1 = K0 >= 1 ? lane_at(hy + (ub - S), ui) : (T)0;
    T hx_k = lane_at(hx + ub, ui), hy_k = lane_at(hy + ub, ui), hz_k = lane_at(hz + ub, ui);
    T hz_jm = lane_at(hz + ub, ujm), hy_im = lane_at(hy + ub, uim), hz_im = lane_at(hz + ub, uim);
    T hx_jm = lane_at(hx + ub, ujm);
    T exo = lane_at(ex + ub, ui), eyo = lane_at(ey + ub, ui), ezo = lane_at(ez + ub, ui);
    T cbv = MAT >= 1 ? lane_at(cb + ub, ui) : (T)0, cav = MAT == 2 ? lane_at(ca + ub, ui) : (T)0;
    // CPML as in the H sweep: gradings of the column once
    Prof<T> prx{}, pry{}, prz{};
    T ps[6] = {};
    bool pz = false;
    if (PML) {
        if (px) prx = pm.ep[0][I];
        if (py) pry = pm.ep[1][J];
    }
    for (int K = K0; K < K1; ++K, ub += S) {
        T n_hx = 0, n_hy = 0, n_hz = 0, n_hz_jm = 0, n_hy_im = 0, n_hz_im = 0, n_hx_jm = 0;
        T n_ex = 0, n_ey = 0, n_ez = 0, n_cb = 0, n_ca = 0;
        if (K + 1 < K1) {
            const long vb = ub + S;
            n_hx = lane_at(hx + vb, ui);
            n_hy = lane_at(hy + vb, ui);
            n_hz = lane_at(hz + vb, ui);
            n_hz_jm = lane_at(hz + vb, ujm);
            n_hy_im = lane_at(hy + vb, uim);
            n_hz_im = lane_at(hz + vb, uim);
            n_hx_jm = lane_at(hx + vb, ujm);
            n_ex = lane_at(ex + vb, ui);
            n_ey = lane_at(ey + vb, ui);
            n_ez = lane_at(ez + vb, ui);
            if (MAT >= 1) n_cb = lane_at(cb + vb, ui);
            if (MAT == 2) n_ca = lane_at(ca + vb, ui);
        }
        if (PML) {
            pz = in_pml(pm.P, g.nz, K);
            if (pz) prz = pm.ep[2][K];
            psi_load(g, pm, pm.psi_e, I, J, K, px, py, pz, ps);
        }
        int qx = 0, qy = 0, qz = 0;
        if (PML) {
            if (px) qx = slab_x(g, pm.P, pm.W, I, J, K);
            if (py) qy = slab_y(g, pm.P, pm.W, I, J, K);
            if (pz) qz = slab_z(g, pm.P, pm.W, I, J, K);
        }
        if (J >= 1 && K >= 1) {  // Ex: (Hz(J) - Hz(J-1)) - Hy(K) + Hy(K-1)
            const T a = hz_k, b = hz_jm, c = hy_k, d = hy_km1;
            T br;
            if (PML && (py || pz)) {
                T t1 = a - b, t2 = c - d;
                if (py) t1 = cpml_reg(ps[0], pm.psi_e[1][0], qy, pry, t1);
                if (pz) t2 = cpml_reg(ps[1], pm.psi_e[2][0], qz, prz, t2);
                br = t1 - t2;
            } else {
                br = a - b - c + d;
            }
            lane_at(ex + ub, ui) = MAT == 0 ? exo + ce * br : MAT == 1 ? exo + cbv * br : cav * exo + cbv * br;
        }
        if (I >= 1 && K >= 1) {  // Ey: (Hx(K) - Hx(K-1)) - Hz(I) + Hz(I-1)
            const T a = hx_k, b = hx_km1, c = hz_k, d = hz_im;
            T br;
            if (PML && (pz || px)) {
                T t1 = a - b, t2 = c - d;
                if (pz) t1 = cpml_reg(ps[2], pm.psi_e[2][1], qz, prz, t1);
                if (px) t2 = cpml_reg(ps[3], pm.psi_e[0][0], qx, prx, t2);
                br = t1 - t2;
            } else {
                br = a - b - c + d;
            }
            lane_at(ey + ub, ui) = MAT == 0 ? eyo + ce * br : MAT == 1 ? eyo + cbv * br : cav * eyo + cbv * br;
        }
        if (I >= 1 && J >= 1) {  // Ez: (Hy(I) - Hy(I-1)) - Hx(J) + Hx(J-1)
            const T a = hy_k, b = hy_im, c = hx_k, d = hx_jm;
            T br;
            if (PML && (px || py)) {
                T t1 = a - b, t2 = c - d;
                if (px) t1 = cpml_reg(ps[4], pm.psi_e[0][1], qx, prx, t1);
                if (py) t2 = cpml_reg(ps[5], pm.psi_e[1][1], qy, pry, t2);
                br = t1 - t2;
            } else {
                br = a - b - c + d;
            }
            lane_at(ez + ub, ui) = MAT == 0 ? ezo + ce * br : MAT == 1 ? ezo + cbv * br : cav * ezo + cbv * br;
        }
        hx_km1 = hx_k;
        hy_km1 = hy_k;
        hx_k = n_hx;
        hy_k = n_hy;
        hz_k = n_hz;
        hz_jm = n_hz_jm;
        hy_im = n_hy_im;
        hz_im = n_hz_im;
        hx_jm = n_hx_jm;
        exo = n_ex;
        eyo = n_ey;
        ezo = n_ez;
        cbv = n_cb;
        cav = n_ca;
    }
}

template <typename T, bool PML>
__device__ __forceinline__ void h_split_body(const GridView &g, T ch, const Fields<T> &f, const PmlView<T> &pm,
                                             int lz, int npw, int bx, int by, int bz)
{
    const int I = split_column(g, pm.P, npw, bx, threadIdx.x);
    // a wavefront is one x-row (workgroup 64 x rows): J is wave-uniform
    const int J = __builtin_amdgcn_readfirstlane(by * blockDim.y + threadIdx.y);
    const int K0 = bz * lz;
    const int K1 = min(K0 + lz, g.nz + 1);
    if (I < 0 || I > g.nx || J > g.ny) return;
    h_column<T, PML>(g, ch, f, pm, I, J, K0, K1);
}

template <typename T, int MAT, bool PML>
__device__ __forceinline__ void e_split_body(const GridView &g, T ce, const Fields<T> &f, const T *__restrict__ ca,
                                             const T *__restrict__ cb, const PmlView<T> &pm, int lz, int npw, int bx,
                                             int by, int bz)
{
    const int I = split_column(g, pm.P, npw, bx, threadIdx.x);
    const int J = __builtin_amdgcn_readfirstlane(by * blockDim.y + threadIdx.y);
    const int K0 = bz * lz;
    const int K1 = min(K0 + lz, g.nz + 1);
    if (I < 0 || I > g.nx || J > g.ny) return;
    e_column<T, MAT, PML>(g, ce, f, ca, cb, pm, I, J, K0, K1);
}

// Workgroup order of the split sweeps.  Workgroups go to the 8 XCDs round
// robin by linear id (workgroup b runs on XCD b % 8), each XCD with an L2 of
// its own.  A tile (workgroup row block by, K chunk bz) has nbx workgroups
// along x, which share cache lines: the x-slab wavefronts' lines with the
// first and last interior ones, and each interior wavefront's I+1 line with
// the next.  The 1-D grid therefore gives all nbx workgroups of a tile to one
// XCD, b = 8 (bx + nbx (t / 8)) + t % 8, while the eight XCDs work on eight
// consecutive tiles, as in plain order.  Returns false for the padding
// workgroups of the last eight tiles.
__device__ __forceinline__ bool split_tile(int b, int nbx, int nby, int ntiles, int &bx, int &by, int &bz)
{
    const int c = b & 7, sl = b >> 3;
    bx = sl % nbx;
    const int t = (sl / nbx) * 8 + c;
    if (t >= ntiles) return false;
    by = t % nby;
    bz = t / nby;
    return true;
}

template <typename T, bool PML>
__global__ __launch_bounds__(512) void k_h_split(GridView g, T ch, Fields<T> f, PmlView<T> pm, int lz, int npw,
                                                 int nbx, int nby, int ntiles)
{
    int bx, by, bz;
    if (!split_tile(blockIdx.x, nbx, nby, ntiles, bx, by, bz)) return;
    h_split_body<T, PML>(g, ch, f, pm, lz, npw, bx, by, bz);
}

// The E sweep runs its workgroups in reverse order, so that it starts where
// the H sweep ended, on the planes the caches still hold (measured 1-6 %
// faster at 256^3; the next H sweep then starts where this one ends).
template <typename T, int MAT, bool PML>
__global__ __launch_bounds__(512) void k_e_split(GridView g, T ce, Fields<T> f, const T *__restrict__ ca,
                                                 const T *__restrict__ cb, PmlView<T> pm, int lz, int npw, int nbx,
                                                 int nby, int ntiles)
{
    int bx, by, bz;
    if (!split_tile(gridDim.x - 1 - blockIdx.x, nbx, nby, ntiles, bx, by, bz)) return;
    e_split_body<T, MAT, PML>(g, ce, f, ca, cb, pm, lz, npw, bx, by, bz);
}

// ---------------------------------------------------------------------------
// Split sweeps over boxes (an alternative to one launch per sweep, selected
// by a tiling flag).  With a CPML the interior box (no CPML cell) runs the
// CPML-free instantiation and the shell around it the CPML one, in a second
// launch.  Measured at 256^3 this is slower than one launch: the shell
// launch is small and cannot hide its latency behind the interior's
// streaming, while the CPML instantiation costs interior cells only ~3%.
// A box covers
//   columns: I in [i0, i1), or (packed) the order of split_column over nbx
//            wavefronts (nbx = npw: the x-slab columns only);
//   rows:    J in [j0, j1), or (jpack) the 2P+1 rows of the two y-slabs;
//   planes:  [k0, k1) in chunks of lz per thread.
// With xrows = R > 0 a wavefront holds the 2P+1 x-slab columns of R
// consecutive rows (lane = r*(2P+1) + l), so that the x-slab box does not run
// wavefronts with 2P+1 of 64 lanes busy; J is then not wave-uniform.
// Blocks of one launch are numbered box by box (x fastest, then y, then z).
// ---------------------------------------------------------------------------
struct SplitBox {
    int i0, i1, j0, j1, k0, k1;
    int packed, jpack, xrows;
    int nbx, nby, base;
    int ilo;  // first column (unpacked boxes start at i0 = ilo rounded down to 64)
};

constexpr int kMaxBoxes = 4;
struct SplitBoxes {
    SplitBox b[kMaxBoxes];
    int n;
};

// Cell column of this thread in the box list; false for idle lanes.  XR:
// the list may hold a box with xrows > 0.
template <bool XR>
__device__ __forceinline__ bool box_column(const GridView &g, int P, int npw, int lz, const SplitBoxes &bs, int &I,
                                           int &J, int &K0, int &K1)
{
    const int blk = blockIdx.x;
    SplitBox b = bs.b[0];
#pragma unroll
    for (int q = 1; q < kMaxBoxes; ++q)
        if (q < bs.n && blk >= bs.b[q].base) b = bs.b[q];
    const int lb = blk - b.base;
    const int bx = lb % b.nbx, by = (lb / b.nbx) % b.nby, bz = lb / (b.nbx * b.nby);
    K0 = b.k0 + bz * lz;
    K1 = min(K0 + lz, b.k1);
    if (XR && b.xrows > 0) {
        const int W = 2 * P + 1;
        const int r = (int)threadIdx.x / W, l = (int)threadIdx.x - r * W;
        if (r >= b.xrows) return false;
        I = l <= P ? l : g.nx - 2 * P + l;
        J = b.j0 + (by * (int)blockDim.y + (int)threadIdx.y) * b.xrows + r;
        return J < b.j1;
    }
    I = b.packed ? split_column(g, P, npw, bx, threadIdx.x) : b.i0 + bx * 64 + (int)threadIdx.x;
    const int jj = by * blockDim.y + threadIdx.y;  // wave-uniform
    int Jv;
    if (b.jpack) {
        if (jj > 2 * P) return false;
        Jv = jj <= P ? jj : g.ny - 2 * P + jj;
    } else {
        Jv = b.j0 + jj;
        if (Jv >= b.j1) return false;
    }
    J = XR ? Jv : __builtin_amdgcn_readfirstlane(Jv);
    if (b.packed) return I >= 0;
    return I >= b.ilo && I < b.i1;
}

template <typename T, bool PML>
__global__ __launch_bounds__(512) void k_h_boxes(GridView g, T ch, Fields<T> f, PmlView<T> pm, int lz, int npw,
                                                 SplitBoxes bs)
{
    int I, J, K0, K1;
    if (!box_column<PML>(g, pm.P, npw, lz, bs, I, J, K0, K1)) return;
    h_column<T, PML>(g, ch, f, pm, I, J, K0, K1);
}

template <typename T, int MAT, bool PML>
__global__ __launch_bounds__(512) void k_e_boxes(GridView g, T ce, Fields<T> f, const T *__restrict__ ca,
                                                 const T *__restrict__ cb, PmlView<T> pm, int lz, int npw,
                                                 SplitBoxes bs)
{
    int I, J, K0, K1;
    if (!box_column<PML>(g, pm.P, npw, lz, bs, I, J, K0, K1)) return;
    e_column<T, MAT, PML>(g, ce, f, ca, cb, pm, I, J, K0, K1);
}

// ---------------------------------------------------------------------------
// Batched split sweeps: many independent grids (shots, models of uneven
// size) in one launch per sweep.  Workgroup b belongs to the member whose
// block range [base, base + 8*nbx*ceil(ntiles/8)) contains it (binary search
// over the members' bases, all multiples of 8); inside its member it takes
// the single-grid order of split_tile and runs the single-grid body.
// ---------------------------------------------------------------------------
template <typename T> struct BatchMember {
    GridView g;
    Fields<T> f;
    PmlView<T> pm;
    const T *ca, *cb;
    T ch, ce;
    int lz, npw, nbx, nby, ntiles;
    long base;
};

template <typename T>
__device__ __forceinline__ int batch_find(const BatchMember<T> *__restrict__ m, int nm, long b)
{
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (m[mid].base <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T, bool PML>
__global__ __launch_bounds__(512) void k_h_batch(const BatchMember<T> *__restrict__ mem, int nm)
{
    const long b = blockIdx.x;
    const BatchMember<T> &m = mem[batch_find(mem, nm, b)];
    int bx, by, bz;
    if (!split_tile((int)(b - m.base), m.nbx, m.nby, m.ntiles, bx, by, bz)) return;
    h_split_body<T, PML>(m.g, m.ch, m.f, m.pm, m.lz, m.npw, bx, by, bz);
}

template <typename T, int MAT, bool PML>
__global__ __launch_bounds__(512) void k_e_batch(const BatchMember<T> *__restrict__ mem, int nm)
{
    const long b = gridDim.x - 1L - blockIdx.x;  // reverse order, as k_e_split
    const BatchMember<T> &m = mem[batch_find(mem, nm, b)];
    int bx, by, bz;
    if (!split_tile((int)(b - m.base), m.nbx, m.nby, m.ntiles, bx, by, bz)) return;
    e_split_body<T, MAT, PML>(m.g, m.ce, m.f, m.ca, m.cb, m.pm, m.lz, m.npw, bx, by, bz);
}

// Sources and receivers of every member (block = member).  The step index
// comes from `step` or, when step_ptr is set (graph replay), from the device
// counter *step_ptr (advanced by k_step_advance after the launch).
struct BatchPost {
    int nsrc;
    const int *src_comp;
    const long *src_u;
    const double *src_val;  // [step][nsrc], step relative to the run's first
    int npts;
    const int *pt_comp;
    const long *pt_u;
    double *rec;            // [row][npts], row = step/tf - 1 (relative)
    int tf, step0;          // tf of the member and its run's first global step
};

template <typename T>
__global__ __launch_bounds__(256) void k_post_batch(const BatchMember<T> *__restrict__ mem,
                                                    const BatchPost *__restrict__ post, int step,
                                                    const int *__restrict__ step_ptr)
{
    const BatchMember<T> &m = mem[blockIdx.x];
    const BatchPost &q = post[blockIdx.x];
    const int n = step_ptr ? *step_ptr : step;  // 1-based step of this run
    T *e[3] = {m.f.ex, m.f.ey, m.f.ez};
    if (threadIdx.x == 0) {
        const double *v = q.src_val + (size_t)(n - 1) * q.nsrc;
        for (int s = 0; s < q.nsrc; ++s) {
            T *p = e[q.src_comp[s]] + q.src_u[s];
            *p = *p - (T)v[s];
        }
    }
    __syncthreads();
    if (q.rec && n % q.tf == 0) {
        double *r = q.rec + (size_t)(n / q.tf - 1) * q.npts;
        for (int p = threadIdx.x; p < q.npts; p += blockDim.x) r[p] = (double)e[q.pt_comp[p]][q.pt_u[p]];
    }
}

__global__ void k_step_advance(int *step_ptr)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *step_ptr = *step_ptr + 1;
}

// ---------------------------------------------------------------------------
// Post-step kernel (one block): source injection in the reference's order
// (all x sources, then y, then z; :222-239), then receiver sampling into the
// device trace buffer (row = slot - 1), if this step records.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_post(Fields<T> f, int nsrc, const int *__restrict__ src_comp,
                                              const long *__restrict__ src_u,
                                              const double *__restrict__ src_val, int npts,
                                              const int *__restrict__ pt_comp,
                                              const long *__restrict__ pt_u, double *rec)
{
    T *e[3] = {f.ex, f.ey, f.ez};
    if (threadIdx.x == 0) {
        for (int q = 0; q < nsrc; ++q) {
            T *p = e[src_comp[q]] + src_u[q];
            *p = *p - (T)src_val[q];
        }
    }
    __syncthreads();
    if (rec) {
        for (int p = threadIdx.x; p < npts; p += blockDim.x) rec[p] = (double)e[pt_comp[p]][pt_u[p]];
    }
}

}  // namespace fdtd
}  // namespace ems

namespace ems {
namespace fdtd {

// ---------------------------------------------------------------------------
// Fused single-sweep kernel.
//
// One launch per time step reads E^n, H^{n-1/2} from the "old" buffers and
// writes H^{n+1/2}, E^{n+1} into the "new" buffers (ping-pong), so every
// field word crosses HBM once in each direction per step.  A workgroup of
// 64 x BY threads (one wave per row J) owns a 63 x (BY-1) tile of (I, J)
// columns over a chunk of lz planes and marches K.  E(K) needs H^{n+1/2} at
// I-1, J-1 and K-1:
//   I-1, J-1 : exchanged through LDS (double-buffered by plane parity, one
//              barrier per plane); lane 0 and row 0 form the halo and
//              recompute H there from the old buffers without storing it;
//   K-1      : carried in registers from the previous plane; the plane below
//              the chunk is recomputed as a halo plane.
// The next plane's loads are issued before the barrier.  CPML psi for H is
// ping-ponged too (halo recomputation reads old values); psi for E is updated
// in place by the owning lane.  Tiles that touch the CPML run a PML=true
// instantiation, the others one without CPML code.  The arithmetic per
// output is that of the split kernels: same bits.
// ---------------------------------------------------------------------------
template <typename T> struct PsiH {
    const T *old_[3][2];
    T *new_[3][2];
};

struct TileSet {
    const int *list;  // tile ids (nullptr: identity)
    int count;        // tiles in this launch
    int ntx, nty;     // tiles along I and J
    int lz;           // planes per tile
    int band;         // 1: XCD-banded workgroup order
    int rev;          // 1: tiles in reverse order (every other step)
};

// Loads of one plane for the fused kernel (own cell, +I, +J neighbours).
template <typename T> struct PlaneVals {
    T ex_k1, ey_k1, ez_k, ez_i1, ey_i1, ez_j1, ex_j1, hx, hy, hz;
};

template <typename T>
__device__ __forceinline__ void load_plane(const Fields<T> &o, const GridView &g, long u, PlaneVals<T> &v)
{
    v.ex_k1 = o.ex[u + g.slab];
    v.ey_k1 = o.ey[u + g.slab];
    v.ez_k = o.ez[u];
    v.ez_i1 = o.ez[u + 1];
    v.ey_i1 = o.ey[u + 1];
    v.ez_j1 = o.ez[u + g.pitch];
    v.ex_j1 = o.ex[u + g.pitch];
    v.hx = o.hx[u];
    v.hy = o.hy[u];
    v.hz = o.hz[u];
}

template <bool NT, typename T> __device__ __forceinline__ void st_f(T *p, T v)
{
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// A workgroup of 64 x BY threads owns a 63 x (BY-1) tile of (I, J) columns
// over a chunk of lz planes; lane 0 and row 0 recompute the I-1 / J-1 halo
// of H, the first iteration the K-1 halo plane.  H^{n+1/2} at I-1 / J-1 is
// exchanged through LDS (double-buffered by plane parity: one barrier per
// plane); the next plane's loads are issued before the barrier.
template <typename T, int BY, int MAT, bool PML, bool PF, bool NT>
__global__ __launch_bounds__(64 * BY) void k_fused(GridView g, T ch, T ce, Fields<T> o, Fields<T> nw,
                                                    const T *__restrict__ ca, const T *__restrict__ cb,
                                                    PmlView<T> pm, PsiH<T> ph, TileSet ts)
{
    __shared__ T sH[2][3][BY][64];
    const int per = gridDim.x >> 3;
    int lb = ts.band ? (blockIdx.x & 7) * per + (blockIdx.x >> 3) : blockIdx.x;
    if (lb >= ts.count) return;  // uniform for the whole workgroup
    if (ts.rev) lb = ts.count - 1 - lb;
    const int tile = ts.list ? ts.list[lb] : lb;
    const int ti = tile % ts.ntx;
    const int tj = (tile / ts.ntx) % ts.nty;
    const int kc = tile / (ts.ntx * ts.nty);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int I = ti * 63 - 1 + tx;
    const int J = __builtin_amdgcn_readfirstlane(tj * (BY - 1) - 1 + ty);  // wave-uniform
    const int K0 = kc * ts.lz;
    const int K1 = min(K0 + ts.lz, g.nz + 1);
    const bool act = I >= 0 && J >= 0 && I <= g.nx && J <= g.ny;
    const bool own = act && tx >= 1 && ty >= 1;
    const int Kst = K0 > 0 ? K0 - 1 : 0;
    // old and new buffers never alias: let the compiler move loads past stores
    T *__restrict__ nex = nw.ex;
    T *__restrict__ ney = nw.ey;
    T *__restrict__ nez = nw.ez;
    T *__restrict__ nhx = nw.hx;
    T *__restrict__ nhy = nw.hy;
    T *__restrict__ nhz = nw.hz;
    long u = g.off + (long)(act ? I : 0) + g.pitch * (act ? J : 0) + g.slab * Kst;
    bool px = false, py = false;
    if (PML) {
        px = in_pml(pm.P, g.nx, I);
        py = in_pml(pm.P, g.ny, J);
    }
    T ex_k = 0, ey_k = 0, hx_km1 = 0, hy_km1 = 0;
    PlaneVals<T> cur{}, nxt{};
    if (act) {
        ex_k = o.ex[u];
        ey_k = o.ey[u];
        load_plane(o, g, u, cur);
    }
    for (int K = Kst; K < K1; ++K) {
        const int par = K & 1;
        const bool wr = own && K >= K0;
        // issue the next plane's loads early: they overlap the barrier and E
        if (PF && act && K + 1 < K1) load_plane(o, g, u + g.slab, nxt);
        T cbu = 0, cau = 0;
        if (MAT >= 1 && wr) cbu = cb[u];
        if (MAT == 2 && wr) cau = ca[u];
        T hxn = 0, hyn = 0, hzn = 0;
        const bool pz = PML && in_pml(pm.P, g.nz, K);
        if (act) {
            // Hx: (Ez(J+1) - Ez(J)) - Ey(K+1) + Ey(K)
            if (PML && (py || pz)) {
                T t1 = cur.ez_j1 - cur.ez_k, t2 = cur.ey_k1 - ey_k;
                if (py) {
                    long q = slab_y(g, pm.P, pm.W, I, J, K);
                    const Prof<T> pr = pm.hp[1][J];
                    T p = pr.b * ph.old_[1][0][q] + pr.a * t1;
                    if (wr) ph.new_[1][0][q] = p;
                    t1 = t1 * pr.k + p;
                }
                if (pz) {
                    long q = slab_z(g, pm.P, pm.W, I, J, K);
                    const Prof<T> pr = pm.hp[2][K];
                    T p = pr.b * ph.old_[2][0][q] + pr.a * t2;
                    if (wr) ph.new_[2][0][q] = p;
                    t2 = t2 * pr.k + p;
                }
                hxn = cur.hx - ch * (t1 - t2);
            } else {
                hxn = cur.hx - ch * (cur.ez_j1 - cur.ez_k - cur.ey_k1 + ey_k);
            }
            // Hy: (Ex(K+1) - Ex(K)) - Ez(I+1) + Ez(I)
            if (PML && (pz || px)) {
                T t1 = cur.ex_k1 - ex_k, t2 = cur.ez_i1 - cur.ez_k;
                if (pz) {
                    long q = slab_z(g, pm.P, pm.W, I, J, K);
                    const Prof<T> pr = pm.hp[2][K];
                    T p = pr.b * ph.old_[2][1][q] + pr.a * t1;
                    if (wr) ph.new_[2][1][q] = p;
                    t1 = t1 * pr.k + p;
                }
                if (px) {
                    long q = slab_x(g, pm.P, pm.W, I, J, K);
                    const Prof<T> pr = pm.hp[0][I];
                    T p = pr.b * ph.old_[0][0][q] + pr.a * t2;
                    if (wr) ph.new_[0][0][q] = p;
                    t2 = t2 * pr.k + p;
                }
                hyn = cur.hy - ch * (t1 - t2);
            } else {
                hyn = cur.hy - ch * (cur.ex_k1 - ex_k - cur.ez_i1 + cur.ez_k);
            }
            // Hz: (Ey(I+1) - Ey(I)) - Ex(J+1) + Ex(J)
            if (PML && (px || py)) {
                T t1 = cur.ey_i1 - ey_k, t2 = cur.ex_j1 - ex_k;
                if (px) {
                    long q = slab_x(g, pm.P, pm.W, I, J, K);
                    const Prof<T> pr = pm.hp[0][I];
                    T p = pr.b * ph.old_[0][1][q] + pr.a * t1;
                    if (wr) ph.new_[0][1][q] = p;
                    t1 = t1 * pr.k + p;
                }
                if (py) {
                    long q = slab_y(g, pm.P, pm.W, I, J, K);
                    const Prof<T> pr = pm.hp[1][J];
                    T p = pr.b * ph.old_[1][1][q] + pr.a * t2;
                    if (wr) ph.new_[1][1][q] = p;
                    t2 = t2 * pr.k + p;
                }
                hzn = cur.hz - ch * (t1 - t2);
            } else {
                hzn = cur.hz - ch * (cur.ey_i1 - ey_k - cur.ex_j1 + ex_k);
            }
        }
        sH[par][0][ty][tx] = hxn;
        sH[par][1][ty][tx] = hyn;
        sH[par][2][ty][tx] = hzn;
        __syncthreads();
        if (wr) {
            if (I >= 1) st_f<NT>(&nhx[u], hxn);
            if (J >= 1) st_f<NT>(&nhy[u], hyn);
            if (K >= 1) st_f<NT>(&nhz[u], hzn);
            const T hz_im1 = sH[par][2][ty][tx - 1], hy_im1 = sH[par][1][ty][tx - 1];
            const T hz_jm1 = sH[par][2][ty - 1][tx], hx_jm1 = sH[par][0][ty - 1][tx];
            if (J >= 1 && K >= 1) {  // Ex: (Hz(J) - Hz(J-1)) - Hy(K) + Hy(K-1)
                T br;
                if (PML && (py || pz)) {
                    T t1 = hzn - hz_jm1, t2 = hyn - hy_km1;
                    if (py) t1 = cpml_term(pm.psi_e[1][0], slab_y(g, pm.P, pm.W, I, J, K), pm.ep[1][J], t1);
                    if (pz) t2 = cpml_term(pm.psi_e[2][0], slab_z(g, pm.P, pm.W, I, J, K), pm.ep[2][K], t2);
                    br = t1 - t2;
                } else {
                    br = hzn - hz_jm1 - hyn + hy_km1;
                }
                st_f<NT>(&nex[u], MAT == 0 ? ex_k + ce * br : MAT == 1 ? ex_k + cbu * br : cau * ex_k + cbu * br);
            }
            if (I >= 1 && K >= 1) {  // Ey: (Hx(K) - Hx(K-1)) - Hz(I) + Hz(I-1)
                T br;
                if (PML && (pz || px)) {
                    T t1 = hxn - hx_km1, t2 = hzn - hz_im1;
                    if (pz) t1 = cpml_term(pm.psi_e[2][1], slab_z(g, pm.P, pm.W, I, J, K), pm.ep[2][K], t1);
                    if (px) t2 = cpml_term(pm.psi_e[0][0], slab_x(g, pm.P, pm.W, I, J, K), pm.ep[0][I], t2);
                    br = t1 - t2;
                } else {
                    br = hxn - hx_km1 - hzn + hz_im1;
                }
                st_f<NT>(&ney[u], MAT == 0 ? ey_k + ce * br : MAT == 1 ? ey_k + cbu * br : cau * ey_k + cbu * br);
            }
            if (I >= 1 && J >= 1) {  // Ez: (Hy(I) - Hy(I-1)) - Hx(J) + Hx(J-1)
                T br;
                if (PML && (px || py)) {
                    T t1 = hyn - hy_im1, t2 = hxn - hx_jm1;
                    if (px) t1 = cpml_term(pm.psi_e[0][1], slab_x(g, pm.P, pm.W, I, J, K), pm.ep[0][I], t1);
                    if (py) t2 = cpml_term(pm.psi_e[1][1], slab_y(g, pm.P, pm.W, I, J, K), pm.ep[1][J], t2);
                    br = t1 - t2;
                } else {
                    br = hyn - hy_im1 - hxn + hx_jm1;
                }
                st_f<NT>(&nez[u], MAT == 0 ? cur.ez_k + ce * br : MAT == 1 ? cur.ez_k + cbu * br : cau * cur.ez_k + cbu * br);
            }
        }
        ex_k = cur.ex_k1;
        ey_k = cur.ey_k1;
        hx_km1 = hxn;
        hy_km1 = hyn;
        u += g.slab;
        if (PF) cur = nxt;
        else if (act && K + 1 < K1) load_plane(o, g, u, cur);
    }
}

// ---------------------------------------------------------------------------
// Z-marching fused sweep: one launch per step, no LDS and no barriers.
//
// Reads E^n, H^{n-1/2} from the old buffers and writes H^{n+1/2}, E^{n+1}
// into the new ones (ping-pong, as k_fused; the H-psi too), so every field
// word crosses HBM about once in each direction per step.  Every wavefront is
// independent: lane l holds column I = 62 t - 1 + l of x-tile t, the wave
// owns the 62 columns of lanes 1..62 and a strip of TY rows J0 .. J0+TY-1,
// and it marches a chunk of lz planes.
//   I+1 / I-1: the next / previous lane's registers by DPP wave shifts; lane
//              0 (column I0-1) only computes H, lane 63 only provides E^n;
//   J+1 / J-1: the strip's rows are unrolled in registers; row J0-1 is an H
//              halo row, row J0+TY provides Ex, Ez;
//   K+1 / K-1: the next plane's E^n is loaded a plane ahead, H^{n+1/2} of the
//              plane below stays in registers; the plane below the chunk is
//              an H halo plane.
// Halo H values are computed from the old buffers (old psi) and never
// stored.  Per output the arithmetic is that of k_fused and the split
// sweeps: the same bits.  Fields are addressed as wave-uniform row bases plus
// the lane's 32-bit byte offset (lane_at).
// ---------------------------------------------------------------------------

// lane l receives v of lane l + 1 (DPP wave_shl:1); lane 63 receives 0
__device__ __forceinline__ float lane_next(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}
// lane l receives v of lane l - 1 (DPP wave_shr:1); lane 0 receives 0
__device__ __forceinline__ float lane_prev(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_next(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, 0x130, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0x130, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_prev(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, 0x138, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0x138, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

struct ZTiles {
    int ntx;    // x-tiles of 62 columns
    int nsg;    // strip groups (NW strips of TY rows each)
    int lz;     // planes per chunk
    int count;  // workgroups
    int per;    // workgroups per XCD: XCD c = b % 8 takes tiles [c per, (c+1) per)
};

template <typename T, int TY, int NW, int MAT, bool PML>
__global__ __launch_bounds__(64 * NW) void k_zmarch(GridView g, T ch, T ce, Fields<T> o, Fields<T> nw,
                                                     const T *__restrict__ ca, const T *__restrict__ cb,
                                                     PmlView<T> pm, PsiH<T> ph, ZTiles zt)
{
    const int lb = (blockIdx.x & 7) * zt.per + (blockIdx.x >> 3);
    if (lb >= zt.count) return;  // uniform for the workgroup
    const int ti = lb % zt.ntx;
    const int sg = (lb / zt.ntx) % (zt.nsg);
    const int kc = lb / (zt.ntx * zt.nsg);
    const int lane = threadIdx.x;
    const int I = ti * 62 - 1 + lane;
    const int J0 = __builtin_amdgcn_readfirstlane((sg * NW + (int)threadIdx.y) * TY);
    if (J0 > g.ny) return;  // uniform for the wave; no barriers follow
    const int K0 = kc * zt.lz;
    const int K1 = min(K0 + zt.lz, g.nz + 1);
    const int Kst = K0 > 0 ? K0 - 1 : 0;
    const bool xin = I >= 0 && I <= g.nx;                  // a column of the update box
    const bool xown = lane >= 1 && lane <= 62 && I <= g.nx;  // H stored and E computed here
    const unsigned lo = (unsigned)(I + 1) * (unsigned)sizeof(T);
    bool px = false;  // x-slab column (update box only)
    Prof<T> hpx{}, epx{};
    if (PML && xin) {
        px = in_pml(pm.P, g.nx, I);
        if (px) {
            hpx = pm.hp[0][I];
            epx = pm.ep[0][I];
        }
    }
    // row r of the strip is J = J0 - 1 + r: r = 0 the H halo row, 1..TY own,
    // TY + 1 the Ex / Ez provider
    bool rv[TY + 2], ro[TY + 2];
#pragma unroll
    for (int r = 0; r < TY + 2; ++r) {
        const int J = J0 - 1 + r;
        rv[r] = J >= 0 && J <= g.ny + 1;
        ro[r] = r >= 1 && r <= TY && J <= g.ny;
    }
    // addresses: a wave-uniform plane base per array (plane K, column -1)
    // plus the lane's byte offset within the plane, per row (32-bit: a plane
    // is under 4 GiB); keeps the scalar registers to one base per array
    unsigned lro[TY + 2];
#pragma unroll
    for (int r = 0; r < TY + 2; ++r) lro[r] = lo + (unsigned)((long)(J0 - 1 + r) * g.pitch) * (unsigned)sizeof(T);
    auto pl = [&](auto *a, int K) { return a + (g.off - 1 + g.slab * (long)K); };
    // Loads run ahead of their use so that a wave keeps a plane's worth of
    // loads in flight: E^n two planes ahead (H(K) needs E(K+1)), H^{n-1/2}
    // and the material one plane ahead.  All loads of a plane are issued
    // before its stores (the old and new buffers never alias, but the
    // compiler cannot see it through the lane addressing).
    T ex0[TY + 2], ey0[TY + 2], ez0[TY + 2];  // E^n of plane K
    T ex1[TY + 2], ey1[TY + 2], ez1[TY + 2];  // E^n of plane K + 1
    T hx0[TY + 1], hy0[TY + 1], hz0[TY + 1];  // H^{n-1/2} of plane K
    T cb0[TY + 1], ca0[TY + 1];               // material of plane K (own rows)
    T hxp[TY + 1], hyp[TY + 1];               // H^{n+1/2} of the plane below, per own row
    auto load_e = [&](int K, T (&ex)[TY + 2], T (&ey)[TY + 2], T (&ez)[TY + 2]) {
#pragma unroll
        for (int r = 0; r < TY + 2; ++r) {
            ex[r] = ey[r] = ez[r] = (T)0;
            if (rv[r] && K <= K1) {  // H(K1 - 1) needs E(K1); K1 <= nz + 1
                ex[r] = lane_at(pl(o.ex, K), lro[r]);
                ez[r] = lane_at(pl(o.ez, K), lro[r]);
                if (r <= TY) ey[r] = lane_at(pl(o.ey, K), lro[r]);
            }
        }
    };
    auto load_h = [&](int K, T (&hx)[TY + 1], T (&hy)[TY + 1], T (&hz)[TY + 1], T (&cbv)[TY + 1],
                      T (&cav)[TY + 1]) {
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
            hx[r] = hy[r] = hz[r] = cbv[r] = cav[r] = (T)0;
            if (rv[r] && K < K1) {
                hx[r] = lane_at(pl(o.hx, K), lro[r]);
                if (r > 0) hy[r] = lane_at(pl(o.hy, K), lro[r]);
                hz[r] = lane_at(pl(o.hz, K), lro[r]);
                if (ro[r] && K >= K0) {
                    if (MAT >= 1) cbv[r] = lane_at(pl(cb, K), lro[r]);
                    if (MAT == 2) cav[r] = lane_at(pl(ca, K), lro[r]);
                }
            }
        }
    };
#pragma unroll
    for (int r = 0; r <= TY; ++r) hxp[r] = hyp[r] = (T)0;
    load_e(Kst, ex0, ey0, ez0);
    load_e(Kst + 1, ex1, ey1, ez1);
    load_h(Kst, hx0, hy0, hz0, cb0, ca0);
    for (int K = Kst; K < K1; ++K) {
        const bool eplane = K >= K0;
        const bool pz = PML && in_pml(pm.P, g.nz, K);
        Prof<T> hpz{}, epz{};
        if (pz) {
            hpz = pm.hp[2][K];
            epz = pm.ep[2][K];
        }
        // in flight while this plane is computed
        T ex2[TY + 2], ey2[TY + 2], ez2[TY + 2];
        T hx1[TY + 1], hy1[TY + 1], hz1[TY + 1], cb1[TY + 1], ca1[TY + 1];
        load_e(K + 2, ex2, ey2, ez2);
        load_h(K + 1, hx1, hy1, hz1, cb1, ca1);
        T hzm = 0, hxm = 0;  // H^{n+1/2} of the row below in this plane
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
            if (!rv[r]) continue;  // uniform; only row 0 of the first strip and rows past ny + 1
            const int J = J0 - 1 + r;
            const bool own = ro[r] && eplane;  // uniform
            const bool wr = own && xown;
            const T ez_i1 = lane_next(ez0[r]);
            const T ey_i1 = lane_next(ey0[r]);
            const T ex_k = ex0[r], ey_k = ey0[r], ez_k = ez0[r];
            const T ez_j1 = ez0[r + 1], ex_j1 = ex0[r + 1];
            const T ex_k1 = ex1[r], ey_k1 = ey1[r];
            const T hxo = hx0[r], hyo = hy0[r], hzo = hz0[r];
            const bool py = PML && in_pml(pm.P, g.ny, J);
            // CPML terms only inside the update box: H at I = nx + 1 or
            // J = ny + 1 is neither stored nor used
            const bool yin = J <= g.ny;
            const bool pyl = py && xin && yin, pzl = pz && xin && yin, pxl = px && yin;
            T hxn, hyn, hzn;
            // Hx: (Ez(J+1) - Ez(J)) - Ey(K+1) + Ey(K)
            if (PML && (pyl || pzl)) {
                T t1 = ez_j1 - ez_k, t2 = ey_k1 - ey_k;
                if (pyl) {
                    const int q = slab_y(g, pm.P, pm.W, I, J, K);
                    const Prof<T> pr = pm.hp[1][J];
                    const T p = pr.b * ph.old_[1][0][q] + pr.a * t1;
                    if (wr) ph.new_[1][0][q] = p;
                    t1 = t1 * pr.k + p;
                }
                if (pzl) {
                    const int q = slab_z(g, pm.P, pm.W, I, J, K);
                    const T p = hpz.b * ph.old_[2][0][q] + hpz.a * t2;
                    if (wr) ph.new_[2][0][q] = p;
                    t2 = t2 * hpz.k + p;
                }
                hxn = hxo - ch * (t1 - t2);
            } else {
                hxn = hxo - ch * (ez_j1 - ez_k - ey_k1 + ey_k);
            }
            // Hy: (Ex(K+1) - Ex(K)) - Ez(I+1) + Ez(I)
            if (PML && (pzl || pxl)) {
                T t1 = ex_k1 - ex_k, t2 = ez_i1 - ez_k;
                if (pzl) {
                    const int q = slab_z(g, pm.P, pm.W, I, J, K);
                    const T p = hpz.b * ph.old_[2][1][q] + hpz.a * t1;
                    if (wr) ph.new_[2][1][q] = p;
                    t1 = t1 * hpz.k + p;
                }
                if (pxl) {
                    const int q = slab_x(g, pm.P, pm.W, I, J, K);
                    const T p = hpx.b * ph.old_[0][0][q] + hpx.a * t2;
                    if (wr) ph.new_[0][0][q] = p;
                    t2 = t2 * hpx.k + p;
                }
                hyn = hyo - ch * (t1 - t2);
            } else {
                hyn = hyo - ch * (ex_k1 - ex_k - ez_i1 + ez_k);
            }
            // Hz: (Ey(I+1) - Ey(I)) - Ex(J+1) + Ex(J)
            if (PML && (pxl || pyl)) {
                T t1 = ey_i1 - ey_k, t2 = ex_j1 - ex_k;
                if (pxl) {
                    const int q = slab_x(g, pm.P, pm.W, I, J, K);
                    const T p = hpx.b * ph.old_[0][1][q] + hpx.a * t1;
                    if (wr) ph.new_[0][1][q] = p;
                    t1 = t1 * hpx.k + p;
                }
                if (pyl) {
                    const int q = slab_y(g, pm.P, pm.W, I, J, K);
                    const Prof<T> pr = pm.hp[1][J];
                    const T p = pr.b * ph.old_[1][1][q] + pr.a * t2;
                    if (wr) ph.new_[1][1][q] = p;
                    t2 = t2 * pr.k + p;
                }
                hzn = hzo - ch * (t1 - t2);
            } else {
                hzn = hzo - ch * (ey_i1 - ey_k - ex_j1 + ex_k);
            }
            if (own) {
                const T hz_im1 = lane_prev(hzn), hy_im1 = lane_prev(hyn);
                if (xown) {
                    const T cbu = cb0[r], cau = ca0[r];
                    if (I >= 1) lane_at(pl(nw.hx, K), lro[r]) = hxn;
                    if (J >= 1) lane_at(pl(nw.hy, K), lro[r]) = hyn;
                    if (K >= 1) lane_at(pl(nw.hz, K), lro[r]) = hzn;
                    const T hy_km1 = hyp[r], hx_km1 = hxp[r];
                    if (J >= 1 && K >= 1) {  // Ex: (Hz(J) - Hz(J-1)) - Hy(K) + Hy(K-1)
                        T br;
                        if (PML && (py || pz)) {
                            T t1 = hzn - hzm, t2 = hyn - hy_km1;
                            if (py) t1 = cpml_term(pm.psi_e[1][0], slab_y(g, pm.P, pm.W, I, J, K), pm.ep[1][J], t1);
                            if (pz) t2 = cpml_term(pm.psi_e[2][0], slab_z(g, pm.P, pm.W, I, J, K), epz, t2);
                            br = t1 - t2;
                        } else {
                            br = hzn - hzm - hyn + hy_km1;
                        }
                        lane_at(pl(nw.ex, K), lro[r]) = MAT == 0 ? ex_k + ce * br : MAT == 1 ? ex_k + cbu * br : cau * ex_k + cbu * br;
                    }
                    if (I >= 1 && K >= 1) {  // Ey: (Hx(K) - Hx(K-1)) - Hz(I) + Hz(I-1)
                        T br;
                        if (PML && (pz || px)) {
                            T t1 = hxn - hx_km1, t2 = hzn - hz_im1;
                            if (pz) t1 = cpml_term(pm.psi_e[2][1], slab_z(g, pm.P, pm.W, I, J, K), epz, t1);
                            if (px) t2 = cpml_term(pm.psi_e[0][0], slab_x(g, pm.P, pm.W, I, J, K), epx, t2);
                            br = t1 - t2;
                        } else {
                            br = hxn - hx_km1 - hzn + hz_im1;
                        }
                        lane_at(pl(nw.ey, K), lro[r]) = MAT == 0 ? ey_k + ce * br : MAT == 1 ? ey_k + cbu * br : cau * ey_k + cbu * br;
                    }
                    if (I >= 1 && J >= 1) {  // Ez: (Hy(I) - Hy(I-1)) - Hx(J) + Hx(J-1)
                        T br;
                        if (PML && (px || py)) {
                            T t1 = hyn - hy_im1, t2 = hxn - hxm;
                            if (px) t1 = cpml_term(pm.psi_e[0][1], slab_x(g, pm.P, pm.W, I, J, K), epx, t1);
                            if (py) t2 = cpml_term(pm.psi_e[1][1], slab_y(g, pm.P, pm.W, I, J, K), pm.ep[1][J], t2);
                            br = t1 - t2;
                        } else {
                            br = hyn - hy_im1 - hxn + hxm;
                        }
                        lane_at(pl(nw.ez, K), lro[r]) = MAT == 0 ? ez_k + ce * br : MAT == 1 ? ez_k + cbu * br : cau * ez_k + cbu * br;
                    }
                }
            }
            hxp[r] = hxn;
            hyp[r] = hyn;
            hzm = hzn;
            hxm = hxn;
        }
#pragma unroll
        for (int r = 0; r < TY + 2; ++r) {
            ex0[r] = ex1[r];
            ey0[r] = ey1[r];
            ez0[r] = ez1[r];
            ex1[r] = ex2[r];
            ey1[r] = ey2[r];
            ez1[r] = ez2[r];
        }
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
            hx0[r] = hx1[r];
            hy0[r] = hy1[r];
            hz0[r] = hz1[r];
            cb0[r] = cb1[r];
            ca0[r] = ca1[r];
        }
    }
}

}  // namespace fdtd
}  // namespace ems

namespace ems {
namespace fdtd {

// ---------------------------------------------------------------------------
// Discrete adjoint of the Yee step (no reference counterpart): the exact
// transpose of one step (H update, then E update, CPML included), applied
// backwards in time, for the gradient of the receiver misfit with respect to
// the per-cell eps_r and sigma.  Over the unified grid the forward E update
// reads backward differences of H and the H update forward differences of E,
// so the transpose is a gather of forward (pass A) and backward (pass B)
// differences:
//   pass A (E update transposed):  mu = Cb e'; per derivative u = +-mu,
//     in the CPML of its axis  p = psiE' + u, tbar = u k + a p, psiE = b p,
//     else tbar = u;  hbar += (tbar(own) - tbar(+1)) + (tbar(own) - tbar(+1)).
//   pass B (H update transposed):  nu = -ch hbar, the same CPML transpose with
//     the H-type profiles, e = Ca e' + (sbar(-1) - sbar(own)) + (sbar(-1) - sbar(own));
//     gradient sums P += e' E^n, Q += e' (E^{n+1} - Ca E^n) (FP64), whence
//     dJ/deps_r = dCa/deps_r P + (dCb/deps_r)/Cb Q and likewise for sigma.
// Every term of a cell outside a component's update range is 0.
// ---------------------------------------------------------------------------
template <typename T> struct Adj3 {
    T *c[3];
};

// psi adjoint slabs (as PmlView; old and new buffers) and CPML gradings
template <typename T> struct AdjPml {
    int P, W;
    const Prof<T> *hp[3], *ep[3];
    const T *pe_old[3][2];
    T *pe_new[3][2];
    const T *ph_old[3][2];
    T *ph_new[3][2];
};

// Material derivative factors on the (nx+1)(ny+1)(nz+1) material grid.
struct GradView {
    const double *dca_e, *dcbn_e, *dca_s, *dcbn_s;
    double *g_eps, *g_sig;
};

// ---------------------------------------------------------------------------
// Adjoint sweeps over columns, in the layout and workgroup order of the
// forward split sweeps (split_column, split_tile; 64 x 4 workgroups).  Each
// thread owns one (I, J) column and ADJ_LZ planes of it.
//   k_adj_h (transpose of the E update): tbar of the three E components at
//     the cell and at its I+1, J+1 and K+1 neighbours, hbar += forward
//     differences, new psi_E adjoints (ping-pong: the neighbours read the old
//     ones).
//   k_adj_e (transpose of the H update): sbar at the cell and at its I-1,
//     J-1 and K-1 neighbours; the adjoint E is updated in place,
//     e <- Ca e + backward differences (each thread reads only its own e);
//     new psi_H adjoints; with GRAD the gradient sums of the cell,
//       P += sum_c e_c E^n_c,   Q += sum_c e_c (E^{n+1}_c - Ca E^n_c),
//     from the adjoint e before its update (FP64; the factors of eps_r and
//     sigma multiply P and Q once, at the end).
// The arithmetic of every term is that of the point-wise definition: terms of
// cells outside a component's update range come out as (signed) zeros, since
// the adjoint fields and psi adjoints are zero there and every material
// coefficient read is a cell of the box.  All loads of a thread's planes are
// issued before its first store.
// Several shots on one model run in one launch: the table holds each shot's
// adjoint state, the shots of a cell are visited in table order, and the
// gradient sums of all shots are added once per cell (shot order).
// ---------------------------------------------------------------------------
constexpr int ADJ_LZ = 2;

template <typename T> struct AdjShot {
    Adj3<T> e, hb;  // adjoint E (in place) and H
    T *seg;         // E^n snapshots of the segment: [slot][3][nalloc]
    AdjPml<T> a;    // psi adjoints of this shot (old, new), shared grading
};

// CPML transpose of one derivative term v at a cell inside the CPML of the
// derivative's axis (profile pr, old psi adjoint p0): v k + a (p0 + v).
template <typename T> __device__ __forceinline__ T adj_term(T v, bool in, T p0, const Prof<T> &pr)
{
    if (!in) return v;
    const T p = p0 + v;
    return v * pr.k + pr.a * p;
}

template <typename T, int MAT, bool PML>
__device__ __forceinline__ void adj_h_column(const GridView &g, T ce, const AdjShot<T> &sh, const T *__restrict__ cb,
                                             int I, int J, int K0, int K1)
{
    constexpr int L = ADJ_LZ;
    const long P = g.pitch, S = g.slab;
    const AdjPml<T> &a = sh.a;
    const T *ex = sh.e.c[0], *ey = sh.e.c[1], *ez = sh.e.c[2];
    T *hx = sh.hb.c[0], *hy = sh.hb.c[1], *hz = sh.hb.c[2];
    const long ub0 = g.off + S * K0;
    const unsigned ui = (unsigned)(I + P * J) * sizeof(T);
    const unsigned uI = ui + (unsigned)sizeof(T), uJ = ui + (unsigned)(P * sizeof(T));
    const bool okI = I + 1 <= g.nx, okJ = J + 1 <= g.ny;
    // planes K0 .. K0+L of the cell (the last one only for the K+1 terms),
    // planes K0 .. K0+L-1 of the neighbours and of hbar
    T lx[L + 1], ly[L + 1], cbo[L + 1], lz[L], zI[L], yI[L], cI[L], zJ[L], xJ[L], cJ[L], hbx[L], hby[L], hbz[L];
#pragma unroll
    for (int k = 0; k <= L; ++k) {
        lx[k] = ly[k] = cbo[k] = (T)0;
        if (K0 + k <= g.nz + 1) {
            const long b = ub0 + S * k;
            lx[k] = lane_at(ex + b, ui);
            ly[k] = lane_at(ey + b, ui);
            if (MAT >= 1) cbo[k] = lane_at(cb + b, ui);
        }
    }
#pragma unroll
    for (int k = 0; k < L; ++k) {
        lz[k] = zI[k] = yI[k] = cI[k] = zJ[k] = xJ[k] = cJ[k] = hbx[k] = hby[k] = hbz[k] = (T)0;
        if (K0 + k < K1) {
            const long b = ub0 + S * k;
            lz[k] = lane_at(ez + b, ui);
            zI[k] = lane_at(ez + b, uI);
            yI[k] = lane_at(ey + b, uI);
            zJ[k] = lane_at(ez + b, uJ);
            xJ[k] = lane_at(ex + b, uJ);
            if (MAT >= 1) {
                cI[k] = lane_at(cb + b, uI);
                cJ[k] = lane_at(cb + b, uJ);
            }
            hbx[k] = lane_at(hx + b, ui);
            hby[k] = lane_at(hy + b, ui);
            hbz[k] = lane_at(hz + b, ui);
        }
    }
    bool px = false, py = false, pxn = false, pyn = false;
    Prof<T> prx{}, pry{}, prxn{}, pryn{};
    int qx = 0, qy = 0, qxn = 0, qyn = 0;
    if (PML) {
        px = in_pml(a.P, g.nx, I);
        py = in_pml(a.P, g.ny, J);
        pxn = okI && in_pml(a.P, g.nx, I + 1);
        pyn = okJ && in_pml(a.P, g.ny, J + 1);
        if (px) prx = a.ep[0][I];
        if (py) pry = a.ep[1][J];
        if (pxn) prxn = a.ep[0][I + 1];
        if (pyn) pryn = a.ep[1][J + 1];
    }
    // the z-derivative terms of plane K0 (tbar of Ey, d = 0, and Ex, d = 1)
    T tby0, tbx1;
    {
        const T mx = (MAT == 0 ? ce : cbo[0]) * lx[0], my = (MAT == 0 ? ce : cbo[0]) * ly[0];
        tby0 = my;
        tbx1 = -mx;
        if (PML && in_pml(a.P, g.nz, K0)) {
            const int qz = slab_z(g, a.P, a.W, I, J, K0);
            const Prof<T> prz = a.ep[2][K0];
            tby0 = adj_term(my, true, a.pe_old[2][1][qz], prz);
            tbx1 = adj_term(-mx, true, a.pe_old[2][0][qz], prz);
        }
    }
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const int K = K0 + k;
        if (K >= K1) break;
        const long b = ub0 + S * k;
        const T c0 = MAT == 0 ? ce : cbo[k], c1 = MAT == 0 ? ce : cbo[k + 1];
        const T mx = c0 * lx[k], my = c0 * ly[k], mz = c0 * lz[k];
        // K+1 terms (zero above the update range)
        T tby0n = (T)0, tbx1n = (T)0;
        if (K + 1 <= g.nz) {
            const T mx1 = c1 * lx[k + 1], my1 = c1 * ly[k + 1];
            tby0n = my1;
            tbx1n = -mx1;
            if (PML && in_pml(a.P, g.nz, K + 1)) {
                const int qz = slab_z(g, a.P, a.W, I, J, K + 1);
                const Prof<T> prz = a.ep[2][K + 1];
                tby0n = adj_term(my1, true, a.pe_old[2][1][qz], prz);
                tbx1n = adj_term(-mx1, true, a.pe_old[2][0][qz], prz);
            }
        }
        // x- and y-derivative terms of the cell
        T ox0 = (T)0, ox1 = (T)0, oy0 = (T)0, oy1 = (T)0;  // old psi: x (Ey, Ez), y (Ex, Ez)
        if (PML) {
            if (px) {
                qx = slab_x(g, a.P, a.W, I, J, K);
                ox0 = a.pe_old[0][0][qx];
                ox1 = a.pe_old[0][1][qx];
            }
            if (py) {
                qy = slab_y(g, a.P, a.W, I, J, K);
                oy0 = a.pe_old[1][0][qy];
                oy1 = a.pe_old[1][1][qy];
            }
        }
        const T tbz0 = adj_term(mz, PML && px, ox1, prx);
        const T tby1 = adj_term(-my, PML && px, ox0, prx);
        const T tbz1 = adj_term(-mz, PML && py, oy1, pry);
        const T tbx0 = adj_term(mx, PML && py, oy0, pry);
        // neighbours I+1 (Ez d = 0, Ey d = 1) and J+1 (Ez d = 1, Ex d = 0)
        T tbz0I = (T)0, tby1I = (T)0, tbz1J = (T)0, tbx0J = (T)0;
        if (okI) {
            const T cc = MAT == 0 ? ce : cI[k];
            const T mzI = cc * zI[k], myI = cc * yI[k];
            T o0 = (T)0, o1 = (T)0;
            if (PML && pxn) {
                qxn = slab_x(g, a.P, a.W, I + 1, J, K);
                o0 = a.pe_old[0][0][qxn];
                o1 = a.pe_old[0][1][qxn];
            }
            tbz0I = adj_term(mzI, PML && pxn, o1, prxn);
            tby1I = adj_term(-myI, PML && pxn, o0, prxn);
        }
        if (okJ) {
            const T cc = MAT == 0 ? ce : cJ[k];
            const T mzJ = cc * zJ[k], mxJ = cc * xJ[k];
            T o0 = (T)0, o1 = (T)0;
            if (PML && pyn) {
                qyn = slab_y(g, a.P, a.W, I, J + 1, K);
                o0 = a.pe_old[1][0][qyn];
                o1 = a.pe_old[1][1][qyn];
            }
            tbz1J = adj_term(-mzJ, PML && pyn, o1, pryn);
            tbx0J = adj_term(mxJ, PML && pyn, o0, pryn);
        }
        if (I >= 1) lane_at(hx + b, ui) = hbx[k] + ((tby0 - tby0n) + (tbz1 - tbz1J));
        if (J >= 1) lane_at(hy + b, ui) = hby[k] + ((tbz0 - tbz0I) + (tbx1 - tbx1n));
        if (K >= 1) lane_at(hz + b, ui) = hbz[k] + ((tbx0 - tbx0J) + (tby1 - tby1I));
        // new psi_E adjoints of the cell's updated components
        if (PML) {
            const bool vx = J >= 1 && K >= 1, vy = I >= 1 && K >= 1, vz = I >= 1 && J >= 1;
            if (px) {
                if (vy) a.pe_new[0][0][qx] = prx.b * (ox0 + -my);
                if (vz) a.pe_new[0][1][qx] = prx.b * (ox1 + mz);
            }
            if (py) {
                if (vx) a.pe_new[1][0][qy] = pry.b * (oy0 + mx);
                if (vz) a.pe_new[1][1][qy] = pry.b * (oy1 + -mz);
            }
            if (in_pml(a.P, g.nz, K)) {
                const int qz = slab_z(g, a.P, a.W, I, J, K);
                const Prof<T> prz = a.ep[2][K];
                if (vx) a.pe_new[2][0][qz] = prz.b * (a.pe_old[2][0][qz] + -mx);
                if (vy) a.pe_new[2][1][qz] = prz.b * (a.pe_old[2][1][qz] + my);
            }
        }
        tby0 = tby0n;
        tbx1 = tbx1n;
    }
}

template <typename T, int MAT, bool PML, bool GRAD>
__device__ __forceinline__ void adj_e_column(const GridView &g, T nch, const AdjShot<T> &sh, const T *__restrict__ ca,
                                             long nalloc, int slot, int I, int J, int K0, int K1,
                                             double (&tp)[ADJ_LZ], double (&tq)[ADJ_LZ])
{
    constexpr int L = ADJ_LZ;
    const long P = g.pitch, S = g.slab;
    const AdjPml<T> &a = sh.a;
    T *ex = sh.e.c[0], *ey = sh.e.c[1], *ez = sh.e.c[2];
    const T *hx = sh.hb.c[0], *hy = sh.hb.c[1], *hz = sh.hb.c[2];
    const T *e0x = sh.seg + (long)slot * 3 * nalloc, *e0y = e0x + nalloc, *e0z = e0y + nalloc;
    const T *e1x = e0z + nalloc, *e1y = e1x + nalloc, *e1z = e1y + nalloc;
    const long ub0 = g.off + S * K0;
    const unsigned ui = (unsigned)(I + P * J) * sizeof(T);
    const unsigned uIm = ui - (I >= 1 ? (unsigned)sizeof(T) : 0u);
    const unsigned uJm = ui - (J >= 1 ? (unsigned)(P * sizeof(T)) : 0u);
    // plane K0-1 of Hx, Hy (the K-1 terms of plane K0) and planes K0 .. K0+L-1
    T hxm = (T)0, hym = (T)0;
    if (K0 >= 1) {
        hxm = lane_at(hx + (ub0 - S), ui);
        hym = lane_at(hy + (ub0 - S), ui);
    }
    T hxo[L], hyo[L], hzo[L], zIm[L], yIm[L], zJm[L], xJm[L], lx[L], ly[L], lz[L], cav[L];
    T fx0[L], fy0[L], fz0[L], fx1[L], fy1[L], fz1[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
        hxo[k] = hyo[k] = hzo[k] = zIm[k] = yIm[k] = zJm[k] = xJm[k] = lx[k] = ly[k] = lz[k] = cav[k] = (T)0;
        fx0[k] = fy0[k] = fz0[k] = fx1[k] = fy1[k] = fz1[k] = (T)0;
        if (K0 + k < K1) {
            const long b = ub0 + S * k;
            hxo[k] = lane_at(hx + b, ui);
            hyo[k] = lane_at(hy + b, ui);
            hzo[k] = lane_at(hz + b, ui);
            zIm[k] = lane_at(hz + b, uIm);
            yIm[k] = lane_at(hy + b, uIm);
            zJm[k] = lane_at(hz + b, uJm);
            xJm[k] = lane_at(hx + b, uJm);
            lx[k] = lane_at(ex + b, ui);
            ly[k] = lane_at(ey + b, ui);
            lz[k] = lane_at(ez + b, ui);
            if (MAT == 2) cav[k] = lane_at(ca + b, ui);
            if (GRAD) {
                fx0[k] = lane_at(e0x + b, ui);
                fy0[k] = lane_at(e0y + b, ui);
                fz0[k] = lane_at(e0z + b, ui);
                fx1[k] = lane_at(e1x + b, ui);
                fy1[k] = lane_at(e1y + b, ui);
                fz1[k] = lane_at(e1z + b, ui);
            }
        }
    }
    bool px = false, py = false, pxm = false, pym = false;
    Prof<T> prx{}, pry{}, prxm{}, prym{};
    int qx = 0, qy = 0, qxm = 0, qym = 0;
    if (PML) {
        px = in_pml(a.P, g.nx, I);
        py = in_pml(a.P, g.ny, J);
        pxm = I >= 1 && in_pml(a.P, g.nx, I - 1);
        pym = J >= 1 && in_pml(a.P, g.ny, J - 1);
        if (px) prx = a.hp[0][I];
        if (py) pry = a.hp[1][J];
        if (pxm) prxm = a.hp[0][I - 1];
        if (pym) prym = a.hp[1][J - 1];
    }
    // the z-derivative terms of plane K0-1 (sbar of Hy, d = 0, and Hx, d = 1)
    T sby0m = (T)0, sbx1m = (T)0;
    if (K0 >= 1) {
        const T nux = nch * hxm, nuy = nch * hym;
        sby0m = nuy;
        sbx1m = -nux;
        if (PML && in_pml(a.P, g.nz, K0 - 1)) {
            const int qz = slab_z(g, a.P, a.W, I, J, K0 - 1);
            const Prof<T> prz = a.hp[2][K0 - 1];
            sby0m = adj_term(nuy, true, a.ph_old[2][1][qz], prz);
            sbx1m = adj_term(-nux, true, a.ph_old[2][0][qz], prz);
        }
    }
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const int K = K0 + k;
        if (K >= K1) break;
        const long b = ub0 + S * k;
        const T nux = nch * hxo[k], nuy = nch * hyo[k], nuz = nch * hzo[k];
        T oz0 = (T)0, oz1 = (T)0, ox0 = (T)0, ox1 = (T)0, oy0 = (T)0, oy1 = (T)0;
        bool pz = false;
        Prof<T> prz{};
        int qz = 0;
        if (PML) {
            pz = in_pml(a.P, g.nz, K);
            if (pz) {
                qz = slab_z(g, a.P, a.W, I, J, K);
                prz = a.hp[2][K];
                oz0 = a.ph_old[2][0][qz];
                oz1 = a.ph_old[2][1][qz];
            }
            if (px) {
                qx = slab_x(g, a.P, a.W, I, J, K);
                ox0 = a.ph_old[0][0][qx];
                ox1 = a.ph_old[0][1][qx];
            }
            if (py) {
                qy = slab_y(g, a.P, a.W, I, J, K);
                oy0 = a.ph_old[1][0][qy];
                oy1 = a.ph_old[1][1][qy];
            }
        }
        const T sby0 = adj_term(nuy, PML && pz, oz1, prz);
        const T sbx1 = adj_term(-nux, PML && pz, oz0, prz);
        const T sbz0 = adj_term(nuz, PML && px, ox1, prx);
        const T sby1 = adj_term(-nuy, PML && px, ox0, prx);
        const T sbz1 = adj_term(-nuz, PML && py, oy1, pry);
        const T sbx0 = adj_term(nux, PML && py, oy0, pry);
        // neighbours I-1 (Hz d = 0, Hy d = 1) and J-1 (Hz d = 1, Hx d = 0)
        T sbz0I = (T)0, sby1I = (T)0, sbz1J = (T)0, sbx0J = (T)0;
        if (I >= 1) {
            T o0 = (T)0, o1 = (T)0;
            if (PML && pxm) {
                qxm = slab_x(g, a.P, a.W, I - 1, J, K);
                o0 = a.ph_old[0][0][qxm];
                o1 = a.ph_old[0][1][qxm];
            }
            sbz0I = adj_term(nch * zIm[k], PML && pxm, o1, prxm);
            sby1I = adj_term(-(nch * yIm[k]), PML && pxm, o0, prxm);
        }
        if (J >= 1) {
            T o0 = (T)0, o1 = (T)0;
            if (PML && pym) {
                qym = slab_y(g, a.P, a.W, I, J - 1, K);
                o0 = a.ph_old[1][0][qym];
                o1 = a.ph_old[1][1][qym];
            }
            sbz1J = adj_term(-(nch * zJm[k]), PML && pym, o1, prym);
            sbx0J = adj_term(nch * xJm[k], PML && pym, o0, prym);
        }
        const bool vx = J >= 1 && K >= 1, vy = I >= 1 && K >= 1, vz = I >= 1 && J >= 1;
        if (GRAD) {
            // gradient sums from the adjoint e before its update
            double sp = 0.0, sq = 0.0;
            const double c = MAT == 2 ? (double)cav[k] : 1.0;
            if (vx) {
                const double l = (double)lx[k], f0 = (double)fx0[k];
                sp = sp + l * f0;
                sq = sq + l * ((double)fx1[k] - (MAT == 2 ? c * f0 : f0));
            }
            if (vy) {
                const double l = (double)ly[k], f0 = (double)fy0[k];
                sp = sp + l * f0;
                sq = sq + l * ((double)fy1[k] - (MAT == 2 ? c * f0 : f0));
            }
            if (vz) {
                const double l = (double)lz[k], f0 = (double)fz0[k];
                sp = sp + l * f0;
                sq = sq + l * ((double)fz1[k] - (MAT == 2 ? c * f0 : f0));
            }
            tp[k] = tp[k] + sp;
            tq[k] = tq[k] + sq;
        }
        if (vx) lane_at(ex + b, ui) = (MAT == 2 ? cav[k] * lx[k] : lx[k]) + ((sby0m - sby0) + (sbz1J - sbz1));
        if (vy) lane_at(ey + b, ui) = (MAT == 2 ? cav[k] * ly[k] : ly[k]) + ((sbx1m - sbx1) + (sbz0I - sbz0));
        if (vz) lane_at(ez + b, ui) = (MAT == 2 ? cav[k] * lz[k] : lz[k]) + ((sbx0J - sbx0) + (sby1I - sby1));
        // new psi_H adjoints of the cell's H components
        if (PML) {
            if (px) {
                if (J >= 1) a.ph_new[0][0][qx] = prx.b * (ox0 + -nuy);
                if (K >= 1) a.ph_new[0][1][qx] = prx.b * (ox1 + nuz);
            }
            if (py) {
                if (I >= 1) a.ph_new[1][0][qy] = pry.b * (oy0 + nux);
                if (K >= 1) a.ph_new[1][1][qy] = pry.b * (oy1 + -nuz);
            }
            if (pz) {
                if (I >= 1) a.ph_new[2][0][qz] = prz.b * (oz0 + -nux);
                if (J >= 1) a.ph_new[2][1][qz] = prz.b * (oz1 + nuy);
            }
        }
        sby0m = sby0;
        sbx1m = sbx1;
    }
}

template <typename T, int MAT, bool PML>
__global__ __launch_bounds__(256) void k_adj_h(GridView g, T ce, const AdjShot<T> *__restrict__ shots, int ns,
                                              const T *__restrict__ cb, int P, int npw, int nbx, int nby, int ntiles)
{
    int bx, by, bz;
    if (!split_tile(blockIdx.x, nbx, nby, ntiles, bx, by, bz)) return;
    const int I = split_column(g, P, npw, bx, threadIdx.x);
    const int J = __builtin_amdgcn_readfirstlane(by * blockDim.y + threadIdx.y);
    const int K0 = bz * ADJ_LZ, K1 = min(K0 + ADJ_LZ, g.nz + 1);
    if (I < 0 || I > g.nx || J > g.ny) return;
    for (int q = 0; q < ns; ++q) adj_h_column<T, MAT, PML>(g, ce, shots[q], cb, I, J, K0, K1);
}

// Reverse workgroup order, as k_e_split: it starts on the planes k_adj_h ended on.
template <typename T, int MAT, bool PML, bool GRAD>
__global__ __launch_bounds__(256) void k_adj_e(GridView g, T nch, const AdjShot<T> *__restrict__ shots, int ns,
                                              const T *__restrict__ ca, double *__restrict__ gp,
                                              double *__restrict__ gq, long nalloc, int slot, int P, int npw, int nbx,
                                              int nby, int ntiles)
{
    int bx, by, bz;
    if (!split_tile((int)(gridDim.x - 1u - blockIdx.x), nbx, nby, ntiles, bx, by, bz)) return;
    const int I = split_column(g, P, npw, bx, threadIdx.x);
    const int J = __builtin_amdgcn_readfirstlane(by * blockDim.y + threadIdx.y);
    const int K0 = bz * ADJ_LZ, K1 = min(K0 + ADJ_LZ, g.nz + 1);
    if (I < 0 || I > g.nx || J > g.ny) return;
    double tp[ADJ_LZ], tq[ADJ_LZ];
#pragma unroll
    for (int k = 0; k < ADJ_LZ; ++k) tp[k] = tq[k] = 0.0;
    for (int q = 0; q < ns; ++q) adj_e_column<T, MAT, PML, GRAD>(g, nch, shots[q], ca, nalloc, slot, I, J, K0, K1, tp, tq);
    if (GRAD) {
        const unsigned ud = (unsigned)(I + g.pitch * J) * sizeof(double);
#pragma unroll
        for (int k = 0; k < ADJ_LZ; ++k) {
            if (K0 + k >= K1) break;
            const long b = g.off + g.slab * (K0 + k);
            lane_at(gp + b, ud) = lane_at(gp + b, ud) + tp[k];
            lane_at(gq + b, ud) = lane_at(gq + b, ud) + tq[k];
        }
    }
}

// Gradient from the sums: g_eps = dCa/deps_r P + (dCb/deps_r)/Cb Q, likewise
// for sigma, on the material grid (P, Q on the unified grid).
__global__ void k_adj_grad(GridView g, const double *__restrict__ gp, const double *__restrict__ gq, GradView gv)
{
    const long nm = (long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const int I = (int)(m % (g.nx + 1));
    const long r = m / (g.nx + 1);
    const int J = (int)(r % (g.ny + 1)), K = (int)(r / (g.ny + 1));
    const long u = g.off + I + g.pitch * J + g.slab * K;
    const double p = gp[u], q = gq[u];
    gv.g_eps[m] = gv.dca_e[m] * p + gv.dcbn_e[m] * q;
    gv.g_sig[m] = gv.dca_s[m] * p + gv.dcbn_s[m] * q;
}

// E of up to kShotArgs shots into their snapshot slots: block row 3q + c
// copies component c of shot q (nalloc elements, a multiple of 32) to
// seg_q + (3 slot + c) nalloc, 16 bytes per thread and iteration.  The
// pointers travel as kernel arguments (the forward fields of some sweeps
// change buffers from step to step).
constexpr int kShotArgs = 64;

template <typename T> struct SnapArgs {
    const T *e[kShotArgs][3];
    T *seg[kShotArgs];
};

template <typename T>
__global__ __launch_bounds__(256) void k_snap(SnapArgs<T> a, long nalloc, int slot)
{
    const int q = blockIdx.y / 3, c = blockIdx.y % 3;
    const float4 *src = reinterpret_cast<const float4 *>(a.e[q][c]);
    float4 *dst = reinterpret_cast<float4 *>(a.seg[q] + ((long)slot * 3 + c) * nalloc);
    const long nv = nalloc * (long)sizeof(T) / 16;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Residual injection of up to kShotArgs shots at recorded row `row`: block q
// adds w_q[row][p] to the adjoint E of shot q at its points (serial: one
// thread, in the misfit's order, for point lists that repeat a point).
template <typename T> struct InjectArgs {
    T *e[kShotArgs][3];
    const int *pt_comp[kShotArgs];
    const long *pt_u[kShotArgs];
    const double *w[kShotArgs];
    int npts[kShotArgs], serial[kShotArgs];
};

template <typename T> __global__ __launch_bounds__(256) void k_adj_inject_shots(InjectArgs<T> a, int row)
{
    const int q = blockIdx.x, np = a.npts[q];
    if (np == 0) return;
    const double *w = a.w[q] + (long)row * np;
    if (a.serial[q]) {
        if (threadIdx.x != 0) return;
        for (int p = 0; p < np; ++p) {
            T *x = a.e[q][a.pt_comp[q][p]] + a.pt_u[q][p];
            *x = *x + (T)w[p];
        }
        return;
    }
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
        T *x = a.e[q][a.pt_comp[q][p]] + a.pt_u[q][p];
        *x = *x + (T)w[p];
    }
}

}  // namespace fdtd
}  // namespace ems
