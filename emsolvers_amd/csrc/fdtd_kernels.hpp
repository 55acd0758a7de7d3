// Yee FDTD device kernels for gfx950.
//
// Device layout ("unified grid"): all six components share one padded index
// space (I, J, K) in [0, nx+1] x [0, ny+1] x [0, nz+1],
//     u(I, J, K) = off + I + pitch*J + slab*K,   slab = pitch*(ny+2),
// with pitch a multiple of 16 elements, so that every row starts on a 64-byte
// (FP32) / 128-byte (FP64) boundary.  Reference arrays map into it as
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
template <typename T> struct PmlView {
    int P, W;
    const T *kh[3], *bh[3], *ah[3], *ke[3], *be[3], *ae[3];
    // psi_h[axis][0/1], psi_e[axis][0/1]:
    //   x: hy_x hz_x | ey_x ez_x    y: hx_y hz_y | ex_y ez_y    z: hx_z hy_z | ex_z ey_z
    T *psi_h[3][2];
    T *psi_e[3][2];
};

__device__ __forceinline__ bool in_pml(int P, int n, int I) { return P > 0 && (I <= P || I >= n - P + 1); }

__device__ __forceinline__ long slab_x(const GridView &g, int P, int W, int I, int J, int K)
{
    int l = I <= P ? I : I - g.nx + 2 * P;
    return l + (long)W * (J + (long)(g.ny + 2) * K);
}
__device__ __forceinline__ long slab_y(const GridView &g, int P, int W, int I, int J, int K)
{
    int l = J <= P ? J : J - g.ny + 2 * P;
    return I + (long)(g.nx + 2) * (l + (long)W * K);
}
__device__ __forceinline__ long slab_z(const GridView &g, int P, int W, int I, int J, int K)
{
    int l = K <= P ? K : K - g.nz + 2 * P;
    return I + (long)(g.nx + 2) * (J + (long)(g.ny + 2) * l);
}

// psi <- b*psi + a*t ; returns t/kappa + psi  (Roden & Gedney recursion)
template <typename T> __device__ __forceinline__ T cpml_term(T *psi, long q, T b, T a, T kinv, T t)
{
    T p = b * psi[q] + a * t;
    psi[q] = p;
    return t * kinv + p;
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
// Split kernels: one thread per (I, J, K), H sweep then E sweep, in place.
// Block 64 x 4, grid (ceil((nx+1)/64), ceil((ny+1)/4), nz+1).
// ---------------------------------------------------------------------------
template <typename T, bool PML>
__global__ __launch_bounds__(256) void k_h_split(GridView g, T ch, Fields<T> f, PmlView<T> pm)
{
    const int I = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y;
    const int K = blockIdx.z;
    if (I > g.nx || J > g.ny || K > g.nz) return;
    const long u = g.off + I + g.pitch * J + g.slab * K;
    bool px = false, py = false, pz = false;
    if (PML) {
        px = in_pml(pm.P, g.nx, I);
        py = in_pml(pm.P, g.ny, J);
        pz = in_pml(pm.P, g.nz, K);
    }
    if (I >= 1) {  // Hx: (Ez(J+1) - Ez(J)) - Ey(K+1) + Ey(K)
        T a = f.ez[u + g.pitch], b = f.ez[u], c = f.ey[u + g.slab], d = f.ey[u];
        T br;
        if (PML && (py || pz)) {
            T t1 = a - b, t2 = c - d;
            if (py) t1 = cpml_term(pm.psi_h[1][0], slab_y(g, pm.P, pm.W, I, J, K), pm.bh[1][J], pm.ah[1][J], pm.kh[1][J], t1);
            if (pz) t2 = cpml_term(pm.psi_h[2][0], slab_z(g, pm.P, pm.W, I, J, K), pm.bh[2][K], pm.ah[2][K], pm.kh[2][K], t2);
            br = t1 - t2;
        } else {
            br = a - b - c + d;
        }
        f.hx[u] -= ch * br;
    }
    if (J >= 1) {  // Hy: (Ex(K+1) - Ex(K)) - Ez(I+1) + Ez(I)
        T a = f.ex[u + g.slab], b = f.ex[u], c = f.ez[u + 1], d = f.ez[u];
        T br;
        if (PML && (pz || px)) {
            T t1 = a - b, t2 = c - d;
            if (pz) t1 = cpml_term(pm.psi_h[2][1], slab_z(g, pm.P, pm.W, I, J, K), pm.bh[2][K], pm.ah[2][K], pm.kh[2][K], t1);
            if (px) t2 = cpml_term(pm.psi_h[0][0], slab_x(g, pm.P, pm.W, I, J, K), pm.bh[0][I], pm.ah[0][I], pm.kh[0][I], t2);
            br = t1 - t2;
        } else {
            br = a - b - c + d;
        }
        f.hy[u] -= ch * br;
    }
    if (K >= 1) {  // Hz: (Ey(I+1) - Ey(I)) - Ex(J+1) + Ex(J)
        T a = f.ey[u + 1], b = f.ey[u], c = f.ex[u + g.pitch], d = f.ex[u];
        T br;
        if (PML && (px || py)) {
            T t1 = a - b, t2 = c - d;
            if (px) t1 = cpml_term(pm.psi_h[0][1], slab_x(g, pm.P, pm.W, I, J, K), pm.bh[0][I], pm.ah[0][I], pm.kh[0][I], t1);
            if (py) t2 = cpml_term(pm.psi_h[1][1], slab_y(g, pm.P, pm.W, I, J, K), pm.bh[1][J], pm.ah[1][J], pm.kh[1][J], t2);
            br = t1 - t2;
        } else {
            br = a - b - c + d;
        }
        f.hz[u] -= ch * br;
    }
}

template <typename T, int MAT, bool PML>
__global__ __launch_bounds__(256) void k_e_split(GridView g, T ce, Fields<T> f,
                                                 const T *__restrict__ ca, const T *__restrict__ cb,
                                                 PmlView<T> pm)
{
    const int I = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y;
    const int K = blockIdx.z;
    if (I > g.nx || J > g.ny || K > g.nz) return;
    const long u = g.off + I + g.pitch * J + g.slab * K;
    bool px = false, py = false, pz = false;
    if (PML) {
        px = in_pml(pm.P, g.nx, I);
        py = in_pml(pm.P, g.ny, J);
        pz = in_pml(pm.P, g.nz, K);
    }
    if (J >= 1 && K >= 1) {  // Ex: (Hz(J) - Hz(J-1)) - Hy(K) + Hy(K-1)
        T a = f.hz[u], b = f.hz[u - g.pitch], c = f.hy[u], d = f.hy[u - g.slab];
        T br;
        if (PML && (py || pz)) {
            T t1 = a - b, t2 = c - d;
            if (py) t1 = cpml_term(pm.psi_e[1][0], slab_y(g, pm.P, pm.W, I, J, K), pm.be[1][J], pm.ae[1][J], pm.ke[1][J], t1);
            if (pz) t2 = cpml_term(pm.psi_e[2][0], slab_z(g, pm.P, pm.W, I, J, K), pm.be[2][K], pm.ae[2][K], pm.ke[2][K], t2);
            br = t1 - t2;
        } else {
            br = a - b - c + d;
        }
        f.ex[u] = e_update<T, MAT>(f.ex[u], br, ce, ca, cb, u);
    }
    if (I >= 1 && K >= 1) {  // Ey: (Hx(K) - Hx(K-1)) - Hz(I) + Hz(I-1)
        T a = f.hx[u], b = f.hx[u - g.slab], c = f.hz[u], d = f.hz[u - 1];
        T br;
        if (PML && (pz || px)) {
            T t1 = a - b, t2 = c - d;
            if (pz) t1 = cpml_term(pm.psi_e[2][1], slab_z(g, pm.P, pm.W, I, J, K), pm.be[2][K], pm.ae[2][K], pm.ke[2][K], t1);
            if (px) t2 = cpml_term(pm.psi_e[0][0], slab_x(g, pm.P, pm.W, I, J, K), pm.be[0][I], pm.ae[0][I], pm.ke[0][I], t2);
            br = t1 - t2;
        } else {
            br = a - b - c + d;
        }
        f.ey[u] = e_update<T, MAT>(f.ey[u], br, ce, ca, cb, u);
    }
    if (I >= 1 && J >= 1) {  // Ez: (Hy(I) - Hy(I-1)) - Hx(J) + Hx(J-1)
        T a = f.hy[u], b = f.hy[u - 1], c = f.hx[u], d = f.hx[u - g.pitch];
        T br;
        if (PML && (px || py)) {
            T t1 = a - b, t2 = c - d;
            if (px) t1 = cpml_term(pm.psi_e[0][1], slab_x(g, pm.P, pm.W, I, J, K), pm.be[0][I], pm.ae[0][I], pm.ke[0][I], t1);
            if (py) t2 = cpml_term(pm.psi_e[1][1], slab_y(g, pm.P, pm.W, I, J, K), pm.be[1][J], pm.ae[1][J], pm.ke[1][J], t2);
            br = t1 - t2;
        } else {
            br = a - b - c + d;
        }
        f.ez[u] = e_update<T, MAT>(f.ez[u], br, ce, ca, cb, u);
    }
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
// field word is read once and written once per step.  A workgroup of
// 64 x BY threads owns a 63 x (BY-1) tile of (I, J) columns and marches up a
// chunk of lz planes in K.  E(K) needs H^{n+1/2} at I-1, J-1 and K-1, so the
// workgroup also computes H on a one-cell halo on the low side: lane 0 (I-1),
// row 0 (J-1) and the plane below the chunk (K0-1); these are recomputed from
// the old buffers and never stored.  H at I-1 / J-1 is exchanged through LDS
// (double-buffered by plane parity: one barrier per plane), H at K-1 and
// E^n at K+1 are carried in registers.  CPML psi for H is ping-ponged as
// well (halo threads read the old values); psi for E is updated in place by
// the owning thread.  Arithmetic per output is identical to the split
// kernels, so both give the same bits.
// ---------------------------------------------------------------------------
template <typename T> struct PsiH {
    const T *old_[3][2];
    T *new_[3][2];
};

template <typename T, int BY, int MAT, bool PML>
__global__ __launch_bounds__(64 * BY) void k_fused(GridView g, T ch, T ce, Fields<T> o, Fields<T> nw,
                                                    const T *__restrict__ ca, const T *__restrict__ cb,
                                                    PmlView<T> pm, PsiH<T> ph, int lz, int ntx, int nty,
                                                    int nblk)
{
    __shared__ T sH[2][3][BY][64];
    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs;
    // give each XCD a contiguous range of tiles so halos are shared in its L2.
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lb >= nblk) return;  // uniform for the whole workgroup
    const int ti = lb % ntx;
    const int tj = (lb / ntx) % nty;
    const int kc = lb / (ntx * nty);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int I = ti * 63 - 1 + tx;
    const int J = tj * (BY - 1) - 1 + ty;
    const int K0 = kc * lz;
    const int K1 = min(K0 + lz, g.nz + 1);
    const bool act = I >= 0 && J >= 0 && I <= g.nx && J <= g.ny;
    const bool own = act && tx >= 1 && ty >= 1;
    const int Kst = K0 > 0 ? K0 - 1 : 0;
    long u = g.off + (long)(act ? I : 0) + g.pitch * (act ? J : 0) + g.slab * Kst;
    bool px = false, py = false;
    if (PML) {
        px = in_pml(pm.P, g.nx, I);
        py = in_pml(pm.P, g.ny, J);
    }
    T ex_k = 0, ey_k = 0, hx_km1 = 0, hy_km1 = 0;
    if (act) {
        ex_k = o.ex[u];
        ey_k = o.ey[u];
    }
    for (int K = Kst; K < K1; ++K) {
        const int par = K & 1;
        const bool wr = own && K >= K0;
        T hxn = 0, hyn = 0, hzn = 0, ez_k = 0, ex_k1 = 0, ey_k1 = 0;
        const bool pz = PML && in_pml(pm.P, g.nz, K);
        if (act) {
            ex_k1 = o.ex[u + g.slab];
            ey_k1 = o.ey[u + g.slab];
            ez_k = o.ez[u];
            const T ez_i1 = o.ez[u + 1], ey_i1 = o.ey[u + 1];
            const T ez_j1 = o.ez[u + g.pitch], ex_j1 = o.ex[u + g.pitch];
            const T hx = o.hx[u], hy = o.hy[u], hz = o.hz[u];
            // Hx: (Ez(J+1) - Ez(J)) - Ey(K+1) + Ey(K)
            if (PML && (py || pz)) {
                T t1 = ez_j1 - ez_k, t2 = ey_k1 - ey_k;
                if (py) {
                    long q = slab_y(g, pm.P, pm.W, I, J, K);
                    T p = pm.bh[1][J] * ph.old_[1][0][q] + pm.ah[1][J] * t1;
                    if (wr) ph.new_[1][0][q] = p;
                    t1 = t1 * pm.kh[1][J] + p;
                }
                if (pz) {
                    long q = slab_z(g, pm.P, pm.W, I, J, K);
                    T p = pm.bh[2][K] * ph.old_[2][0][q] + pm.ah[2][K] * t2;
                    if (wr) ph.new_[2][0][q] = p;
                    t2 = t2 * pm.kh[2][K] + p;
                }
                hxn = hx - ch * (t1 - t2);
            } else {
                hxn = hx - ch * (ez_j1 - ez_k - ey_k1 + ey_k);
            }
            // Hy: (Ex(K+1) - Ex(K)) - Ez(I+1) + Ez(I)
            if (PML && (pz || px)) {
                T t1 = ex_k1 - ex_k, t2 = ez_i1 - ez_k;
                if (pz) {
                    long q = slab_z(g, pm.P, pm.W, I, J, K);
                    T p = pm.bh[2][K] * ph.old_[2][1][q] + pm.ah[2][K] * t1;
                    if (wr) ph.new_[2][1][q] = p;
                    t1 = t1 * pm.kh[2][K] + p;
                }
                if (px) {
                    long q = slab_x(g, pm.P, pm.W, I, J, K);
                    T p = pm.bh[0][I] * ph.old_[0][0][q] + pm.ah[0][I] * t2;
                    if (wr) ph.new_[0][0][q] = p;
                    t2 = t2 * pm.kh[0][I] + p;
                }
                hyn = hy - ch * (t1 - t2);
            } else {
                hyn = hy - ch * (ex_k1 - ex_k - ez_i1 + ez_k);
            }
            // Hz: (Ey(I+1) - Ey(I)) - Ex(J+1) + Ex(J)
            if (PML && (px || py)) {
                T t1 = ey_i1 - ey_k, t2 = ex_j1 - ex_k;
                if (px) {
                    long q = slab_x(g, pm.P, pm.W, I, J, K);
                    T p = pm.bh[0][I] * ph.old_[0][1][q] + pm.ah[0][I] * t1;
                    if (wr) ph.new_[0][1][q] = p;
                    t1 = t1 * pm.kh[0][I] + p;
                }
                if (py) {
                    long q = slab_y(g, pm.P, pm.W, I, J, K);
                    T p = pm.bh[1][J] * ph.old_[1][1][q] + pm.ah[1][J] * t2;
                    if (wr) ph.new_[1][1][q] = p;
                    t2 = t2 * pm.kh[1][J] + p;
                }
                hzn = hz - ch * (t1 - t2);
            } else {
                hzn = hz - ch * (ey_i1 - ey_k - ex_j1 + ex_k);
            }
        }
        sH[par][0][ty][tx] = hxn;
        sH[par][1][ty][tx] = hyn;
        sH[par][2][ty][tx] = hzn;
        __syncthreads();
        if (wr) {
            if (I >= 1) nw.hx[u] = hxn;
            if (J >= 1) nw.hy[u] = hyn;
            if (K >= 1) nw.hz[u] = hzn;
            const T hz_im1 = sH[par][2][ty][tx - 1], hy_im1 = sH[par][1][ty][tx - 1];
            const T hz_jm1 = sH[par][2][ty - 1][tx], hx_jm1 = sH[par][0][ty - 1][tx];
            if (J >= 1 && K >= 1) {  // Ex: (Hz(J) - Hz(J-1)) - Hy(K) + Hy(K-1)
                T br;
                if (PML && (py || pz)) {
                    T t1 = hzn - hz_jm1, t2 = hyn - hy_km1;
                    if (py) t1 = cpml_term(pm.psi_e[1][0], slab_y(g, pm.P, pm.W, I, J, K), pm.be[1][J], pm.ae[1][J], pm.ke[1][J], t1);
                    if (pz) t2 = cpml_term(pm.psi_e[2][0], slab_z(g, pm.P, pm.W, I, J, K), pm.be[2][K], pm.ae[2][K], pm.ke[2][K], t2);
                    br = t1 - t2;
                } else {
                    br = hzn - hz_jm1 - hyn + hy_km1;
                }
                nw.ex[u] = e_update<T, MAT>(ex_k, br, ce, ca, cb, u);
            }
            if (I >= 1 && K >= 1) {  // Ey: (Hx(K) - Hx(K-1)) - Hz(I) + Hz(I-1)
                T br;
                if (PML && (pz || px)) {
                    T t1 = hxn - hx_km1, t2 = hzn - hz_im1;
                    if (pz) t1 = cpml_term(pm.psi_e[2][1], slab_z(g, pm.P, pm.W, I, J, K), pm.be[2][K], pm.ae[2][K], pm.ke[2][K], t1);
                    if (px) t2 = cpml_term(pm.psi_e[0][0], slab_x(g, pm.P, pm.W, I, J, K), pm.be[0][I], pm.ae[0][I], pm.ke[0][I], t2);
                    br = t1 - t2;
                } else {
                    br = hxn - hx_km1 - hzn + hz_im1;
                }
                nw.ey[u] = e_update<T, MAT>(ey_k, br, ce, ca, cb, u);
            }
            if (I >= 1 && J >= 1) {  // Ez: (Hy(I) - Hy(I-1)) - Hx(J) + Hx(J-1)
                T br;
                if (PML && (px || py)) {
                    T t1 = hyn - hy_im1, t2 = hxn - hx_jm1;
                    if (px) t1 = cpml_term(pm.psi_e[0][1], slab_x(g, pm.P, pm.W, I, J, K), pm.be[0][I], pm.ae[0][I], pm.ke[0][I], t1);
                    if (py) t2 = cpml_term(pm.psi_e[1][1], slab_y(g, pm.P, pm.W, I, J, K), pm.be[1][J], pm.ae[1][J], pm.ke[1][J], t2);
                    br = t1 - t2;
                } else {
                    br = hyn - hy_im1 - hxn + hx_jm1;
                }
                nw.ez[u] = e_update<T, MAT>(ez_k, br, ce, ca, cb, u);
            }
        }
        ex_k = ex_k1;
        ey_k = ey_k1;
        hx_km1 = hxn;
        hy_km1 = hyn;
        u += g.slab;
    }
}

}  // namespace fdtd
}  // namespace ems
