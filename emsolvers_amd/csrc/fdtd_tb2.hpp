// Two-step temporally blocked Yee sweep for gfx950 (EMS_KERNEL_TB2).
//
// One launch advances the fields by two full time steps:
//     H^{n+1/2}, E^{n+1}  (level 1)   then   H^{n+3/2}, E^{n+2}  (level 2),
// reading E^n, H^{n-1/2} (and the CPML psi of step n) from the "old" buffers
// and writing only the level-2 fields and psi into the "new" ones (ping-pong).
// Level-1 values never leave the chip, so a field word crosses HBM about once
// in each direction per two steps instead of per step.
//
// Decomposition.  A workgroup owns a tile of txw x tyw (I, J) columns (one
// thread per column, threads numbered x-fastest) over a chunk [K0, K1) of
// planes, and marches K upward.  Neighbours in I and J come from LDS; the
// dependency cone of two steps needs a halo of two columns on the low side
// and one on the high side, so the tile's core (where level-2 values are
// stored) is tx in [2, txw-2], ty in [2, tyw-2] and tiles step by txw-3,
// tyw-3.  Halo columns recompute level 1 (and level-2 H) from the old buffers;
// they never store.  Validity shrinks level by level:
//     H1 everywhere (E^n at I+1, J+1 comes from global memory),
//     E1 for tx, ty >= 1,  H2 for 1 <= tx, ty <= max-1,  E2 for 2 <= tx, ty <= max-1,
// and invalid values are never stored nor fed into valid ones.  With XL (64
// columns: one wavefront per tile row) E^n at I+1 is the next lane's value
// (DPP), two loads fewer per plane; lane 63 then has no valid H1, so the
// high halo is two columns: core tx in [2, 61], tiles step by 60.
//
// K march.  At plane K a thread computes H1(K), E1(K) and, lagging behind so
// that one barrier per plane suffices, H2(K-1) and E2(K-2).  The LDS holds,
// per plane parity, the H1, E1 and H2 planes of every thread (slots sH1, sE1,
// sH2); besides exchanging neighbours they keep each thread's own values of
// the last two planes, so registers carry only E^n and the loads in flight:
//     a. H1(K) from E^n(K), E^n(K+1) and E^n(K) at I+1, J+1      -> sH1[K&1]
//     b. barrier
//     e. E2(K-2) from sE1[K&1] (E1(K-2), own), sH2[K&1] (H2(K-2), own and at
//        I-1, J-1), sH2[(K-1)&1] (H2(K-3), own) (+ sources of step n+2)
//     c. E1(K) from H1(K), sH1[(K-1)&1] (H1(K-1), own) and sH1[K&1] at I-1, J-1
//        (+ sources of step n+1)                                  -> sE1[K&1]
//     d. H2(K-1) from sH1[(K-1)&1], sE1[(K-1)&1] (own and at I+1, J+1) and
//        E1(K)                                                    -> sH2[(K-1)&1]
// A slot is overwritten only after the barrier that follows its last read by
// other threads, and after its own thread has read it.  Level-1 planes run over
// [K0-2, K1], level-2 H over [K0-1, K1-1] and level-2 E over [K0, K1-1].
//
// CPML.  psi is private to its cell, but halo columns of neighbouring tiles
// recompute the same cells, so both the H- and the E-psi are ping-ponged:
// level 1 reads the old psi, keeps its result in registers (H-psi for one
// plane, E-psi for two) and level 2 stores the new psi from the core.
//
// Arithmetic per output is that of the split sweeps (same brackets, same
// order, no contraction) and the sources and receivers are those of k_post,
// so two steps of this kernel equal two split steps bit for bit.
#pragma once

#include <type_traits>

#include "fdtd_kernels.hpp"

namespace ems {
namespace fdtd {

constexpr int kTbMaxThreads = 1024;

// Threads per workgroup for a field type: the nine double-buffered exchange
// planes (dynamic LDS, 18 words per thread) take at most 72 KiB in either
// precision, so two workgroups fit the 160 KiB of LDS of a CU.
template <typename T> constexpr int tb_max_threads() { return sizeof(T) == 4 ? 1024 : 512; }

__device__ __forceinline__ long uidx_dev(const GridView &g, int I, int J, int K)
{
    return g.off + I + g.pitch * J + g.slab * (long)K;
}

// Psi of both field types, old (read) and new (written).
template <typename T> struct PsiSet {
    const T *oh[3][2], *oe[3][2];
    T *nh[3][2], *ne[3][2];
};

// A box of cells whose level-2 values one launch stores: cores [i0, i0+ni) x
// [j0, j0+nj) x [k0, k0+nk), tiled by ntx x nty tiles and nkc chunks.
struct TbBox {
    int i0, j0, k0, ni, nj, nk;
    int ntx, nty, nkc;
    int count;  // ntx * nty * nkc
};

// One launch: up to two boxes with one tile shape (the second box's tiles
// follow the first's).
struct TbTiles {
    int txw, tyw;        // thread tile (columns along I, rows along J)
    int lz;              // planes per chunk
    int count;           // tiles in this launch
    int per;             // tiles per XCD band
    TbBox box[2];
};

// Sources and receivers of the two steps, sorted by plane (CSR over K).
struct TbPost {
    const int *s_kptr;   // [nz+2]: sources of plane K are [s_kptr[K], s_kptr[K+1])
    const int *s_comp;
    const long *s_u;
    const int *s_idx;    // index into the per-step value rows
    const double *v1, *v2;
    const int *p_kptr;   // receiver points, as the sources
    const int *p_comp;
    const long *p_u;
    const int *p_idx;    // column of the raw sample row
    double *rec1, *rec2; // raw sample rows of steps n+1, n+2 (nullptr: no record)
    int s_k0, s_k1;      // planes holding sources (s_k1 < s_k0: none)
    int p_k0, p_k1;      // planes holding receiver points
};

// One H update of a cell (the split sweep's expressions).  psi: in the old
// values of the cell's H-psi, out the new ones, order
//   [0] hx_y [1] hx_z [2] hy_z [3] hy_x [4] hz_x [5] hz_y
template <typename T, bool PML>
__device__ __forceinline__ void tb_h(T &hx, T &hy, T &hz, T ez_j1, T ez_k, T ey_k1, T ey_k, T ex_k1, T ex_k,
                                     T ez_i1, T ey_i1, T ex_j1, T ch, bool ux, bool uy, bool uz, bool px, bool py,
                                     bool pz, const Prof<T> &prx, const Prof<T> &pry, const Prof<T> &prz,
                                     T (&psi)[6])
{
    if (!PML) {
        // branch-free: masks as selects (the masks differ between lanes only
        // at the low faces of the grid)
        const T hxn = hx - ch * (ez_j1 - ez_k - ey_k1 + ey_k);
        const T hyn = hy - ch * (ex_k1 - ex_k - ez_i1 + ez_k);
        const T hzn = hz - ch * (ey_i1 - ey_k - ex_j1 + ex_k);
        hx = ux ? hxn : hx;
        hy = uy ? hyn : hy;
        hz = uz ? hzn : hz;
        return;
    }
    if (ux) {  // Hx: (Ez(J+1) - Ez(J)) - Ey(K+1) + Ey(K)
        if (py || pz) {
            T t1 = ez_j1 - ez_k, t2 = ey_k1 - ey_k;
            if (py) {
                const T p = pry.b * psi[0] + pry.a * t1;
                psi[0] = p;
                t1 = t1 * pry.k + p;
            }
            if (pz) {
                const T p = prz.b * psi[1] + prz.a * t2;
                psi[1] = p;
                t2 = t2 * prz.k + p;
            }
            hx = hx - ch * (t1 - t2);
        } else {
            hx = hx - ch * (ez_j1 - ez_k - ey_k1 + ey_k);
        }
    }
    if (uy) {  // Hy: (Ex(K+1) - Ex(K)) - Ez(I+1) + Ez(I)
        if (pz || px) {
            T t1 = ex_k1 - ex_k, t2 = ez_i1 - ez_k;
            if (pz) {
                const T p = prz.b * psi[2] + prz.a * t1;
                psi[2] = p;
                t1 = t1 * prz.k + p;
            }
            if (px) {
                const T p = prx.b * psi[3] + prx.a * t2;
                psi[3] = p;
                t2 = t2 * prx.k + p;
            }
            hy = hy - ch * (t1 - t2);
        } else {
            hy = hy - ch * (ex_k1 - ex_k - ez_i1 + ez_k);
        }
    }
    if (uz) {  // Hz: (Ey(I+1) - Ey(I)) - Ex(J+1) + Ex(J)
        if (px || py) {
            T t1 = ey_i1 - ey_k, t2 = ex_j1 - ex_k;
            if (px) {
                const T p = prx.b * psi[4] + prx.a * t1;
                psi[4] = p;
                t1 = t1 * prx.k + p;
            }
            if (py) {
                const T p = pry.b * psi[5] + pry.a * t2;
                psi[5] = p;
                t2 = t2 * pry.k + p;
            }
            hz = hz - ch * (t1 - t2);
        } else {
            hz = hz - ch * (ey_i1 - ey_k - ex_j1 + ex_k);
        }
    }
}

// One E update of a cell.  psi order
//   [0] ex_y [1] ex_z [2] ey_z [3] ey_x [4] ez_x [5] ez_y
template <typename T, int MAT, bool PML>
__device__ __forceinline__ void tb_e(T &ex, T &ey, T &ez, T hx, T hy, T hz, T hx_km1, T hy_km1, T hz_im1, T hy_im1,
                                     T hz_jm1, T hx_jm1, T ce, T cbv, T cav, bool ux, bool uy, bool uz, bool px,
                                     bool py, bool pz, const Prof<T> &prx, const Prof<T> &pry, const Prof<T> &prz,
                                     T (&psi)[6])
{
    auto upd = [&](T e, T br) { return MAT == 0 ? e + ce * br : MAT == 1 ? e + cbv * br : cav * e + cbv * br; };
    if (!PML) {
        const T exn = upd(ex, hz - hz_jm1 - hy + hy_km1);
        const T eyn = upd(ey, hx - hx_km1 - hz + hz_im1);
        const T ezn = upd(ez, hy - hy_im1 - hx + hx_jm1);
        ex = ux ? exn : ex;
        ey = uy ? eyn : ey;
        ez = uz ? ezn : ez;
        return;
    }
    if (ux) {  // Ex: (Hz(J) - Hz(J-1)) - Hy(K) + Hy(K-1)
        T br;
        if (py || pz) {
            T t1 = hz - hz_jm1, t2 = hy - hy_km1;
            if (py) {
                const T p = pry.b * psi[0] + pry.a * t1;
                psi[0] = p;
                t1 = t1 * pry.k + p;
            }
            if (pz) {
                const T p = prz.b * psi[1] + prz.a * t2;
                psi[1] = p;
                t2 = t2 * prz.k + p;
            }
            br = t1 - t2;
        } else {
            br = hz - hz_jm1 - hy + hy_km1;
        }
        ex = upd(ex, br);
    }
    if (uy) {  // Ey: (Hx(K) - Hx(K-1)) - Hz(I) + Hz(I-1)
        T br;
        if (pz || px) {
            T t1 = hx - hx_km1, t2 = hz - hz_im1;
            if (pz) {
                const T p = prz.b * psi[2] + prz.a * t1;
                psi[2] = p;
                t1 = t1 * prz.k + p;
            }
            if (px) {
                const T p = prx.b * psi[3] + prx.a * t2;
                psi[3] = p;
                t2 = t2 * prx.k + p;
            }
            br = t1 - t2;
        } else {
            br = hx - hx_km1 - hz + hz_im1;
        }
        ey = upd(ey, br);
    }
    if (uz) {  // Ez: (Hy(I) - Hy(I-1)) - Hx(J) + Hx(J-1)
        T br;
        if (px || py) {
            T t1 = hy - hy_im1, t2 = hx - hx_jm1;
            if (px) {
                const T p = prx.b * psi[4] + prx.a * t1;
                psi[4] = p;
                t1 = t1 * prx.k + p;
            }
            if (py) {
                const T p = pry.b * psi[5] + pry.a * t2;
                psi[5] = p;
                t2 = t2 * pry.k + p;
            }
            br = t1 - t2;
        } else {
            br = hy - hy_im1 - hx + hx_jm1;
        }
        ez = upd(ez, br);
    }
}

// psi of one cell from the slabs (old buffers), in the order of tb_h / tb_e:
//   ps[axis][0/1] = x: (y-comp, z-comp) | y: (x-comp, z-comp) | z: (x-comp, y-comp)
template <typename T>
__device__ __forceinline__ void tb_psi_load(const GridView &g, int P, int W, const T *const (&ps)[3][2], int I, int J,
                                            int K, bool px, bool py, bool pz, T (&v)[6])
{
    if (py) {
        const int q = slab_y(g, P, W, I, J, K);
        v[0] = ps[1][0][q];
        v[5] = ps[1][1][q];
    }
    if (pz) {
        const int q = slab_z(g, P, W, I, J, K);
        v[1] = ps[2][0][q];
        v[2] = ps[2][1][q];
    }
    if (px) {
        const int q = slab_x(g, P, W, I, J, K);
        v[3] = ps[0][0][q];
        v[4] = ps[0][1][q];
    }
}

template <typename T>
__device__ __forceinline__ void tb_psi_store(const GridView &g, int P, int W, T *const (&ps)[3][2], int I, int J,
                                             int K, bool px, bool py, bool pz, const T (&v)[6])
{
    if (py) {
        const int q = slab_y(g, P, W, I, J, K);
        ps[1][0][q] = v[0];
        ps[1][1][q] = v[5];
    }
    if (pz) {
        const int q = slab_z(g, P, W, I, J, K);
        ps[2][0][q] = v[1];
        ps[2][1][q] = v[2];
    }
    if (px) {
        const int q = slab_x(g, P, W, I, J, K);
        ps[0][0][q] = v[3];
        ps[0][1][q] = v[4];
    }
}

// Sources of one step on the E values of the cell at unified index u.
template <typename T>
__device__ __forceinline__ void tb_sources(const TbPost &pp, const double *v, int K, long u, T &ex, T &ey, T &ez)
{
    for (int q = pp.s_kptr[K]; q < pp.s_kptr[K + 1]; ++q) {
        if (pp.s_u[q] != u) continue;
        const T d = (T)v[pp.s_idx[q]];
        const int c = pp.s_comp[q];
        // selects, not an indexed array (which would live in scratch memory)
        ex = c == 0 ? ex - d : ex;
        ey = c == 1 ? ey - d : ey;
        ez = c == 2 ? ez - d : ez;
    }
}

template <typename T>
__device__ __forceinline__ void tb_receivers(const TbPost &pp, double *rec, int K, long u, T ex, T ey, T ez)
{
    for (int q = pp.p_kptr[K]; q < pp.p_kptr[K + 1]; ++q) {
        if (pp.p_u[q] != u) continue;
        const int c = pp.p_comp[q];
        T e = ez;
        e = c == 1 ? ey : e;
        e = c == 0 ? ex : e;
        rec[pp.p_idx[q]] = (double)e;
    }
}

// Everything one plane K of the march reads from global memory: the own
// E^n of planes K and K+1, E^n at I+1 / J+1, H^{n-1/2}, the material of
// plane K (for E1) and of plane K-2 (for E2).
template <typename T> struct TbSet {
    T ex, ey, ez, ex1, ey1;
    T ez_i1, ey_i1, ez_j1, ex_j1;
    T hx, hy, hz;
    T cb, ca, cb2, ca2;
};

// Array bases (element (I, J, K) = base + (I + pitch J + slab K)) and the
// byte strides; a whole array is under 4 GiB (checked on the host), so a
// lane addresses any element by a 32-bit byte offset from a wave-uniform base.
template <typename T> struct TbArrays {
    const T *ex, *ey, *ez, *hx, *hy, *hz, *ca, *cb;
    T *nex, *ney, *nez, *nhx, *nhy, *nhz;
    unsigned di, dj, dk;  // bytes per step in I, J, K
};

// Loads of the two-step sweep are unconditional (clamped addresses): a load
// under a per-lane condition merged with a default makes the compiler wait
// for it where it is issued, which would serialise the march.
// ui, ui_i1, ui_j1: the lane's cell and its I+1 / J+1 neighbours in a plane,
// clamped to the allocated layer n+1.
template <typename T, int MAT, bool XL>
__device__ __forceinline__ void tb_load(const TbArrays<T> &a, int nz, unsigned ui, unsigned ui_i1, unsigned ui_j1,
                                        int K, TbSet<T> &v)
{
    // every plane index clamped to the allocated planes 0 .. nz+1 (planes
    // beyond the grid read the ghost plane nz+1; their values are not used)
    const int Kc = min(max(K, 0), nz + 1);
    const int K1c = min(max(K + 1, 0), nz + 1);
    const int K2c = min(max(K - 2, 0), nz + 1);
    const unsigned pk = (unsigned)Kc * a.dk;
    const unsigned u0 = ui + pk, u1 = ui + (unsigned)K1c * a.dk, u2 = ui + (unsigned)K2c * a.dk;
    v.ex = lane_at(a.ex, u0);
    v.ey = lane_at(a.ey, u0);
    v.ez = lane_at(a.ez, u0);
    v.ex1 = lane_at(a.ex, u1);
    v.ey1 = lane_at(a.ey, u1);
    if (XL) {  // from the next lane's ez, ey
        v.ez_i1 = v.ey_i1 = (T)0;
    } else {
        v.ez_i1 = lane_at(a.ez, ui_i1 + pk);
        v.ey_i1 = lane_at(a.ey, ui_i1 + pk);
    }
    v.ez_j1 = lane_at(a.ez, ui_j1 + pk);
    v.ex_j1 = lane_at(a.ex, ui_j1 + pk);
    v.hx = lane_at(a.hx, u0);
    v.hy = lane_at(a.hy, u0);
    v.hz = lane_at(a.hz, u0);
    v.cb = v.ca = v.cb2 = v.ca2 = (T)0;
    if (MAT >= 1) {
        v.cb = lane_at(a.cb, u0);
        v.cb2 = lane_at(a.cb, u2);
    }
    if (MAT == 2) {
        v.ca = lane_at(a.ca, u0);
        v.ca2 = lane_at(a.ca, u2);
    }
}

template <int V> using tb_par = std::integral_constant<int, V>;

// LDS: nine exchange planes (H1, E1, H2 x 3 components) per plane parity,
// MAXT entries each (compile-time offsets, so that one address register
// serves most slots), after a pad of txw + 1 entries that the out-of-tile
// neighbours of halo lanes may read; passed as dynamic shared memory.
// PM: CPML axes the launch's tiles can meet (bit 0 x, bit 1 y, bit 2 z; 0:
// none), so that the instantiations for the CPML-free interior and for the
// slabs of the CPML shell carry only the psi of their own axes.
// MAXT: the largest workgroup the launch uses (1024 or 512), which sets the
// register budget (128 or 256 VGPRs).
template <typename T, int MAT, int PM, int MAXT, bool XL>
__global__ __launch_bounds__(MAXT) void k_tb2(GridView g, T ch, T ce, TbArrays<T> a, PmlView<T> pm, PsiSet<T> ps,
                                              TbTiles tt, TbPost pp)
{
    constexpr bool PML = PM != 0;
    extern __shared__ __align__(16) unsigned char tb_lds[];
    constexpr int NTP = MAXT;
    T *const L = reinterpret_cast<T *>(tb_lds) + tt.txw + 1;
    // XCD bands: workgroup b runs on XCD b % 8, which takes the contiguous
    // tile range [c per, (c+1) per), so that halo rows stay in its L2
    const int lb = (blockIdx.x & 7) * tt.per + (blockIdx.x >> 3);
    if (lb >= tt.count) return;  // uniform for the workgroup
    const bool second = lb >= tt.box[0].count;
    const TbBox &bx = tt.box[second ? 1 : 0];
    const int tile = lb - (second ? tt.box[0].count : 0);
    const int ti = tile % bx.ntx;
    const int tj = (tile / bx.ntx) % bx.nty;
    const int kc = tile / (bx.ntx * bx.nty);
    const int t = threadIdx.x;
    const int NT = tt.txw * tt.tyw;
    const bool tv = t < NT;
    const int tx = tv ? t % tt.txw : 0, ty = tv ? t / tt.txw : 0;
    // XL (a tile row is one wavefront): E^n at I+1 comes from the next lane,
    // so lane 63 has no valid H1 and the high halo is two columns
    constexpr int HX = XL ? 4 : 3;
    const int I = bx.i0 + ti * (tt.txw - HX) - 2 + tx;
    const int J = bx.j0 + tj * (tt.tyw - 3) - 2 + ty;
    const bool inb = tv && I >= 0 && I <= g.nx && J >= 0 && J <= g.ny;
    // the ghost layer I = nx+1 / J = ny+1 holds values (never updated) that
    // the cells I = nx / J = ny read: its threads pass their old values on
    const bool ing = tv && I >= 0 && I <= g.nx + 1 && J >= 0 && J <= g.ny + 1;
    // stored: the box's cells of the update box and of its ghost layer n+1
    const bool core = ing && tx >= 2 && tx <= tt.txw - (HX - 1) && ty >= 2 && ty <= tt.tyw - 2 && I < bx.i0 + bx.ni &&
                      J < bx.j0 + bx.nj;
    // LDS neighbours (out-of-tile ones, in the pads or other slots, only feed
    // halo lanes, whose values are invalid anyway)
    const int tim = t - 1, tip = t + 1, tjm = t - tt.txw, tjp = t + tt.txw;
    const int K0 = bx.k0 + kc * tt.lz;
    const int K1 = min(K0 + tt.lz, bx.k0 + bx.nk);
    const int Ks = max(K0 - 2, 0);
    // update masks of the column (no update outside the update box)
    const bool cI = inb && I >= 1, cJ = inb && J >= 1, c0 = inb;
    // threads outside the update box read a clamped cell: on the ghost layer
    // their own cell, whose values pass unchanged through both levels; further
    // out the values enter the exchange as zeros.  Nothing outside the box is
    // stored.
    const int Ic = min(max(I, 0), g.nx + 1), Jc = min(max(J, 0), g.ny + 1);
    const unsigned ui = (unsigned)Ic * a.di + (unsigned)Jc * a.dj;
    const unsigned ui_i1 = (unsigned)min(Ic + 1, g.nx + 1) * a.di + (unsigned)Jc * a.dj;
    const unsigned ui_j1 = (unsigned)Ic * a.di + (unsigned)min(Jc + 1, g.ny + 1) * a.dj;
    // CPML: the column's x and y gradings
    bool px = false, py = false;
    Prof<T> hpx{}, hpy{}, epx{}, epy{};
    if (PML && inb) {
        px = (PM & 1) && in_pml(pm.P, g.nx, I);
        py = (PM & 2) && in_pml(pm.P, g.ny, J);
        if (px) {
            hpx = pm.hp[0][I];
            epx = pm.ep[0][I];
        }
        if (py) {
            hpy = pm.hp[1][J];
            epy = pm.ep[1][J];
        }
    }
    // CPML psi carried between levels: H-psi of plane K-1, E-psi of K-1, K-2
    T psh1m[6] = {}, pse1m[6] = {}, pse1mm[6] = {};

    // slots of planes before the chunk's first are read (never used) by the
    // first iterations: start from zeros
#pragma unroll
    for (int q = 0; q < 18; ++q) L[q * NTP + t] = (T)0;

    // One plane of the march (plane parity a compile-time constant, so the
    // exchange slots are fixed LDS offsets).  Only the loaded set is carried
    // in registers; the level-1 and level-2 values of the last planes stay in
    // the LDS slots, from which each thread reads its own entries back
    // (H1(K-1), E1(K-1), E1(K-2), H2(K-2), H2(K-3)).
    auto plane = [&](auto parc, const int K, const TbSet<T> &v) {
        constexpr int par = decltype(parc)::value, p1 = par ^ 1;
        auto slot = [&](int l, int pr, int c) { return L + ((l * 2 + pr) * 3 + c) * NTP; };
        // Uniform plane conditions.  Every level is produced on its planes;
        // its updates run on planes of the update box only (K <= nz), the
        // ghost plane nz+1 passing its values on.  Lanes outside the box
        // compute garbage that is masked where it enters the exchange.
        const bool kin = K <= g.nz;
        const bool do1 = K <= K1;                          // H1 at plane K
        const bool doe1 = do1 && K >= K0 - 1;              // E1 at plane K
        const bool doh2 = K >= K0 && K <= K1 && K >= 1;    // H2 at plane K-1
        const bool doe2 = K >= K0 + 2;                     // E2 at plane K-2 (<= K1-1 by the loop bound)

        // ---- a. H1(K) ------------------------------------------------------
        T hx1 = v.hx, hy1 = v.hy, hz1 = v.hz;
        T psh[6] = {};
        const bool pz = (PM & 4) && inb && kin && in_pml(pm.P, g.nz, K);
        if (do1 && kin) {
            Prof<T> hpz{};
            if (PML && pz) hpz = pm.hp[2][K];
            if (PML) tb_psi_load(g, pm.P, pm.W, ps.oh, I, J, K, px, py, pz, psh);
            const T ez_i1 = XL ? lane_next(v.ez) : v.ez_i1, ey_i1 = XL ? lane_next(v.ey) : v.ey_i1;
            tb_h<T, PML>(hx1, hy1, hz1, v.ez_j1, v.ez, v.ey1, v.ey, v.ex1, v.ex, ez_i1, ey_i1, v.ex_j1, ch, cI, cJ,
                         c0 && K >= 1, px, py, pz, hpx, hpy, hpz, psh);
        }
        slot(0, par, 0)[t] = ing ? hx1 : (T)0;
        slot(0, par, 1)[t] = ing ? hy1 : (T)0;
        slot(0, par, 2)[t] = ing ? hz1 : (T)0;
        __syncthreads();

        // ---- e. E2(K-2): from H2(K-2) (own and at I-1 / J-1), H2(K-3), E1(K-2)
        if (doe2) {
            const int Km = K - 2;
            const bool pzm = (PM & 4) && inb && Km <= g.nz && in_pml(pm.P, g.nz, Km);
            T pvs[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) pvs[q] = pse1mm[q];
            T ex2 = slot(1, par, 0)[t], ey2 = slot(1, par, 1)[t], ez2 = slot(1, par, 2)[t];
            if (Km <= g.nz) {
                Prof<T> epz{};
                if (PML && pzm) epz = pm.ep[2][Km];
                tb_e<T, MAT, PML>(ex2, ey2, ez2, slot(2, par, 0)[t], slot(2, par, 1)[t], slot(2, par, 2)[t],
                                  slot(2, p1, 0)[t], slot(2, p1, 1)[t], slot(2, par, 2)[tim], slot(2, par, 1)[tim],
                                  slot(2, par, 2)[tjm], slot(2, par, 0)[tjm], ce, v.cb2, v.ca2, cJ && Km >= 1,
                                  cI && Km >= 1, cI && cJ, px, py, pzm, epx, epy, epz, pvs);
            }
            if (Km >= pp.s_k0 && Km <= pp.s_k1) tb_sources(pp, pp.v2, Km, uidx_dev(g, I, J, Km), ex2, ey2, ez2);
            if (core) {
                // every component, updated or not (sources reach cells that
                // are not updated; the other buffer is two steps old)
                const unsigned uo = ui + (unsigned)Km * a.dk;
                lane_at(a.nex, uo) = ex2;
                lane_at(a.ney, uo) = ey2;
                lane_at(a.nez, uo) = ez2;
                if (PML) tb_psi_store(g, pm.P, pm.W, ps.ne, I, J, Km, px, py, pzm, pvs);
                if (pp.rec2 && Km >= pp.p_k0 && Km <= pp.p_k1)
                    tb_receivers(pp, pp.rec2, Km, uidx_dev(g, I, J, Km), ex2, ey2, ez2);
            }
        }

        // ---- c. E1(K): from H1(K) (own and at I-1 / J-1), H1(K-1) ---------------
        const T hx1m = slot(0, p1, 0)[t], hy1m = slot(0, p1, 1)[t], hz1m = slot(0, p1, 2)[t];
        const T ex1m = slot(1, p1, 0)[t], ey1m = slot(1, p1, 1)[t], ez1m = slot(1, p1, 2)[t];
        T ex1 = v.ex, ey1 = v.ey, ez1 = v.ez;
        T pse[6] = {};
        if (doe1) {
            if (kin) {
                Prof<T> epz{};
                if (PML && pz) epz = pm.ep[2][K];
                if (PML) tb_psi_load(g, pm.P, pm.W, ps.oe, I, J, K, px, py, pz, pse);
                tb_e<T, MAT, PML>(ex1, ey1, ez1, hx1, hy1, hz1, hx1m, hy1m, slot(0, par, 2)[tim],
                                  slot(0, par, 1)[tim], slot(0, par, 2)[tjm], slot(0, par, 0)[tjm], ce, v.cb, v.ca,
                                  cJ && K >= 1, cI && K >= 1, cI && cJ, px, py, pz, epx, epy, epz, pse);
            }
            if (K >= pp.s_k0 && K <= pp.s_k1) tb_sources(pp, pp.v1, K, uidx_dev(g, I, J, K), ex1, ey1, ez1);
            if (pp.rec1 && K >= K0 && K < K1 && K >= pp.p_k0 && K <= pp.p_k1 && core)
                tb_receivers(pp, pp.rec1, K, uidx_dev(g, I, J, K), ex1, ey1, ez1);
        }
        // E1(K-2) of this slot was read above (by this thread only)
        slot(1, par, 0)[t] = ing ? ex1 : (T)0;
        slot(1, par, 1)[t] = ing ? ey1 : (T)0;
        slot(1, par, 2)[t] = ing ? ez1 : (T)0;

        // ---- d. H2(K-1): from H1(K-1), E1(K-1) (own and at I+1 / J+1), E1(K) -----
        T hx2 = hx1m, hy2 = hy1m, hz2 = hz1m;
        if (doh2) {
            const int Km = K - 1;
            const bool pzm = (PM & 4) && inb && Km <= g.nz && in_pml(pm.P, g.nz, Km);
            T pvs[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) pvs[q] = psh1m[q];
            if (Km <= g.nz) {
                Prof<T> hpz{};
                if (PML && pzm) hpz = pm.hp[2][Km];
                tb_h<T, PML>(hx2, hy2, hz2, slot(1, p1, 2)[tjp], ez1m, ey1, ey1m, ex1, ex1m, slot(1, p1, 2)[tip],
                             slot(1, p1, 1)[tip], slot(1, p1, 0)[tjp], ch, cI, cJ, c0 && Km >= 1, px, py, pzm, hpx,
                             hpy, hpz, pvs);
            }
            if (core && Km >= K0) {
                const unsigned uo = ui + (unsigned)Km * a.dk;
                lane_at(a.nhx, uo) = hx2;
                lane_at(a.nhy, uo) = hy2;
                lane_at(a.nhz, uo) = hz2;
                if (PML) tb_psi_store(g, pm.P, pm.W, ps.nh, I, J, Km, px, py, pzm, pvs);
            }
        }
        // H2(K-3) of this slot was read above (by this thread only)
        slot(2, p1, 0)[t] = ing ? hx2 : (T)0;
        slot(2, p1, 1)[t] = ing ? hy2 : (T)0;
        slot(2, p1, 2)[t] = ing ? hz2 : (T)0;

        if (PML) {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                pse1mm[q] = pse1m[q];
                pse1m[q] = pse[q];
                psh1m[q] = psh[q];
            }
        }
    };

    // Planes Ks .. K1+1, double-buffered: the loads of the next plane are in
    // flight while this one is computed, and the two sets alternate roles (no
    // register copies of data in flight).  Set A serves even planes, B odd.
    const int Ke = K1 + 1;
    TbSet<T> A, B;
    int K = Ks;
    if (K & 1) {
        tb_load<T, MAT, XL>(a, g.nz, ui, ui_i1, ui_j1, K, B);
        tb_load<T, MAT, XL>(a, g.nz, ui, ui_i1, ui_j1, K + 1, A);
        plane(tb_par<1>{}, K, B);
        ++K;
    } else {
        tb_load<T, MAT, XL>(a, g.nz, ui, ui_i1, ui_j1, K, A);
    }
    for (; K + 1 <= Ke; K += 2) {
        tb_load<T, MAT, XL>(a, g.nz, ui, ui_i1, ui_j1, K + 1, B);
        plane(tb_par<0>{}, K, A);
        tb_load<T, MAT, XL>(a, g.nz, ui, ui_i1, ui_j1, K + 2, A);
        plane(tb_par<1>{}, K + 1, B);
    }
    if (K <= Ke) plane(tb_par<0>{}, K, A);
}

}  // namespace fdtd
}  // namespace ems
