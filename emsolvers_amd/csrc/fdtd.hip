// Yee FDTD solver: C ABI over the gfx950 kernels of fdtd_kernels.hpp.
//
// Drop-in for the reference class FemSimulation of src/fdtd/
// (fdtd_simulation.hpp:9-80): same constructor arguments, sources,
// observation lists, run() semantics and trace layout.  The whole domain is
// device resident; traces stay on the device until they are read.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "fdtd_kernels.hpp"
#include "fdtd_tb2.hpp"

using namespace ems;
using namespace ems::fdtd;

namespace {

constexpr long kPitchAlign = 32;  // elements: rows start on 128-byte lines (256 B in FP64)
constexpr long kFrontPad = 64;    // keeps (I-1) reads of row 0 inside the allocation

struct Src {
    int comp, i, j, k;
    long u;         // unified index
    double eps;     // permittivity at the source cell
    double scale;   // 1 + sigma*dt/(2 eps) at the source cell
};

using ObsLists = std::array<std::vector<std::array<int, 3>>, 3>;

}  // namespace

struct ems_fdtd {
    ems_fdtd_desc d{};
    GridView g{};
    size_t esz = 8;
    size_t nalloc = 0;  // elements per unified array
    DevBuf f[6];
    DevBuf ca, cb;
    int mat_mode = 0;  // 0 uniform, 1 Cb, 2 Ca + Cb
    std::vector<double> eps_r, sigma;  // material grid copies (empty = default)
    // CPML
    int npml = 0;
    double pml_order = 3.0, pml_r0 = 1e-6, pml_kmax = 1.0, pml_amax = 0.0;
    DevBuf prof;    // packed CPML grading tables
    DevBuf psi[12];
    // ping-pong kernels (fused, z-march, two-step): second field set and
    // second psi set (same indexing as psi; fused and z-march swap the H-psi)
    DevBuf f2[6];
    DevBuf psi2[12];
    bool alt_ready = false;
    int tile_by = 8;   // workgroup rows of the fused kernel (4, 8 or 16)
    int tile_lz = 4;   // planes per workgroup of the fused kernel
    int tile_flags = 1;  // bit 0: prefetch the next plane, bit 1: non-temporal stores
    int split_by = 0;    // split kernels: workgroup rows (0: 4 with a CPML, else 8)
    int split_lz = 0;    // split kernels: planes per thread (0: 4 in FP64 with a CPML, else 2)
    int fused_rev = 0;   // fused kernel: tiles of this step in reverse order
    int z_ty = 4, z_lz = 8, z_nw = 1;  // z-march kernel: strip rows, planes per chunk, waves per workgroup
    int tb_txw = 0, tb_tyw = 0, tb_lz = 0;  // two-step kernel: thread tile columns, rows, planes per chunk (0: default)
    // two-step kernel: sources and receiver points sorted by plane (CSR over K)
    DevBuf d_tb_skptr, d_tb_scomp, d_tb_su, d_tb_sidx, d_tb_pkptr, d_tb_pcomp, d_tb_pu, d_tb_pidx;
    int tb_sk[2] = {0, -1}, tb_pk[2] = {0, -1};  // first and last plane holding sources / receiver points
    // two-step kernel with a CPML: the shell launches run on a second stream,
    // concurrently with the interior one (fork / join events)
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
    // fused-kernel tile lists (interior / CPML-touching)
    DevBuf d_tiles_inner, d_tiles_pml;
    int n_inner = 0, n_pml = 0, ntx = 0, nty = 0;
    int tiles_key[3] = {0, 0, 0};
    bool tiles_ok = false;
    // sources / receivers
    std::vector<Src> src;
    std::vector<ObsLists> obs;
    int wavelet = EMS_WAVELET_HERTZ;
    double f0 = 1.0e9, t0 = -1.0, amp = 1.0;
    // traces (host, reference semantics) and device staging
    std::vector<double> traces;  // [nobs][nslot][3]
    int nslot = 0;
    double t_now = 0.0;
    int kernel = EMS_KERNEL_AUTO;
    hipStream_t stream = nullptr;
    DevBuf d_src_comp, d_src_u, d_src_val, d_pt_comp, d_pt_u, d_rec;
    bool dirty_pml = true;
    // misfit gradient: checkpoint interval (0 = ceil(sqrt(steps))) and buffers
    int ckpt_interval = 0;
    size_t grad_mem = 0;  // cap on the gradient's snapshot and checkpoint bytes (0: none)
    struct Adj *adj = nullptr;
    long mat_version = 0;  // set by every set_material call (unique over all handles)
    ~ems_fdtd();
};

namespace {

// The unified grid of an nx x ny x nz handle whose cell I = 0 lies `shift`
// elements into a 128-byte line (rows keep the line-aligned pitch).
GridView make_grid(int nx, int ny, int nz, long shift = 0)
{
    GridView g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.pitch = ((long)nx + 2 + shift + kPitchAlign - 1) / kPitchAlign * kPitchAlign;
    g.slab = g.pitch * (ny + 2);
    g.off = kFrontPad + shift;
    return g;
}

inline long uidx(const GridView &g, int I, int J, int K) { return g.off + I + g.pitch * J + g.slab * K; }

// Reference-layout dims and unified shift of component w.
void ref_dims(const GridView &g, int w, int dims[3], int shift[3])
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int D[6][3] = {{nx + 1, ny + 2, nz + 2}, {nx + 2, ny + 1, nz + 2}, {nx + 2, ny + 2, nz + 1},
                         {nx, ny + 1, nz + 1},     {nx + 1, ny, nz + 1},     {nx + 1, ny + 1, nz}};
    const int S[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 3; ++a) {
        dims[a] = D[w][a];
        shift[a] = S[w][a];
    }
}

size_t ref_size(const GridView &g, int w)
{
    int d[3], s[3];
    ref_dims(g, w, d, s);
    return (size_t)d[0] * d[1] * d[2];
}

long ref_to_unified(const GridView &g, int comp, int i, int j, int k)
{
    int d[3], s[3];
    ref_dims(g, comp, d, s);
    return uidx(g, i + s[0], j + s[1], k + s[2]);
}

bool ref_in_range(const GridView &g, int comp, int i, int j, int k)
{
    int d[3], s[3];
    ref_dims(g, comp, d, s);
    return i >= 0 && j >= 0 && k >= 0 && i < d[0] && j < d[1] && k < d[2];
}

// src/fdtd/fdtd_simulation.cpp:24-36 with the same operation order.
double hertz(double t, double domain_size, double f0)
{
    return -4.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) *
           std::exp(-2.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) * (t - 1.0 / f0)) / domain_size /
           domain_size / domain_size * 0.001;
}

double wavelet_value(const ems_fdtd *s, double t)
{
    const double t0 = s->t0 < 0 ? std::sqrt(2.0) / s->f0 : s->t0;
    if (s->wavelet == EMS_WAVELET_RICKER) {
        double a = M_PI * M_PI * s->f0 * s->f0 * (t - t0) * (t - t0);
        return s->amp * ((1.0 - 2.0 * a) * std::exp(-a));
    }
    if (s->wavelet == EMS_WAVELET_GAUSS) {
        double a = M_PI * s->f0 * (t - t0);
        return s->amp * std::exp(-a * a);
    }
    return hertz(t, s->d.h, s->f0);
}

// CPML grading at physical coordinate x on an axis of n cells.
void cpml_profile(const ems_fdtd *s, int n, double x, double &kinv, double &b, double &a)
{
    const int P = s->npml;
    const double h = s->d.h, L = P * h;
    double d = 0.0;
    if (x < L) d = L - x;
    if (x > n * h - L) d = std::max(d, x - (n * h - L));
    if (P <= 0 || d <= 0.0) {
        kinv = 1.0;
        b = 1.0;
        a = 0.0;
        return;
    }
    const double m = s->pml_order;
    const double eta = std::sqrt(s->d.mu / s->d.eps);
    const double smax = -(m + 1.0) * std::log(s->pml_r0) / (2.0 * eta * L);
    const double r = d / L;
    const double gr = std::pow(r, m);
    const double sig = smax * gr;
    const double kap = 1.0 + (s->pml_kmax - 1.0) * gr;
    const double alp = s->pml_amax * (1.0 - r);
    kinv = 1.0 / kap;
    b = std::exp(-(sig / kap + alp) * s->d.dt / s->d.eps);
    a = (sig > 0.0) ? sig / (sig * kap + kap * kap * alp) * (b - 1.0) : 0.0;
}

template <typename T> ems_status upload_as(DevBuf &dst, const std::vector<double> &v)
{
    std::vector<T> tmp(v.begin(), v.end());
    EMS_TRY(dst.reserve(tmp.size() * sizeof(T)));
    EMS_HIP(hipMemcpy(dst.p, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return EMS_OK;
}

ems_status setup_pml(ems_fdtd *s)
{
    if (!s->dirty_pml) return EMS_OK;
    s->dirty_pml = false;
    for (auto &p : s->psi) p.release();
    for (auto &p : s->psi2) p.release();
    s->prof.release();
    if (s->npml <= 0) return EMS_OK;
    const GridView &g = s->g;
    const int nn[3] = {g.nx, g.ny, g.nz};
    const int P = s->npml, W = 2 * P + 1;
    // packed grading tables: [type H/E][axis][position], each (b, a, 1/kappa, 0)
    const int len = std::max({g.nx, g.ny, g.nz}) + 2;
    std::vector<double> tab((size_t)2 * 3 * len * 4, 0.0);
    for (int ax = 0; ax < 3; ++ax)
        for (int I = 0; I < nn[ax] + 2; ++I) {
            double kh, bh, ah, ke, be, ae;
            cpml_profile(s, nn[ax], I * s->d.h, kh, bh, ah);
            cpml_profile(s, nn[ax], (I - 0.5) * s->d.h, ke, be, ae);
            double *th = &tab[(((size_t)0 * 3 + ax) * len + I) * 4];
            double *te = &tab[(((size_t)1 * 3 + ax) * len + I) * 4];
            th[0] = bh; th[1] = ah; th[2] = kh;
            te[0] = be; te[1] = ae; te[2] = ke;
        }
    if (s->d.precision == EMS_F32) EMS_TRY(upload_as<float>(s->prof, tab));
    else EMS_TRY(upload_as<double>(s->prof, tab));
    const size_t nslab[3] = {(size_t)W * (g.ny + 2) * (g.nz + 2), (size_t)(g.nx + 2) * W * (g.nz + 2),
                             (size_t)(g.nx + 2) * (g.ny + 2) * W};
    for (int ax = 0; ax < 3; ++ax)
        for (int q = 0; q < 4; ++q) {
            DevBuf &b = s->psi[ax * 4 + q];
            EMS_TRY(b.reserve(nslab[ax] * s->esz));
            EMS_HIP(hipMemset(b.p, 0, nslab[ax] * s->esz));
        }
    for (int ax = 0; ax < 3; ++ax)
        for (int q = 0; q < 4; ++q) {
            DevBuf &b = s->psi2[ax * 4 + q];
            EMS_TRY(b.reserve(nslab[ax] * s->esz));
            EMS_HIP(hipMemset(b.p, 0, nslab[ax] * s->esz));
        }
    return EMS_OK;
}

void swap_buf(DevBuf &a, DevBuf &b)
{
    std::swap(a.p, b.p);
    std::swap(a.bytes, b.bytes);
}

// Second field set for the fused kernel, initialised as a copy of the
// current one so that entries no kernel writes (ghosts) agree in both.
ems_status ensure_alt(ems_fdtd *s)
{
    if (s->alt_ready) return EMS_OK;
    for (int w = 0; w < 6; ++w) {
        EMS_TRY(s->f2[w].reserve(s->nalloc * s->esz));
        EMS_HIP(hipMemcpy(s->f2[w].p, s->f[w].p, s->nalloc * s->esz, hipMemcpyDeviceToDevice));
    }
    s->alt_ready = true;
    return EMS_OK;
}

// Row shift for the two-step kernel's CPML x-slabs (FP32): their tiles read
// the columns [0, P+5) and [nx-P-2, nx+2) of every row; the shift puts each
// range into as few 128-byte lines as possible (at 256^3 with P = 10, one
// line each instead of one and two: the x-slabs fetched 341 MB per two steps
// unshifted).  FP64 rows (16 words per line) gain nothing from a shift.
long row_shift(int nx, int P, size_t esz)
{
    if (P <= 0 || esz != 4) return 0;
    auto lines = [](long a, long b) { return (b - 1) / kPitchAlign - a / kPitchAlign + 1; };
    long best = 0, bl = 1L << 30;
    for (long sh = 0; sh < kPitchAlign; ++sh) {
        const long l = lines(sh, sh + P + 5) + lines(sh + nx - P - 2, sh + nx + 2);
        if (l < bl) {
            bl = l;
            best = sh;
        }
    }
    return best;
}

// Moves every unified array of the handle to the layout with row shift
// `shift` (cells I in [0, nx+1] of every row; pads zeroed).  Cached unified
// indices and the adjoint buffers are rebuilt.
ems_status relayout(ems_fdtd *s, long shift)
{
    const GridView g0 = s->g;
    if (g0.off - kFrontPad == shift) return EMS_OK;
    const GridView g1 = make_grid(g0.nx, g0.ny, g0.nz, shift);
    const size_t n1 = (size_t)(g1.off + g1.slab * (g0.nz + 2) + kFrontPad);
    const size_t e = s->esz;
    // every moved array is built completely before any is replaced, so that
    // an error leaves the handle as it was
    DevBuf *arrays[14] = {&s->ca, &s->cb};
    for (int w = 0; w < 6; ++w) {
        arrays[2 + w] = &s->f[w];
        arrays[8 + w] = &s->f2[w];
    }
    DevBuf moved[14];
    for (int q = 0; q < 14; ++q) {
        const DevBuf &b = *arrays[q];
        if (!b.p) continue;
        EMS_TRY(moved[q].reserve(n1 * e));
        EMS_HIP(hipMemset(moved[q].p, 0, n1 * e));
        EMS_HIP(hipMemcpy2D(static_cast<char *>(moved[q].p) + g1.off * e, g1.pitch * e,
                            static_cast<const char *>(b.p) + g0.off * e, g0.pitch * e, (size_t)(g0.nx + 2) * e,
                            (size_t)(g0.ny + 2) * (g0.nz + 2), hipMemcpyDeviceToDevice));
    }
    for (int q = 0; q < 14; ++q)
        if (arrays[q]->p) *arrays[q] = std::move(moved[q]);
    s->g = g1;
    s->nalloc = n1;
    for (auto &q : s->src) q.u = ref_to_unified(s->g, q.comp, q.i, q.j, q.k);
    s->tiles_ok = false;
    delete s->adj;
    s->adj = nullptr;
    return EMS_OK;
}

template <typename T> PmlView<T> pml_view(ems_fdtd *s)
{
    PmlView<T> v{};
    v.P = s->npml > 0 ? s->npml : 0;
    v.W = 2 * v.P + 1;
    if (s->npml <= 0) return v;
    const int len = std::max({s->g.nx, s->g.ny, s->g.nz}) + 2;
    const Prof<T> *t = reinterpret_cast<const Prof<T> *>(s->prof.p);
    for (int ax = 0; ax < 3; ++ax) {
        v.hp[ax] = t + (size_t)(0 * 3 + ax) * len;
        v.ep[ax] = t + (size_t)(1 * 3 + ax) * len;
        v.psi_h[ax][0] = s->psi[ax * 4 + 0].as<T>();
        v.psi_h[ax][1] = s->psi[ax * 4 + 1].as<T>();
        v.psi_e[ax][0] = s->psi[ax * 4 + 2].as<T>();
        v.psi_e[ax][1] = s->psi[ax * 4 + 3].as<T>();
    }
    return v;
}

template <typename T> Fields<T> fields(ems_fdtd *s)
{
    return Fields<T>{s->f[0].as<T>(), s->f[1].as<T>(), s->f[2].as<T>(),
                     s->f[3].as<T>(), s->f[4].as<T>(), s->f[5].as<T>()};
}

// Boxes of the split sweeps with a CPML of P cells: the interior box (no
// CPML cell) and the shell around it (z-slabs over whole planes, y-slabs over
// the interior planes, x-slab columns over the interior rows and planes).
// False if the grid has no interior box.
bool split_boxes(const GridView &g, int P, int by, int lz, SplitBoxes &inner, int &n_inner, SplitBoxes &shell,
                 int &n_shell)
{
    if (g.nx < 2 * P + 1 || g.ny < 2 * P + 1 || g.nz < 2 * P + 1) return false;
    const int npw = (2 * P + 1 + 63) / 64;
    const int nfull = split_row_waves(g, P);
    auto add = [&](SplitBoxes &bs, int &total, SplitBox b) {
        const int rows = b.jpack ? 2 * P + 1 : b.j1 - b.j0;
        const int per = b.xrows > 0 ? by * b.xrows : by;  // rows per workgroup
        b.nby = (rows + per - 1) / per;
        b.base = total;
        bs.b[bs.n++] = b;
        total += b.nbx * b.nby * ((b.k1 - b.k0 + lz - 1) / lz);
    };
    inner.n = shell.n = 0;
    n_inner = n_shell = 0;
    const int Ki0 = P + 1, Ki1 = g.nz - P + 1, Ji0 = P + 1, Ji1 = g.ny - P + 1;
    // x-slab columns of several rows per wavefront (2P+1 <= 64 always here)
    const int xrows = 64 / (2 * P + 1);
    //              i0     i1            j0   j1        k0            k1        packed jpack xrows nbx
    // interior columns P+1 .. nx-P on wavefronts aligned to 64 columns (as split_column)
    const int ia = (P + 1) / 64 * 64;
    SplitBox ib{ia, g.nx - P + 1, Ji0, Ji1, Ki0, Ki1, 0, 0, 0, (g.nx - P + 1 - ia + 63) / 64, 0, 0};
    ib.ilo = P + 1;
    add(inner, n_inner, ib);
    add(shell, n_shell, {0, 0, 0, g.ny + 1, 0, P + 1, 1, 0, 0, nfull, 0, 0});
    add(shell, n_shell, {0, 0, 0, g.ny + 1, g.nz - P + 1, g.nz + 1, 1, 0, 0, nfull, 0, 0});
    add(shell, n_shell, {0, 0, 0, 0, Ki0, Ki1, 1, 1, 0, nfull, 0, 0});
    if (xrows > 0) add(shell, n_shell, {0, 0, Ji0, Ji1, Ki0, Ki1, 0, 0, xrows, 1, 0, 0});
    else add(shell, n_shell, {0, 0, Ji0, Ji1, Ki0, Ki1, 1, 0, 0, npw, 0, 0});
    return true;
}

template <typename T, int MAT, bool PML>
void launch_split(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm)
{
    const GridView &g = s->g;
    // measured at 256^3: workgroups of 4 rows with a CPML (4 % faster than 8
    // in FP32 and FP64; 2 rows are 1 % slower in FP32), of 8 without (1 %
    // faster than 4); planes per thread 2 (FP32 with a CPML 2 % faster than
    // 4, without one 4 %), except FP64 with a CPML: 4 (1 % faster)
    const int by = s->split_by > 0 ? s->split_by : PML ? 4 : 8;
    const int lz = s->split_lz > 0 ? s->split_lz : PML && sizeof(T) == 8 ? 4 : 2;
    if (PML && (s->tile_flags & 4)) {
        SplitBoxes inner, shell;
        int ni, ns;
        if (split_boxes(g, s->npml, by, lz, inner, ni, shell, ns)) {
            const int npw = (2 * s->npml + 1 + 63) / 64;
            dim3 blk(64, by, 1);
            Fields<T> f = fields<T>(s);
            k_h_boxes<T, false><<<ni, blk, 0, s->stream>>>(g, ch, f, pm, lz, npw, inner);
            k_h_boxes<T, true><<<ns, blk, 0, s->stream>>>(g, ch, f, pm, lz, npw, shell);
            k_e_boxes<T, MAT, false><<<ni, blk, 0, s->stream>>>(g, ce, f, s->ca.as<T>(), s->cb.as<T>(), pm, lz, npw,
                                                                inner);
            k_e_boxes<T, MAT, true><<<ns, blk, 0, s->stream>>>(g, ce, f, s->ca.as<T>(), s->cb.as<T>(), pm, lz, npw,
                                                               shell);
            return;
        }
    }
    // CPML: the x-slab columns get wavefronts of their own
    const int P = PML ? s->npml : 0;
    const int npw = PML ? (2 * P + 1 + 63) / 64 : 0;
    const int nbx = split_row_waves(g, P);
    dim3 blk(64, by, 1);
    const int nby = (g.ny + 1 + by - 1) / by, ntiles = nby * ((g.nz + 1 + lz - 1) / lz);
    const unsigned grd = 8u * (unsigned)nbx * (unsigned)((ntiles + 7) / 8);  // see split_tile
    Fields<T> f = fields<T>(s);
    k_h_split<T, PML><<<grd, blk, 0, s->stream>>>(g, ch, f, pm, lz, npw, nbx, nby, ntiles);
    k_e_split<T, MAT, PML><<<grd, blk, 0, s->stream>>>(g, ce, f, s->ca.as<T>(), s->cb.as<T>(), pm, lz, npw, nbx, nby,
                                                        ntiles);
}

// Tile lists of the fused kernel: tiles whose cells (halos included) touch
// the CPML, and the others.  Rebuilt when the tiling or the CPML changes.
ems_status build_tiles(ems_fdtd *s)
{
    const GridView &g = s->g;
    const int R = s->tile_by, lz = s->tile_lz, P = s->npml;
    if (s->tiles_ok && s->tiles_key[0] == R && s->tiles_key[1] == lz && s->tiles_key[2] == P) return EMS_OK;
    const int ntx = (g.nx + 1 + 62) / 63;
    const int rows = R - 1;  // owned rows per workgroup (row 0 is the halo)
    const int nty = (g.ny + 1 + rows - 1) / rows;
    const int nkc = (g.nz + 1 + lz - 1) / lz;
    std::vector<int> inner, pml;
    auto touches = [&](int a, int b, int n) { return P > 0 && (a <= P || b >= n - P + 1); };
    for (int kc = 0; kc < nkc; ++kc)
        for (int tj = 0; tj < nty; ++tj)
            for (int ti = 0; ti < ntx; ++ti) {
                const int i0 = std::max(ti * 63 - 1, 0), i1 = std::min(ti * 63 + 62, g.nx);
                const int j0 = std::max(tj * rows - 1, 0), j1 = std::min(tj * rows + rows - 1, g.ny);
                const int k0 = std::max(kc * lz - 1, 0), k1 = std::min(kc * lz + lz - 1, g.nz);
                const int id = ti + ntx * (tj + nty * kc);
                if (touches(i0, i1, g.nx) || touches(j0, j1, g.ny) || touches(k0, k1, g.nz)) pml.push_back(id);
                else inner.push_back(id);
            }
    s->ntx = ntx;
    s->nty = nty;
    s->n_inner = (int)inner.size();
    s->n_pml = (int)pml.size();
    if (!inner.empty()) {
        EMS_TRY(s->d_tiles_inner.reserve(sizeof(int) * inner.size()));
        EMS_HIP(hipMemcpy(s->d_tiles_inner.p, inner.data(), sizeof(int) * inner.size(), hipMemcpyHostToDevice));
    }
    if (!pml.empty()) {
        EMS_TRY(s->d_tiles_pml.reserve(sizeof(int) * pml.size()));
        EMS_HIP(hipMemcpy(s->d_tiles_pml.p, pml.data(), sizeof(int) * pml.size(), hipMemcpyHostToDevice));
    }
    s->tiles_key[0] = R;
    s->tiles_key[1] = lz;
    s->tiles_key[2] = P;
    s->tiles_ok = true;
    return EMS_OK;
}

template <typename T, int BY, int MAT, bool PML, bool PF, bool NT>
void launch_tiles(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm, const PsiH<T> &ph, const int *list, int count)
{
    if (count <= 0) return;
    // Workgroup order, measured at 256^3: XCD-banded (XCD c = b % 8 takes one
    // contiguous eighth of the tiles) is faster in FP32 and slower in FP64
    // than plain order; the tiles run in reverse every other step, so that a
    // step starts where the previous one ended, on what the caches hold
    // (2-3 % faster in both precisions).
    TileSet ts{list, count, s->ntx, s->nty, s->tile_lz, sizeof(T) == 4 ? 1 : 0, s->fused_rev};
    const int grid = (count + 7) / 8 * 8;
    Fields<T> o = fields<T>(s);
    Fields<T> nw{s->f2[0].as<T>(), s->f2[1].as<T>(), s->f2[2].as<T>(),
                 s->f2[3].as<T>(), s->f2[4].as<T>(), s->f2[5].as<T>()};
    k_fused<T, BY, MAT, PML, PF, NT><<<grid, dim3(64, BY, 1), 0, s->stream>>>(s->g, ch, ce, o, nw, s->ca.as<T>(),
                                                                      s->cb.as<T>(), pm, ph, ts);
}

template <typename T, int R, int MAT, bool PML>
void launch_fused(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm)
{
    PsiH<T> ph{};
    if (PML)
        for (int ax = 0; ax < 3; ++ax)
            for (int q = 0; q < 2; ++q) {
                ph.old_[ax][q] = s->psi[ax * 4 + q].as<T>();
                ph.new_[ax][q] = s->psi2[ax * 4 + q].as<T>();
            }
    // interior tiles without CPML code, then the CPML tiles; both read the
    // old buffers and write disjoint cells of the new ones
    switch (s->tile_flags & 3) {
    case 0:
        launch_tiles<T, R, MAT, false, false, false>(s, ch, ce, pm, ph, s->d_tiles_inner.as<int>(), s->n_inner);
        if (PML) launch_tiles<T, R, MAT, true, false, false>(s, ch, ce, pm, ph, s->d_tiles_pml.as<int>(), s->n_pml);
        break;
    case 1:
        launch_tiles<T, R, MAT, false, true, false>(s, ch, ce, pm, ph, s->d_tiles_inner.as<int>(), s->n_inner);
        if (PML) launch_tiles<T, R, MAT, true, true, false>(s, ch, ce, pm, ph, s->d_tiles_pml.as<int>(), s->n_pml);
        break;
    case 2:
        launch_tiles<T, R, MAT, false, false, true>(s, ch, ce, pm, ph, s->d_tiles_inner.as<int>(), s->n_inner);
        if (PML) launch_tiles<T, R, MAT, true, false, true>(s, ch, ce, pm, ph, s->d_tiles_pml.as<int>(), s->n_pml);
        break;
    default:
        launch_tiles<T, R, MAT, false, true, true>(s, ch, ce, pm, ph, s->d_tiles_inner.as<int>(), s->n_inner);
        if (PML) launch_tiles<T, R, MAT, true, true, true>(s, ch, ce, pm, ph, s->d_tiles_pml.as<int>(), s->n_pml);
        break;
    }
    s->fused_rev ^= 1;
    // the new buffers become current
    for (int w = 0; w < 6; ++w) swap_buf(s->f[w], s->f2[w]);
    if (PML)
        for (int ax = 0; ax < 3; ++ax)
            for (int q = 0; q < 2; ++q) swap_buf(s->psi[ax * 4 + q], s->psi2[ax * 4 + q]);
}

template <typename T, int MAT, bool PML> void launch_fused_by(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm)
{
    if (s->tile_by == 16) launch_fused<T, 16, MAT, PML>(s, ch, ce, pm);
    else launch_fused<T, 8, MAT, PML>(s, ch, ce, pm);
}

// EMS_KERNEL_AUTO: measured on MI355X (eps_r model, ms per step), the
// two-step sweep is fastest on large grids, in both precisions:
//     256^3 FP32  no CPML  0.193 (split 0.267)   CPML 10  0.267 (split 0.324)
//     256^3 FP64  no CPML  0.354 (fused 0.498)   CPML 10  0.485 (split 0.622)
//     192^3 FP32  CPML 10  0.152 (split 0.162)
// but not on small ones, where the CPML shell and short launches dominate
// (FP32 CPML 10: 128^3 0.068 vs 0.051, 96^3 0.055 vs 0.035, 64^3 0.052 vs
// 0.021).  So it is chosen when the CPML-free interior spans at least 144
// cells on every axis and the field arrays stay under 4 GiB (32-bit byte
// offsets); otherwise the split sweeps run, or the fused sweep in FP64
// without a CPML (split 0.38 vs fused 0.57 ms FP32 with a CPML, fused 0.50
// vs split 0.53 ms FP64 without, at 256^3).
bool tb2_fits(const ems_fdtd *s) { return (double)s->g.slab * (s->g.nz + 2) * s->esz < 4294967296.0; }

int kernel_kind(const ems_fdtd *s)
{
    if (s->kernel != EMS_KERNEL_AUTO) return s->kernel;
    const int shell = s->npml > 0 ? 2 * s->npml + 4 : 0;
    const int inner = std::min({s->g.nx, s->g.ny, s->g.nz}) - shell;
    if (tb2_fits(s) && inner >= 144) return EMS_KERNEL_TB2;
    return s->npml <= 0 && s->d.precision == EMS_F64 ? EMS_KERNEL_FUSED : EMS_KERNEL_SPLIT;
}

// The fused and z-march sweeps ping-pong between two field sets; only the
// fused sweep uses tile lists.
bool use_fused(const ems_fdtd *s) { return kernel_kind(s) != EMS_KERNEL_SPLIT; }
bool use_tiles(const ems_fdtd *s) { return kernel_kind(s) == EMS_KERNEL_FUSED; }

template <typename T, int TY, int NW, int MAT, bool PML>
void launch_zmarch_t(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm)
{
    const GridView &g = s->g;
    PsiH<T> ph{};
    if (PML)
        for (int ax = 0; ax < 3; ++ax)
            for (int q = 0; q < 2; ++q) {
                ph.old_[ax][q] = s->psi[ax * 4 + q].as<T>();
                ph.new_[ax][q] = s->psi2[ax * 4 + q].as<T>();
            }
    ZTiles zt;
    zt.ntx = (g.nx + 1 + 61) / 62;
    const int nstrip = (g.ny + 1 + TY - 1) / TY;
    zt.nsg = (nstrip + NW - 1) / NW;
    zt.lz = s->z_lz;
    const int nkc = (g.nz + 1 + zt.lz - 1) / zt.lz;
    zt.count = zt.ntx * zt.nsg * nkc;
    zt.per = (zt.count + 7) / 8;
    Fields<T> o = fields<T>(s);
    Fields<T> nw{s->f2[0].as<T>(), s->f2[1].as<T>(), s->f2[2].as<T>(),
                 s->f2[3].as<T>(), s->f2[4].as<T>(), s->f2[5].as<T>()};
    k_zmarch<T, TY, NW, MAT, PML><<<zt.per * 8, dim3(64, NW, 1), 0, s->stream>>>(s->g, ch, ce, o, nw, s->ca.as<T>(),
                                                                                s->cb.as<T>(), pm, ph, zt);
    for (int w = 0; w < 6; ++w) swap_buf(s->f[w], s->f2[w]);
    if (PML)
        for (int ax = 0; ax < 3; ++ax)
            for (int q = 0; q < 2; ++q) swap_buf(s->psi[ax * 4 + q], s->psi2[ax * 4 + q]);
}

template <typename T, int MAT, bool PML> void launch_zmarch(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm)
{
    if (s->z_ty == 2) {
        if (s->z_nw == 1) launch_zmarch_t<T, 2, 1, MAT, PML>(s, ch, ce, pm);
        else launch_zmarch_t<T, 2, 4, MAT, PML>(s, ch, ce, pm);
    } else {
        if (s->z_nw == 1) launch_zmarch_t<T, 4, 1, MAT, PML>(s, ch, ce, pm);
        else launch_zmarch_t<T, 4, 4, MAT, PML>(s, ch, ce, pm);
    }
}

// Tile of the two-step kernel: the handle's setting or the default for the
// precision (FP32 64 x 16 threads and 12 planes, FP64 64 x 8 and 16 planes:
// one wavefront per tile row; measured best at 256^3 with a 10-cell CPML
// among 64 x 8..16, 68 x 7..15, 36 x 14..28, 32 x 16, 28 x 18, 24 x 21,
// 44 x 11..23, 100 x 10, 128 x 8, 132 x 7 and 8..32 planes).
template <typename T> void tb_tile(const ems_fdtd *s, int &txw, int &tyw, int &lz)
{
    txw = s->tb_txw > 0 ? s->tb_txw : 64;
    tyw = s->tb_tyw > 0 ? s->tb_tyw : sizeof(T) == 4 ? 16 : 8;
    lz = s->tb_lz > 0 ? s->tb_lz : sizeof(T) == 4 ? 12 : 16;
}

// Launches of one two-step advance.  Without a CPML: one launch over the grid.
// With a CPML of P cells: the interior box, whose tiles (with their 2 + 1
// halo and the 2 + 1 planes below / above a chunk) meet no CPML cell, without
// CPML code; the shell in three launches of two slabs each, whose
// instantiations carry only the psi of the axes they meet:
//     z-slabs  K < P+3 or K > n-P-1, whole planes                 (x, y, z)
//     x-slabs  the other planes, I < P+3 or I > n-P-1, all rows   (x, y)
//     y-slabs  the other planes and columns, J < P+3 or J > n-P-1 (y)
// Every cell of the update box is stored by exactly one launch.
// x-halo of a tile of width txw in the interior or z-slab launches: 4 for
// one wavefront per row (I+1 neighbours by DPP), else 3
inline int tb_hx(int txw) { return txw == 64 ? 4 : 3; }

struct TbLaunch {
    int pm;  // CPML axes (bit 0 x, 1 y, 2 z)
    TbTiles tt;
};

template <typename T> std::vector<TbLaunch> tb_plan(const ems_fdtd *s)
{
    const GridView &g = s->g;
    int txw, tyw, lz;
    tb_tile<T>(s, txw, tyw, lz);
    // shell launches use workgroups of at most 512 threads (so that their
    // CPML instantiations get 256 registers)
    const int maxt = 512;
    // a tile row of one wavefront (64 columns, the interior and z-slabs) has
    // a two-column high halo in x
    auto box = [&](int i0, int i1, int j0, int j1, int k0, int k1, int bw, int bh, int bz, int hx = 3) {
        TbBox b{};
        b.i0 = i0;
        b.j0 = j0;
        b.k0 = k0;
        b.ni = i1 - i0;
        b.nj = j1 - j0;
        b.nk = k1 - k0;
        b.ntx = (b.ni + bw - hx - 1) / (bw - hx);
        b.nty = (b.nj + bh - 4) / (bh - 3);
        b.nkc = (b.nk + bz - 1) / bz;
        b.count = b.ni > 0 && b.nj > 0 && b.nk > 0 ? b.ntx * b.nty * b.nkc : 0;
        return b;
    };
    auto launch = [&](int pm, int bw, int bh, int bz, TbBox b0, TbBox b1) {
        TbLaunch l{};
        l.pm = pm;
        l.tt.txw = bw;
        l.tt.tyw = bh;
        l.tt.lz = bz;
        l.tt.box[0] = b0;
        l.tt.box[1] = b1;
        l.tt.count = b0.count + b1.count;
        l.tt.per = (l.tt.count + 7) / 8;
        return l;
    };
    std::vector<TbLaunch> v;
    const int P = s->npml > 0 ? s->npml : 0;
    // the boxes cover the update box and its ghost layer n+1, whose cells are
    // never updated but may receive sources
    const int n1x = g.nx + 2, n1y = g.ny + 2, n1z = g.nz + 2;
    const TbBox none{};
    if (P == 0) {
        v.push_back(launch(0, txw, tyw, lz, box(0, n1x, 0, n1y, 0, n1z, txw, tyw, lz, tb_hx(txw)), none));
        return v;
    }
    // interior cores [P+3, n-P-1] on each axis
    const int m = P + 3;
    const int ex = g.nx - P, ey = g.ny - P, ez = g.nz - P;  // one past the interior cores
    if (ex <= m || ey <= m || ez <= m) {
        const int zw = txw * tyw <= maxt ? txw : 36, zh = txw * tyw <= maxt ? tyw : 14;
        v.push_back(launch(7, zw, zh, lz, box(0, n1x, 0, n1y, 0, n1z, zw, zh, lz, tb_hx(zw)), none));
        return v;
    }
    v.push_back(launch(0, txw, tyw, lz, box(m, ex, m, ey, m, ez, txw, tyw, lz, tb_hx(txw)), none));
    // z-slabs: one chunk per slab
    const int zl = std::max(m, n1z - ez);
    const int zw = txw * tyw <= maxt ? txw : 36, zh = txw * tyw <= maxt ? tyw : 14;
    v.push_back(launch(7, zw, zh, zl, box(0, n1x, 0, n1y, 0, m, zw, zh, zl, tb_hx(zw)),
                       box(0, n1x, 0, n1y, ez, n1z, zw, zh, zl, tb_hx(zw))));
    // x-slabs: narrow tiles across I
    const int xw = std::max(m, n1x - ex) + 3;
    const int xh = std::min(64, maxt / xw);
    v.push_back(launch(3, xw, xh, lz, box(0, m, 0, n1y, m, ez, xw, xh, lz), box(ex, n1x, 0, n1y, m, ez, xw, xh, lz)));
    // y-slabs: narrow tiles across J
    const int yh = std::max(m, n1y - ey) + 3;
    const int yw = std::min(64, maxt / yh);
    v.push_back(launch(2, yw, yh, lz, box(m, ex, 0, m, m, ez, yw, yh, lz), box(m, ex, ey, n1y, m, ez, yw, yh, lz)));
    return v;
}

// Two steps in one advance (EMS_KERNEL_TB2): every launch reads the current
// sets and writes the other field set and the other psi set (all twelve
// slabs), which become current.
template <typename T, int MAT, int PM, int MAXT, bool XL>
void launch_tb2_t(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm, const TbPost &pp, const TbTiles &tt,
                  hipStream_t st)
{
    const GridView &g = s->g;
    if (tt.count == 0) return;
    PsiSet<T> ps{};
    if (PM != 0)
        for (int ax = 0; ax < 3; ++ax)
            for (int q = 0; q < 2; ++q) {
                ps.oh[ax][q] = s->psi[ax * 4 + q].as<T>();
                ps.oe[ax][q] = s->psi[ax * 4 + 2 + q].as<T>();
                ps.nh[ax][q] = s->psi2[ax * 4 + q].as<T>();
                ps.ne[ax][q] = s->psi2[ax * 4 + 2 + q].as<T>();
            }
    TbArrays<T> ar{};
    const long off = g.off;
    ar.ex = s->f[0].as<T>() + off;
    ar.ey = s->f[1].as<T>() + off;
    ar.ez = s->f[2].as<T>() + off;
    ar.hx = s->f[3].as<T>() + off;
    ar.hy = s->f[4].as<T>() + off;
    ar.hz = s->f[5].as<T>() + off;
    ar.nex = s->f2[0].as<T>() + off;
    ar.ney = s->f2[1].as<T>() + off;
    ar.nez = s->f2[2].as<T>() + off;
    ar.nhx = s->f2[3].as<T>() + off;
    ar.nhy = s->f2[4].as<T>() + off;
    ar.nhz = s->f2[5].as<T>() + off;
    ar.ca = s->ca.p ? s->ca.as<T>() + off : nullptr;
    ar.cb = s->cb.p ? s->cb.as<T>() + off : nullptr;
    ar.di = (unsigned)sizeof(T);
    ar.dj = (unsigned)(g.pitch * sizeof(T));
    ar.dk = (unsigned)(g.slab * sizeof(T));
    const int threads = (tt.txw * tt.tyw + 63) / 64 * 64;
    const size_t lds = ((size_t)18 * MAXT + 2 * (tt.txw + 1)) * sizeof(T);
    static std::once_flag attr;  // per instantiation: allow the full exchange size
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tb2<T, MAT, PM, MAXT, XL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (18 * MAXT + 2 * (MAXT + 1)) * (int)sizeof(T));
    });
    k_tb2<T, MAT, PM, MAXT, XL><<<tt.per * 8, threads, lds, st>>>(g, ch, ce, ar, pm, ps, tt, pp);
}

template <typename T, int MAT>
void launch_tb2_m(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm, const TbPost &pp, const TbLaunch &l, hipStream_t st)
{
    // the CPML-free interior of an FP32 grid may use workgroups of 1024
    // threads; everything else at most 512
    constexpr int M0 = sizeof(T) == 4 ? 1024 : 512;
    const bool xl = tb_hx(l.tt.txw) == 4;
    switch (l.pm) {
    case 0:
        if (xl) launch_tb2_t<T, MAT, 0, M0, true>(s, ch, ce, pm, pp, l.tt, st);
        else launch_tb2_t<T, MAT, 0, M0, false>(s, ch, ce, pm, pp, l.tt, st);
        break;
    case 2: launch_tb2_t<T, MAT, 2, 512, false>(s, ch, ce, pm, pp, l.tt, st); break;
    case 3: launch_tb2_t<T, MAT, 3, 512, false>(s, ch, ce, pm, pp, l.tt, st); break;
    default:
        if (xl) launch_tb2_t<T, MAT, 7, 512, true>(s, ch, ce, pm, pp, l.tt, st);
        else launch_tb2_t<T, MAT, 7, 512, false>(s, ch, ce, pm, pp, l.tt, st);
        break;
    }
}

template <typename T> void launch_tb2_one(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm, const TbPost &pp,
                                          const TbLaunch &l, hipStream_t st)
{
    switch (s->mat_mode) {
    case 0: launch_tb2_m<T, 0>(s, ch, ce, pm, pp, l, st); break;
    case 1: launch_tb2_m<T, 1>(s, ch, ce, pm, pp, l, st); break;
    default: launch_tb2_m<T, 2>(s, ch, ce, pm, pp, l, st); break;
    }
}

// The side streams and fork / join events of a two-step handle with a CPML
// (created once, at the handle's set-up).
ems_status tb_streams(ems_fdtd *s)
{
    if (s->npml <= 0) return EMS_OK;
    if (!s->ev_fork) EMS_HIP(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    for (int q = 0; q < 3; ++q) {
        if (!s->side[q]) EMS_HIP(hipStreamCreateWithFlags(&s->side[q], hipStreamNonBlocking));
        if (!s->ev_join[q]) EMS_HIP(hipEventCreateWithFlags(&s->ev_join[q], hipEventDisableTiming));
    }
    return EMS_OK;
}

// One two-step advance.  With a CPML each shell launch goes to a stream of
// its own (three streams of the handle, default priority) while the interior
// runs on the handle's stream: they read the same old sets and write disjoint
// cells, and the shell's small launches would otherwise leave most of the GPU
// idle (measured at 256^3 FP32: 0.32 ms per step with the four launches in
// sequence, 0.27 with the shell on one second stream).
template <typename T> ems_status launch_tb2(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm, const TbPost &pp)
{
    const std::vector<TbLaunch> plan = tb_plan<T>(s);
    if (plan.size() > 1) {
        const size_t ns = std::min<size_t>(plan.size() - 1, 3);
        EMS_TRY(tb_streams(s));
        EMS_HIP(hipEventRecord(s->ev_fork, s->stream));
        for (size_t q = 0; q < ns; ++q) EMS_HIP(hipStreamWaitEvent(s->side[q], s->ev_fork, 0));
        for (size_t q = 1; q < plan.size(); ++q) launch_tb2_one<T>(s, ch, ce, pm, pp, plan[q], s->side[(q - 1) % ns]);
        EMS_HIP(hipGetLastError());
        for (size_t q = 0; q < ns; ++q) EMS_HIP(hipEventRecord(s->ev_join[q], s->side[q]));
        launch_tb2_one<T>(s, ch, ce, pm, pp, plan[0], s->stream);
        EMS_HIP(hipGetLastError());
        for (size_t q = 0; q < ns; ++q) EMS_HIP(hipStreamWaitEvent(s->stream, s->ev_join[q], 0));
    } else {
        launch_tb2_one<T>(s, ch, ce, pm, pp, plan[0], s->stream);
        EMS_HIP(hipGetLastError());
    }
    for (int w = 0; w < 6; ++w) swap_buf(s->f[w], s->f2[w]);
    if (s->npml > 0)
        for (int q = 0; q < 12; ++q) swap_buf(s->psi[q], s->psi2[q]);
    return EMS_OK;
}

// Single steps of a two-step handle (odd step counts, the gradient's
// recompute, ems_fdtd_run_static) take the split sweeps: they work in place on
// the current sets, which the two-step launches keep current.
template <typename T, int MAT, bool PML> void launch_variant(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm)
{
    switch (kernel_kind(s)) {
    case EMS_KERNEL_FUSED: launch_fused_by<T, MAT, PML>(s, ch, ce, pm); break;
    case EMS_KERNEL_ZMARCH: launch_zmarch<T, MAT, PML>(s, ch, ce, pm); break;
    default: launch_split<T, MAT, PML>(s, ch, ce, pm); break;
    }
}

// One Yee step; the psi pointers of pm are refreshed by the caller after a
// fused step (the H-psi buffers swap).
template <typename T> void launch_step(ems_fdtd *s, T ch, T ce, const PmlView<T> &pm)
{
    const bool pml = s->npml > 0;
    switch (s->mat_mode * 2 + (pml ? 1 : 0)) {
    case 0: launch_variant<T, 0, false>(s, ch, ce, pm); break;
    case 1: launch_variant<T, 0, true>(s, ch, ce, pm); break;
    case 2: launch_variant<T, 1, false>(s, ch, ce, pm); break;
    case 3: launch_variant<T, 1, true>(s, ch, ce, pm); break;
    case 4: launch_variant<T, 2, false>(s, ch, ce, pm); break;
    default: launch_variant<T, 2, true>(s, ch, ce, pm); break;
    }
}

// Host side of run(): trace resize (saved_electric_field_ semantics), the
// per-step source values and the receiver tables.  Shared by single and
// batched runs.
struct RunPlan {
    int nsrc = 0, npts = 0, nrec = 0;
    double t_end = 0.0;
};

// The two-step kernel finds a plane's sources and receiver points through
// CSR lists over K (stable: sources of one cell keep their order).  sc, su:
// the sources in value-row order; pc, pu: the receiver points in raw-sample
// column order.
ems_status tb_lists(ems_fdtd *s, const std::vector<int> &sc, const std::vector<long> &su, const std::vector<int> &pc,
                    const std::vector<long> &pu)
{
    const GridView &g = s->g;
    auto plane_of = [&](long u) { return (int)((u - g.off) / g.slab); };
    auto csr = [&](int n, const std::vector<int> &comp, const std::vector<long> &uu, DevBuf &kptr, DevBuf &dc,
                   DevBuf &du, DevBuf &di, int (&krange)[2]) -> ems_status {
        krange[0] = 0;
        krange[1] = -1;
        for (int q = 0; q < n; ++q) {
            const int K = plane_of(uu[q]);
            krange[0] = q == 0 ? K : std::min(krange[0], K);
            krange[1] = q == 0 ? K : std::max(krange[1], K);
        }
        std::vector<int> order(n);
        for (int q = 0; q < n; ++q) order[q] = q;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return plane_of(uu[a]) < plane_of(uu[b]); });
        std::vector<int> kp(g.nz + 3, 0), c2(std::max(n, 1)), i2(std::max(n, 1));
        std::vector<long> u2(std::max(n, 1));
        for (int q = 0; q < n; ++q) {
            const int o = order[q];
            c2[q] = comp[o];
            u2[q] = uu[o];
            i2[q] = o;
            kp[plane_of(uu[o]) + 1]++;
        }
        for (int K = 0; K < g.nz + 2; ++K) kp[K + 1] += kp[K];
        EMS_TRY(kptr.reserve(sizeof(int) * kp.size()));
        EMS_TRY(dc.reserve(sizeof(int) * c2.size()));
        EMS_TRY(du.reserve(sizeof(long) * u2.size()));
        EMS_TRY(di.reserve(sizeof(int) * i2.size()));
        EMS_HIP(hipMemcpy(kptr.p, kp.data(), sizeof(int) * kp.size(), hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(dc.p, c2.data(), sizeof(int) * c2.size(), hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(du.p, u2.data(), sizeof(long) * u2.size(), hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(di.p, i2.data(), sizeof(int) * i2.size(), hipMemcpyHostToDevice));
        return EMS_OK;
    };
    EMS_TRY(csr((int)sc.size(), sc, su, s->d_tb_skptr, s->d_tb_scomp, s->d_tb_su, s->d_tb_sidx, s->tb_sk));
    EMS_TRY(csr((int)pc.size(), pc, pu, s->d_tb_pkptr, s->d_tb_pcomp, s->d_tb_pu, s->d_tb_pidx, s->tb_pk));
    return EMS_OK;
}

template <typename T> ems_status run_prepare(ems_fdtd *s, int num_steps, RunPlan &rp, bool fused_setup)
{
    const int tf = s->d.tf;
    // trace slots: saved_electric_field_ resize semantics (old values kept)
    const int nslot_new = num_steps / tf + 1;
    const int nobs = (int)s->obs.size();
    if (nslot_new > s->nslot) {
        std::vector<double> t((size_t)nobs * nslot_new * 3, 0.0);
        for (int o = 0; o < nobs; ++o)
            for (int q = 0; q < s->nslot; ++q)
                for (int c = 0; c < 3; ++c)
                    t[((size_t)o * nslot_new + q) * 3 + c] = s->traces[((size_t)o * s->nslot + q) * 3 + c];
        s->traces.swap(t);
        s->nslot = nslot_new;
    }
    rp = RunPlan();
    rp.t_end = s->t_now;
    // the handle's own device state (CPML, fused-kernel ping-pong set and
    // tiles) is set up even by a zero-step run, which so prepares a handle
    // for ems_fdtd_run_static
    EMS_TRY(setup_pml(s));
    if (fused_setup && use_fused(s)) EMS_TRY(ensure_alt(s));
    if (fused_setup && use_tiles(s)) EMS_TRY(build_tiles(s));
    if (fused_setup && kernel_kind(s) == EMS_KERNEL_TB2) EMS_TRY(tb_streams(s));
    if (num_steps <= 0) return EMS_OK;

    // sources sorted by component (x list, then y, then z; :222-239)
    std::vector<Src> src = s->src;
    std::stable_sort(src.begin(), src.end(), [](const Src &a, const Src &b) { return a.comp < b.comp; });
    const int nsrc = (int)src.size();
    std::vector<int> sc(nsrc);
    std::vector<long> su(nsrc);
    for (int q = 0; q < nsrc; ++q) {
        sc[q] = src[q].comp;
        su[q] = src[q].u;
    }
    // per-step injected values (s(t)*dt)/eps [/scale], t accumulated as current_time_
    std::vector<double> sval((size_t)num_steps * std::max(nsrc, 1), 0.0);
    double t = s->t_now;
    for (int st = 0; st < num_steps; ++st) {
        double w = wavelet_value(s, t);
        for (int q = 0; q < nsrc; ++q) {
            double v = w * s->d.dt / src[q].eps;
            if (src[q].scale != 1.0) v = v / src[q].scale;
            sval[(size_t)st * nsrc + q] = v;
        }
        t += s->d.dt;
    }
    // receiver points
    std::vector<int> pc;
    std::vector<long> pu;
    for (auto &o : s->obs)
        for (int c = 0; c < 3; ++c)
            for (auto &p : o[c]) {
                pc.push_back(c);
                pu.push_back(ref_to_unified(s->g, c, p[0], p[1], p[2]));
            }
    const int npts = (int)pc.size();
    const int nrec = num_steps / tf;
    EMS_TRY(s->d_src_comp.reserve(sizeof(int) * std::max(nsrc, 1)));
    EMS_TRY(s->d_src_u.reserve(sizeof(long) * std::max(nsrc, 1)));
    EMS_TRY(s->d_src_val.reserve(sizeof(double) * sval.size()));
    EMS_TRY(s->d_pt_comp.reserve(sizeof(int) * std::max(npts, 1)));
    EMS_TRY(s->d_pt_u.reserve(sizeof(long) * std::max(npts, 1)));
    EMS_TRY(s->d_rec.reserve(sizeof(double) * std::max((size_t)nrec * npts, (size_t)1)));
    if (nsrc) {
        EMS_HIP(hipMemcpy(s->d_src_comp.p, sc.data(), sizeof(int) * nsrc, hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(s->d_src_u.p, su.data(), sizeof(long) * nsrc, hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(s->d_src_val.p, sval.data(), sizeof(double) * sval.size(), hipMemcpyHostToDevice));
    }
    if (npts) {
        EMS_HIP(hipMemcpy(s->d_pt_comp.p, pc.data(), sizeof(int) * npts, hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(s->d_pt_u.p, pu.data(), sizeof(long) * npts, hipMemcpyHostToDevice));
    }
    if (kernel_kind(s) == EMS_KERNEL_TB2 && num_steps >= 2) EMS_TRY(tb_lists(s, sc, su, pc, pu));
    rp.nsrc = nsrc;
    rp.npts = npts;
    rp.nrec = nrec;
    rp.t_end = t;
    return EMS_OK;
}

// After the device steps of a run: combine the raw receiver samples into the
// traces exactly as the reference (saved += p for each point, then /= count).
ems_status run_finish(ems_fdtd *s, const RunPlan &rp)
{
    s->t_now = rp.t_end;
    const int nobs = (int)s->obs.size();
    if (rp.nrec > 0 && nobs > 0) {
        std::vector<double> raw((size_t)rp.nrec * rp.npts);
        if (rp.npts)
            EMS_HIP(hipMemcpy(raw.data(), s->d_rec.p, sizeof(double) * raw.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < rp.nrec; ++r) {
            const int slot = r + 1;
            int p = 0;
            for (int o = 0; o < nobs; ++o)
                for (int c = 0; c < 3; ++c) {
                    double &dst = s->traces[((size_t)o * s->nslot + slot) * 3 + c];
                    const size_t cnt = s->obs[o][c].size();
                    for (size_t q = 0; q < cnt; ++q, ++p) dst += raw[(size_t)r * rp.npts + p];
                    dst /= (double)cnt;
                }
        }
    }
    return EMS_OK;
}

// Checks of a run on the two-step kernel (tile and array size).
template <typename T> ems_status tb2_check(const ems_fdtd *s)
{
    if (kernel_kind(s) != EMS_KERNEL_TB2) return EMS_OK;
    int txw, tyw, lz;
    tb_tile<T>(s, txw, tyw, lz);
    if (txw * tyw > tb_max_threads<T>()) return invalid("two-step tile: at most 512 threads in FP64");
    if (!tb2_fits(s)) return invalid("two-step kernel: a field array reaches 4 GiB (32-bit offsets); use the split sweeps");
    return EMS_OK;
}

// The launches of a prepared run, all on the handle's stream (the two-step
// kernel's CPML shell also on its side streams, joined back).  ev: when not
// null, 2 (num_steps + 1) events; [2 l, 2 l + 1] are recorded around the Yee
// launch(es) of launch group l (one step, or two for the two-step kernel).
template <typename T>
ems_status run_loop(ems_fdtd *s, int num_steps, const RunPlan &rp, std::vector<hipEvent_t> *ev, int &ngroups)
{
    const int tf = s->d.tf, nsrc = rp.nsrc, npts = rp.npts;
    const T ch = (T)(s->d.dt / s->d.mu / s->d.h);
    const T ce = (T)(s->d.dt / s->d.eps / s->d.h);
    const bool tb2 = kernel_kind(s) == EMS_KERNEL_TB2;
    ngroups = 0;
    for (int step = 1; step <= num_steps;) {
        if (ev) EMS_HIP(hipEventRecord((*ev)[2 * (size_t)ngroups], s->stream));
        const PmlView<T> pm = pml_view<T>(s);
        if (tb2 && step + 1 <= num_steps) {
            // steps `step` and `step + 1`, sources and receivers inside
            TbPost pp{};
            pp.s_kptr = s->d_tb_skptr.as<int>();
            pp.s_comp = s->d_tb_scomp.as<int>();
            pp.s_u = s->d_tb_su.as<long>();
            pp.s_idx = s->d_tb_sidx.as<int>();
            pp.v1 = s->d_src_val.as<double>() + (size_t)(step - 1) * nsrc;
            pp.v2 = s->d_src_val.as<double>() + (size_t)step * nsrc;
            pp.p_kptr = s->d_tb_pkptr.as<int>();
            pp.p_comp = s->d_tb_pcomp.as<int>();
            pp.p_u = s->d_tb_pu.as<long>();
            pp.p_idx = s->d_tb_pidx.as<int>();
            pp.s_k0 = s->tb_sk[0];
            pp.s_k1 = s->tb_sk[1];
            pp.p_k0 = s->tb_pk[0];
            pp.p_k1 = s->tb_pk[1];
            if (npts && step % tf == 0) pp.rec1 = s->d_rec.as<double>() + (size_t)(step / tf - 1) * npts;
            if (npts && (step + 1) % tf == 0) pp.rec2 = s->d_rec.as<double>() + (size_t)((step + 1) / tf - 1) * npts;
            EMS_TRY(launch_tb2<T>(s, ch, ce, pm, pp));
            if (ev) EMS_HIP(hipEventRecord((*ev)[2 * (size_t)ngroups + 1], s->stream));
            EMS_HIP(hipGetLastError());
            ++ngroups;
            step += 2;
            continue;
        }
        launch_step<T>(s, ch, ce, pm);
        const Fields<T> f = fields<T>(s);
        if (ev) EMS_HIP(hipEventRecord((*ev)[2 * (size_t)ngroups + 1], s->stream));
        EMS_HIP(hipGetLastError());
        ++ngroups;
        double *rec = nullptr;
        if (npts && step % tf == 0) rec = s->d_rec.as<double>() + (size_t)(step / tf - 1) * npts;
        if (nsrc || rec) {
            k_post<T><<<1, 256, 0, s->stream>>>(f, nsrc, s->d_src_comp.as<int>(), s->d_src_u.as<long>(),
                                                s->d_src_val.as<double>() + (size_t)(step - 1) * nsrc, npts,
                                                s->d_pt_comp.as<int>(), s->d_pt_u.as<long>(), rec);
            EMS_HIP(hipGetLastError());
        }
        ++step;
    }
    return EMS_OK;
}

// Runs num_steps steps; in timed mode records HIP events around the Yee
// kernels of every step and around the whole sequence.
template <typename T>
ems_status run_impl(ems_fdtd *s, int num_steps, double *total_ms, double *kernel_ms)
{
    RunPlan rp;
    EMS_TRY(run_prepare<T>(s, num_steps, rp, true));
    if (num_steps <= 0) return EMS_OK;
    EMS_TRY(tb2_check<T>(s));
    const bool timed = total_ms != nullptr;
    // events: [2 l, 2 l + 1] around launch group l; the last pair brackets all
    std::vector<hipEvent_t> ev;
    int ngroups = 0;
    if (timed) {
        ev.resize(2 * (size_t)num_steps + 2, nullptr);
        for (auto &e : ev) EMS_HIP(hipEventCreate(&e));
        EMS_HIP(hipEventRecord(ev[2 * (size_t)num_steps], s->stream));
    }
    EMS_TRY(run_loop<T>(s, num_steps, rp, timed ? &ev : nullptr, ngroups));
    if (timed) {
        EMS_HIP(hipEventRecord(ev[2 * (size_t)num_steps + 1], s->stream));
        EMS_HIP(hipEventSynchronize(ev[2 * (size_t)num_steps + 1]));
        double ksum = 0.0;
        for (int st = 0; st < ngroups; ++st) {
            float ms = 0.f;
            EMS_HIP(hipEventElapsedTime(&ms, ev[2 * (size_t)st], ev[2 * (size_t)st + 1]));
            ksum += ms;
        }
        float ms = 0.f;
        EMS_HIP(hipEventElapsedTime(&ms, ev[2 * (size_t)num_steps], ev[2 * (size_t)num_steps + 1]));
        *total_ms = ms;
        *kernel_ms = ksum;
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    EMS_HIP(hipStreamSynchronize(s->stream));
    return run_finish(s, rp);
}

// Autotuner of the two-step kernel: times candidate tiles (the tile shape
// sets how many workgroups share a CU: 64 x 16 FP32 threads run one per CU,
// 64 x 12 and 64 x 8 two or three) and planes per chunk on zeroed scratch
// field and psi sets swapped in for the trials, without sources or
// receivers; the handle's state, traces and time are untouched.  Keeps and
// reports the fastest.
template <typename T> ems_status autotune_tb2_impl(ems_fdtd *s, int trial, int *cols, int *rows, int *planes)
{
    const int kernel0 = s->kernel;
    s->kernel = EMS_KERNEL_TB2;
    RunPlan rp;
    ems_status st0 = run_prepare<T>(s, 0, rp, true);
    if (st0 == EMS_OK && !tb2_fits(s))
        st0 = invalid("two-step kernel: a field array reaches 4 GiB (32-bit offsets)");
    if (st0 != EMS_OK) {
        s->kernel = kernel0;
        return st0;
    }
    // the handle's sets, restored on every exit
    struct Held {
        ems_fdtd *s;
        int kernel0;
        DevBuf f[6], f2[6], psi[12], psi2[12];
        bool held = false;
        ~Held()
        {
            if (held) {
                for (int w = 0; w < 6; ++w) {
                    s->f[w] = std::move(f[w]);
                    s->f2[w] = std::move(f2[w]);
                }
                for (int q = 0; q < 12; ++q) {
                    s->psi[q] = std::move(psi[q]);
                    s->psi2[q] = std::move(psi2[q]);
                }
            }
            s->kernel = kernel0;
        }
    } h{s, kernel0};
    DevBuf sf[6], sf2[6], sp[12], sp2[12];
    const size_t fb = s->nalloc * s->esz;
    for (int w = 0; w < 6; ++w) {
        EMS_TRY(sf[w].reserve(fb));
        EMS_TRY(sf2[w].reserve(fb));
        EMS_HIP(hipMemset(sf[w].p, 0, fb));
        EMS_HIP(hipMemset(sf2[w].p, 0, fb));
    }
    for (int q = 0; q < 12; ++q) {
        if (s->psi[q].bytes) {
            EMS_TRY(sp[q].reserve(s->psi[q].bytes));
            EMS_HIP(hipMemset(sp[q].p, 0, s->psi[q].bytes));
        }
        if (s->psi2[q].bytes) {
            EMS_TRY(sp2[q].reserve(s->psi2[q].bytes));
            EMS_HIP(hipMemset(sp2[q].p, 0, s->psi2[q].bytes));
        }
    }
    for (int w = 0; w < 6; ++w) {
        h.f[w] = std::move(s->f[w]);
        h.f2[w] = std::move(s->f2[w]);
        s->f[w] = std::move(sf[w]);
        s->f2[w] = std::move(sf2[w]);
    }
    for (int q = 0; q < 12; ++q) {
        h.psi[q] = std::move(s->psi[q]);
        h.psi2[q] = std::move(s->psi2[q]);
        s->psi[q] = std::move(sp[q]);
        s->psi2[q] = std::move(sp2[q]);
    }
    h.held = true;
    const T ch = (T)(s->d.dt / s->d.mu / s->d.h);
    const T ce = (T)(s->d.dt / s->d.eps / s->d.h);
    TbPost pp{};
    pp.s_k0 = pp.p_k0 = 0;
    pp.s_k1 = pp.p_k1 = -1;  // no sources, no receivers
    const int c4[][2] = {{64, 16}, {64, 12}, {64, 8}, {68, 15}};
    const int c8[][2] = {{64, 8}, {64, 6}, {36, 14}};
    const int lzs[] = {8, 12, 16};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    EMS_HIP(hipEventCreate(&e0));
    EMS_HIP(hipEventCreate(&e1));
    struct Ev {
        hipEvent_t a, b;
        ~Ev()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev{e0, e1};
    const int nc = sizeof(T) == 4 ? 4 : 3;
    const int pairs = std::max(1, trial / 2);
    float best = 1e30f;
    int bc = 0, br = 0, bl = 0;
    for (int c = 0; c < nc; ++c)
        for (int lz : lzs) {
            s->tb_txw = sizeof(T) == 4 ? c4[c][0] : c8[c][0];
            s->tb_tyw = sizeof(T) == 4 ? c4[c][1] : c8[c][1];
            s->tb_lz = lz;
            const PmlView<T> pm = pml_view<T>(s);
            EMS_TRY(launch_tb2<T>(s, ch, ce, pm, pp));  // warm-up
            EMS_HIP(hipEventRecord(e0, s->stream));
            for (int q = 0; q < pairs; ++q) EMS_TRY(launch_tb2<T>(s, ch, ce, pm, pp));
            EMS_HIP(hipEventRecord(e1, s->stream));
            EMS_HIP(hipEventSynchronize(e1));
            EMS_HIP(hipGetLastError());
            float ms = 0.f;
            EMS_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                bc = s->tb_txw;
                br = s->tb_tyw;
                bl = lz;
            }
        }
    EMS_HIP(hipStreamSynchronize(s->stream));
    s->tb_txw = bc;
    s->tb_tyw = br;
    s->tb_lz = bl;
    if (cols) *cols = bc;
    if (rows) *rows = br;
    if (planes) *planes = bl;
    return EMS_OK;
}

ems_status set_material_impl(ems_fdtd *s)
{
    const GridView &g = s->g;
    const bool has_eps = !s->eps_r.empty(), has_sig = !s->sigma.empty();
    // source cells see the local material
    for (auto &q : s->src) {
        size_t m = (size_t)q.i + (size_t)(g.nx + 1) * (q.j + (size_t)(g.ny + 1) * q.k);
        bool in = q.i <= g.nx && q.j <= g.ny && q.k <= g.nz;
        double er = (has_eps && in) ? s->eps_r[m] : 1.0;
        double sg = (has_sig && in) ? s->sigma[m] : 0.0;
        q.eps = s->d.eps * er;
        q.scale = 1.0 + sg * s->d.dt / (2.0 * q.eps);
    }
    if (!has_eps && !has_sig) {
        s->mat_mode = 0;
        s->ca.release();
        s->cb.release();
        return EMS_OK;
    }
    std::vector<double> ca(s->nalloc, 1.0), cb(s->nalloc, 0.0);
    for (int K = 0; K <= g.nz; ++K)
        for (int J = 0; J <= g.ny; ++J)
            for (int I = 0; I <= g.nx; ++I) {
                size_t m = (size_t)I + (size_t)(g.nx + 1) * (J + (size_t)(g.ny + 1) * K);
                double er = has_eps ? s->eps_r[m] : 1.0;
                double sg = has_sig ? s->sigma[m] : 0.0;
                double e = s->d.eps * er;
                double hs = sg * s->d.dt / (2.0 * e);
                long u = uidx(g, I, J, K);
                ca[u] = (1.0 - hs) / (1.0 + hs);
                cb[u] = s->d.dt / e / s->d.h / (1.0 + hs);
            }
    s->mat_mode = has_sig ? 2 : 1;
    if (s->d.precision == EMS_F32) {
        EMS_TRY(upload_as<float>(s->cb, cb));
        if (has_sig) EMS_TRY(upload_as<float>(s->ca, ca));
    } else {
        EMS_TRY(upload_as<double>(s->cb, cb));
        if (has_sig) EMS_TRY(upload_as<double>(s->ca, ca));
    }
    if (!has_sig) s->ca.release();
    return EMS_OK;
}

template <typename T> ems_status get_field_t(ems_fdtd *s, int w, double *out)
{
    std::vector<T> host(s->nalloc);
    EMS_HIP(hipStreamSynchronize(s->stream));
    EMS_HIP(hipMemcpy(host.data(), s->f[w].p, s->nalloc * sizeof(T), hipMemcpyDeviceToHost));
    int d[3], sh[3];
    ref_dims(s->g, w, d, sh);
    size_t n = 0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) out[n++] = (double)host[uidx(s->g, i + sh[0], j + sh[1], k + sh[2])];
    return EMS_OK;
}

template <typename T> ems_status set_field_t(ems_fdtd *s, int w, const double *in)
{
    std::vector<T> host(s->nalloc);
    EMS_HIP(hipStreamSynchronize(s->stream));
    EMS_HIP(hipMemcpy(host.data(), s->f[w].p, s->nalloc * sizeof(T), hipMemcpyDeviceToHost));
    int d[3], sh[3];
    ref_dims(s->g, w, d, sh);
    size_t n = 0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) host[uidx(s->g, i + sh[0], j + sh[1], k + sh[2])] = (T)in[n++];
    EMS_HIP(hipMemcpy(s->f[w].p, host.data(), s->nalloc * sizeof(T), hipMemcpyHostToDevice));
    // keep the ping-pong partner identical: entries no kernel writes (the
    // ghost ring) must hold the same values in both buffers
    if (s->alt_ready)
        EMS_HIP(hipMemcpy(s->f2[w].p, host.data(), s->nalloc * sizeof(T), hipMemcpyHostToDevice));
    return EMS_OK;
}

}  // namespace

// =========================================================================
// Misfit gradient by the discrete adjoint with sqrt(N) checkpointing
// (no reference counterpart).
// =========================================================================
struct Adj {
    DevBuf fin;                 // the final forward state
    void *segp = nullptr;       // this shot's E snapshots in the gradient store: [slot][3][nalloc]
    DevBuf e[3], h[3];          // adjoint fields (e updated in place)
    DevBuf pe[2][6], ph[2][6];  // psi adjoints (ping-pong), compact slabs
    DevBuf fac, grad;           // 4 x nmat derivative factors, 2 x nmat gradient (FP64)
    DevBuf sums;                // gradient sums P, Q (FP64, unified grid)
    DevBuf tab;                 // shot tables of the adjoint sweeps, one per psi parity
    long fac_version = -1;      // material version the factors were built for
    DevBuf w;                   // residual weights [rows][npts]
};

ems_fdtd::~ems_fdtd()
{
    delete adj;
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    for (int q = 0; q < 3; ++q) {
        if (ev_join[q]) (void)hipEventDestroy(ev_join[q]);
        if (side[q]) (void)hipStreamDestroy(side[q]);
    }
}

namespace {

size_t state_bytes(const ems_fdtd *s)
{
    size_t b = 6 * s->nalloc * s->esz;
    for (int p = 0; p < 12; ++p) b += s->psi[p].bytes;
    return b;
}

// The state (6 fields, then the 12 psi slabs that exist) to or from dst.
ems_status copy_state(ems_fdtd *s, void *dst, bool save)
{
    char *q = static_cast<char *>(dst);
    const size_t fb = s->nalloc * s->esz;
    for (int w = 0; w < 6; ++w, q += fb)
        EMS_HIP(hipMemcpyAsync(save ? q : s->f[w].p, save ? s->f[w].p : q, fb, hipMemcpyDeviceToDevice, s->stream));
    for (int p = 0; p < 12; ++p) {
        if (!s->psi[p].bytes) continue;
        EMS_HIP(hipMemcpyAsync(save ? q : s->psi[p].p, save ? s->psi[p].p : q, s->psi[p].bytes,
                               hipMemcpyDeviceToDevice, s->stream));
        q += s->psi[p].bytes;
    }
    return EMS_OK;
}

// The current E of the shots `mem` (one grid) into snapshot slot `slot`,
// one launch per kShotArgs shots.
template <typename T> ems_status snap_shots(const std::vector<ems_fdtd *> &mem, int slot, hipStream_t st)
{
    const long nalloc = (long)mem[0]->nalloc;
    const long nv = nalloc * (long)sizeof(T) / 16;
    const unsigned gx = (unsigned)std::min<long>((nv + 255) / 256, 512);
    for (size_t q0 = 0; q0 < mem.size(); q0 += kShotArgs) {
        const int cnt = (int)std::min<size_t>(kShotArgs, mem.size() - q0);
        SnapArgs<T> a{};
        for (int q = 0; q < cnt; ++q) {
            const ems_fdtd *s = mem[q0 + q];
            for (int c = 0; c < 3; ++c) a.e[q][c] = s->f[c].as<T>();
            a.seg[q] = static_cast<T *>(s->adj->segp);
        }
        k_snap<T><<<dim3(gx, 3u * cnt), 256, 0, st>>>(a, nalloc, slot);
    }
    EMS_HIP(hipGetLastError());
    return EMS_OK;
}

// Residual injection arguments of the shots `mem`, kShotArgs per launch
// (residual weights w and point lists in place).
template <typename T>
std::vector<InjectArgs<T>> inject_args(const std::vector<ems_fdtd *> &mem, const std::vector<int> &npts,
                                       const std::vector<int> &serial)
{
    std::vector<InjectArgs<T>> out;
    for (size_t q0 = 0; q0 < mem.size(); q0 += kShotArgs) {
        const int cnt = (int)std::min<size_t>(kShotArgs, mem.size() - q0);
        InjectArgs<T> a{};
        for (int q = 0; q < cnt; ++q) {
            const ems_fdtd *s = mem[q0 + q];
            for (int c = 0; c < 3; ++c) a.e[q][c] = s->adj->e[c].as<T>();
            a.pt_comp[q] = s->d_pt_comp.as<int>();
            a.pt_u[q] = s->d_pt_u.as<long>();
            a.w[q] = s->adj->w.as<double>();
            a.npts[q] = npts[q0 + q];
            a.serial[q] = serial[q0 + q];
        }
        out.push_back(a);
    }
    return out;
}

template <typename T> ems_status inject_shots(const std::vector<InjectArgs<T>> &args, int nshots, int row, hipStream_t st)
{
    for (size_t i = 0; i < args.size(); ++i) {
        const int cnt = (int)std::min<size_t>(kShotArgs, (size_t)nshots - i * kShotArgs);
        k_adj_inject_shots<T><<<cnt, 256, 0, st>>>(args[i], row);
    }
    EMS_HIP(hipGetLastError());
    return EMS_OK;
}

// ---------------------------------------------------------------------------
// Checkpoint schedule of a gradient over N steps.  The adjoint of step n
// needs E^n and E^{n+1}; they come from a snapshot store that holds the E of
// Ks+1 consecutive steps per shot, filled by running a segment of Ks steps
// forward from a checkpoint (full state).  Two levels:
//   level 1: checkpoints every K1 steps, written in the first forward pass;
//   level 2: checkpoints every Ks steps inside one level-1 segment, written
//            when that segment is (re)run.
// Backwards over level-1 segments: rerun the segment from its checkpoint
// writing level-2 checkpoints (the last segment is not rerun: the first pass
// wrote them), then backwards over its Ks-step sub-segments: rerun each from
// its checkpoint into the snapshot store (the last one is not rerun) and run
// its adjoint steps.  K1 = N is the one-level scheme.  The last segment of a
// level needs no checkpoint.
// ---------------------------------------------------------------------------
struct GradPlan {
    int K1 = 1, Ks = 1;
};

int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Forward steps of a plan (first pass included) and its store bytes, for
// snapshot bytes `snap` and checkpoint bytes `ck` per step / state of all shots.
double plan_work(int N, const GradPlan &p)
{
    const int n1 = ceil_div(N, p.K1), last1 = N - (n1 - 1) * p.K1;
    auto l2 = [&](int len) { return (double)(ceil_div(len, p.Ks) - 1) * p.Ks; };  // steps rerun inside a segment
    return (double)N + (double)(N - last1) + (n1 - 1) * l2(p.K1) + l2(last1);
}
double plan_bytes(int N, const GradPlan &p, double snap, double ck)
{
    const int n1 = ceil_div(N, p.K1);
    const int n2 = ceil_div(std::min(p.K1, N), p.Ks);
    return (p.Ks + 1.0) * snap + ((n1 - 1) + (n2 - 1)) * ck;
}

// The plan with the fewest forward steps (then the fewest bytes) within
// `budget`; a set interval gives the one-level plan with Ks = interval.
// Nothing fits: the plan with the fewest bytes (its allocation then reports
// the shortage).
GradPlan plan_gradient(int N, int interval, double snap, double ck, double budget)
{
    GradPlan best, least;
    if (interval > 0) {
        best.K1 = N;
        best.Ks = std::min(interval, N);
        return best;
    }
    double bw = 0.0, bb = 0.0, lb = 0.0;
    bool found = false, any = false;
    auto consider = [&](int K1, int Ks) {
        const GradPlan p{K1, Ks};
        const double b = plan_bytes(N, p, snap, ck);
        if (!any || b < lb) {
            least = p;
            lb = b;
            any = true;
        }
        if (b > budget) return;
        const double w = plan_work(N, p);
        if (!found || w < bw || (w == bw && b < bb)) {
            best = p;
            bw = w;
            bb = b;
            found = true;
        }
    };
    for (int Ks = 1; Ks <= N; ++Ks) {
        consider(N, Ks);
        for (long K1 = 2L * Ks; K1 < N; K1 += Ks) consider((int)K1, Ks);
    }
    return found ? best : least;
}

// One store per GPU for the snapshots and checkpoints of the gradient that
// runs on it: grown on demand and kept for later calls (allocating and
// freeing stores of tens of GB per call costs more than the calls' host
// work); gradients on one GPU take turns on it.
struct GradStore {
    std::mutex mu;
    DevBuf buf;  // one call's snapshots ([shot][slot][3][nalloc]), then its checkpoints
};

constexpr int kStoreDevices = 64;
GradStore g_grad_store[kStoreDevices];

ems_status grad_store(GradStore **out)
{
    int dev = -1;
    EMS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kStoreDevices) return invalid("gradient: device index out of range");
    *out = &g_grad_store[dev];
    return EMS_OK;
}

// Plan and store of a gradient over N steps of the shots `mem` (one grid),
// with the store's lock held by the caller: the budget is 3/4 of the device
// memory that is free or already held by the store, or the first shot's cap
// if that is lower.  Sets every shot's
// snapshot base; *ck_base = the checkpoint area ([idx][shot] states, level-1
// indices first, then n1-1 .. level-2).
ems_status grad_plan_for(const GradStore &gs, const ems_fdtd *s, int shots, int N, GradPlan &plan, double &bytes)
{
    const double snap = shots * 3.0 * (double)s->nalloc * (double)s->esz;
    const double ck = shots * (double)state_bytes(s);
    size_t fr = 0, tot = 0;
    EMS_HIP(hipMemGetInfo(&fr, &tot));
    double budget = 0.75 * ((double)fr + (double)gs.buf.bytes);
    if (s->grad_mem > 0) budget = std::min(budget, (double)s->grad_mem);
    plan = plan_gradient(N, s->ckpt_interval, snap, ck, budget);
    bytes = plan_bytes(N, plan, snap, ck);
    if (s->grad_mem > 0 && s->ckpt_interval <= 0 && bytes > (double)s->grad_mem)
        return invalid("gradient: no checkpoint schedule of " + std::to_string(N) + " steps fits the memory cap of " +
                       std::to_string(s->grad_mem) + " bytes");
    return EMS_OK;
}

ems_status grad_prepare(GradStore &gs, const std::vector<ems_fdtd *> &mem, int N, GradPlan &plan, char **ck_base)
{
    const ems_fdtd *s = mem[0];
    double bytes = 0.0;
    EMS_TRY(grad_plan_for(gs, s, (int)mem.size(), N, plan, bytes));
    const size_t per_shot = (size_t)(plan.Ks + 1) * 3 * s->nalloc * s->esz;
    const int n1 = ceil_div(N, plan.K1), n2 = ceil_div(std::min(plan.K1, N), plan.Ks);
    const size_t nck = (size_t)std::max(0, n1 - 1) + (size_t)std::max(0, n2 - 1);
    const size_t snap_bytes = (per_shot * mem.size() + 255) / 256 * 256;
    const size_t want = snap_bytes + nck * mem.size() * state_bytes(s);
    if (gs.buf.bytes < want) {
        // grow (the old store is released first, so the two are never held together)
        gs.buf.release();
        if (gs.buf.reserve(want) != EMS_OK) {
            set_error("gradient: the snapshot and checkpoint store (" + std::to_string(want >> 20) +
                      " MiB for segments of " + std::to_string(plan.Ks) + " steps) does not fit in device memory");
            return EMS_ERR_NO_MEMORY;
        }
    }
    for (size_t q = 0; q < mem.size(); ++q) mem[q]->adj->segp = static_cast<char *>(gs.buf.p) + q * per_shot;
    *ck_base = static_cast<char *>(gs.buf.p) + snap_bytes;
    return EMS_OK;
}

// Drives the schedule.  Ops:
//   adv(n, record, slot): forward step n -> n+1 of every shot (receivers when
//       record), then E^{n+1} into snapshot slot `slot` when slot >= 0
//   snap(slot): the current E of every shot into snapshot slot `slot`
//   save(idx) / load(idx): the state of every shot to / from checkpoint idx
//   adjoint(n, slot): adjoint step n, E^n in slot `slot`, E^{n+1} in slot + 1
// adv2 of schedules without paired steps
struct NoPairs {};

template <typename Ops> struct GradSchedule {
    Ops &op;
    GradPlan p;
    int N;
    int l2_base() const { return std::max(0, ceil_div(N, p.K1) - 1); }
    // steps [a, b) without snapshots: in pairs where the ops have them
    ems_status plain(int a, int b, bool record)
    {
        int n = a;
        if constexpr (!std::is_same<decltype(op.adv2), NoPairs>::value)
            for (; n + 1 < b; n += 2) EMS_TRY(op.adv2(n, record));
        for (; n < b; ++n) EMS_TRY(op.adv(n, record, -1));
        return EMS_OK;
    }
    // a level-1 segment [a, b) from its start state: level-2 checkpoints and
    // the last sub-segment's snapshots
    ems_status level2_forward(int a, int b, bool record)
    {
        const int n2 = ceil_div(b - a, p.Ks);
        for (int q = 0; q < n2; ++q) {
            const int c = a + q * p.Ks, d = std::min(b, c + p.Ks);
            const bool last = q == n2 - 1;
            if (!last) {
                EMS_TRY(op.save(l2_base() + q));
                EMS_TRY(plain(c, d, record));
                continue;
            }
            EMS_TRY(op.snap(0));
            for (int n = c; n < d; ++n) EMS_TRY(op.adv(n, record, n + 1 - c));
        }
        return EMS_OK;
    }
    ems_status first_pass()
    {
        const int n1 = ceil_div(N, p.K1);
        for (int q = 0; q < n1; ++q) {
            const int a = q * p.K1, b = std::min(N, a + p.K1);
            if (q < n1 - 1) {
                EMS_TRY(op.save(q));
                EMS_TRY(plain(a, b, true));
            } else {
                EMS_TRY(level2_forward(a, b, true));
            }
        }
        return EMS_OK;
    }
    ems_status backward()
    {
        const int n1 = ceil_div(N, p.K1);
        for (int q1 = n1 - 1; q1 >= 0; --q1) {
            const int a = q1 * p.K1, b = std::min(N, a + p.K1);
            if (q1 < n1 - 1) {
                EMS_TRY(op.load(q1));
                EMS_TRY(level2_forward(a, b, false));
            }
            const int n2 = ceil_div(b - a, p.Ks);
            for (int q2 = n2 - 1; q2 >= 0; --q2) {
                const int c = a + q2 * p.Ks, d = std::min(b, c + p.Ks);
                if (q2 < n2 - 1) {
                    EMS_TRY(op.load(l2_base() + q2));
                    EMS_TRY(op.snap(0));
                    for (int n = c; n < d; ++n) EMS_TRY(op.adv(n, false, n + 1 - c));
                }
                for (int n = d - 1; n >= c; --n) EMS_TRY(op.adjoint(n, n - c));
            }
        }
        return EMS_OK;
    }
};

// Material versions, unique over all handles (keys of cached gradient factors).
long next_version()
{
    static std::atomic<long> versions{0};
    return ++versions;
}

template <typename T> AdjPml<T> adj_pml(ems_fdtd *s, int cur)
{
    AdjPml<T> a{};
    a.P = s->npml > 0 ? s->npml : 0;
    a.W = 2 * a.P + 1;
    if (s->npml <= 0) return a;
    const PmlView<T> v = pml_view<T>(s);
    Adj *A = s->adj;
    for (int ax = 0; ax < 3; ++ax) {
        a.hp[ax] = v.hp[ax];
        a.ep[ax] = v.ep[ax];
        for (int q = 0; q < 2; ++q) {
            a.pe_old[ax][q] = A->pe[cur][ax * 2 + q].as<T>();
            a.pe_new[ax][q] = A->pe[1 - cur][ax * 2 + q].as<T>();
            a.ph_old[ax][q] = A->ph[cur][ax * 2 + q].as<T>();
            a.ph_new[ax][q] = A->ph[1 - cur][ax * 2 + q].as<T>();
        }
    }
    return a;
}

// Adjoint state of one shot for the sweeps, psi adjoints of parity `par` old.
template <typename T> AdjShot<T> adj_shot(ems_fdtd *s, int par)
{
    Adj *A = s->adj;
    AdjShot<T> sh{};
    sh.e = Adj3<T>{{A->e[0].as<T>(), A->e[1].as<T>(), A->e[2].as<T>()}};
    sh.hb = Adj3<T>{{A->h[0].as<T>(), A->h[1].as<T>(), A->h[2].as<T>()}};
    sh.seg = static_cast<T *>(A->segp);
    sh.a = adj_pml<T>(s, par);
    return sh;
}

// Launch geometry of the adjoint sweeps: the split sweeps' tiling with 4-row
// workgroups and ADJ_LZ planes per thread.
struct AdjGeom {
    int P, npw, nbx, nby, ntiles;
    unsigned grd;
};

AdjGeom adj_geom(const GridView &g, int npml)
{
    AdjGeom a{};
    a.P = npml > 0 ? npml : 0;
    a.npw = a.P > 0 ? (2 * a.P + 1 + 63) / 64 : 0;
    a.nbx = split_row_waves(g, a.P);
    a.nby = (g.ny + 1 + 3) / 4;
    a.ntiles = a.nby * ((g.nz + 1 + ADJ_LZ - 1) / ADJ_LZ);
    a.grd = 8u * (unsigned)a.nbx * (unsigned)((a.ntiles + 7) / 8);  // see split_tile
    return a;
}

// Both adjoint sweeps of one step for the ns shots of the device table sh,
// accumulating the gradient sums of E snapshot slots `slot` and slot + 1.
template <typename T, int MAT, bool PML>
void adj_sweeps_t(const GridView &g, const AdjGeom &ag, hipStream_t st, T ce, T nch, const AdjShot<T> *sh, int ns,
                  const T *ca, const T *cb, double *gp, double *gq, long nalloc, int slot)
{
    const dim3 blk(64, 4, 1);
    k_adj_h<T, MAT, PML><<<ag.grd, blk, 0, st>>>(g, ce, sh, ns, cb, ag.P, ag.npw, ag.nbx, ag.nby, ag.ntiles);
    k_adj_e<T, MAT, PML, true><<<ag.grd, blk, 0, st>>>(g, nch, sh, ns, ca, gp, gq, nalloc, slot, ag.P, ag.npw, ag.nbx,
                                                       ag.nby, ag.ntiles);
}

template <typename T>
void adj_sweeps(int mat, bool pml, const GridView &g, const AdjGeom &ag, hipStream_t st, T ce, T nch,
                const AdjShot<T> *sh, int ns, const T *ca, const T *cb, double *gp, double *gq, long nalloc, int slot)
{
    switch (mat * 2 + (pml ? 1 : 0)) {
    case 0: adj_sweeps_t<T, 0, false>(g, ag, st, ce, nch, sh, ns, ca, cb, gp, gq, nalloc, slot); break;
    case 1: adj_sweeps_t<T, 0, true>(g, ag, st, ce, nch, sh, ns, ca, cb, gp, gq, nalloc, slot); break;
    case 2: adj_sweeps_t<T, 1, false>(g, ag, st, ce, nch, sh, ns, ca, cb, gp, gq, nalloc, slot); break;
    case 3: adj_sweeps_t<T, 1, true>(g, ag, st, ce, nch, sh, ns, ca, cb, gp, gq, nalloc, slot); break;
    case 4: adj_sweeps_t<T, 2, false>(g, ag, st, ce, nch, sh, ns, ca, cb, gp, gq, nalloc, slot); break;
    default: adj_sweeps_t<T, 2, true>(g, ag, st, ce, nch, sh, ns, ca, cb, gp, gq, nalloc, slot); break;
    }
}

ems_status zero(DevBuf &b, size_t bytes, hipStream_t st)
{
    EMS_TRY(b.reserve(bytes));
    EMS_HIP(hipMemsetAsync(b.p, 0, bytes, st));
    return EMS_OK;
}

// Zero adjoint fields and psi adjoints (both parities) of one shot.
ems_status adj_zero(ems_fdtd *s, hipStream_t st)
{
    Adj *A = s->adj;
    const size_t fb = s->nalloc * s->esz;
    for (int c = 0; c < 3; ++c) {
        EMS_TRY(zero(A->e[c], fb, st));
        EMS_TRY(zero(A->h[c], fb, st));
    }
    if (s->npml > 0)
        for (int ax = 0; ax < 3; ++ax)
            for (int q = 0; q < 2; ++q)
                for (int b = 0; b < 2; ++b) {
                    EMS_TRY(zero(A->pe[b][ax * 2 + q], s->psi[ax * 4 + 2 + q].bytes, st));
                    EMS_TRY(zero(A->ph[b][ax * 2 + q], s->psi[ax * 4 + q].bytes, st));
                }
    return EMS_OK;
}

// Material derivative factors (FP64, material grid):
//   s = sigma dt/(2 eps), dCa/deps_r = 2 s/(eps_r (1+s)^2), (dCb/deps_r)/Cb = -1/(eps_r (1+s)),
//   dCa/dsigma = -dt/(eps (1+s)^2), (dCb/dsigma)/Cb = -dt/(2 eps (1+s)).
// The source amplitude dt/(eps_src (1+s_src)) equals h Cb at the source
// cell, so with E^{n+1} - Ca E^n = Cb curl H - src its derivative is carried
// by the (dCb/Cb)(E^{n+1} - Ca E^n) term.
std::vector<double> material_factors(const ems_fdtd *s)
{
    const GridView &g = s->g;
    const size_t nm = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    std::vector<double> f(4 * nm);
    for (size_t m = 0; m < nm; ++m) {
        const double er = s->eps_r.empty() ? 1.0 : s->eps_r[m];
        const double sg = s->sigma.empty() ? 0.0 : s->sigma[m];
        const double e = s->d.eps * er;
        const double hs = sg * s->d.dt / (2.0 * e);
        const double op = 1.0 + hs;
        f[m] = 2.0 * hs / (er * op * op);
        f[nm + m] = -1.0 / (er * op);
        f[2 * nm + m] = -s->d.dt / (e * op * op);
        f[3 * nm + m] = -s->d.dt / (2.0 * e * op);
    }
    return f;
}

template <typename A, typename B, typename C, typename D, typename E, typename P = NoPairs> struct GradOps {
    A adv;
    B snap;
    C save;
    D load;
    E adjoint;
    P adv2;  // steps n, n+1 without snapshots (the two-step kernel), or NoPairs
};
template <typename A, typename B, typename C, typename D, typename E>
GradOps<A, B, C, D, E> grad_ops(A a, B b, C c, D d, E e)
{
    return GradOps<A, B, C, D, E>{a, b, c, d, e, NoPairs{}};
}
template <typename A, typename B, typename C, typename D, typename E, typename P>
GradOps<A, B, C, D, E, P> grad_ops(A a, B b, C c, D d, E e, P p)
{
    return GradOps<A, B, C, D, E, P>{a, b, c, d, e, p};
}

template <typename T>
ems_status gradient_impl(ems_fdtd *s, int N, const double *d_obs, double *misfit, double *g_eps, double *g_sig)
{
    const GridView &g = s->g;
    const int tf = s->d.tf;
    EMS_TRY(setup_pml(s));
    if (!s->adj) s->adj = new Adj();
    Adj *A = s->adj;
    hipStream_t st = s->stream;
    // zero forward state at t = 0
    for (int w = 0; w < 6; ++w) EMS_HIP(hipMemsetAsync(s->f[w].p, 0, s->nalloc * s->esz, st));
    for (int p = 0; p < 12; ++p)
        if (s->psi[p].bytes) EMS_HIP(hipMemsetAsync(s->psi[p].p, 0, s->psi[p].bytes, st));
    if (s->alt_ready)
        for (int w = 0; w < 6; ++w) EMS_HIP(hipMemsetAsync(s->f2[w].p, 0, s->nalloc * s->esz, st));
    // sources (x, then y, then z) and their per-step values
    std::vector<Src> src = s->src;
    std::stable_sort(src.begin(), src.end(), [](const Src &a, const Src &b) { return a.comp < b.comp; });
    const int nsrc = (int)src.size();
    std::vector<int> sc(nsrc);
    std::vector<long> su(nsrc);
    for (int q = 0; q < nsrc; ++q) {
        sc[q] = src[q].comp;
        su[q] = src[q].u;
    }
    std::vector<double> sval((size_t)N * std::max(nsrc, 1), 0.0);
    double t = 0.0;
    for (int n = 0; n < N; ++n) {
        const double wv = wavelet_value(s, t);
        for (int q = 0; q < nsrc; ++q) {
            double v = wv * s->d.dt / src[q].eps;
            if (src[q].scale != 1.0) v = v / src[q].scale;
            sval[(size_t)n * nsrc + q] = v;
        }
        t += s->d.dt;
    }
    std::vector<int> pc;
    std::vector<long> pu;
    for (auto &o : s->obs)
        for (int c = 0; c < 3; ++c)
            for (auto &p : o[c]) {
                pc.push_back(c);
                pu.push_back(ref_to_unified(g, c, p[0], p[1], p[2]));
            }
    const int npts = (int)pc.size();
    const int nrec = N / tf;
    EMS_TRY(s->d_src_comp.reserve(sizeof(int) * std::max(nsrc, 1)));
    EMS_TRY(s->d_src_u.reserve(sizeof(long) * std::max(nsrc, 1)));
    EMS_TRY(s->d_src_val.reserve(sizeof(double) * sval.size()));
    EMS_TRY(s->d_pt_comp.reserve(sizeof(int) * std::max(npts, 1)));
    EMS_TRY(s->d_pt_u.reserve(sizeof(long) * std::max(npts, 1)));
    EMS_TRY(s->d_rec.reserve(sizeof(double) * std::max((size_t)nrec * npts, (size_t)1)));
    if (nsrc) {
        EMS_HIP(hipMemcpy(s->d_src_comp.p, sc.data(), sizeof(int) * nsrc, hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(s->d_src_u.p, su.data(), sizeof(long) * nsrc, hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(s->d_src_val.p, sval.data(), sizeof(double) * sval.size(), hipMemcpyHostToDevice));
    }
    if (npts) {
        EMS_HIP(hipMemcpy(s->d_pt_comp.p, pc.data(), sizeof(int) * npts, hipMemcpyHostToDevice));
        EMS_HIP(hipMemcpy(s->d_pt_u.p, pu.data(), sizeof(long) * npts, hipMemcpyHostToDevice));
    }
    const T ch = (T)(s->d.dt / s->d.mu / s->d.h);
    const T ce = (T)(s->d.dt / s->d.eps / s->d.h);
    if (use_fused(s)) EMS_TRY(ensure_alt(s));
    if (use_tiles(s)) EMS_TRY(build_tiles(s));
    // one forward step n -> n+1 (sources; receivers when `record`)
    auto fstep = [&](int n, bool record) -> ems_status {
        const PmlView<T> pm = pml_view<T>(s);
        launch_step<T>(s, ch, ce, pm);
        const Fields<T> f = fields<T>(s);
        double *rec = nullptr;
        if (record && npts && (n + 1) % tf == 0) rec = s->d_rec.as<double>() + (size_t)((n + 1) / tf - 1) * npts;
        if (nsrc || rec)
            k_post<T><<<1, 256, 0, st>>>(f, nsrc, s->d_src_comp.as<int>(), s->d_src_u.as<long>(),
                                         s->d_src_val.as<double>() + (size_t)n * nsrc, npts, s->d_pt_comp.as<int>(),
                                         s->d_pt_u.as<long>(), rec);
        EMS_HIP(hipGetLastError());
        return EMS_OK;
    };
    // steps n, n+1 by the two-step kernel where a single run would take it
    // (the same bits as two single steps; sources and receivers inside)
    const bool tb2 = kernel_kind(s) == EMS_KERNEL_TB2 && N >= 2;
    if (tb2) {
        EMS_TRY(tb2_check<T>(s));
        EMS_TRY(tb_streams(s));
        EMS_TRY(tb_lists(s, sc, su, pc, pu));
    }
    auto fpair = [&](int n, bool record) -> ems_status {
        if (!tb2) {
            EMS_TRY(fstep(n, record));
            return fstep(n + 1, record);
        }
        TbPost pp{};
        pp.s_kptr = s->d_tb_skptr.as<int>();
        pp.s_comp = s->d_tb_scomp.as<int>();
        pp.s_u = s->d_tb_su.as<long>();
        pp.s_idx = s->d_tb_sidx.as<int>();
        pp.v1 = s->d_src_val.as<double>() + (size_t)n * nsrc;
        pp.v2 = s->d_src_val.as<double>() + (size_t)(n + 1) * nsrc;
        pp.p_kptr = s->d_tb_pkptr.as<int>();
        pp.p_comp = s->d_tb_pcomp.as<int>();
        pp.p_u = s->d_tb_pu.as<long>();
        pp.p_idx = s->d_tb_pidx.as<int>();
        pp.s_k0 = s->tb_sk[0];
        pp.s_k1 = s->tb_sk[1];
        pp.p_k0 = s->tb_pk[0];
        pp.p_k1 = s->tb_pk[1];
        if (record && npts && (n + 1) % tf == 0) pp.rec1 = s->d_rec.as<double>() + (size_t)((n + 1) / tf - 1) * npts;
        if (record && npts && (n + 2) % tf == 0) pp.rec2 = s->d_rec.as<double>() + (size_t)((n + 2) / tf - 1) * npts;
        EMS_TRY(launch_tb2<T>(s, ch, ce, pml_view<T>(s), pp));
        EMS_HIP(hipGetLastError());
        return EMS_OK;
    };
    // gradient buffers outside the store first: the store takes its share of
    // the memory that is left
    const size_t sb = state_bytes(s);
    const size_t nm = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    EMS_TRY(A->fin.reserve(sb));
    EMS_TRY(A->grad.reserve(sizeof(double) * 2 * nm));
    EMS_TRY(A->sums.reserve(sizeof(double) * 2 * s->nalloc));
    EMS_TRY(A->fac.reserve(sizeof(double) * 4 * nm));
    EMS_TRY(A->w.reserve(sizeof(double) * (size_t)std::max(nrec, 1) * std::max(npts, 1)));
    EMS_TRY(adj_zero(s, st));
    GradStore *gs = nullptr;
    EMS_TRY(grad_store(&gs));
    std::lock_guard<std::mutex> store_lock(gs->mu);
    GradPlan plan;
    char *ckb = nullptr;
    EMS_TRY(grad_prepare(*gs, std::vector<ems_fdtd *>{s}, N, plan, &ckb));
    // adjoint launch data (set once the residual weights are known)
    int serial = 0, cur = 0;
    double *gp = nullptr, *gq = nullptr;
    const AdjGeom ag = adj_geom(g, s->npml);
    const T nch = -ch;
    const std::vector<ems_fdtd *> one{s};
    std::vector<InjectArgs<T>> inj;
    auto ops = grad_ops(
        [&](int n, bool record, int slot) -> ems_status {
            EMS_TRY(fstep(n, record));
            if (slot >= 0) EMS_TRY(snap_shots<T>(one, slot, st));
            return EMS_OK;
        },
        [&](int slot) { return snap_shots<T>(one, slot, st); },
        [&](int idx) { return copy_state(s, ckb + (size_t)idx * sb, true); },
        [&](int idx) { return copy_state(s, ckb + (size_t)idx * sb, false); },
        [&](int n, int slot) -> ems_status {
            if (npts && (n + 1) % tf == 0) EMS_TRY(inject_shots<T>(inj, 1, (n + 1) / tf - 1, st));
            adj_sweeps<T>(s->mat_mode, s->npml > 0, g, ag, st, ce, nch, A->tab.as<AdjShot<T>>() + cur, 1,
                          s->ca.as<T>(), s->cb.as<T>(), gp, gq, (long)s->nalloc, slot);
            EMS_HIP(hipGetLastError());
            cur = 1 - cur;
            return EMS_OK;
        },
        fpair);
    GradSchedule<decltype(ops)> sched{ops, plan, N};
    EMS_TRY(sched.first_pass());
    EMS_TRY(copy_state(s, A->fin.p, true));
    EMS_HIP(hipStreamSynchronize(st));
    // traces (the reference's averaging) and the misfit
    const int nobs = (int)s->obs.size();
    const int nslot = N / tf + 1;
    std::vector<double> raw((size_t)nrec * npts);
    if (!raw.empty()) EMS_HIP(hipMemcpy(raw.data(), s->d_rec.p, sizeof(double) * raw.size(), hipMemcpyDeviceToHost));
    std::vector<double> tr((size_t)nobs * nslot * 3, 0.0);
    for (int r = 0; r < nrec; ++r) {
        int p = 0;
        for (int o = 0; o < nobs; ++o)
            for (int c = 0; c < 3; ++c) {
                double v = 0.0;
                const size_t cnt = s->obs[o][c].size();
                for (size_t q = 0; q < cnt; ++q, ++p) v += raw[(size_t)r * npts + p];
                tr[((size_t)o * nslot + r + 1) * 3 + c] = v / (double)cnt;
            }
    }
    double J = 0.0;
    for (int o = 0; o < nobs; ++o)
        for (int sl = 1; sl < nslot; ++sl)
            for (int c = 0; c < 3; ++c) {
                if (s->obs[o][c].empty()) continue;
                const double r = tr[((size_t)o * nslot + sl) * 3 + c] - d_obs[((size_t)o * nslot + sl) * 3 + c];
                J += 0.5 * r * r;
            }
    // residual weights per recorded row and point: r / count
    std::vector<double> w((size_t)std::max(nrec, 1) * std::max(npts, 1), 0.0);
    for (int r = 0; r < nrec; ++r) {
        int p = 0;
        for (int o = 0; o < nobs; ++o)
            for (int c = 0; c < 3; ++c) {
                const size_t cnt = s->obs[o][c].size();
                const double res =
                    cnt ? tr[((size_t)o * nslot + r + 1) * 3 + c] - d_obs[((size_t)o * nslot + r + 1) * 3 + c] : 0.0;
                for (size_t q = 0; q < cnt; ++q, ++p) w[(size_t)r * npts + p] = res / (double)cnt;
            }
    }
    // points that repeat are injected serially (the misfit's order)
    {
        std::vector<std::pair<int, long>> keys;
        for (int p = 0; p < npts; ++p) keys.emplace_back(pc[p], pu[p]);
        std::sort(keys.begin(), keys.end());
        serial = std::adjacent_find(keys.begin(), keys.end()) != keys.end() ? 1 : 0;
    }
    EMS_HIP(hipMemcpy(A->w.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    inj = inject_args<T>(one, std::vector<int>{npts}, std::vector<int>{serial});
    if (A->fac_version != s->mat_version) {
        const std::vector<double> fac = material_factors(s);
        EMS_TRY(A->fac.reserve(sizeof(double) * fac.size()));
        EMS_HIP(hipMemcpy(A->fac.p, fac.data(), sizeof(double) * fac.size(), hipMemcpyHostToDevice));
        A->fac_version = s->mat_version;
    }
    EMS_TRY(zero(A->sums, sizeof(double) * 2 * s->nalloc, st));
    const AdjShot<T> tab[2] = {adj_shot<T>(s, 0), adj_shot<T>(s, 1)};
    EMS_TRY(A->tab.reserve(sizeof(tab)));
    EMS_HIP(hipMemcpy(A->tab.p, tab, sizeof(tab), hipMemcpyHostToDevice));
    GradView gv{A->fac.as<double>(), A->fac.as<double>() + nm, A->fac.as<double>() + 2 * nm,
                A->fac.as<double>() + 3 * nm, A->grad.as<double>(), A->grad.as<double>() + nm};
    gp = A->sums.as<double>();
    gq = gp + s->nalloc;
    EMS_TRY(sched.backward());
    k_adj_grad<<<(unsigned)((nm + 255) / 256), 256, 0, st>>>(g, gp, gq, gv);
    EMS_HIP(hipGetLastError());
    EMS_TRY(copy_state(s, A->fin.p, false));
    EMS_HIP(hipStreamSynchronize(st));
    // straight into the caller's arrays (no staging copy)
    if (g_eps) EMS_HIP(hipMemcpy(g_eps, A->grad.p, sizeof(double) * nm, hipMemcpyDeviceToHost));
    if (g_sig) EMS_HIP(hipMemcpy(g_sig, A->grad.as<double>() + nm, sizeof(double) * nm, hipMemcpyDeviceToHost));
    if (misfit) *misfit = J;
    s->traces.swap(tr);
    s->nslot = nslot;
    s->t_now = t;
    return EMS_OK;
}

}  // namespace

// =========================================================================
// Batched runs: many handles (shots, models of uneven size) advanced together
// with one launch per sweep (no reference counterpart).
// =========================================================================
struct ems_batch {
    std::vector<ems_fdtd *> mem;
    hipStream_t stream = nullptr;
    int by = 8, lz = 2;      // workgroup rows, planes per thread
    int graph_steps = 0;     // steps per captured graph (0 = plain launches)
    int grad_group = 0;      // shots per gradient group (0: automatic)
    DevBuf d_mem, d_post, d_step;
    DevBuf d_post2, d_shots[2], fac, grad, sums;  // gradient: recompute posts, shot tables, shared FP64 arrays
    const ems_fdtd *fac_owner = nullptr;     // factors cached for this member's material version
    long fac_version = -1;
    ~ems_batch()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Give a handle the coefficient arrays of a more general material mode; the
// results do not change (Cb = dt/eps/h and Ca = 1 reproduce the uniform
// update bit for bit).
template <typename T> ems_status ensure_mode(ems_fdtd *s, int mode)
{
    if (s->mat_mode >= mode) return EMS_OK;
    if (s->mat_mode == 0) {
        std::vector<double> cb(s->nalloc, s->d.dt / s->d.eps / s->d.h);
        EMS_TRY(upload_as<T>(s->cb, cb));
    }
    if (mode == 2) {
        std::vector<double> ca(s->nalloc, 1.0);
        EMS_TRY(upload_as<T>(s->ca, ca));
    }
    s->mat_mode = mode;
    return EMS_OK;
}

template <typename T>
BatchMember<T> member_of(ems_fdtd *s, bool pml, int by, int lz, long base)
{
    BatchMember<T> m{};
    const GridView &g = s->g;
    m.g = g;
    m.f = fields<T>(s);
    m.pm = pml_view<T>(s);
    m.ca = s->ca.as<T>();
    m.cb = s->cb.as<T>();
    m.ch = (T)(s->d.dt / s->d.mu / s->d.h);
    m.ce = (T)(s->d.dt / s->d.eps / s->d.h);
    const int P = pml ? std::max(s->npml, 0) : 0;
    m.lz = lz;
    m.npw = P > 0 ? (2 * P + 1 + 63) / 64 : 0;
    m.nbx = split_row_waves(g, P);
    m.nby = (g.ny + 1 + by - 1) / by;
    m.ntiles = m.nby * ((g.nz + 1 + lz - 1) / lz);
    m.base = base;
    return m;
}

// workgroups of a member: whole groups of 8 tiles (split_tile), so every base is a multiple of 8
long member_blocks(int nbx, int ntiles) { return 8L * nbx * ((ntiles + 7) / 8); }

template <typename T>
ems_status build_members(ems_batch *b, bool pml, std::vector<BatchMember<T>> &ms, long &total)
{
    ms.clear();
    total = 0;
    for (auto *s : b->mem) {
        BatchMember<T> m = member_of<T>(s, pml, b->by, b->lz, total);
        total += member_blocks(m.nbx, m.ntiles);
        ms.push_back(m);
    }
    EMS_TRY(b->d_mem.reserve(sizeof(BatchMember<T>) * ms.size()));
    EMS_HIP(hipMemcpy(b->d_mem.p, ms.data(), sizeof(BatchMember<T>) * ms.size(), hipMemcpyHostToDevice));
    return EMS_OK;
}

template <typename T, int MAT, bool PML>
void batch_sweeps(ems_batch *b, long total, int nm)
{
    if (b->by < 1 || 64 * b->by > 512) return;  // the sweeps are compiled for at most 512 threads
    const dim3 blk(64, b->by, 1);
    const BatchMember<T> *m = b->d_mem.as<BatchMember<T>>();
    k_h_batch<T, PML><<<(unsigned)total, blk, 0, b->stream>>>(m, nm);
    k_e_batch<T, MAT, PML><<<(unsigned)total, blk, 0, b->stream>>>(m, nm);
}

template <typename T> void batch_sweeps_any(ems_batch *b, int mat, bool pml, long total, int nm)
{
    switch (mat * 2 + (pml ? 1 : 0)) {
    case 0: batch_sweeps<T, 0, false>(b, total, nm); break;
    case 1: batch_sweeps<T, 0, true>(b, total, nm); break;
    case 2: batch_sweeps<T, 1, false>(b, total, nm); break;
    case 3: batch_sweeps<T, 1, true>(b, total, nm); break;
    case 4: batch_sweeps<T, 2, false>(b, total, nm); break;
    default: batch_sweeps<T, 2, true>(b, total, nm); break;
    }
}

struct BatchModes {
    int mat = 0;
    bool pml = false;
};

template <typename T> ems_status batch_modes(ems_batch *b, BatchModes &md)
{
    md = BatchModes();
    for (auto *s : b->mem) {
        EMS_TRY(setup_pml(s));
        md.mat = std::max(md.mat, s->mat_mode);
        md.pml = md.pml || s->npml > 0;
    }
    for (auto *s : b->mem) EMS_TRY(ensure_mode<T>(s, md.mat));
    return EMS_OK;
}

// Restores a batch's member list when the batched sweeps of a subset end.
struct MemberSubset {
    ems_batch *b;
    std::vector<ems_fdtd *> all;
    ~MemberSubset() { b->mem = all; }
};

// Members whose single runs take the two-step sweep (grids with a large
// CPML-free interior) keep it in a batch: each runs its own launch sequence
// on its own stream, forked from and joined back into the batch stream,
// concurrently with the others.
bool batch_own(const ems_fdtd *s, int num_steps) { return num_steps >= 2 && kernel_kind(s) == EMS_KERNEL_TB2; }

// A batch run: the members that keep the two-step sweep on their own streams,
// the others through the batched sweeps (one launch per sweep for all of
// them) on the batch stream.  Every member's fields and traces equal its
// single run's bit for bit.
template <typename T> ems_status batch_run_impl(ems_batch *b, int num_steps, double *total_ms)
{
    std::vector<ems_fdtd *> all = b->mem, own, shared;
    for (auto *s : all) (batch_own(s, num_steps) ? own : shared).push_back(s);
    std::vector<RunPlan> rp_own(own.size()), rp(shared.size());
    for (size_t q = 0; q < own.size(); ++q) {
        EMS_TRY(run_prepare<T>(own[q], num_steps, rp_own[q], true));
        EMS_TRY(tb2_check<T>(own[q]));
    }
    for (size_t q = 0; q < shared.size(); ++q) EMS_TRY(run_prepare<T>(shared[q], num_steps, rp[q], false));
    if (num_steps <= 0 || all.empty()) {
        if (total_ms) *total_ms = 0.0;
        return EMS_OK;
    }
    // the shared members' descriptors (before any launch: these copies are
    // synchronous)
    const int nm = (int)shared.size();
    BatchModes md;
    long total = 0;
    if (nm > 0) {
        MemberSubset subset{b, all};
        b->mem = shared;
        EMS_TRY(batch_modes<T>(b, md));
        std::vector<BatchMember<T>> ms;
        EMS_TRY(build_members<T>(b, md.pml, ms, total));
        std::vector<BatchPost> post(nm);
        for (int q = 0; q < nm; ++q) {
            ems_fdtd *s = shared[q];
            BatchPost &p = post[q];
            p.nsrc = rp[q].nsrc;
            p.src_comp = s->d_src_comp.as<int>();
            p.src_u = s->d_src_u.as<long>();
            p.src_val = s->d_src_val.as<double>();
            p.npts = rp[q].npts;
            p.pt_comp = s->d_pt_comp.as<int>();
            p.pt_u = s->d_pt_u.as<long>();
            p.rec = (rp[q].npts && rp[q].nrec) ? s->d_rec.as<double>() : nullptr;
            p.tf = s->d.tf;
            p.step0 = 0;
        }
        EMS_TRY(b->d_post.reserve(sizeof(BatchPost) * nm));
        EMS_HIP(hipMemcpy(b->d_post.p, post.data(), sizeof(BatchPost) * nm, hipMemcpyHostToDevice));
        EMS_TRY(b->d_step.reserve(sizeof(int)));
        const int one = 1;
        EMS_HIP(hipMemcpy(b->d_step.p, &one, sizeof(int), hipMemcpyHostToDevice));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<hipEvent_t> joined(own.size(), nullptr);
    struct Events {
        hipEvent_t &e0, &e1;
        std::vector<hipEvent_t> &j;
        ~Events()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            for (auto e : j)
                if (e) (void)hipEventDestroy(e);
        }
    } events{e0, e1, joined};
    EMS_HIP(hipEventCreate(&e0));
    if (total_ms) EMS_HIP(hipEventCreate(&e1));
    for (auto &e : joined) EMS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    EMS_HIP(hipEventRecord(e0, b->stream));
    for (size_t q = 0; q < own.size(); ++q) {
        EMS_HIP(hipStreamWaitEvent(own[q]->stream, e0, 0));
        int ngroups = 0;
        EMS_TRY(run_loop<T>(own[q], num_steps, rp_own[q], nullptr, ngroups));
        EMS_HIP(hipEventRecord(joined[q], own[q]->stream));
    }
    if (nm > 0) {
        const BatchMember<T> *dm = b->d_mem.as<BatchMember<T>>();
        const BatchPost *dp = b->d_post.as<BatchPost>();
        int done = 0;
        const int S = b->graph_steps;
        if (S > 0 && num_steps >= S) {
            // one captured graph of S steps, replayed; the step index lives on the device
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            EMS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            for (int q = 0; q < S; ++q) {
                batch_sweeps_any<T>(b, md.mat, md.pml, total, nm);
                k_post_batch<T><<<nm, 256, 0, b->stream>>>(dm, dp, 0, b->d_step.as<int>());
                k_step_advance<<<1, 64, 0, b->stream>>>(b->d_step.as<int>());
            }
            EMS_HIP(hipStreamEndCapture(b->stream, &graph));
            EMS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            const int reps = num_steps / S;
            for (int r = 0; r < reps; ++r) EMS_HIP(hipGraphLaunch(exec, b->stream));
            done = reps * S;
            EMS_HIP(hipStreamSynchronize(b->stream));
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
        }
        for (int step = done + 1; step <= num_steps; ++step) {
            batch_sweeps_any<T>(b, md.mat, md.pml, total, nm);
            k_post_batch<T><<<nm, 256, 0, b->stream>>>(dm, dp, step, nullptr);
        }
        EMS_HIP(hipGetLastError());
    }
    for (auto e : joined) EMS_HIP(hipStreamWaitEvent(b->stream, e, 0));
    if (total_ms) {
        EMS_HIP(hipEventRecord(e1, b->stream));
        EMS_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        EMS_HIP(hipEventElapsedTime(&ms, e0, e1));
        *total_ms = ms;
    }
    EMS_HIP(hipStreamSynchronize(b->stream));
    for (size_t q = 0; q < own.size(); ++q) EMS_TRY(run_finish(own[q], rp_own[q]));
    for (size_t q = 0; q < shared.size(); ++q) EMS_TRY(run_finish(shared[q], rp[q]));
    return EMS_OK;
}

// Autotuner: times the batched sweeps for workgroup rows x planes per thread
// candidates on zeroed scratch fields (the members' own state is untouched;
// zero inputs stay zero) and keeps the fastest.
template <typename T> ems_status batch_autotune_impl(ems_batch *b, int trial, int *rows, int *planes)
{
    if (b->mem.empty()) return invalid("empty batch");
    // only the members that run through the batched sweeps
    MemberSubset subset{b, b->mem};
    std::vector<ems_fdtd *> shared;
    for (auto *s : subset.all)
        if (!batch_own(s, 2)) shared.push_back(s);
    b->mem = shared;
    const int nm = (int)b->mem.size();
    if (nm == 0) {
        if (rows) *rows = b->by;
        if (planes) *planes = b->lz;
        return EMS_OK;
    }
    BatchModes md;
    EMS_TRY(batch_modes<T>(b, md));
    size_t fmax = 0, pmax = 0;
    for (auto *s : b->mem) {
        fmax = std::max(fmax, s->nalloc * s->esz);
        for (int p = 0; p < 12; ++p) pmax = std::max(pmax, s->psi[p].bytes);
    }
    DevBuf scratch;
    const size_t sbytes = 6 * fmax + 12 * std::max(pmax, (size_t)256);
    EMS_TRY(scratch.reserve(sbytes));
    EMS_HIP(hipMemset(scratch.p, 0, sbytes));
    char *base = static_cast<char *>(scratch.p);
    // workgroups of 64 x rows threads: rows <= 8 keeps them within the
    // sweeps' 512-thread launch bound
    const int cand_by[3] = {2, 4, 8}, cand_lz[4] = {1, 2, 4, 8};
    float best = 1e30f;
    int bb = b->by, bl = b->lz;
    hipEvent_t e0, e1;
    EMS_HIP(hipEventCreate(&e0));
    EMS_HIP(hipEventCreate(&e1));
    for (int by : cand_by)
        for (int lz : cand_lz) {
            b->by = by;
            b->lz = lz;
            std::vector<BatchMember<T>> ms;
            long total = 0;
            for (auto *s : b->mem) {
                BatchMember<T> m = member_of<T>(s, md.pml, by, lz, total);
                total += member_blocks(m.nbx, m.ntiles);
                T *f[6];
                for (int w = 0; w < 6; ++w) f[w] = reinterpret_cast<T *>(base + w * fmax);
                m.f = Fields<T>{f[0], f[1], f[2], f[3], f[4], f[5]};
                for (int ax = 0; ax < 3; ++ax)
                    for (int q = 0; q < 2; ++q) {
                        m.pm.psi_h[ax][q] = reinterpret_cast<T *>(base + 6 * fmax + (ax * 4 + q) * pmax);
                        m.pm.psi_e[ax][q] = reinterpret_cast<T *>(base + 6 * fmax + (ax * 4 + 2 + q) * pmax);
                    }
                ms.push_back(m);
            }
            EMS_TRY(b->d_mem.reserve(sizeof(BatchMember<T>) * ms.size()));
            EMS_HIP(hipMemcpy(b->d_mem.p, ms.data(), sizeof(BatchMember<T>) * ms.size(), hipMemcpyHostToDevice));
            batch_sweeps_any<T>(b, md.mat, md.pml, total, nm);  // warm-up
            EMS_HIP(hipEventRecord(e0, b->stream));
            for (int q = 0; q < trial; ++q) batch_sweeps_any<T>(b, md.mat, md.pml, total, nm);
            EMS_HIP(hipEventRecord(e1, b->stream));
            EMS_HIP(hipEventSynchronize(e1));
            EMS_HIP(hipGetLastError());
            float ms_ = 0.f;
            EMS_HIP(hipEventElapsedTime(&ms_, e0, e1));
            if (ms_ < best) {
                best = ms_;
                bb = by;
                bl = lz;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->by = bb;
    b->lz = bl;
    if (rows) *rows = bb;
    if (planes) *planes = bl;
    return EMS_OK;
}

// Traces of a fresh run from the raw samples, the misfit and the residual
// weights (r / count per point and recorded row); shared by the gradients.
void misfit_of(const ems_fdtd *s, int N, const RunPlan &rp, const std::vector<double> &raw, const double *d_obs,
               std::vector<double> &tr, double &J, std::vector<double> &w, int &serial,
               const std::vector<int> &pc, const std::vector<long> &pu)
{
    const int tf = s->d.tf, nobs = (int)s->obs.size(), nslot = N / tf + 1, npts = rp.npts, nrec = rp.nrec;
    tr.assign((size_t)nobs * nslot * 3, 0.0);
    for (int r = 0; r < nrec; ++r) {
        int p = 0;
        for (int o = 0; o < nobs; ++o)
            for (int c = 0; c < 3; ++c) {
                double v = 0.0;
                const size_t cnt = s->obs[o][c].size();
                for (size_t q = 0; q < cnt; ++q, ++p) v += raw[(size_t)r * npts + p];
                tr[((size_t)o * nslot + r + 1) * 3 + c] = v / (double)cnt;
            }
    }
    J = 0.0;
    for (int o = 0; o < nobs; ++o)
        for (int sl = 1; sl < nslot; ++sl)
            for (int c = 0; c < 3; ++c) {
                if (s->obs[o][c].empty()) continue;
                const double r = tr[((size_t)o * nslot + sl) * 3 + c] - d_obs[((size_t)o * nslot + sl) * 3 + c];
                J += 0.5 * r * r;
            }
    w.assign((size_t)std::max(nrec, 1) * std::max(npts, 1), 0.0);
    for (int r = 0; r < nrec; ++r) {
        int p = 0;
        for (int o = 0; o < nobs; ++o)
            for (int c = 0; c < 3; ++c) {
                const size_t cnt = s->obs[o][c].size();
                const double res =
                    cnt ? tr[((size_t)o * nslot + r + 1) * 3 + c] - d_obs[((size_t)o * nslot + r + 1) * 3 + c] : 0.0;
                for (size_t q = 0; q < cnt; ++q, ++p) w[(size_t)r * npts + p] = res / (double)cnt;
            }
    }
    std::vector<std::pair<int, long>> keys;
    for (int p = 0; p < npts; ++p) keys.emplace_back(pc[p], pu[p]);
    std::sort(keys.begin(), keys.end());
    serial = std::adjacent_find(keys.begin(), keys.end()) != keys.end() ? 1 : 0;
}

bool same_model(const ems_fdtd *a, const ems_fdtd *b)
{
    return a->d.nx == b->d.nx && a->d.ny == b->d.ny && a->d.nz == b->d.nz && a->d.h == b->d.h && a->d.dt == b->d.dt &&
           a->d.eps == b->d.eps && a->d.mu == b->d.mu && a->d.tf == b->d.tf && a->d.precision == b->d.precision &&
           a->npml == b->npml && a->pml_order == b->pml_order && a->pml_r0 == b->pml_r0 &&
           a->pml_kmax == b->pml_kmax && a->pml_amax == b->pml_amax && a->eps_r == b->eps_r && a->sigma == b->sigma;
}

// Misfits of all members of b (shots on one model) and their gradient sums,
// added to P = gp, Q = gq (unified grid, FP64): batched forward and
// recompute sweeps, each shot's own checkpoints, and adjoint passes over all
// shots that add the shots' sums to P and Q once per point and step.
template <typename T>
ems_status batch_gradient_group(ems_batch *b, int N, const double *const *d_obs, double *misfits, double &Jsum,
                                double *gp_out, double *gq_out)
{
    const int ns = (int)b->mem.size();
    ems_fdtd *s0 = b->mem[0];
    for (auto *s : b->mem)
        if (!same_model(s0, s)) return invalid("batch gradient: members must share grid, material, CPML, dt and tf");
    const GridView &g = s0->g;
    const int tf = s0->d.tf;
    BatchModes md;
    EMS_TRY(batch_modes<T>(b, md));
    std::vector<hipStream_t> saved(ns);
    for (int q = 0; q < ns; ++q) {
        saved[q] = b->mem[q]->stream;
        b->mem[q]->stream = b->stream;
    }
    struct Restore {
        ems_batch *b;
        std::vector<hipStream_t> &sv;
        ~Restore()
        {
            for (size_t q = 0; q < b->mem.size(); ++q) b->mem[q]->stream = sv[q];
        }
    } restore{b, saved};
    hipStream_t st = b->stream;
    // fresh state and run tables per shot
    std::vector<RunPlan> rp(ns);
    std::vector<std::vector<int>> pcs(ns);
    std::vector<std::vector<long>> pus(ns);
    for (int q = 0; q < ns; ++q) {
        ems_fdtd *s = b->mem[q];
        for (int w = 0; w < 6; ++w) EMS_HIP(hipMemsetAsync(s->f[w].p, 0, s->nalloc * s->esz, st));
        for (int p = 0; p < 12; ++p)
            if (s->psi[p].bytes) EMS_HIP(hipMemsetAsync(s->psi[p].p, 0, s->psi[p].bytes, st));
        if (s->alt_ready)
            for (int w = 0; w < 6; ++w) EMS_HIP(hipMemsetAsync(s->f2[w].p, 0, s->nalloc * s->esz, st));
        EMS_HIP(hipStreamSynchronize(st));
        s->t_now = 0.0;
        s->traces.clear();
        s->nslot = 0;
        EMS_TRY(run_prepare<T>(s, N, rp[q], false));
        for (auto &o : s->obs)
            for (int c = 0; c < 3; ++c)
                for (auto &p : o[c]) {
                    pcs[q].push_back(c);
                    pus[q].push_back(ref_to_unified(g, c, p[0], p[1], p[2]));
                }
    }
    std::vector<BatchMember<T>> ms;
    long total = 0;
    EMS_TRY(build_members<T>(b, md.pml, ms, total));
    std::vector<BatchPost> post(ns), post2(ns);
    for (int q = 0; q < ns; ++q) {
        ems_fdtd *s = b->mem[q];
        BatchPost &p = post[q];
        p.nsrc = rp[q].nsrc;
        p.src_comp = s->d_src_comp.as<int>();
        p.src_u = s->d_src_u.as<long>();
        p.src_val = s->d_src_val.as<double>();
        p.npts = rp[q].npts;
        p.pt_comp = s->d_pt_comp.as<int>();
        p.pt_u = s->d_pt_u.as<long>();
        p.rec = (rp[q].npts && rp[q].nrec) ? s->d_rec.as<double>() : nullptr;
        p.tf = tf;
        p.step0 = 0;
        post2[q] = p;
        post2[q].rec = nullptr;
    }
    EMS_TRY(b->d_post.reserve(sizeof(BatchPost) * ns));
    EMS_TRY(b->d_post2.reserve(sizeof(BatchPost) * ns));
    EMS_HIP(hipMemcpy(b->d_post.p, post.data(), sizeof(BatchPost) * ns, hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(b->d_post2.p, post2.data(), sizeof(BatchPost) * ns, hipMemcpyHostToDevice));
    const BatchMember<T> *dm = b->d_mem.as<BatchMember<T>>();
    // gradient buffers outside the store first
    const size_t sb = state_bytes(s0);
    for (int q = 0; q < ns; ++q) {
        ems_fdtd *s = b->mem[q];
        if (!s->adj) s->adj = new Adj();
        EMS_TRY(s->adj->fin.reserve(sb));
        EMS_TRY(s->adj->w.reserve(sizeof(double) * (size_t)std::max(rp[q].nrec, 1) * std::max(rp[q].npts, 1)));
        EMS_TRY(adj_zero(s, st));
    }
    GradStore *gs = nullptr;
    EMS_TRY(grad_store(&gs));
    std::lock_guard<std::mutex> store_lock(gs->mu);
    GradPlan plan;
    char *ckb = nullptr;
    EMS_TRY(grad_prepare(*gs, b->mem, N, plan, &ckb));
    // adjoint launch data (set once the residual weights are known)
    std::vector<int> serial(ns, 0);
    int par = 0;
    double *gp = nullptr, *gq = nullptr;
    const AdjGeom ag = adj_geom(g, s0->npml);
    const T ce = (T)(s0->d.dt / s0->d.eps / s0->d.h);
    const T nch = -(T)(s0->d.dt / s0->d.mu / s0->d.h);
    auto snap_all = [&](int slot) { return snap_shots<T>(b->mem, slot, st); };
    std::vector<InjectArgs<T>> inj;
    auto state_all = [&](int idx, bool save) -> ems_status {
        for (int q = 0; q < ns; ++q) EMS_TRY(copy_state(b->mem[q], ckb + ((size_t)idx * ns + q) * sb, save));
        return EMS_OK;
    };
    auto ops = grad_ops(
        [&](int n, bool record, int slot) -> ems_status {
            batch_sweeps_any<T>(b, md.mat, md.pml, total, ns);
            k_post_batch<T><<<ns, 256, 0, st>>>(dm, (record ? b->d_post : b->d_post2).as<BatchPost>(), n + 1, nullptr);
            EMS_HIP(hipGetLastError());
            if (slot >= 0) EMS_TRY(snap_all(slot));
            return EMS_OK;
        },
        snap_all, [&](int idx) { return state_all(idx, true); }, [&](int idx) { return state_all(idx, false); },
        [&](int n, int slot) -> ems_status {
            if ((n + 1) % tf == 0) EMS_TRY(inject_shots<T>(inj, ns, (n + 1) / tf - 1, st));
            adj_sweeps<T>(md.mat, md.pml, g, ag, st, ce, nch, b->d_shots[par].as<AdjShot<T>>(), ns, s0->ca.as<T>(),
                          s0->cb.as<T>(), gp, gq, (long)s0->nalloc, slot);
            EMS_HIP(hipGetLastError());
            par = 1 - par;
            return EMS_OK;
        });
    GradSchedule<decltype(ops)> sched{ops, plan, N};
    EMS_TRY(sched.first_pass());
    for (auto *s : b->mem) EMS_TRY(copy_state(s, s->adj->fin.p, true));
    EMS_HIP(hipStreamSynchronize(st));
    EMS_HIP(hipGetLastError());
    // traces, misfits and residual weights
    Jsum = 0.0;
    for (int q = 0; q < ns; ++q) {
        ems_fdtd *s = b->mem[q];
        std::vector<double> raw((size_t)rp[q].nrec * rp[q].npts), tr, w;
        if (!raw.empty())
            EMS_HIP(hipMemcpy(raw.data(), s->d_rec.p, sizeof(double) * raw.size(), hipMemcpyDeviceToHost));
        double J = 0.0;
        misfit_of(s, N, rp[q], raw, d_obs[q], tr, J, w, serial[q], pcs[q], pus[q]);
        if (misfits) misfits[q] = J;
        Jsum += J;
        EMS_HIP(hipMemcpy(s->adj->w.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
        s->traces.swap(tr);
        s->nslot = N / tf + 1;
        s->t_now = rp[q].t_end;
    }
    {
        std::vector<int> npts(ns);
        for (int q = 0; q < ns; ++q) npts[q] = rp[q].npts;
        inj = inject_args<T>(b->mem, npts, serial);
    }
    // the two parity tables of the shots' adjoint state
    std::vector<AdjShot<T>> tab[2];
    for (auto *s : b->mem)
        for (int pp = 0; pp < 2; ++pp) tab[pp].push_back(adj_shot<T>(s, pp));
    for (int pp = 0; pp < 2; ++pp) {
        EMS_TRY(b->d_shots[pp].reserve(sizeof(AdjShot<T>) * ns));
        EMS_HIP(hipMemcpy(b->d_shots[pp].p, tab[pp].data(), sizeof(AdjShot<T>) * ns, hipMemcpyHostToDevice));
    }
    gp = gp_out;
    gq = gq_out;
    EMS_TRY(sched.backward());
    for (auto *s : b->mem) EMS_TRY(copy_state(s, s->adj->fin.p, false));
    EMS_HIP(hipStreamSynchronize(st));
    return EMS_OK;
}

// Shots per gradient group: the set size, else the most shots whose E
// snapshots of all N+1 steps take at most half of the device's memory (no
// step is then recomputed), at least 1.  Depends on the device and the
// workload only, so that the summation order (group by group) does too.
int gradient_group(const ems_batch *b, int N)
{
    const int ns = (int)b->mem.size();
    if (b->grad_group > 0) return std::min(b->grad_group, ns);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        (void)hipGetLastError();
        return ns;
    }
    const ems_fdtd *s = b->mem[0];
    const double per_shot = (N + 1.0) * 3.0 * (double)s->nalloc * (double)s->esz;
    return (int)std::max(1.0, std::min((double)ns, std::floor(0.5 * (double)tot / per_shot)));
}

// Gradient of the summed misfit of all members: the members in groups of
// gradient_group shots (as even as possible, in member order), each group's
// sums added to one P, Q, then the eps_r / sigma factors once.
template <typename T>
ems_status batch_gradient_impl(ems_batch *b, int N, const double *const *d_obs, double *misfits, double *Jtot,
                               double *g_eps, double *g_sig)
{
    const int ns = (int)b->mem.size();
    ems_fdtd *s0 = b->mem[0];
    for (auto *s : b->mem)
        if (!same_model(s0, s)) return invalid("batch gradient: members must share grid, material, CPML, dt and tf");
    const GridView &g = s0->g;
    const size_t nm = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    hipStream_t st = b->stream;
    EMS_TRY(b->grad.reserve(sizeof(double) * 2 * nm));
    EMS_TRY(b->fac.reserve(sizeof(double) * 4 * nm));
    EMS_TRY(zero(b->sums, sizeof(double) * 2 * s0->nalloc, st));
    double *gp = b->sums.as<double>(), *gq = gp + s0->nalloc;
    const int gsz = gradient_group(b, N);
    const int ngroups = (ns + gsz - 1) / gsz;
    std::vector<double> mis(ns, 0.0);
    double Jg = 0.0;
    if (ngroups == 1) {
        EMS_TRY(batch_gradient_group<T>(b, N, d_obs, mis.data(), Jg, gp, gq));
    } else {
        for (int k = 0; k < ngroups; ++k) {
            const int q0 = (int)((long)k * ns / ngroups), q1 = (int)((long)(k + 1) * ns / ngroups);
            ems_batch sub;
            sub.mem.assign(b->mem.begin() + q0, b->mem.begin() + q1);
            sub.stream = st;
            sub.by = b->by;
            sub.lz = b->lz;
            struct Lend {  // the parent's stream, not the group's to destroy
                ems_batch &b;
                ~Lend() { b.stream = nullptr; }
            } lend{sub};
            EMS_TRY(batch_gradient_group<T>(&sub, N, d_obs + q0, mis.data() + q0, Jg, gp, gq));
        }
    }
    double Jsum = 0.0;  // member order, whatever the groups
    for (int q = 0; q < ns; ++q) Jsum += mis[q];
    if (misfits) std::copy(mis.begin(), mis.end(), misfits);
    if (Jtot) *Jtot = Jsum;
    if (b->fac_owner != s0 || b->fac_version != s0->mat_version) {
        const std::vector<double> fac = material_factors(s0);
        EMS_HIP(hipMemcpy(b->fac.p, fac.data(), sizeof(double) * fac.size(), hipMemcpyHostToDevice));
        b->fac_owner = s0;
        b->fac_version = s0->mat_version;
    }
    GradView gv{b->fac.as<double>(), b->fac.as<double>() + nm, b->fac.as<double>() + 2 * nm,
                b->fac.as<double>() + 3 * nm, b->grad.as<double>(), b->grad.as<double>() + nm};
    k_adj_grad<<<(unsigned)((nm + 255) / 256), 256, 0, st>>>(g, gp, gq, gv);
    EMS_HIP(hipGetLastError());
    EMS_HIP(hipStreamSynchronize(st));
    // straight into the caller's arrays (no staging copy)
    if (g_eps) EMS_HIP(hipMemcpy(g_eps, b->grad.p, sizeof(double) * nm, hipMemcpyDeviceToHost));
    if (g_sig) EMS_HIP(hipMemcpy(g_sig, b->grad.as<double>() + nm, sizeof(double) * nm, hipMemcpyDeviceToHost));
    return EMS_OK;
}

ems_status batch_check(ems_batch *b)
{
    if (!b) return invalid("null batch");
    for (auto *s : b->mem)
        if (s->d.precision != b->mem[0]->d.precision) return invalid("batch members must share one precision");
    return EMS_OK;
}

}  // namespace

// =========================================================================
// ems_fdtd_run_static: ems_fdtd_run without device allocation
// =========================================================================
//
// The per-call staging of a run (source component and index lists, the
// per-step source values, the receiver point lists and the raw receiver
// samples) lives in one file-scope device array per GPU instead of the
// handle's grow-only buffers.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU (two threads, two streams) race; calls that
// follow each other do not (each call waits for its kernels before it
// returns).  Every byte a call reads from the array was written by the same
// call first, so nothing needs clearing.
namespace {

constexpr size_t kStaticScratchBytes = (size_t)8 << 20;  // per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current GPU, looked up once per GPU
// (the array's host-side name is no device address).
ems_status static_scratch(void **out)
{
    int dev = -1;
    EMS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return invalid("ems_fdtd_run_static: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        EMS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("ems_fdtd_run_static: no device address for the static scratch array");
            return EMS_ERR_HIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return EMS_OK;
}

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Copies n values produced by gen(i) to dst through a fixed host buffer.
template <typename V, typename Gen> ems_status stage_values(V *dst, size_t n, Gen gen)
{
    constexpr size_t kChunk = 512;
    V buf[kChunk];
    for (size_t i0 = 0; i0 < n; i0 += kChunk) {
        const size_t m = std::min(kChunk, n - i0);
        for (size_t i = 0; i < m; ++i) buf[i] = gen(i0 + i);
        EMS_HIP(hipMemcpy(dst + i0, buf, sizeof(V) * m, hipMemcpyHostToDevice));
    }
    return EMS_OK;
}

template <typename T> ems_status run_static_impl(ems_fdtd *s, int num_steps)
{
    if (s->npml > 0 && s->dirty_pml)
        return (set_error("ems_fdtd_run_static: the CPML is not set up yet (run ems_fdtd_run(s, 0) first)"),
                EMS_ERR_STATE);
    if ((use_fused(s) && !s->alt_ready) ||
        (use_tiles(s) && (!s->tiles_ok || s->tiles_key[0] != s->tile_by || s->tiles_key[1] != s->tile_lz ||
                          s->tiles_key[2] != s->npml)))
        return (set_error("ems_fdtd_run_static: the fused kernel is not set up yet (run ems_fdtd_run(s, 0) first)"),
                EMS_ERR_STATE);
    const int tf = s->d.tf;
    // sources in the order of run_prepare (stable by component)
    int nsrc = (int)s->src.size();
    auto src_at = [&](int q) -> const Src & {
        int k = 0;
        for (int c = 0; c < 3; ++c)
            for (const Src &x : s->src)
                if (x.comp == c && k++ == q) return x;
        return s->src[0];
    };
    size_t npts = 0;
    for (auto &o : s->obs)
        for (int c = 0; c < 3; ++c) npts += o[c].size();
    const size_t nrec = (size_t)(num_steps / tf);
    const size_t b_sc = align256(sizeof(int) * nsrc), b_su = align256(sizeof(long) * nsrc);
    const size_t b_pc = align256(sizeof(int) * npts), b_pu = align256(sizeof(long) * npts);
    const size_t b_sv = align256(sizeof(double) * (size_t)num_steps * nsrc);
    const size_t b_rec = align256(sizeof(double) * nrec * npts);
    const size_t need = b_sc + b_su + b_pc + b_pu + b_sv + b_rec;
    if (need > kStaticScratchBytes)
        return invalid("ems_fdtd_run_static: this call needs " + std::to_string(need) +
                       " bytes of staging, more than the static maximum of " + std::to_string(kStaticScratchBytes) +
                       " (fewer steps per call, sources or receiver points)");
    void *base = nullptr;
    EMS_TRY(static_scratch(&base));
    unsigned char *bp = static_cast<unsigned char *>(base);
    int *d_sc = reinterpret_cast<int *>(bp);
    long *d_su = reinterpret_cast<long *>(bp + b_sc);
    int *d_pc = reinterpret_cast<int *>(bp + b_sc + b_su);
    long *d_pu = reinterpret_cast<long *>(bp + b_sc + b_su + b_pc);
    double *d_sv = reinterpret_cast<double *>(bp + b_sc + b_su + b_pc + b_pu);
    double *d_rec = reinterpret_cast<double *>(bp + b_sc + b_su + b_pc + b_pu + b_sv);

    // trace slots: the resize semantics of ems_fdtd_run (host bookkeeping)
    const int nslot_new = num_steps / tf + 1;
    const int nobs = (int)s->obs.size();
    if (nslot_new > s->nslot) {
        std::vector<double> t((size_t)nobs * nslot_new * 3, 0.0);
        for (int o = 0; o < nobs; ++o)
            for (int q = 0; q < s->nslot; ++q)
                for (int c = 0; c < 3; ++c)
                    t[((size_t)o * nslot_new + q) * 3 + c] = s->traces[((size_t)o * s->nslot + q) * 3 + c];
        s->traces.swap(t);
        s->nslot = nslot_new;
    }
    if (num_steps <= 0) return EMS_OK;
    EMS_TRY(stage_values(d_sc, (size_t)nsrc, [&](size_t q) { return src_at((int)q).comp; }));
    EMS_TRY(stage_values(d_su, (size_t)nsrc, [&](size_t q) { return src_at((int)q).u; }));
    // per-step injected values, t accumulated as current_time_ (as run_prepare)
    const double t0 = s->t_now, dt = s->d.dt;
    double t_end = t0;
    for (int st = 0; st < num_steps; ++st) t_end += dt;
    if (nsrc) {
        double t = t0;
        int st_cur = 0;
        double w = wavelet_value(s, t);
        EMS_TRY(stage_values(d_sv, (size_t)num_steps * nsrc, [&](size_t i) {
            const int st = (int)(i / nsrc), q = (int)(i % nsrc);
            while (st_cur < st) {  // values are generated in step order
                t += dt;
                ++st_cur;
                w = wavelet_value(s, t);
            }
            const Src &x = src_at(q);
            double v = w * dt / x.eps;
            if (x.scale != 1.0) v = v / x.scale;
            return v;
        }));
    }
    // receiver points in observation order
    auto point_at = [&](size_t i, int &comp) -> long {
        size_t k = 0;
        for (auto &o : s->obs)
            for (int c = 0; c < 3; ++c) {
                if (i < k + o[c].size()) {
                    const auto &p = o[c][i - k];
                    comp = c;
                    return ref_to_unified(s->g, c, p[0], p[1], p[2]);
                }
                k += o[c].size();
            }
        comp = 0;
        return 0;
    };
    EMS_TRY(stage_values(d_pc, npts, [&](size_t i) { int c; point_at(i, c); return c; }));
    EMS_TRY(stage_values(d_pu, npts, [&](size_t i) { int c; return point_at(i, c); }));

    const T ch = (T)(s->d.dt / s->d.mu / s->d.h);
    const T ce = (T)(s->d.dt / s->d.eps / s->d.h);
    for (int step = 1; step <= num_steps; ++step) {
        const PmlView<T> pm = pml_view<T>(s);
        launch_step<T>(s, ch, ce, pm);
        const Fields<T> f = fields<T>(s);
        EMS_HIP(hipGetLastError());
        double *rec = nullptr;
        if (npts && step % tf == 0) rec = d_rec + (size_t)(step / tf - 1) * npts;
        if (nsrc || rec) {
            k_post<T><<<1, 256, 0, s->stream>>>(f, nsrc, d_sc, d_su, d_sv + (size_t)(step - 1) * nsrc, (int)npts,
                                                d_pc, d_pu, rec);
            EMS_HIP(hipGetLastError());
        }
    }
    EMS_HIP(hipStreamSynchronize(s->stream));
    // combine the raw samples into the traces as run_finish does, one
    // receiver row at a time through a fixed host buffer
    s->t_now = t_end;
    if (nrec > 0 && nobs > 0) {
        constexpr size_t kRow = 512;
        double row[kRow];
        for (size_t r = 0; r < nrec; ++r) {
            const int slot = (int)r + 1;
            size_t p = 0, have = 0, at = 0;
            auto next = [&]() -> ems_status {
                if (at == have) {
                    have = std::min(kRow, npts - p);
                    EMS_HIP(hipMemcpy(row, d_rec + r * npts + p, sizeof(double) * have, hipMemcpyDeviceToHost));
                    at = 0;
                }
                return EMS_OK;
            };
            for (int o = 0; o < nobs; ++o)
                for (int c = 0; c < 3; ++c) {
                    double &dst = s->traces[((size_t)o * s->nslot + slot) * 3 + c];
                    const size_t cnt = s->obs[o][c].size();
                    for (size_t q = 0; q < cnt; ++q, ++p) {
                        EMS_TRY(next());
                        dst += row[at++];
                    }
                    dst /= (double)cnt;
                }
        }
    }
    return EMS_OK;
}

}  // namespace

// =========================================================================
// C ABI
// =========================================================================
extern "C" {

ems_status ems_fdtd_create(const ems_fdtd_desc *desc, ems_fdtd **out)
{
    clear_error();
    if (!desc || !out) return invalid("ems_fdtd_create: null argument");
    *out = nullptr;
    if (desc->nx <= 0 || desc->ny <= 0 || desc->nz <= 0) return invalid("grid sizes must be positive");
    if (desc->tf <= 0) return invalid("time_frequency must be >= 1");
    if (!(desc->h > 0) || !(desc->dt > 0) || !(desc->eps > 0) || !(desc->mu > 0))
        return invalid("h, dt, eps and mu must be positive");
    if (desc->precision != EMS_F64 && desc->precision != EMS_F32) return invalid("unknown precision");
    {  // the split sweeps address a plane with 32-bit byte offsets (any row shift)
        const GridView g = make_grid(desc->nx, desc->ny, desc->nz, kPitchAlign - 1);
        if ((double)g.slab * (desc->precision == EMS_F32 ? 4 : 8) >= 4294967296.0)
            return invalid("grid plane too large (one z-plane must be under 4 GiB)");
    }
    if (ems_device_count() <= 0) {
        set_error("no HIP device available");
        return EMS_ERR_HIP;
    }
    ems_fdtd *s = new ems_fdtd();
    s->d = *desc;
    s->mat_version = next_version();
    s->esz = desc->precision == EMS_F32 ? 4 : 8;
    s->g = make_grid(desc->nx, desc->ny, desc->nz);
    s->nalloc = (size_t)(s->g.off + s->g.slab * (desc->nz + 2) + kFrontPad);
    for (int w = 0; w < 6; ++w) {
        ems_status st = s->f[w].reserve(s->nalloc * s->esz);
        if (st != EMS_OK) {
            delete s;
            return st;
        }
        hipError_t e = hipMemset(s->f[w].p, 0, s->nalloc * s->esz);
        if (e != hipSuccess) {
            delete s;
            return hip_fail(e, "hipMemset", __FILE__, __LINE__);
        }
    }
    *out = s;
    return EMS_OK;
}

void ems_fdtd_destroy(ems_fdtd *s) { delete s; }

ems_status ems_fdtd_add_source(ems_fdtd *s, int comp, int i, int j, int k)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (comp < 0 || comp > 2) return invalid("source component must be 0, 1 or 2");
    if (!ref_in_range(s->g, comp, i, j, k)) return invalid("source index outside the field array");
    Src q{comp, i, j, k, ref_to_unified(s->g, comp, i, j, k), s->d.eps, 1.0};
    s->src.push_back(q);
    if (!s->eps_r.empty() || !s->sigma.empty()) return set_material_impl(s);
    return EMS_OK;
}

ems_status ems_fdtd_add_observation(ems_fdtd *s, const int counts[3], const int *pts)
{
    clear_error();
    if (!s || !counts) return invalid("null argument");
    ObsLists o;
    int p = 0;
    for (int c = 0; c < 3; ++c) {
        if (counts[c] < 0) return invalid("negative point count");
        for (int q = 0; q < counts[c]; ++q, ++p) {
            std::array<int, 3> v{pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
            if (!ref_in_range(s->g, c, v[0], v[1], v[2])) return invalid("observation point outside the field array");
            o[c].push_back(v);
        }
    }
    s->obs.push_back(o);
    // a new observation set starts with an empty trace (reference push_back)
    const int nobs = (int)s->obs.size();
    std::vector<double> t((size_t)nobs * s->nslot * 3, 0.0);
    std::copy(s->traces.begin(), s->traces.end(), t.begin());
    s->traces.swap(t);
    return EMS_OK;
}

ems_status ems_fdtd_set_material(ems_fdtd *s, const double *eps_r, const double *sigma)
{
    clear_error();
    if (!s) return invalid("null handle");
    const size_t n = (size_t)(s->g.nx + 1) * (s->g.ny + 1) * (s->g.nz + 1);
    for (size_t q = 0; eps_r && q < n; ++q)
        if (!(eps_r[q] > 0)) return invalid("eps_r must be positive");
    for (size_t q = 0; sigma && q < n; ++q)
        if (!(sigma[q] >= 0)) return invalid("sigma must be non-negative");
    if (eps_r) s->eps_r.assign(eps_r, eps_r + n);
    else s->eps_r.clear();
    if (sigma) s->sigma.assign(sigma, sigma + n);
    else s->sigma.clear();
    s->mat_version = next_version();  // unique over all handles: keys the cached gradient factors
    return set_material_impl(s);
}

ems_status ems_fdtd_set_cpml(ems_fdtd *s, int ncells, double order, double r0, double kappa_max,
                             double alpha_max)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (ncells < 0) return invalid("CPML thickness must be >= 0");
    const int nmin = std::min({s->g.nx, s->g.ny, s->g.nz});
    if (ncells > 0 && 2 * ncells + 1 > nmin) return invalid("CPML thicker than half the grid");
    {  // psi slabs are indexed with 32-bit integers
        const long W = 2L * ncells + 1, ax = s->g.nx + 2L, ay = s->g.ny + 2L, az = s->g.nz + 2L;
        if (ncells > 0 && std::max({W * ay * az, ax * W * az, ax * ay * W}) >= (1L << 31))
            return invalid("CPML slabs too large (2^31 entries per slab)");
    }
    if (ncells > 0 && (!(r0 > 0 && r0 < 1) || !(order >= 0) || !(kappa_max >= 1) || !(alpha_max >= 0)))
        return invalid("invalid CPML parameters");
    s->npml = ncells;
    s->pml_order = order;
    s->pml_r0 = r0;
    s->pml_kmax = kappa_max;
    s->pml_amax = alpha_max;
    s->dirty_pml = true;
    // the layout the two-step kernel's CPML slabs prefer, when it runs
    return relayout(s, kernel_kind(s) == EMS_KERNEL_TB2 ? row_shift(s->g.nx, ncells, s->esz) : 0);
}

ems_status ems_fdtd_set_wavelet(ems_fdtd *s, int kind, double f0, double t0, double amp)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (kind < 0 || kind > 2) return invalid("unknown wavelet");
    if (!(f0 > 0)) return invalid("f0 must be positive");
    s->wavelet = kind;
    s->f0 = f0;
    s->t0 = t0;
    s->amp = amp;
    return EMS_OK;
}

ems_status ems_fdtd_set_kernel(ems_fdtd *s, int variant)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (variant < EMS_KERNEL_AUTO || variant > EMS_KERNEL_TB2) return invalid("unknown kernel variant");
    s->kernel = variant;
    return EMS_OK;
}

ems_status ems_fdtd_set_tb2(ems_fdtd *s, int cols, int rows, int planes)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (!(cols == 0 && rows == 0) && (cols < 5 || rows < 5 || cols * rows > kTbMaxThreads))
        return invalid("two-step tile: cols, rows >= 5 and cols * rows <= 1024 (512 in FP64), or both 0");
    if (planes < 0) return invalid("two-step planes per workgroup must be >= 1 (0: default)");
    s->tb_txw = cols;
    s->tb_tyw = rows;
    s->tb_lz = planes;
    return EMS_OK;
}

ems_status ems_fdtd_autotune_tb2(ems_fdtd *s, int trial_steps, int *cols, int *rows, int *planes)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (trial_steps < 1) return invalid("trial_steps must be >= 1");
    return s->d.precision == EMS_F32 ? autotune_tb2_impl<float>(s, trial_steps, cols, rows, planes)
                                     : autotune_tb2_impl<double>(s, trial_steps, cols, rows, planes);
}

ems_status ems_fdtd_set_zmarch(ems_fdtd *s, int rows, int planes, int waves)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (rows != 2 && rows != 4) return invalid("z-march strip rows must be 2 or 4");
    if (planes < 1) return invalid("z-march planes per chunk must be >= 1");
    if (waves != 1 && waves != 4) return invalid("z-march waves per workgroup must be 1 or 4");
    s->z_ty = rows;
    s->z_lz = planes;
    s->z_nw = waves;
    return EMS_OK;
}

ems_status ems_fdtd_set_tiling(ems_fdtd *s, int rows, int planes, int flags)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (rows != 8 && rows != 16) return invalid("tile rows must be 8 or 16");
    if (flags < 0 || flags > 7) return invalid("tile flags must be 0..7");
    s->tile_flags = flags;
    if (planes < 1) return invalid("tile planes must be >= 1");
    s->tile_by = rows;
    s->tile_lz = planes;
    s->split_lz = planes;
    s->tiles_ok = false;
    return EMS_OK;
}

ems_status ems_fdtd_run(ems_fdtd *s, int num_steps)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return s->d.precision == EMS_F32 ? run_impl<float>(s, num_steps, nullptr, nullptr)
                                     : run_impl<double>(s, num_steps, nullptr, nullptr);
}

ems_status ems_fdtd_run_static(ems_fdtd *s, int num_steps)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return s->d.precision == EMS_F32 ? run_static_impl<float>(s, num_steps) : run_static_impl<double>(s, num_steps);
}

ems_status ems_fdtd_run_timed(ems_fdtd *s, int num_steps, double *total_ms, double *kernel_ms)
{
    clear_error();
    if (!s || !total_ms || !kernel_ms) return invalid("null argument");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return s->d.precision == EMS_F32 ? run_impl<float>(s, num_steps, total_ms, kernel_ms)
                                     : run_impl<double>(s, num_steps, total_ms, kernel_ms);
}

ems_status ems_fdtd_trace_shape(const ems_fdtd *s, int *nobs, int *nslot)
{
    clear_error();
    if (!s || !nobs || !nslot) return invalid("null argument");
    *nobs = (int)s->obs.size();
    *nslot = s->nslot;
    return EMS_OK;
}

ems_status ems_fdtd_get_traces(ems_fdtd *s, double *out, size_t count)
{
    clear_error();
    if (!s || (!out && count)) return invalid("null argument");
    if (count != s->traces.size()) return invalid("trace buffer has the wrong size");
    std::copy(s->traces.begin(), s->traces.end(), out);
    return EMS_OK;
}

ems_status ems_fdtd_field_size(const ems_fdtd *s, int which, size_t *n)
{
    clear_error();
    if (!s || !n) return invalid("null argument");
    if (which < 0 || which > 5) return invalid("field selector must be 0..5");
    *n = ref_size(s->g, which);
    return EMS_OK;
}

ems_status ems_fdtd_get_field(ems_fdtd *s, int which, double *out, size_t count)
{
    clear_error();
    if (!s || !out) return invalid("null argument");
    if (which < 0 || which > 5) return invalid("field selector must be 0..5");
    if (count != ref_size(s->g, which)) return invalid("field buffer has the wrong size");
    return s->d.precision == EMS_F32 ? get_field_t<float>(s, which, out) : get_field_t<double>(s, which, out);
}

ems_status ems_fdtd_set_field(ems_fdtd *s, int which, const double *in, size_t count)
{
    clear_error();
    if (!s || !in) return invalid("null argument");
    if (which < 0 || which > 5) return invalid("field selector must be 0..5");
    if (count != ref_size(s->g, which)) return invalid("field buffer has the wrong size");
    return s->d.precision == EMS_F32 ? set_field_t<float>(s, which, in) : set_field_t<double>(s, which, in);
}

ems_status ems_fdtd_get_time(const ems_fdtd *s, double *t)
{
    clear_error();
    if (!s || !t) return invalid("null argument");
    *t = s->t_now;
    return EMS_OK;
}

// src/fdtd/fdtd_simulation.cpp:308-331 (file name, header and default
// 6-significant-digit stream formatting)
ems_status ems_fdtd_save_results(ems_fdtd *s, int num_steps, double c, const char *dir)
{
    (void)c;
    clear_error();
    if (!s) return invalid("null handle");
    const int nobs = (int)s->obs.size();
    for (int i = 0; i < nobs; ++i) {
        std::string name = "results_" + std::to_string(i) + "_" + std::to_string(s->d.dt) + "_" +
                           std::to_string(s->d.h) + "_" + std::to_string(num_steps) + ".csv";
        if (dir && dir[0]) name = std::string(dir) + "/" + name;
        std::ofstream file(name);
        if (!file.is_open()) {
            set_error("Error: Could not open file " + name);
            return EMS_ERR_IO;
        }
        file << "time,Ex,Ey,Ez" << std::endl;
        for (int j = 0; j < s->nslot; ++j) {
            const double *v = &s->traces[((size_t)i * s->nslot + j) * 3];
            file << j * s->d.tf * s->d.dt << "," << v[0] << "," << v[1] << "," << v[2] << std::endl;
        }
    }
    return EMS_OK;
}

ems_status ems_fdtd_device_bytes(const ems_fdtd *s, size_t *bytes)
{
    clear_error();
    if (!s || !bytes) return invalid("null argument");
    size_t b = 0;
    for (auto &f : s->f) b += f.bytes;
    b += s->ca.bytes + s->cb.bytes + s->prof.bytes;
    for (auto &p : s->psi) b += p.bytes;
    *bytes = b;
    return EMS_OK;
}

int ems_fdtd_check_params(const double domain_sizes[3], const int grid_num[3], double duration,
                          double time_step, double domain_size, double c)
{
    if (domain_sizes[0] <= 0 || domain_sizes[1] <= 0 || domain_sizes[2] <= 0) return 0;
    if (grid_num[0] <= 0 || grid_num[1] <= 0 || grid_num[2] <= 0) return 0;
    if (duration <= 0) return 0;
    if (time_step <= 0) return 0;
    if (std::round(domain_sizes[0] / grid_num[0]) != std::round(domain_sizes[1] / grid_num[1]) ||
        std::round(domain_sizes[0] / grid_num[0]) != std::round(domain_sizes[2] / grid_num[2]))
        return 0;
    if (c * time_step >= domain_size) return 0;
    return 1;
}

ems_status ems_fdtd_set_checkpoint_interval(ems_fdtd *s, int interval)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (interval < 0) return invalid("checkpoint interval must be >= 0");
    s->ckpt_interval = interval;
    return EMS_OK;
}

ems_status ems_fdtd_set_gradient_memory(ems_fdtd *s, size_t bytes)
{
    if (!s) return invalid("null handle");
    s->grad_mem = bytes;
    return EMS_OK;
}

ems_status ems_fdtd_gradient_plan(ems_fdtd *s, int num_steps, int *interval1, int *interval, size_t *bytes)
{
    if (!s) return invalid("null handle");
    if (num_steps < 1) return invalid("gradient plan: num_steps must be >= 1");
    EMS_TRY(setup_pml(s));  // the checkpoints hold the CPML state
    GradStore *gs = nullptr;
    EMS_TRY(grad_store(&gs));
    std::lock_guard<std::mutex> lock(gs->mu);
    GradPlan p;
    double b = 0.0;
    EMS_TRY(grad_plan_for(*gs, s, 1, num_steps, p, b));
    if (interval1) *interval1 = p.K1;
    if (interval) *interval = p.Ks;
    if (bytes) *bytes = (size_t)b;
    return EMS_OK;
}

ems_status ems_fdtd_gradient(ems_fdtd *s, int num_steps, const double *d_obs, size_t count, double *misfit,
                             double *grad_eps_r, double *grad_sigma)
{
    clear_error();
    if (!s || !d_obs) return invalid("null argument");
    if (num_steps <= 0) return invalid("num_steps must be >= 1");
    const size_t need = s->obs.size() * (size_t)(num_steps / s->d.tf + 1) * 3;
    if (count != need) return invalid("d_obs must hold nobs * (num_steps/tf + 1) * 3 values");
    return s->esz == 4 ? gradient_impl<float>(s, num_steps, d_obs, misfit, grad_eps_r, grad_sigma)
                       : gradient_impl<double>(s, num_steps, d_obs, misfit, grad_eps_r, grad_sigma);
}

ems_status ems_batch_create(ems_batch **out)
{
    clear_error();
    if (!out) return invalid("null argument");
    *out = nullptr;
    if (ems_device_count() <= 0) {
        set_error("no HIP device available");
        return EMS_ERR_HIP;
    }
    ems_batch *b = new ems_batch();
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    *out = b;
    return EMS_OK;
}

void ems_batch_destroy(ems_batch *b) { delete b; }

ems_status ems_batch_add(ems_batch *b, ems_fdtd *member)
{
    clear_error();
    if (!b || !member) return invalid("null argument");
    for (auto *s : b->mem)
        if (s == member) return invalid("handle already in the batch");
    if (!b->mem.empty() && b->mem[0]->d.precision != member->d.precision)
        return invalid("batch members must share one precision");
    b->mem.push_back(member);
    return EMS_OK;
}

ems_status ems_batch_set_tiling(ems_batch *b, int rows, int planes)
{
    clear_error();
    if (!b) return invalid("null batch");
    if (rows != 2 && rows != 4 && rows != 8) return invalid("rows must be 2, 4 or 8");
    if (planes < 1) return invalid("planes must be >= 1");
    b->by = rows;
    b->lz = planes;
    return EMS_OK;
}

ems_status ems_batch_autotune(ems_batch *b, int trial_steps, int *rows, int *planes)
{
    clear_error();
    EMS_TRY(batch_check(b));
    if (trial_steps < 1) return invalid("trial_steps must be >= 1");
    if (b->mem.empty()) return invalid("empty batch");
    return b->mem[0]->esz == 4 ? batch_autotune_impl<float>(b, trial_steps, rows, planes)
                               : batch_autotune_impl<double>(b, trial_steps, rows, planes);
}

ems_status ems_batch_set_graph(ems_batch *b, int steps_per_graph)
{
    clear_error();
    if (!b) return invalid("null batch");
    if (steps_per_graph < 0) return invalid("steps_per_graph must be >= 0");
    b->graph_steps = steps_per_graph;
    return EMS_OK;
}

ems_status ems_batch_run(ems_batch *b, int num_steps)
{
    clear_error();
    EMS_TRY(batch_check(b));
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    if (b->mem.empty()) return EMS_OK;
    return b->mem[0]->esz == 4 ? batch_run_impl<float>(b, num_steps, nullptr)
                               : batch_run_impl<double>(b, num_steps, nullptr);
}

ems_status ems_batch_run_timed(ems_batch *b, int num_steps, double *total_ms)
{
    clear_error();
    EMS_TRY(batch_check(b));
    if (!total_ms) return invalid("null argument");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    if (b->mem.empty()) return invalid("empty batch");
    return b->mem[0]->esz == 4 ? batch_run_impl<float>(b, num_steps, total_ms)
                               : batch_run_impl<double>(b, num_steps, total_ms);
}

ems_status ems_batch_set_gradient_group(ems_batch *b, int shots)
{
    clear_error();
    if (!b) return invalid("null batch");
    if (shots < 0) return invalid("gradient group must be >= 0");
    b->grad_group = shots;
    return EMS_OK;
}

ems_status ems_batch_gradient(ems_batch *b, int num_steps, const double *const *d_obs, double *misfits,
                              double *misfit_total, double *grad_eps_r, double *grad_sigma)
{
    clear_error();
    EMS_TRY(batch_check(b));
    if (!d_obs) return invalid("null argument");
    if (num_steps <= 0) return invalid("num_steps must be >= 1");
    if (b->mem.empty()) return invalid("empty batch");
    for (size_t q = 0; q < b->mem.size(); ++q)
        if (!d_obs[q]) return invalid("null d_obs entry");
    return b->mem[0]->esz == 4
               ? batch_gradient_impl<float>(b, num_steps, d_obs, misfits, misfit_total, grad_eps_r, grad_sigma)
               : batch_gradient_impl<double>(b, num_steps, d_obs, misfits, misfit_total, grad_eps_r, grad_sigma);
}

}  // extern "C"
