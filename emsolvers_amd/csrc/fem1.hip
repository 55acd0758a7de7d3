// First-order stabilized hex FEM (src/fem leapfrog, src/fem_abm
// Adams-Bashforth-Moulton) behind the C ABI; one GPU owns the whole domain.
#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "fem1_kernels.hpp"
#include "fem_kernels.hpp"

using namespace ems;
using namespace ems::fem1;

namespace {
#include "fem1_tables.inc"
}

struct ems_fem1 {
    ems_fem1_desc d{};
    Hex1 g{};
    size_t esz = 8;
    double K[144] = {};
    DevBuf dK, dKt;
    DevBuf dKs, dKts;  // merged stencils [3][33] of K and K^T
    int elem_kind = EMS_FEM_ELEM_AUTO;  // STENCIL (AUTO) or the element row dots (ROWS)
    // E/H: 3 rotating slots; Ep/Hp: predicted fields; E3/H3: scratch
    DevBuf E[3][3], H[3][3], Ep[3], Hp[3], E3[3], H3[3];
    int t0 = 0, t1 = 1, t2 = 2;
    std::vector<SrcOp> ops;
    std::vector<std::array<int, 3>> dipoles;
    std::vector<std::array<int, 3>> cnt;      // per observation, points per component
    std::vector<std::vector<long>> oidx;      // per observation*3
    std::vector<double> traces;               // [nobs][nslot][3]
    int nslot = 0;
    double t_now = 0.0;
    hipStream_t stream = nullptr;
    DevBuf d_ops, d_val, d_ocnt, d_oidx, d_ooff, d_tr;
};

namespace {

bool abm(const ems_fem1 *s) { return s->d.stepper == EMS_FEM1_ABM; }

long fidx(const Hex1 &g, int comp, int i, int j, int k)
{
    const long Nx = g.Nx, Ny = g.Ny;
    if (comp == 0) return i + j * Nx + k * Nx * (Ny + 1);
    if (comp == 1) return i + j * (Nx + 1) + k * (Nx + 1) * Ny;
    return i + j * (Nx + 1) + k * (Nx + 1) * (Ny + 1);
}

bool in_comp(const Hex1 &g, int comp, int i, int j, int k)
{
    if (i < 0 || j < 0 || k < 0) return false;
    return i <= g.Nx - (comp == 0) && j <= g.Ny - (comp == 1) && k <= g.Nz - (comp == 2);
}

bool pec_host(const Hex1 &g, int comp, int i, int j, int k)
{
    if (comp == 0) return j == 0 || j == g.Ny || k == 0 || k == g.Nz;
    if (comp == 1) return i == 0 || i == g.Nx || k == 0 || k == g.Nz;
    return i == 0 || i == g.Nx || j == 0 || j == g.Ny;
}

// src/fem/fem_simulation.cpp:29-39
double sf_fem(double t, double mu, double h)
{
    const double f0 = 1.0e9;
    return 4.0 * M_PI * M_PI * f0 * f0 * (1.0 - 4.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) * (t - 1.0 / f0)) *
           std::exp(-2.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) * (t - 1.0 / f0)) * mu / h / h / h * 0.001;
}

// src/fem_abm/fem_simulation.cpp:31-39
double sf_abm(double t, double h)
{
    const double f0 = 1.0e9;
    return -4.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) *
           std::exp(-2.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) * (t - 1.0 / f0)) / h / h / h * 0.001;
}

void table(const signed short *num, int den, double *out)
{
    for (int q = 0; q < 144; ++q) out[q] = (double)num[q] / (double)den;
}

// X_s = X + s * _X entrywise in double (Amat_Eigen + s * _Amat_Eigen)
void stabilized(const signed short *a, int da, const signed short *b, int db, double s, double *out)
{
    double x[144], y[144];
    table(a, da, x);
    table(b, db, y);
    for (int q = 0; q < 144; ++q) out[q] = x[q] + s * y[q];
}

// Element matrix of the step (src/fem/fem_simulation.cpp:125-434,
// src/fem_abm/fem_simulation.cpp:143-246).  fem: K = c C_s B_s^-1 A_s with
// c = 0.5*dt*dt/mu/eps/h/h; the product is formed in long double
// (Gauss-Jordan with partial pivoting on B_s) and rounded once.  fem_abm:
// K = A_s.
ems_status element_matrix(const ems_fem1_desc *d, double *K)
{
    if (d->stepper == EMS_FEM1_ABM) {
        stabilized(kfem_abm_Amat, kfem_abm_Amat_den, kfem_abm_sAmat, kfem_abm_sAmat_den, d->stab, K);
        return EMS_OK;
    }
    double A[144], B[144], C[144];
    if (d->use_ofem) {
        stabilized(kfem_ofem1_Amat, kfem_ofem1_Amat_den, kfem_ofem1_sAmat, kfem_ofem1_sAmat_den, d->stab, A);
        stabilized(kfem_ofem1_Bmat, kfem_ofem1_Bmat_den, kfem_ofem1_sBmat, kfem_ofem1_sBmat_den, d->stab, B);
        stabilized(kfem_ofem1_Cmat, kfem_ofem1_Cmat_den, kfem_ofem1_sCmat, kfem_ofem1_sCmat_den, d->stab, C);
    } else {
        stabilized(kfem_ofem0_Amat, kfem_ofem0_Amat_den, kfem_ofem0_sAmat, kfem_ofem0_sAmat_den, d->stab, A);
        stabilized(kfem_ofem0_Bmat, kfem_ofem0_Bmat_den, kfem_ofem0_sBmat, kfem_ofem0_sBmat_den, d->stab, B);
        stabilized(kfem_ofem0_Cmat, kfem_ofem0_Cmat_den, kfem_ofem0_sCmat, kfem_ofem0_sCmat_den, d->stab, C);
    }
    // [B | A] -> [I | B^-1 A]
    long double M[12][24];
    for (int r = 0; r < 12; ++r)
        for (int c = 0; c < 12; ++c) {
            M[r][c] = B[r * 12 + c];
            M[r][12 + c] = A[r * 12 + c];
        }
    for (int col = 0; col < 12; ++col) {
        int piv = col;
        for (int r = col + 1; r < 12; ++r)
            if (std::fabs((double)M[r][col]) > std::fabs((double)M[piv][col])) piv = r;
        if (M[piv][col] == 0.0L) return invalid("stabilized B matrix is singular for this stabilization factor");
        if (piv != col)
            for (int c = 0; c < 24; ++c) std::swap(M[piv][c], M[col][c]);
        const long double pv = M[col][col];
        for (int c = 0; c < 24; ++c) M[col][c] /= pv;
        for (int r = 0; r < 12; ++r) {
            if (r == col || M[r][col] == 0.0L) continue;
            const long double f = M[r][col];
            for (int c = 0; c < 24; ++c) M[r][c] -= f * M[col][c];
        }
    }
    const double sc = 0.5 * d->dt * d->dt / d->mu / d->eps / d->h / d->h;
    for (int r = 0; r < 12; ++r)
        for (int c = 0; c < 12; ++c) {
            long double acc = 0.0L;
            for (int q = 0; q < 12; ++q) acc += (long double)C[r * 12 + q] * M[q][12 + c];
            K[r * 12 + c] = (double)((long double)sc * acc);
        }
    return EMS_OK;
}

ems_status validate(const ems_fem1_desc *d)
{
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return invalid("grid sizes must be positive");
    if (d->tf <= 0) return invalid("time_frequency must be >= 1");
    if (!(d->h > 0) || !(d->dt > 0) || !(d->eps > 0) || !(d->mu > 0))
        return invalid("h, dt, eps and mu must be positive");
    if (d->stepper != EMS_FEM1_LEAPFROG && d->stepper != EMS_FEM1_ABM) return invalid("unknown stepper");
    if (d->use_ofem != 0 && d->use_ofem != 1) return invalid("use_ofem must be 0 or 1");
    if (d->stepper == EMS_FEM1_ABM && d->use_ofem) return invalid("fem_abm has no use_ofem variant");
    if (!std::isfinite(d->stab)) return invalid("stabilization factor must be finite");
    if (d->precision != EMS_F64 && d->precision != EMS_F32) return invalid("unknown precision");
    return EMS_OK;
}

template <typename T> CF3<T> cf(DevBuf *b)
{
    return CF3<T>{{b[0].as<T>(), b[1].as<T>(), b[2].as<T>()}};
}
template <typename T> F3<T> mf(DevBuf *b) { return F3<T>{{b[0].as<T>(), b[1].as<T>(), b[2].as<T>()}}; }

template <int COMP> dim3 grid_of(const Hex1 &g)
{
    const int dx = g.Nx + (COMP != 0), dy = g.Ny + (COMP != 1), dz = g.Nz + (COMP != 2);
    return dim3((dx + 63) / 64, (dy + 3) / 4, dz);
}

bool stencil(const ems_fem1 *s) { return s->elem_kind != EMS_FEM_ELEM_ROWS; }

// The leapfrog update E2 = (0 - K E1) + (2 E1 - E0) with PEC zeros is the
// fem_v2 one on the same layouts: with the merged stencil it runs on that
// family's z-marching kernel (no sources there: k_fem1_post adds them), 64 x 8
// points per workgroup, 16 planes (FP32 8), as measured fastest at 128^3.
fem::FemGrid z_grid(const ems_fem1 *s)
{
    fem::FemGrid g{};
    g.p = 1;
    g.Nx = g.nex = s->g.Nx;
    g.Ny = g.ney = s->g.Ny;
    g.Nz = g.nez = s->g.Nz;
    g.NX = s->g.n[0];
    g.NY = s->g.n[1];
    g.NZ = s->g.n[2];
    g.D = 4;
    g.n = 12;
    return g;
}

// Largest per-component count of face edges (k_fem1_abm_faces).
long face_edges(const Hex1 &g)
{
    const long N[3] = {g.Nx, g.Ny, g.Nz};
    long m = 0;
    for (int c = 0; c < 3; ++c) {
        const int a1 = c == 0 ? 1 : 0, a2 = c == 2 ? 1 : 2;
        m = std::max(m, N[c] * (2 * (N[a1] + 1) + 2 * (N[a2] - 1)));
    }
    return m;
}

template <typename T> void launch_leap_stencil(ems_fem1 *s)
{
    const fem::FemGrid g = z_grid(s);
    const int lz = sizeof(T) == 4 ? 8 : 16;
    auto c3 = [](DevBuf *b) { return fem::CComp3<T>{b[0].as<T>(), b[1].as<T>(), b[2].as<T>()}; };
    const fem::CComp3<T> e0 = c3(s->E[s->t0]), e1 = c3(s->E[s->t1]);
    DevBuf *o = s->E[s->t2];
    const fem::Comp3<T> e2{o[0].as<T>(), o[1].as<T>(), o[2].as<T>()};
    const dim3 grd((g.Nx + 63) / 64, (g.Ny + 7) / 8, (g.Nz + lz - 1) / lz);
    fem::k_fem1z<T, 8, false, true><<<grd, dim3(64, 8, 1), 0, s->stream>>>(
        g, s->dK.as<T>(), s->dKs.as<T>(), e1, e1, e0, e2, e2, lz, (T)0, (T)0, (T)1, 0, nullptr, nullptr);
}

template <typename T> void launch_leap(ems_fem1 *s)
{
    if (stencil(s)) return launch_leap_stencil<T>(s);
    const dim3 blk(64, 4, 1);
    const CF3<T> e0 = cf<T>(s->E[s->t0]), e1 = cf<T>(s->E[s->t1]);
    const F3<T> e2 = mf<T>(s->E[s->t2]);
    const T *K = s->dK.as<T>();
    k_fem1_leap<T, 0><<<grid_of<0>(s->g), blk, 0, s->stream>>>(s->g, K, e0, e1, e2);
    k_fem1_leap<T, 1><<<grid_of<1>(s->g), blk, 0, s->stream>>>(s->g, K, e0, e1, e2);
    k_fem1_leap<T, 2><<<grid_of<2>(s->g), blk, 0, s->stream>>>(s->g, K, e0, e1, e2);
}

template <typename T, bool CORR> void launch_abm(ems_fem1 *s)
{
    const long nmax = std::max(s->g.n[0], std::max(s->g.n[1], s->g.n[2]));
    const dim3 sg((unsigned)((nmax + 255) / 256), 3, 1);
    const T dt = (T)s->d.dt, eps = (T)s->d.eps, mu = (T)s->d.mu, h = (T)s->d.h;
    k_fem1_scratch<T, CORR><<<sg, 256, 0, s->stream>>>(s->g, cf<T>(s->E[s->t0]), cf<T>(s->E[s->t1]), cf<T>(s->Ep),
                                                       cf<T>(s->H[s->t0]), cf<T>(s->H[s->t1]), cf<T>(s->Hp),
                                                       mf<T>(s->E3), mf<T>(s->H3), dt, eps, mu, h);
    const dim3 blk(64, 4, 1);
    const CF3<T> e3 = cf<T>(s->E3), h3 = cf<T>(s->H3), e1 = cf<T>(s->E[s->t1]), h1 = cf<T>(s->H[s->t1]);
    const F3<T> eo = CORR ? mf<T>(s->E[s->t2]) : mf<T>(s->Ep);
    const F3<T> ho = CORR ? mf<T>(s->H[s->t2]) : mf<T>(s->Hp);
    const T *K = s->dK.as<T>(), *Kt = s->dKt.as<T>(), *cK = s->dKs.as<T>(), *cKt = s->dKts.as<T>();
    if (stencil(s)) {
        // interior edges: the z-marching stencil pass per side; face edges:
        // the element row dots
        const fem::FemGrid fg = z_grid(s);
        const int lz = sizeof(T) == 4 ? 8 : 16;
        const dim3 grd((fg.Nx + 63) / 64, (fg.Ny + 7) / 8, (fg.Nz + lz - 1) / lz);
        auto c3 = [](const CF3<T> &f) { return fem::CComp3<T>{f.c[0], f.c[1], f.c[2]}; };
        auto m3 = [](const F3<T> &f) { return fem::Comp3<T>{f.c[0], f.c[1], f.c[2]}; };
        fem::k_fem1z_sum<T, 8><<<grd, dim3(64, 8, 1), 0, s->stream>>>(fg, cK, c3(h3), c3(h1), m3(ho), lz);
        fem::k_fem1z_sum<T, 8><<<grd, dim3(64, 8, 1), 0, s->stream>>>(fg, cKt, c3(e3), c3(e1), m3(eo), lz);
        const long nf = face_edges(s->g);
        k_fem1_abm_faces<T, CORR><<<dim3((unsigned)((nf + 255) / 256), 3, 1), 256, 0, s->stream>>>(
            s->g, K, Kt, e3, h3, e1, h1, eo, ho);
        return;
    }
    k_fem1_abm<T, 0, CORR><<<grid_of<0>(s->g), blk, 0, s->stream>>>(s->g, K, Kt, e3, h3, e1, h1, eo, ho);
    k_fem1_abm<T, 1, CORR><<<grid_of<1>(s->g), blk, 0, s->stream>>>(s->g, K, Kt, e3, h3, e1, h1, eo, ho);
    k_fem1_abm<T, 2, CORR><<<grid_of<2>(s->g), blk, 0, s->stream>>>(s->g, K, Kt, e3, h3, e1, h1, eo, ho);
}

template <typename T> void launch_post(ems_fem1 *s, F3<T> e, const double *val, int skip_pec, int row)
{
    const int nobs = (int)s->cnt.size();
    k_fem1_post<T><<<1, 256, 0, s->stream>>>(e, (int)s->ops.size(), s->d_ops.as<SrcOp>(), val, skip_pec, nobs,
                                             s->d_ocnt.as<int>(), s->d_oidx.as<long>(), s->d_ooff.as<long>(),
                                             s->d_tr.as<double>(), s->nslot, row);
}

template <typename T> ems_status run_impl(ems_fem1 *s, int num_steps, double *total_ms, double *kernel_ms)
{
    const int tf = s->d.tf;
    const int nobs = (int)s->cnt.size();
    // saved_electric_field_[i].resize(round(num_steps / tf) + 1), old values kept
    const int nslot = num_steps / tf + 1;
    if (nslot != s->nslot) {
        std::vector<double> t((size_t)nobs * nslot * 3, 0.0);
        for (int o = 0; o < nobs; ++o)
            for (int q = 0; q < std::min(nslot, s->nslot); ++q)
                for (int c = 0; c < 3; ++c) t[((size_t)o * nslot + q) * 3 + c] = s->traces[((size_t)o * s->nslot + q) * 3 + c];
        s->traces.swap(t);
        s->nslot = nslot;
    }
    if (num_steps <= 0) return EMS_OK;
    // per-step source values: [step][stage][point, dipole]
    const bool A = abm(s);
    const int nst = A ? 2 : 1;
    const double dt = s->d.dt, eps = s->d.eps, mu = s->d.mu, h = s->d.h;
    std::vector<double> val((size_t)num_steps * nst * 2);
    double t = s->t_now;
    for (int st = 0; st < num_steps; ++st) {
        double *v = &val[(size_t)st * nst * 2];
        if (!A) {
            v[0] = sf_fem(t, mu, h) * dt * dt / mu / eps / h / h / h;
            v[1] = sf_fem(t, mu, h) * dt * dt / mu / eps / 8.0;
        } else {
            v[0] = (3.0 * sf_abm(t, h) - sf_abm(t - dt, h)) * dt / 2.0 / eps / h / h / h;
            v[1] = -(3.0 * sf_abm(t, h) - sf_abm(t - dt, h)) * dt / 16.0 / eps;
            v[2] = (5.0 * sf_abm(t + dt, h) + 8.0 * sf_abm(t, h) - sf_abm(t - dt, h)) * dt / 12.0 / eps / h / h / h;
            v[3] = -(5.0 * sf_abm(t + dt, h) + 8.0 * sf_abm(t, h) - sf_abm(t - dt, h)) * dt / 96.0 / eps;
        }
        t += dt;
    }
    EMS_TRY(s->d_val.reserve(sizeof(double) * val.size()));
    EMS_HIP(hipMemcpy(s->d_val.p, val.data(), sizeof(double) * val.size(), hipMemcpyHostToDevice));
    EMS_TRY(s->d_tr.reserve(sizeof(double) * std::max(s->traces.size(), (size_t)1)));
    if (!s->traces.empty())
        EMS_HIP(hipMemcpy(s->d_tr.p, s->traces.data(), sizeof(double) * s->traces.size(), hipMemcpyHostToDevice));

    const bool timed = total_ms != nullptr;
    std::vector<hipEvent_t> ev;
    if (timed) {
        ev.resize(2 * (size_t)num_steps + 2, nullptr);
        for (auto &e : ev) EMS_HIP(hipEventCreate(&e));
        EMS_HIP(hipEventRecord(ev[2 * (size_t)num_steps], s->stream));
    }
    const double *dv = s->d_val.as<double>();
    for (int st = 1; st <= num_steps; ++st) {
        const int row = (st % tf == 0) ? st / tf : -1;
        const double *v = dv + (size_t)(st - 1) * nst * 2;
        if (timed) EMS_HIP(hipEventRecord(ev[2 * (size_t)(st - 1)], s->stream));
        if (!A) {
            launch_leap<T>(s);
            if (timed) EMS_HIP(hipEventRecord(ev[2 * (size_t)(st - 1) + 1], s->stream));
            launch_post<T>(s, mf<T>(s->E[s->t2]), v, 1, row);
        } else {
            launch_abm<T, false>(s);
            launch_post<T>(s, mf<T>(s->Ep), v, 0, -1);
            launch_abm<T, true>(s);
            if (timed) EMS_HIP(hipEventRecord(ev[2 * (size_t)(st - 1) + 1], s->stream));
            launch_post<T>(s, mf<T>(s->E[s->t2]), v + 2, 1, row);
        }
        EMS_HIP(hipGetLastError());
        const int o0 = s->t0;
        s->t0 = s->t1;
        s->t1 = s->t2;
        s->t2 = o0;
    }
    if (timed) {
        EMS_HIP(hipEventRecord(ev[2 * (size_t)num_steps + 1], s->stream));
        EMS_HIP(hipEventSynchronize(ev[2 * (size_t)num_steps + 1]));
        double ksum = 0;
        for (int st = 0; st < num_steps; ++st) {
            float ms = 0;
            EMS_HIP(hipEventElapsedTime(&ms, ev[2 * (size_t)st], ev[2 * (size_t)st + 1]));
            ksum += ms;
        }
        float ms = 0;
        EMS_HIP(hipEventElapsedTime(&ms, ev[2 * (size_t)num_steps], ev[2 * (size_t)num_steps + 1]));
        *total_ms = ms;
        *kernel_ms = ksum;
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    EMS_HIP(hipStreamSynchronize(s->stream));
    s->t_now = t;
    if (!s->traces.empty())
        EMS_HIP(hipMemcpy(s->traces.data(), s->d_tr.p, sizeof(double) * s->traces.size(), hipMemcpyDeviceToHost));
    return EMS_OK;
}

ems_status upload_ops(ems_fem1 *s)
{
    if (s->ops.empty()) return EMS_OK;
    EMS_TRY(s->d_ops.reserve(sizeof(SrcOp) * s->ops.size()));
    EMS_HIP(hipMemcpy(s->d_ops.p, s->ops.data(), sizeof(SrcOp) * s->ops.size(), hipMemcpyHostToDevice));
    return EMS_OK;
}

// Rebuild the source-op list in the reference's order: x, y, z point
// sources, then every dipole's pattern (src/fem/fem_simulation.cpp:613-740).
ems_status rebuild_ops(ems_fem1 *s, const std::vector<SrcOp> &points)
{
    std::vector<SrcOp> ops = points;
    std::stable_sort(ops.begin(), ops.end(), [](const SrcOp &a, const SrcOp &b) { return a.comp < b.comp; });
    const bool A = abm(s);
    const double sh = s->d.stab / 2.0;
    for (auto &dp : s->dipoles) {
        const int i = dp[0], j = dp[1], k = dp[2];
        auto add = [&](int comp, int a, int b, int c, double coef, int sub) {
            SrcOp o{};
            o.comp = comp;
            o.idx = fidx(s->g, comp, a, b, c);
            o.coef = coef;
            o.kind = 1;
            o.subtract = sub;
            o.pec = pec_host(s->g, comp, a, b, c);
            ops.push_back(o);
        };
        if (!A && s->d.use_ofem) {
            add(0, i, j + 1, k + 1, 8.0, 0);
            continue;
        }
        // x-edges of the 2 x 2 element block (conn_0 .. conn_3)
        add(0, i, j, k, 1.0 / 8.0, 0);
        add(0, i, j + 1, k, 3.0 / 4.0, 0);
        add(0, i, j, k + 1, 3.0 / 4.0, 0);
        add(0, i, j + 1, k + 1, 9.0 / 2.0, 0);
        add(0, i, j + 2, k, 1.0 / 8.0, 0);
        add(0, i, j + 2, k + 1, 3.0 / 4.0, 0);
        add(0, i, j, k + 2, 1.0 / 8.0, 0);
        add(0, i, j + 1, k + 2, 3.0 / 4.0, 0);
        add(0, i, j + 2, k + 2, 1.0 / 8.0, 0);
        // stabilization terms on y- and z-edges
        add(1, i, j, k, sh, 0);
        add(1, i + 1, j, k, sh, 0);
        add(1, i, j + 1, k, sh, 0);
        add(1, i + 1, j + 1, k, sh, 0);
        add(1, i, j, k + 2, sh, 1);
        add(1, i + 1, j, k + 2, sh, 1);
        add(1, i, j + 1, k + 2, sh, 1);
        add(1, i + 1, j + 1, k + 2, sh, 1);
        add(2, i, j, k, sh, 1);
        add(2, i + 1, j, k, sh, 1);
        add(2, i, j + 2, k, sh, 0);
        add(2, i + 1, j + 2, k, sh, 0);
        add(2, i, j, k + 1, sh, 1);
        add(2, i + 1, j, k + 1, sh, 1);
        add(2, i, j + 2, k + 1, sh, 0);
        add(2, i + 1, j + 2, k + 1, sh, 0);
    }
    s->ops.swap(ops);
    return upload_ops(s);
}

std::vector<SrcOp> points_of(const ems_fem1 *s)
{
    std::vector<SrcOp> p;
    for (auto &o : s->ops)
        if (o.kind == 0) p.push_back(o);
    return p;
}

ems_status upload_obs(ems_fem1 *s)
{
    std::vector<int> cnt;
    std::vector<long> idx, off;
    for (size_t o = 0; o < s->cnt.size(); ++o)
        for (int q = 0; q < 3; ++q) {
            cnt.push_back(s->cnt[o][q]);
            off.push_back((long)idx.size());
            idx.insert(idx.end(), s->oidx[o * 3 + q].begin(), s->oidx[o * 3 + q].end());
        }
    if (idx.empty()) idx.push_back(0);
    EMS_TRY(s->d_ocnt.reserve(sizeof(int) * cnt.size()));
    EMS_TRY(s->d_ooff.reserve(sizeof(long) * off.size()));
    EMS_TRY(s->d_oidx.reserve(sizeof(long) * idx.size()));
    EMS_HIP(hipMemcpy(s->d_ocnt.p, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(s->d_ooff.p, off.data(), sizeof(long) * off.size(), hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(s->d_oidx.p, idx.data(), sizeof(long) * idx.size(), hipMemcpyHostToDevice));
    return EMS_OK;
}

DevBuf *field_buf(ems_fem1 *s, int field, int slot)
{
    const int sl = slot == 0 ? s->t0 : s->t1;
    return field < 3 ? &s->E[sl][field] : &s->H[sl][field - 3];
}

template <typename T> ems_status upload_k(ems_fem1 *s)
{
    std::vector<T> k(144), kt(144);
    for (int r = 0; r < 12; ++r)
        for (int c = 0; c < 12; ++c) {
            k[r * 12 + c] = (T)s->K[r * 12 + c];
            kt[r * 12 + c] = (T)s->K[c * 12 + r];
        }
    EMS_TRY(s->dK.reserve(sizeof(T) * 144));
    EMS_TRY(s->dKt.reserve(sizeof(T) * 144));
    EMS_HIP(hipMemcpy(s->dK.p, k.data(), sizeof(T) * 144, hipMemcpyHostToDevice));
    EMS_HIP(hipMemcpy(s->dKt.p, kt.data(), sizeof(T) * 144, hipMemcpyHostToDevice));
    // merged stencils (src/fem numbering: y edges 4 + 2a + c), from FP64
    double ktd[144], c[3 * kZ1Terms];
    for (int r = 0; r < 12; ++r)
        for (int q = 0; q < 12; ++q) ktd[r * 12 + q] = s->K[q * 12 + r];
    DevBuf *dst[2] = {&s->dKs, &s->dKts};
    const double *src[2] = {s->K, ktd};
    for (int w = 0; w < 2; ++w) {
        if (!z1_stencil_coefs(src[w], true, c)) return not_supported("first-order stencil term missing");
        std::vector<T> ct(c, c + 3 * kZ1Terms);
        EMS_TRY(dst[w]->reserve(sizeof(T) * ct.size()));
        EMS_HIP(hipMemcpy(dst[w]->p, ct.data(), sizeof(T) * ct.size(), hipMemcpyHostToDevice));
    }
    return EMS_OK;
}

ems_status alloc_zero(DevBuf &b, size_t n)
{
    EMS_TRY(b.reserve(n));
    EMS_HIP(hipMemset(b.p, 0, n));
    return EMS_OK;
}

ems_status init_device(ems_fem1 *s)
{
    EMS_TRY(s->esz == 4 ? upload_k<float>(s) : upload_k<double>(s));
    const bool A = abm(s);
    for (int q = 0; q < 3; ++q) {
        const size_t nb = s->g.n[q] * s->esz;
        for (int sl = 0; sl < 3; ++sl) {
            EMS_TRY(alloc_zero(s->E[sl][q], nb));
            if (A) EMS_TRY(alloc_zero(s->H[sl][q], nb));
        }
        if (A) {
            EMS_TRY(alloc_zero(s->Ep[q], nb));
            EMS_TRY(alloc_zero(s->Hp[q], nb));
            EMS_TRY(alloc_zero(s->E3[q], nb));
            EMS_TRY(alloc_zero(s->H3[q], nb));
        }
    }
    return upload_obs(s);
}

}  // namespace

extern "C" {

ems_status ems_fem1_element_matrix(const ems_fem1_desc *desc, double *out)
{
    clear_error();
    if (!desc || !out) return invalid("null argument");
    EMS_TRY(validate(desc));
    return element_matrix(desc, out);
}

ems_status ems_fem1_create(const ems_fem1_desc *desc, ems_fem1 **out)
{
    clear_error();
    if (!desc || !out) return invalid("ems_fem1_create: null argument");
    *out = nullptr;
    EMS_TRY(validate(desc));
    if (ems_device_count() <= 0) {
        set_error("no HIP device available");
        return EMS_ERR_HIP;
    }
    ems_fem1 *s = new ems_fem1();
    s->d = *desc;
    s->esz = desc->precision == EMS_F32 ? 4 : 8;
    Hex1 &g = s->g;
    g.Nx = desc->nx;
    g.Ny = desc->ny;
    g.Nz = desc->nz;
    const long Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    g.n[0] = Nx * (Ny + 1) * (Nz + 1);
    g.n[1] = (Nx + 1) * Ny * (Nz + 1);
    g.n[2] = (Nx + 1) * (Ny + 1) * Nz;
    ems_status st = element_matrix(desc, s->K);
    if (st == EMS_OK) st = init_device(s);
    if (st != EMS_OK) {
        delete s;
        return st;
    }
    *out = s;
    return EMS_OK;
}

void ems_fem1_destroy(ems_fem1 *s) { delete s; }

ems_status ems_fem1_add_source(ems_fem1 *s, int comp, int i, int j, int k)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (comp < 0 || comp > 2) return invalid("source component must be 0, 1 or 2");
    if (!in_comp(s->g, comp, i, j, k)) return invalid("source outside the field array");
    std::vector<SrcOp> p = points_of(s);
    SrcOp o{};
    o.comp = comp;
    o.idx = fidx(s->g, comp, i, j, k);
    o.kind = 0;
    o.subtract = abm(s) ? 1 : 0;
    o.pec = pec_host(s->g, comp, i, j, k);
    p.push_back(o);
    return rebuild_ops(s, p);
}

ems_status ems_fem1_add_dipole_x(ems_fem1 *s, int i, int j, int k)
{
    clear_error();
    if (!s) return invalid("null handle");
    // the pattern touches elements (i, j..j+1, k..k+1) and edges up to j+2, k+2
    if (i < 0 || j < 0 || k < 0 || i + 1 > s->g.Nx || j + 2 > s->g.Ny || k + 2 > s->g.Nz)
        return invalid("dipole pattern outside the mesh");
    s->dipoles.push_back({i, j, k});
    return rebuild_ops(s, points_of(s));
}

ems_status ems_fem1_add_observation(ems_fem1 *s, const int counts[3], const int *pts)
{
    clear_error();
    if (!s || !counts) return invalid("null argument");
    std::array<int, 3> c = {counts[0], counts[1], counts[2]};
    if (c[0] < 0 || c[1] < 0 || c[2] < 0) return invalid("negative point count");
    if ((c[0] + c[1] + c[2]) > 0 && !pts) return invalid("null point list");
    std::vector<std::vector<long>> ix(3);
    int p = 0;
    for (int q = 0; q < 3; ++q)
        for (int a = 0; a < c[q]; ++a, ++p) {
            const int *v = pts + 3 * p;
            if (!in_comp(s->g, q, v[0], v[1], v[2])) return invalid("observation point outside the field array");
            ix[q].push_back(fidx(s->g, q, v[0], v[1], v[2]));
        }
    s->cnt.push_back(c);
    for (int q = 0; q < 3; ++q) s->oidx.push_back(ix[q]);
    std::vector<double> t(s->cnt.size() * (size_t)s->nslot * 3, 0.0);
    std::copy(s->traces.begin(), s->traces.end(), t.begin());
    s->traces.swap(t);
    return upload_obs(s);
}

ems_status ems_fem1_run(ems_fem1 *s, int num_steps)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return s->esz == 4 ? run_impl<float>(s, num_steps, nullptr, nullptr)
                       : run_impl<double>(s, num_steps, nullptr, nullptr);
}

ems_status ems_fem1_run_timed(ems_fem1 *s, int num_steps, double *total_ms, double *kernel_ms)
{
    clear_error();
    if (!s || !total_ms || !kernel_ms) return invalid("null argument");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return s->esz == 4 ? run_impl<float>(s, num_steps, total_ms, kernel_ms)
                       : run_impl<double>(s, num_steps, total_ms, kernel_ms);
}

ems_status ems_fem1_trace_shape(const ems_fem1 *s, int *nobs, int *nslot)
{
    clear_error();
    if (!s || !nobs || !nslot) return invalid("null argument");
    *nobs = (int)s->cnt.size();
    *nslot = s->nslot;
    return EMS_OK;
}

ems_status ems_fem1_get_traces(ems_fem1 *s, double *out, size_t count)
{
    clear_error();
    if (!s || (!out && count)) return invalid("null argument");
    if (count != s->traces.size()) return invalid("trace buffer has the wrong size");
    std::copy(s->traces.begin(), s->traces.end(), out);
    return EMS_OK;
}

ems_status ems_fem1_field_size(const ems_fem1 *s, int field, size_t *n)
{
    clear_error();
    if (!s || !n) return invalid("null argument");
    if (field < 0 || field > 5 || (field > 2 && !abm(s))) return invalid("field must be 0..2 (fem) or 0..5 (fem_abm)");
    *n = (size_t)s->g.n[field % 3];
    return EMS_OK;
}

ems_status ems_fem1_get_field(ems_fem1 *s, int field, int slot, double *out, size_t count)
{
    clear_error();
    if (!s || !out) return invalid("null argument");
    if (field < 0 || field > 5 || (field > 2 && !abm(s))) return invalid("field must be 0..2 (fem) or 0..5 (fem_abm)");
    if (slot < 0 || slot > 1) return invalid("slot must be 0 or 1");
    if (count != (size_t)s->g.n[field % 3]) return invalid("field buffer has the wrong size");
    EMS_HIP(hipStreamSynchronize(s->stream));
    DevBuf *b = field_buf(s, field, slot);
    if (s->esz == 8) {
        EMS_HIP(hipMemcpy(out, b->p, count * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t(count);
        EMS_HIP(hipMemcpy(t.data(), b->p, count * 4, hipMemcpyDeviceToHost));
        std::copy(t.begin(), t.end(), out);
    }
    return EMS_OK;
}

ems_status ems_fem1_set_field(ems_fem1 *s, int field, int slot, const double *in, size_t count)
{
    clear_error();
    if (!s || !in) return invalid("null argument");
    if (field < 0 || field > 5 || (field > 2 && !abm(s))) return invalid("field must be 0..2 (fem) or 0..5 (fem_abm)");
    if (slot < 0 || slot > 1) return invalid("slot must be 0 or 1");
    if (count != (size_t)s->g.n[field % 3]) return invalid("field buffer has the wrong size");
    EMS_HIP(hipStreamSynchronize(s->stream));
    DevBuf *b = field_buf(s, field, slot);
    if (s->esz == 8) {
        EMS_HIP(hipMemcpy(b->p, in, count * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(in, in + count);
        EMS_HIP(hipMemcpy(b->p, t.data(), count * 4, hipMemcpyHostToDevice));
    }
    return EMS_OK;
}

ems_status ems_fem1_get_time(const ems_fem1 *s, double *t)
{
    clear_error();
    if (!s || !t) return invalid("null argument");
    *t = s->t_now;
    return EMS_OK;
}

ems_status ems_fem1_matrix(const ems_fem1 *s, double *out)
{
    clear_error();
    if (!s || !out) return invalid("null argument");
    std::copy(s->K, s->K + 144, out);
    return EMS_OK;
}

// saveResults: src/fem/fem_simulation.cpp:900-924 (every observation writes
// `path`, the last one wins) and src/fem_abm/fem_simulation.cpp:1005-1030
// (results_<i>_<dt>_<h>_<num_steps>.csv in directory `path`, "" = cwd);
// default stream precision.
ems_status ems_fem1_save_results(ems_fem1 *s, int num_steps, const char *path)
{
    clear_error();
    if (!s || !path) return invalid("null argument");
    const int nobs = (int)s->cnt.size();
    for (int i = 0; i < nobs; ++i) {
        std::string name = path;
        if (abm(s)) {
            name = "results_" + std::to_string(i) + "_" + std::to_string(s->d.dt) + "_" + std::to_string(s->d.h) + "_" +
                   std::to_string(num_steps) + ".csv";
            if (path[0]) name = std::string(path) + "/" + name;
        }
        std::ofstream file(name);
        if (!file.is_open()) {
            set_error("Error: Could not open file " + name);
            return EMS_ERR_IO;
        }
        file << "time,Ex,Ey,Ez" << std::endl;
        for (int j = 0; j < s->nslot; ++j) {
            const double *v = &s->traces[((size_t)i * s->nslot + j) * 3];
            file << (size_t)j * s->d.tf * s->d.dt << "," << v[0] << "," << v[1] << "," << v[2] << std::endl;
        }
    }
    return EMS_OK;
}

ems_status ems_fem1_set_element_product(ems_fem1 *s, int kind)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (kind != EMS_FEM_ELEM_AUTO && kind != EMS_FEM_ELEM_ROWS && kind != EMS_FEM_ELEM_STENCIL)
        return invalid("element product must be AUTO, ROWS or STENCIL");
    s->elem_kind = kind;
    return EMS_OK;
}

ems_status ems_fem1_dof_count(const ems_fem1 *s, size_t *n)
{
    clear_error();
    if (!s || !n) return invalid("null argument");
    *n = (size_t)(s->g.n[0] + s->g.n[1] + s->g.n[2]);
    return EMS_OK;
}

}  // extern "C"
