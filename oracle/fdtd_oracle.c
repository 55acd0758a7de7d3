/*
 * FP64 CPU oracle for the explicit Yee FDTD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference FDTD solver (src/fdtd/fdtd_simulation.cpp) used as the checker
 * for the HIP implementation.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it; the product library never links,
 * loads or calls it.
 *
 * Restated reference behaviour (bit-exact operation order):
 *   - field layouts with the 1-cell zero ghost ring
 *       (src/fdtd/fdtd_simulation.cpp:55-88, index algebra :119-220)
 *   - H update, bracket evaluated as ((a - b) - c) + d, coefficient
 *       (dt / mu) / h                                     (:119-169)
 *   - E update over the interior, coefficient (dt / eps) / h  (:170-220)
 *   - source injection E[src] -= (s(t) * dt) / eps with t = current_time_
 *       before the increment                              (:222-239, :304)
 *   - source_function, the Hertz-dipole derivative-of-Gaussian
 *       with f0 = 1 GHz and dl = 1e-3                       (:24-36)
 *   - receiver averaging: saved[slot][c] += p for each point, then /= count,
 *       slot = step / tf when step % tf == 0              (:262-303)
 *   - check_params (1-D CFL test c*dt < h)                (:333-366)
 * The reference writes Hy with a k-stride of Nx*(Ny+1) (:144) but reads it
 * with (Nx+1)*Ny (:182-183, :214-215).  The two agree for Nx == Ny, which is
 * the only case the reference handles correctly; this restatement uses the
 * read stride everywhere.
 *
 * Extensions (absent from the reference; their nearest reference hook is
 * the scalar coefficient of :119-220):
 *   - per-cell eps_r / sigma on the (Nx+1)(Ny+1)(Nz+1) material grid.  The
 *     value at (i,j,k) applies to Ex/Ey/Ez[k][j][i] (the three faces on the
 *     upper side of one cell); E = Ca*E + Cb*curl with
 *     Ca = (1 - s)/(1 + s), Cb = ((dt/eps)/h)/(1 + s), s = sigma*dt/(2 eps),
 *     eps = eps_bg*eps_r.  eps_r = 1, sigma = 0 gives the reference exactly.
 *   - CPML (Roden & Gedney 2000, CFS form) on all six faces, P cells thick,
 *     polynomial grading of order m, sigma_max = -(m+1) ln R0 / (2 eta P h).
 *   - Ricker and Gaussian source wavelets.
 * Pinned by: golden vectors from the compiled reference (vacuum/PEC), and by
 * closed forms (cavity resonances, energy conservation, Fresnel reflection,
 * lossy decay) in tests/test_oracle_fdtd.py.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int nx, ny, nz;
    double h, dt, eps, mu;
    int tf;
    /* wavelet: 0 = reference Hertz dipole, 1 = Ricker, 2 = Gaussian pulse */
    int wavelet;
    double f0, t0, amp;
    /* CPML */
    int npml;
    double pml_order, pml_r0, pml_kappa_max, pml_alpha_max;
} ems_oracle_fdtd_cfg;

typedef struct {
    ems_oracle_fdtd_cfg c;
    size_t nex, ney, nez, nhx, nhy, nhz, nmat;
    double *ex, *ey, *ez, *hx, *hy, *hz;
    double *ca, *cb; /* per material cell, NULL = uniform */
    int has_ca;
    /* sources */
    int nsrc;
    int *src_comp, *src_idx;
    double *src_scale; /* 1 / (1 + s) at the source cell */
    double *src_eps;   /* eps at the source cell */
    /* receivers: nobs lists of 3 point lists */
    int nobs;
    int *obs_count; /* [nobs][3] */
    int *obs_off;   /* [nobs][3] offset into obs_idx */
    int *obs_idx;
    double *traces; /* [nobs][nslot][3] */
    int nslot;
    double t_now;
    /* CPML profiles and psi (full-size unified-grid arrays) */
    double *kinvh[3], *bh[3], *ah[3], *kinve[3], *be[3], *ae[3];
    double *psi[12];
    size_t nuni;
} ems_oracle_fdtd;

/* ---- reference layouts (src/fdtd/fdtd_simulation.cpp:57-68, 119-220) ---- */
#define EX(i, j, k) ((size_t)(i) + (size_t)(j) * (nx + 1) + (size_t)(k) * (nx + 1) * (ny + 2))
#define EY(i, j, k) ((size_t)(i) + (size_t)(j) * (nx + 2) + (size_t)(k) * (nx + 2) * (ny + 1))
#define EZ(i, j, k) ((size_t)(i) + (size_t)(j) * (nx + 2) + (size_t)(k) * (nx + 2) * (ny + 2))
#define HX(i, j, k) ((size_t)(i) + (size_t)(j) * nx + (size_t)(k) * nx * (ny + 1))
#define HY(i, j, k) ((size_t)(i) + (size_t)(j) * (nx + 1) + (size_t)(k) * (nx + 1) * ny)
#define HZ(i, j, k) ((size_t)(i) + (size_t)(j) * (nx + 1) + (size_t)(k) * (nx + 1) * (ny + 1))
/* material grid: (nx+1)(ny+1)(nz+1), value (i,j,k) applies to E*[k][j][i] */
#define MAT(i, j, k) ((size_t)(i) + (size_t)(j) * (nx + 1) + (size_t)(k) * (nx + 1) * (ny + 1))
/* unified grid used for the CPML psi arrays: (nx+2)(ny+2)(nz+2) */
#define UNI(i, j, k) ((size_t)(i) + (size_t)(j) * (nx + 2) + (size_t)(k) * (nx + 2) * (ny + 2))

/* src/fdtd/fdtd_simulation.cpp:24-36, same operation order; f0 is a
 * parameter (the reference fixes it to 1 GHz). */
double ems_oracle_hertz(double t, double domain_size, double f0)
{
    return -4.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) *
           exp(-2.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) * (t - 1.0 / f0)) /
           domain_size / domain_size / domain_size * 0.001;
}

/* Ricker wavelet, the form left as a comment at src/fdtd/fdtd_simulation.cpp:27-32 */
double ems_oracle_ricker(double t, double f0, double t0)
{
    double a = M_PI * M_PI * f0 * f0 * (t - t0) * (t - t0);
    return (1.0 - 2.0 * a) * exp(-a);
}

/* Gaussian pulse exp(-(pi f0 (t - t0))^2) */
double ems_oracle_gauss(double t, double f0, double t0)
{
    double a = M_PI * f0 * (t - t0);
    return exp(-a * a);
}

static double source_value(const ems_oracle_fdtd_cfg *c, double t)
{
    switch (c->wavelet) {
    case 1:
        return c->amp * ems_oracle_ricker(t, c->f0, c->t0);
    case 2:
        return c->amp * ems_oracle_gauss(t, c->f0, c->t0);
    default:
        return ems_oracle_hertz(t, c->h, c->f0);
    }
}

/* src/fdtd/fdtd_simulation.cpp:333-366 */
int ems_oracle_check_params(const double *domain_sizes, const int *grid_num, double duration,
                            double time_step, double domain_size, double c)
{
    if (domain_sizes[0] <= 0 || domain_sizes[1] <= 0 || domain_sizes[2] <= 0) return 0;
    if (grid_num[0] <= 0 || grid_num[1] <= 0 || grid_num[2] <= 0) return 0;
    if (duration <= 0) return 0;
    if (time_step <= 0) return 0;
    if (round(domain_sizes[0] / grid_num[0]) != round(domain_sizes[1] / grid_num[1]) ||
        round(domain_sizes[0] / grid_num[0]) != round(domain_sizes[2] / grid_num[2]))
        return 0;
    if (c * time_step >= domain_size) return 0;
    return 1;
}

/* ------------------------------------------------------------------ CPML */
static void cpml_profile(const ems_oracle_fdtd_cfg *c, int n, double x, double *kinv, double *b,
                         double *a)
{
    const int P = c->npml;
    const double L = P * c->h;
    double d = 0.0;
    if (x < L) d = L - x;
    if (x > n * c->h - L) {
        double d2 = x - (n * c->h - L);
        if (d2 > d) d = d2;
    }
    if (P <= 0 || d <= 0.0) {
        *kinv = 1.0;
        *b = 1.0;
        *a = 0.0;
        return;
    }
    const double m = c->pml_order;
    const double eta = sqrt(c->mu / c->eps);
    const double smax = -(m + 1.0) * log(c->pml_r0) / (2.0 * eta * L);
    const double r = d / L;
    const double g = pow(r, m);
    const double sig = smax * g;
    const double kap = 1.0 + (c->pml_kappa_max - 1.0) * g;
    const double alp = c->pml_alpha_max * (1.0 - r);
    *kinv = 1.0 / kap;
    *b = exp(-(sig / kap + alp) * c->dt / c->eps);
    *a = (sig > 0.0) ? sig / (sig * kap + kap * kap * alp) * (*b - 1.0) : 0.0;
}

/* x-region of the PML: positions I <= P or I >= N - P + 1 (both the integer
 * positions I*h and the half positions (I - 1/2)*h inside the layer). */
static int in_pml(int P, int n, int I) { return P > 0 && (I <= P || I >= n - P + 1); }

static int alloc_cpml(ems_oracle_fdtd *s)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    const int nn[3] = {c->nx, c->ny, c->nz};
    for (int ax = 0; ax < 3; ++ax) {
        int len = nn[ax] + 2;
        s->kinvh[ax] = malloc(sizeof(double) * len);
        s->bh[ax] = malloc(sizeof(double) * len);
        s->ah[ax] = malloc(sizeof(double) * len);
        s->kinve[ax] = malloc(sizeof(double) * len);
        s->be[ax] = malloc(sizeof(double) * len);
        s->ae[ax] = malloc(sizeof(double) * len);
        if (!s->kinvh[ax] || !s->bh[ax] || !s->ah[ax] || !s->kinve[ax] || !s->be[ax] || !s->ae[ax])
            return -1;
        for (int I = 0; I < len; ++I) {
            /* H-type derivative positions sit at I*h, E-type at (I - 1/2)*h */
            cpml_profile(c, nn[ax], I * c->h, &s->kinvh[ax][I], &s->bh[ax][I], &s->ah[ax][I]);
            cpml_profile(c, nn[ax], (I - 0.5) * c->h, &s->kinve[ax][I], &s->be[ax][I],
                         &s->ae[ax][I]);
        }
    }
    for (int p = 0; p < 12; ++p) {
        s->psi[p] = calloc(s->nuni, sizeof(double));
        if (!s->psi[p]) return -1;
    }
    return 0;
}

/* psi slots: H: 0 hx_y, 1 hx_z, 2 hy_z, 3 hy_x, 4 hz_x, 5 hz_y
 *            E: 6 ex_y, 7 ex_z, 8 ey_z, 9 ey_x, 10 ez_x, 11 ez_y */

/* ------------------------------------------------------------- lifecycle */
ems_oracle_fdtd *ems_oracle_fdtd_create(const ems_oracle_fdtd_cfg *cfg)
{
    if (!cfg || cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0 || cfg->tf <= 0) return NULL;
    ems_oracle_fdtd *s = calloc(1, sizeof(*s));
    if (!s) return NULL;
    s->c = *cfg;
    const size_t nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    s->nex = (nx + 1) * (ny + 2) * (nz + 2);
    s->ney = (nx + 2) * (ny + 1) * (nz + 2);
    s->nez = (nx + 2) * (ny + 2) * (nz + 1);
    s->nhx = nx * (ny + 1) * (nz + 1);
    s->nhy = (nx + 1) * ny * (nz + 1);
    s->nhz = (nx + 1) * (ny + 1) * nz;
    s->nmat = (nx + 1) * (ny + 1) * (nz + 1);
    s->nuni = (nx + 2) * (ny + 2) * (nz + 2);
    s->ex = calloc(s->nex, sizeof(double));
    s->ey = calloc(s->ney, sizeof(double));
    s->ez = calloc(s->nez, sizeof(double));
    s->hx = calloc(s->nhx, sizeof(double));
    s->hy = calloc(s->nhy, sizeof(double));
    s->hz = calloc(s->nhz, sizeof(double));
    if (!s->ex || !s->ey || !s->ez || !s->hx || !s->hy || !s->hz) goto fail;
    if (cfg->npml > 0 && alloc_cpml(s) != 0) goto fail;
    s->t_now = 0.0;
    return s;
fail:
    free(s->ex);
    free(s->ey);
    free(s->ez);
    free(s->hx);
    free(s->hy);
    free(s->hz);
    free(s);
    return NULL;
}

void ems_oracle_fdtd_adjoint_release(ems_oracle_fdtd *s);

void ems_oracle_fdtd_destroy(ems_oracle_fdtd *s)
{
    if (!s) return;
    ems_oracle_fdtd_adjoint_release(s);
    free(s->ex);
    free(s->ey);
    free(s->ez);
    free(s->hx);
    free(s->hy);
    free(s->hz);
    free(s->ca);
    free(s->cb);
    free(s->src_comp);
    free(s->src_idx);
    free(s->src_scale);
    free(s->src_eps);
    free(s->obs_count);
    free(s->obs_off);
    free(s->obs_idx);
    free(s->traces);
    for (int ax = 0; ax < 3; ++ax) {
        free(s->kinvh[ax]);
        free(s->bh[ax]);
        free(s->ah[ax]);
        free(s->kinve[ax]);
        free(s->be[ax]);
        free(s->ae[ax]);
    }
    for (int p = 0; p < 12; ++p) free(s->psi[p]);
    free(s);
}

/* per-cell eps_r / sigma (either may be NULL = 1 / 0) */
int ems_oracle_fdtd_set_material(ems_oracle_fdtd *s, const double *eps_r, const double *sigma)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    if (!s->ca) s->ca = malloc(sizeof(double) * s->nmat);
    if (!s->cb) s->cb = malloc(sizeof(double) * s->nmat);
    if (!s->ca || !s->cb) return -1;
    s->has_ca = sigma != NULL;
    for (size_t n = 0; n < s->nmat; ++n) {
        double er = eps_r ? eps_r[n] : 1.0;
        double sg = sigma ? sigma[n] : 0.0;
        double e = c->eps * er;
        double hs = sg * c->dt / (2.0 * e);
        s->ca[n] = (1.0 - hs) / (1.0 + hs);
        s->cb[n] = c->dt / e / c->h / (1.0 + hs);
    }
    return 0;
}

/* setSource_{x,y,z}: src/fdtd/fdtd_simulation.cpp:90-106 */
/* OpenMP threads of the oracle's parallel loops (timing legs only; the
 * results do not depend on it: every loop writes disjoint entries) */
void ems_oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int ems_oracle_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int ems_oracle_fdtd_add_source(ems_oracle_fdtd *s, int comp, int i, int j, int k)
{
    const size_t nx = s->c.nx, ny = s->c.ny;
    size_t idx = comp == 0 ? EX(i, j, k) : comp == 1 ? EY(i, j, k) : EZ(i, j, k);
    s->src_comp = realloc(s->src_comp, sizeof(int) * (s->nsrc + 1));
    s->src_idx = realloc(s->src_idx, sizeof(int) * (s->nsrc + 1));
    s->src_scale = realloc(s->src_scale, sizeof(double) * (s->nsrc + 1));
    s->src_eps = realloc(s->src_eps, sizeof(double) * (s->nsrc + 1));
    s->src_comp[s->nsrc] = comp;
    s->src_idx[s->nsrc] = (int)idx;
    s->src_scale[s->nsrc] = 1.0;
    s->src_eps[s->nsrc] = s->c.eps;
    s->nsrc++;
    return 0;
}

/* material seen by the sources: must be called after set_material */
int ems_oracle_fdtd_bind_source_material(ems_oracle_fdtd *s, const double *eps_r,
                                         const double *sigma)
{
    const size_t nx = s->c.nx, ny = s->c.ny;
    for (int q = 0; q < s->nsrc; ++q) {
        int comp = s->src_comp[q];
        size_t idx = (size_t)s->src_idx[q];
        size_t pitch = comp == 0 ? nx + 1 : nx + 2;
        size_t rows = comp == 1 ? ny + 1 : ny + 2;
        size_t i = idx % pitch, j = (idx / pitch) % rows, k = idx / (pitch * rows);
        size_t m = MAT(i, j, k);
        double er = eps_r ? eps_r[m] : 1.0;
        double sg = sigma ? sigma[m] : 0.0;
        double e = s->c.eps * er;
        s->src_eps[q] = e;
        s->src_scale[q] = 1.0 + sg * s->c.dt / (2.0 * e);
    }
    return 0;
}

/* setObservationPoint: counts[3] points in the x, y and z lists, pts [n][3] */
int ems_oracle_fdtd_add_observation(ems_oracle_fdtd *s, const int *counts, const int *pts)
{
    const size_t nx = s->c.nx, ny = s->c.ny;
    int total_old = 0;
    for (int o = 0; o < s->nobs; ++o)
        total_old += s->obs_count[3 * o] + s->obs_count[3 * o + 1] + s->obs_count[3 * o + 2];
    int add = counts[0] + counts[1] + counts[2];
    s->obs_count = realloc(s->obs_count, sizeof(int) * 3 * (s->nobs + 1));
    s->obs_off = realloc(s->obs_off, sizeof(int) * 3 * (s->nobs + 1));
    s->obs_idx = realloc(s->obs_idx, sizeof(int) * (total_old + add + 1));
    int off = total_old, p = 0;
    for (int c = 0; c < 3; ++c) {
        s->obs_count[3 * s->nobs + c] = counts[c];
        s->obs_off[3 * s->nobs + c] = off;
        for (int q = 0; q < counts[c]; ++q, ++p) {
            int i = pts[3 * p], j = pts[3 * p + 1], k = pts[3 * p + 2];
            size_t idx = c == 0 ? EX(i, j, k) : c == 1 ? EY(i, j, k) : EZ(i, j, k);
            s->obs_idx[off++] = (int)idx;
        }
    }
    s->nobs++;
    return 0;
}

/* --------------------------------------------------------------- stepping */
static void update_h(ems_oracle_fdtd *s)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    const int nx = c->nx, ny = c->ny, nz = c->nz, P = c->npml;
    const double ch = c->dt / c->mu / c->h;
    double *ex = s->ex, *ey = s->ey, *ez = s->ez, *hx = s->hx, *hy = s->hy, *hz = s->hz;
    /* Hx: k in [0,nz], j in [0,ny], i in [0,nx)  (:119-135) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz + 1; ++k)
        for (int j = 0; j < ny + 1; ++j)
            for (int i = 0; i < nx; ++i) {
                double a = ez[EZ(i + 1, j + 1, k)], b = ez[EZ(i + 1, j, k)];
                double cc = ey[EY(i + 1, j, k + 1)], d = ey[EY(i + 1, j, k)];
                /* unified index of Hx[k][j][i] is (i+1, j, k) */
                int py = in_pml(P, ny, j), pz = in_pml(P, nz, k);
                if (py || pz) {
                    size_t u = UNI(i + 1, j, k);
                    double t1 = a - b, t2 = cc - d;
                    if (py) {
                        s->psi[0][u] = s->bh[1][j] * s->psi[0][u] + s->ah[1][j] * t1;
                        t1 = t1 * s->kinvh[1][j] + s->psi[0][u];
                    }
                    if (pz) {
                        s->psi[1][u] = s->bh[2][k] * s->psi[1][u] + s->ah[2][k] * t2;
                        t2 = t2 * s->kinvh[2][k] + s->psi[1][u];
                    }
                    hx[HX(i, j, k)] -= ch * (t1 - t2);
                } else {
                    hx[HX(i, j, k)] -= ch * (a - b - cc + d);
                }
            }
    /* Hy: k in [0,nz], j in [0,ny), i in [0,nx]  (:136-152) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz + 1; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx + 1; ++i) {
                double a = ex[EX(i, j + 1, k + 1)], b = ex[EX(i, j + 1, k)];
                double cc = ez[EZ(i + 1, j + 1, k)], d = ez[EZ(i, j + 1, k)];
                /* unified index (i, j+1, k) */
                int pz = in_pml(P, nz, k), px = in_pml(P, nx, i);
                if (pz || px) {
                    size_t u = UNI(i, j + 1, k);
                    double t1 = a - b, t2 = cc - d;
                    if (pz) {
                        s->psi[2][u] = s->bh[2][k] * s->psi[2][u] + s->ah[2][k] * t1;
                        t1 = t1 * s->kinvh[2][k] + s->psi[2][u];
                    }
                    if (px) {
                        s->psi[3][u] = s->bh[0][i] * s->psi[3][u] + s->ah[0][i] * t2;
                        t2 = t2 * s->kinvh[0][i] + s->psi[3][u];
                    }
                    hy[HY(i, j, k)] -= ch * (t1 - t2);
                } else {
                    hy[HY(i, j, k)] -= ch * (a - b - cc + d);
                }
            }
    /* Hz: k in [0,nz), j in [0,ny], i in [0,nx]  (:153-169) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny + 1; ++j)
            for (int i = 0; i < nx + 1; ++i) {
                double a = ey[EY(i + 1, j, k + 1)], b = ey[EY(i, j, k + 1)];
                double cc = ex[EX(i, j + 1, k + 1)], d = ex[EX(i, j, k + 1)];
                /* unified index (i, j, k+1) */
                int px = in_pml(P, nx, i), py = in_pml(P, ny, j);
                if (px || py) {
                    size_t u = UNI(i, j, k + 1);
                    double t1 = a - b, t2 = cc - d;
                    if (px) {
                        s->psi[4][u] = s->bh[0][i] * s->psi[4][u] + s->ah[0][i] * t1;
                        t1 = t1 * s->kinvh[0][i] + s->psi[4][u];
                    }
                    if (py) {
                        s->psi[5][u] = s->bh[1][j] * s->psi[5][u] + s->ah[1][j] * t2;
                        t2 = t2 * s->kinvh[1][j] + s->psi[5][u];
                    }
                    hz[HZ(i, j, k)] -= ch * (t1 - t2);
                } else {
                    hz[HZ(i, j, k)] -= ch * (a - b - cc + d);
                }
            }
}

static void update_e(ems_oracle_fdtd *s)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    const int nx = c->nx, ny = c->ny, nz = c->nz, P = c->npml;
    const double ce = c->dt / c->eps / c->h;
    const double *ca = s->has_ca ? s->ca : NULL, *cb = s->cb;
    double *ex = s->ex, *ey = s->ey, *ez = s->ez, *hx = s->hx, *hy = s->hy, *hz = s->hz;
    /* Ex: k in [1,nz], j in [1,ny], i in [0,nx]  (:170-186) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < nz + 1; ++k)
        for (int j = 1; j < ny + 1; ++j)
            for (int i = 0; i < nx + 1; ++i) {
                double a = hz[HZ(i, j, k - 1)], b = hz[HZ(i, j - 1, k - 1)];
                double cc = hy[HY(i, j - 1, k)], d = hy[HY(i, j - 1, k - 1)];
                double br;
                int py = in_pml(P, ny, j), pz = in_pml(P, nz, k);
                if (py || pz) {
                    size_t u = UNI(i, j, k);
                    double t1 = a - b, t2 = cc - d;
                    if (py) {
                        s->psi[6][u] = s->be[1][j] * s->psi[6][u] + s->ae[1][j] * t1;
                        t1 = t1 * s->kinve[1][j] + s->psi[6][u];
                    }
                    if (pz) {
                        s->psi[7][u] = s->be[2][k] * s->psi[7][u] + s->ae[2][k] * t2;
                        t2 = t2 * s->kinve[2][k] + s->psi[7][u];
                    }
                    br = t1 - t2;
                } else {
                    br = a - b - cc + d;
                }
                size_t e = EX(i, j, k);
                if (cb) {
                    size_t m = MAT(i, j, k);
                    ex[e] = (ca ? ca[m] * ex[e] : ex[e]) + cb[m] * br;
                } else {
                    ex[e] += ce * br;
                }
            }
    /* Ey: k in [1,nz], j in [0,ny], i in [1,nx]  (:187-203) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < nz + 1; ++k)
        for (int j = 0; j < ny + 1; ++j)
            for (int i = 1; i < nx + 1; ++i) {
                double a = hx[HX(i - 1, j, k)], b = hx[HX(i - 1, j, k - 1)];
                double cc = hz[HZ(i, j, k - 1)], d = hz[HZ(i - 1, j, k - 1)];
                double br;
                int pz = in_pml(P, nz, k), px = in_pml(P, nx, i);
                if (pz || px) {
                    size_t u = UNI(i, j, k);
                    double t1 = a - b, t2 = cc - d;
                    if (pz) {
                        s->psi[8][u] = s->be[2][k] * s->psi[8][u] + s->ae[2][k] * t1;
                        t1 = t1 * s->kinve[2][k] + s->psi[8][u];
                    }
                    if (px) {
                        s->psi[9][u] = s->be[0][i] * s->psi[9][u] + s->ae[0][i] * t2;
                        t2 = t2 * s->kinve[0][i] + s->psi[9][u];
                    }
                    br = t1 - t2;
                } else {
                    br = a - b - cc + d;
                }
                size_t e = EY(i, j, k);
                if (cb) {
                    size_t m = MAT(i, j, k);
                    ey[e] = (ca ? ca[m] * ey[e] : ey[e]) + cb[m] * br;
                } else {
                    ey[e] += ce * br;
                }
            }
    /* Ez: k in [0,nz], j in [1,ny], i in [1,nx]  (:204-220) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz + 1; ++k)
        for (int j = 1; j < ny + 1; ++j)
            for (int i = 1; i < nx + 1; ++i) {
                double a = hy[HY(i, j - 1, k)], b = hy[HY(i - 1, j - 1, k)];
                double cc = hx[HX(i - 1, j, k)], d = hx[HX(i - 1, j - 1, k)];
                double br;
                int px = in_pml(P, nx, i), py = in_pml(P, ny, j);
                if (px || py) {
                    size_t u = UNI(i, j, k);
                    double t1 = a - b, t2 = cc - d;
                    if (px) {
                        s->psi[10][u] = s->be[0][i] * s->psi[10][u] + s->ae[0][i] * t1;
                        t1 = t1 * s->kinve[0][i] + s->psi[10][u];
                    }
                    if (py) {
                        s->psi[11][u] = s->be[1][j] * s->psi[11][u] + s->ae[1][j] * t2;
                        t2 = t2 * s->kinve[1][j] + s->psi[11][u];
                    }
                    br = t1 - t2;
                } else {
                    br = a - b - cc + d;
                }
                size_t e = EZ(i, j, k);
                if (cb) {
                    size_t m = MAT(i, j, k);
                    ez[e] = (ca ? ca[m] * ez[e] : ez[e]) + cb[m] * br;
                } else {
                    ez[e] += ce * br;
                }
            }
}

/* run(num_steps): src/fdtd/fdtd_simulation.cpp:242-306.  traces are kept
 * across calls like saved_electric_field_ (resize keeps old values). */
int ems_oracle_fdtd_run(ems_oracle_fdtd *s, int num_steps)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    int nslot = num_steps / c->tf + 1;
    if (nslot > s->nslot) {
        double *t = calloc((size_t)s->nobs * nslot * 3, sizeof(double));
        if (!t) return -1;
        for (int o = 0; o < s->nobs; ++o)
            for (int q = 0; q < s->nslot && q < nslot; ++q)
                for (int cc = 0; cc < 3; ++cc)
                    t[((size_t)o * nslot + q) * 3 + cc] =
                        s->traces[((size_t)o * s->nslot + q) * 3 + cc];
        free(s->traces);
        s->traces = t;
        s->nslot = nslot;
    }
    double *fld[3] = {s->ex, s->ey, s->ez};
    for (int step = 1; step < num_steps + 1; ++step) {
        update_h(s);
        update_e(s);
        for (int q = 0; q < s->nsrc; ++q) {
            double v = source_value(c, s->t_now) * c->dt / s->src_eps[q];
            if (s->src_scale[q] != 1.0) v = v / s->src_scale[q];
            fld[s->src_comp[q]][s->src_idx[q]] -= v;
        }
        if (step % c->tf == 0) {
            int slot = step / c->tf;
            for (int o = 0; o < s->nobs; ++o)
                for (int cc = 0; cc < 3; ++cc) {
                    double *dst = &s->traces[((size_t)o * s->nslot + slot) * 3 + cc];
                    int cnt = s->obs_count[3 * o + cc], off = s->obs_off[3 * o + cc];
                    for (int q = 0; q < cnt; ++q) *dst += fld[cc][s->obs_idx[off + q]];
                    *dst /= (double)cnt;
                }
        }
        s->t_now += c->dt;
    }
    return 0;
}

/* ----------------------------------------------------------- accessors */
int ems_oracle_fdtd_nslot(const ems_oracle_fdtd *s) { return s->nslot; }
int ems_oracle_fdtd_nobs(const ems_oracle_fdtd *s) { return s->nobs; }
double ems_oracle_fdtd_time(const ems_oracle_fdtd *s) { return s->t_now; }

size_t ems_oracle_fdtd_field_size(const ems_oracle_fdtd *s, int which)
{
    const size_t n[6] = {s->nex, s->ney, s->nez, s->nhx, s->nhy, s->nhz};
    return (which >= 0 && which < 6) ? n[which] : 0;
}

static double *field_ptr(ems_oracle_fdtd *s, int which)
{
    double *p[6] = {s->ex, s->ey, s->ez, s->hx, s->hy, s->hz};
    return (which >= 0 && which < 6) ? p[which] : NULL;
}

int ems_oracle_fdtd_get_field(ems_oracle_fdtd *s, int which, double *out)
{
    double *p = field_ptr(s, which);
    if (!p) return -1;
    memcpy(out, p, sizeof(double) * ems_oracle_fdtd_field_size(s, which));
    return 0;
}

int ems_oracle_fdtd_set_field(ems_oracle_fdtd *s, int which, const double *in)
{
    double *p = field_ptr(s, which);
    if (!p) return -1;
    memcpy(p, in, sizeof(double) * ems_oracle_fdtd_field_size(s, which));
    return 0;
}

int ems_oracle_fdtd_get_traces(const ems_oracle_fdtd *s, double *out)
{
    memcpy(out, s->traces, sizeof(double) * (size_t)s->nobs * s->nslot * 3);
    return 0;
}

/* one bare step (no source, no recording) for operator tests */
int ems_oracle_fdtd_bare_step(ems_oracle_fdtd *s)
{
    update_h(s);
    update_e(s);
    return 0;
}

/* ======================================================================
 * Discrete adjoint (no reference counterpart).
 *
 * The forward step maps the state (E, H, psi_H, psi_E) linearly; its exact
 * transpose, applied backwards in time, gives the gradient of the receiver
 * misfit J = 1/2 sum (trace - d_obs)^2 with respect to the per-cell eps_r and
 * sigma.  Everything is written over the unified grid (I, J, K), in which
 * E*(I,J,K) = E*[K][J][I] and Hx/Hy/Hz(I,J,K) = H*[K][J][I-1] / [K][J-1][I] /
 * [K-1][J][I]; the forward E update then reads backward differences of H and
 * the H update forward differences of E.  Adjoint step (operation order is
 * part of the definition; the GPU follows it):
 *  pass A (transpose of the E update), for every updated E cell:
 *      mu = Cb e'                 u1 = mu, u2 = -mu
 *      per derivative: in the CPML along its axis
 *          p = psiE' + u,  tbar = u * kinv + a * p,  psiE = b * p
 *      else tbar = u
 *      e = Ca e'
 *    and for every H cell, e.g. hbar_z += (tb1_Ex(I,J,K) - tb1_Ex(I,J+1,K))
 *                                      + (tb2_Ey(I,J,K) - tb2_Ey(I+1,J,K))
 *  pass B (transpose of the H update), for every H cell:
 *      nu = -ch hbar,  v1 = nu, v2 = -nu, the same CPML transpose with the
 *      H-type profiles and psiH, and for every updated E cell, e.g.
 *      ebar_z += (sb1_Hx(I,J-1,K) - sb1_Hx(I,J,K)) + (sb2_Hy(I-1,J,K) - sb2_Hy(I,J,K))
 *  Terms of cells outside the update ranges are 0.
 * ====================================================================== */

typedef struct {
    double *e[3], *h[3], *psi[12];
} adj_state;

static adj_state *adj_of(ems_oracle_fdtd *s);

/* per-oracle adjoint storage, kept in a side table keyed by the handle */
#define MAX_ADJ 64
static struct {
    ems_oracle_fdtd *owner;
    adj_state st;
} g_adj[MAX_ADJ];

static adj_state *adj_of(ems_oracle_fdtd *s)
{
    int free_slot = -1;
    for (int q = 0; q < MAX_ADJ; ++q) {
        if (g_adj[q].owner == s) return &g_adj[q].st;
        if (!g_adj[q].owner && free_slot < 0) free_slot = q;
    }
    if (free_slot < 0) return NULL;
    adj_state *a = &g_adj[free_slot].st;
    memset(a, 0, sizeof(*a));
    const size_t n[6] = {s->nex, s->ney, s->nez, s->nhx, s->nhy, s->nhz};
    for (int c = 0; c < 3; ++c) {
        a->e[c] = calloc(n[c], sizeof(double));
        a->h[c] = calloc(n[3 + c], sizeof(double));
    }
    for (int p = 0; p < 12; ++p) a->psi[p] = calloc(s->nuni, sizeof(double));
    g_adj[free_slot].owner = s;
    return a;
}

void ems_oracle_fdtd_adjoint_release(ems_oracle_fdtd *s)
{
    for (int q = 0; q < MAX_ADJ; ++q)
        if (g_adj[q].owner == s) {
            adj_state *a = &g_adj[q].st;
            for (int c = 0; c < 3; ++c) {
                free(a->e[c]);
                free(a->h[c]);
            }
            for (int p = 0; p < 12; ++p) free(a->psi[p]);
            g_adj[q].owner = NULL;
        }
}

/* update ranges over the unified grid */
static int e_in(const ems_oracle_fdtd_cfg *c, int comp, int I, int J, int K)
{
    if (I < 0 || J < 0 || K < 0 || I > c->nx + 1 || J > c->ny + 1 || K > c->nz + 1) return 0;
    if (comp == 0) return I <= c->nx && J >= 1 && J <= c->ny && K >= 1 && K <= c->nz;
    if (comp == 1) return I >= 1 && I <= c->nx && J <= c->ny && K >= 1 && K <= c->nz;
    return I >= 1 && I <= c->nx && J >= 1 && J <= c->ny && K <= c->nz;
}
static int h_in(const ems_oracle_fdtd_cfg *c, int comp, int I, int J, int K)
{
    if (I < 0 || J < 0 || K < 0) return 0;
    if (comp == 0) return I >= 1 && I <= c->nx && J <= c->ny && K <= c->nz;
    if (comp == 1) return I <= c->nx && J >= 1 && J <= c->ny && K <= c->nz;
    return I <= c->nx && J <= c->ny && K >= 1 && K <= c->nz;
}
static size_t e_ref(const ems_oracle_fdtd_cfg *c, int comp, int I, int J, int K)
{
    const size_t nx = c->nx, ny = c->ny;
    return comp == 0 ? EX(I, J, K) : comp == 1 ? EY(I, J, K) : EZ(I, J, K);
}
static size_t h_ref(const ems_oracle_fdtd_cfg *c, int comp, int I, int J, int K)
{
    const size_t nx = c->nx, ny = c->ny;
    return comp == 0 ? HX(I - 1, J, K) : comp == 1 ? HY(I, J - 1, K) : HZ(I, J, K - 1);
}

/* derivative axes: E comp c, derivative d (0: t1, 1: t2) -> axis; psi slot */
static const int E_AXIS[3][2] = {{1, 2}, {2, 0}, {0, 1}};
static const int H_AXIS[3][2] = {{1, 2}, {2, 0}, {0, 1}};

static double cb_at(const ems_oracle_fdtd *s, int I, int J, int K)
{
    const size_t nx = s->c.nx, ny = s->c.ny;
    return s->cb ? s->cb[MAT(I, J, K)] : s->c.dt / s->c.eps / s->c.h;
}
static double ca_at(const ems_oracle_fdtd *s, int I, int J, int K)
{
    const size_t nx = s->c.nx, ny = s->c.ny;
    return (s->cb && s->has_ca) ? s->ca[MAT(I, J, K)] : 1.0;
}

/* tbar of E cell (comp, I, J, K), derivative d, from the adjoint e' and the
 * old psiE; 0 outside the update range */
static double tbar_e(ems_oracle_fdtd *s, adj_state *a, int comp, int d, int I, int J, int K)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    const size_t nx = c->nx, ny = c->ny;
    if (!e_in(c, comp, I, J, K)) return 0.0;
    const double mu = cb_at(s, I, J, K) * a->e[comp][e_ref(c, comp, I, J, K)];
    const double u = d == 0 ? mu : -mu;
    const int ax = E_AXIS[comp][d];
    const int n[3] = {c->nx, c->ny, c->nz};
    const int pos[3] = {I, J, K};
    if (!in_pml(c->npml, n[ax], pos[ax])) return u;
    const double p = a->psi[6 + 2 * comp + d][UNI(I, J, K)] + u;
    return u * s->kinve[ax][pos[ax]] + s->ae[ax][pos[ax]] * p;
}

static double sbar_h(ems_oracle_fdtd *s, adj_state *a, int comp, int d, int I, int J, int K)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    const size_t nx = c->nx, ny = c->ny;
    if (!h_in(c, comp, I, J, K)) return 0.0;
    const double nu = -(c->dt / c->mu / c->h) * a->h[comp][h_ref(c, comp, I, J, K)];
    const double v = d == 0 ? nu : -nu;
    const int ax = H_AXIS[comp][d];
    const int n[3] = {c->nx, c->ny, c->nz};
    const int pos[3] = {I, J, K};
    if (!in_pml(c->npml, n[ax], pos[ax])) return v;
    const double p = a->psi[2 * comp + d][UNI(I, J, K)] + v;
    return v * s->kinvh[ax][pos[ax]] + s->ah[ax][pos[ax]] * p;
}

/* the transpose of one bare step (update_h then update_e) applied to the
 * adjoint state; `gp`, `gq` (may be NULL): accumulate the gradient sums of
 * every material cell from e' and E^n = en[3], E^{n+1} = e1[3] (reference
 * layouts), in component order:
 *     P += sum_c e'_c E^n_c,   Q += sum_c e'_c (E^{n+1}_c - Ca E^n_c) */
static void adjoint_step(ems_oracle_fdtd *s, adj_state *a, double *const *en, double *const *e1, double *gp,
                         double *gq)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t nxs = nx, nys = ny;
    const int n3[3] = {nx, ny, nz};
    /* ---- pass A: tbar with the old psiE, then hbar gathers, psiE, e = Ca e' */
    double *tb[3][2];
    for (int q = 0; q < 3; ++q)
        for (int d = 0; d < 2; ++d) tb[q][d] = calloc(s->nuni, sizeof(double));
    for (int K = 0; K <= nz + 1; ++K)
        for (int J = 0; J <= ny + 1; ++J)
            for (int I = 0; I <= nx + 1; ++I)
                for (int q = 0; q < 3; ++q)
                    for (int d = 0; d < 2; ++d) tb[q][d][UNI(I, J, K)] = tbar_e(s, a, q, d, I, J, K);
    /* gradient sums (use e' = the adjoint of E^{n+1}) */
    if (gp && en && e1) {
        for (int K = 0; K <= nz; ++K)
            for (int J = 0; J <= ny; ++J)
                for (int I = 0; I <= nx; ++I) {
                    double tp = 0.0, tq = 0.0;
                    size_t m = I + J * (nxs + 1) + (size_t)K * (nxs + 1) * (nys + 1);
                    for (int q = 0; q < 3; ++q) {
                        if (!e_in(c, q, I, J, K)) continue;
                        size_t r = e_ref(c, q, I, J, K);
                        double lam = a->e[q][r];
                        double ca = ca_at(s, I, J, K);
                        double de = e1[q][r] - ca * en[q][r];
                        tp = tp + lam * en[q][r];
                        tq = tq + lam * de;
                    }
                    gp[m] += tp;
                    gq[m] += tq;
                }
    }
    for (int K = 0; K <= nz + 1; ++K)
        for (int J = 0; J <= ny + 1; ++J)
            for (int I = 0; I <= nx + 1; ++I) {
                size_t u = UNI(I, J, K);
                /* hbar gathers (forward differences of tbar) */
                if (h_in(c, 0, I, J, K)) {
                    double *hb = &a->h[0][h_ref(c, 0, I, J, K)];
                    double A = tb[1][0][u], B = K <= nz ? tb[1][0][UNI(I, J, K + 1)] : 0.0;
                    double C = tb[2][1][u], D = J <= ny ? tb[2][1][UNI(I, J + 1, K)] : 0.0;
                    *hb = *hb + ((A - B) + (C - D));
                }
                if (h_in(c, 1, I, J, K)) {
                    double *hb = &a->h[1][h_ref(c, 1, I, J, K)];
                    double A = tb[2][0][u], B = I <= nx ? tb[2][0][UNI(I + 1, J, K)] : 0.0;
                    double C = tb[0][1][u], D = K <= nz ? tb[0][1][UNI(I, J, K + 1)] : 0.0;
                    *hb = *hb + ((A - B) + (C - D));
                }
                if (h_in(c, 2, I, J, K)) {
                    double *hb = &a->h[2][h_ref(c, 2, I, J, K)];
                    double A = tb[0][0][u], B = J <= ny ? tb[0][0][UNI(I, J + 1, K)] : 0.0;
                    double C = tb[1][1][u], D = I <= nx ? tb[1][1][UNI(I + 1, J, K)] : 0.0;
                    *hb = *hb + ((A - B) + (C - D));
                }
            }
    for (int K = 0; K <= nz + 1; ++K)
        for (int J = 0; J <= ny + 1; ++J)
            for (int I = 0; I <= nx + 1; ++I)
                for (int q = 0; q < 3; ++q) {
                    if (!e_in(c, q, I, J, K)) continue;
                    size_t r = e_ref(c, q, I, J, K);
                    double mu = cb_at(s, I, J, K) * a->e[q][r];
                    const int pos[3] = {I, J, K};
                    for (int d = 0; d < 2; ++d) {
                        int ax = E_AXIS[q][d];
                        if (!in_pml(c->npml, n3[ax], pos[ax])) continue;
                        double uu = d == 0 ? mu : -mu;
                        double *ps = &a->psi[6 + 2 * q + d][UNI(I, J, K)];
                        *ps = s->be[ax][pos[ax]] * (*ps + uu);
                    }
                    a->e[q][r] = ca_at(s, I, J, K) * a->e[q][r];
                }
    for (int q = 0; q < 3; ++q)
        for (int d = 0; d < 2; ++d) free(tb[q][d]);
    /* ---- pass B: sbar with the old psiH, then ebar gathers, psiH */
    double *sb[3][2];
    for (int q = 0; q < 3; ++q)
        for (int d = 0; d < 2; ++d) sb[q][d] = calloc(s->nuni, sizeof(double));
    for (int K = 0; K <= nz + 1; ++K)
        for (int J = 0; J <= ny + 1; ++J)
            for (int I = 0; I <= nx + 1; ++I)
                for (int q = 0; q < 3; ++q)
                    for (int d = 0; d < 2; ++d) sb[q][d][UNI(I, J, K)] = sbar_h(s, a, q, d, I, J, K);
    for (int K = 0; K <= nz + 1; ++K)
        for (int J = 0; J <= ny + 1; ++J)
            for (int I = 0; I <= nx + 1; ++I) {
                size_t u = UNI(I, J, K);
                if (e_in(c, 0, I, J, K)) {
                    double *eb = &a->e[0][e_ref(c, 0, I, J, K)];
                    double A = sb[1][0][UNI(I, J, K - 1)], B = sb[1][0][u];
                    double C = sb[2][1][UNI(I, J - 1, K)], D = sb[2][1][u];
                    *eb = *eb + ((A - B) + (C - D));
                }
                if (e_in(c, 1, I, J, K)) {
                    double *eb = &a->e[1][e_ref(c, 1, I, J, K)];
                    double A = sb[0][1][UNI(I, J, K - 1)], B = sb[0][1][u];
                    double C = sb[2][0][UNI(I - 1, J, K)], D = sb[2][0][u];
                    *eb = *eb + ((A - B) + (C - D));
                }
                if (e_in(c, 2, I, J, K)) {
                    double *eb = &a->e[2][e_ref(c, 2, I, J, K)];
                    double A = sb[0][0][UNI(I, J - 1, K)], B = sb[0][0][u];
                    double C = sb[1][1][UNI(I - 1, J, K)], D = sb[1][1][u];
                    *eb = *eb + ((A - B) + (C - D));
                }
            }
    for (int K = 0; K <= nz + 1; ++K)
        for (int J = 0; J <= ny + 1; ++J)
            for (int I = 0; I <= nx + 1; ++I)
                for (int q = 0; q < 3; ++q) {
                    if (!h_in(c, q, I, J, K)) continue;
                    double nu = -(c->dt / c->mu / c->h) * a->h[q][h_ref(c, q, I, J, K)];
                    const int pos[3] = {I, J, K};
                    for (int d = 0; d < 2; ++d) {
                        int ax = H_AXIS[q][d];
                        if (!in_pml(c->npml, n3[ax], pos[ax])) continue;
                        double vv = d == 0 ? nu : -nu;
                        double *ps = &a->psi[2 * q + d][UNI(I, J, K)];
                        *ps = s->bh[ax][pos[ax]] * (*ps + vv);
                    }
                }
    for (int q = 0; q < 3; ++q)
        for (int d = 0; d < 2; ++d) free(sb[q][d]);
}

int ems_oracle_fdtd_adjoint_bare_step(ems_oracle_fdtd *s)
{
    adj_state *a = adj_of(s);
    if (!a) return -1;
    adjoint_step(s, a, NULL, NULL, NULL, NULL);
    return 0;
}

/* adjoint arrays: which 0..5 = ebar/hbar (reference layouts), 6..17 = psi
 * adjoint slot which-6 (unified grid) */
int ems_oracle_fdtd_get_adjoint(ems_oracle_fdtd *s, int which, double *out)
{
    adj_state *a = adj_of(s);
    if (!a) return -1;
    if (which < 3) memcpy(out, a->e[which], sizeof(double) * ems_oracle_fdtd_field_size(s, which));
    else if (which < 6) memcpy(out, a->h[which - 3], sizeof(double) * ems_oracle_fdtd_field_size(s, which));
    else memcpy(out, a->psi[which - 6], sizeof(double) * s->nuni);
    return 0;
}

int ems_oracle_fdtd_set_adjoint(ems_oracle_fdtd *s, int which, const double *in)
{
    adj_state *a = adj_of(s);
    if (!a) return -1;
    if (which < 3) memcpy(a->e[which], in, sizeof(double) * ems_oracle_fdtd_field_size(s, which));
    else if (which < 6) memcpy(a->h[which - 3], in, sizeof(double) * ems_oracle_fdtd_field_size(s, which));
    else memcpy(a->psi[which - 6], in, sizeof(double) * s->nuni);
    return 0;
}

/* forward CPML psi (slot 0..11, unified grid) */
int ems_oracle_fdtd_get_psi(ems_oracle_fdtd *s, int slot, double *out)
{
    if (!s->psi[slot]) return -1;
    memcpy(out, s->psi[slot], sizeof(double) * s->nuni);
    return 0;
}
int ems_oracle_fdtd_set_psi(ems_oracle_fdtd *s, int slot, const double *in)
{
    if (!s->psi[slot]) return -1;
    memcpy(s->psi[slot], in, sizeof(double) * s->nuni);
    return 0;
}
size_t ems_oracle_fdtd_uni_size(const ems_oracle_fdtd *s) { return s->nuni; }

/* Material derivative factors per material cell (eps_r, sigma may be NULL):
 *   s = sigma dt / (2 eps),  Ca = (1 - s)/(1 + s),  Cb = (dt/eps/h)/(1 + s)
 *   dCa/deps_r = 2 s / (eps_r (1 + s)^2),   (dCb/deps_r)/Cb = -1 / (eps_r (1 + s))
 *   dCa/dsigma = -dt / (eps (1 + s)^2),     (dCb/dsigma)/Cb = -dt / (2 eps (1 + s))
 * The source amplitude dt/(eps_src (1 + s_src)) = h Cb at the source cell, so
 * with E^{n+1} - Ca E^n = Cb (curl H) - src its derivative folds into the
 * second term: dE^{n+1} = dCa E^n + (dCb/Cb) (E^{n+1} - Ca E^n). */
void ems_oracle_fdtd_material_derivs(const ems_oracle_fdtd_cfg *c, size_t nmat, const double *eps_r,
                                     const double *sigma, double *dca_e, double *dcbn_e, double *dca_s,
                                     double *dcbn_s)
{
    for (size_t m = 0; m < nmat; ++m) {
        double er = eps_r ? eps_r[m] : 1.0, sg = sigma ? sigma[m] : 0.0;
        double e = c->eps * er;
        double hs = sg * c->dt / (2.0 * e);
        double op = 1.0 + hs;
        dca_e[m] = 2.0 * hs / (er * op * op);
        dcbn_e[m] = -1.0 / (er * op);
        dca_s[m] = -c->dt / (e * op * op);
        dcbn_s[m] = -c->dt / (2.0 * e * op);
    }
}

/* Misfit and gradient of one shot from a zero state at t = 0.
 * d_obs: [nobs][nslot][3] with nslot = nsteps/tf + 1 (slot 0 unused);
 * components whose point list is empty are not part of the misfit.
 * Outputs: *J, geps/gsig on the material grid (overwritten), traces. */
int ems_oracle_fdtd_gradient(ems_oracle_fdtd *s, int nsteps, const double *eps_r, const double *sigma,
                             const double *d_obs, double *J, double *geps, double *gsig)
{
    const ems_oracle_fdtd_cfg *c = &s->c;
    adj_state *a = adj_of(s);
    if (!a) return -1;
    /* zero forward and adjoint state */
    double *fl[6] = {s->ex, s->ey, s->ez, s->hx, s->hy, s->hz};
    for (int q = 0; q < 6; ++q) memset(fl[q], 0, sizeof(double) * ems_oracle_fdtd_field_size(s, q));
    for (int p = 0; p < 12; ++p)
        if (s->psi[p]) memset(s->psi[p], 0, sizeof(double) * s->nuni);
    for (int q = 0; q < 3; ++q) {
        memset(a->e[q], 0, sizeof(double) * ems_oracle_fdtd_field_size(s, q));
        memset(a->h[q], 0, sizeof(double) * ems_oracle_fdtd_field_size(s, q + 3));
    }
    for (int p = 0; p < 12; ++p) memset(a->psi[p], 0, sizeof(double) * s->nuni);
    free(s->traces);
    s->traces = NULL;
    s->nslot = 0;
    s->t_now = 0.0;
    /* forward, keeping E at every step */
    const size_t ne[3] = {s->nex, s->ney, s->nez};
    double **hist = malloc(sizeof(double *) * 3 * (size_t)(nsteps + 1));
    for (int n = 0; n <= nsteps; ++n)
        for (int q = 0; q < 3; ++q) hist[3 * n + q] = malloc(sizeof(double) * ne[q]);
    for (int q = 0; q < 3; ++q) memcpy(hist[q], fl[q], sizeof(double) * ne[q]);
    for (int n = 1; n <= nsteps; ++n) {
        ems_oracle_fdtd_run(s, 1);
        /* run(1) records slot 1 for step 1 only when tf == 1; record by hand */
        for (int q = 0; q < 3; ++q) memcpy(hist[3 * n + q], fl[q], sizeof(double) * ne[q]);
    }
    const int nslot = nsteps / c->tf + 1;
    /* traces from the stored history (the reference's averaging) */
    double *tr = calloc((size_t)s->nobs * nslot * 3 + 1, sizeof(double));
    for (int slot = 1; slot < nslot; ++slot) {
        const int n = slot * c->tf;
        for (int o = 0; o < s->nobs; ++o)
            for (int cc = 0; cc < 3; ++cc) {
                int cnt = s->obs_count[3 * o + cc], off = s->obs_off[3 * o + cc];
                double v = 0.0;
                for (int q = 0; q < cnt; ++q) v += hist[3 * n + cc][s->obs_idx[off + q]];
                tr[((size_t)o * nslot + slot) * 3 + cc] = v / (double)cnt;
            }
    }
    free(s->traces);
    s->traces = tr;
    s->nslot = nslot;
    /* misfit */
    double jsum = 0.0;
    for (int o = 0; o < s->nobs; ++o)
        for (int slot = 1; slot < nslot; ++slot)
            for (int cc = 0; cc < 3; ++cc) {
                if (s->obs_count[3 * o + cc] == 0) continue;
                double r = tr[((size_t)o * nslot + slot) * 3 + cc] - d_obs[((size_t)o * nslot + slot) * 3 + cc];
                jsum += 0.5 * r * r;
            }
    *J = jsum;
    /* material derivative factors */
    const size_t nm = s->nmat;
    double *dca_e = malloc(sizeof(double) * nm), *dcbn_e = malloc(sizeof(double) * nm);
    double *dca_s = malloc(sizeof(double) * nm), *dcbn_s = malloc(sizeof(double) * nm);
    ems_oracle_fdtd_material_derivs(c, nm, eps_r, sigma, dca_e, dcbn_e, dca_s, dcbn_s);
    double *gp = calloc(nm, sizeof(double)), *gq = calloc(nm, sizeof(double));
    /* backward */
    for (int n = nsteps - 1; n >= 0; --n) {
        const int step = n + 1;
        if (step % c->tf == 0) {
            const int slot = step / c->tf;
            for (int o = 0; o < s->nobs; ++o)
                for (int cc = 0; cc < 3; ++cc) {
                    int cnt = s->obs_count[3 * o + cc], off = s->obs_off[3 * o + cc];
                    if (cnt == 0) continue;
                    double r = tr[((size_t)o * nslot + slot) * 3 + cc] - d_obs[((size_t)o * nslot + slot) * 3 + cc];
                    double w = r / (double)cnt;
                    for (int q = 0; q < cnt; ++q) a->e[cc][s->obs_idx[off + q]] += w;
                }
        }
        adjoint_step(s, a, &hist[3 * n], &hist[3 * (n + 1)], gp, gq);
    }
    /* dJ/deps_r = dCa/deps_r P + (dCb/deps_r)/Cb Q, likewise for sigma */
    for (size_t m = 0; m < nm; ++m) {
        geps[m] = dca_e[m] * gp[m] + dcbn_e[m] * gq[m];
        gsig[m] = dca_s[m] * gp[m] + dcbn_s[m] * gq[m];
    }
    free(gp);
    free(gq);
    for (int n = 0; n <= nsteps; ++n)
        for (int q = 0; q < 3; ++q) free(hist[3 * n + q]);
    free(hist);
    free(dca_e);
    free(dcbn_e);
    free(dca_s);
    free(dcbn_s);
    return 0;
}
