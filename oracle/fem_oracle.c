/*
 * FP64 CPU oracle for the order-p structured-hex edge FEM family:
 * fem_v2 (3-level leapfrog), symplectic_2 and symplectic_4.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation; only
 * tests/, smoke() and bench.py's cpu_baseline load it.
 *
 * It restates, with one rank owning the whole domain (dims = 1 1 1, so every
 * MPI exchange is guarded off), the serial CPU semantics of:
 *   - connectivity            src/fem_v2/fem_simulation.cpp:318-358
 *   - outer-element list      src/fem_v2/fem_simulation.cpp:359-395
 *   - element operator scale  src/fem_v2/fem_simulation.cpp:458-490 (leapfrog)
 *                             src/symplectic_2/fem_simulation.cpp:458-490
 *   - EbE products            src/fem_v2/fem_kernels.hpp:13-175 (outer elements
 *                             first in list order, then inner elements k,j,i;
 *                             temp accumulated from 0 over m; -= for fem_v2,
 *                             += for the symplectic family)
 *   - leapfrog step           src/fem_v2/fem_simulation.cpp:560-655
 *   - symplectic stage        src/symplectic_2/fem_simulation.cpp:557-672
 *   - symplectic_4 stages     src/symplectic_4/fem_simulation.cpp:1278-1359
 *   - PEC boundary            src/fem_v2/fem_simulation.cpp:1115-1247 and the
 *                             symplectic variant (zeroes U and P)
 *   - receivers / run loop    src/fem_v2/fem_simulation.cpp:1249-1322,
 *                             src/symplectic_2/fem_simulation.cpp:1278-1351
 *   - source functions        src/fem_v2/fem_simulation.cpp:77-87
 * Only z sources are injected (the reference stores x/y sources but never
 * uses them).  The raw element matrix is passed in as parsed from the kmat
 * CSV files.
 * Pinned by: the kmat fixtures (symmetry, nullspace), the discrete energy of
 * the leapfrog and symplectic schemes, PEC-cavity eigenfrequencies and time
 * reversibility (tests/test_oracle_fem.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { FEM_LEAPFROG = 0, FEM_SYMP2 = 1, FEM_SYMP4 = 2 };

typedef struct {
    int order;
    int nx, ny, nz; /* grid_size (fine cells), divisible by order */
    double h, dt, eps, mu;
    int tf;
    int stepper;
} ems_oracle_femv2_cfg;

typedef struct {
    ems_oracle_femv2_cfg c;
    int p, dof, n; /* order, DOFs per component per element, 3*dof */
    int nex, ney, nez;
    size_t NX, NY, NZ; /* entries per slot */
    double *ex, *ey, *ez; /* 3 slots each */
    int *cx, *cy, *cz;    /* connectivity [elements][dof] */
    int *outer;
    int nouter;
    double *K; /* scaled n x n */
    int t0, t1, t2;
    double t_now;
    int nsrc;
    int *src; /* z sources (flat index) */
    int nobs;
    int *obs; /* [nobs][3] */
    double *traces; /* [nobs][nslot][4] */
    int nslot;
} ems_oracle_femv2;

/* src/fem_v2/fem_simulation.cpp:77-87 */
double ems_oracle_fem_source(double t)
{
    const double freq = 1.0e9;
    const double chi = 1.0 / freq;
    double delay = t - chi;
    double zeta = 2.0 * M_PI * M_PI * freq * freq;
    return 2.0 * zeta * (2.0 * zeta * delay * delay - 1.0) * exp(0.0 - zeta * delay * delay) * 0.001;
}

static void build_mesh(ems_oracle_femv2 *s)
{
    const int p = s->p;
    const int Nx = s->c.nx, Ny = s->c.ny;
    const int nex = s->nex, ney = s->ney, nez = s->nez;
    for (int k = 0; k < nez; ++k)
        for (int j = 0; j < ney; ++j)
            for (int i = 0; i < nex; ++i) {
                int idx = (i + j * nex + k * nex * ney) * s->dof;
                int ie = p * (i + j * Nx + k * Nx * Ny);
                for (int n = 0; n < p + 1; ++n)
                    for (int m = 0; m < p + 1; ++m)
                        for (int l = 0; l < p; ++l)
                            s->cx[idx + l + m * p + n * p * (p + 1)] =
                                ie + k * p * Nx + l + m * Nx + n * (Ny + 1) * Nx;
                for (int n = 0; n < p + 1; ++n)
                    for (int m = 0; m < p; ++m)
                        for (int l = 0; l < p + 1; ++l)
                            s->cy[idx + l + m * (p + 1) + n * p * (p + 1)] =
                                ie + j * p + k * p * Ny + l + m * (Nx + 1) + n * Ny * (Nx + 1);
                for (int n = 0; n < p; ++n)
                    for (int m = 0; m < p + 1; ++m)
                        for (int l = 0; l < p + 1; ++l)
                            s->cz[idx + l + m * (p + 1) + n * (p + 1) * (p + 1)] =
                                ie + j * p + k * p * (Nx + Ny + 1) + l + m * (Nx + 1) + n * (Ny + 1) * (Nx + 1);
            }
    /* outer-element list (src/fem_v2/fem_simulation.cpp:359-395) */
    for (int j = 0; j < ney; ++j)
        for (int i = 0; i < nex; ++i) {
            s->outer[i + j * nex] = i + j * nex;
            s->outer[nex * ney + i + j * nex] = nex * ney * (nez - 1) + i + j * nex;
        }
    for (int k = 1; k < nez - 1; ++k)
        for (int j = 0; j < ney; ++j) {
            s->outer[2 * nex * ney + j + (k - 1) * ney] = j * nex + k * nex * ney;
            s->outer[2 * nex * ney + ney * (nez - 2) + j + (k - 1) * ney] =
                (nex - 1) + j * nex + k * nex * ney;
        }
    for (int i = 1; i < nex - 1; ++i)
        for (int k = 1; k < nez - 1; ++k) {
            s->outer[2 * nex * ney + 2 * ney * (nez - 2) + k - 1 + (i - 1) * (nez - 2)] = i + k * nex * ney;
            s->outer[2 * nex * ney + 2 * ney * (nez - 2) + (nez - 2) * (nex - 2) + k - 1 + (i - 1) * (nez - 2)] =
                nex * (ney - 1) + i + k * nex * ney;
        }
}

ems_oracle_femv2 *ems_oracle_femv2_create(const ems_oracle_femv2_cfg *cfg, const double *kraw)
{
    const int p = cfg->order;
    if (p < 1 || p > 3) return NULL;
    if (cfg->nx % p || cfg->ny % p || cfg->nz % p) return NULL;
    if (cfg->nx / p < 2 || cfg->ny / p < 2 || cfg->nz / p < 2) return NULL;
    ems_oracle_femv2 *s = calloc(1, sizeof(*s));
    s->c = *cfg;
    s->p = p;
    s->dof = p * (p + 1) * (p + 1);
    s->n = 3 * s->dof;
    s->nex = cfg->nx / p;
    s->ney = cfg->ny / p;
    s->nez = cfg->nz / p;
    const size_t Nx = cfg->nx, Ny = cfg->ny, Nz = cfg->nz;
    s->NX = Nx * (Ny + 1) * (Nz + 1);
    s->NY = (Nx + 1) * Ny * (Nz + 1);
    s->NZ = (Nx + 1) * (Ny + 1) * Nz;
    s->ex = calloc(3 * s->NX, sizeof(double));
    s->ey = calloc(3 * s->NY, sizeof(double));
    s->ez = calloc(3 * s->NZ, sizeof(double));
    size_t ne = (size_t)s->nex * s->ney * s->nez;
    s->cx = malloc(sizeof(int) * ne * s->dof);
    s->cy = malloc(sizeof(int) * ne * s->dof);
    s->cz = malloc(sizeof(int) * ne * s->dof);
    const int a = s->nex, b = s->ney, cc = s->nez;
    s->nouter = 2 * (a * b + b * cc + cc * a) - 4 * (a + b + cc) + 8;
    s->outer = calloc(s->nouter, sizeof(int));
    s->K = malloc(sizeof(double) * s->n * s->n);
    build_mesh(s);
    /* operator scale: src/fem_v2/fem_simulation.cpp:476 (leapfrog),
     * src/symplectic_2/fem_simulation.cpp:473 (symplectic) */
    double sc;
    const double h = cfg->h, dt = cfg->dt;
    if (cfg->stepper == FEM_LEAPFROG)
        sc = 4.0 * dt * dt / (h * p) / (h * p) / cfg->mu / cfg->eps / (8.0 / p / p / p);
    else
        sc = 4.0 / (h * p) / (h * p) / cfg->mu / cfg->eps;
    for (int q = 0; q < s->n * s->n; ++q) s->K[q] = sc * kraw[q];
    if (cfg->stepper == FEM_LEAPFROG) {
        s->t0 = 0;
        s->t1 = 1;
        s->t2 = 2;
    } else { /* U = slot 0, P = slot 1, tmp = slot 2 */
        s->t0 = 0;
        s->t1 = 1;
        s->t2 = 2;
    }
    return s;
}

void ems_oracle_femv2_destroy(ems_oracle_femv2 *s)
{
    if (!s) return;
    free(s->ex);
    free(s->ey);
    free(s->ez);
    free(s->cx);
    free(s->cy);
    free(s->cz);
    free(s->outer);
    free(s->K);
    free(s->src);
    free(s->obs);
    free(s->traces);
    free(s);
}

const double *ems_oracle_femv2_matrix(const ems_oracle_femv2 *s) { return s->K; }

/* setSource_z (src/fem_v2/fem_simulation.cpp:524-538) */
int ems_oracle_femv2_add_source_z(ems_oracle_femv2 *s, int i, int j, int k)
{
    const int Nx = s->c.nx, Ny = s->c.ny, Nz = s->c.nz;
    if (i < 0 || j < 0 || k < 0 || i > Nx || j > Ny || k > Nz) return 0; /* filtered */
    s->src = realloc(s->src, sizeof(int) * (s->nsrc + 1));
    s->src[s->nsrc++] = i + j * (Nx + 1) + k * (Nx + 1) * (Ny + 1);
    return 1;
}

/* setObservationPoint (src/fem_v2/fem_simulation.cpp:540-558) */
int ems_oracle_femv2_add_observation(ems_oracle_femv2 *s, int i, int j, int k)
{
    const int Nx = s->c.nx, Ny = s->c.ny, Nz = s->c.nz;
    if (i < 0 || j < 0 || k < 0 || i > Nx || j > Ny || k > Nz) return 0;
    s->obs = realloc(s->obs, sizeof(int) * 3 * (s->nobs + 1));
    int *o = s->obs + 3 * s->nobs;
    o[0] = i + j * Nx + k * Nx * (Ny + 1);
    o[1] = i + j * (Nx + 1) + k * (Nx + 1) * Ny;
    o[2] = i + j * (Nx + 1) + k * (Nx + 1) * (Ny + 1);
    s->nobs++;
    return 1;
}

/* one EbE element: gather slot `in`, scatter +/- into slot `out` */
static void ebe_element(ems_oracle_femv2 *s, int e, int in, int out, double sign)
{
    const int D = s->dof, n = s->n;
    double loc[144];
    const int *cx = s->cx + (size_t)e * D, *cy = s->cy + (size_t)e * D, *cz = s->cz + (size_t)e * D;
    double *ex = s->ex, *ey = s->ey, *ez = s->ez;
    const size_t ix = in * s->NX, iy = in * s->NY, iz = in * s->NZ;
    const size_t ox = out * s->NX, oy = out * s->NY, oz = out * s->NZ;
    for (int l = 0; l < D; ++l) {
        loc[l] = ex[ix + cx[l]];
        loc[l + D] = ey[iy + cy[l]];
        loc[l + 2 * D] = ez[iz + cz[l]];
    }
    for (int l = 0; l < n; ++l) {
        double temp = 0;
        for (int m = 0; m < n; ++m) temp += s->K[l * n + m] * loc[m];
        double *dst = l < D ? &ex[ox + cx[l]] : l < 2 * D ? &ey[oy + cy[l - D]] : &ez[oz + cz[l - 2 * D]];
        if (sign < 0) *dst -= temp;
        else *dst += temp;
    }
}

static void ebe_outer(ems_oracle_femv2 *s, int in, int out, double sign)
{
    for (int q = 0; q < s->nouter; ++q) ebe_element(s, s->outer[q], in, out, sign);
}

static void ebe_inner(ems_oracle_femv2 *s, int in, int out, double sign)
{
    const int nex = s->nex, ney = s->ney, nez = s->nez;
    for (int k = 1; k < nez - 1; ++k)
        for (int j = 1; j < ney - 1; ++j)
            for (int i = 1; i < nex - 1; ++i) ebe_element(s, i + j * nex + k * nex * ney, in, out, sign);
}

/* zero tangential E on the six faces of slot `sl` (src/fem_v2/fem_simulation.cpp:1115-1247) */
static void pec(ems_oracle_femv2 *s, int sl)
{
    const int Nx = s->c.nx, Ny = s->c.ny, Nz = s->c.nz;
    double *ex = s->ex + sl * s->NX, *ey = s->ey + sl * s->NY, *ez = s->ez + sl * s->NZ;
    for (int j = 0; j < Ny; ++j)
        for (int k = 0; k < Nz + 1; ++k) {
            ey[j * (Nx + 1) + k * (Nx + 1) * Ny] = 0.0;
            ey[j * (Nx + 1) + k * (Nx + 1) * Ny + Nx] = 0.0;
        }
    for (int k = 0; k < Nz; ++k)
        for (int j = 0; j < Ny + 1; ++j) {
            ez[j * (Nx + 1) + k * (Nx + 1) * (Ny + 1)] = 0.0;
            ez[j * (Nx + 1) + k * (Nx + 1) * (Ny + 1) + Nx] = 0.0;
        }
    for (int k = 0; k < Nz; ++k)
        for (int i = 0; i < Nx + 1; ++i) {
            ez[i + k * (Nx + 1) * (Ny + 1)] = 0.0;
            ez[i + k * (Nx + 1) * (Ny + 1) + (Nx + 1) * Ny] = 0.0;
        }
    for (int i = 0; i < Nx; ++i)
        for (int k = 0; k < Nz + 1; ++k) {
            ex[i + k * Nx * (Ny + 1)] = 0.0;
            ex[i + k * Nx * (Ny + 1) + Nx * Ny] = 0.0;
        }
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny + 1; ++j) {
            ex[i + j * Nx] = 0.0;
            ex[i + j * Nx + Nx * (Ny + 1) * Nz] = 0.0;
        }
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx + 1; ++i) {
            ey[i + j * (Nx + 1)] = 0.0;
            ey[i + j * (Nx + 1) + (Nx + 1) * Ny * Nz] = 0.0;
        }
}

static void zero_slot(ems_oracle_femv2 *s, int sl)
{
    memset(s->ex + sl * s->NX, 0, sizeof(double) * s->NX);
    memset(s->ey + sl * s->NY, 0, sizeof(double) * s->NY);
    memset(s->ez + sl * s->NZ, 0, sizeof(double) * s->NZ);
}

/* src/fem_v2/fem_simulation.cpp:560-655 */
static void leapfrog_step(ems_oracle_femv2 *s)
{
    const int s0 = s->t0, s1 = s->t1, s2 = s->t2;
    zero_slot(s, s2);
    ebe_outer(s, s1, s2, -1.0);
    ebe_inner(s, s1, s2, -1.0);
    double *a[3] = {s->ex, s->ey, s->ez};
    const size_t N[3] = {s->NX, s->NY, s->NZ};
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < N[c]; ++i)
            a[c][s2 * N[c] + i] += 2.0 * a[c][s1 * N[c] + i] - a[c][s0 * N[c] + i];
    const double h = s->c.h;
    for (int q = 0; q < s->nsrc; ++q) {
        double temp = ems_oracle_fem_source(s->t_now) * s->c.dt * s->c.dt / s->c.eps / h / h / h;
        s->ez[s2 * s->NZ + s->src[q]] -= temp;
    }
    pec(s, s2);
}

/* src/symplectic_2/fem_simulation.cpp:557-672 */
static void symplectic_stage(ems_oracle_femv2 *s, double deltat, double offset)
{
    const int p = s->p;
    const double m = 8.0 / p / p / p;
    double *a[3] = {s->ex, s->ey, s->ez};
    const size_t N[3] = {s->NX, s->NY, s->NZ};
    zero_slot(s, 2);
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < N[c]; ++i) a[c][i] += 0.5 * deltat * a[c][N[c] + i] / m;
    ebe_outer(s, 0, 2, +1.0);
    const double hp = s->c.h * p;
    for (int q = 0; q < s->nsrc; ++q) {
        double temp = ems_oracle_fem_source(s->t_now + offset + 0.5 * deltat) * 8.0 / s->c.eps / hp / hp / hp;
        s->ez[2 * s->NZ + s->src[q]] += temp;
    }
    ebe_inner(s, 0, 2, +1.0);
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < N[c]; ++i) {
            a[c][N[c] + i] -= deltat * a[c][2 * N[c] + i];
            a[c][i] += 0.5 * deltat * a[c][N[c] + i] / m;
        }
    pec(s, 0);
    pec(s, 1);
}

static void record(ems_oracle_femv2 *s, int slot_idx, int sl)
{
    const int Nx = s->c.nx, Ny = s->c.ny;
    const double *ex = s->ex + sl * s->NX, *ey = s->ey + sl * s->NY, *ez = s->ez + sl * s->NZ;
    for (int o = 0; o < s->nobs; ++o) {
        const int *ob = s->obs + 3 * o;
        int x0 = ob[0] - 1, x1 = ob[0], x2 = ob[0] + Nx * (Ny + 1) - 1, x3 = ob[0] + Nx * (Ny + 1);
        int y0 = ob[1] - (Nx + 1), y1 = ob[1], y2 = ob[1] + (Nx + 1) * (Ny - 1), y3 = ob[1] + (Nx + 1) * Ny;
        double *t = s->traces + ((size_t)o * s->nslot + slot_idx) * 4;
        t[0] = s->t_now;
        t[1] = (ex[x0] + ex[x1] + ex[x2] + ex[x3]) / 4.0;
        t[2] = (ey[y0] + ey[y1] + ey[y2] + ey[y3]) / 4.0;
        t[3] = ez[ob[2]];
    }
}

/* run (src/fem_v2/fem_simulation.cpp:1249-1322, symplectic :1278-1359) */
/* one symplectic stage with explicit (deltat, offset), for reversibility and
 * order-of-accuracy tests */
int ems_oracle_femv2_stage(ems_oracle_femv2 *s, double deltat, double offset)
{
    symplectic_stage(s, deltat, offset);
    return 0;
}

int ems_oracle_femv2_run(ems_oracle_femv2 *s, int num_steps)
{
    const int tf = s->c.tf;
    int nslot = (num_steps - 1) / tf + 1;
    if (nslot > s->nslot) {
        double *t = calloc((size_t)s->nobs * nslot * 4, sizeof(double));
        for (int o = 0; o < s->nobs; ++o)
            for (int q = 0; q < s->nslot; ++q)
                memcpy(t + ((size_t)o * nslot + q) * 4, s->traces + ((size_t)o * s->nslot + q) * 4,
                       4 * sizeof(double));
        free(s->traces);
        s->traces = t;
        s->nslot = nslot;
    }
    const double x1 = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0));
    const double x2 = 1.0 - 2.0 * x1;
    for (int step = 0; step < num_steps; ++step) {
        if (s->c.stepper == FEM_LEAPFROG) {
            leapfrog_step(s);
            if (step % tf == 0) record(s, step / tf, s->t1);
            s->t_now += s->c.dt;
            s->t0 = (s->t0 + 1) % 3;
            s->t1 = (s->t1 + 1) % 3;
            s->t2 = (s->t2 + 1) % 3;
        } else {
            if (step % tf == 0) record(s, step / tf, 0);
            if (s->c.stepper == FEM_SYMP2) {
                symplectic_stage(s, s->c.dt, 0.0);
            } else {
                symplectic_stage(s, x1 * s->c.dt, 0.0);
                symplectic_stage(s, x2 * s->c.dt, x1 * s->c.dt);
                symplectic_stage(s, x1 * s->c.dt, (x1 + x2) * s->c.dt);
            }
            s->t_now += s->c.dt;
        }
    }
    return 0;
}

int ems_oracle_femv2_nslot(const ems_oracle_femv2 *s) { return s->nslot; }
int ems_oracle_femv2_nobs(const ems_oracle_femv2 *s) { return s->nobs; }
double ems_oracle_femv2_time(const ems_oracle_femv2 *s) { return s->t_now; }

int ems_oracle_femv2_get_traces(const ems_oracle_femv2 *s, double *out)
{
    memcpy(out, s->traces, sizeof(double) * (size_t)s->nobs * s->nslot * 4);
    return 0;
}

size_t ems_oracle_femv2_size(const ems_oracle_femv2 *s, int comp)
{
    return comp == 0 ? s->NX : comp == 1 ? s->NY : s->NZ;
}

/* logical slot: leapfrog 0 = E^{n-1}, 1 = E^n (the current step's input);
 * symplectic 0 = U, 1 = P.  */
static int phys_slot(const ems_oracle_femv2 *s, int logical)
{
    if (s->c.stepper != FEM_LEAPFROG) return logical;
    return logical == 0 ? s->t0 : logical == 1 ? s->t1 : s->t2;
}

int ems_oracle_femv2_get(const ems_oracle_femv2 *s, int comp, int logical, double *out)
{
    int sl = phys_slot(s, logical);
    const double *a = comp == 0 ? s->ex : comp == 1 ? s->ey : s->ez;
    size_t N = ems_oracle_femv2_size(s, comp);
    memcpy(out, a + sl * N, sizeof(double) * N);
    return 0;
}

int ems_oracle_femv2_set(ems_oracle_femv2 *s, int comp, int logical, const double *in)
{
    int sl = phys_slot(s, logical);
    double *a = comp == 0 ? s->ex : comp == 1 ? s->ey : s->ez;
    size_t N = ems_oracle_femv2_size(s, comp);
    memcpy(a + sl * N, in, sizeof(double) * N);
    return 0;
}

/* y = K x over the whole mesh (outer then inner elements, accumulated from 0
 * with +=), for operator tests; x and y are [NX | NY | NZ] vectors. */
int ems_oracle_femv2_apply(ems_oracle_femv2 *s, const double *x, double *y)
{
    double *save_x = malloc(sizeof(double) * s->NX * 3);
    double *save_y = malloc(sizeof(double) * s->NY * 3);
    double *save_z = malloc(sizeof(double) * s->NZ * 3);
    memcpy(save_x, s->ex, sizeof(double) * s->NX * 3);
    memcpy(save_y, s->ey, sizeof(double) * s->NY * 3);
    memcpy(save_z, s->ez, sizeof(double) * s->NZ * 3);
    memcpy(s->ex, x, sizeof(double) * s->NX);
    memcpy(s->ey, x + s->NX, sizeof(double) * s->NY);
    memcpy(s->ez, x + s->NX + s->NY, sizeof(double) * s->NZ);
    zero_slot(s, 2);
    ebe_outer(s, 0, 2, +1.0);
    ebe_inner(s, 0, 2, +1.0);
    memcpy(y, s->ex + 2 * s->NX, sizeof(double) * s->NX);
    memcpy(y + s->NX, s->ey + 2 * s->NY, sizeof(double) * s->NY);
    memcpy(y + s->NX + s->NY, s->ez + 2 * s->NZ, sizeof(double) * s->NZ);
    memcpy(s->ex, save_x, sizeof(double) * s->NX * 3);
    memcpy(s->ey, save_y, sizeof(double) * s->NY * 3);
    memcpy(s->ez, save_z, sizeof(double) * s->NZ * 3);
    free(save_x);
    free(save_y);
    free(save_z);
    return 0;
}
