/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py use it; the product path never does.
 *
 * Serial FP64 restatement of the first-order stabilized hex edge-element
 * solvers of the reference:
 *   - src/fem      (stepper 0): leapfrog E^{n+1} = 2E^n - E^{n-1} - K E^n,
 *                   K = 1/2 dt^2/(mu eps h^2) C_s B_s^-1 A_s (12 x 12),
 *                   fem_simulation.cpp:466-740 (step, sources), :743-833 (PEC),
 *                   :835-898 (run / receivers);
 *   - src/fem_abm  (stepper 1): Adams-Bashforth-2 predictor and
 *                   Adams-Moulton-3 corrector on co-located E and H,
 *                   E side with K^T, H side with K (K = A_s),
 *                   fem_simulation.cpp:278-833 (step), :835-943 (PEC, E only),
 *                   :944-1003 (run).
 * The element loop is written in scatter form over elements in the
 * reference's serial order (k, j, i), exactly like the reference's own loop
 * without atomics, so its sums happen in the serial order.  K is an input:
 * the Python side computes it from the reference's tables.
 *
 * Parity: the reference needs Eigen (absent here), so it is not compiled;
 * this restatement is pinned by invariants in tests/test_oracle_fem1.py.
 */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nx, ny, nz;
    double h, dt, eps, mu;
    int tf;
    int stepper; /* 0: fem (leapfrog), 1: fem_abm */
    double stab;
    int use_ofem;
} ems_oracle_fem1_cfg;

typedef struct {
    ems_oracle_fem1_cfg c;
    double K[144];
    long n[3];
    double *E[4][3], *H[4][3];
    int t0, t1, t2;
    long *src[3];
    int nsrc[3];
    int *dip; /* [ndip][3] */
    int ndip;
    int nobs;
    int *obs_cnt; /* [nobs][3] */
    long **obs_idx; /* [nobs*3] -> flat indices */
    double *tr;     /* [nobs][nslot][3] */
    int nslot;
    double t;
} ems_oracle_fem1;

static long fx(const ems_oracle_fem1 *s, int i, int j, int k)
{
    return i + (long)j * s->c.nx + (long)k * s->c.nx * (s->c.ny + 1);
}
static long fy(const ems_oracle_fem1 *s, int i, int j, int k)
{
    return i + (long)j * (s->c.nx + 1) + (long)k * (s->c.nx + 1) * s->c.ny;
}
static long fz(const ems_oracle_fem1 *s, int i, int j, int k)
{
    return i + (long)j * (s->c.nx + 1) + (long)k * (s->c.nx + 1) * (s->c.ny + 1);
}

/* src/fem/fem_simulation.cpp:83-121: the 4 x-, y- and z-edges of element (i,j,k) */
static void conn(const ems_oracle_fem1 *s, int i, int j, int k, long cx[4], long cy[4], long cz[4])
{
    cx[0] = fx(s, i, j, k);
    cx[1] = fx(s, i, j + 1, k);
    cx[2] = fx(s, i, j, k + 1);
    cx[3] = fx(s, i, j + 1, k + 1);
    cy[0] = fy(s, i, j, k);
    cy[1] = fy(s, i, j, k + 1);
    cy[2] = fy(s, i + 1, j, k);
    cy[3] = fy(s, i + 1, j, k + 1);
    cz[0] = fz(s, i, j, k);
    cz[1] = fz(s, i + 1, j, k);
    cz[2] = fz(s, i, j + 1, k);
    cz[3] = fz(s, i + 1, j + 1, k);
}

/* src/fem/fem_simulation.cpp:29-39 */
static double sf_fem(double t, double mu, double h)
{
    const double f0 = 1.0e9;
    return 4.0 * M_PI * M_PI * f0 * f0 * (1.0 - 4.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) * (t - 1.0 / f0)) *
           exp(-2.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) * (t - 1.0 / f0)) * mu / h / h / h * 0.001;
}

/* src/fem_abm/fem_simulation.cpp:31-39 */
static double sf_abm(double t, double h)
{
    const double f0 = 1.0e9;
    return -4.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) *
           exp(-2.0 * M_PI * M_PI * f0 * f0 * (t - 1.0 / f0) * (t - 1.0 / f0)) / h / h / h * 0.001;
}

double ems_oracle_fem1_source(int stepper, double t, double mu, double h)
{
    return stepper == 0 ? sf_fem(t, mu, h) : sf_abm(t, h);
}

ems_oracle_fem1 *ems_oracle_fem1_create(const ems_oracle_fem1_cfg *c, const double *K)
{
    ems_oracle_fem1 *s = (ems_oracle_fem1 *)calloc(1, sizeof(*s));
    s->c = *c;
    memcpy(s->K, K, sizeof(s->K));
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    s->n[0] = nx * (ny + 1) * (nz + 1);
    s->n[1] = (nx + 1) * ny * (nz + 1);
    s->n[2] = (nx + 1) * (ny + 1) * nz;
    for (int sl = 0; sl < 4; ++sl)
        for (int q = 0; q < 3; ++q) {
            s->E[sl][q] = (double *)calloc(s->n[q], sizeof(double));
            s->H[sl][q] = (double *)calloc(s->n[q], sizeof(double));
        }
    s->t0 = 0;
    s->t1 = 1;
    s->t2 = 2;
    return s;
}

void ems_oracle_fem1_destroy(ems_oracle_fem1 *s)
{
    if (!s) return;
    for (int sl = 0; sl < 4; ++sl)
        for (int q = 0; q < 3; ++q) {
            free(s->E[sl][q]);
            free(s->H[sl][q]);
        }
    for (int q = 0; q < 3; ++q) free(s->src[q]);
    free(s->dip);
    for (int o = 0; o < 3 * s->nobs; ++o) free(s->obs_idx[o]);
    free(s->obs_idx);
    free(s->obs_cnt);
    free(s->tr);
    free(s);
}

/* setSource_x/y/z: flat index in the component's array */
void ems_oracle_fem1_add_source(ems_oracle_fem1 *s, int comp, int i, int j, int k)
{
    long idx = comp == 0 ? fx(s, i, j, k) : comp == 1 ? fy(s, i, j, k) : fz(s, i, j, k);
    s->src[comp] = (long *)realloc(s->src[comp], sizeof(long) * (s->nsrc[comp] + 1));
    s->src[comp][s->nsrc[comp]++] = idx;
}

void ems_oracle_fem1_add_dipole_x(ems_oracle_fem1 *s, int i, int j, int k)
{
    s->dip = (int *)realloc(s->dip, sizeof(int) * 3 * (s->ndip + 1));
    s->dip[3 * s->ndip] = i;
    s->dip[3 * s->ndip + 1] = j;
    s->dip[3 * s->ndip + 2] = k;
    s->ndip++;
}

/* setObservationPoint: counts[3] points per component, pts[(sum counts)][3] */
void ems_oracle_fem1_add_observation(ems_oracle_fem1 *s, const int *counts, const int *pts)
{
    const int o = s->nobs;
    s->obs_cnt = (int *)realloc(s->obs_cnt, sizeof(int) * 3 * (o + 1));
    s->obs_idx = (long **)realloc(s->obs_idx, sizeof(long *) * 3 * (o + 1));
    int p = 0;
    for (int q = 0; q < 3; ++q) {
        s->obs_cnt[3 * o + q] = counts[q];
        s->obs_idx[3 * o + q] = (long *)malloc(sizeof(long) * (counts[q] > 0 ? counts[q] : 1));
        for (int a = 0; a < counts[q]; ++a, ++p) {
            const int *v = pts + 3 * p;
            s->obs_idx[3 * o + q][a] = q == 0 ? fx(s, v[0], v[1], v[2]) : q == 1 ? fy(s, v[0], v[1], v[2])
                                                                                 : fz(s, v[0], v[1], v[2]);
        }
    }
    /* saved_electric_field_.push_back(empty): the new row has no slots until run() resizes */
    double *nt = (double *)calloc((size_t)(o + 1) * s->nslot * 3 + 1, sizeof(double));
    if (s->tr) memcpy(nt, s->tr, sizeof(double) * (size_t)o * s->nslot * 3);
    free(s->tr);
    s->tr = nt;
    s->nobs = o + 1;
}

/* dipole source pattern, src/fem/fem_simulation.cpp:637-736 and
   src/fem_abm/fem_simulation.cpp:429-478: x-edges of the 2x2 element
   block around the dipole, then the stabilization terms on y and z edges */
static void apply_dipoles(ems_oracle_fem1 *s, double **E, double temp, double stab, int ofem)
{
    for (int d = 0; d < s->ndip; ++d) {
        const int i = s->dip[3 * d], j = s->dip[3 * d + 1], k = s->dip[3 * d + 2];
        long c0x[4], c0y[4], c0z[4], c1x[4], c1y[4], c1z[4], c2x[4], c2y[4], c2z[4], c3x[4], c3y[4], c3z[4];
        conn(s, i, j, k, c0x, c0y, c0z);
        conn(s, i, j + 1, k, c1x, c1y, c1z);
        conn(s, i, j, k + 1, c2x, c2y, c2z);
        conn(s, i, j + 1, k + 1, c3x, c3y, c3z);
        if (ofem) {
            E[0][c0x[3]] += 8.0 * temp;
            continue;
        }
        E[0][c0x[0]] += 1.0 / 8.0 * temp;
        E[0][c0x[1]] += 3.0 / 4.0 * temp;
        E[0][c0x[2]] += 3.0 / 4.0 * temp;
        E[0][c0x[3]] += 9.0 / 2.0 * temp;
        E[0][c1x[1]] += 1.0 / 8.0 * temp;
        E[0][c1x[3]] += 3.0 / 4.0 * temp;
        E[0][c2x[2]] += 1.0 / 8.0 * temp;
        E[0][c2x[3]] += 3.0 / 4.0 * temp;
        E[0][c3x[3]] += 1.0 / 8.0 * temp;
        E[1][c0y[0]] += stab / 2.0 * temp;
        E[1][c0y[2]] += stab / 2.0 * temp;
        E[1][c1y[0]] += stab / 2.0 * temp;
        E[1][c1y[2]] += stab / 2.0 * temp;
        E[1][c2y[1]] -= stab / 2.0 * temp;
        E[1][c2y[3]] -= stab / 2.0 * temp;
        E[1][c3y[1]] -= stab / 2.0 * temp;
        E[1][c3y[3]] -= stab / 2.0 * temp;
        E[2][c0z[0]] -= stab / 2.0 * temp;
        E[2][c0z[1]] -= stab / 2.0 * temp;
        E[2][c1z[2]] += stab / 2.0 * temp;
        E[2][c1z[3]] += stab / 2.0 * temp;
        E[2][c2z[0]] -= stab / 2.0 * temp;
        E[2][c2z[1]] -= stab / 2.0 * temp;
        E[2][c3z[2]] += stab / 2.0 * temp;
        E[2][c3z[3]] += stab / 2.0 * temp;
    }
}

/* PEC: tangential E on the outer faces -> 0 (src/fem/fem_simulation.cpp:743-805) */
static void pec(ems_oracle_fem1 *s, double **E)
{
    const int nx = s->c.nx, ny = s->c.ny, nz = s->c.nz;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i)
                if (j == 0 || j == ny || k == 0 || k == nz) E[0][fx(s, i, j, k)] = 0.0;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i)
                if (i == 0 || i == nx || k == 0 || k == nz) E[1][fy(s, i, j, k)] = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i)
                if (i == 0 || i == nx || j == 0 || j == ny) E[2][fz(s, i, j, k)] = 0.0;
}

/* out -= K x_e (sign -1, rows of K: fem) or out += K^T x_e / K x_e
   (sign +1: fem_abm E / H side) over all elements in serial order */
static void ebe(ems_oracle_fem1 *s, double **x, double **out, int transpose, int subtract)
{
    const int nx = s->c.nx, ny = s->c.ny, nz = s->c.nz;
    const double *K = s->K;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long cx[4], cy[4], cz[4];
                conn(s, i, j, k, cx, cy, cz);
                double e[12];
                for (int l = 0; l < 4; ++l) {
                    e[l] = x[0][cx[l]];
                    e[4 + l] = x[1][cy[l]];
                    e[8 + l] = x[2][cz[l]];
                }
                if (subtract) {
                    for (int l = 0; l < 12; ++l) {
                        double temp = 0;
                        for (int m = 0; m < 12; ++m) temp += K[l * 12 + m] * e[m];
                        long n = l < 4 ? cx[l] : l < 8 ? cy[l - 4] : cz[l - 8];
                        out[l / 4][n] -= temp;
                    }
                } else {
                    for (int l = 0; l < 4; ++l) {
                        double tx = 0, ty = 0, tz = 0;
                        for (int m = 0; m < 12; ++m) {
                            if (transpose) {
                                tx += K[m * 12 + l] * e[m];
                                ty += K[m * 12 + l + 4] * e[m];
                                tz += K[m * 12 + l + 8] * e[m];
                            } else {
                                tx += K[l * 12 + m] * e[m];
                                ty += K[(l + 4) * 12 + m] * e[m];
                                tz += K[(l + 8) * 12 + m] * e[m];
                            }
                        }
                        out[0][cx[l]] += tx;
                        out[1][cy[l]] += ty;
                        out[2][cz[l]] += tz;
                    }
                }
            }
}

static void add_points(ems_oracle_fem1 *s, double **E, double v, int subtract)
{
    for (int q = 0; q < 3; ++q)
        for (int a = 0; a < s->nsrc[q]; ++a) {
            if (subtract) E[q][s->src[q][a]] -= v;
            else E[q][s->src[q][a]] += v;
        }
}

static void step_fem(ems_oracle_fem1 *s)
{
    const ems_oracle_fem1_cfg *c = &s->c;
    double **E0 = s->E[s->t0], **E1 = s->E[s->t1], **E2 = s->E[s->t2];
    for (int q = 0; q < 3; ++q) memset(E2[q], 0, sizeof(double) * s->n[q]);
    ebe(s, E1, E2, 0, 1);
    for (int q = 0; q < 3; ++q)
        for (long i = 0; i < s->n[q]; ++i) E2[q][i] += 2.0 * E1[q][i] - E0[q][i];
    const double t = s->t;
    add_points(s, E2, sf_fem(t, c->mu, c->h) * c->dt * c->dt / c->mu / c->eps / c->h / c->h / c->h, 0);
    if (s->ndip) {
        const double temp = sf_fem(t, c->mu, c->h) * c->dt * c->dt / c->mu / c->eps / 8.0;
        apply_dipoles(s, E2, temp, c->stab, c->use_ofem);
    }
    pec(s, E2);
}

static void step_abm(ems_oracle_fem1 *s)
{
    const ems_oracle_fem1_cfg *c = &s->c;
    double **E0 = s->E[s->t0], **E1 = s->E[s->t1], **E2 = s->E[s->t2], **E3 = s->E[3];
    double **H0 = s->H[s->t0], **H1 = s->H[s->t1], **H2 = s->H[s->t2], **H3 = s->H[3];
    const double dt = c->dt, eps = c->eps, mu = c->mu, h = c->h, t = s->t;
    /* predictor, E */
    for (int q = 0; q < 3; ++q)
        for (long i = 0; i < s->n[q]; ++i) {
            E3[q][i] = (3.0 * H1[q][i] - H0[q][i]) * dt / 8.0 / eps / h;
            E2[q][i] = 0.0;
        }
    ebe(s, E3, E2, 1, 0);
    for (int q = 0; q < 3; ++q)
        for (long i = 0; i < s->n[q]; ++i) E2[q][i] += E1[q][i];
    add_points(s, E2, (3.0 * sf_abm(t, h) - sf_abm(t - dt, h)) * dt / 2.0 / eps / h / h / h, 1);
    if (s->ndip) apply_dipoles(s, E2, -(3.0 * sf_abm(t, h) - sf_abm(t - dt, h)) * dt / 16.0 / eps, c->stab, 0);
    /* predictor, H */
    for (int q = 0; q < 3; ++q)
        for (long i = 0; i < s->n[q]; ++i) {
            H3[q][i] = -(3.0 * E1[q][i] - E0[q][i]) * dt / 8.0 / mu / h;
            H2[q][i] = 0.0;
        }
    ebe(s, H3, H2, 0, 0);
    for (int q = 0; q < 3; ++q)
        for (long i = 0; i < s->n[q]; ++i) H2[q][i] += H1[q][i];
    /* corrector */
    for (int q = 0; q < 3; ++q)
        for (long i = 0; i < s->n[q]; ++i) {
            E3[q][i] = (5.0 * H2[q][i] + 8.0 * H1[q][i] - H0[q][i]) * dt / 48.0 / eps / h;
            H3[q][i] = -(5.0 * E2[q][i] + 8.0 * E1[q][i] - E0[q][i]) * dt / 48.0 / mu / h;
            E2[q][i] = 0.0;
            H2[q][i] = 0.0;
        }
    ebe(s, E3, E2, 1, 0);
    ebe(s, H3, H2, 0, 0);
    for (int q = 0; q < 3; ++q)
        for (long i = 0; i < s->n[q]; ++i) {
            E2[q][i] += E1[q][i];
            H2[q][i] += H1[q][i];
        }
    add_points(s, E2, (5.0 * sf_abm(t + dt, h) + 8.0 * sf_abm(t, h) - sf_abm(t - dt, h)) * dt / 12.0 / eps / h / h / h,
               1);
    if (s->ndip)
        apply_dipoles(s, E2,
                      -(5.0 * sf_abm(t + dt, h) + 8.0 * sf_abm(t, h) - sf_abm(t - dt, h)) * dt / 96.0 / eps,
                      c->stab, 0);
    pec(s, E2);
}

/* FemSimulation::run (src/fem/fem_simulation.cpp:835-898) */
void ems_oracle_fem1_run(ems_oracle_fem1 *s, int num_steps)
{
    const int tf = s->c.tf;
    const int nslot = num_steps / tf + 1; /* round(num_steps / tf) + 1, integer division */
    if (nslot != s->nslot) {
        double *nt = (double *)calloc((size_t)s->nobs * nslot * 3 + 1, sizeof(double));
        for (int o = 0; o < s->nobs; ++o)
            for (int q = 0; q < nslot && q < s->nslot; ++q)
                for (int cc = 0; cc < 3; ++cc) nt[((size_t)o * nslot + q) * 3 + cc] = s->tr[((size_t)o * s->nslot + q) * 3 + cc];
        free(s->tr);
        s->tr = nt;
        s->nslot = nslot;
    }
    for (int step = 1; step < num_steps + 1; ++step) {
        if (s->c.stepper == 0) step_fem(s);
        else step_abm(s);
        if (step % tf == 0) {
            double **E2 = s->E[s->t2];
            const int r = step / tf;
            for (int o = 0; o < s->nobs; ++o)
                for (int q = 0; q < 3; ++q) {
                    double *v = &s->tr[((size_t)o * nslot + r) * 3 + q];
                    for (int a = 0; a < s->obs_cnt[3 * o + q]; ++a) *v += E2[q][s->obs_idx[3 * o + q][a]];
                    *v /= (double)(size_t)s->obs_cnt[3 * o + q];
                }
        }
        s->t += s->c.dt;
        s->t0 = (s->t0 + 1) % 3;
        s->t1 = (s->t1 + 1) % 3;
        s->t2 = (s->t2 + 1) % 3;
    }
}

/* one application of the element operator: y = -sum_e K_e x (subtract=1) or
   +sum_e K_e x / K_e^T x; for operator tests */
void ems_oracle_fem1_apply(ems_oracle_fem1 *s, const double *x, double *y, int transpose, int subtract)
{
    double *xx[3], *yy[3];
    long o = 0;
    for (int q = 0; q < 3; ++q) {
        xx[q] = (double *)(x + o);
        yy[q] = y + o;
        memset(yy[q], 0, sizeof(double) * s->n[q]);
        o += s->n[q];
    }
    ebe(s, xx, yy, transpose, subtract);
}

long ems_oracle_fem1_size(const ems_oracle_fem1 *s, int comp) { return s->n[comp]; }
int ems_oracle_fem1_nslot(const ems_oracle_fem1 *s) { return s->nslot; }
int ems_oracle_fem1_nobs(const ems_oracle_fem1 *s) { return s->nobs; }
double ems_oracle_fem1_time(const ems_oracle_fem1 *s) { return s->t; }

void ems_oracle_fem1_get_traces(const ems_oracle_fem1 *s, double *out)
{
    memcpy(out, s->tr, sizeof(double) * (size_t)s->nobs * s->nslot * 3);
}

/* field (0..2 = Ex, Ey, Ez; 3..5 = Hx, Hy, Hz), logical slot 0 = E^{n-1},
   1 = E^n (the latest step), as the reference's time_0_ / time_1_ after a step */
static double *field_ptr(ems_oracle_fem1 *s, int field, int slot)
{
    const int sl = slot == 0 ? s->t0 : s->t1;
    return field < 3 ? s->E[sl][field] : s->H[sl][field - 3];
}

void ems_oracle_fem1_get_field(ems_oracle_fem1 *s, int field, int slot, double *out)
{
    memcpy(out, field_ptr(s, field, slot), sizeof(double) * s->n[field % 3]);
}

void ems_oracle_fem1_set_field(ems_oracle_fem1 *s, int field, int slot, const double *in)
{
    memcpy(field_ptr(s, field, slot), in, sizeof(double) * s->n[field % 3]);
}
