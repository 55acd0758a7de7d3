/*
 * EMSolvers for MI355X — C ABI of the explicit time-domain Maxwell hot path.
 *
 * One opaque handle per solver instance; the handle owns its device buffers
 * (no process globals, several handles may live in one process).  Every call
 * returns an ems_status; ems_last_error() gives the message of the last
 * failure on the calling thread.  A handle is externally synchronised: do not
 * call into one handle from two threads at once.
 *
 * Field arrays crossing this boundary are FP64 in the reference layouts
 * (x fastest):
 *   Ex[k][j][i]  (nx+1)(ny+2)(nz+2)     Hx[k][j][i]  nx(ny+1)(nz+1)
 *   Ey[k][j][i]  (nx+2)(ny+1)(nz+2)     Hy[k][j][i]  (nx+1)ny(nz+1)
 *   Ez[k][j][i]  (nx+2)(ny+2)(nz+1)     Hz[k][j][i]  (nx+1)(ny+1)nz
 * as allocated by src/fdtd/fdtd_simulation.cpp:55-88.
 */
#ifndef EMS_EMS_H
#define EMS_EMS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int ems_status;
enum {
    EMS_OK = 0,
    EMS_ERR_INVALID_ARG = 1,
    EMS_ERR_CFL = 2,
    EMS_ERR_IO = 3,
    EMS_ERR_HIP = 4,
    EMS_ERR_NOT_SUPPORTED = 5,
    EMS_ERR_NO_MEMORY = 6,
    EMS_ERR_STATE = 7
};

enum { EMS_F64 = 0, EMS_F32 = 1 };

/* field selectors for get/set_field */
enum { EMS_EX = 0, EMS_EY = 1, EMS_EZ = 2, EMS_HX = 3, EMS_HY = 4, EMS_HZ = 5 };

/* source wavelets */
enum {
    EMS_WAVELET_HERTZ = 0,  /* src/fdtd/fdtd_simulation.cpp:24-36 (default)          */
    EMS_WAVELET_RICKER = 1, /* (1 - 2 a) exp(-a), a = (pi f0 (t - t0))^2, times amp  */
    EMS_WAVELET_GAUSS = 2   /* exp(-(pi f0 (t - t0))^2), times amp                   */
};

/* kernel variants for ems_fdtd_set_kernel */
enum { EMS_KERNEL_AUTO = 0, EMS_KERNEL_SPLIT = 1, EMS_KERNEL_FUSED = 2, EMS_KERNEL_ZMARCH = 3, EMS_KERNEL_TB2 = 4 };

/* Last error message of the calling thread ("" if none). */
const char *ems_last_error(void);

/* Library version string. */
const char *ems_version(void);

/* Number of visible HIP devices (0 when none; never fails). */
int ems_device_count(void);

/* ------------------------------------------------------------------ FDTD */

typedef struct ems_fdtd ems_fdtd;

/* Constructor arguments of FemSimulation (src/fdtd/fdtd_simulation.hpp:13-14)
 * plus the arithmetic precision of the device fields. */
typedef struct {
    int nx, ny, nz;   /* grid_size: cells per axis                         */
    double h;         /* domain_size: cell edge [m]                        */
    double dt;        /* time_step [s]                                     */
    double eps;       /* permittivity [F/m] (background)                   */
    double mu;        /* permeability [H/m]                                */
    int tf;           /* time_frequency: record every tf steps (>= 1)      */
    int precision;    /* EMS_F64 or EMS_F32                                */
} ems_fdtd_desc;

/* FemSimulation::FemSimulation (src/fdtd/fdtd_simulation.cpp:38-53).
 * EMS_ERR_INVALID_ARG also when one padded z-plane of a field would reach
 * 4 GiB (the kernels address a plane with 32-bit offsets). */
ems_status ems_fdtd_create(const ems_fdtd_desc *desc, ems_fdtd **out);
void ems_fdtd_destroy(ems_fdtd *s);

/* setSource_x / _y / _z (src/fdtd/fdtd_simulation.cpp:90-106):
 * comp 0/1/2 selects the E component, (i,j,k) indexes its reference array. */
ems_status ems_fdtd_add_source(ems_fdtd *s, int comp, int i, int j, int k);

/* setObservationPoint (src/fdtd/fdtd_simulation.cpp:108-112): counts[c]
 * points for component c, pts = the concatenated [n][3] (i,j,k) triples. */
ems_status ems_fdtd_add_observation(ems_fdtd *s, const int counts[3], const int *pts);

/* Per-cell relative permittivity and conductivity on the (nx+1)(ny+1)(nz+1)
 * material grid (x fastest); value (i,j,k) applies to Ex/Ey/Ez[k][j][i].
 * Either pointer may be NULL (eps_r = 1, sigma = 0).  No reference
 * counterpart: extends the scalar coefficient of fdtd_simulation.cpp:170-220. */
ems_status ems_fdtd_set_material(ems_fdtd *s, const double *eps_r, const double *sigma);

/* CPML on all six faces, ncells thick (0 disables).  No reference counterpart.
 * EMS_ERR_INVALID_ARG when 2*ncells+1 exceeds the smallest grid size, for
 * invalid grading parameters, or when one psi slab would hold 2^31 entries
 * or more (32-bit slab indices). */
ems_status ems_fdtd_set_cpml(ems_fdtd *s, int ncells, double order, double r0,
                             double kappa_max, double alpha_max);

/* Source wavelet (default EMS_WAVELET_HERTZ with f0 = 1 GHz, as the
 * reference).  t0 < 0 selects sqrt(2)/f0. */
ems_status ems_fdtd_set_wavelet(ems_fdtd *s, int kind, double f0, double t0, double amp);

/* EMS_KERNEL_AUTO (default: EMS_KERNEL_TB2 when a padded field array is
 * under 4 GiB, else the split sweeps, or the fused sweep in FP64 without a
 * CPML), EMS_KERNEL_SPLIT (an H sweep and an E sweep
 * in place), EMS_KERNEL_FUSED (one tiled sweep with LDS exchange, ping-pong
 * buffers), EMS_KERNEL_ZMARCH (one sweep of independent wavefronts: DPP
 * lane shifts in x, unrolled rows in y, register z-march; ping-pong
 * buffers) or EMS_KERNEL_TB2 (two time steps per launch: overlapped tiles
 * with a halo of 2 + 1 columns, LDS exchange, a lagged z-march; sources and
 * receivers inside; ping-pong field and psi sets; odd remaining steps take
 * the split sweeps).  Results do not depend on it. */
ems_status ems_fdtd_set_kernel(ems_fdtd *s, int variant);

/* Tiling of EMS_KERNEL_TB2: thread tile of cols x rows (I, J) columns,
 * cols, rows >= 5 and cols * rows <= 1024 (512 in FP64; a larger FP64 tile
 * makes ems_fdtd_run fail with EMS_ERR_INVALID_ARG), its stored core
 * (cols - 3) x (rows - 3), and planes per workgroup (>= 1).  0 for
 * cols = rows or planes selects the default (FP32 64 x 16 and 12 planes,
 * FP64 64 x 8 and 16 planes).  With a CPML the tile applies to the CPML-free
 * interior, and to the z-slabs of the shell if it has at most 512 threads
 * (36 x 14 otherwise); the x- and y-slabs take narrow tiles of their own.
 * Results do not depend on it.  No reference counterpart. */
ems_status ems_fdtd_set_tb2(ems_fdtd *s, int cols, int rows, int planes);

/* Autotuner of EMS_KERNEL_TB2: times candidate tiles (FP32 64 x 16, 64 x 12,
 * 64 x 8 and 68 x 15 threads; FP64 64 x 8, 64 x 6 and 36 x 14; the tile sets
 * how many workgroups share a CU) with 8, 12 and 16 planes per chunk, for
 * about trial_steps steps each, on zeroed scratch fields swapped in for the
 * trials (no sources or receivers; the handle's fields, traces and time are
 * untouched), keeps the fastest as by ems_fdtd_set_tb2 and reports it
 * (outputs may be NULL).  Results do not depend on it.  No reference
 * counterpart. */
ems_status ems_fdtd_autotune_tb2(ems_fdtd *s, int trial_steps, int *cols, int *rows, int *planes);

/* Tiling of EMS_KERNEL_ZMARCH: rows per wavefront strip (2 or 4), planes
 * per chunk (>= 1), wavefronts per workgroup (1 or 4).  Defaults 4, 8, 1.
 * No reference counterpart. */
ems_status ems_fdtd_set_zmarch(ems_fdtd *s, int rows, int planes, int waves);

/* Kernel tiling: fused-kernel workgroup rows (8 or 16; 63 x (rows-1) cells
 * per plane), planes per workgroup or per split-kernel thread (>= 1) and
 * flags (bit 0: prefetch the next plane into registers, bit 1: non-temporal
 * stores, bit 2: split sweeps with a CPML as an interior launch plus a
 * CPML-shell launch per sweep instead of one launch).  Results do not depend
 * on it.  Defaults: rows 8, 4 planes per fused workgroup, 2 planes per split
 * thread (4 in FP64 with a CPML), flags 1; split workgroups have 4 rows with
 * a CPML and 8 without. */
ems_status ems_fdtd_set_tiling(ems_fdtd *s, int rows, int planes, int flags);

/* FemSimulation::run (src/fdtd/fdtd_simulation.cpp:242-306): advances
 * num_steps steps on the device and records the receivers every tf steps
 * into device-resident traces; synchronous. */
ems_status ems_fdtd_run(ems_fdtd *s, int num_steps);

/* Runs num_steps steps and measures them with HIP events on the handle's
 * stream: *total_ms = whole step sequence, *kernel_ms = summed duration of
 * the Yee update kernel(s) alone.  Same results as ems_fdtd_run. */
ems_status ems_fdtd_run_timed(ems_fdtd *s, int num_steps, double *total_ms, double *kernel_ms);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same steps, traces and fields as ems_fdtd_run, bit for bit,
 * but the per-call staging (source and receiver lists, the per-step source
 * values, the raw receiver samples) lives in a static 8 MiB device array per
 * GPU instead of the handle's buffers: no hipMalloc, hipFree or other
 * device-allocating call.  The only host allocation is the growth of the
 * trace array, as in ems_fdtd_run.  EMS_ERR_INVALID_ARG when one call's
 * staging exceeds 8 MiB (e.g. 32 receiver points x 32768 recorded steps);
 * EMS_ERR_STATE until the handle's own device state exists: call
 * ems_fdtd_run(s, 0) once after set_cpml / set_kernel / set_tiling.  All
 * calls on one GPU share its array: calls that overlap on one GPU (two
 * threads, two streams) race; consecutive calls do not (each returns after
 * its kernels finish). */
ems_status ems_fdtd_run_static(ems_fdtd *s, int num_steps);

/* Trace shape: nobs observation sets, nslot = num_steps/tf + 1 slots. */
ems_status ems_fdtd_trace_shape(const ems_fdtd *s, int *nobs, int *nslot);

/* saved_electric_field_ as [nobs][nslot][3] FP64 (count = nobs*nslot*3). */
ems_status ems_fdtd_get_traces(ems_fdtd *s, double *out, size_t count);

/* Number of entries of field `which` in its reference layout. */
ems_status ems_fdtd_field_size(const ems_fdtd *s, int which, size_t *n);

/* Copy a field out of / into the device, reference layout, FP64. */
ems_status ems_fdtd_get_field(ems_fdtd *s, int which, double *out, size_t count);
ems_status ems_fdtd_set_field(ems_fdtd *s, int which, const double *in, size_t count);

/* current_time_ of the reference (sum of dt over the steps taken). */
ems_status ems_fdtd_get_time(const ems_fdtd *s, double *t);

/* FemSimulation::saveResults (src/fdtd/fdtd_simulation.cpp:308-331): writes
 * results_<i>_<dt>_<h>_<num_steps>.csv into dir ("" = cwd). */
ems_status ems_fdtd_save_results(ems_fdtd *s, int num_steps, double c, const char *dir);

/* Device bytes held by the handle. */
ems_status ems_fdtd_device_bytes(const ems_fdtd *s, size_t *bytes);

/* check_params (src/fdtd/fdtd_simulation.cpp:333-366): 1 valid, 0 invalid. */
int ems_fdtd_check_params(const double domain_sizes[3], const int grid_num[3], double duration,
                          double time_step, double domain_size, double c);

/* Misfit gradient (no reference counterpart; extends run() of
 * src/fdtd/fdtd_simulation.cpp:242-306 with its discrete adjoint).
 * Runs num_steps steps from a zero state at t = 0 with the handle's sources,
 * receivers, material and CPML, and returns
 *   J = 1/2 sum over observations o, recorded slots 1..num_steps/tf and
 *       components c with a non-empty point list of (trace - d_obs)^2
 * and dJ/deps_r, dJ/dsigma on the (nx+1)(ny+1)(nz+1) material grid (either
 * output may be NULL), accumulated in FP64.  d_obs is [nobs][num_steps/tf+1][3]
 * (slot 0 unused).  Forward states are recovered from checkpoints by
 * recomputing one segment at a time (see ems_fdtd_set_checkpoint_interval
 * and ems_fdtd_set_gradient_memory).  Afterwards the
 * handle holds the final forward state, time num_steps*dt and the run's
 * traces (not accumulated onto earlier ones). */
ems_status ems_fdtd_gradient(ems_fdtd *s, int num_steps, const double *d_obs, size_t count,
                             double *misfit, double *grad_eps_r, double *grad_sigma);

/* Checkpoint interval of ems_fdtd_gradient and ems_batch_gradient: one
 * level of checkpoints every `interval` steps, E snapshots of interval+1
 * steps per shot, every segment but the last run twice.  0 (the default):
 * chosen per call within the memory cap (ems_fdtd_set_gradient_memory).
 * The interval does not change any result. */
ems_status ems_fdtd_set_checkpoint_interval(ems_fdtd *s, int interval);

/* Cap on the bytes of E snapshots and checkpoints that ems_fdtd_gradient
 * (and ems_batch_gradient, with its first member's cap) may use (0, the
 * default: 3/4 of the device memory that is free or already held by the
 * gradient store of the GPU).  With the interval left at 0 the gradient takes
 * the schedule with the fewest recomputed steps within the cap: one level of
 * checkpoints, or two (every segment rerun once more) when one does not fit.
 * A cap that no schedule fits makes the gradient fail (EMS_ERR_INVALID_ARG).
 * The store is one per GPU, kept from call to call and shared by all
 * handles (gradients on one GPU take turns on it).  Does not change any
 * result. */
ems_status ems_fdtd_set_gradient_memory(ems_fdtd *s, size_t bytes);

/* The schedule ems_fdtd_gradient would take now for num_steps steps:
 * *interval1 = level-1 checkpoint interval (num_steps: one level),
 * *interval = snapshot segment length, *bytes = its snapshot and checkpoint
 * bytes (outputs may be NULL). */
ems_status ems_fdtd_gradient_plan(ems_fdtd *s, int num_steps, int *interval1, int *interval, size_t *bytes);

/* ------------------------------------------------------- batched FDTD runs
 * No reference counterpart (the reference runs one shot per process).  A
 * batch advances many FDTD handles (shots and models of uneven size, one
 * precision) together: one launch per sweep covers every member, on the
 * batch's own stream, except for members whose single runs take the
 * two-step sweep (EMS_KERNEL_TB2 set, or chosen by EMS_KERNEL_AUTO on grids
 * with a large CPML-free interior): these run their own launch sequences,
 * forked from and joined back into the batch stream, concurrently with the
 * batched sweeps.  Each member keeps its own sources, receivers,
 * material, CPML and wavelet, and after ems_batch_run(b, n) it is in exactly
 * the state ems_fdtd_run(member, n) would have left it in (same traces and
 * fields, bit for bit).  Members must outlive the batch. */
typedef struct ems_batch ems_batch;

ems_status ems_batch_create(ems_batch **out);
void ems_batch_destroy(ems_batch *b);
ems_status ems_batch_add(ems_batch *b, ems_fdtd *member);

/* Workgroup rows (2, 4, 8) and planes per thread of the batched sweeps. */
ems_status ems_batch_set_tiling(ems_batch *b, int rows, int planes);

/* Times the candidate tilings (trial_steps each, on zeroed scratch fields:
 * the members' state is untouched) and keeps the fastest; reports it. */
ems_status ems_batch_autotune(ems_batch *b, int trial_steps, int *rows, int *planes);

/* Capture steps_per_graph steps into one HIP graph and replay it (0 = plain
 * launches); the step index is kept in device memory. */
ems_status ems_batch_set_graph(ems_batch *b, int steps_per_graph);

ems_status ems_batch_run(ems_batch *b, int num_steps);

/* As ems_batch_run; *total_ms = HIP-event time of the whole step sequence. */
ems_status ems_batch_run_timed(ems_batch *b, int num_steps, double *total_ms);

/* Misfit gradient of several shots on one model (members must share grid,
 * material, CPML, dt and tf): as ems_fdtd_gradient for every member, with
 * d_obs[q] the observed traces of member q; returns each member's misfit
 * (misfits may be NULL), their sum, and the gradient of the sum with respect
 * to eps_r and sigma.  The members run in groups (ems_batch_set_gradient_group),
 * one after the other; in a group the forward and recompute sweeps of all
 * shots are batched, and each adjoint pass runs over all of them in one
 * launch that adds their FP64 gradient sums once per cell.  The summation
 * order, and so the last bits of the gradient, depend on the group size. */
ems_status ems_batch_gradient(ems_batch *b, int num_steps, const double *const *d_obs, double *misfits,
                              double *misfit_total, double *grad_eps_r, double *grad_sigma);

/* Shots per group of ems_batch_gradient (member order, groups as even as
 * possible).  0, the default: the most shots whose E snapshots of every step
 * fit in half of the device's memory (so that no step is recomputed), at
 * least one; a function of the device and the workload, not of the memory
 * free at the call. */
ems_status ems_batch_set_gradient_group(ems_batch *b, int shots);

/* --------------------------------------------------------- structured-hex FEM
 * The fem_v2 / symplectic_2 / symplectic_4 family: order-p edge elements on a
 * structured hex grid (grid_size = fine sub-edges per axis, multiple of p),
 * lumped scalar mass, PEC outer boundary, one GPU owning the whole domain.
 * Fields are FP64 in the reference layouts (src/fem_v2/fem_simulation.cpp:
 * 169-171): Ex Nx(Ny+1)(Nz+1), Ey (Nx+1)Ny(Nz+1), Ez (Nx+1)(Ny+1)Nz, x fastest. */

enum { EMS_FEM_LEAPFROG = 0, EMS_FEM_SYMPLECTIC2 = 1, EMS_FEM_SYMPLECTIC4 = 2 };

typedef struct ems_fem ems_fem;

/* FemSimulation constructor arguments (src/fem_v2/fem_simulation.hpp:14-24)
 * without the MPI topology (one GPU owns the whole domain). */
typedef struct {
    int order;        /* 1, 2 or 3                                            */
    int nx, ny, nz;   /* grid_size (fine sub-edges per axis)                  */
    double h;         /* domain_size: fine spacing [m]                        */
    double dt;        /* time_step [s]                                        */
    double eps, mu;   /* permittivity, permeability                           */
    int tf;           /* time_frequency                                       */
    int stepper;      /* EMS_FEM_LEAPFROG (fem_v2) / _SYMPLECTIC2 / _SYMPLECTIC4 */
    int precision;    /* EMS_F64 or EMS_F32                                   */
} ems_fem_desc;

/* loadMatrixFromCSV (src/fem_v2/fem_simulation.cpp:1477-1524): quoted
 * rationals "a/b".  out == NULL only reports the row count in *rows. */
ems_status ems_fem_load_kmat(const char *path, double *out, size_t count, int *rows);

/* FemSimulation::FemSimulation + calculateElementStiffnessMatrix
 * (src/fem_v2/fem_simulation.cpp:89-160, :458-490): kmat is the raw n x n
 * element matrix (n = 3 p (p+1)^2), scaled here as the reference scales it. */
ems_status ems_fem_create(const ems_fem_desc *desc, const double *kmat, ems_fem **out);
void ems_fem_destroy(ems_fem *s);

/* setSource_x/_y/_z (src/fem_v2/fem_simulation.cpp:492-538); like the
 * reference only z sources are injected. */
ems_status ems_fem_add_source(ems_fem *s, int comp, int i, int j, int k);

/* setObservationPoint (src/fem_v2/fem_simulation.cpp:540-558). */
ems_status ems_fem_add_observation(ems_fem *s, int i, int j, int k);

/* run (src/fem_v2/fem_simulation.cpp:1249-1322; symplectic_2 :1278-1351;
 * symplectic_4 :1278-1359). */
ems_status ems_fem_run(ems_fem *s, int num_steps);

/* As ems_fem_run, with HIP-event timing of the step kernels (see FDTD). */
ems_status ems_fem_run_timed(ems_fem *s, int num_steps, double *total_ms, double *kernel_ms);

/* Traces [nobs][nslot][4] = (time, Ex, Ey, Ez) as saved_electric_field_. */
ems_status ems_fem_trace_shape(const ems_fem *s, int *nobs, int *nslot);
ems_status ems_fem_get_traces(ems_fem *s, double *out, size_t count);

/* Field access: comp 0/1/2; slot 0/1 = E^{n-1}/E^n (leapfrog) or U/P
 * (symplectic). */
ems_status ems_fem_field_size(const ems_fem *s, int comp, size_t *n);
ems_status ems_fem_get_field(ems_fem *s, int comp, int slot, double *out, size_t count);
ems_status ems_fem_set_field(ems_fem *s, int comp, int slot, const double *in, size_t count);
ems_status ems_fem_get_time(const ems_fem *s, double *t);

/* saveResults (src/fem_v2/fem_simulation.cpp:1324-1349). */
ems_status ems_fem_save_results(ems_fem *s, int num_steps, const char *filename);

/* Total edge DOFs (NX + NY + NZ). */
ems_status ems_fem_dof_count(const ems_fem *s, size_t *n);

/* First-order kernels (no reference counterpart): rows = 4, 8 or 16 selects
 * the z-marching kernel that updates all three components in one pass with a
 * 64 x rows workgroup and `planes` planes per workgroup (16 rows for the FP32
 * merged stencil only, 8 otherwise); rows = 0 one gather kernel per
 * component; rows = -1 (default) the measured choice per stepper, precision
 * and element product.  Results do not depend on it, except that rows = 0
 * always takes the element row dots (see ems_fem_set_element_product). */
ems_status ems_fem_set_tiling(ems_fem *s, int rows, int planes);

/* Element products of orders 2 and 3 (no reference counterpart): ROWS one
 * thread per element forming its n row dots in the reference's order (bit
 * for bit the reference's sums); MFMA the product of the element matrix with
 * 16 elements' local vectors on matrix cores (16 x 16 x 4 steps, fused
 * multiply-adds: the same results to rounding, not bit for bit); AUTO (the
 * default) MFMA at order 3 and ROWS at order 2, as measured fastest (order 3
 * 4.9-5.3x faster in FP64, 2.9x in FP32; order 2 within +-5 %).  Order 1
 * (z-march tiling, rows 4 or 8): STENCIL (order 1 only, EMS_ERR_INVALID_ARG
 * otherwise) sums each DOF's 4 element rows into one 33-point stencil per
 * component (coefficients formed once from the scaled operator: the same
 * results to rounding; 1.9-3.2x faster at 128^3), AUTO is STENCIL, ROWS and
 * MFMA take the 4 element row dots in the reference's order (bit for bit).
 * The per-component kernels (rows 0) always take the row dots. */
enum { EMS_FEM_ELEM_AUTO = 0, EMS_FEM_ELEM_ROWS = 1, EMS_FEM_ELEM_MFMA = 2, EMS_FEM_ELEM_STENCIL = 3 };
ems_status ems_fem_set_element_product(ems_fem *s, int kind);

/* ------------------------------------------ first-order stabilized hex FEM
 * src/fem (leapfrog, element matrix 1/2 dt^2/(mu eps h^2) C_s B_s^-1 A_s with
 * stabilization factor s, Hertz-dipole source) and src/fem_abm
 * (Adams-Bashforth-2 predictor / Adams-Moulton-3 corrector on co-located E
 * and H, element matrix A_s).  Layouts as the hex FEM above with nx, ny, nz
 * cells; fields 0..2 = Ex, Ey, Ez and (fem_abm only) 3..5 = Hx, Hy, Hz. */

enum { EMS_FEM1_LEAPFROG = 0, EMS_FEM1_ABM = 1 };

typedef struct ems_fem1 ems_fem1;

/* Constructor arguments of src/fem/fem_simulation.hpp:13-15 (stab, use_ofem)
 * and src/fem_abm/fem_simulation.hpp:13-14 (stab is the file-scope
 * stabilization_factor = 0.05 of src/fem_abm/fem_simulation.cpp:29). */
typedef struct {
    int nx, ny, nz;   /* grid_size: cells per axis                     */
    double h;         /* domain_size [m]                               */
    double dt;        /* time_step [s]                                 */
    double eps, mu;   /* permittivity, permeability                    */
    int tf;           /* time_frequency                                */
    int stepper;      /* EMS_FEM1_LEAPFROG (src/fem) or EMS_FEM1_ABM   */
    double stab;      /* stabilization factor                          */
    int use_ofem;     /* src/fem: 0 conventional, 1 orthogonal basis   */
    int precision;    /* EMS_F64 or EMS_F32                            */
} ems_fem1_desc;

/* calculateElementStiffnessMatrix (src/fem/fem_simulation.cpp:125-434,
 * src/fem_abm/fem_simulation.cpp:143-246): the 12 x 12 row-major element
 * matrix the step uses.  Host only (no device needed). */
ems_status ems_fem1_element_matrix(const ems_fem1_desc *desc, double *out144);

/* FemSimulation::FemSimulation (src/fem/fem_simulation.cpp:41-62,
 * src/fem_abm/fem_simulation.cpp:41-59). */
ems_status ems_fem1_create(const ems_fem1_desc *desc, ems_fem1 **out);
void ems_fem1_destroy(ems_fem1 *s);

/* setSource_x/_y/_z (src/fem/fem_simulation.cpp:436-452). */
ems_status ems_fem1_add_source(ems_fem1 *s, int comp, int i, int j, int k);

/* setSourceDipole_x (src/fem/fem_simulation.cpp:454-457). */
ems_status ems_fem1_add_dipole_x(ems_fem1 *s, int i, int j, int k);

/* setObservationPoint (src/fem/fem_simulation.cpp:459-464): counts[c]
 * points of component c, pts the concatenated [n][3] triples. */
ems_status ems_fem1_add_observation(ems_fem1 *s, const int counts[3], const int *pts);

/* run (src/fem/fem_simulation.cpp:835-898, src/fem_abm/fem_simulation.cpp:944-1003). */
ems_status ems_fem1_run(ems_fem1 *s, int num_steps);
ems_status ems_fem1_run_timed(ems_fem1 *s, int num_steps, double *total_ms, double *kernel_ms);

/* saved_electric_field_ as [nobs][nslot][3], nslot = num_steps/tf + 1. */
ems_status ems_fem1_trace_shape(const ems_fem1 *s, int *nobs, int *nslot);
ems_status ems_fem1_get_traces(ems_fem1 *s, double *out, size_t count);

/* Field access, slot 0/1 = time_0_/time_1_ (the previous / latest step). */
ems_status ems_fem1_field_size(const ems_fem1 *s, int field, size_t *n);
ems_status ems_fem1_get_field(ems_fem1 *s, int field, int slot, double *out, size_t count);
ems_status ems_fem1_set_field(ems_fem1 *s, int field, int slot, const double *in, size_t count);
ems_status ems_fem1_get_time(const ems_fem1 *s, double *t);

/* The element matrix of this handle (row-major 144 doubles). */
ems_status ems_fem1_matrix(const ems_fem1 *s, double *out144);

/* saveResults: src/fem writes every observation to `path` (the file name;
 * the last one wins, src/fem/fem_simulation.cpp:900-924); src/fem_abm writes
 * results_<i>_<dt>_<h>_<num_steps>.csv into directory `path` ("" = cwd,
 * src/fem_abm/fem_simulation.cpp:1005-1030). */
ems_status ems_fem1_save_results(ems_fem1 *s, int num_steps, const char *path);

/* Total edge DOFs per field set. */
ems_status ems_fem1_dof_count(const ems_fem1 *s, size_t *n);

/* Element operator application (no reference counterpart): STENCIL (and AUTO,
 * the default) sums each edge's 4 element rows into one 33-point stencil per
 * component, formed once from the FP64 element matrix (and its transpose for
 * the fem_abm E update), on edges whose transverse coordinates are inside
 * the grid; edges on the faces keep the element row dots.  The same results
 * as the reference to rounding.  ROWS: the element row dots in the
 * reference's order everywhere (bit for bit).  EMS_FEM_ELEM_MFMA is
 * EMS_ERR_INVALID_ARG. */
ems_status ems_fem1_set_element_product(ems_fem1 *s, int kind);

/* --------------------------------- tetrahedral Nedelec (Whitney) elements
 * No reference counterpart: the unstructured-mesh extension of the edge-FEM
 * time steppers (src/symplectic_2, src/symplectic_4, src/fem_abm) on the
 * Kuhn 6-tetrahedron split of an nx x ny x nz cube grid of spacing h, with
 * the consistent mass matrix solved by conjugate gradients on the device and
 * PEC outer faces.  Edge numbering: 7 families (x, y, z, x+y, x+z, y+z,
 * x+y+z directions), family f edges at base vertex (I, J, K) numbered
 * off_f + I + cx (J + cy K), cx = nx + 1 - dx_f, cy = ny + 1 - dy_f.
 * Faces (order 2): a face with vertices p < q < r is numbered by the edge
 * families of (q - p, r - p), 12 families in the order (x, x+y), (x, x+z),
 * (x, x+y+z), (y, x+y), (y, y+z), (y, x+y+z), (z, x+z), (z, y+z),
 * (z, x+y+z), (x+y, x+y+z), (x+z, x+y+z), (y+z, x+y+z), then base vertex
 * p as for edges with the extent of r - p.  Dofs: order 1 = edges; order 2
 * = Whitney dofs (edge e), edge-gradient dofs (ne + e), face dofs
 * (2 ne + 2 face + kappa). */

enum { EMS_TET_SYMPLECTIC2 = 0, EMS_TET_SYMPLECTIC4 = 1, EMS_TET_ABM = 2 };

/* Operator of the tetrahedral solver: assembled CSR matrices, or
 * element-by-element products with the 6 shared element-type matrices (no
 * assembled matrix; per-tetrahedron products, then a fixed-order gather per
 * dof).  EBE_MFMA forms the mass operator's element products with matrix
 * cores (16x16x4 MFMA over 16 tetrahedra of one type).  EBE_ROWS applies the
 * mass operator in one pass per row (each row forms the local products of
 * its incident tetrahedra; no per-tetrahedron buffer), with the same bits
 * as EBE.  AUTO:
 * element-by-element (measured faster than CSR at both orders). */
enum { EMS_TET_OP_AUTO = 0, EMS_TET_OP_CSR = 1, EMS_TET_OP_EBE = 2, EMS_TET_OP_EBE_MFMA = 3, EMS_TET_OP_EBE_ROWS = 4 };

typedef struct ems_tet ems_tet;

typedef struct {
    int nx, ny, nz;   /* cubes per axis (6 tetrahedra each)             */
    double h;         /* cube edge length [m]                           */
    double dt;        /* time step [s]                                  */
    double eps, mu;   /* background permittivity, permeability          */
    int tf;           /* receiver sampling: every tf steps              */
    int stepper;      /* EMS_TET_SYMPLECTIC2 / _SYMPLECTIC4 / _ABM      */
    int order;        /* element order: 1 (Whitney) or 2                */
    int precision;    /* EMS_F64 or EMS_F32                             */
} ems_tet_desc;

/* Element mass M = int w_l . w_m and curl-curl S = int curl w_l . curl w_m
 * (nl x nl row-major, nl = 6 for order 1, 20 for order 2) of Kuhn
 * tetrahedron `type` (0..5: the axis permutation of its vertex path) of a
 * cube of edge h.  Local dofs: order 1: the Whitney functions of local edges
 * (01, 02, 03, 12, 13, 23); order 2: those 6, the edge gradients
 * grad(l_a l_b) of the same edges, then per local face (012, 013, 023, 123)
 * l_a w_bc and l_b w_ac.  Host only (no device needed). */
ems_status ems_tet_element_matrices(int order, int type, double h, double *M, double *S);

/* The assembled CSR matrices of a mesh (host only): rp (dofs + 1), col, S,
 * M_eps, M_sig (nnz each), PEC rows as identity rows of the mass-type
 * matrices.  Call with capacity 0 to get *nnz, then with arrays of that size. */
ems_status ems_tet_assemble(const ems_tet_desc *desc, const double *eps_r, const double *sigma, int *rp, int *col,
                            double *S, double *Me, double *Ms, size_t capacity, size_t *nnz);

ems_status ems_tet_create(const ems_tet_desc *desc, ems_tet **out);
void ems_tet_destroy(ems_tet *s);

/* Relative permittivity and conductivity per tetrahedron (6 nx ny nz
 * values, tetrahedron 6 (i + nx (j + ny k)) + t); NULL = 1 / 0.  At rest
 * only (before the first step, or after ems_tet_gradient); the next run
 * reassembles and keeps a field set with ems_tet_set_field. */
ems_status ems_tet_set_material(ems_tet *s, const double *eps_r, const double *sigma);

/* Electric dipole of moment p(t) d at pos, p(t) = amp exp(-(pi f0 (t - t0))^2):
 * the load p'(t) (w_i(pos) . d) on the edges of the enclosing tetrahedron. */
ems_status ems_tet_add_dipole(ems_tet *s, const double pos[3], const double dir[3], double f0, double t0,
                              double amp);

/* Receiver of the three components of E at pos. */
ems_status ems_tet_add_receiver(ems_tet *s, const double pos[3]);

/* Change the time step before the first step (rebuilds the stage solve
 * matrices on the device; no reassembly). */
ems_status ems_tet_set_time_step(ems_tet *s, double dt);

/* Operator choice (EMS_TET_OP_*), at rest only; the next run rebuilds (a
 * field set with ems_tet_set_field is kept). */
ems_status ems_tet_set_operator(ems_tet *s, int op);

/* Advances num_steps steps.  A run from rest (no step taken yet, or after
 * ems_tet_gradient) sizes the traces to exactly num_steps/tf + 1 slots; later
 * runs append.  If a run fails after device work began (a mass solve that
 * does not converge returns EMS_ERR_STATE, or a HIP error), the field state
 * is undefined: the step count, time and traces keep the values of the last
 * completed call, and run returns EMS_ERR_STATE until both E and dE/dt are
 * set again with ems_tet_set_field (or a gradient succeeds, which leaves the
 * handle at rest). */
ems_status ems_tet_run(ems_tet *s, int num_steps);
ems_status ems_tet_run_timed(ems_tet *s, int num_steps, double *total_ms);

/* Misfit gradient (symplectic steppers): runs num_steps steps from rest at
 * t = 0 with the handle's dipoles and receivers and returns J = 1/2 sum over
 * slots 1..num_steps/tf of (trace - d_obs)^2 (d_obs [nobs][num_steps/tf +
 * 1][3], slot 0 unused) and dJ/deps_r, dJ/dsigma per tetrahedron (either may
 * be NULL), FP64.  Lossless models: the forward states are recovered by
 * running the integrator backwards (no checkpoints); *drift (may be NULL) =
 * the largest entry of the recovered initial state relative to the final
 * state's.  Lossy models (any sigma > 0): checkpoints every
 * ceil(sqrt(num_steps)) steps, each segment rerun once with its stages' states
 * kept (*drift = 0).  Afterwards the handle is at rest at t = 0 and holds the
 * run's traces.  No reference counterpart (extends run() of
 * src/symplectic_4/fem_simulation.cpp:1278-1359). */
ems_status ems_tet_gradient(ems_tet *s, int num_steps, const double *d_obs, size_t count, double *misfit,
                            double *grad_eps_r, double *grad_sigma, double *drift);

/* Receiver traces [nobs][nslot][3], slot q = step q tf (slot 0 = t 0). */
ems_status ems_tet_trace_shape(const ems_tet *s, int *nobs, int *nslot);
ems_status ems_tet_get_traces(ems_tet *s, double *out, size_t count);

ems_status ems_tet_dof_count(const ems_tet *s, size_t *n);

/* Edge coefficients of E (which = 0) or dE/dt (which = 1), dof_count values. */
ems_status ems_tet_get_field(ems_tet *s, int which, double *out, size_t count);
ems_status ems_tet_set_field(ems_tet *s, int which, const double *in, size_t count);
ems_status ems_tet_get_time(const ems_tet *s, double *t);

/* Squared angular frequency w_max^2 = lambda_max(M_eps^-1 S) estimated by
 * `iters` power iterations (the Rayleigh quotient of the last iterate, a lower
 * bound).  Stable steps: dt < kappa / w_max with kappa = 2 (symplectic_2),
 * 1.5735 (symplectic_4), 1.2005 (ABM), the limits of each stepper on
 * y'' = -w^2 y. */
ems_status ems_tet_max_frequency(ems_tet *s, int iters, double *omega2);

/* Mass solver: EMS_TET_SOLVER_AUTO (default) or EMS_TET_SOLVER_CG
 * (Jacobi-preconditioned conjugate gradients).  AUTO solves FP64 order-1
 * systems by Jacobi with Chebyshev acceleration (spectral bounds of each
 * solve matrix by Lanczos, a fixed dot-free iteration count per solve from
 * the Chebyshev bound, CG for a matrix whose final residual misses the
 * tolerance) and everything else by CG.  Both stop at the same relative
 * tolerance (1e-14 FP64, 1e-7 FP32). */
enum { EMS_TET_SOLVER_AUTO = 0, EMS_TET_SOLVER_CG = 1 };
ems_status ems_tet_set_solver(ems_tet *s, int solver);

/* Iterations of the last mass solve and the largest so far. */
ems_status ems_tet_solver_stats(const ems_tet *s, int *last_iters, int *max_iters);

#ifdef __cplusplus
}
#endif

#endif /* EMS_EMS_H */
