// Structured-hex edge FEM (fem_v2, symplectic_2, symplectic_4) behind the C ABI.
//
// Drop-in for the reference class FemSimulation of src/fem_v2 and
// src/symplectic_{2,4} with one GPU owning the whole domain: the MPI
// interface sums of the reference (fem_simulation.cpp:657-1113) vanish.
#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "fem_kernels.hpp"

using namespace ems;
using namespace ems::fem;

struct ems_fem {
    ems_fem_desc d{};
    FemGrid g{};
    size_t esz = 8;
    std::vector<double> kscaled;  // n x n, FP64
    DevBuf K, off;
    DevBuf Kp;          // orders 2, 3: K in the MFMA A-operand layout (k_fem_elem_mfma)
    DevBuf Kst;         // order 1: the merged stencil coefficients [3][33] of k_fem1z
    int elem_kind = EMS_FEM_ELEM_AUTO;  // orders 2, 3: element products (per element thread or matrix cores)
    // leapfrog: slot[s][comp]; symplectic: U = slot 0, P = slot 1, and the
    // ping-pong partners slot 2 (U), slot 3 (P)
    DevBuf f[5][3];  // slot 4: drifted field of a symplectic stage (order >= 2)
    DevBuf etmp;     // orders 2, 3: per-element row dots [n][nelem] (k_fem_elem)
    int t0 = 0, t1 = 1, t2 = 2;  // leapfrog rotation
    std::vector<long> src_z;     // z sources (flat Ez index)
    int nsrc_xy = 0;             // x/y sources: stored by the reference, never injected
    std::vector<std::array<long, 9>> obs;
    std::vector<double> traces;  // [nobs][nslot][4]
    int nslot = 0;
    double t_now = 0.0;
    hipStream_t stream = nullptr;
    DevBuf d_src, d_sval, d_obs, d_rec;
    // first order: z-marching all-component kernel with BY rows (4 or 8) and
    // lz planes per workgroup; 0 rows: one gather kernel per component;
    // 16 rows: the FP32 stencil only (8 otherwise); -1: as measured fastest
    // at 128^3 on MI355X: merged stencil FP64 8 x 16, FP32 leapfrog 8 x 8,
    // FP32 symplectic stages 16 x 8; element row dots leapfrog 8 x 8,
    // symplectic stages 4 x 4
    int z1_by = -1, z1_lz = 8;
};

namespace {

// src/fem_v2/fem_simulation.cpp:77-87
double fem_source(double t)
{
    const double freq = 1.0e9;
    const double chi = 1.0 / freq;
    double delay = t - chi;
    double zeta = 2.0 * M_PI * M_PI * freq * freq;
    return 2.0 * zeta * (2.0 * zeta * delay * delay - 1.0) * std::exp(0.0 - zeta * delay * delay) * 0.001;
}

size_t comp_size(const ems_fem *s, int c) { return c == 0 ? s->g.NX : c == 1 ? s->g.NY : s->g.NZ; }

template <typename T> CComp3<T> cc3(ems_fem *s, int slot)
{
    return CComp3<T>{s->f[slot][0].as<T>(), s->f[slot][1].as<T>(), s->f[slot][2].as<T>()};
}
template <typename T> Comp3<T> c3(ems_fem *s, int slot)
{
    return Comp3<T>{s->f[slot][0].as<T>(), s->f[slot][1].as<T>(), s->f[slot][2].as<T>()};
}

void swap_slots(ems_fem *s, int a, int b)
{
    for (int c = 0; c < 3; ++c) {
        std::swap(s->f[a][c].p, s->f[b][c].p);
        std::swap(s->f[a][c].bytes, s->f[b][c].bytes);
    }
}

// 3-D launch over component COMP's array: workgroup 64 x 4
template <int COMP> dim3 comp_grid(const ems_fem *s)
{
    const int dx = s->g.Nx + (COMP != 0), dy = s->g.Ny + (COMP != 1), dz = s->g.Nz + (COMP != 2);
    return dim3((dx + 63) / 64, (dy + 3) / 4, dz);
}

// First order, one z-marching launch for all components (k_fem1z).
template <typename T, bool DRIFT>
void launch_z1(ems_fem *s, CComp3<T> u, CComp3<T> pm, CComp3<T> other, Comp3<T> o1, Comp3<T> o2, T deltat, T hd,
               T mm, const double *sval)
{
    const FemGrid &g = s->g;
    const bool stn = s->elem_kind == EMS_FEM_ELEM_STENCIL || s->elem_kind == EMS_FEM_ELEM_AUTO;
    int by = s->z1_by > 0 ? s->z1_by : !stn ? (DRIFT ? 4 : 8) : (DRIFT && sizeof(T) == 4) ? 16 : 8;
    if (by == 16 && !(stn && sizeof(T) == 4)) by = 8;  // 16 rows: FP32 stencil only (1024-thread register budget)
    const int lz = s->z1_by > 0 ? s->z1_lz : !stn ? (DRIFT ? 4 : 8) : sizeof(T) == 4 ? 8 : 16;
    const int ns = (int)s->src_z.size();
    const dim3 grd((g.Nx + 63) / 64, (g.Ny + by - 1) / by, (g.Nz + lz - 1) / lz);
    const T *K = s->K.as<T>(), *stc = s->Kst.as<T>();
    const long *si = s->d_src.as<long>();
    if constexpr (sizeof(T) == 4) {
        if (by == 16) {
            k_fem1z<T, 16, DRIFT, true><<<grd, dim3(64, 16, 1), 0, s->stream>>>(g, K, stc, u, pm, other, o1, o2, lz,
                                                                               deltat, hd, mm, ns, si, sval);
            return;
        }
    }
    if (by == 4 && stn)
        k_fem1z<T, 4, DRIFT, true><<<grd, dim3(64, 4, 1), 0, s->stream>>>(g, K, stc, u, pm, other, o1, o2, lz,
                                                                         deltat, hd, mm, ns, si, sval);
    else if (by == 4)
        k_fem1z<T, 4, DRIFT, false><<<grd, dim3(64, 4, 1), 0, s->stream>>>(g, K, stc, u, pm, other, o1, o2, lz,
                                                                          deltat, hd, mm, ns, si, sval);
    else if (stn)
        k_fem1z<T, 8, DRIFT, true><<<grd, dim3(64, 8, 1), 0, s->stream>>>(g, K, stc, u, pm, other, o1, o2, lz,
                                                                         deltat, hd, mm, ns, si, sval);
    else
        k_fem1z<T, 8, DRIFT, false><<<grd, dim3(64, 8, 1), 0, s->stream>>>(g, K, stc, u, pm, other, o1, o2, lz,
                                                                          deltat, hd, mm, ns, si, sval);
}

// Row dots of every element (orders 2, 3) into etmp: per-element threads
// (the oracle's sums) or matrix cores.
template <typename T, int P> void launch_elem(ems_fem *s, CComp3<T> x)
{
    const long nelem = (long)s->g.nex * s->g.ney * s->g.nez;
    T *et = s->etmp.as<T>();
    const bool mfma = s->elem_kind == EMS_FEM_ELEM_MFMA || (s->elem_kind == EMS_FEM_ELEM_AUTO && P == 3);
    if (mfma)
        k_fem_elem_mfma<T, P><<<(unsigned)((nelem + 63) / 64), 256, 0, s->stream>>>(s->g, s->Kp.as<T>(),
                                                                                   s->off.as<int>(), x, et);
    else
        k_fem_elem<T, P><<<(unsigned)((nelem + kElemWG<P> - 1) / kElemWG<P>), kElemWG<P>, 0, s->stream>>>(
            s->g, s->K.as<T>(), s->off.as<int>(), x, et);
}

template <typename T, int P>
void launch_leapfrog(ems_fem *s, const double *sval)
{
    const dim3 blk(64, 4, 1);
    const CComp3<T> e0 = cc3<T>(s, s->t0), e1 = cc3<T>(s, s->t1);
    const Comp3<T> e2 = c3<T>(s, s->t2);
    const int ns = (int)s->src_z.size();
    if (P == 1 && s->z1_by != 0) {
        launch_z1<T, false>(s, e1, e1, e0, e2, e2, (T)0, (T)0, (T)1, sval);
        return;
    }
    if (P == 1) {  // fields staged in LDS
        k_fem_leapfrog1<T, 0><<<comp_grid<0>(s), blk, 0, s->stream>>>(s->g, s->K.as<T>(), e0, e1, e2, ns,
                                                                     s->d_src.as<long>(), sval);
        k_fem_leapfrog1<T, 1><<<comp_grid<1>(s), blk, 0, s->stream>>>(s->g, s->K.as<T>(), e0, e1, e2, ns,
                                                                     s->d_src.as<long>(), sval);
        k_fem_leapfrog1<T, 2><<<comp_grid<2>(s), blk, 0, s->stream>>>(s->g, s->K.as<T>(), e0, e1, e2, ns,
                                                                     s->d_src.as<long>(), sval);
        return;
    }
    if constexpr (P >= 2) {  // element by element: row dots once per element, then the gathers
        T *et = s->etmp.as<T>();
        launch_elem<T, P>(s, e1);
        k_fem_leapfrog<T, P, 0><<<comp_grid<0>(s), blk, 0, s->stream>>>(s->g, e0, e1, e2, ns, s->d_src.as<long>(), sval,
                                                                         et);
        k_fem_leapfrog<T, P, 1><<<comp_grid<1>(s), blk, 0, s->stream>>>(s->g, e0, e1, e2, ns, s->d_src.as<long>(), sval,
                                                                         et);
        k_fem_leapfrog<T, P, 2><<<comp_grid<2>(s), blk, 0, s->stream>>>(s->g, e0, e1, e2, ns, s->d_src.as<long>(), sval,
                                                                         et);
    }
}

template <typename T, int P>
void launch_stage(ems_fem *s, double deltat, const double *sval)
{
    const double mm = 8.0 / s->g.p / s->g.p / s->g.p;
    const dim3 blk(64, 4, 1);
    const CComp3<T> u = cc3<T>(s, 0), pm = cc3<T>(s, 1);
    const Comp3<T> uo = c3<T>(s, 2), po = c3<T>(s, 3);
    const int ns = (int)s->src_z.size();
    if (P == 1 && s->z1_by != 0) {
        launch_z1<T, true>(s, u, pm, pm, uo, po, (T)deltat, (T)(0.5 * deltat), (T)mm, sval);
        swap_slots(s, 0, 2);
        swap_slots(s, 1, 3);
        return;
    }
    if (P == 1) {  // fields staged in LDS
        k_fem_symplectic1<T, 0><<<comp_grid<0>(s), blk, 0, s->stream>>>(
            s->g, s->K.as<T>(), u, pm, uo, po, (T)deltat, (T)(0.5 * deltat), (T)mm, ns, s->d_src.as<long>(), sval);
        k_fem_symplectic1<T, 1><<<comp_grid<1>(s), blk, 0, s->stream>>>(
            s->g, s->K.as<T>(), u, pm, uo, po, (T)deltat, (T)(0.5 * deltat), (T)mm, ns, s->d_src.as<long>(), sval);
        k_fem_symplectic1<T, 2><<<comp_grid<2>(s), blk, 0, s->stream>>>(
            s->g, s->K.as<T>(), u, pm, uo, po, (T)deltat, (T)(0.5 * deltat), (T)mm, ns, s->d_src.as<long>(), sval);
        swap_slots(s, 0, 2);
        swap_slots(s, 1, 3);
        return;
    }
    if constexpr (P >= 2) {  // the drifted field once per stage, then element by element
        const long nmax = std::max({s->g.NX, s->g.NY, s->g.NZ});
        k_fem_drift<T><<<dim3((unsigned)((nmax + 255) / 256), 3, 1), 256, 0, s->stream>>>(
            s->g, u, pm, c3<T>(s, 4), (T)(0.5 * deltat), (T)mm);
        const CComp3<T> ud = cc3<T>(s, 4);
        T *et = s->etmp.as<T>();
        launch_elem<T, P>(s, ud);
        k_fem_symplectic<T, P, 0><<<comp_grid<0>(s), blk, 0, s->stream>>>(
            s->g, u, pm, uo, po, (T)deltat, (T)(0.5 * deltat), (T)mm, ns, s->d_src.as<long>(), sval, et);
        k_fem_symplectic<T, P, 1><<<comp_grid<1>(s), blk, 0, s->stream>>>(
            s->g, u, pm, uo, po, (T)deltat, (T)(0.5 * deltat), (T)mm, ns, s->d_src.as<long>(), sval, et);
        k_fem_symplectic<T, P, 2><<<comp_grid<2>(s), blk, 0, s->stream>>>(
            s->g, u, pm, uo, po, (T)deltat, (T)(0.5 * deltat), (T)mm, ns, s->d_src.as<long>(), sval, et);
        swap_slots(s, 0, 2);
        swap_slots(s, 1, 3);
    }
}

template <typename T> void step_leapfrog(ems_fem *s, const double *sval)
{
    switch (s->g.p) {
    case 1: launch_leapfrog<T, 1>(s, sval); break;
    case 2: launch_leapfrog<T, 2>(s, sval); break;
    default: launch_leapfrog<T, 3>(s, sval); break;
    }
}

template <typename T> void step_stage(ems_fem *s, double deltat, const double *sval)
{
    switch (s->g.p) {
    case 1: launch_stage<T, 1>(s, deltat, sval); break;
    case 2: launch_stage<T, 2>(s, deltat, sval); break;
    default: launch_stage<T, 3>(s, deltat, sval); break;
    }
}

template <typename T> ems_status record(ems_fem *s, int slot, size_t row)
{
    const int nobs = (int)s->obs.size();
    if (!nobs) return EMS_OK;
    k_fem_record<T><<<(nobs + 63) / 64, 64, 0, s->stream>>>(cc3<T>(s, slot), nobs, s->d_obs.as<long>(),
                                                            s->d_rec.as<double>() + row * 9 * nobs);
    EMS_HIP(hipGetLastError());
    return EMS_OK;
}

template <typename T> ems_status run_impl(ems_fem *s, int num_steps, double *total_ms, double *kernel_ms)
{
    const int tf = s->d.tf;
    // saved_electric_field_ resize (src/fem_v2/fem_simulation.cpp:1251-1255)
    const int nslot = (num_steps - 1) / tf + 1;
    const int nobs = (int)s->obs.size();
    if (nslot != s->nslot) {
        std::vector<double> t((size_t)nobs * std::max(nslot, 0) * 4, 0.0);
        for (int o = 0; o < nobs; ++o)
            for (int q = 0; q < std::min(nslot, s->nslot); ++q)
                for (int c = 0; c < 4; ++c)
                    t[((size_t)o * nslot + q) * 4 + c] = s->traces[((size_t)o * s->nslot + q) * 4 + c];
        s->traces.swap(t);
        s->nslot = std::max(nslot, 0);
    }
    if (num_steps <= 0) return EMS_OK;
    const bool leap = s->d.stepper == EMS_FEM_LEAPFROG;
    const int nstage = leap ? 1 : s->d.stepper == EMS_FEM_SYMPLECTIC2 ? 1 : 3;
    const double x1 = 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0));
    const double x2 = 1.0 - 2.0 * x1;
    const double dt = s->d.dt, h = s->d.h, hp = s->d.h * s->g.p;
    // per-(step, stage) source value and the host time stamps
    std::vector<double> sval((size_t)num_steps * nstage);
    std::vector<double> stamp(num_steps);
    double t = s->t_now;
    for (int st = 0; st < num_steps; ++st) {
        stamp[st] = t;
        if (leap) {
            sval[st] = fem_source(t) * dt * dt / s->d.eps / h / h / h;
        } else if (nstage == 1) {
            sval[st] = fem_source(t + 0.0 + 0.5 * dt) * 8.0 / s->d.eps / hp / hp / hp;
        } else {
            const double D[3] = {x1 * dt, x2 * dt, x1 * dt};
            const double O[3] = {0.0, x1 * dt, (x1 + x2) * dt};
            for (int q = 0; q < 3; ++q)
                sval[(size_t)st * 3 + q] = fem_source(t + O[q] + 0.5 * D[q]) * 8.0 / s->d.eps / hp / hp / hp;
        }
        t += dt;
    }
    const int nrec = (num_steps - 1) / tf + 1;
    EMS_TRY(s->d_sval.reserve(sizeof(double) * sval.size()));
    EMS_HIP(hipMemcpy(s->d_sval.p, sval.data(), sizeof(double) * sval.size(), hipMemcpyHostToDevice));
    EMS_TRY(s->d_rec.reserve(sizeof(double) * std::max((size_t)nrec * nobs * 9, (size_t)1)));

    const bool timed = total_ms != nullptr;
    std::vector<hipEvent_t> ev;
    if (timed) {
        ev.resize(2 * (size_t)num_steps + 2, nullptr);
        for (auto &e : ev) EMS_HIP(hipEventCreate(&e));
        EMS_HIP(hipEventRecord(ev[2 * (size_t)num_steps], s->stream));
    }
    const double *dsv = s->d_sval.as<double>();
    for (int st = 0; st < num_steps; ++st) {
        if (leap) {
            if (timed) EMS_HIP(hipEventRecord(ev[2 * (size_t)st], s->stream));
            step_leapfrog<T>(s, dsv + st);
            if (timed) EMS_HIP(hipEventRecord(ev[2 * (size_t)st + 1], s->stream));
            EMS_HIP(hipGetLastError());
            // the reference records E^n (slot time_1_) after the update
            if (st % tf == 0) EMS_TRY(record<T>(s, s->t1, st / tf));
            const int o0 = s->t0;
            s->t0 = s->t1;
            s->t1 = s->t2;
            s->t2 = o0;
        } else {
            // the symplectic family records U before the step
            if (st % tf == 0) EMS_TRY(record<T>(s, 0, st / tf));
            if (timed) EMS_HIP(hipEventRecord(ev[2 * (size_t)st], s->stream));
            if (nstage == 1) {
                step_stage<T>(s, dt, dsv + st);
            } else {
                step_stage<T>(s, x1 * dt, dsv + (size_t)st * 3);
                step_stage<T>(s, x2 * dt, dsv + (size_t)st * 3 + 1);
                step_stage<T>(s, x1 * dt, dsv + (size_t)st * 3 + 2);
            }
            if (timed) EMS_HIP(hipEventRecord(ev[2 * (size_t)st + 1], s->stream));
            EMS_HIP(hipGetLastError());
        }
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
    if (nobs) {
        std::vector<double> raw((size_t)nrec * nobs * 9);
        EMS_HIP(hipMemcpy(raw.data(), s->d_rec.p, sizeof(double) * raw.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < nrec && r < s->nslot; ++r)
            for (int o = 0; o < nobs; ++o) {
                const double *v = &raw[((size_t)r * nobs + o) * 9];
                double *dst = &s->traces[((size_t)o * s->nslot + r) * 4];
                dst[0] = stamp[(size_t)r * tf];
                dst[1] = (v[0] + v[1] + v[2] + v[3]) / 4.0;
                dst[2] = (v[4] + v[5] + v[6] + v[7]) / 4.0;
                dst[3] = v[8];
            }
    }
    return EMS_OK;
}

template <typename T> ems_status upload_kmat(ems_fem *s)
{
    std::vector<T> k(s->kscaled.begin(), s->kscaled.end());
    EMS_TRY(s->K.reserve(sizeof(T) * k.size()));
    EMS_HIP(hipMemcpy(s->K.p, k.data(), sizeof(T) * k.size(), hipMemcpyHostToDevice));
    if (s->g.p >= 2) {  // the A-operand layout of k_fem_elem_mfma
        const int n = (int)std::lround(std::sqrt((double)k.size()));
        const int mt = (n + 15) / 16, ks = (n + 3) / 4;
        std::vector<T> kp((size_t)mt * ks * 64, (T)0);
        for (int q = 0; q < mt; ++q)
            for (int c = 0; c < ks; ++c)
                for (int l = 0; l < 64; ++l) {
                    const int r = 16 * q + (l & 15), m = 4 * c + (l >> 4);
                    if (r < n && m < n) kp[((size_t)q * ks + c) * 64 + l] = k[(size_t)r * n + m];
                }
        EMS_TRY(s->Kp.reserve(sizeof(T) * kp.size()));
        EMS_HIP(hipMemcpy(s->Kp.p, kp.data(), sizeof(T) * kp.size(), hipMemcpyHostToDevice));
    } else {
        // order 1: each component's 4 element rows summed into the terms of
        // the merged stencil
        std::vector<double> c(3 * kZ1Terms);
        if (!z1_stencil_coefs(s->kscaled.data(), false, c.data()))
            return not_supported("order-1 stencil term missing");
        std::vector<T> ct(c.begin(), c.end());
        EMS_TRY(s->Kst.reserve(sizeof(T) * ct.size()));
        EMS_HIP(hipMemcpy(s->Kst.p, ct.data(), sizeof(T) * ct.size(), hipMemcpyHostToDevice));
    }
    return EMS_OK;
}


int logical_slot(const ems_fem *s, int logical)
{
    if (s->d.stepper != EMS_FEM_LEAPFROG) return logical;  // 0 = U, 1 = P
    return logical == 0 ? s->t0 : logical == 1 ? s->t1 : -1;
}

// src/fem_v2/fem_simulation.cpp:1477-1493
double parse_value(const std::string &str)
{
    size_t start = str.find_first_not_of(" \t\r\n\"");
    size_t end = str.find_last_not_of(" \t\r\n\"");
    if (start == std::string::npos) return 0.0;
    std::string t = str.substr(start, end - start + 1);
    size_t slash = t.find('/');
    if (slash != std::string::npos) return std::stod(t.substr(0, slash)) / std::stod(t.substr(slash + 1));
    return std::stod(t);
}

}  // namespace

extern "C" {

ems_status ems_fem_load_kmat(const char *path, double *out, size_t count, int *rows)
{
    clear_error();
    if (!path) return invalid("null path");
    std::ifstream file(path);
    if (!file.is_open()) {
        set_error(std::string("Cannot open file: ") + path);
        return EMS_ERR_IO;
    }
    std::vector<std::vector<double>> m;
    std::string line;
    try {
        while (std::getline(file, line)) {
            std::vector<double> row;
            std::stringstream ss(line);
            std::string cell;
            while (std::getline(ss, cell, ',')) row.push_back(parse_value(cell));
            if (!row.empty()) m.push_back(row);
        }
    } catch (const std::exception &e) {
        set_error(std::string("malformed kmat file ") + path + ": " + e.what());
        return EMS_ERR_IO;
    }
    const size_t n = m.size();
    for (auto &r : m)
        if (r.size() != n) {
            set_error(std::string("kmat file is not square: ") + path);
            return EMS_ERR_IO;
        }
    if (rows) *rows = (int)n;
    if (!out) return EMS_OK;
    if (count != n * n) return invalid("kmat buffer has the wrong size");
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) out[i * n + j] = m[i][j];
    return EMS_OK;
}

ems_status ems_fem_create(const ems_fem_desc *desc, const double *kmat, ems_fem **out)
{
    clear_error();
    if (!desc || !kmat || !out) return invalid("ems_fem_create: null argument");
    *out = nullptr;
    const int p = desc->order;
    if (p < 1 || p > 3) return invalid("order must be 1, 2 or 3");
    if (desc->nx <= 0 || desc->ny <= 0 || desc->nz <= 0) return invalid("grid sizes must be positive");
    if (desc->nx % p || desc->ny % p || desc->nz % p) return invalid("grid sizes must be multiples of the order");
    if (desc->nx / p < 2 || desc->ny / p < 2 || desc->nz / p < 2)
        return invalid("at least two elements per axis are required");
    if (desc->tf <= 0) return invalid("time_frequency must be >= 1");
    if (!(desc->h > 0) || !(desc->dt > 0) || !(desc->eps > 0) || !(desc->mu > 0))
        return invalid("h, dt, eps and mu must be positive");
    if (desc->stepper < EMS_FEM_LEAPFROG || desc->stepper > EMS_FEM_SYMPLECTIC4) return invalid("unknown stepper");
    if (desc->precision != EMS_F64 && desc->precision != EMS_F32) return invalid("unknown precision");
    if (ems_device_count() <= 0) {
        set_error("no HIP device available");
        return EMS_ERR_HIP;
    }
    ems_fem *s = new ems_fem();
    s->d = *desc;
    s->esz = desc->precision == EMS_F32 ? 4 : 8;
    FemGrid &g = s->g;
    g.p = p;
    g.Nx = desc->nx;
    g.Ny = desc->ny;
    g.Nz = desc->nz;
    g.nex = desc->nx / p;
    g.ney = desc->ny / p;
    g.nez = desc->nz / p;
    const long Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    g.NX = Nx * (Ny + 1) * (Nz + 1);
    g.NY = (Nx + 1) * Ny * (Nz + 1);
    g.NZ = (Nx + 1) * (Ny + 1) * Nz;
    const long a = g.nex, b = g.ney, c = g.nez;
    g.OE = 2 * (a * b + b * c + c * a) - 4 * (a + b + c) + 8;
    g.D = p * (p + 1) * (p + 1);
    g.n = 3 * g.D;
    const int n = g.n;
    // element operator scale (fem_v2: fem_simulation.cpp:476; symplectic: :473)
    const double h = desc->h, dt = desc->dt;
    double sc;
    if (desc->stepper == EMS_FEM_LEAPFROG)
        sc = 4.0 * dt * dt / (h * p) / (h * p) / desc->mu / desc->eps / (8.0 / p / p / p);
    else
        sc = 4.0 / (h * p) / (h * p) / desc->mu / desc->eps;
    s->kscaled.resize((size_t)n * n);
    for (int q = 0; q < n * n; ++q) s->kscaled[q] = sc * kmat[q];
    // local DOF -> offset from the element base
    std::vector<int> off(n);
    for (int nn = 0; nn < p + 1; ++nn)
        for (int m = 0; m < p + 1; ++m)
            for (int l = 0; l < p; ++l) off[l + m * p + nn * p * (p + 1)] = (int)(l + m * Nx + nn * Nx * (Ny + 1));
    for (int nn = 0; nn < p + 1; ++nn)
        for (int m = 0; m < p; ++m)
            for (int l = 0; l < p + 1; ++l)
                off[g.D + l + m * (p + 1) + nn * p * (p + 1)] = (int)(l + m * (Nx + 1) + nn * (Nx + 1) * Ny);
    for (int nn = 0; nn < p; ++nn)
        for (int m = 0; m < p + 1; ++m)
            for (int l = 0; l < p + 1; ++l)
                off[2 * g.D + l + m * (p + 1) + nn * (p + 1) * (p + 1)] =
                    (int)(l + m * (Nx + 1) + nn * (Nx + 1) * (Ny + 1));
    ems_status st = s->off.reserve(sizeof(int) * n);
    if (st == EMS_OK && hipMemcpy(s->off.p, off.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess)
        st = hip_fail(hipGetLastError(), "hipMemcpy", __FILE__, __LINE__);
    if (st == EMS_OK) st = desc->precision == EMS_F32 ? upload_kmat<float>(s) : upload_kmat<double>(s);
    const int nslots = desc->stepper == EMS_FEM_LEAPFROG ? 3 : desc->order >= 2 ? 5 : 4;
    if (st == EMS_OK && desc->order >= 2) st = s->etmp.reserve((size_t)g.n * g.nex * g.ney * g.nez * s->esz);
    for (int sl = 0; sl < nslots && st == EMS_OK; ++sl)
        for (int cc = 0; cc < 3 && st == EMS_OK; ++cc) {
            st = s->f[sl][cc].reserve(comp_size(s, cc) * s->esz);
            if (st == EMS_OK && hipMemset(s->f[sl][cc].p, 0, comp_size(s, cc) * s->esz) != hipSuccess)
                st = hip_fail(hipGetLastError(), "hipMemset", __FILE__, __LINE__);
        }
    if (st != EMS_OK) {
        delete s;
        return st;
    }
    *out = s;
    return EMS_OK;
}

void ems_fem_destroy(ems_fem *s) { delete s; }

ems_status ems_fem_add_source(ems_fem *s, int comp, int i, int j, int k)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (comp < 0 || comp > 2) return invalid("source component must be 0, 1 or 2");
    const long Nx = s->g.Nx, Ny = s->g.Ny, Nz = s->g.Nz;
    if (comp != 2) {
        // setSource_x / _y: stored by the reference but never injected
        s->nsrc_xy++;
        return EMS_OK;
    }
    if (i < 0 || j < 0 || k < 0 || i > Nx || j > Ny || k >= Nz) return invalid("source outside the Ez array");
    s->src_z.push_back(i + j * (Nx + 1) + k * (Nx + 1) * (Ny + 1));
    EMS_TRY(s->d_src.reserve(sizeof(long) * s->src_z.size()));
    EMS_HIP(hipMemcpy(s->d_src.p, s->src_z.data(), sizeof(long) * s->src_z.size(), hipMemcpyHostToDevice));
    return EMS_OK;
}

ems_status ems_fem_add_observation(ems_fem *s, int i, int j, int k)
{
    clear_error();
    if (!s) return invalid("null handle");
    const long Nx = s->g.Nx, Ny = s->g.Ny;
    const long o0 = i + j * Nx + k * Nx * (Ny + 1);
    const long o1 = i + j * (Nx + 1) + k * (Nx + 1) * Ny;
    const long o2 = i + j * (Nx + 1) + k * (Nx + 1) * (Ny + 1);
    std::array<long, 9> q = {o0 - 1, o0, o0 + Nx * (Ny + 1) - 1, o0 + Nx * (Ny + 1),
                             o1 - (Nx + 1), o1, o1 + (Nx + 1) * (Ny - 1), o1 + (Nx + 1) * Ny, o2};
    for (int a = 0; a < 9; ++a) {
        const long lim = a < 4 ? s->g.NX : a < 8 ? s->g.NY : s->g.NZ;
        if (i < 0 || j < 0 || k < 0 || q[a] < 0 || q[a] >= lim)
            return invalid("observation point reads outside the field arrays");
    }
    s->obs.push_back(q);
    std::vector<long> flat;
    for (auto &o : s->obs) flat.insert(flat.end(), o.begin(), o.end());
    EMS_TRY(s->d_obs.reserve(sizeof(long) * flat.size()));
    EMS_HIP(hipMemcpy(s->d_obs.p, flat.data(), sizeof(long) * flat.size(), hipMemcpyHostToDevice));
    std::vector<double> t(s->obs.size() * (size_t)s->nslot * 4, 0.0);
    std::copy(s->traces.begin(), s->traces.end(), t.begin());
    s->traces.swap(t);
    return EMS_OK;
}

ems_status ems_fem_run(ems_fem *s, int num_steps)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return s->d.precision == EMS_F32 ? run_impl<float>(s, num_steps, nullptr, nullptr)
                                     : run_impl<double>(s, num_steps, nullptr, nullptr);
}

ems_status ems_fem_run_timed(ems_fem *s, int num_steps, double *total_ms, double *kernel_ms)
{
    clear_error();
    if (!s || !total_ms || !kernel_ms) return invalid("null argument");
    if (num_steps < 0) return invalid("num_steps must be >= 0");
    return s->d.precision == EMS_F32 ? run_impl<float>(s, num_steps, total_ms, kernel_ms)
                                     : run_impl<double>(s, num_steps, total_ms, kernel_ms);
}

ems_status ems_fem_trace_shape(const ems_fem *s, int *nobs, int *nslot)
{
    clear_error();
    if (!s || !nobs || !nslot) return invalid("null argument");
    *nobs = (int)s->obs.size();
    *nslot = s->nslot;
    return EMS_OK;
}

ems_status ems_fem_get_traces(ems_fem *s, double *out, size_t count)
{
    clear_error();
    if (!s || (!out && count)) return invalid("null argument");
    if (count != s->traces.size()) return invalid("trace buffer has the wrong size");
    std::copy(s->traces.begin(), s->traces.end(), out);
    return EMS_OK;
}

ems_status ems_fem_field_size(const ems_fem *s, int comp, size_t *n)
{
    clear_error();
    if (!s || !n) return invalid("null argument");
    if (comp < 0 || comp > 2) return invalid("component must be 0, 1 or 2");
    *n = comp_size(s, comp);
    return EMS_OK;
}

ems_status ems_fem_get_field(ems_fem *s, int comp, int slot, double *out, size_t count)
{
    clear_error();
    if (!s || !out) return invalid("null argument");
    if (comp < 0 || comp > 2) return invalid("component must be 0, 1 or 2");
    const int sl = logical_slot(s, slot);
    if (slot < 0 || slot > 1 || sl < 0) return invalid("slot must be 0 or 1");
    if (count != comp_size(s, comp)) return invalid("field buffer has the wrong size");
    EMS_HIP(hipStreamSynchronize(s->stream));
    if (s->esz == 8) {
        EMS_HIP(hipMemcpy(out, s->f[sl][comp].p, count * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t(count);
        EMS_HIP(hipMemcpy(t.data(), s->f[sl][comp].p, count * 4, hipMemcpyDeviceToHost));
        std::copy(t.begin(), t.end(), out);
    }
    return EMS_OK;
}

ems_status ems_fem_set_field(ems_fem *s, int comp, int slot, const double *in, size_t count)
{
    clear_error();
    if (!s || !in) return invalid("null argument");
    if (comp < 0 || comp > 2) return invalid("component must be 0, 1 or 2");
    const int sl = logical_slot(s, slot);
    if (slot < 0 || slot > 1 || sl < 0) return invalid("slot must be 0 or 1");
    if (count != comp_size(s, comp)) return invalid("field buffer has the wrong size");
    EMS_HIP(hipStreamSynchronize(s->stream));
    if (s->esz == 8) {
        EMS_HIP(hipMemcpy(s->f[sl][comp].p, in, count * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(in, in + count);
        EMS_HIP(hipMemcpy(s->f[sl][comp].p, t.data(), count * 4, hipMemcpyHostToDevice));
    }
    return EMS_OK;
}

ems_status ems_fem_get_time(const ems_fem *s, double *t)
{
    clear_error();
    if (!s || !t) return invalid("null argument");
    *t = s->t_now;
    return EMS_OK;
}

// FemSimulation::saveResults (src/fem_v2/fem_simulation.cpp:1324-1349):
// every observation writes the same file name (a reference quirk: the last
// one wins); values in scientific notation with 15 digits.
ems_status ems_fem_save_results(ems_fem *s, int num_steps, const char *filename)
{
    (void)num_steps;
    clear_error();
    if (!s || !filename) return invalid("null argument");
    for (size_t i = 0; i < s->obs.size(); ++i) {
        std::ofstream file(filename);
        if (!file.is_open()) {
            set_error(std::string("Error: Could not open file ") + filename);
            return EMS_ERR_IO;
        }
        file << std::setprecision(15) << std::scientific;
        file << "time,Ex,Ey,Ez" << std::endl;
        for (int j = 0; j < s->nslot; ++j) {
            const double *v = &s->traces[(i * s->nslot + j) * 4];
            file << v[0] << "," << v[1] << "," << v[2] << "," << v[3] << std::endl;
        }
    }
    return EMS_OK;
}

ems_status ems_fem_set_tiling(ems_fem *s, int rows, int planes)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (rows != -1 && rows != 0 && rows != 4 && rows != 8 && rows != 16)
        return invalid("tile rows must be -1, 0, 4, 8 or 16");
    if (planes < 1) return invalid("tile planes must be >= 1");
    s->z1_by = rows;
    s->z1_lz = planes;
    return EMS_OK;
}

ems_status ems_fem_set_element_product(ems_fem *s, int kind)
{
    clear_error();
    if (!s) return invalid("null handle");
    if (kind != EMS_FEM_ELEM_AUTO && kind != EMS_FEM_ELEM_ROWS && kind != EMS_FEM_ELEM_MFMA &&
        kind != EMS_FEM_ELEM_STENCIL)
        return invalid("unknown element product");
    if (kind == EMS_FEM_ELEM_STENCIL && s->g.p != 1) return invalid("the merged stencil is first order only");
    s->elem_kind = kind;
    return EMS_OK;
}

ems_status ems_fem_dof_count(const ems_fem *s, size_t *n)
{
    clear_error();
    if (!s || !n) return invalid("null argument");
    *n = (size_t)(s->g.NX + s->g.NY + s->g.NZ);
    return EMS_OK;
}

}  // extern "C"
