// C++ interface of the tetrahedral Nedelec edge-element solver (ems_tet_* of
// ems.h): orders 1 and 2 on the Kuhn split of a structured block of cubes,
// symplectic_2 / symplectic_4 / ABM steppers (the reference's stepper set,
// src/symplectic_2, src/symplectic_4 and src/fem_abm, here on tetrahedra),
// dipole sources, point receivers, and the discrete-adjoint misfit gradient.
// The method names follow the Python binding emsolvers_amd.tet.TetSimulation.
// Header-only, no HIP types; link with libems.so.  A failing C-ABI call
// throws std::runtime_error with ems_last_error().
#pragma once

#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "ems.h"

namespace ems {

inline void tet_check(ems_status s)
{
    if (s != EMS_OK) throw std::runtime_error(std::string("libems: ") + ems_last_error());
}

namespace tet {

class TetSimulation {
public:
    // grid: cubes per axis (6 tetrahedra each); h: cube edge [m]; stepper:
    // EMS_TET_SYMPLECTIC2 / _SYMPLECTIC4 / _ABM; order 1 or 2.
    TetSimulation(const std::array<int, 3> &grid, double h, double dt, double eps, double mu, int tf, int stepper,
                  int order = 1, int precision = EMS_F64)
        : grid_(grid)
    {
        ems_tet_desc d{grid[0], grid[1], grid[2], h, dt, eps, mu, tf, stepper, order, precision};
        tet_check(ems_tet_create(&d, &h_));
    }
    ~TetSimulation() { ems_tet_destroy(h_); }
    TetSimulation(const TetSimulation &) = delete;
    TetSimulation &operator=(const TetSimulation &) = delete;

    // eps_r and sigma per tetrahedron (6 nx ny nz each; empty = 1 / 0), at rest
    void setMaterial(const std::vector<double> &eps_r, const std::vector<double> &sigma = {})
    {
        tet_check(ems_tet_set_material(h_, eps_r.empty() ? nullptr : eps_r.data(),
                                       sigma.empty() ? nullptr : sigma.data()));
    }
    // dipole of moment amp exp(-(pi f0 (t - t0))^2) along dir at pos
    void addDipole(const std::array<double, 3> &pos, const std::array<double, 3> &dir, double f0, double t0,
                   double amp = 1.0)
    {
        tet_check(ems_tet_add_dipole(h_, pos.data(), dir.data(), f0, t0, amp));
    }
    void addReceiver(const std::array<double, 3> &pos) { tet_check(ems_tet_add_receiver(h_, pos.data())); }
    void setTimeStep(double dt) { tet_check(ems_tet_set_time_step(h_, dt)); }
    void setOperator(int op) { tet_check(ems_tet_set_operator(h_, op)); }  // EMS_TET_OP_*
    void setSolver(int solver) { tet_check(ems_tet_set_solver(h_, solver)); }  // EMS_TET_SOLVER_*

    void run(int num_steps) { tet_check(ems_tet_run(h_, num_steps)); }
    double runTimed(int num_steps)
    {
        double ms = 0.0;
        tet_check(ems_tet_run_timed(h_, num_steps, &ms));
        return ms;
    }

    // w_max (rad/s) of M_eps^-1 S by `iters` power iterations (a lower bound)
    double maxFrequency(int iters = 200)
    {
        double w2 = 0.0;
        tet_check(ems_tet_max_frequency(h_, iters, &w2));
        return std::sqrt(w2);
    }

    // Misfit J of a run of num_steps from rest and dJ/deps_r, dJ/dsigma per
    // tetrahedron (symplectic steppers, lossless models); d_obs:
    // [receiver][slot][3] flattened, slots 0..num_steps/tf.  *drift (may be
    // null): the reconstruction drift of the backward integration.
    double gradient(int num_steps, const std::vector<double> &d_obs, std::vector<double> *grad_eps_r,
                    std::vector<double> *grad_sigma, double *drift = nullptr)
    {
        const size_t nt = tetCount();
        if (grad_eps_r) grad_eps_r->assign(nt, 0.0);
        if (grad_sigma) grad_sigma->assign(nt, 0.0);
        double J = 0.0;
        tet_check(ems_tet_gradient(h_, num_steps, d_obs.data(), d_obs.size(), &J,
                                   grad_eps_r ? grad_eps_r->data() : nullptr,
                                   grad_sigma ? grad_sigma->data() : nullptr, drift));
        return J;
    }

    // traces[receiver][slot] = {Ex, Ey, Ez}, slot q at step q tf
    std::vector<std::vector<std::array<double, 3>>> traces() const
    {
        int nobs = 0, nslot = 0;
        tet_check(ems_tet_trace_shape(h_, &nobs, &nslot));
        std::vector<double> flat((size_t)nobs * nslot * 3);
        tet_check(ems_tet_get_traces(h_, flat.data(), flat.size()));
        std::vector<std::vector<std::array<double, 3>>> out(nobs, std::vector<std::array<double, 3>>(nslot));
        for (int o = 0; o < nobs; ++o)
            for (int s = 0; s < nslot; ++s)
                for (int c = 0; c < 3; ++c) out[o][s][c] = flat[((size_t)o * nslot + s) * 3 + c];
        return out;
    }

    size_t dofCount() const
    {
        size_t n = 0;
        tet_check(ems_tet_dof_count(h_, &n));
        return n;
    }
    size_t tetCount() const { return (size_t)6 * grid_[0] * grid_[1] * grid_[2]; }
    // edge coefficients of E (which = 0) or dE/dt (which = 1)
    std::vector<double> field(int which = 0) const
    {
        std::vector<double> v(dofCount());
        tet_check(ems_tet_get_field(h_, which, v.data(), v.size()));
        return v;
    }
    void setField(int which, const std::vector<double> &values)
    {
        tet_check(ems_tet_set_field(h_, which, values.data(), values.size()));
    }
    double currentTime() const
    {
        double t = 0.0;
        tet_check(ems_tet_get_time(h_, &t));
        return t;
    }
    // {iterations of the last mass solve, largest so far}
    std::array<int, 2> solverStats() const
    {
        int last = 0, mx = 0;
        tet_check(ems_tet_solver_stats(h_, &last, &mx));
        return {last, mx};
    }
    ems_tet *handle() const { return h_; }

private:
    std::array<int, 3> grid_;
    ems_tet *h_ = nullptr;
};

}  // namespace tet
}  // namespace ems
