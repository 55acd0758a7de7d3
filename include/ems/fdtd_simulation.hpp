// C++ drop-in for the reference FDTD class (src/fdtd/fdtd_simulation.hpp:9-84).
//
// Same constructor, method names, argument types and output files as the
// reference FemSimulation; the work runs on the GPU through the C ABI of
// ems.h (link with libems.so).  Header-only, no HIP types.
//
//   #include "fdtd_simulation.hpp"      ->   #include <ems/fdtd_simulation.hpp>
//   FemSimulation sim(...)              ->   ems::fdtd::FemSimulation sim(...)
//
// Errors: the reference has none in these calls; a failing C-ABI call throws
// std::runtime_error with ems_last_error().  saveResults keeps the
// reference behaviour (message on std::cerr, return).
#pragma once

#include <array>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ems.h"

namespace ems {

inline void check(ems_status s)
{
    if (s != EMS_OK) throw std::runtime_error(std::string("libems: ") + ems_last_error());
}

namespace fdtd {

class FemSimulation {
public:
    FemSimulation(const std::array<int, 3> &grid_size, double domain_size, double time_step,
                  double permittivity, double permeability, int time_frequency, int precision = EMS_F64)
    {
        ems_fdtd_desc d{grid_size[0], grid_size[1], grid_size[2], domain_size, time_step, permittivity,
                        permeability, time_frequency, precision};
        check(ems_fdtd_create(&d, &h_));
        grid_ = grid_size;
    }
    ~FemSimulation() { ems_fdtd_destroy(h_); }
    FemSimulation(const FemSimulation &) = delete;
    FemSimulation &operator=(const FemSimulation &) = delete;

    void run(int num_steps) { check(ems_fdtd_run(h_, num_steps)); }

    void setSource_x(const std::array<int, 3> &p) { check(ems_fdtd_add_source(h_, 0, p[0], p[1], p[2])); }
    void setSource_y(const std::array<int, 3> &p) { check(ems_fdtd_add_source(h_, 1, p[0], p[1], p[2])); }
    void setSource_z(const std::array<int, 3> &p) { check(ems_fdtd_add_source(h_, 2, p[0], p[1], p[2])); }

    void setObservationPoint(const std::array<std::vector<std::array<int, 3>>, 3> &position)
    {
        int counts[3];
        std::vector<int> pts;
        for (int c = 0; c < 3; ++c) {
            counts[c] = (int)position[c].size();
            for (auto &p : position[c]) pts.insert(pts.end(), p.begin(), p.end());
        }
        check(ems_fdtd_add_observation(h_, counts, pts.data()));
    }

    void saveResults(int num_steps, double c)
    {
        if (ems_fdtd_save_results(h_, num_steps, c, "") != EMS_OK) std::cerr << ems_last_error() << std::endl;
    }

    // ---- extensions (no reference counterpart) ----
    void setMaterial(const std::vector<double> &eps_r, const std::vector<double> &sigma = {})
    {
        check(ems_fdtd_set_material(h_, eps_r.empty() ? nullptr : eps_r.data(),
                                    sigma.empty() ? nullptr : sigma.data()));
    }
    void setCPML(int cells, double order = 3.0, double r0 = 1e-6, double kappa_max = 1.0, double alpha_max = 0.0)
    {
        check(ems_fdtd_set_cpml(h_, cells, order, r0, kappa_max, alpha_max));
    }
    void setWavelet(int kind, double f0, double t0 = -1.0, double amp = 1.0)
    {
        check(ems_fdtd_set_wavelet(h_, kind, f0, t0, amp));
    }
    // saved_electric_field_: [observation][slot] = {Ex, Ey, Ez}
    std::vector<std::vector<std::array<double, 3>>> savedElectricField() const
    {
        int nobs = 0, nslot = 0;
        check(ems_fdtd_trace_shape(h_, &nobs, &nslot));
        std::vector<double> flat((size_t)nobs * nslot * 3);
        check(ems_fdtd_get_traces(h_, flat.data(), flat.size()));
        std::vector<std::vector<std::array<double, 3>>> out(nobs, std::vector<std::array<double, 3>>(nslot));
        for (int o = 0; o < nobs; ++o)
            for (int s = 0; s < nslot; ++s)
                for (int c = 0; c < 3; ++c) out[o][s][c] = flat[((size_t)o * nslot + s) * 3 + c];
        return out;
    }
    std::vector<double> field(int which) const
    {
        size_t n = 0;
        check(ems_fdtd_field_size(h_, which, &n));
        std::vector<double> v(n);
        check(ems_fdtd_get_field(h_, which, v.data(), n));
        return v;
    }
    void setField(int which, const std::vector<double> &values)
    {
        check(ems_fdtd_set_field(h_, which, values.data(), values.size()));
    }
    double currentTime() const
    {
        double t = 0.0;
        check(ems_fdtd_get_time(h_, &t));
        return t;
    }

    // Misfit J = 1/2 sum (trace - d_obs)^2 of a run of num_steps from rest
    // and dJ/deps_r, dJ/dsigma on the (nx+1)(ny+1)(nz+1) material grid
    // (ems_fdtd_gradient).  d_obs: [observation][slot][3] flattened, slots
    // 0..num_steps/tf.  Either gradient pointer may be null.
    double gradient(int num_steps, const std::vector<double> &d_obs, std::vector<double> *grad_eps_r,
                    std::vector<double> *grad_sigma)
    {
        const size_t nm = materialSize();
        if (grad_eps_r) grad_eps_r->assign(nm, 0.0);
        if (grad_sigma) grad_sigma->assign(nm, 0.0);
        double J = 0.0;
        check(ems_fdtd_gradient(h_, num_steps, d_obs.data(), d_obs.size(), &J,
                                grad_eps_r ? grad_eps_r->data() : nullptr, grad_sigma ? grad_sigma->data() : nullptr));
        return J;
    }
    void setCheckpointInterval(int interval) { check(ems_fdtd_set_checkpoint_interval(h_, interval)); }
    void setGradientMemory(size_t bytes) { check(ems_fdtd_set_gradient_memory(h_, bytes)); }
    // {level-1 interval, segment length, bytes} of the schedule gradient(num_steps) would take now
    std::array<size_t, 3> gradientPlan(int num_steps)
    {
        int k1 = 0, ks = 0;
        size_t b = 0;
        check(ems_fdtd_gradient_plan(h_, num_steps, &k1, &ks, &b));
        return {(size_t)k1, (size_t)ks, b};
    }
    // points of the material grid, (nx+1)(ny+1)(nz+1)
    size_t materialSize() const { return (size_t)(grid_[0] + 1) * (grid_[1] + 1) * (grid_[2] + 1); }
    ems_fdtd *handle() const { return h_; }

private:
    std::array<int, 3> grid_{};
    ems_fdtd *h_ = nullptr;
};

// Many handles advanced together (ems_batch_*): one launch per sweep for all
// members, optional HIP-graph replay, batched gradients of shots on one
// model.  Members stay owned by the caller and must outlive the batch.
class Batch {
public:
    Batch() { check(ems_batch_create(&b_)); }
    ~Batch() { ems_batch_destroy(b_); }
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    void add(FemSimulation &sim)
    {
        check(ems_batch_add(b_, sim.handle()));
        mem_.push_back(&sim);
    }
    void setTiling(int rows, int planes) { check(ems_batch_set_tiling(b_, rows, planes)); }
    std::array<int, 2> autotune(int trial_steps = 5)
    {
        int r = 0, p = 0;
        check(ems_batch_autotune(b_, trial_steps, &r, &p));
        return {r, p};
    }
    void setGraph(int steps_per_graph) { check(ems_batch_set_graph(b_, steps_per_graph)); }
    void setGradientGroup(int shots) { check(ems_batch_set_gradient_group(b_, shots)); }
    void run(int num_steps) { check(ems_batch_run(b_, num_steps)); }

    // Sum of the members' misfits (each in *misfits when given) and the
    // gradient of the sum; d_obs[q] as for FemSimulation::gradient.
    double gradient(int num_steps, const std::vector<std::vector<double>> &d_obs, std::vector<double> *misfits,
                    std::vector<double> *grad_eps_r, std::vector<double> *grad_sigma)
    {
        if (d_obs.size() != mem_.size()) throw std::runtime_error("libems: one d_obs per batch member");
        std::vector<const double *> ptrs;
        for (auto &d : d_obs) ptrs.push_back(d.data());
        const size_t nm = mem_.empty() ? 0 : mem_[0]->materialSize();
        if (misfits) misfits->assign(mem_.size(), 0.0);
        if (grad_eps_r) grad_eps_r->assign(nm, 0.0);
        if (grad_sigma) grad_sigma->assign(nm, 0.0);
        double J = 0.0;
        check(ems_batch_gradient(b_, num_steps, ptrs.data(), misfits ? misfits->data() : nullptr, &J,
                                 grad_eps_r ? grad_eps_r->data() : nullptr,
                                 grad_sigma ? grad_sigma->data() : nullptr));
        return J;
    }

private:
    ems_batch *b_ = nullptr;
    std::vector<FemSimulation *> mem_;
};

// check_params (src/fdtd/fdtd_simulation.cpp:333-366)
inline bool check_params(const std::array<double, 3> &domain_sizes, const std::array<int, 3> &grid_num,
                         double duration, double time_step, double domain_size, double c)
{
    return ems_fdtd_check_params(domain_sizes.data(), grid_num.data(), duration, time_step, domain_size, c) != 0;
}

}  // namespace fdtd
}  // namespace ems
