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
    ems_fdtd *handle() const { return h_; }

private:
    ems_fdtd *h_ = nullptr;
};

// check_params (src/fdtd/fdtd_simulation.cpp:333-366)
inline bool check_params(const std::array<double, 3> &domain_sizes, const std::array<int, 3> &grid_num,
                         double duration, double time_step, double domain_size, double c)
{
    return ems_fdtd_check_params(domain_sizes.data(), grid_num.data(), duration, time_step, domain_size, c) != 0;
}

}  // namespace fdtd
}  // namespace ems
