// C++ drop-in for the reference first-order stabilized hex FEM classes:
//   ems::fem::FemSimulation      src/fem/fem_simulation.hpp:10-40
//   ems::fem_abm::FemSimulation  src/fem_abm/fem_simulation.hpp:10-40
//
// Same constructor arguments and methods.  setSourceDipole_y / _z are
// declared but never defined in the reference (src/fem/fem_simulation.hpp:
// 27-28); they are not offered here.  Link with libems.so.
#pragma once

#include <array>
#include <iostream>
#include <string>
#include <vector>

#include "ems.h"
#include "fdtd_simulation.hpp"  // ems::check

namespace ems {
namespace detail {

class Fem1Base {
public:
    ~Fem1Base() { ems_fem1_destroy(h_); }
    Fem1Base(const Fem1Base &) = delete;
    Fem1Base &operator=(const Fem1Base &) = delete;

    void run(int num_steps) { check(ems_fem1_run(h_, num_steps)); }
    void setSource_x(const std::array<int, 3> &p) { check(ems_fem1_add_source(h_, 0, p[0], p[1], p[2])); }
    void setSource_y(const std::array<int, 3> &p) { check(ems_fem1_add_source(h_, 1, p[0], p[1], p[2])); }
    void setSource_z(const std::array<int, 3> &p) { check(ems_fem1_add_source(h_, 2, p[0], p[1], p[2])); }
    void setSourceDipole_x(const std::array<int, 3> &p) { check(ems_fem1_add_dipole_x(h_, p[0], p[1], p[2])); }
    // Not in the reference: EMS_FEM_ELEM_ROWS gives the reference's sums bit
    // for bit (the default merged stencil sums in another order; see
    // ems_fem1_set_element_product).
    void setElementProduct(int kind) { check(ems_fem1_set_element_product(h_, kind)); }
    void setObservationPoint(const std::array<std::vector<std::array<int, 3>>, 3> &position)
    {
        const int counts[3] = {(int)position[0].size(), (int)position[1].size(), (int)position[2].size()};
        std::vector<int> pts;
        for (const auto &l : position)
            for (const auto &p : l) pts.insert(pts.end(), p.begin(), p.end());
        check(ems_fem1_add_observation(h_, counts, pts.empty() ? nullptr : pts.data()));
    }
    // saved_electric_field_: [observation][slot] = {Ex, Ey, Ez}
    std::vector<std::vector<std::array<double, 3>>> savedElectricField() const
    {
        int nobs = 0, nslot = 0;
        check(ems_fem1_trace_shape(h_, &nobs, &nslot));
        std::vector<double> flat((size_t)nobs * nslot * 3);
        check(ems_fem1_get_traces(h_, flat.data(), flat.size()));
        std::vector<std::vector<std::array<double, 3>>> out(nobs, std::vector<std::array<double, 3>>(nslot));
        for (int o = 0; o < nobs; ++o)
            for (int s = 0; s < nslot; ++s)
                for (int c = 0; c < 3; ++c) out[o][s][c] = flat[((size_t)o * nslot + s) * 3 + c];
        return out;
    }
    ems_fem1 *handle() const { return h_; }

protected:
    Fem1Base(const std::array<int, 3> &g, double h, double dt, double eps, double mu, int tf, int stepper,
             double stab, int use_ofem, int precision)
    {
        ems_fem1_desc d{g[0], g[1], g[2], h, dt, eps, mu, tf, stepper, stab, use_ofem, precision};
        check(ems_fem1_create(&d, &h_));
    }
    void save(int num_steps, const std::string &path)
    {
        if (ems_fem1_save_results(h_, num_steps, path.c_str()) != EMS_OK) std::cerr << ems_last_error() << std::endl;
    }
    ems_fem1 *h_ = nullptr;
};

}  // namespace detail

namespace fem {
class FemSimulation : public detail::Fem1Base {
public:
    FemSimulation(const std::array<int, 3> &grid_size, double domain_size, double time_step, double permittivity,
                  double permeability, int time_frequency, double stabilization_factor, int use_ofem,
                  int precision = EMS_F64)
        : Fem1Base(grid_size, domain_size, time_step, permittivity, permeability, time_frequency, EMS_FEM1_LEAPFROG,
                   stabilization_factor, use_ofem, precision)
    {
    }
    void saveResults(int num_steps, const std::string &filename) { save(num_steps, filename); }
};
}  // namespace fem

namespace fem_abm {
class FemSimulation : public detail::Fem1Base {
public:
    // src/fem_abm/fem_simulation.cpp:29: stabilization_factor = 0.05
    FemSimulation(const std::array<int, 3> &grid_size, double domain_size, double time_step, double permittivity,
                  double permeability, int time_frequency, int precision = EMS_F64)
        : Fem1Base(grid_size, domain_size, time_step, permittivity, permeability, time_frequency, EMS_FEM1_ABM, 0.05,
                   0, precision)
    {
    }
    void saveResults(int num_steps) { save(num_steps, ""); }
};
}  // namespace fem_abm

}  // namespace ems
