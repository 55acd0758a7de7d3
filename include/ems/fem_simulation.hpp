// C++ drop-in for the reference order-p hex FEM classes:
//   ems::fem_v2::FemSimulation        src/fem_v2/fem_simulation.hpp:10-60
//   ems::symplectic_2::FemSimulation  src/symplectic_2/fem_simulation.hpp
//   ems::symplectic_4::FemSimulation  src/symplectic_4/fem_simulation.hpp
//
// Same constructor arguments.  The reference constructor also takes the MPI
// topology (dims, coords, rank, 18 communicators, 18 ranks); these are
// accepted and ignored when dims = {1, 1, 1}: one GPU owns the whole domain
// and there are no interface sums.  Like the reference, the constructor reads
// kmat/<order>.csv (use_ofem = 1) or kmat/lamped<order>.csv (use_ofem = 0)
// relative to the current directory and throws std::runtime_error when it
// cannot (src/fem_v2/fem_simulation.cpp:1500).  Link with libems.so.
#pragma once

#include <array>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ems.h"
#include "fdtd_simulation.hpp"  // ems::check

namespace ems {
namespace detail {

template <int STEPPER> class HexFem {
public:
    HexFem(int order, const std::array<int, 3> &grid_size, double domain_size, double time_step,
           double permittivity, double permeability, int time_frequency, int use_ofem,
           int precision = EMS_F64)
    {
        init(order, grid_size, domain_size, time_step, permittivity, permeability, time_frequency, use_ofem,
             precision);
    }
    // Reference signature with the MPI topology; only dims = {1,1,1} is served.
    template <typename... MpiArgs>
    HexFem(int order, const std::array<int, 3> &grid_size, double domain_size, double time_step,
           double permittivity, double permeability, int time_frequency, int use_ofem,
           const std::array<int, 3> &dims, const std::array<int, 3> &coords, int rank, MpiArgs &&...)
    {
        (void)coords;
        (void)rank;
        if (dims[0] * dims[1] * dims[2] != 1)
            throw std::runtime_error("libems: one GPU owns the whole domain; use dims = {1, 1, 1}");
        init(order, grid_size, domain_size, time_step, permittivity, permeability, time_frequency, use_ofem,
             EMS_F64);
    }
    ~HexFem() { ems_fem_destroy(h_); }
    HexFem(const HexFem &) = delete;
    HexFem &operator=(const HexFem &) = delete;

    void run(int num_steps) { check(ems_fem_run(h_, num_steps)); }
    void setSource_x(const std::array<int, 3> &p) { check(ems_fem_add_source(h_, 0, p[0], p[1], p[2])); }
    void setSource_y(const std::array<int, 3> &p) { check(ems_fem_add_source(h_, 1, p[0], p[1], p[2])); }
    void setSource_z(const std::array<int, 3> &p) { check(ems_fem_add_source(h_, 2, p[0], p[1], p[2])); }
    void setObservationPoint(const std::array<int, 3> &p) { check(ems_fem_add_observation(h_, p[0], p[1], p[2])); }
    // Not in the reference: EMS_FEM_ELEM_ROWS gives the reference's sums bit
    // for bit at every order (the defaults at orders 1 and 3 sum in another
    // order; see ems_fem_set_element_product).
    void setElementProduct(int kind) { check(ems_fem_set_element_product(h_, kind)); }
    void saveResults(int num_steps, const std::string &filename)
    {
        if (ems_fem_save_results(h_, num_steps, filename.c_str()) != EMS_OK)
            std::cerr << ems_last_error() << std::endl;
    }
    // saved_electric_field_: [observation][slot] = {time, Ex, Ey, Ez}
    std::vector<std::vector<std::array<double, 4>>> savedElectricField() const
    {
        int nobs = 0, nslot = 0;
        check(ems_fem_trace_shape(h_, &nobs, &nslot));
        std::vector<double> flat((size_t)nobs * nslot * 4);
        check(ems_fem_get_traces(h_, flat.data(), flat.size()));
        std::vector<std::vector<std::array<double, 4>>> out(nobs, std::vector<std::array<double, 4>>(nslot));
        for (int o = 0; o < nobs; ++o)
            for (int s = 0; s < nslot; ++s)
                for (int c = 0; c < 4; ++c) out[o][s][c] = flat[((size_t)o * nslot + s) * 4 + c];
        return out;
    }
    ems_fem *handle() const { return h_; }

private:
    void init(int order, const std::array<int, 3> &g, double h, double dt, double eps, double mu, int tf,
              int use_ofem, int precision)
    {
        const std::string file =
            use_ofem == 0 ? "kmat/lamped" + std::to_string(order) + ".csv" : "kmat/" + std::to_string(order) + ".csv";
        int rows = 0;
        if (ems_fem_load_kmat(file.c_str(), nullptr, 0, &rows) != EMS_OK)
            throw std::runtime_error("Cannot open file: " + file);
        std::vector<double> k((size_t)rows * rows);
        check(ems_fem_load_kmat(file.c_str(), k.data(), k.size(), nullptr));
        ems_fem_desc d{order, g[0], g[1], g[2], h, dt, eps, mu, tf, STEPPER, precision};
        check(ems_fem_create(&d, k.data(), &h_));
    }
    ems_fem *h_ = nullptr;
};

}  // namespace detail

namespace fem_v2 {
using FemSimulation = detail::HexFem<EMS_FEM_LEAPFROG>;
}
namespace symplectic_2 {
using FemSimulation = detail::HexFem<EMS_FEM_SYMPLECTIC2>;
}
namespace symplectic_4 {
using FemSimulation = detail::HexFem<EMS_FEM_SYMPLECTIC4>;
}

}  // namespace ems
