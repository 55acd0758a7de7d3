// Command-line driver with the argv of the reference order-p hex FEM mains
// (src/fem_v2/main.cpp:11-263; src/symplectic_2/main.cpp and
// src/symplectic_4/main.cpp are the same program over their classes):
//
//   <prog> <scale> <use_ofem> <order> <domain_size_x> <domain_size_y>
//          <domain_size_z> <base_num> <duration> <source_position_x>
//          <source_position_y> <source_position_z> <observation_position_x>
//          <observation_position_y> <observation_position_z> <dim_x> <dim_y> <dim_z>
//
// One GPU owns the whole domain, so the MPI topology must be 1 1 1.  The
// element operator is read from kmat/<order>.csv (use_ofem = 1) or
// kmat/lamped<order>.csv relative to the current directory, as the
// reference does.  Writes (o)fem<order>_<scale>_<compressed parameters>.csv.
//
// Built three times: EMS_HEX_STEPPER = 0 (fem_v2 leapfrog), 1 (symplectic_2),
// 2 (symplectic_4).
#include <array>
#include <string>

#include <ems/fem_simulation.hpp>

#include "driver_util.hpp"

#ifndef EMS_HEX_STEPPER
#define EMS_HEX_STEPPER 0
#endif

#if EMS_HEX_STEPPER == 0
using Simulation = ems::fem_v2::FemSimulation;
#elif EMS_HEX_STEPPER == 1
using Simulation = ems::symplectic_2::FemSimulation;
#else
using Simulation = ems::symplectic_4::FemSimulation;
#endif

int main(int argc, char **argv)
{
    using namespace ems::driver;
    if (argc != 18) {
        std::cout << "Usage: " << argv[0]
                  << " <scale> <use_ofem> <order> <domain_size_x> <domain_size_y> <domain_size_z> <base_num>"
                     " <duration> <source_position_x> <source_position_y> <source_position_z>"
                     " <observation_position_x> <observation_position_y> <observation_position_z>"
                     " <dim_x> <dim_y> <dim_z>"
                  << std::endl;
        std::cout << "  scale: Grid size multiplier (e.g., 1 for 100x100x100)" << std::endl;
        std::cout << "  use_ofem: Use OFEM (1) or standard FEM (0)" << std::endl;
        std::cout << "  order: element order" << std::endl;
        std::cout << "  domain_size_x, domain_size_y, domain_size_z: Size of the domain in each dimension" << std::endl;
        std::cout << "  base_num: Base grid number" << std::endl;
        std::cout << "  duration: Simulation duration" << std::endl;
        std::cout << "  source_position_x, source_position_y, source_position_z: Source position coordinates"
                  << std::endl;
        std::cout << "  observation_position_x, observation_position_y, observation_position_z: Observation "
                     "position coordinates"
                  << std::endl;
        std::cout << "  dim_x, dim_y, dim_z: process grid (one GPU: 1 1 1)" << std::endl;
        return 1;
    }
    const int scale = std::stoi(argv[1]);
    const int use_ofem = std::stoi(argv[2]);
    const int order = std::stoi(argv[3]);
    const double dsx = std::stod(argv[4]), dsy = std::stod(argv[5]), dsz = std::stod(argv[6]);
    const int base_num = std::stoi(argv[7]);
    const double duration = std::stod(argv[8]);
    const double spx = std::stod(argv[9]), spy = std::stod(argv[10]), spz = std::stod(argv[11]);
    const double opx = std::stod(argv[12]), opy = std::stod(argv[13]), opz = std::stod(argv[14]);
    const std::array<int, 3> dims = {std::stoi(argv[15]), std::stoi(argv[16]), std::stoi(argv[17])};

    const double domain_sizes[3] = {dsx, dsy, dsz};
    const int grid[3] = {scale * base_num, scale * base_num, scale * base_num};
    const double h = domain_sizes[0] / grid[0];
    const std::array<int, 3> source = {ref_round(spx / h), ref_round(spy / h), ref_round(spz / h)};
    const std::array<int, 3> observation = {ref_round(opx / h), ref_round(opy / h), ref_round(opz / h)};
    // constants of this main (src/fem_v2/main.cpp:33-34)
    const double eps = 1.0 * 8.8541878128e-12, mu = 1.0 * 1.25663706212e-6;
    const double c = 1.0 / std::sqrt(eps * mu);
    const double dt = h / (c * std::sqrt(3.0)) / 2.0;
    const int time_frequency = scale * 2;
    const int num_steps = static_cast<int>(std::floor(duration / dt) + 1);
    const std::string filename = (use_ofem ? "ofem" : "fem") + std::to_string(order) + "_" + std::to_string(scale) +
                                 "_" + compress(dsx) + "_" + compress(dsy) + "_" + compress(dsz) + "_" +
                                 compress(duration) + "_" + compress(spx) + "_" + compress(spy) + "_" +
                                 compress(spz) + "_" + compress(opx) + "_" + compress(opy) + "_" + compress(opz) +
                                 ".csv";
    if (!check_params(domain_sizes, grid, duration, dt, h, c)) {
        std::cerr << "Invalid parameters" << std::endl;
        return 1;
    }
    if (dims[0] * dims[1] * dims[2] != 1) {
        std::cerr << "one GPU owns the whole domain: the process grid must be 1 1 1" << std::endl;
        return 1;
    }
    print_device();
    try {
        double t0 = wall_seconds();
        Simulation simulation(order, {grid[0], grid[1], grid[2]}, h, dt, eps, mu, time_frequency, use_ofem);
        std::cout << "rank: 0 Simulation object created in " << wall_seconds() - t0 << " seconds" << std::endl;
        simulation.setSource_z(source);
        simulation.setObservationPoint(observation);
        std::cout << "rank: 0 Starting simulation..." << std::endl;
        t0 = wall_seconds();
        simulation.run(num_steps);
        std::cout << "rank: 0 Simulation completed in " << wall_seconds() - t0 << " seconds" << std::endl;
        simulation.saveResults(num_steps, filename);
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    std::cout << "rank: 0 Simulation completed." << std::endl;
    return 0;
}
