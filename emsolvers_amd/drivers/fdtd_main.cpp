// Command-line driver with the parameters of the reference FDTD main
// (src/fdtd/main.cpp:12-96): 1 m cube, 500^3 cells, 780 steps over 3 ns,
// x source at 0.49 m, receivers around 0.51 m; writes
// results_<i>_<dt>_<h>_<N>.csv into the current directory.
//
//   ems_fdtd          the reference run
//   ems_fdtd <n>      the same set-up on n^3 cells and round(780 n / 500) steps
#include <array>
#include <string>
#include <vector>

#include <ems/fdtd_simulation.hpp>

#include "driver_util.hpp"

int main(int argc, char **argv)
{
    using namespace ems::driver;
    const int n = argc > 1 ? std::stoi(argv[1]) : 500;
    if (n <= 0) {
        std::cerr << "grid number must be positive" << std::endl;
        return 1;
    }
    print_device();
    const double domain_sizes[3] = {1.0, 1.0, 1.0};
    const int grid[3] = {n, n, n};
    const double duration = 3.0e-9;
    const int num_steps = n == 500 ? 780 : static_cast<int>(std::lround(780.0 * n / 500.0));
    const double eps = 8.854187817e-12 * 1.0, mu = 1.25663706212e-6 * 1.0;
    const std::array<double, 3> src_pos = {0.49, 0.49, 0.49}, obs_pos = {0.51, 0.51, 0.51};
    const int time_frequency = 1;
    const double h = domain_sizes[0] / grid[0];
    const double dt = duration / num_steps;
    const std::array<int, 3> source = {ref_round(src_pos[0] / h) - 1, ref_round(src_pos[1] / h),
                                       ref_round(src_pos[2] / h)};
    using Sets = std::array<std::vector<std::array<int, 3>>, 3>;
    const Sets obs = observation_sets<Sets>(ref_round(obs_pos[0] / h), ref_round(obs_pos[1] / h),
                                            ref_round(obs_pos[2] / h));
    const double c = 1.0 / std::sqrt(eps * mu);
    if (!check_params(domain_sizes, grid, duration, dt, h, c)) {
        std::cerr << "Invalid parameters" << std::endl;
        return 1;
    }
    try {
        double t0 = wall_seconds();
        ems::fdtd::FemSimulation simulation({n, n, n}, h, dt, eps, mu, time_frequency);
        std::cout << "Simulation object created in " << wall_seconds() - t0 << " seconds" << std::endl;
        simulation.setSource_x(source);
        simulation.setObservationPoint(obs);
        std::cout << "Starting simulation..." << std::endl;
        t0 = wall_seconds();
        simulation.run(num_steps);
        std::cout << "Simulation completed in " << wall_seconds() - t0 << " seconds" << std::endl;
        simulation.saveResults(num_steps, c);
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    std::cout << "Simulation completed. Results saved to results.csv" << std::endl;
    return 0;
}
