// Command-line driver with the argv and parameters of the reference
// stabilized-FEM main (src/fem/main.cpp:9-134):
//
//   ems_fem [<n> <stabilization_factor> <use_ofem> [<base>]]
//
// without arguments n = 1, stabilization 0.3, use_ofem 0.  4 m cube of
// base*n cells per axis (base = 200 as in the reference), 2*base*n steps
// over 8 ns, receivers recorded every n steps, Hertz x-dipole at 1.5 m,
// receivers around 2.5 m.  Writes fem_<stab>_<n>.csv or ofem_<stab>_<n>.csv.
// The optional <base> (not in the reference) shrinks the same set-up.
#include <array>
#include <string>
#include <vector>

#include <ems/fem1_simulation.hpp>

#include "driver_util.hpp"

int main(int argc, char **argv)
{
    using namespace ems::driver;
    int n = 1, use_ofem = 0, base = 200;
    double stab = 0.3;
    if (argc >= 4) {
        n = std::stoi(argv[1]);
        stab = std::stod(argv[2]);
        use_ofem = std::stoi(argv[3]);
        if (argc >= 5) base = std::stoi(argv[4]);
    }
    if (n <= 0 || base <= 0) {
        std::cerr << "n and base must be positive" << std::endl;
        return 1;
    }
    const std::string filename = (use_ofem ? "ofem_" : "fem_") + stab_string(stab) + "_" + std::to_string(n) + ".csv";
    print_device();
    const double domain_sizes[3] = {4.0, 4.0, 4.0};
    const int g = base * n;
    const int grid[3] = {g, g, g};
    const double duration = 0.8e-8;
    const int num_steps = 2 * base * n;
    const double eps = 8.854187817e-12 * 1.0, mu = 1.25663706212e-6 * 1.0;
    const std::array<double, 3> src_pos = {1.50, 1.50, 1.50}, obs_pos = {2.50, 2.50, 2.50};
    const int time_frequency = n;
    const double h = domain_sizes[0] / grid[0];
    const double dt = duration / num_steps;
    const std::array<int, 3> dipole = {ref_round(src_pos[0] / h) - 1, ref_round(src_pos[1] / h) - 1,
                                       ref_round(src_pos[2] / h) - 1};
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
        ems::fem::FemSimulation simulation({g, g, g}, h, dt, eps, mu, time_frequency, stab, use_ofem);
        std::cout << "Simulation object created in " << wall_seconds() - t0 << " seconds" << std::endl;
        simulation.setSourceDipole_x(dipole);
        simulation.setObservationPoint(obs);
        std::cout << "Starting simulation..." << std::endl;
        t0 = wall_seconds();
        simulation.run(num_steps);
        std::cout << "Simulation completed in " << wall_seconds() - t0 << " seconds" << std::endl;
        simulation.saveResults(num_steps, filename);
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    std::cout << "Simulation completed. Results saved to " << filename << std::endl;
    return 0;
}
