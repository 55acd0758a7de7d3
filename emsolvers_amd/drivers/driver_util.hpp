// Helpers shared by the reference-compatible command-line drivers.
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include <ems/ems.h>

namespace ems {
namespace driver {

// static_cast<int>(std::round(v)), the reference's index rule (half away from zero)
inline int ref_round(double v) { return static_cast<int>(std::round(v)); }

// Number for file names: scientific with 6 digits, mantissa without trailing
// zeros (and without a bare '.'), exponent dropped when it is e+00
// (behaviour of compress, src/fem_v2/fem_simulation.cpp:1446-1475).
inline std::string compress(double value)
{
    std::ostringstream os;
    os << std::scientific << std::setprecision(6) << value;
    const std::string s = os.str();
    const size_t e = s.find_first_of("eE");
    if (e == std::string::npos) return s;
    std::string mant = s.substr(0, e), expo = s.substr(e);
    if (expo == "e+00" || expo == "E+00") expo.clear();
    mant.erase(mant.find_last_not_of('0') + 1);
    if (!mant.empty() && mant.back() == '.') mant.pop_back();
    return mant + expo;
}

// Stabilization factor in the fem file name: fixed with 3 decimals, trailing
// zeros and a bare '.' removed (src/fem/main.cpp:26-33).
inline std::string stab_string(double v)
{
    std::ostringstream os;
    os << std::fixed << std::setprecision(3) << v;
    std::string s = os.str();
    s.erase(s.find_last_not_of('0') + 1);
    if (!s.empty() && s.back() == '.') s.pop_back();
    return s;
}

// check_params of every reference solver (src/fdtd/fdtd_simulation.cpp:333-366,
// src/fem/fem_simulation.cpp:934-967, src/fem_v2/fem_simulation.cpp:1411-1444
// are the same test), with the reference's messages.
inline bool check_params(const double domain_sizes[3], const int grid_num[3], double duration, double time_step,
                         double domain_size, double c)
{
    if (ems_fdtd_check_params(domain_sizes, grid_num, duration, time_step, domain_size, c)) return true;
    if (domain_sizes[0] > 0 && domain_sizes[1] > 0 && domain_sizes[2] > 0 && grid_num[0] > 0 && grid_num[1] > 0 &&
        grid_num[2] > 0 && duration > 0 && time_step > 0) {
        if (std::round(domain_sizes[0] / grid_num[0]) != std::round(domain_sizes[1] / grid_num[1]) ||
            std::round(domain_sizes[0] / grid_num[0]) != std::round(domain_sizes[2] / grid_num[2])) {
            std::cerr << "Invalid grid size" << std::endl;
        } else {
            std::cerr << "Invalid time step" << std::endl;
            std::cerr << "c * time_step: " << c * time_step << std::endl;
            std::cerr << "domain_size: " << domain_size << std::endl;
        }
    }
    return false;
}

// Observation sets of the fdtd / fem / fem_abm mains around point index
// (i, j, k) = round(position / h): Ex at (i-1, j, k); Ey at the four
// (i-1|i, j-1|j, k); Ez at the four (i-1|i, j, k-1|k)
// (src/fdtd/main.cpp:32-61, src/fem/main.cpp:67-96).
template <typename Sets> inline Sets observation_sets(int i, int j, int k)
{
    Sets s;
    s[0] = {{i - 1, j, k}};
    s[1] = {{i - 1, j - 1, k}, {i - 1, j, k}, {i, j - 1, k}, {i, j, k}};
    s[2] = {{i - 1, j, k - 1}, {i - 1, j, k}, {i, j, k - 1}, {i, j, k}};
    return s;
}

inline double wall_seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline void print_device()
{
    std::cout << "libems " << ems_version() << ", HIP devices = " << ems_device_count() << " (using device 0)"
              << std::endl;
}

}  // namespace driver
}  // namespace ems
