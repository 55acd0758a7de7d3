// Driver for the compiled reference FDTD class (src/fdtd/fdtd_simulation.cpp).
//
// TEST INFRASTRUCTURE ONLY: it is linked with the reference sources where
// they lie (see oracle/Makefile) and used to write golden vectors and to time
// the reference CPU path.  The reference class is compiled with
// -Dprivate=public so that the full-precision traces
// (saved_electric_field_) and the six field vectors can be read; the CSV
// writer of the reference keeps only 6 significant digits.
//
// stdin:  nx ny nz / h dt eps mu tf nsteps / nsrc, then "comp i j k" lines /
//         nobs, then per observation "cx cy cz" and cx+cy+cz "i j k" lines
// argv[1]: output file (raw little-endian): int32 nobs, nslot; double traces
//          [nobs][nslot][3]; then Ex, Ey, Ez, Hx, Hy, Hz in reference layout.
//          "-" skips the output (timing only).
// stdout: "run_seconds <t>"
#include <omp.h>

#include <cstdio>
#include <iostream>

#include "fdtd_simulation.hpp"

int main(int argc, char **argv)
{
    std::array<int, 3> n;
    double h, dt, eps, mu;
    int tf, nsteps, nsrc, nobs;
    if (!(std::cin >> n[0] >> n[1] >> n[2] >> h >> dt >> eps >> mu >> tf >> nsteps >> nsrc)) return 2;
    FemSimulation sim(n, h, dt, eps, mu, tf);
    for (int s = 0; s < nsrc; ++s) {
        int c;
        std::array<int, 3> p;
        std::cin >> c >> p[0] >> p[1] >> p[2];
        if (c == 0) sim.setSource_x(p);
        else if (c == 1) sim.setSource_y(p);
        else sim.setSource_z(p);
    }
    std::cin >> nobs;
    for (int o = 0; o < nobs; ++o) {
        int cnt[3];
        std::cin >> cnt[0] >> cnt[1] >> cnt[2];
        std::array<std::vector<std::array<int, 3>>, 3> pts;
        for (int c = 0; c < 3; ++c)
            for (int q = 0; q < cnt[c]; ++q) {
                std::array<int, 3> p;
                std::cin >> p[0] >> p[1] >> p[2];
                pts[c].push_back(p);
            }
        sim.setObservationPoint(pts);
    }
    double t0 = omp_get_wtime();
    sim.run(nsteps);
    double t1 = omp_get_wtime();
    std::printf("run_seconds %.9f\n", t1 - t0);
    if (argc > 1 && std::string(argv[1]) != "-") {
        FILE *f = std::fopen(argv[1], "wb");
        if (!f) return 3;
        int hdr[2] = {(int)sim.saved_electric_field_.size(),
                      sim.saved_electric_field_.empty() ? 0 : (int)sim.saved_electric_field_[0].size()};
        std::fwrite(hdr, sizeof(int), 2, f);
        for (auto &obs : sim.saved_electric_field_)
            for (auto &v : obs) std::fwrite(v.data(), sizeof(double), 3, f);
        const std::vector<double> *fl[6] = {&sim.electric_field_x_, &sim.electric_field_y_,
                                            &sim.electric_field_z_, &sim.magnetic_field_x_,
                                            &sim.magnetic_field_y_, &sim.magnetic_field_z_};
        for (auto *v : fl) std::fwrite(v->data(), sizeof(double), v->size(), f);
        std::fclose(f);
    }
    return 0;
}
