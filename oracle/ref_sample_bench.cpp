// ref_sample_bench.cpp -- CPU baseline of the sample scheme: the reference's
// own RK3_step (sample/weno5_rk3.cpp:148-188) on its Bunsen-flame case
// (sample/main.cpp:89-141) for a given grid and step count, OpenMP threads as
// set by OMP_NUM_THREADS.  Prints one JSON line.  TEST INFRASTRUCTURE ONLY.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "field.hpp"
#include "weno5_rk3.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s N STEPS [FREQ]\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]), steps = std::atoi(argv[2]);
    const int freq = argc > 3 ? std::atoi(argv[3]) : 10;
    const double L = 0.02, dx = L / (n - 1), dy = L / (n - 1), dt = 0.0000025, SL = 0.4;
    const double u0 = 0.0, v0 = 0.8, eu = 0.05, ev = 0.05;
    Field2D G(n, n), u(n, n), v(n, n), G0(n, n);
    const double R = 0.005, slope = std::tan((90.0 - 30.0) * M_PI / 180.0), X0 = 0.01, Y0 = 0.0;
    for (int j = 0; j < n + 6; ++j)
        for (int i = 0; i < n + 6; ++i) {
            double x = (i - 3) * dx, y = (j - 3) * dy;
            double d = x < X0 ? slope * (x - (X0 - R)) - (y - Y0) : -slope * (x - (X0 + R)) - (y - Y0);
            G.data[j][i] = -d;
            G0.data[j][i] = -d;
        }
    double t = 0.0;
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) {
        G.apply_anchored_bc(G0);
        G.apply_periodic();
        for (int j = 0; j < n + 6; ++j)
            for (int i = 0; i < n + 6; ++i) {
                u.data[j][i] = u0 + eu * v0 * cos(freq * 2 * M_PI * (t));
                v.data[j][i] = v0 + ev * v0 * cos(freq * 2 * M_PI * (t));
            }
        RK3_step(G, u, v, dx, dy, dt, SL, G0);
        G.apply_periodic();
        G.apply_anchored_bc(G0);
        t += dt;
    }
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double chk = 0.0;
    for (int j = 3; j < n + 3; ++j)
        for (int i = 3; i < n + 3; ++i) chk += G.data[j][i];
    std::printf("{\"n\": %d, \"steps\": %d, \"threads\": %d, \"seconds\": %.6f, \"cell_updates_per_s\": %.6e, "
                "\"checksum\": %.17g}\n",
                n, steps, omp_get_max_threads(), sec, (double)n * n * steps / sec, chk);
    return 0;
}
