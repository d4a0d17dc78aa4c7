// ref_bench.cpp -- times the reference's own 3D CPU solver under MPI.
//
// TEST/BENCH INFRASTRUCTURE ONLY (the CPU baseline of bench.py).  Compiled by
// oracle/Makefile against Local_code/level-set_MPI_3D/include where it lies;
// no numerics of its own.  Each rank owns a y-slab from the reference's
// setupMPIDomainDecomposition (config.h:242-265) and the timed loop follows
// runStandaloneSimulation3D (src/main.cpp:285-340): rk3TimeStep3D, pointer
// swap, reinitializeWithSwap3D every `ri` steps after step 0, max-over-ranks
// MPI_Wtime.  Sizes are runtime values (the reference functions read them
// from SimParams).  The velocity and G are the plane-with-modes workload:
// G = x - 0.3 - sum a sin(2 pi (ky y + kz z) + ph), u = (1,0,0) + solenoidal
// modes (fixed table below); the timing does not depend on the exact values.
//
// usage: mpirun -np P ref_bench3d NX NY NZ STEPS RI [S_L]
#include <mpi.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "config.h"
#include "weno5.h"
#include "boundary.h"
#include "rk3.h"
#include "reinitialization.h"

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc < 6) {
        if (rank == 0) std::fprintf(stderr, "usage: ref_bench3d NX NY NZ STEPS RI [S_L]\n");
        MPI_Finalize();
        return 2;
    }
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]);
    const int steps = std::atoi(argv[4]), ri = std::atoi(argv[5]);
    const double s_l = argc > 6 ? std::atof(argv[6]) : 0.08;

    SimParams p;
    std::memset(&p, 0, sizeof(p));
    p.nx = nx; p.ny = ny; p.nz = nz; p.nghost = 3;
    p.nx_total = nx + 6; p.ny_total = ny + 6; p.nz_total = nz + 6;
    p.dx = 1.0 / (nx - 1); p.dy = 1.0 / (ny - 1); p.dz = 1.0 / (nz - 1);
    p.local_ny = ny; p.local_ny_total = ny + 6;
    setupMPIDomainDecomposition(p, rank, size);
    p.s_l = s_l;
    p.epsilon = 1.0e-10;
    p.enable_reinit = ri > 0;
    p.reinit_interval = ri > 0 ? ri : 1;
    p.reinit_iterations = 2;
    p.reinit_dtau = 0.25 * std::fmin(p.dx, std::fmin(p.dy, p.dz));
    p.reinit_beta = 0.5;
    p.dt = 0.5 * p.dx / (1.0 + 0.4 + s_l);

    const long long n = (long long)p.nx_total * p.local_ny_total * p.nz_total;
    double* G = new double[n]();
    double* Gn = new double[n]();
    double* G1 = new double[n]();
    double* G2 = new double[n]();
    double* R = new double[n]();
    double* u = new double[n]();
    double* v = new double[n]();
    double* w = new double[n]();
    double* tmp = new double[n]();
    double* p0 = new double[n]();
    double* rt = new double[n]();
    int* fl = new int[n]();

    const double PI = 3.14159265358979323846;
    const double ga[4] = {0.012, 0.007, 0.015, 0.004};
    const int gk[4][2] = {{1, 2}, {3, 1}, {2, 5}, {7, 4}};
    for (int k = 0; k < p.nz_total; k++)
        for (int j = 0; j < p.local_ny_total; j++)
            for (int i = 0; i < p.nx_total; i++) {
                double x, y, z;
                indexToCoord3D(i, j, k, 0.0, 0.0, 0.0, p.dx, p.dy, p.dz, 3, p.y_start, x, y, z);
                double eta = 0.0;
                for (int m = 0; m < 4; m++) eta += ga[m] * std::sin(2.0 * PI * (gk[m][0] * y + gk[m][1] * z) + m);
                long long c = idx3(i, j, k, p.nx_total, p.local_ny_total);
                G[c] = x - 0.3 - eta;
                const double s1 = 0.1 * std::sin(2.0 * PI * (y + 2.0 * z));
                const double s2 = 0.1 * std::sin(2.0 * PI * (2.0 * x - z) + 1.0);
                u[c] = 1.0 + s2 * (1.0 / std::sqrt(5.0));
                v[c] = s1;
                w[c] = s2 * (2.0 / std::sqrt(5.0));
            }
    applyBoundaryConditions3D(G, p);

    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = MPI_Wtime();
    for (int step = 0; step < steps; step++) {
        rk3TimeStep3D(G, Gn, G1, G2, R, u, v, w, p);
        double* t = G; G = Gn; Gn = t;
        if (p.enable_reinit && step > 0 && step % p.reinit_interval == 0)
            reinitializeWithSwap3D(&G, &tmp, p0, rt, fl, p);
    }
    const double el = MPI_Wtime() - t0;
    double mx = 0.0;
    MPI_Reduce(&el, &mx, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    if (rank == 0) {
        const double cells = (double)nx * ny * nz;
        std::printf("{\"ranks\": %d, \"grid\": [%d, %d, %d], \"steps\": %d, \"reinit_interval\": %d, "
                    "\"seconds\": %.6f, \"cell_updates_per_s\": %.6e}\n",
                    size, nx, ny, nz, steps, ri, mx, cells * steps / mx);
    }
    MPI_Finalize();
    return 0;
}
