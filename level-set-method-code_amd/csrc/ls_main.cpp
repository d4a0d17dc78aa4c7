// ls_main.cpp -- command-line driver (the reference's runStandaloneSimulation
// on one MI355X).  Flags follow Local_code/level-set_MPI_2D/src/main.cpp:89-154
// and Local_code/level-set_MPI_3D/src/main.cpp:82-147; the loop follows
// main.cpp:310-354 (3D :285-334): sliver last step, velocity update at the
// start of each 3D step, RK3 step, pointer swap, HCR-2 after step > 0 with
// step % ri == 0, diagnostics every OUTPUT_INTERVAL steps.
//
// usage: ls_solve [-dim 2|3] [-n N] [-t zalesak|circle|sphere|swirl|plane]
//                 [-T t] [-cfl c] [-sl s] [-u u] [-v v] [-w w] [-reinit|-no-reinit]
//                 [-ri N] [-riter N] [-mode exact|fast] [-o dir] [-no-output] [-q]
//                 [-dump-crossings]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ls_reference_api.hpp"

static void usage(const char* prog) {
    std::printf("Usage: %s [options]\n"
                "  -dim <2|3>     dimension (default 2)\n"
                "  -n <N>         interior points per axis (default 201 in 2D, 64 in 3D)\n"
                "  -t <test>      zalesak (2D default), circle, sphere (3D default), swirl, plane\n"
                "  -T <time>      final time\n  -cfl <val>     CFL number\n"
                "  -sl <val>      laminar flame speed\n  -u/-v/-w <val> constant velocity\n"
                "  -reinit / -no-reinit, -ri <N>, -riter <N>   HCR-2 settings\n"
                "  -mode <exact|fast>  arithmetic mode (default exact)\n"
                "  -o <dir>       output directory (default ./output)\n"
                "  -no-output     no G files\n  -q             quiet\n"
                "  -dump-crossings  write phi_0, r_tilde and the crossing flags of the last\n"
                "                 reinitialisation to <dir>/{phi_0,r_tilde,flag}.raw (padded, no header)\n", prog);
}

int main(int argc, char** argv) {
    int dim = 2, n = -1, mode = LS_MODE_EXACT, ri = 10, riter = 2;
    std::string test, outdir = "./output";
    double T = -1.0, cfl = -1.0, sl = 0.0, u = 0.0, v = 0.0, w = 0.0;
    bool have_u = false, reinit = false, save = true, verbose = true, dump = false;
    for (int a = 1; a < argc; a++) {
        auto next = [&](const char* f) -> const char* {
            if (a + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", f); std::exit(1); }
            return argv[++a];
        };
        if (!std::strcmp(argv[a], "-h") || !std::strcmp(argv[a], "--help")) { usage(argv[0]); return 0; }
        else if (!std::strcmp(argv[a], "-dim")) dim = std::atoi(next("-dim"));
        else if (!std::strcmp(argv[a], "-n")) n = std::atoi(next("-n"));
        else if (!std::strcmp(argv[a], "-t")) test = next("-t");
        else if (!std::strcmp(argv[a], "-T")) T = std::atof(next("-T"));
        else if (!std::strcmp(argv[a], "-cfl")) cfl = std::atof(next("-cfl"));
        else if (!std::strcmp(argv[a], "-sl")) sl = std::atof(next("-sl"));
        else if (!std::strcmp(argv[a], "-u")) { u = std::atof(next("-u")); have_u = true; }
        else if (!std::strcmp(argv[a], "-v")) { v = std::atof(next("-v")); have_u = true; }
        else if (!std::strcmp(argv[a], "-w")) { w = std::atof(next("-w")); have_u = true; }
        else if (!std::strcmp(argv[a], "-reinit")) reinit = true;
        else if (!std::strcmp(argv[a], "-no-reinit")) reinit = false;
        else if (!std::strcmp(argv[a], "-ri")) ri = std::atoi(next("-ri"));
        else if (!std::strcmp(argv[a], "-riter")) riter = std::atoi(next("-riter"));
        else if (!std::strcmp(argv[a], "-mode")) mode = !std::strcmp(next("-mode"), "fast") ? LS_MODE_FAST : LS_MODE_EXACT;
        else if (!std::strcmp(argv[a], "-o")) outdir = next("-o");
        else if (!std::strcmp(argv[a], "-no-output")) save = false;
        else if (!std::strcmp(argv[a], "-q")) verbose = false;
        else if (!std::strcmp(argv[a], "-dump-crossings")) dump = true;
        else { std::fprintf(stderr, "Unknown option: %s\n", argv[a]); usage(argv[0]); return 1; }
    }
    if (test.empty()) test = dim == 3 ? "sphere" : "zalesak";
    if (n < 0) n = dim == 3 ? 64 : 201;

    SimParams sp = initDefaultParams(dim, n);
    sp.mode = mode;
    sp.s_l = sl;
    sp.enable_reinit = reinit;
    sp.reinit_interval = ri;
    sp.reinit_iterations = riter;
    if (T > 0.0) sp.t_final = T;
    if (cfl > 0.0) sp.cfl = cfl;
    if (have_u) { sp.u_const = u; sp.v_const = v; sp.w_const = w; }
    int case_id;
    bool deform = false;
    if (test == "zalesak") case_id = LS_CASE_ZALESAK2D;
    else if (test == "circle") case_id = LS_CASE_CIRCLE2D;
    else if (test == "sphere") { case_id = LS_CASE_SPHERE_DEFORM3D; deform = true; }
    else if (test == "swirl") case_id = LS_CASE_SPHERE_SWIRL3D;
    else if (test == "plane") case_id = LS_CASE_PLANE_MODES3D;
    else { std::fprintf(stderr, "unknown test case %s\n", test.c_str()); return 1; }
    if (dim == 3 && !deform) sp.dt = 0.0;  // CFL-based unless the deformation test's fixed DT
    if (test == "circle" && !have_u) sp.u_const = 0.5;
    if (dim == 3 && !deform && !have_u) sp.u_const = 1.0;
    sp.dt = calculateTimeStep(sp);

    ls_params_t p = lsref::to_params(sp);
    const size_t N = (size_t)ls_field_size(&p);
    const size_t bytes = N * sizeof(double);
    std::vector<double*> f(11, nullptr);
    for (auto& ptr : f) {
        if (hipMalloc((void**)&ptr, bytes) != hipSuccess || hipMemset(ptr, 0, bytes) != hipSuccess) {
            std::fprintf(stderr, "device allocation of %zu bytes failed\n", bytes);
            return 1;
        }
    }
    int* flag = nullptr;
    if (hipMalloc((void**)&flag, N * sizeof(int)) != hipSuccess || hipMemset(flag, 0, N * sizeof(int)) != hipSuccess)
        return 1;
    double *G = f[0], *G_new = f[1], *G_1 = f[2], *G_2 = f[3], *G_init = f[4], *du = f[5], *dv = f[6],
           *dw = f[7], *tmp = f[8], *p0 = f[9], *rt = f[10];
    lsref::check(ls_init_case(&p, case_id, 3, sp.t_final, G, du, dv, dim == 3 ? dw : nullptr, nullptr), "init");
    if (test == "circle" || (dim == 3 && (test == "swirl" || test == "plane") && have_u)) {
        std::vector<double> c(N);
        const double vals[3] = {sp.u_const, sp.v_const, sp.w_const};
        double* dst[3] = {du, dv, dw};
        for (int q = 0; q < dim; q++) {
            std::fill(c.begin(), c.end(), vals[q]);
            hipMemcpy(dst[q], c.data(), bytes, hipMemcpyHostToDevice);
        }
    }
    applyBoundaryConditions(G, sp);
    hipMemcpy(G_init, G, bytes, hipMemcpyDeviceToDevice);
    const double a0 = computeInterfaceArea(G, sp);
    std::printf("G-Equation Level-Set Solver %dD (MI355X, %s mode)\n", dim, mode == LS_MODE_FAST ? "fast" : "exact");
    std::printf("  Grid size: %d^%d, dx: %.6e, dt: %.6e, T: %.4f, S_L: %.4f, reinit: %s\n", n, dim, sp.dx,
                sp.dt, sp.t_final, sp.s_l, reinit ? "Enabled" : "Disabled");
    std::printf("Initial interface %s: %.6f\n", dim == 3 ? "volume" : "area", a0);
    if (save || dump) {
        std::string cmd = "mkdir -p " + outdir;
        if (std::system(cmd.c_str()) != 0) std::fprintf(stderr, "could not create %s\n", outdir.c_str());
        if (save) saveFieldBinary((outdir + "/G_initial.bin").c_str(), G, sp);
    }

    const int output_interval = dim == 3 ? 50 : 10;
    double t = 0.0;
    int step = 0;
    const int max_steps = (int)(sp.t_final / sp.dt) + 100;
    hipDeviceSynchronize();
    const auto t0 = std::chrono::steady_clock::now();
    while (t < sp.t_final && step < max_steps) {
        if (t + sp.dt > sp.t_final) sp.dt = sp.t_final - t;
        if (deform) {
            sp.current_time = t;
            updateDeformationVelocity(du, dv, dw, sp);
        }
        rk3TimeStep(G, G_new, G_1, G_2, nullptr, du, dv, sp, dim == 3 ? dw : nullptr);
        std::swap(G, G_new);
        if (sp.enable_reinit && step > 0 && step % sp.reinit_interval == 0)
            reinitializeWithSwap(&G, &tmp, p0, rt, flag, sp);
        t += sp.dt;
        step++;
        if (step % output_interval == 0 && verbose) {
            const double l2 = computeL2Error(G, G_init, sp);
            const double ar = computeInterfaceArea(G, sp);
            std::printf("Step %6d: t = %.6f, dt = %.6e, L2 error = %.6e, %s = %.6f\n", step, t, sp.dt, l2,
                        dim == 3 ? "Volume" : "Area", ar);
        }
    }
    hipDeviceSynchronize();
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double l2 = computeL2Error(G, G_init, sp);
    const double a1 = computeInterfaceArea(G, sp);
    std::printf("Final Results:\n  L2 Error: %.6e\n  Initial %s: %.6f\n  Final %s: %.6f\n", l2,
                dim == 3 ? "Volume" : "Area", a0, dim == 3 ? "Volume" : "Area", a1);
    std::printf("Timing Results:\n  Total steps: %d\n  Wall time: %.3f s\n  Time per step: %.3f ms\n", step, el,
                el / step * 1e3);
    if (save) {
        saveFieldBinary((outdir + "/G_final.bin").c_str(), G, sp);
        // the 3D driver also writes a VTK copy (Local_code/level-set_MPI_3D/src/main.cpp:377-380)
        if (dim == 3) saveFieldVTK3D((outdir + "/G_final.vtk").c_str(), G, sp);
    }
    if (dump) {
        // the arrays the shim's computeInterfaceCrossings filled at the last
        // reinitializeWithSwap, as the host sees them
        auto raw = [&](const char* name, const void* dev, size_t sz) {
            std::vector<char> h(sz);
            if (hipMemcpy(h.data(), dev, sz, hipMemcpyDeviceToHost) != hipSuccess) return false;
            FILE* fp = std::fopen((outdir + "/" + name).c_str(), "wb");
            if (!fp) return false;
            const bool ok = std::fwrite(h.data(), 1, sz, fp) == sz;
            return std::fclose(fp) == 0 && ok;
        };
        if (!raw("phi_0.raw", p0, bytes) || !raw("r_tilde.raw", rt, bytes) || !raw("flag.raw", flag, N * sizeof(int))) {
            std::fprintf(stderr, "could not write the crossing arrays to %s\n", outdir.c_str());
            return 1;
        }
    }
    for (auto ptr : f) hipFree(ptr);
    hipFree(flag);
    return 0;
}
