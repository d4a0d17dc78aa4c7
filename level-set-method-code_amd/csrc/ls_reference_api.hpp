// ls_reference_api.hpp -- the reference's host interface, on the gfx950 solver.
//
// Drop-in for C++ code written against Local_code/level-set_MPI_2D/include and
// Local_code/level-set_MPI_3D/include: same function names, argument order and
// meaning, with every field pointer a DEVICE pointer (the caller allocates
// with hipMalloc, zero-filled like the reference's new double[]()).  The
// arithmetic is the exact mode (bitwise equal to the contraction-free CPU
// build) unless params.mode selects LS_MODE_FAST.  Like the reference, errors
// print a message and exit(EXIT_FAILURE).
//
//   rk3TimeStep / rk3TimeStep3D          rk3.h:168-187 / 3D rk3.h:131-150
//   applyBoundaryConditions[3D]          boundary.h:217-229 / 3D boundary.h:250-261
//   reinitializeWithSwap[3D]             reinitialization.h:321-360 / 3D :199-228
//   computeInterfaceCrossings[3D]        reinitialization.h:75-130 / 3D :46-89
//   calculateTimeStep, initDefaultParams config.h:166-263 / 3D config.h:180-293
//   computeL2Error[3D], computeInterfaceArea, computeInterfaceVolume, saveFieldBinary[3D],
//   saveFieldVTK[3D]                     io.h:34-246 / 3D io.h:34-288
//
// Differences a caller can observe: G_rhs is not written (the RHS lives only
// in registers inside the fused stage kernel), and the single-GPU solver has
// no MPI decomposition (rank 0 of 1; y_start = 0).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ls_api.h"

// Runtime mirror of SimParams (config.h:123-161, 3D config.h:134-175) plus the
// gfx950 selections: dim, per-axis boundary kinds, arithmetic mode, stream.
struct SimParams {
    int nx, ny, nz;
    int nx_total, ny_total, nz_total;
    int nghost;
    double dx, dy, dz;
    double x_min, x_max, y_min, y_max, z_min, z_max;
    int rank, num_procs, local_ny, local_ny_total, y_start, neighbor_below, neighbor_above;
    double s_l, u_const, v_const, w_const;
    double dt, t_final, cfl, current_time;
    bool enable_reinit;
    int reinit_interval, reinit_iterations;
    double reinit_dtau, reinit_beta;
    double epsilon, interface_band;
    int dim;
    int bc[3];
    int mode;
    hipStream_t stream;
};

namespace lsref {

inline void die(const char* what) {
    std::fprintf(stderr, "level-set solver error in %s: %s\n", what, ls_last_error());
    std::exit(EXIT_FAILURE);
}

inline ls_params_t to_params(const SimParams& sp) {
    ls_params_t p;
    std::memset(&p, 0, sizeof(p));
    p.dim = sp.dim;
    p.nx = sp.nx;
    p.ny = sp.ny;
    p.nz = sp.dim == 3 ? sp.nz : 1;
    p.nghost = sp.nghost;
    for (int a = 0; a < 3; a++) p.bc[a] = sp.bc[a];
    p.mode = sp.mode;
    p.velocity = LS_VEL_ARRAYS;
    p.dx = sp.dx;
    p.dy = sp.dy;
    p.dz = sp.dim == 3 ? sp.dz : 1.0;
    p.x0 = sp.x_min;
    p.y0 = sp.y_min;
    p.z0 = sp.z_min;
    p.s_l = sp.s_l;
    p.eps = sp.epsilon;
    p.reinit_interval = sp.enable_reinit ? sp.reinit_interval : 0;
    p.reinit_iters = sp.reinit_iterations;
    p.reinit_dtau = sp.reinit_dtau;
    p.reinit_beta = sp.reinit_beta;
    p.vel_period = sp.t_final;
    return p;
}

// One solver per thread, rebuilt when the parameters change.
inline ls_solver_t* solver_for(const SimParams& sp) {
    thread_local ls_solver_t* s = nullptr;
    thread_local ls_params_t last;
    ls_params_t p = to_params(sp);
    if (s && std::memcmp(&p, &last, sizeof(p)) == 0) return s;
    if (s) ls_destroy(s);
    s = nullptr;
    if (ls_create(&p, &s) != LS_OK) die("ls_create");
    last = p;
    return s;
}

inline void check(int rc, const char* what) {
    if (rc != LS_OK) die(what);
}

}  // namespace lsref

// initDefaultParams with the grid size as a runtime argument (the reference
// takes NX/NY/NZ from constexpr constants).  2D: periodic x, zero-gradient y;
// 3D: periodic.
inline SimParams initDefaultParams(int dim = 2, int n = 201) {
    SimParams sp;
    std::memset(&sp, 0, sizeof(sp));
    ls_params_t p;
    lsref::check(ls_default_params(&p, dim, n), "initDefaultParams");
    sp.dim = dim;
    sp.nx = sp.ny = n;
    sp.nz = dim == 3 ? n : 1;
    sp.nghost = 3;
    sp.nx_total = n + 6;
    sp.ny_total = n + 6;
    sp.nz_total = dim == 3 ? n + 6 : 1;
    sp.dx = p.dx;
    sp.dy = p.dy;
    sp.dz = p.dz;
    sp.x_max = sp.y_max = sp.z_max = 1.0;
    sp.num_procs = 1;
    sp.local_ny = n;
    sp.local_ny_total = n + 6;
    sp.neighbor_below = sp.neighbor_above = -1;
    sp.cfl = 0.2;
    sp.t_final = dim == 3 ? 1.5 : 6.283185;
    sp.dt = dim == 3 ? 0.001 : 0.0;
    sp.u_const = dim == 3 ? 0.0 : 1.0;
    sp.reinit_interval = p.reinit_interval;
    sp.reinit_iterations = p.reinit_iters;
    sp.reinit_dtau = p.reinit_dtau;
    sp.reinit_beta = p.reinit_beta;
    sp.epsilon = p.eps;
    sp.interface_band = 2.0;
    for (int a = 0; a < 3; a++) sp.bc[a] = p.bc[a];
    sp.mode = LS_MODE_EXACT;
    sp.stream = nullptr;
    return sp;
}

// calculateTimeStep (2D config.h:251-263; 3D config.h:272-293 where a positive
// params.dt wins).
inline double calculateTimeStep(const SimParams& sp) {
    if (sp.dim == 3 && sp.dt > 0.0) return sp.dt;
    ls_params_t p = lsref::to_params(sp);
    return ls_cfl_dt(&p, sp.cfl, sp.u_const, sp.v_const, sp.w_const);
}

inline void applyBoundaryConditions(double* G, SimParams& sp) {
    lsref::check(ls_fill_ghosts(lsref::solver_for(sp), G, sp.stream), "applyBoundaryConditions");
}
inline void applyBoundaryConditions3D(double* G, SimParams& sp) { applyBoundaryConditions(G, sp); }

// rk3TimeStep: G_1 and G_2 receive the ghost-filled stage fields as in the
// reference; G_rhs is not used.
inline void rk3TimeStep(double* G_n, double* G_new, double* G_1, double* G_2, double* G_rhs,
                        const double* u, const double* v, SimParams& sp, const double* w = nullptr) {
    (void)G_rhs;
    ls_solver_t* s = lsref::solver_for(sp);
    lsref::check(ls_rk3_stage(s, 1, G_n, G_n, G_1, u, v, w, nullptr, sp.dt, sp.stream), "rk3 stage 1");
    lsref::check(ls_rk3_stage(s, 2, G_1, G_n, G_2, u, v, w, nullptr, sp.dt, sp.stream), "rk3 stage 2");
    lsref::check(ls_rk3_stage(s, 3, G_2, G_n, G_new, u, v, w, nullptr, sp.dt, sp.stream), "rk3 stage 3");
}
inline void rk3TimeStep3D(double* G_n, double* G_new, double* G_1, double* G_2, double* G_rhs,
                          const double* u, const double* v, const double* w, SimParams& sp) {
    rk3TimeStep(G_n, G_new, G_1, G_2, G_rhs, u, v, sp, w);
}

inline void computeInterfaceCrossings(const double* phi_0, double* r_tilde, int* flag, SimParams& sp) {
    lsref::check(ls_hcr2_crossings(lsref::solver_for(sp), phi_0, r_tilde, reinterpret_cast<int32_t*>(flag),
                                   sp.stream), "computeInterfaceCrossings");
}
inline void computeInterfaceCrossings3D(const double* phi_0, double* r_tilde, int* flag, SimParams& sp) {
    computeInterfaceCrossings(phi_0, r_tilde, flag, sp);
}

// reinitializeWithSwap: copy phi -> phi_0, ghost fill, crossings, then
// reinit_iterations steps ping-ponging phi / phi_temp; the caller's pointers
// are swapped exactly as the reference swaps them.
inline void reinitializeWithSwap(double** phi_ptr, double** phi_temp_ptr, double* phi_0, double* r_tilde,
                                 int* flag, SimParams& sp) {
    if (!sp.enable_reinit) return;
    ls_solver_t* s = lsref::solver_for(sp);
    ls_params_t p = lsref::to_params(sp);
    const size_t bytes = (size_t)ls_field_size(&p) * sizeof(double);
    double* phi = *phi_ptr;
    double* tmp = *phi_temp_ptr;
    if (hipMemcpyAsync(phi_0, phi, bytes, hipMemcpyDeviceToDevice, sp.stream) != hipSuccess)
        lsref::die("reinitializeWithSwap copy");
    lsref::check(ls_fill_ghosts(s, phi_0, sp.stream), "reinitializeWithSwap ghosts");
    computeInterfaceCrossings(phi_0, r_tilde, flag, sp);
    for (int it = 0; it < sp.reinit_iterations; it++) {
        lsref::check(ls_hcr2_step(s, phi, phi_0, tmp, sp.stream), "reinitStep");
        double* t = phi;
        phi = tmp;
        tmp = t;
    }
    *phi_ptr = phi;
    *phi_temp_ptr = tmp;
}
inline void reinitializeWithSwap3D(double** phi_ptr, double** phi_temp_ptr, double* phi_0, double* r_tilde,
                                   int* flag, SimParams& sp) {
    reinitializeWithSwap(phi_ptr, phi_temp_ptr, phi_0, r_tilde, flag, sp);
}

// updateDeformationVelocity (3D initial_conditions.h:278-280): field at
// sp.current_time with period sp.t_final.
inline void updateDeformationVelocity(double* u, double* v, double* w, SimParams& sp) {
    ls_params_t p = lsref::to_params(sp);
    lsref::check(ls_deformation_velocity(&p, sp.current_time, sp.t_final, u, v, w, sp.stream),
                 "updateDeformationVelocity");
}

inline double computeL2Error(const double* G, const double* G_ref, SimParams& sp) {
    double l2 = 0.0, area = 0.0;
    lsref::check(ls_diag(lsref::solver_for(sp), G, G_ref, &l2, &area, sp.stream), "computeL2Error");
    return l2;
}
inline double computeL2Error3D(const double* G, const double* G_ref, SimParams& sp) {
    return computeL2Error(G, G_ref, sp);
}
inline double computeInterfaceArea(const double* G, SimParams& sp) {
    double area = 0.0;
    lsref::check(ls_diag(lsref::solver_for(sp), G, nullptr, nullptr, &area, sp.stream), "computeInterfaceArea");
    return area;
}
inline double computeInterfaceVolume(const double* G, SimParams& sp) { return computeInterfaceArea(G, sp); }

// saveFieldBinary: device field -> host -> the reference file format.
inline bool saveFieldBinary(const char* filename, const double* G, SimParams& sp) {
    ls_params_t p = lsref::to_params(sp);
    std::vector<double> h((size_t)ls_field_size(&p));
    if (hipMemcpyAsync(h.data(), G, h.size() * sizeof(double), hipMemcpyDeviceToHost, sp.stream) != hipSuccess ||
        hipStreamSynchronize(sp.stream) != hipSuccess)
        return false;
    return ls_write_G(&p, h.data(), filename) == LS_OK;
}
inline bool saveFieldBinary3D(const char* filename, const double* G, SimParams& sp, double = 0.0) {
    return saveFieldBinary(filename, G, sp);
}

// saveFieldVTK[3D]: device field -> host -> legacy ASCII VTK of the interior.
inline bool saveFieldVTK(const char* filename, const double* G, SimParams& sp) {
    ls_params_t p = lsref::to_params(sp);
    std::vector<double> h((size_t)ls_field_size(&p));
    if (hipMemcpyAsync(h.data(), G, h.size() * sizeof(double), hipMemcpyDeviceToHost, sp.stream) != hipSuccess ||
        hipStreamSynchronize(sp.stream) != hipSuccess)
        return false;
    return ls_write_G_vtk(&p, h.data(), filename) == LS_OK;
}
inline bool saveFieldVTK3D(const char* filename, const double* G, SimParams& sp) {
    return saveFieldVTK(filename, G, sp);
}
