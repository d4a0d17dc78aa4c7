// ref_harness.cpp -- builds the reference's own CPU solver into oracle/_ref.
//
// TEST INFRASTRUCTURE ONLY.  This translation unit contains no numerics of
// its own: it includes the reference headers where they lie
// (Local_code/level-set_MPI_2D/include or Local_code/level-set_MPI_3D/include,
// selected by -DREF_DIM=2|3 and -I), fills a SimParams with runtime sizes
// instead of the compile-time NX/NY/NZ, and exposes the reference functions
// through extern "C" entry points so tests can call them with ctypes.
//
// The reference functions read only runtime SimParams fields (nx, local_ny,
// nx_total, ...), never the constexpr grid constants, so runtime sizes are
// exact.  One MPI rank: the 2D y-halo exchange is a no-op
// (Local_code/level-set_MPI_2D/include/config.h:242-243) and the 3D y-halo is
// a self-exchange that realises the periodic wrap
// (Local_code/level-set_MPI_3D/include/config.h:263-264).
#include <mpi.h>
#include <cstdint>
#include <cstring>
#include <cstdlib>

#include "config.h"
#include "weno5.h"
#include "boundary.h"
#include "rk3.h"
#include "reinitialization.h"
#include "io.h"
#include "initial_conditions.h"

// Same memory layout as lso_grid in ls_oracle.c.
struct ref_grid {
    int32_t dim, nx, ny, nz, ng;
    int32_t bc[3];
    double dx, dy, dz, s_l, eps;
    int32_t reinit_iters, pad0;
    double reinit_dtau, reinit_beta;
};

static void ensure_mpi() {
    int ok = 0;
    MPI_Initialized(&ok);
    if (!ok) MPI_Init(nullptr, nullptr);
}

static SimParams mk(const ref_grid* g) {
    SimParams p;
    std::memset(&p, 0, sizeof(p));
    p.nx = g->nx;
    p.ny = g->ny;
    p.nghost = g->ng;
    p.nx_total = g->nx + 2 * g->ng;
    p.ny_total = g->ny + 2 * g->ng;
    p.dx = g->dx;
    p.dy = g->dy;
#if REF_DIM == 3
    p.nz = g->nz;
    p.nz_total = g->nz + 2 * g->ng;
    p.dz = g->dz;
#endif
    p.local_ny = g->ny;
    p.local_ny_total = g->ny + 2 * g->ng;
    setupMPIDomainDecomposition(p, 0, 1);
    p.s_l = g->s_l;
    p.epsilon = g->eps;
    p.enable_reinit = true;
    p.reinit_iterations = g->reinit_iters;
    p.reinit_dtau = g->reinit_dtau;
    p.reinit_beta = g->reinit_beta;
    return p;
}

// Ghost fill composed from the reference's own BC functions.
// 2D: exchange -> x (periodic | zero-gradient) -> y (zero-gradient | linear
// extrapolation), as applyBoundaryConditions (boundary.h:217-229) orders them.
// 3D: applyBoundaryConditions3D (fully periodic, boundary.h:250-261).
static int ref_bc(double* G, SimParams& p, const ref_grid* g) {
#if REF_DIM == 2
    exchangeGhostCells(G, p);
    if (g->bc[0] == 0) applyPeriodicBC_X(G, p.nx, p.local_ny, p.nghost, p.nx_total);
    else if (g->bc[0] == 1) applyZeroGradientBC_X(G, p.nx, p.local_ny, p.nghost, p.nx_total);
    else return -1;
    if (g->bc[1] == 1) applyZeroGradientBC_Y(G, p.nx, p.local_ny, p.nghost, p.nx_total, true, true);
    else if (g->bc[1] == 2) applyExtrapolationBC_Y(G, p.nx, p.local_ny, p.nghost, p.nx_total, true, true);
    else return -1;
    return 0;
#else
    if (g->bc[0] != 0 || g->bc[1] != 0 || g->bc[2] != 0) return -1;
    applyBoundaryConditions3D(G, p);
    return 0;
#endif
}

static int64_t padded(const ref_grid* g) {
    int64_t n = (int64_t)(g->nx + 2 * g->ng) * (g->ny + 2 * g->ng);
    if (g->dim == 3) n *= (g->nz + 2 * g->ng);
    return n;
}

static bool default_bc(const ref_grid* g) {
#if REF_DIM == 2
    return g->bc[0] == 0 && g->bc[1] == 1;
#else
    return g->bc[0] == 0 && g->bc[1] == 0 && g->bc[2] == 0;
#endif
}

extern "C" {

int ref_fill_ghosts(const ref_grid* g, double* G) {
    ensure_mpi();
    SimParams p = mk(g);
    return ref_bc(G, p, g);
}

int ref_rhs(const ref_grid* g, const double* G, double* rhs, const double* u, const double* v,
            const double* w) {
    ensure_mpi();
    SimParams p = mk(g);
#if REF_DIM == 2
    (void)w;
    computeRHS(G, rhs, u, v, p);
#else
    computeRHS3D(G, rhs, u, v, w, p);
#endif
    return 0;
}

// One RK3 step.  direct != 0 calls the reference's rk3TimeStep[3D] (default
// BCs only); otherwise the step is composed from computeRHS + rk3StageK +
// ref_bc.  work: G1, G2, rhs (3 padded fields, rhs zeroed here).
int ref_rk3_step(const ref_grid* g, double* Gn, double* Gnew, const double* u, const double* v,
                 const double* w, double dt, double* work, int direct) {
    ensure_mpi();
    SimParams p = mk(g);
    p.dt = dt;
    int64_t n = padded(g);
    double* G1 = work;
    double* G2 = work + n;
    double* R = work + 2 * n;
    std::memset(R, 0, n * sizeof(double));
    if (direct) {
        if (!default_bc(g)) return -1;
#if REF_DIM == 2
        (void)w;
        rk3TimeStep(Gn, Gnew, G1, G2, R, u, v, p);
#else
        rk3TimeStep3D(Gn, Gnew, G1, G2, R, u, v, w, p);
#endif
        return 0;
    }
#if REF_DIM == 2
    (void)w;
    computeRHS(Gn, R, u, v, p);
    rk3Stage1(Gn, R, G1, p.dt, p);
    if (ref_bc(G1, p, g)) return -1;
    computeRHS(G1, R, u, v, p);
    rk3Stage2(Gn, G1, R, G2, p.dt, p);
    ref_bc(G2, p, g);
    computeRHS(G2, R, u, v, p);
    rk3Stage3(Gn, G2, R, Gnew, p.dt, p);
    ref_bc(Gnew, p, g);
#else
    computeRHS3D(Gn, R, u, v, w, p);
    rk3Stage1_3D(Gn, R, G1, p.dt, p);
    if (ref_bc(G1, p, g)) return -1;
    computeRHS3D(G1, R, u, v, w, p);
    rk3Stage2_3D(Gn, G1, R, G2, p.dt, p);
    ref_bc(G2, p, g);
    computeRHS3D(G2, R, u, v, w, p);
    rk3Stage3_3D(Gn, G2, R, Gnew, p.dt, p);
    ref_bc(Gnew, p, g);
#endif
    return 0;
}

// HCR-2 on G in place.  direct != 0 calls reinitializeWithSwap[3D] (default
// BCs); otherwise it is composed from the reference's crossings/step
// functions with ref_bc.  work: 3 padded doubles; flag: padded int32.
int ref_hcr2(const ref_grid* g, double* G, double* work, int* flag, int direct) {
    ensure_mpi();
    SimParams p = mk(g);
    int64_t n = padded(g);
    double* tmp = work;
    double* p0 = work + n;
    double* rt = work + 2 * n;
    double* a = G;
    double* b = tmp;
    if (direct) {
        if (!default_bc(g)) return -1;
#if REF_DIM == 2
        reinitializeWithSwap(&a, &b, p0, rt, flag, p);
#else
        reinitializeWithSwap3D(&a, &b, p0, rt, flag, p);
#endif
    } else {
        std::memcpy(p0, G, n * sizeof(double));
        if (ref_bc(p0, p, g)) return -1;
#if REF_DIM == 2
        computeInterfaceCrossings(p0, rt, flag, p);
#else
        computeInterfaceCrossings3D(p0, rt, flag, p);
#endif
        for (int it = 0; it < p.reinit_iterations; it++) {
#if REF_DIM == 2
            reinitStep(a, b, p0, rt, flag, p.reinit_dtau, p);
#else
            reinitStep3D(a, b, p0, rt, flag, p.reinit_dtau, p);
#endif
            double* t = a; a = b; b = t;
            ref_bc(a, p, g);
        }
    }
    if (a != G) std::memcpy(G, a, n * sizeof(double));
    return 0;
}

// computeInterfaceCrossings[3D] on phi_0 as given (its ghosts already
// filled by the caller): r~ and the crossing bit mask on interior cells.
int ref_crossings(const ref_grid* g, const double* p0, double* rt, int* flag) {
    ensure_mpi();
    SimParams p = mk(g);
#if REF_DIM == 2
    computeInterfaceCrossings(p0, rt, flag, p);
#else
    computeInterfaceCrossings3D(p0, rt, flag, p);
#endif
    return 0;
}

// Time loop with the semantics of main.cpp:310-354 (3D :287-334).
// deform_T > 0: recompute the 3D deformation velocity with the reference's
// updateDeformationVelocity at the start of each step.  work: 7 padded
// doubles; flag: padded int32.  Returns the step count.
int64_t ref_advance(const ref_grid* g, double* G, double* u, double* v, double* w, double dt,
                    double t_final, int64_t max_steps, int32_t reinit_interval, double deform_T,
                    double* work, int* flag, double* t_out, double* dt_last) {
    ensure_mpi();
    SimParams p = mk(g);
    int64_t n = padded(g);
    double* Gnew = work + 3 * n;
    double* rwork = work + 4 * n;
    double t = 0.0;
    int64_t step = 0;
    const bool direct = default_bc(g);
    while (t < t_final && step < max_steps) {
        if (t + dt > t_final) dt = t_final - t;
#if REF_DIM == 3
        if (deform_T > 0.0) {
            p.current_time = t;
            p.t_final = deform_T;
            updateDeformationVelocity(u, v, w, p);
            applyVelocityBoundaryConditions3D(u, v, w, p);
        }
#else
        (void)deform_T;
#endif
        if (ref_rk3_step(g, G, Gnew, u, v, w, dt, work, direct ? 1 : 0)) return -1;
        std::memcpy(G, Gnew, n * sizeof(double));
        if (reinit_interval > 0 && step > 0 && step % reinit_interval == 0)
            ref_hcr2(g, G, rwork, flag, direct ? 1 : 0);
        t += dt;
        step++;
    }
    if (t_out) *t_out = t;
    if (dt_last) *dt_last = dt;
    return step;
}

double ref_l2(const ref_grid* g, const double* G, const double* G0) {
    ensure_mpi();
    SimParams p = mk(g);
#if REF_DIM == 2
    return computeL2Error(G, G0, p);
#else
    return computeL2Error3D(G, G0, p);
#endif
}

double ref_area(const ref_grid* g, const double* G) {
    ensure_mpi();
    SimParams p = mk(g);
#if REF_DIM == 2
    return computeInterfaceArea(G, p);
#else
    return computeInterfaceVolume(G, p);
#endif
}

int ref_write_G(const ref_grid* g, const double* G, const char* path) {
    ensure_mpi();
    SimParams p = mk(g);
#if REF_DIM == 2
    return saveFieldBinary(path, G, p) ? 0 : -1;
#else
    return saveFieldBinary3D(path, G, p, 0.0) ? 0 : -1;
#endif
}

// The reference's VTK writer (x_min = y_min = z_min = 0 here).
int ref_write_vtk(const ref_grid* g, const double* G, const char* path) {
    ensure_mpi();
    SimParams p = mk(g);
#if REF_DIM == 2
    return saveFieldVTK(path, G, p) ? 0 : -1;
#else
    return saveFieldVTK3D(path, G, p) ? 0 : -1;
#endif
}

// Reference initial conditions on the padded grid (x0 = y0 = z0 = 0).
// 2D: case 1 circle (0.25,0.5) r 0.15, case 2 Zalesak disk + rotation
// (omega = 2 pi / t_final).  3D: case 0 sphere-deformation test (G only;
// velocity via updateDeformationVelocity), case 1 Zalesak sphere + rotation.
int ref_init_case(const ref_grid* g, int which, double t_final, double u_c, double v_c,
                  double* G, double* u, double* v, double* w) {
    ensure_mpi();
    SimParams p = mk(g);
    p.x_min = 0.0; p.x_max = 1.0; p.y_min = 0.0; p.y_max = 1.0;
    p.t_final = t_final;
    p.u_const = u_c;
    p.v_const = v_c;
#if REF_DIM == 2
    (void)w;
    if (which == 1) initCircleTestCase(G, u, v, p);
    else if (which == 2) initZalesakTestCase(G, u, v, p);
    else return -1;
#else
    p.z_min = 0.0; p.z_max = 1.0;
    if (which == 0) initSphereTestCase(G, u, v, w, p);
    else if (which == 1) initZalesak3DTestCase(G, u, v, w, p);
    else return -1;
#endif
    return 0;
}

}  // extern "C"
