/*
 * ls_api.h -- C ABI of the MI355X (gfx950) G-equation level-set solver.
 *
 *   dG/dt + u . grad G = S_L |grad G|
 *
 * advanced with upwinded WENO-5 on the effective velocity
 * u_eff = u - S_L grad_c G / sqrt(|grad_c G|^2 + eps), TVD-RK3 (Shu-Osher) and
 * HCR-2 constrained reinitialisation, on a padded FP64 grid (3 ghost layers
 * per side, x fastest).
 *
 * Conventions
 *   - Every entry point returns LS_OK (0) or a negative LS_ERR_* code; nothing
 *     calls exit().  ls_last_error() returns a message for the last failure on
 *     the calling thread.
 *   - All field pointers are DEVICE pointers to caller-owned padded arrays of
 *     nx_t*ny_t[*nz_t] doubles, nx_t = nx + 2*nghost, index (k*ny_t + j)*nx_t + i.
 *     The caller allocates them; the solver never frees them.  Solver-owned
 *     workspace is allocated by ls_create and freed by ls_destroy.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Work is stream-ordered; entry points return without synchronising
 *     unless their comment says otherwise.
 *   - One solver handle per host thread at a time; calls that share a handle
 *     from two threads race.
 *
 * Reference seams replaced (juhoon196/Level-set-method-code):
 *   SimParams / initDefaultParams / calculateTimeStep
 *       Local_code/level-set_MPI_2D/include/config.h:123-263
 *       Local_code/level-set_MPI_3D/include/config.h:134-293
 *   rk3TimeStep(G_n, G_new, G_1, G_2, G_rhs, u, v, params)
 *       Local_code/level-set_MPI_2D/include/rk3.h:168-187
 *   rk3TimeStep3D(..., u, v, w, params)
 *       Local_code/level-set_MPI_3D/include/rk3.h:131-150
 *   applyBoundaryConditions / applyBoundaryConditions3D
 *       Local_code/level-set_MPI_2D/include/boundary.h:217-229
 *       Local_code/level-set_MPI_3D/include/boundary.h:250-261
 *   reinitializeWithSwap / reinitializeWithSwap3D
 *       Local_code/level-set_MPI_2D/include/reinitialization.h:321-360
 *       Local_code/level-set_MPI_3D/include/reinitialization.h:199-228
 *   time loop of runStandaloneSimulation[3D]
 *       Local_code/level-set_MPI_2D/src/main.cpp:310-354
 *       Local_code/level-set_MPI_3D/src/main.cpp:287-334
 *   computeL2Error / computeInterfaceArea / computeInterfaceVolume
 *       Local_code/level-set_MPI_2D/include/io.h:194-246
 *       Local_code/level-set_MPI_3D/include/io.h:229-288
 *   saveFieldBinary / saveFieldBinary3D
 *       Local_code/level-set_MPI_2D/include/io.h:34-114
 *       Local_code/level-set_MPI_3D/include/io.h:34-136
 */
#ifndef LS_API_H
#define LS_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LS_OK 0
#define LS_ERR_INVALID (-1)     /* bad parameter or shape */
#define LS_ERR_NOMEM (-2)       /* device allocation failed */
#define LS_ERR_DEVICE (-3)      /* HIP runtime error */
#define LS_ERR_IO (-4)          /* file could not be written */
#define LS_ERR_NONFINITE (-5)   /* a NaN or Inf was found by ls_check_finite */

/* Boundary kinds per axis (both sides of an axis share the kind). */
#define LS_BC_PERIODIC 0        /* boundary.h:98-114: ghost g <- interior n+g */
#define LS_BC_ZERO_GRAD 1       /* boundary.h:126-168: copy of the edge cell */
#define LS_BC_LIN_EXTRAP 2      /* boundary.h:177-200: y axis only */

/* Arithmetic modes. */
#define LS_MODE_EXACT 0         /* the reference's operation order, no FMA,
                                   correctly rounded divisions: bitwise equal
                                   to the contraction-free CPU build */
#define LS_MODE_FAST 1          /* same scheme, reorganised arithmetic
                                   (shared smoothness indicators, reciprocal
                                   multiplies, FMA); within 1e-12 relative */

/* Velocity handling inside ls_advance. */
#define LS_VEL_ARRAYS 0         /* caller arrays, constant in time */
#define LS_VEL_DEFORM3D 1       /* recompute the 3D deformation field at the
                                   start of every step, period vel_period
                                   (3D main.cpp:296-299, initial_conditions.h:232-280) */

typedef struct ls_params {
    int32_t dim;                /* 2 or 3 */
    int32_t nx, ny, nz;         /* interior points; nz ignored in 2D */
    int32_t nghost;             /* must be 3 (WENO-5 stencil) */
    int32_t bc[3];              /* LS_BC_* for x, y, z */
    int32_t mode;               /* LS_MODE_* */
    int32_t velocity;           /* LS_VEL_* */
    double dx, dy, dz;          /* spacing; reference: L/(N-1) */
    double x0, y0, z0;          /* coordinate of the first interior point */
    double s_l;                 /* laminar flame speed (scalar) */
    double eps;                 /* 1e-10: S_L gate and |grad| regulariser */
    int32_t reinit_interval;    /* HCR-2 after steps ri+1, 2ri+1, ...; 0 = off */
    int32_t reinit_iters;       /* pseudo-time iterations (reference default 2) */
    double reinit_dtau;         /* 0.25*min(dx,dy[,dz]) in the reference */
    double reinit_beta;         /* 0.5 */
    double vel_period;          /* T for LS_VEL_DEFORM3D */
} ls_params_t;

typedef struct ls_solver ls_solver_t;

/* Fill `p` with the reference defaults for a dim-D grid of n^dim interior
 * points on the unit box: dx = 1/(n-1), eps 1e-10, HCR-2 dtau 0.25*dx,
 * beta 0.5, 2 iterations, interval 10 (2D: periodic x + zero-gradient y;
 * 3D: periodic), exact mode.  Replaces initDefaultParams (config.h:166-213). */
int ls_default_params(ls_params_t* p, int32_t dim, int32_t n);

/* dt = cfl*min(dx,dy[,dz]) / (|u|+|v|[+|w|]+S_L) with the reference's
 * fallbacks for a vanishing speed (2D: eps, 3D: 2.0).  Replaces
 * calculateTimeStep (2D config.h:251-263, 3D config.h:272-293). */
double ls_cfl_dt(const ls_params_t* p, double cfl, double u, double v, double w);

/* Create a solver on the current HIP device; allocates two padded fields of
 * workspace plus reduction scratch.  The solver stays bound to that device:
 * call it with the same device current and with fields that live there. */
int ls_create(const ls_params_t* p, ls_solver_t** out);
int ls_destroy(ls_solver_t* s);
/* Number of doubles in one padded field. */
int64_t ls_field_size(const ls_params_t* p);

/* Fill every ghost cell of G from its interior (applyBoundaryConditions). */
int ls_fill_ghosts(ls_solver_t* s, double* G, void* stream);

/* One TVD-RK3 step G_new = RK3(G_n) with time step dt, ghosts filled after
 * every stage (rk3TimeStep / rk3TimeStep3D).  w is ignored in 2D.  sl_field
 * NULL uses the scalar s_l; otherwise a padded per-cell flame speed gated per
 * cell by s_l > eps.  G_new must not alias G_n. */
int ls_rk3_step(ls_solver_t* s, const double* G_n, double* G_new, const double* u,
                const double* v, const double* w, const double* sl_field, double dt,
                void* stream);

/* One fused RK3 stage: out = combine_k(G_n, G_s, dt * RHS(G_s)) on interior
 * cells followed by a ghost fill of out, with RHS the u_eff WENO-5 operator:
 *   k=1: G_n + dt*r      k=2: 0.75 G_n + 0.25 G_s + 0.25 dt r
 *   k=3: (1/3) G_n + (2/3) G_s + (2/3) dt r      (G_s = G_n for k = 1).
 * Replaces computeRHS + rk3Stage1/2/3 + applyBoundaryConditions
 * (Local_code/level-set_MPI_2D/include/rk3.h:54-150,
 *  Local_code/level-set_MPI_3D/include/rk3.h:31-125).  out must not alias
 * G_s or G_n. */
int ls_rk3_stage(ls_solver_t* s, int32_t stage, const double* G_s, const double* G_n,
                 double* out, const double* u, const double* v, const double* w,
                 const double* sl_field, double dt, void* stream);

/* HCR-2 reinitialisation of G in place: reinit_iters pseudo-time steps with
 * phi_0 = G and a ghost fill after each (reinitializeWithSwap). */
int ls_reinit_hcr2(ls_solver_t* s, double* G, void* stream);

/* One HCR-2 pseudo-time iteration phi_new = step(phi_n; phi_0) over all
 * interior cells, then a ghost fill of phi_new (reinitStep +
 * applyBoundaryConditions, reinitialization.h:200-263 and :344-355; 3D
 * :136-193, :489-497).  Crossing flags and r~ are derived from phi_0 inside
 * the kernel.  phi_new must not alias phi_n or phi_0. */
int ls_hcr2_step(ls_solver_t* s, const double* phi_n, const double* phi_0, double* phi_new,
                 void* stream);

/* The HCR-2 set-up arrays of computeInterfaceCrossings (reinitialization.h:75-130,
 * 3D :46-89) for callers that read them: r_tilde (padded doubles) and the
 * crossing bit mask (padded int32: 1 x-, 2 x+, 4 y-, 8 y+, 16 z-, 32 z+) on
 * interior cells; ghost entries are left untouched. */
int ls_hcr2_crossings(ls_solver_t* s, const double* phi_0, double* r_tilde, int32_t* flag,
                      void* stream);

/* The reference time loop on G in place:
 *   while (t < t_final && step < step0 + max_steps):
 *     if t + dt > t_final: dt = t_final - t
 *     [LS_VEL_DEFORM3D: u,v,w <- deformation field at time t]
 *     G <- RK3(G);  if ri > 0 && step > 0 && step % ri == 0: HCR-2(G)
 *     t += dt; step++
 * step counts from step0 (the global step index of the first step).  On
 * return *t_io holds the final time, *steps_done the number of steps and
 * *dt_last the last dt used (any of them may be NULL except t_io).
 * With ri == 0 and velocity arrays, runs of full steps replay a hipGraph that
 * the solver captures on a private stream on first use and keeps (for the
 * same pointers and dt) until ls_destroy; the results are those of the
 * launches themselves.  Synchronises the stream before returning. */
int ls_advance(ls_solver_t* s, double* G, double* u, double* v, double* w,
               const double* sl_field, double dt, double t_final, int64_t step0,
               int64_t max_steps, double* t_io, int64_t* steps_done, double* dt_last,
               void* stream);

/* Diagnostics over interior cells: *l2 = sqrt(mean((G-G0)^2)) (G0 may be
 * NULL, then l2 is not computed) and *area = count(G < 0) * dx*dy[*dz].
 * Deterministic summation order.  Synchronises the stream. */
int ls_diag(ls_solver_t* s, const double* G, const double* G0, double* l2, double* area,
            void* stream);

/* LS_OK if every interior value of G is finite, LS_ERR_NONFINITE otherwise.
 * Synchronises the stream. */
int ls_check_finite(ls_solver_t* s, const double* G, void* stream);

/* ls_diag without a solver handle and without allocating.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and summation order as ls_diag; the results
 * go to DEVICE memory, stream-ordered, with no synchronisation:
 *   out[0] = sqrt(mean((G-G0)^2)) over interior cells (0 when G0 is NULL),
 *   out[1] = count(G < 0) * dx*dy[*dz],
 *   out[2] = number of non-finite interior values of G,
 * bitwise the values ls_diag and ls_check_finite see.  The partial sums live
 * in a static device array of fixed size (1025 partials, 24,600 bytes; the
 * reduction runs 1024 blocks that stride over the interior, so its size does
 * not grow with the grid and every grid ls_create accepts fits), one per GPU,
 * kept until the process ends.  Calls that overlap on one GPU (two streams or
 * two host threads) race on it.  Rejects the parameters ls_create rejects. */
int ls_diag_noalloc(const ls_params_t* p, const double* G, const double* G0, double* out, void* stream);

/* Device check of the division and square-root sequences (no counterpart in
 * the reference; the tests call it): out[i] = a[i] / b[i] (kinds 0-2) as the
 * exact-mode kernels compute it on the GPU, for device arrays of n doubles.
 *   kind 0: rdiv_u, the sequence used for quotients with a divisor of their
 *           own (b in [2^-1000, 2^1000], a = 0 or |a| in [2^-960, 2^960],
 *           |a|/b in [2^-1000, 2^700]);
 *   kind 1: qdiv_u with y = rdiv_u(1, b), the sequence of quotients sharing a
 *           divisor (b in [2^-300, 2^300], a = 0 or |a| in [2^-960, 2^960],
 *           |a|/b in [2^-1000, 2^1000]);
 *   kind 2: div6_u, a / 6 (b unused; a = 0 or |a| in [2^-960, 2^1000]);
 *   kind 3: sqrt_u, the square-root sequence of the fast HCR-2 (b unused;
 *           a in [2^-767, 2^1000]): out[i] = sqrt(a[i]);
 *   kind 4: sqrt_u(max(a, 2^-767)) - 1, the fast HCR-2's |grad phi| - 1 (b
 *           unused; a = 0 or in (0, 2^1000]): out[i] = sqrt(a[i]) - 1.
 * Inside those ranges each must equal the IEEE result bit for bit.
 * Stream-ordered, no synchronisation. */
int ls_div_check(const double* a, const double* b, double* out, int64_t n, int32_t kind, void* stream);

/* Write a HOST copy of a padded field in the reference's binary G format:
 * int32 {nx, ny, nghost} (2D) or {nx, ny, nz, nghost} (3D), then the padded
 * doubles with the y-ghost rows replaced by the adjacent interior row. */
int ls_write_G(const ls_params_t* p, const double* host_G, const char* path);

/* Write a HOST copy of a padded field as legacy ASCII VTK structured points of
 * the interior (x fastest, then y, then z), as saveFieldVTK
 * (Local_code/level-set_MPI_2D/include/io.h:119-180) and saveFieldVTK3D
 * (Local_code/level-set_MPI_3D/include/io.h:145-215) write it: ORIGIN is
 * (x0, y0[, z0]), SPACING (dx, dy[, dz]), one value per line in %e. */
int ls_write_G_vtk(const ls_params_t* p, const double* host_G, const char* path);

/* Synthetic inputs generated on the device (padded arrays, ghosts included).
 *   LS_CASE_CIRCLE2D        circle r=0.25 at (0.5,0.5), u=(0.5,0)
 *   LS_CASE_SPHERE_SWIRL3D  sphere r=0.2 at (0.5,0.5,0.5),
 *                           u=(sin pi x cos pi y, -cos pi x sin pi y, 0.2 sin 2 pi z)
 *   LS_CASE_PLANE_MODES3D   G = x-0.3-eta(y,z), eta = 8 seeded modes a~U[0,0.02],
 *                           k in {1..8}^2; u = (1,0,0) + 16 seeded solenoidal
 *                           modes of amplitude 0.1, integer |k| in [1,4]
 *   LS_CASE_ZALESAK2D       reference Zalesak disk (0.5,0.75) r 0.15, slot
 *                           0.05 x 0.25, rotation omega = 2 pi / period
 *   LS_CASE_SPHERE_DEFORM3D reference sphere (0.35,0.35,0.35) r 0.15 with the
 *                           deformation velocity at t = 0, period `period`
 *   LS_CASE_FLAME2D         seeded circle or plane G0, u = uniform + 8 seeded
 *                           modes of amplitude 0.1 (one ensemble member)
 * w may be NULL in 2D.  `seed` selects the seeded draws (SplitMix64). */
#define LS_CASE_CIRCLE2D 1
#define LS_CASE_SPHERE_SWIRL3D 2
#define LS_CASE_PLANE_MODES3D 3
#define LS_CASE_ZALESAK2D 4
#define LS_CASE_SPHERE_DEFORM3D 5
#define LS_CASE_FLAME2D 6
int ls_init_case(const ls_params_t* p, int32_t case_id, uint64_t seed, double period,
                 double* G, double* u, double* v, double* w, void* stream);

/* The time-dependent deformation field of the 3D sphere test at time t,
 * period `period`, on the whole padded arrays (updateDeformationVelocity,
 * Local_code/level-set_MPI_3D/include/initial_conditions.h:232-280). */
int ls_deformation_velocity(const ls_params_t* p, double t, double period, double* u, double* v,
                            double* w, void* stream);

/* ---- Batched engine: many independent 2D flames of uneven sizes ----------
 * Every member's padded field lives in one packed device buffer at element
 * offset ls_batch_offset(b, m) (layout of one member as for a single grid).
 * Each RK3 stage is one launch over all members' tiles (the kernel writes the
 * ghost cells itself); ls_batch_advance captures nsteps steps into a hipGraph
 * on first use and replays it while the arguments stay the same.  No HCR-2.
 * The single-grid path it batches is rk3TimeStep (rk3.h:168-187) run as a
 * fixed number of steps with a per-member dt. */
typedef struct ls_batch_member {
    int32_t nx, ny;     /* interior points, >= 6 */
    int32_t bc[2];      /* LS_BC_PERIODIC or LS_BC_ZERO_GRAD for x and y */
    double dx, dy;
    double s_l;         /* flame speed of this member */
    double dt;          /* time step of this member */
} ls_batch_member_t;

typedef struct ls_batch ls_batch_t;

int ls_batch_create(const ls_batch_member_t* members, int32_t count, int32_t mode, ls_batch_t** out);
int ls_batch_destroy(ls_batch_t* b);
/* Total packed doubles, and the element offset of member m. */
int64_t ls_batch_size(const ls_batch_t* b);
int64_t ls_batch_offset(const ls_batch_t* b, int32_t member);
/* Ghost fill of every member of a packed field. */
int ls_batch_fill_ghosts(ls_batch_t* b, double* G, void* stream);
/* nsteps RK3 steps of every member, G in place (packed).  use_graph != 0
 * captures and replays a hipGraph.  Does not synchronise. */
int ls_batch_advance(ls_batch_t* b, double* G, const double* u, const double* v, int32_t nsteps,
                     int32_t use_graph, void* stream);
/* Schedule shapes 0..7: the y-march, one wavefront per 64-column strip and
 * chunk of 32, 16, 4 or 8 rows (shape % 4, both modes), with the kernel's own
 * register budget (shapes 0-3) or the budget of one more wavefront per SIMD
 * (shapes 4-7).  ls_batch_autotune times probe_steps steps of each (two
 * passes) on a scratch copy and keeps the fastest (*best_ms = its time per
 * step).  The result does not depend on the shape.  Synchronises the stream. */
int ls_batch_set_tile(ls_batch_t* b, int32_t shape);
int32_t ls_batch_get_tile(const ls_batch_t* b);
int ls_batch_autotune(ls_batch_t* b, const double* G, const double* u, const double* v,
                      int32_t probe_steps, double* best_ms, void* stream);

/* ---- Discrete adjoint of the 3D solver ----------------------------------
 * No counterpart in the reference, which has no derivatives; the adjoint
 * differentiates the same discrete operators that ls_rk3_stage and
 * ls_hcr2_step apply (rk3.h:31-125, weno5.h:28-248, reinitialization.h:46-193,
 * boundary.h:250-261, all under Local_code/level-set_MPI_3D/include).
 *
 * Loss L = sum over interior cells of (H_e(G_T) - H_e(G_target))^2, with G_T
 * the state after nsteps RK3 steps from G0 (HCR-2 after steps ri, 2ri, ...
 * as in ls_advance with step0 = 0) and the smoothed Heaviside
 * H_e(x) = 0 (x < -e), 1 (x > e), (1 + x/e + sin(pi x/e)/pi)/2 otherwise.
 *
 * ls_adjoint_create binds a context to a 3D solver (periodic or zero-gradient
 * boundaries) and allocates its workspace: 26 padded fields, plus one per
 * checkpoint (every ckpt steps), ckpt + 1 segment states, 2 ckpt stage
 * fields, ckpt pre-reinitialisation states and reinit_iters - 1 HCR-2
 * iterates. */
typedef struct ls_adjoint ls_adjoint_t;
int ls_adjoint_create(ls_solver_t* s, int32_t nsteps, int32_t ckpt, ls_adjoint_t** out);
int ls_adjoint_destroy(ls_adjoint_t* a);
/* Device bytes held by the context. */
int64_t ls_adjoint_workspace_bytes(const ls_adjoint_t* a);
/* L for a given final state G (no time stepping).  Synchronises the stream. */
int ls_adjoint_loss(ls_solver_t* s, const double* G, const double* G_target, double eps_h,
                    double* loss, void* stream);
/* Forward sweep with checkpoints, then the reverse sweep.  Writes *loss and,
 * when gu/gv/gw are given, dL/du, dL/dv, dL/dw (interior cells), dL/dS_L
 * (gsl, per cell; the sum over cells when sl_field is NULL and the scalar
 * s_l is used) and dL/dG0 (gG0, ghosts folded into the interior).  All
 * gradient pointers NULL: loss only.  sl_field as in ls_rk3_stage.
 * Synchronises the stream. */
int ls_adjoint_run(ls_adjoint_t* a, const double* G0, const double* u, const double* v,
                   const double* w, const double* sl_field, const double* G_target,
                   double eps_h, double dt, double* loss, double* gG0, double* gu, double* gv,
                   double* gw, double* gsl, void* stream);

/* Many parameter sets at once (no counterpart in the reference): set q runs
 * ls_adjoint_run on context q % nctx, with that context's stream
 * (streams[c], NULL = all on the null stream).  Each context gets a host
 * thread of its own and takes its sets in order, so up to nctx sets run
 * concurrently on the GPU; results are those of ls_adjoint_run called set by
 * set.  The contexts must be distinct and bound to distinct solvers; sets
 * sharing a context must not share output buffers with sets of another
 * context that run at the same time.  Returns the error of the first failing
 * set (its index in the message).  Synchronises every stream used. */
typedef struct ls_adjoint_set {
    const double* G0;
    const double* u;
    const double* v;
    const double* w;
    const double* sl_field;     /* NULL: the solver's scalar s_l */
    const double* G_target;
    double eps_h, dt;
    double* loss;               /* host */
    double *gG0, *gu, *gv, *gw, *gsl;   /* device, or all NULL (loss only) */
} ls_adjoint_set_t;
int ls_adjoint_run_sets(ls_adjoint_t* const* ctx, void* const* streams, int32_t nctx,
                        const ls_adjoint_set_t* sets, int32_t nsets);

/* Parameter maps of the flame-inversion problem (no counterpart in the
 * reference).  S_L(x) is the trilinear interpolant of nc^3 coarse node values
 * (device, x fastest) over the interior box; the velocity is the 3D swirl
 * field of LS_CASE_SPHERE_SWIRL3D plus sum_q amps[q] * a_q sin(2 pi k_q.x + ph_q)
 * over nmodes (<= 16) solenoidal unit modes drawn from mode_seed.  Either
 * output group may be NULL.  ls_flame_project applies the transposed maps to
 * the per-cell adjoints (interior cells) and synchronises the stream. */
int ls_flame_fields(const ls_params_t* p, uint64_t mode_seed, int32_t nc, const double* sl_coarse,
                    int32_t nmodes, const double* amps, double* sl_field, double* u, double* v,
                    double* w, void* stream);
int ls_flame_project(const ls_params_t* p, uint64_t mode_seed, int32_t nc, int32_t nmodes,
                     const double* gsl, const double* gu, const double* gv, const double* gw,
                     double* g_coarse, double* g_amps, void* stream);

/* ---- The sample/ scheme (2D, local Lax-Friedrichs) ------------------------
 * The reference's second 2D solver (sample/): WENO-5 with its own smoothness
 * indicators, an LLF Hamiltonian, RK3 in the forms 0.75 G + 0.25 (G1 + dt r)
 * and 1/3 G + 2/3 (G2 + dt r), the anchored Bunsen boundary followed by the
 * periodic fill after every stage.  Fields are padded (ny+6) x (nx+6)
 * doubles, x fastest (Field2D::data[j][i] at j*(nx+6)+i).  Arithmetic equals
 * the reference's CPU build bit for bit. */
typedef struct ls_sample ls_sample_t;
/* (nx+6)*(ny+6), or -1 when nx or ny < 6. */
int64_t ls_sample_field_size(int32_t nx, int32_t ny);
int ls_sample_create(int32_t nx, int32_t ny, double dx, double dy, ls_sample_t** out);
int ls_sample_destroy(ls_sample_t* s);
/* compute_geqn_rhs (sample/weno5_rk3.cpp:91-145): rhs on interior cells,
 * zero elsewhere. */
int ls_sample_rhs(ls_sample_t* s, const double* G, const double* u, const double* v, double S_L,
                  double* rhs, void* stream);
/* RK3_step (sample/weno5_rk3.cpp:148-188), G in place. */
int ls_sample_rk3_step(ls_sample_t* s, double* G, const double* u, const double* v, double dt,
                       double S_L, const double* G_initial, void* stream);
/* Field2D::apply_periodic and apply_anchored_bc (sample/field.cpp:16-63). */
int ls_sample_apply_periodic(ls_sample_t* s, double* G, void* stream);
int ls_sample_apply_anchored_bc(ls_sample_t* s, double* G, const double* G_initial, void* stream);
/* Field2D::reinitialize (sample/field.cpp:66-130). */
int ls_sample_reinitialize(ls_sample_t* s, double* G, const double* G_initial, int32_t num_iters,
                           void* stream);
/* compute_first_order_upwind (sample/first_order_upwind.cpp:6-71), G in place. */
int ls_sample_first_order_upwind(ls_sample_t* s, double* G, const double* u, const double* v,
                                 double dt, double S_L, void* stream);
/* The Bunsen-flame initial field of sample/main.cpp:95-116 (whole padded array). */
int ls_sample_bunsen(ls_sample_t* s, double R, double flame_angle_deg, double X0, double Y0,
                     double* G, void* stream);
/* Uniform forced velocity u = u0 + eps_u v0 cos(2 pi freq t),
 * v = v0 + eps_v v0 cos(2 pi freq t) (sample/main.cpp:131-136). */
int ls_sample_forced_velocity(ls_sample_t* s, double t, int32_t freq, double u0, double v0,
                              double eps_u, double eps_v, double* u, double* v, void* stream);
/* nsteps iterations of the sample/main.cpp:126-148 loop body from time t0
 * (boundary, forced velocity, RK3_step, boundary, t += dt); *t_out receives
 * the final time.  Synchronises the stream. */
int ls_sample_advance(ls_sample_t* s, double* G, double* u, double* v, const double* G_initial,
                      double S_L, double dt, double t0, int64_t nsteps, int32_t freq, double u0,
                      double v0, double eps_u, double eps_v, double* t_out, void* stream);
/* save_G_to_bin (sample/main.cpp:34-51): interior rows of a HOST padded
 * field, no header. */
int ls_sample_write_G(int32_t nx, int32_t ny, const double* host_G, const char* path);

/* Message for the last error on this thread (never NULL). */
const char* ls_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* LS_API_H */
