/*
 * ls_oracle.c -- CPU restatement of the reference G-equation level-set solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke entry
 * point and the bench's CPU-baseline leg may load it; the product path (the
 * HIP library) never does.
 *
 * What it restates (reference = juhoon196/Level-set-method-code, CPU/MPI variants):
 *   - WENO-5 reconstruction with the reference's smoothness indicators
 *       Local_code/level-set_MPI_2D/include/weno5.h:39-122
 *       Local_code/level-set_MPI_3D/include/weno5.h:28-81
 *   - upwinded WENO-5 axis derivatives
 *       Local_code/level-set_MPI_2D/include/weno5.h:135-219
 *       Local_code/level-set_MPI_3D/include/weno5.h:90-196
 *   - effective-velocity RHS  -(u_eff . grad_w G)
 *       Local_code/level-set_MPI_2D/include/rk3.h:54-98
 *       Local_code/level-set_MPI_3D/include/rk3.h:31-81
 *   - Shu-Osher TVD-RK3 stage combines over the whole padded array
 *       Local_code/level-set_MPI_2D/include/rk3.h:103-187
 *       Local_code/level-set_MPI_3D/include/rk3.h:87-150
 *   - ghost fill: periodic / zero-gradient / linear extrapolation
 *       Local_code/level-set_MPI_2D/include/boundary.h:98-229
 *       Local_code/level-set_MPI_3D/include/boundary.h:37-261
 *   - HCR-2 constrained reinitialisation (crossings, first-order Godunov
 *     |grad phi|, stability constraint, pseudo-time step, swap semantics)
 *       Local_code/level-set_MPI_2D/include/reinitialization.h:37-360
 *       Local_code/level-set_MPI_2D/include/weno5.h:234-267
 *       Local_code/level-set_MPI_3D/include/reinitialization.h:46-228
 *       Local_code/level-set_MPI_3D/include/weno5.h:202-248
 *   - time-loop semantics (sliver last step, reinit trigger before step++)
 *       Local_code/level-set_MPI_2D/src/main.cpp:310-354
 *       Local_code/level-set_MPI_3D/src/main.cpp:287-334
 *   - diagnostics (RMS of G-G0 over the interior, cell-counted area/volume)
 *       Local_code/level-set_MPI_2D/include/io.h:194-258
 *       Local_code/level-set_MPI_3D/include/io.h:229-300
 *   - G output file format
 *       Local_code/level-set_MPI_2D/include/io.h:34-114
 *       Local_code/level-set_MPI_3D/include/io.h:34-136
 *
 * Floating-point expressions keep the reference's C++ evaluation order
 * (left-to-right, no reassociation).  Build with -ffp-contract=off to obtain
 * the contraction-free bit pattern that the GPU exact mode reproduces.
 *
 * Sizes are runtime values.  Layout: padded FP64, x fastest,
 * index = (k*nyt + j)*nxt + i with 3 ghost layers per side (2D: nzt = 1).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LSO_BC_PERIODIC 0
#define LSO_BC_ZERO_GRAD 1
#define LSO_BC_LIN_EXTRAP 2

typedef struct {
    int32_t dim;          /* 2 or 3 */
    int32_t nx, ny, nz;   /* interior points (nz = 1 in 2D) */
    int32_t ng;           /* ghost layers, must be 3 */
    int32_t bc[3];        /* per axis x, y, z */
    double dx, dy, dz;
    double s_l;           /* scalar laminar flame speed */
    double eps;           /* 1e-10 gate and |grad| regulariser */
    int32_t reinit_iters; /* HCR-2 pseudo-time iterations */
    int32_t pad0;
    double reinit_dtau;   /* 0.25*min(dx,dy[,dz]) in the reference */
    double reinit_beta;   /* 0.5 */
} lso_grid;

static inline int64_t nxt_(const lso_grid* g) { return g->nx + 2 * g->ng; }
static inline int64_t nyt_(const lso_grid* g) { return g->ny + 2 * g->ng; }
static inline int64_t nzt_(const lso_grid* g) { return g->dim == 3 ? g->nz + 2 * g->ng : 1; }

int64_t lso_padded_size(const lso_grid* g) { return nxt_(g) * nyt_(g) * nzt_(g); }

/* ---- WENO-5 (weno5.h:39-122) --------------------------------------------- */

static inline void lso_beta(const double* v, double* b) {
    double d0 = v[2] - 2.0 * v[1] + v[0];
    b[0] = (13.0 / 12.0) * d0 * d0 + 0.25 * (v[2] - v[0]) * (v[2] - v[0]);
    double d1 = v[3] - 2.0 * v[2] + v[1];
    b[1] = (13.0 / 12.0) * d1 * d1 + 0.25 * (v[3] - v[1]) * (v[3] - v[1]);
    double d2 = v[4] - 2.0 * v[3] + v[2];
    b[2] = (13.0 / 12.0) * d2 * d2 + 0.25 * (v[4] - v[2]) * (v[4] - v[2]);
}

#define LSO_WENO_EPS 1.0e-6

double lso_weno_left(const double* v) {
    double b[3];
    lso_beta(v, b);
    double a0 = 0.1 / ((LSO_WENO_EPS + b[0]) * (LSO_WENO_EPS + b[0]));
    double a1 = 0.6 / ((LSO_WENO_EPS + b[1]) * (LSO_WENO_EPS + b[1]));
    double a2 = 0.3 / ((LSO_WENO_EPS + b[2]) * (LSO_WENO_EPS + b[2]));
    double s = a0 + a1 + a2;
    double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
    double p0 = (2.0 * v[0] - 7.0 * v[1] + 11.0 * v[2]) / 6.0;
    double p1 = (-v[1] + 5.0 * v[2] + 2.0 * v[3]) / 6.0;
    double p2 = (2.0 * v[2] + 5.0 * v[3] - v[4]) / 6.0;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

double lso_weno_right(const double* v) {
    double m[5] = {v[4], v[3], v[2], v[1], v[0]};
    double b[3];
    lso_beta(m, b);
    double a0 = 0.3 / ((LSO_WENO_EPS + b[0]) * (LSO_WENO_EPS + b[0]));
    double a1 = 0.6 / ((LSO_WENO_EPS + b[1]) * (LSO_WENO_EPS + b[1]));
    double a2 = 0.1 / ((LSO_WENO_EPS + b[2]) * (LSO_WENO_EPS + b[2]));
    double s = a0 + a1 + a2;
    double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
    double p0 = (11.0 * v[2] - 7.0 * v[3] + 2.0 * v[4]) / 6.0;
    double p1 = (2.0 * v[1] + 5.0 * v[2] - v[3]) / 6.0;
    double p2 = (-v[0] + 5.0 * v[1] + 2.0 * v[2]) / 6.0;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

/* Upwinded derivative along one axis (weno5.h:135-171).  G points at the
 * cell, s is the stride of the axis. */
static inline double lso_weno_deriv(const double* G, int64_t s, double vel, double h) {
    double v[5] = {G[-2 * s], G[-s], G[0], G[s], G[2 * s]};
    if (vel >= 0.0) {
        double vl[5] = {G[-3 * s], G[-2 * s], G[-s], G[0], G[s]};
        double gl = lso_weno_left(vl);
        double gr = lso_weno_left(v);
        return (gr - gl) / h;
    } else {
        double vr[5] = {G[-s], G[0], G[s], G[2 * s], G[3 * s]};
        double gl = lso_weno_right(v);
        double gr = lso_weno_right(vr);
        return (gr - gl) / h;
    }
}

/* ---- RHS (rk3.h:54-98, 3D rk3.h:31-81) ----------------------------------- */

/* rhs is written on interior cells only; ghost entries are left untouched
 * (the reference allocates rhs zero-filled and never writes its ghosts).
 * sl == NULL uses the scalar g->s_l; otherwise a per-cell flame speed with
 * the same s_l > eps gate per cell. */
void lso_rhs(const lso_grid* g, const double* G, double* rhs, const double* u,
             const double* v, const double* w, const double* sl) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g);
    const int64_t sx = 1, sy = nxt, sz = nxt * nyt;
    const int ng = g->ng;
    const int kz0 = g->dim == 3 ? ng : 0;
    const int kz1 = g->dim == 3 ? g->nz + ng : 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = kz0; k < kz1; k++) {
        for (int j = ng; j < g->ny + ng; j++) {
            for (int i = ng; i < g->nx + ng; i++) {
                int64_t c = (int64_t)k * sz + (int64_t)j * sy + i;
                const double* P = G + c;
                double s_l = sl ? sl[c] : g->s_l;
                double gx = 0.0, gy = 0.0, gz = 0.0, gm = 1.0;
                if (g->dim == 3) {
                    if (s_l > g->eps) {
                        gx = (P[sx] - P[-sx]) / (2.0 * g->dx);
                        gy = (P[sy] - P[-sy]) / (2.0 * g->dy);
                        gz = (P[sz] - P[-sz]) / (2.0 * g->dz);
                        gm = sqrt(gx * gx + gy * gy + gz * gz + g->eps);
                    }
                    double ue = u[c] - s_l * gx / gm;
                    double ve = v[c] - s_l * gy / gm;
                    double we = w[c] - s_l * gz / gm;
                    double dgx = lso_weno_deriv(P, sx, ue, g->dx);
                    double dgy = lso_weno_deriv(P, sy, ve, g->dy);
                    double dgz = lso_weno_deriv(P, sz, we, g->dz);
                    rhs[c] = -(ue * dgx + ve * dgy + we * dgz);
                } else {
                    if (s_l > g->eps) {
                        gx = (P[sx] - P[-sx]) / (2.0 * g->dx);
                        gy = (P[sy] - P[-sy]) / (2.0 * g->dy);
                        gm = sqrt(gx * gx + gy * gy + g->eps);
                    }
                    double ue = u[c] - s_l * gx / gm;
                    double ve = v[c] - s_l * gy / gm;
                    double dgx = lso_weno_deriv(P, sx, ue, g->dx);
                    double dgy = lso_weno_deriv(P, sy, ve, g->dy);
                    rhs[c] = -(ue * dgx + ve * dgy);
                }
            }
        }
    }
}

/* ---- RK3 stage combines (rk3.h:103-150) ----------------------------------- */

void lso_stage(int stage, int64_t n, const double* Gn, const double* Gs,
               const double* rhs, double* out, double dt) {
    if (stage == 1) {
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < n; q++) out[q] = Gn[q] + dt * rhs[q];
    } else if (stage == 2) {
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < n; q++) out[q] = 0.75 * Gn[q] + 0.25 * Gs[q] + 0.25 * dt * rhs[q];
    } else {
        const double a0 = 1.0 / 3.0, a2 = 2.0 / 3.0, b3 = 2.0 / 3.0;
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < n; q++) out[q] = a0 * Gn[q] + a2 * Gs[q] + b3 * dt * rhs[q];
    }
}

/* ---- ghost fill (boundary.h) --------------------------------------------- */

/* Axis order x, then z, then y.  Periodic and zero-gradient fills are pure
 * index maps, so their order does not change any value; linear extrapolation
 * (reference: y axis only, boundary.h:177-200) runs last so that corner
 * ghosts extrapolate x/z-filled columns, as the reference's
 * exchange -> x -> y order does (boundary.h:217-229). */
static void fill_axis(const lso_grid* g, double* G, int axis) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g), nzt = nzt_(g);
    const int ng = g->ng;
    int64_t n, stride, o1n, o1s, o2n, o2s;
    if (axis == 0) { n = g->nx; stride = 1; o1n = nyt; o1s = nxt; o2n = nzt; o2s = nxt * nyt; }
    else if (axis == 1) { n = g->ny; stride = nxt; o1n = nxt; o1s = 1; o2n = nzt; o2s = nxt * nyt; }
    else { n = g->nz; stride = nxt * nyt; o1n = nxt; o1s = 1; o2n = nyt; o2s = nxt; }
    const int bc = g->bc[axis];
    for (int64_t b = 0; b < o2n; b++) {
        for (int64_t a = 0; a < o1n; a++) {
            double* L = G + a * o1s + b * o2s;
            for (int q = 0; q < ng; q++) {
                if (bc == LSO_BC_PERIODIC) {
                    /* boundary.h:98-114: ghost q <- n+q, ghost n+ng+q <- ng+q */
                    L[q * stride] = L[(n + q) * stride];
                    L[(n + ng + q) * stride] = L[(ng + q) * stride];
                } else if (bc == LSO_BC_ZERO_GRAD) {
                    /* boundary.h:126-168 */
                    L[q * stride] = L[ng * stride];
                    L[(n + ng + q) * stride] = L[(n + ng - 1) * stride];
                } else {
                    /* boundary.h:177-200 */
                    double b0 = L[ng * stride], b1 = L[(ng + 1) * stride];
                    double sb = b0 - b1;
                    L[(ng - 1 - q) * stride] = b0 + (q + 1) * sb;
                    double t0 = L[(n + ng - 1) * stride], t1 = L[(n + ng - 2) * stride];
                    double st = t0 - t1;
                    L[(n + ng + q) * stride] = t0 + (q + 1) * st;
                }
            }
        }
    }
}

void lso_fill_ghosts(const lso_grid* g, double* G) {
    fill_axis(g, G, 0);
    if (g->dim == 3) fill_axis(g, G, 2);
    fill_axis(g, G, 1);
}

/* ---- one RK3 step (rk3.h:168-187) ---------------------------------------- */

/* work: 3 padded fields (G1, G2, rhs).  rhs ghosts are zeroed here, matching
 * the reference's zero-initialised allocation. */
void lso_rk3_step(const lso_grid* g, const double* Gn, double* Gnew, const double* u,
                  const double* v, const double* w, const double* sl, double dt,
                  double* work) {
    const int64_t n = lso_padded_size(g);
    double* G1 = work;
    double* G2 = work + n;
    double* R = work + 2 * n;
    memset(R, 0, (size_t)n * sizeof(double));
    lso_rhs(g, Gn, R, u, v, w, sl);
    lso_stage(1, n, Gn, NULL, R, G1, dt);
    lso_fill_ghosts(g, G1);
    lso_rhs(g, G1, R, u, v, w, sl);
    lso_stage(2, n, Gn, G1, R, G2, dt);
    lso_fill_ghosts(g, G2);
    lso_rhs(g, G2, R, u, v, w, sl);
    lso_stage(3, n, Gn, G2, R, Gnew, dt);
    lso_fill_ghosts(g, Gnew);
}

/* ---- HCR-2 (reinitialization.h) ------------------------------------------ */

/* reinitialization.h:75-130 (3D :46-89).  Flags: 1 x-, 2 x+, 4 y-, 8 y+,
 * 16 z-, 32 z+. */
void lso_crossings(const lso_grid* g, const double* p0, double* rt, int32_t* flag) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g);
    const int64_t sx = 1, sy = nxt, sz = nxt * nyt;
    const int ng = g->ng;
    const int kz0 = g->dim == 3 ? ng : 0;
    const int kz1 = g->dim == 3 ? g->nz + ng : 1;
    const int nnb = g->dim == 3 ? 6 : 4;
    const int64_t off[6] = {-sx, sx, -sy, sy, -sz, sz};
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = kz0; k < kz1; k++) {
        for (int j = ng; j < g->ny + ng; j++) {
            for (int i = ng; i < g->nx + ng; i++) {
                int64_t c = (int64_t)k * sz + (int64_t)j * sy + i;
                double pc = p0[c];
                double sum = 0.0;
                int cnt = 0, f = 0;
                for (int q = 0; q < nnb; q++) {
                    double pn = p0[c + off[q]];
                    if (pc * pn < 0.0) { sum += pn; cnt++; f |= 1 << q; }
                }
                rt[c] = 0.0;
                flag[c] = f;
                if (cnt > 0 && fabs(sum) > 1e-15) rt[c] = pc / sum;
            }
        }
    }
}

/* weno5.h:234-267 (3D :202-248): first-order Godunov |grad phi|. */
static inline double lso_godunov(const double* P, int dim, int64_t sx, int64_t sy, int64_t sz,
                                 double sgn, double dx, double dy, double dz) {
    double xm = (P[0] - P[-sx]) / dx, xp = (P[sx] - P[0]) / dx;
    double ym = (P[0] - P[-sy]) / dy, yp = (P[sy] - P[0]) / dy;
    double zm = 0.0, zp = 0.0;
    if (dim == 3) { zm = (P[0] - P[-sz]) / dz; zp = (P[sz] - P[0]) / dz; }
    double ax, bx, ay, by, az, bz;
    if (sgn > 0.0) {
        ax = fmax(xm, 0.0); bx = fmin(xp, 0.0);
        ay = fmax(ym, 0.0); by = fmin(yp, 0.0);
        az = fmax(zm, 0.0); bz = fmin(zp, 0.0);
    } else {
        ax = fmin(xm, 0.0); bx = fmax(xp, 0.0);
        ay = fmin(ym, 0.0); by = fmax(yp, 0.0);
        az = fmin(zm, 0.0); bz = fmax(zp, 0.0);
    }
    double s;
    if (dim == 3) s = fmax(ax * ax, bx * bx) + fmax(ay * ay, by * by) + fmax(az * az, bz * bz);
    else s = fmax(ax * ax, bx * bx) + fmax(ay * ay, by * by);
    return sqrt(s);
}

/* reinitialization.h:200-263 (3D :136-193): one pseudo-time step over all
 * interior cells.  dx is used for the smoothed sign and the forcing in 3D
 * as well, as in the reference. */
void lso_reinit_step(const lso_grid* g, const double* pn, double* pnew, const double* p0,
                     const double* rt, const int32_t* flag) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g);
    const int64_t sx = 1, sy = nxt, sz = nxt * nyt;
    const int ng = g->ng;
    const int kz0 = g->dim == 3 ? ng : 0;
    const int kz1 = g->dim == 3 ? g->nz + ng : 1;
    const int nnb = g->dim == 3 ? 6 : 4;
    const int64_t off[6] = {-sx, sx, -sy, sy, -sz, sz};
    const double dtau = g->reinit_dtau;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = kz0; k < kz1; k++) {
        for (int j = ng; j < g->ny + ng; j++) {
            for (int i = ng; i < g->nx + ng; i++) {
                int64_t c = (int64_t)k * sz + (int64_t)j * sy + i;
                double phi = pn[c];
                double phi0 = p0[c];
                double S = phi0 / sqrt(phi0 * phi0 + g->dx * g->dx);
                double gm = lso_godunov(pn + c, g->dim, sx, sy, sz, S, g->dx, g->dy, g->dz);
                double term = -S * (gm - 1.0);
                double forcing = 0.0;
                int f = flag[c];
                if (f != 0) {
                    /* stability constraint, reinitialization.h:152-189 */
                    int stable = ((pn[c] > 0.0) == (p0[c] > 0.0));
                    for (int q = 0; q < nnb && stable; q++)
                        if (f & (1 << q)) stable = ((pn[c + off[q]] > 0.0) == (p0[c + off[q]] > 0.0));
                    if (stable) {
                        double sum = 0.0;
                        for (int q = 0; q < nnb; q++)
                            if (f & (1 << q)) sum += pn[c + off[q]];
                        double r = rt[c];
                        forcing = g->reinit_beta * (r * sum - phi) / g->dx;
                    }
                }
                pnew[c] = phi + dtau * (term + forcing);
            }
        }
    }
}

/* reinitializeWithSwap (reinitialization.h:321-360, 3D :199-228).  The result
 * is written back into G.  work: 3 padded doubles (phi_tmp, phi_0, r_tilde);
 * flag: padded int32. */
void lso_hcr2(const lso_grid* g, double* G, double* work, int32_t* flag) {
    const int64_t n = lso_padded_size(g);
    double* tmp = work;
    double* p0 = work + n;
    double* rt = work + 2 * n;
    memcpy(p0, G, (size_t)n * sizeof(double));
    lso_fill_ghosts(g, p0);
    lso_crossings(g, p0, rt, flag);
    double* a = G;
    double* b = tmp;
    for (int it = 0; it < g->reinit_iters; it++) {
        lso_reinit_step(g, a, b, p0, rt, flag);
        double* t = a; a = b; b = t;
        lso_fill_ghosts(g, a);
    }
    if (a != G) memcpy(G, a, (size_t)n * sizeof(double));
}

/* ---- velocity for the 3D sphere-deformation test -------------------------
 * initial_conditions.h:232-273 (3D): u = 2 sin^2(pi x) sin(2 pi y) sin(2 pi z) c,
 * v = -sin(2 pi x) sin^2(pi y) sin(2 pi z) c, w = -sin(2 pi x) sin(2 pi y) sin^2(pi z) c,
 * c = cos(pi t / T), evaluated on the whole padded array at x = x0 + (i-3) dx. */
void lso_deformation_velocity(const lso_grid* g, double x0, double y0, double z0, double t,
                              double T, double* u, double* v, double* w) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g), nzt = nzt_(g);
    const double PI = 3.14159265358979323846;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nzt; k++) {
        for (int64_t j = 0; j < nyt; j++) {
            for (int64_t i = 0; i < nxt; i++) {
                double x = x0 + (double)((int)i - g->ng) * g->dx;
                double y = y0 + (double)((int)j - g->ng) * g->dy;
                double z = z0 + (double)((int)k - g->ng) * g->dz;
                double tf = cos(PI * t / T);
                double spx = sin(PI * x), s2px = sin(2.0 * PI * x);
                double spy = sin(PI * y), s2py = sin(2.0 * PI * y);
                double spz = sin(PI * z), s2pz = sin(2.0 * PI * z);
                int64_t c = (k * nyt + j) * nxt + i;
                u[c] = 2.0 * spx * spx * s2py * s2pz * tf;
                v[c] = -s2px * spy * spy * s2pz * tf;
                w[c] = -s2px * s2py * spz * spz * tf;
            }
        }
    }
}

/* ---- time loop (main.cpp:310-354, 3D :287-334) -----------------------------
 * Runs while t < t_final (and step < max_steps): clamps the last dt to land
 * on t_final, advances one RK3 step, runs HCR-2 when reinit_interval > 0,
 * step > 0 and step % reinit_interval == 0 (step counted before increment),
 * then t += dt, step++.  deform_T > 0 recomputes the 3D deformation velocity
 * at the start of every step (3D main.cpp:296-299).  Returns the step count;
 * *t_out receives the final time and *dt_last the last dt used.
 * work: 6 padded doubles; flag: padded int32. */
int64_t lso_advance(const lso_grid* g, double* G, double* u, double* v, double* w,
                    const double* sl, double dt, double t_final, int64_t max_steps,
                    int32_t reinit_interval, double deform_T, double* work, int32_t* flag,
                    double* t_out, double* dt_last) {
    const int64_t n = lso_padded_size(g);
    double* Gnew = work + 3 * n;
    double t = 0.0;
    int64_t step = 0;
    while (t < t_final && step < max_steps) {
        if (t + dt > t_final) dt = t_final - t;
        if (deform_T > 0.0) lso_deformation_velocity(g, 0.0, 0.0, 0.0, t, deform_T, u, v, w);
        lso_rk3_step(g, G, Gnew, u, v, w, sl, dt, work);
        memcpy(G, Gnew, (size_t)n * sizeof(double));
        if (reinit_interval > 0 && step > 0 && step % reinit_interval == 0)
            lso_hcr2(g, G, work, flag);
        t += dt;
        step++;
    }
    if (t_out) *t_out = t;
    if (dt_last) *dt_last = dt;
    return step;
}

double lso_l2(const lso_grid* g, const double* G, const double* G0);
double lso_area(const lso_grid* g, const double* G);

/* The reference's timed region (Local_code/level-set_MPI_2D/src/main.cpp:308-357,
 * 3D :285-336) on `nthreads` OpenMP threads, for the CPU baseline: nsteps RK3
 * steps of a fixed dt, HCR-2 after step > 0 with step % reinit_interval == 0,
 * and every diag_interval steps the in-loop diagnostics the reference computes
 * inside its timed region (L2 against G0 and the area / volume); no file
 * output.  Returns the wall seconds of the loop (omp_get_wtime, the analogue
 * of the reference's MPI_Wtime; with one process there is no max over ranks);
 * diag[0], diag[1] receive the last L2 and area.  work: 6 padded doubles;
 * flag: padded int32. */
double lso_timed_loop(const lso_grid* g, double* G, const double* G0, double* u, double* v, double* w,
                      double dt, int64_t nsteps, int32_t reinit_interval, int32_t diag_interval,
                      int32_t nthreads, double* work, int32_t* flag, double* diag) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const double t0 = omp_get_wtime();
#else
    (void)nthreads;
    const double t0 = 0.0;
#endif
    const int64_t n = lso_padded_size(g);
    double* Gnew = work + 3 * n;
    diag[0] = diag[1] = 0.0;
    for (int64_t step = 0; step < nsteps; step++) {
        lso_rk3_step(g, G, Gnew, u, v, w, NULL, dt, work);
        memcpy(G, Gnew, (size_t)n * sizeof(double));
        if (reinit_interval > 0 && step > 0 && step % reinit_interval == 0) lso_hcr2(g, G, work, flag);
        if (diag_interval > 0 && (step + 1) % diag_interval == 0) {
            diag[0] = lso_l2(g, G, G0);
            diag[1] = lso_area(g, G);
        }
    }
#ifdef _OPENMP
    return omp_get_wtime() - t0;
#else
    return t0;
#endif
}

/* ---- diagnostics (io.h:194-258) ------------------------------------------ */

double lso_l2(const lso_grid* g, const double* G, const double* G0) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g);
    const int ng = g->ng;
    const int kz0 = g->dim == 3 ? ng : 0;
    const int kz1 = g->dim == 3 ? g->nz + ng : 1;
    double sum = 0.0;
    int64_t cnt = 0;
    for (int k = kz0; k < kz1; k++)
        for (int j = ng; j < g->ny + ng; j++)
            for (int i = ng; i < g->nx + ng; i++) {
                int64_t c = ((int64_t)k * nyt + j) * nxt + i;
                double d = G[c] - G0[c];
                sum += d * d;
                cnt++;
            }
    return sqrt(sum / (double)cnt);
}

/* Area (2D) or volume (3D) of G < 0, accumulated cell by cell as io.h:225-246. */
double lso_area(const lso_grid* g, const double* G) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g);
    const int ng = g->ng;
    const int kz0 = g->dim == 3 ? ng : 0;
    const int kz1 = g->dim == 3 ? g->nz + ng : 1;
    const double cell = g->dim == 3 ? g->dx * g->dy * g->dz : g->dx * g->dy;
    double acc = 0.0;
    for (int k = kz0; k < kz1; k++)
        for (int j = ng; j < g->ny + ng; j++)
            for (int i = ng; i < g->nx + ng; i++)
                if (G[((int64_t)k * nyt + j) * nxt + i] < 0.0) acc += cell;
    return acc;
}

/* ---- G output (io.h:34-114, 3D io.h:34-136) --------------------------------
 * int32 header {nx, ny, nghost} (2D) or {nx, ny, nz, nghost} (3D), then the
 * padded field with the y-ghost rows replaced by the adjacent interior row. */
int lso_write_G(const lso_grid* g, const double* G, const char* path) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g), nzt = nzt_(g);
    const int ng = g->ng;
    FILE* fp = fopen(path, "wb");
    if (!fp) return -1;
    int32_t hdr[4];
    int nh;
    if (g->dim == 3) { hdr[0] = g->nx; hdr[1] = g->ny; hdr[2] = g->nz; hdr[3] = ng; nh = 4; }
    else { hdr[0] = g->nx; hdr[1] = g->ny; hdr[2] = ng; nh = 3; }
    fwrite(hdr, sizeof(int32_t), (size_t)nh, fp);
    double* row = (double*)malloc((size_t)nxt * sizeof(double));
    for (int64_t k = 0; k < nzt; k++) {
        for (int64_t j = 0; j < nyt; j++) {
            int64_t js = j;
            if (j < ng) js = ng;
            if (j >= g->ny + ng) js = g->ny + ng - 1;
            memcpy(row, G + (k * nyt + js) * nxt, (size_t)nxt * sizeof(double));
            fwrite(row, sizeof(double), (size_t)nxt, fp);
        }
    }
    free(row);
    fclose(fp);
    return 0;
}

/* saveFieldVTK (Local_code/level-set_MPI_2D/include/io.h:119-180) and
 * saveFieldVTK3D (Local_code/level-set_MPI_3D/include/io.h:145-215): legacy
 * ASCII VTK structured points of the interior, x fastest. */
int lso_write_vtk(const lso_grid* g, const double* G, double x0, double y0, double z0, const char* path) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g);
    const int ng = g->ng;
    FILE* fp = fopen(path, "w");
    if (!fp) return -1;
    fprintf(fp, "# vtk DataFile Version 3.0\n");
    fprintf(fp, g->dim == 3 ? "G-equation Level-Set Field 3D\n" : "G-equation Level-Set Field\n");
    fprintf(fp, "ASCII\n");
    fprintf(fp, "DATASET STRUCTURED_POINTS\n");
    if (g->dim == 3) {
        fprintf(fp, "DIMENSIONS %d %d %d\n", g->nx, g->ny, g->nz);
        fprintf(fp, "ORIGIN %f %f %f\n", x0, y0, z0);
        fprintf(fp, "SPACING %f %f %f\n", g->dx, g->dy, g->dz);
        fprintf(fp, "POINT_DATA %d\n", g->nx * g->ny * g->nz);
    } else {
        fprintf(fp, "DIMENSIONS %d %d 1\n", g->nx, g->ny);
        fprintf(fp, "ORIGIN %f %f 0\n", x0, y0);
        fprintf(fp, "SPACING %f %f 1\n", g->dx, g->dy);
        fprintf(fp, "POINT_DATA %d\n", g->nx * g->ny);
    }
    fprintf(fp, "SCALARS G double 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    const int nz = g->dim == 3 ? g->nz : 1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < g->ny; j++)
            for (int i = 0; i < g->nx; i++) {
                const int64_t kk = g->dim == 3 ? k + ng : 0;
                fprintf(fp, "%e\n", G[(kk * nyt + j + ng) * nxt + i + ng]);
            }
    fclose(fp);
    return 0;
}

/* ---- reference initial conditions (log known-answer cases) ----------------
 * Coordinates x = (i - ng) * dx on the unit box (config.h:278-294).
 * Zalesak disk: Local_code/level-set_MPI_2D/include/initial_conditions.h:91-117
 * with the rotation field of :209-227,260-281 (omega = 2 pi / T).
 * Sphere:       Local_code/level-set_MPI_3D/include/initial_conditions.h:52-58. */
static double lso_zalesak(double x, double y, double cx, double cy, double r, double sw, double sl) {
    double ddx = x - cx;
    double ddy = y - cy;
    double disk = sqrt(ddx * ddx + ddy * ddy) - r;
    double hw = sw / 2.0;
    double bot = cy - r;
    double top = bot + sl;
    double sx = fabs(x - cx) - hw;
    double sy = fmax(bot - y, y - top);
    double slot;
    if (sx < 0.0 && sy < 0.0) slot = fmax(sx, sy);
    else if (sx < 0.0) slot = sy;
    else if (sy < 0.0) slot = sx;
    else slot = sqrt(sx * sx + sy * sy);
    return fmax(disk, -slot);
}

void lso_ic_zalesak2d(const lso_grid* g, double T, double* G, double* u, double* v) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g);
    const double PI = 3.14159265358979323846;
    const double om = 2.0 * PI / T;
    for (int64_t j = 0; j < nyt; j++)
        for (int64_t i = 0; i < nxt; i++) {
            double x = 0.0 + (double)((int)i - g->ng) * g->dx;
            double y = 0.0 + (double)((int)j - g->ng) * g->dy;
            int64_t c = j * nxt + i;
            G[c] = lso_zalesak(x, y, 0.5, 0.75, 0.15, 0.05, 0.25);
            u[c] = -om * (y - 0.5);
            v[c] = om * (x - 0.5);
        }
}

void lso_ic_sphere3d(const lso_grid* g, double cx, double cy, double cz, double r, double* G) {
    const int64_t nxt = nxt_(g), nyt = nyt_(g), nzt = nzt_(g);
    for (int64_t k = 0; k < nzt; k++)
        for (int64_t j = 0; j < nyt; j++)
            for (int64_t i = 0; i < nxt; i++) {
                double x = 0.0 + (double)((int)i - g->ng) * g->dx;
                double y = 0.0 + (double)((int)j - g->ng) * g->dy;
                double z = 0.0 + (double)((int)k - g->ng) * g->dz;
                double a = x - cx, b = y - cy, d = z - cz;
                G[(k * nyt + j) * nxt + i] = sqrt(a * a + b * b + d * d) - r;
            }
}
