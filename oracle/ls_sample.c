/*
 * ls_sample.c -- CPU restatement of the reference's `sample/` 2D solver.
 *
 * TEST INFRASTRUCTURE ONLY (part of liblsoracle.so): the checker of the GPU
 * sample-scheme kernels; the product path never loads it.
 *
 * What it restates (juhoon196/Level-set-method-code, sample/):
 *   - smoothness indicator beta(f0,f1,f2) = 13/12 (f0-2f1+f2)^2 +
 *     1/4 (f0-4f1+3f2)^2, pow(.,2) evaluated as a product (sample/weno5_rk3.cpp:11-14)
 *   - weno5_left / weno5_right, linear weights (0.1,0.6,0.3) on both
 *     (sample/weno5_rk3.cpp:17-50)
 *   - x/y derivatives dG_L, dG_R (sample/weno5_rk3.cpp:53-87).  dG_R passes
 *     an already reversed stencil to weno5_right, which reverses it again, so
 *     dG_R at i equals dG_L at i+1 bit for bit; restated as written.
 *   - local Lax-Friedrichs Hamiltonian compute_geqn_rhs (sample/weno5_rk3.cpp:91-145)
 *   - RK3_step with the sample's stage forms 0.75 G + 0.25 (G1 + dt r) and
 *     1/3 G + 2/3 (G2 + dt r), anchored BC then periodic after every stage
 *     (sample/weno5_rk3.cpp:148-188)
 *   - Field2D::apply_periodic, apply_anchored_bc, reinitialize
 *     (sample/field.cpp:16-130)
 *   - compute_first_order_upwind (sample/first_order_upwind.cpp:6-71),
 *     including its loop bound that reaches the first ghost row and column
 *   - the Bunsen initial condition, the forced velocity and save_G_to_bin
 *     (sample/main.cpp:34-51, 95-116, 131-136)
 * Expressions keep the reference's C++ evaluation order; build with
 * -ffp-contract=off (the sample's own flags, -O3 on x86-64, do not contract).
 *
 * Layout: one padded array per field, (ny+6) rows of (nx+6) doubles, x
 * fastest: data[j][i] of the reference is a[j*(nx+6) + i].
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define LSS_WENO_EPS 1e-6
#define LSS_PI 3.14159265358979323846  /* M_PI */

static inline double sq(double x) { return x * x; }
static inline double dmax(double a, double b) { return (a < b) ? b : a; }  /* std::max */
static inline double dmin(double a, double b) { return (b < a) ? b : a; }  /* std::min */

double lss_beta(double f0, double f1, double f2) {
    return (13.0 / 12.0) * sq(f0 - 2.0 * f1 + f2) + (1.0 / 4.0) * sq(f0 - 4.0 * f1 + 3.0 * f2);
}

double lss_weno_left(double v0, double v1, double v2, double v3, double v4) {
    double b0 = lss_beta(v0, v1, v2);
    double b1 = lss_beta(v1, v2, v3);
    double b2 = lss_beta(v2, v3, v4);
    double a0 = 0.1 / ((LSS_WENO_EPS + b0) * (LSS_WENO_EPS + b0));
    double a1 = 0.6 / ((LSS_WENO_EPS + b1) * (LSS_WENO_EPS + b1));
    double a2 = 0.3 / ((LSS_WENO_EPS + b2) * (LSS_WENO_EPS + b2));
    double sum = a0 + a1 + a2;
    double w0 = a0 / sum, w1 = a1 / sum, w2 = a2 / sum;
    return w0 * (2.0 * v0 - 7.0 * v1 + 11.0 * v2) / 6.0 + w1 * (-v1 + 5.0 * v2 + 2.0 * v3) / 6.0 +
           w2 * (2.0 * v2 + 5.0 * v3 - v4) / 6.0;
}

double lss_weno_right(double v0, double v1, double v2, double v3, double v4) {
    double b0 = lss_beta(v4, v3, v2);
    double b1 = lss_beta(v3, v2, v1);
    double b2 = lss_beta(v2, v1, v0);
    double a0 = 0.1 / ((LSS_WENO_EPS + b0) * (LSS_WENO_EPS + b0));
    double a1 = 0.6 / ((LSS_WENO_EPS + b1) * (LSS_WENO_EPS + b1));
    double a2 = 0.3 / ((LSS_WENO_EPS + b2) * (LSS_WENO_EPS + b2));
    double sum = a0 + a1 + a2;
    double w0 = a0 / sum, w1 = a1 / sum, w2 = a2 / sum;
    return w0 * (2.0 * v4 - 7.0 * v3 + 11.0 * v2) / 6.0 + w1 * (-v3 + 5.0 * v2 + 2.0 * v1) / 6.0 +
           w2 * (2.0 * v2 + 5.0 * v1 - v0) / 6.0;
}

/* dG_L and dG_R along a stride s at the cell G points to (weno5_rk3.cpp:53-87). */
static inline void lss_derivs(const double* G, int64_t s, double h, double* dl, double* dr) {
    *dl = (lss_weno_left(G[-2 * s], G[-s], G[0], G[s], G[2 * s]) -
           lss_weno_left(G[-3 * s], G[-2 * s], G[-s], G[0], G[s])) / h;
    *dr = (lss_weno_right(G[3 * s], G[2 * s], G[s], G[0], G[-s]) -
           lss_weno_right(G[2 * s], G[s], G[0], G[-s], G[-2 * s])) / h;
}

/* compute_geqn_rhs: rhs on interior cells, zero elsewhere. */
void lss_rhs(int nx, int ny, const double* G, const double* u, const double* v, double* rhs,
             double dx, double dy, double S_L) {
    const int64_t nxt = nx + 6;
    memset(rhs, 0, (size_t)(ny + 6) * nxt * sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 3; j <= ny + 2; ++j) {
        for (int i = 3; i <= nx + 2; ++i) {
            const int64_t c = (int64_t)j * nxt + i;
            double Gdxm, Gdxp, Gdym, Gdyp;
            lss_derivs(G + c, 1, dx, &Gdxm, &Gdxp);
            lss_derivs(G + c, nxt, dy, &Gdym, &Gdyp);
            double Gdx_mean = 0.5 * (Gdxm + Gdxp);
            double Gdy_mean = 0.5 * (Gdym + Gdyp);
            double transport = u[c] * Gdx_mean + v[c] * Gdy_mean;
            double propagation = S_L * sqrt(Gdx_mean * Gdx_mean + Gdy_mean * Gdy_mean);
            double s_xm = u[c] - S_L * Gdxm / sqrt(Gdxm * Gdxm + Gdym * Gdym + 1e-14);
            double s_xp = u[c] - S_L * Gdxp / sqrt(Gdxp * Gdxp + Gdyp * Gdyp + 1e-14);
            double alpha_x = dmax(fabs(s_xm), fabs(s_xp));
            double s_ym = v[c] - S_L * Gdym / sqrt(Gdxm * Gdxm + Gdym * Gdym + 1e-14);
            double s_yp = v[c] - S_L * Gdyp / sqrt(Gdxp * Gdxp + Gdyp * Gdyp + 1e-14);
            double alpha_y = dmax(fabs(s_ym), fabs(s_yp));
            double alpha = dmax(alpha_x, alpha_y);
            double diss_x = alpha * (Gdxp - Gdxm) * 0.5;
            double diss_y = alpha * (Gdyp - Gdym) * 0.5;
            double H = transport - propagation - (diss_x + diss_y);
            rhs[c] = -H;
        }
    }
}

/* Field2D::apply_periodic (field.cpp:16-35): columns first over all rows,
 * then rows over all columns. */
void lss_apply_periodic(int nx, int ny, double* a) {
    const int64_t nxt = nx + 6, nyt = ny + 6;
    for (int64_t j = 0; j < nyt; ++j) {
        double* r = a + j * nxt;
        for (int k = 0; k < 3; ++k) {
            r[k] = r[nx + k];
            r[nx + 3 + k] = r[3 + k];
        }
    }
    for (int64_t i = 0; i < nxt; ++i) {
        for (int k = 0; k < 3; ++k) {
            a[k * nxt + i] = a[(ny + k) * nxt + i];
            a[(ny + 3 + k) * nxt + i] = a[(3 + k) * nxt + i];
        }
    }
}

/* Field2D::apply_anchored_bc (field.cpp:37-63).  Note that row 3 (the first
 * interior row) is overwritten by the initial field and row ny+2 (the last
 * interior row) by a linear extrapolation. */
void lss_apply_anchored(int nx, int ny, double* a, const double* init) {
    const int64_t nxt = nx + 6, nyt = ny + 6;
#define A(j, i) a[(int64_t)(j) * nxt + (i)]
    for (int64_t i = 0; i < nxt; ++i) {
        A(3, i) = init[3 * nxt + i];
        A(2, i) = 2.0 * A(3, i) - A(4, i);
        A(1, i) = 2.0 * A(2, i) - A(3, i);
        A(0, i) = 2.0 * A(1, i) - A(2, i);
    }
    for (int64_t i = 0; i < nxt; ++i) {
        A(ny + 2, i) = 2.0 * A(ny + 1, i) - A(ny, i);
        A(ny + 3, i) = 2.0 * A(ny + 2, i) - A(ny + 1, i);
        A(ny + 4, i) = 2.0 * A(ny + 3, i) - A(ny + 2, i);
        A(ny + 5, i) = 2.0 * A(ny + 4, i) - A(ny + 3, i);
    }
    for (int64_t j = 0; j < nyt; ++j) {
        A(j, 2) = 2.0 * A(j, 3) - A(j, 4);
        A(j, 1) = 2.0 * A(j, 2) - A(j, 3);
        A(j, 0) = 2.0 * A(j, 1) - A(j, 2);
        A(j, nx + 3) = 2.0 * A(j, nx + 2) - A(j, nx + 1);
        A(j, nx + 4) = 2.0 * A(j, nx + 3) - A(j, nx + 2);
        A(j, nx + 5) = 2.0 * A(j, nx + 4) - A(j, nx + 3);
    }
#undef A
}

/* RK3_step (weno5_rk3.cpp:148-188), G updated in place.  work: 3 padded
 * fields (G1, G2, rhs). */
void lss_rk3_step(int nx, int ny, double* G, const double* u, const double* v, double dx, double dy,
                  double dt, double S_L, const double* G_initial, double* work) {
    const int64_t nxt = nx + 6, n = (int64_t)(ny + 6) * nxt;
    double* G1 = work;
    double* G2 = work + n;
    double* rhs = work + 2 * n;
    memset(G1, 0, (size_t)n * sizeof(double));
    memset(G2, 0, (size_t)n * sizeof(double));
    lss_rhs(nx, ny, G, u, v, rhs, dx, dy, S_L);
    for (int j = 3; j <= ny + 2; ++j)
        for (int i = 3; i <= nx + 2; ++i) {
            const int64_t c = (int64_t)j * nxt + i;
            G1[c] = G[c] + dt * rhs[c];
        }
    lss_apply_anchored(nx, ny, G1, G_initial);
    lss_apply_periodic(nx, ny, G1);
    lss_rhs(nx, ny, G1, u, v, rhs, dx, dy, S_L);
    for (int j = 3; j <= ny + 2; ++j)
        for (int i = 3; i <= nx + 2; ++i) {
            const int64_t c = (int64_t)j * nxt + i;
            G2[c] = 0.75 * G[c] + 0.25 * (G1[c] + dt * rhs[c]);
        }
    lss_apply_anchored(nx, ny, G2, G_initial);
    lss_apply_periodic(nx, ny, G2);
    lss_rhs(nx, ny, G2, u, v, rhs, dx, dy, S_L);
    for (int j = 3; j <= ny + 2; ++j)
        for (int i = 3; i <= nx + 2; ++i) {
            const int64_t c = (int64_t)j * nxt + i;
            G[c] = (1.0 / 3.0) * G[c] + (2.0 / 3.0) * (G2[c] + dt * rhs[c]);
        }
    lss_apply_anchored(nx, ny, G, G_initial);
    lss_apply_periodic(nx, ny, G);
}

/* Field2D::reinitialize (field.cpp:66-130): first-order sign-distance
 * iterations with dtau = 0.2 min(dx,dy) and the anchored/extrapolated
 * boundary after each iteration.  work: 1 padded field. */
void lss_reinitialize(int nx, int ny, double* a, const double* init, double dx, double dy,
                      int num_iters, double* work) {
    const int64_t nxt = nx + 6, nyt = ny + 6, n = nyt * nxt;
    const double dtau = 0.2 * dmin(dx, dy);
    double* rhs = work;
#define P(j, i) a[(int64_t)(j) * nxt + (i)]
    for (int iter = 0; iter < num_iters; ++iter) {
        memset(rhs, 0, (size_t)n * sizeof(double));
        for (int j = 3; j <= ny + 2; ++j) {
            for (int i = 3; i <= nx + 2; ++i) {
                const double i0 = init[(int64_t)j * nxt + i];
                double s = i0 / sqrt(i0 * i0 + dx * dx);
                double dGdx_pos = (P(j, i) - P(j, i - 1)) / dx;
                double dGdx_neg = (P(j, i + 1) - P(j, i)) / dx;
                double dGdy_pos = (P(j, i) - P(j - 1, i)) / dy;
                double dGdy_neg = (P(j + 1, i) - P(j, i)) / dy;
                double grad;
                if (s > 0.0) {
                    double px = dmax(dmax(dGdx_pos, 0.0), -dmin(dGdx_neg, 0.0));
                    double py = dmax(dmax(dGdy_pos, 0.0), -dmin(dGdy_neg, 0.0));
                    grad = sqrt(px * px + py * py);
                } else {
                    double qx = dmax(dmax(-dGdx_pos, 0.0), dmin(dGdx_neg, 0.0));
                    double qy = dmax(dmax(-dGdy_pos, 0.0), dmin(dGdy_neg, 0.0));
                    grad = sqrt(qx * qx + qy * qy);
                }
                rhs[(int64_t)j * nxt + i] = -s * (grad - 1.0);
            }
        }
        for (int j = 3; j <= ny + 2; ++j)
            for (int i = 3; i <= nx + 2; ++i) P(j, i) += dtau * rhs[(int64_t)j * nxt + i];
        for (int64_t i = 0; i < nxt; ++i) {
            P(3, i) = init[3 * nxt + i];
            P(2, i) = 2.0 * P(3, i) - P(4, i);
            P(1, i) = 2.0 * P(2, i) - P(3, i);
            P(0, i) = 2.0 * P(1, i) - P(2, i);
        }
        for (int64_t i = 0; i < nxt; ++i) {
            P(ny + 2, i) = 2.0 * P(ny + 1, i) - P(ny, i);
            P(ny + 3, i) = 2.0 * P(ny + 2, i) - P(ny + 1, i);
            P(ny + 4, i) = 2.0 * P(ny + 3, i) - P(ny + 2, i);
            P(ny + 5, i) = 2.0 * P(ny + 4, i) - P(ny + 3, i);
        }
        for (int64_t j = 0; j < nyt; ++j) {
            P(j, 2) = 2.0 * P(j, 3) - P(j, 4);
            P(j, 1) = 2.0 * P(j, 2) - P(j, 3);
            P(j, 0) = 2.0 * P(j, 1) - P(j, 2);
            P(j, nx + 3) = 2.0 * P(j, nx + 2) - P(j, nx + 1);
            P(j, nx + 4) = 2.0 * P(j, nx + 3) - P(j, nx + 2);
            P(j, nx + 5) = 2.0 * P(j, nx + 4) - P(j, nx + 3);
        }
    }
#undef P
}

/* compute_first_order_upwind (first_order_upwind.cpp:6-71): explicit Euler
 * step over rows/columns 3 .. n+3 (one ghost layer beyond the interior on
 * the high side, as the reference's loop bounds give).  work: 1 field. */
void lss_first_order_upwind(int nx, int ny, double* G, const double* u, const double* v, double dx,
                            double dy, double dt, double S_L, double* work) {
    const int64_t nxt = nx + 6, n = (int64_t)(ny + 6) * nxt;
    const double EPS = 1e-10;
    double* Gn = work;
    memcpy(Gn, G, (size_t)n * sizeof(double));
    for (int j = 3; j <= ny + 3; ++j) {
        for (int i = 3; i <= nx + 3; ++i) {
            const int64_t c = (int64_t)j * nxt + i;
            double Gx_mns = (G[c] - G[c - 1]) / dx;
            double Gx_pls = (G[c + 1] - G[c]) / dx;
            double Gy_mns = (G[c] - G[c - nxt]) / dy;
            double Gy_pls = (G[c + nxt] - G[c]) / dy;
            double dG_dx = 0.5 * (Gx_mns + Gx_pls);
            double dG_dy = 0.5 * (Gy_mns + Gy_pls);
            double mag_sq = dG_dx * dG_dx + dG_dy * dG_dy;
            double uf, vf;
            if (mag_sq > EPS * EPS) {
                double mag = sqrt(mag_sq);
                uf = u[c] - S_L * dG_dx / mag;
                vf = v[c] - S_L * dG_dy / mag;
            } else {
                uf = u[c];
                vf = v[c];
            }
            double a_pls = dmax(uf, 0.0), a_mns = dmin(uf, 0.0);
            double b_pls = dmax(vf, 0.0), b_mns = dmin(vf, 0.0);
            Gn[c] = G[c] - dt * (a_pls * Gx_mns + a_mns * Gx_pls + b_pls * Gy_mns + b_mns * Gy_pls);
        }
    }
    for (int j = 3; j <= ny + 3; ++j)
        for (int i = 3; i <= nx + 3; ++i) G[(int64_t)j * nxt + i] = Gn[(int64_t)j * nxt + i];
}

/* The Bunsen-flame initial field of sample/main.cpp:95-116 on the whole
 * padded array (G = -distance, x = (i-3) dx, y = (j-3) dy). */
void lss_bunsen(int nx, int ny, double dx, double dy, double R, double flame_angle_deg, double X0,
                double Y0, double* G) {
    const int64_t nxt = nx + 6;
    const double slope = tan((90.0 - flame_angle_deg) * LSS_PI / 180.0);
    for (int j = 0; j < ny + 6; ++j)
        for (int i = 0; i < nx + 6; ++i) {
            double x = (i - 3) * dx;
            double y = (j - 3) * dy;
            double distance;
            if (x < X0)
                distance = slope * (x - (X0 - R)) - (y - Y0);
            else
                distance = -slope * (x - (X0 + R)) - (y - Y0);
            G[(int64_t)j * nxt + i] = -distance;
        }
}

/* save_G_to_bin (sample/main.cpp:34-51): interior rows, no header. */
int lss_write_G(int nx, int ny, const double* G, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    const int64_t nxt = nx + 6;
    for (int j = 3; j <= ny + 2; ++j)
        if (fwrite(G + (int64_t)j * nxt + 3, sizeof(double), (size_t)nx, f) != (size_t)nx) {
            fclose(f);
            return -1;
        }
    return fclose(f) == 0 ? 0 : -1;
}
