/*
 * ls_tangent.cpp -- forward-mode (tangent) restatement of the 3D solver.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU discrete adjoint.  Every
 * operator of ls_oracle.c's 3D path is written once over a scalar type T and
 * run with T = Dual (value, directional derivative), so that one forward run
 * gives the loss L and dL/de along a direction e = (dG0, du, dv, dw, dS_L).
 * The value part performs the same double operations in the same order as
 * ls_oracle.c (built with -ffp-contract=off), so G_T is bitwise the oracle's.
 *
 * Restated from (juhoon196/Level-set-method-code, Local_code/level-set_MPI_3D):
 *   weno5.h:28-81      smoothness indicators, left/right WENO-5 faces
 *   weno5.h:90-196     upwinded axis derivatives (branch on the value of u_eff)
 *   rk3.h:31-81        u_eff RHS, rk3.h:87-150 stage combines
 *   boundary.h:37-261  periodic / zero-gradient ghost fill
 *   reinitialization.h:46-228, weno5.h:202-248   HCR-2
 *   src/main.cpp:287-334   step loop (reinit after step n when n > 0 and
 *                          n % ri == 0)
 * Non-smooth points take one fixed side, the same one the GPU adjoint takes:
 * max(x, 0) and min(x, 0) pass x only when strictly beyond 0; the Godunov
 * axis term takes a^2 when a^2 >= b^2; sqrt has derivative 0 at 0.  The loss
 * is sum over interior cells of (H_e(G_T) - H_e(G_target))^2 with the
 * smoothed Heaviside H_e.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

struct Dual {
    double v, d;
    Dual() : v(0.0), d(0.0) {}
    Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants promote
    Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
inline Dual sqrt(Dual a) {
    const double r = ::sqrt(a.v);
    return Dual(r, r > 0.0 ? a.d / (2.0 * r) : 0.0);
}
inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
inline double val(Dual a) { return a.v; }
inline double val(double a) { return a; }
inline double fabs_(double a) { return ::fabs(a); }

// Same-layout grid description as lso_grid in ls_oracle.c.
struct Grid {
    int32_t dim, nx, ny, nz, ng;
    int32_t bc[3];
    double dx, dy, dz, s_l, eps;
    int32_t reinit_iters, pad0;
    double reinit_dtau, reinit_beta;
};

constexpr double kWenoEps = 1.0e-6;
constexpr double kPi = 3.14159265358979323846;

template <class T>
struct Solver {
    const Grid& g;
    int64_t nxt, nyt, nzt, sy, sz, n;
    explicit Solver(const Grid& gr) : g(gr) {
        nxt = g.nx + 2 * g.ng;
        nyt = g.ny + 2 * g.ng;
        nzt = g.nz + 2 * g.ng;
        sy = nxt;
        sz = nxt * nyt;
        n = sz * nzt;
    }

    static void beta(const T* v, T* b) {
        T d0 = v[2] - 2.0 * v[1] + v[0];
        b[0] = (13.0 / 12.0) * d0 * d0 + 0.25 * (v[2] - v[0]) * (v[2] - v[0]);
        T d1 = v[3] - 2.0 * v[2] + v[1];
        b[1] = (13.0 / 12.0) * d1 * d1 + 0.25 * (v[3] - v[1]) * (v[3] - v[1]);
        T d2 = v[4] - 2.0 * v[3] + v[2];
        b[2] = (13.0 / 12.0) * d2 * d2 + 0.25 * (v[4] - v[2]) * (v[4] - v[2]);
    }
    static T weno_left(const T* v) {
        T b[3];
        beta(v, b);
        T a0 = 0.1 / ((kWenoEps + b[0]) * (kWenoEps + b[0]));
        T a1 = 0.6 / ((kWenoEps + b[1]) * (kWenoEps + b[1]));
        T a2 = 0.3 / ((kWenoEps + b[2]) * (kWenoEps + b[2]));
        T s = a0 + a1 + a2;
        T w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
        T p0 = (2.0 * v[0] - 7.0 * v[1] + 11.0 * v[2]) / 6.0;
        T p1 = (-v[1] + 5.0 * v[2] + 2.0 * v[3]) / 6.0;
        T p2 = (2.0 * v[2] + 5.0 * v[3] - v[4]) / 6.0;
        return w0 * p0 + w1 * p1 + w2 * p2;
    }
    static T weno_right(const T* v) {
        T m[5] = {v[4], v[3], v[2], v[1], v[0]};
        T b[3];
        beta(m, b);
        T a0 = 0.3 / ((kWenoEps + b[0]) * (kWenoEps + b[0]));
        T a1 = 0.6 / ((kWenoEps + b[1]) * (kWenoEps + b[1]));
        T a2 = 0.1 / ((kWenoEps + b[2]) * (kWenoEps + b[2]));
        T s = a0 + a1 + a2;
        T w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
        T p0 = (11.0 * v[2] - 7.0 * v[3] + 2.0 * v[4]) / 6.0;
        T p1 = (2.0 * v[1] + 5.0 * v[2] - v[3]) / 6.0;
        T p2 = (-v[0] + 5.0 * v[1] + 2.0 * v[2]) / 6.0;
        return w0 * p0 + w1 * p1 + w2 * p2;
    }
    static T deriv(const T* G, int64_t s, T vel, double h) {
        T v[5] = {G[-2 * s], G[-s], G[0], G[s], G[2 * s]};
        if (val(vel) >= 0.0) {
            T vl[5] = {G[-3 * s], G[-2 * s], G[-s], G[0], G[s]};
            T gl = weno_left(vl);
            T gr = weno_left(v);
            return (gr - gl) / h;
        }
        T vr[5] = {G[-s], G[0], G[s], G[2 * s], G[3 * s]};
        T gl = weno_right(v);
        T gr = weno_right(vr);
        return (gr - gl) / h;
    }

    int64_t at(int i, int j, int k) const { return (int64_t)k * sz + (int64_t)j * sy + i; }

    void rhs(const T* G, T* R, const T* u, const T* v, const T* w, const T* sl) const {
        const int ng = g.ng;
        for (int k = ng; k < g.nz + ng; k++)
            for (int j = ng; j < g.ny + ng; j++)
                for (int i = ng; i < g.nx + ng; i++) {
                    const int64_t c = at(i, j, k);
                    const T* P = G + c;
                    T s_l = sl ? sl[c] : T(g.s_l);
                    T gx = 0.0, gy = 0.0, gz = 0.0, gm = 1.0;
                    if (val(s_l) > g.eps) {
                        gx = (P[1] - P[-1]) / (2.0 * g.dx);
                        gy = (P[sy] - P[-sy]) / (2.0 * g.dy);
                        gz = (P[sz] - P[-sz]) / (2.0 * g.dz);
                        gm = sqrt(gx * gx + gy * gy + gz * gz + g.eps);
                    }
                    T ue = u[c] - s_l * gx / gm;
                    T ve = v[c] - s_l * gy / gm;
                    T we = w[c] - s_l * gz / gm;
                    T dgx = deriv(P, 1, ue, g.dx);
                    T dgy = deriv(P, sy, ve, g.dy);
                    T dgz = deriv(P, sz, we, g.dz);
                    R[c] = -(ue * dgx + ve * dgy + we * dgz);
                }
    }

    void stage(int st, const T* Gn, const T* Gs, const T* R, T* out, double dt) const {
        if (st == 1) {
            for (int64_t q = 0; q < n; q++) out[q] = Gn[q] + dt * R[q];
        } else if (st == 2) {
            for (int64_t q = 0; q < n; q++) out[q] = 0.75 * Gn[q] + 0.25 * Gs[q] + 0.25 * dt * R[q];
        } else {
            const double a0 = 1.0 / 3.0, a2 = 2.0 / 3.0, b3 = 2.0 / 3.0;
            for (int64_t q = 0; q < n; q++) out[q] = a0 * Gn[q] + a2 * Gs[q] + b3 * dt * R[q];
        }
    }

    void fill_axis(T* G, int axis) const {
        const int ng = g.ng;
        int64_t len, stride, o1n, o1s, o2n, o2s;
        if (axis == 0) { len = g.nx; stride = 1; o1n = nyt; o1s = nxt; o2n = nzt; o2s = sz; }
        else if (axis == 1) { len = g.ny; stride = nxt; o1n = nxt; o1s = 1; o2n = nzt; o2s = sz; }
        else { len = g.nz; stride = sz; o1n = nxt; o1s = 1; o2n = nyt; o2s = nxt; }
        const bool per = g.bc[axis] == 0;
        for (int64_t b = 0; b < o2n; b++)
            for (int64_t a = 0; a < o1n; a++) {
                T* L = G + a * o1s + b * o2s;
                for (int q = 0; q < ng; q++) {
                    if (per) {
                        L[q * stride] = L[(len + q) * stride];
                        L[(len + ng + q) * stride] = L[(ng + q) * stride];
                    } else {
                        L[q * stride] = L[ng * stride];
                        L[(len + ng + q) * stride] = L[(len + ng - 1) * stride];
                    }
                }
            }
    }
    void fill(T* G) const {
        fill_axis(G, 0);
        fill_axis(G, 2);
        fill_axis(G, 1);
    }

    void rk3(const T* Gn, T* Gnew, const T* u, const T* v, const T* w, const T* sl, double dt,
             std::vector<T>& work) const {
        T* G1 = work.data();
        T* G2 = G1 + n;
        T* R = G2 + n;
        for (int64_t q = 0; q < n; q++) R[q] = T(0.0);
        rhs(Gn, R, u, v, w, sl);
        stage(1, Gn, nullptr, R, G1, dt);
        fill(G1);
        rhs(G1, R, u, v, w, sl);
        stage(2, Gn, G1, R, G2, dt);
        fill(G2);
        rhs(G2, R, u, v, w, sl);
        stage(3, Gn, G2, R, Gnew, dt);
        fill(Gnew);
    }

    // max(a,0) / min(a,0) pass a only when strictly beyond 0
    static T pos(T a) { return val(a) > 0.0 ? a : T(0.0); }
    static T neg(T a) { return val(a) < 0.0 ? a : T(0.0); }
    static T bigger(T a2, T b2) { return val(a2) >= val(b2) ? a2 : b2; }

    T godunov(const T* P, double sgn) const {
        T xm = (P[0] - P[-1]) / g.dx, xp = (P[1] - P[0]) / g.dx;
        T ym = (P[0] - P[-sy]) / g.dy, yp = (P[sy] - P[0]) / g.dy;
        T zm = (P[0] - P[-sz]) / g.dz, zp = (P[sz] - P[0]) / g.dz;
        T ax, bx, ay, by, az, bz;
        if (sgn > 0.0) {
            ax = pos(xm); bx = neg(xp);
            ay = pos(ym); by = neg(yp);
            az = pos(zm); bz = neg(zp);
        } else {
            ax = neg(xm); bx = pos(xp);
            ay = neg(ym); by = pos(yp);
            az = neg(zm); bz = pos(zp);
        }
        T s = bigger(ax * ax, bx * bx) + bigger(ay * ay, by * by) + bigger(az * az, bz * bz);
        return sqrt(s);
    }

    void hcr2(T* G, std::vector<T>& work) const {
        T* tmp = work.data();
        T* p0 = tmp + n;
        T* rt = p0 + n;
        std::vector<int32_t> flag(n, 0);
        for (int64_t q = 0; q < n; q++) p0[q] = G[q];
        fill(p0);
        const int ng = g.ng;
        const int64_t off[6] = {-1, 1, -sy, sy, -sz, sz};
        for (int k = ng; k < g.nz + ng; k++)
            for (int j = ng; j < g.ny + ng; j++)
                for (int i = ng; i < g.nx + ng; i++) {
                    const int64_t c = at(i, j, k);
                    T pc = p0[c];
                    T sum = 0.0;
                    int cnt = 0, f = 0;
                    for (int q = 0; q < 6; q++) {
                        T pn = p0[c + off[q]];
                        if (val(pc) * val(pn) < 0.0) {
                            sum = sum + pn;
                            cnt++;
                            f |= 1 << q;
                        }
                    }
                    rt[c] = 0.0;
                    flag[c] = f;
                    if (cnt > 0 && fabs_(val(sum)) > 1e-15) rt[c] = pc / sum;
                }
        T* a = G;
        T* b = tmp;
        for (int it = 0; it < g.reinit_iters; it++) {
            for (int k = ng; k < g.nz + ng; k++)
                for (int j = ng; j < g.ny + ng; j++)
                    for (int i = ng; i < g.nx + ng; i++) {
                        const int64_t c = at(i, j, k);
                        T phi = a[c];
                        T phi0 = p0[c];
                        T S = phi0 / sqrt(phi0 * phi0 + g.dx * g.dx);
                        T gm = godunov(a + c, val(S));
                        T term = -S * (gm - 1.0);
                        T forcing = 0.0;
                        const int f = flag[c];
                        if (f != 0) {
                            bool stable = (val(a[c]) > 0.0) == (val(p0[c]) > 0.0);
                            for (int q = 0; q < 6 && stable; q++)
                                if (f & (1 << q)) stable = (val(a[c + off[q]]) > 0.0) == (val(p0[c + off[q]]) > 0.0);
                            if (stable) {
                                T sum = 0.0;
                                for (int q = 0; q < 6; q++)
                                    if (f & (1 << q)) sum = sum + a[c + off[q]];
                                forcing = g.reinit_beta * (rt[c] * sum - phi) / g.dx;
                            }
                        }
                        b[c] = phi + g.reinit_dtau * (term + forcing);
                    }
            T* t = a;
            a = b;
            b = t;
            fill(a);
        }
        if (a != G)
            for (int64_t q = 0; q < n; q++) G[q] = a[q];
    }
};

Dual heaviside(Dual x, double e) {
    if (x.v < -e) return Dual(0.0);
    if (x.v > e) return Dual(1.0);
    const double s = ::sin(kPi * x.v / e), c = ::cos(kPi * x.v / e);
    return Dual(0.5 * (1.0 + x.v / e + s / kPi), 0.5 * (1.0 + c) / e * x.d);
}

}  // namespace

extern "C" {

/* Runs nsteps RK3 steps (+ HCR-2 after step n when ri > 0, n > 0, n % ri == 0)
 * from G0 with tangents along (dG0, du, dv, dw, dsl) (each may be NULL = 0;
 * sl NULL = scalar s_l).  Writes the final field's values to G_out (may be
 * NULL), the loss to *loss and its directional derivative to *dloss. */
int lst_run(const Grid* g, const double* G0, const double* dG0, const double* u, const double* du,
            const double* v, const double* dv, const double* w, const double* dw, const double* sl,
            const double* dsl, const double* G_target, double eps_h, double dt, int32_t nsteps,
            int32_t ri, double* G_out, double* loss, double* dloss) {
    if (!g || g->dim != 3 || g->ng != 3) return 1;
    for (int a = 0; a < 3; a++)
        if (g->bc[a] != 0 && g->bc[a] != 1) return 1;
    Solver<Dual> S(*g);
    const int64_t n = S.n;
    auto mk = [&](const double* x, const double* dx) {
        std::vector<Dual> r(n);
        for (int64_t q = 0; q < n; q++) r[q] = Dual(x[q], dx ? dx[q] : 0.0);
        return r;
    };
    std::vector<Dual> G = mk(G0, dG0), U = mk(u, du), V = mk(v, dv), W = mk(w, dw);
    std::vector<Dual> SL;
    if (sl) SL = mk(sl, dsl);
    std::vector<Dual> Gn(n), work(3 * n);
    S.fill(G.data());
    for (int step = 0; step < nsteps; step++) {
        S.rk3(G.data(), Gn.data(), U.data(), V.data(), W.data(), sl ? SL.data() : nullptr, dt, work);
        G.swap(Gn);
        if (ri > 0 && step > 0 && step % ri == 0) S.hcr2(G.data(), work);
    }
    Dual L = 0.0;
    const int ng = g->ng;
    for (int k = ng; k < g->nz + ng; k++)
        for (int j = ng; j < g->ny + ng; j++)
            for (int i = ng; i < g->nx + ng; i++) {
                const int64_t c = S.at(i, j, k);
                Dual r = heaviside(G[c], eps_h) - heaviside(Dual(G_target[c]), eps_h);
                L = L + r * r;
            }
    if (G_out)
        for (int64_t q = 0; q < n; q++) G_out[q] = G[q].v;
    *loss = L.v;
    *dloss = L.d;
    return 0;
}

}  // extern "C"
