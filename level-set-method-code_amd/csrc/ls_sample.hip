// ls_sample.hip -- the reference's sample/ 2D G-equation scheme on the GPU.
//
// Scheme (sample/weno5_rk3.cpp, sample/field.cpp, sample/main.cpp):
//   * WENO-5 with beta(f0,f1,f2) = 13/12 (f0-2f1+f2)^2 + 1/4 (f0-4f1+3f2)^2 and
//     linear weights (0.1, 0.6, 0.3) (weno5_rk3.cpp:11-50);
//   * dG_L = (L(i-2..i+2) - L(i-3..i+1))/h and dG_R as the reference computes
//     it, which is bit for bit dG_L one cell to the right (weno5_right is
//     handed a reversed stencil and reverses it again, weno5_rk3.cpp:62-65):
//     each cell therefore needs three faces per axis, L at i-1/2, i+1/2 and
//     i+3/2;
//   * local Lax-Friedrichs Hamiltonian H = u.g - S_L|g| - alpha (g+ - g-)/2,
//     g the mean of the two derivatives, rhs = -H (weno5_rk3.cpp:91-145);
//   * RK3 with the stage forms G1 = G + dt r, G2 = 0.75 G + 0.25 (G1 + dt r),
//     G = 1/3 G + 2/3 (G2 + dt r), each followed by the anchored boundary and
//     the periodic fill (weno5_rk3.cpp:148-188).
// Anchored-then-periodic (field.cpp:16-63) leaves: interior row 3 = the
// initial field's row 3, interior row ny+2 = 2 row(ny+1) - row(ny), and every
// ghost cell a doubly periodic image of the interior.  The stage kernel
// writes exactly that itself (k_sample_ymarch): a wavefront marching up a
// strip has rows ny and ny+1 of its output when it reaches row ny+2, and
// every interior cell writes its own periodic images.  One launch per stage.
//
// Arithmetic keeps the reference's evaluation order; the library is compiled
// with -ffp-contract=off and IEEE division and sqrt, so results equal the
// reference's CPU build bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "ls_api.h"
#include "ls_crdiv.hpp"
#include "ls_internal.hpp"

namespace {

constexpr int kSX = 64, kSY = 4;
constexpr double kWenoEps = 1e-6;

struct SGeom {
    int32_t nx, ny, nxt, nyt;
    double dx, dy;
    double rdx, rdy;  // RN(1/dx), RN(1/dy) for lsx::qdiv
    int32_t crdiv;    // the spacings admit the FMA division sequences (ls_crdiv.hpp)
    int32_t pad;
};

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min

__device__ __forceinline__ double sbeta(double f0, double f1, double f2) {
    const double p = f0 - 2.0 * f1 + f2;
    const double q = f0 - 4.0 * f1 + 3.0 * f2;
    return (13.0 / 12.0) * (p * p) + (1.0 / 4.0) * (q * q);
}

// The part of sface after the weight numerators a0..a2.
template <int D>
__device__ __forceinline__ double sface_tail(double v0, double v1, double v2, double v3, double v4, double a0,
                                             double a1, double a2) {
    const double sum = a0 + a1 + a2;
    const double ys = D == lsx::kDivIEEE ? 0.0 : lsx::dr<D>(1.0, sum);
    const double w0 = lsx::dq<D>(a0, sum, ys), w1 = lsx::dq<D>(a1, sum, ys), w2 = lsx::dq<D>(a2, sum, ys);
    return lsx::d6<D>(w0 * (2.0 * v0 - 7.0 * v1 + 11.0 * v2)) + lsx::d6<D>(w1 * (-v1 + 5.0 * v2 + 2.0 * v3)) +
           lsx::d6<D>(w2 * (2.0 * v2 + 5.0 * v3 - v4));
}

// weno5_left of the reference (sample/weno5_rk3.cpp:17-32); D is the division
// flavour (ls_crdiv.hpp), every choice giving the same bits.
template <int D = lsx::kDivIEEE>
__device__ __forceinline__ double sface(double v0, double v1, double v2, double v3, double v4) {
    const double b0 = sbeta(v0, v1, v2);
    const double b1 = sbeta(v1, v2, v3);
    const double b2 = sbeta(v2, v3, v4);
    const double a0 = lsx::dr<D>(0.1, (kWenoEps + b0) * (kWenoEps + b0));
    const double a1 = lsx::dr<D>(0.6, (kWenoEps + b1) * (kWenoEps + b1));
    const double a2 = lsx::dr<D>(0.3, (kWenoEps + b2) * (kWenoEps + b2));
    return sface_tail<D>(v0, v1, v2, v3, v4, a0, a1, a2);
}

// The sub-stencil centred at f1: t = (eps + beta)^2, the divisor of its weight
// numerator, and rcp_u(t), the part of that quotient's rdiv_u that reads t
// alone.  A sub-stencil serves three faces, so the march computes this once.
__device__ __forceinline__ double2 ssub(double f0, double f1, double f2) {
    const double b = sbeta(f0, f1, f2);
    const double t = (kWenoEps + b) * (kWenoEps + b);
    return make_double2(t, lsx::rcp_u(t));
}

// sface<kDivFma> from the sub-stencil data of its centres v1, v2, v3: the same
// operations on the same values, so the same bits.
__device__ __forceinline__ double sface_t(double v0, double v1, double v2, double v3, double v4, double2 s0,
                                          double2 s1, double2 s2) {
    const double a0 = lsx::rdiv_y(0.1, s0.x, s0.y);
    const double a1 = lsx::rdiv_y(0.6, s1.x, s1.y);
    const double a2 = lsx::rdiv_y(0.3, s2.x, s2.y);
    return sface_tail<lsx::kDivFma>(v0, v1, v2, v3, v4, a0, a1, a2);
}

// compute_geqn_rhs at one interior cell (sample/weno5_rk3.cpp:111-143) from
// its three faces per axis, L at i-1/2, i+1/2, i+3/2 (dG_L and dG_R of
// weno5_rk3.cpp:53-87 are their differences over h).
template <int D = lsx::kDivIEEE>
__device__ __forceinline__ double srhs_f(const SGeom& g, double fmx, double f0x, double fpx, double fmy, double f0y,
                                         double fpy, double uc, double vc, double S_L) {
    const double Gdxm = lsx::dq<D>(f0x - fmx, g.dx, g.rdx), Gdxp = lsx::dq<D>(fpx - f0x, g.dx, g.rdx);
    const double Gdym = lsx::dq<D>(f0y - fmy, g.dy, g.rdy), Gdyp = lsx::dq<D>(fpy - f0y, g.dy, g.rdy);
    const double Gdx_mean = 0.5 * (Gdxm + Gdxp);
    const double Gdy_mean = 0.5 * (Gdym + Gdyp);
    const double transport = uc * Gdx_mean + vc * Gdy_mean;
    const double propagation = S_L * sqrt(Gdx_mean * Gdx_mean + Gdy_mean * Gdy_mean);
    const double nm = sqrt(Gdxm * Gdxm + Gdym * Gdym + 1e-14);
    const double np = sqrt(Gdxp * Gdxp + Gdyp * Gdyp + 1e-14);
    const double ynm = D == lsx::kDivIEEE ? 0.0 : lsx::dr<D>(1.0, nm);
    const double ynp = D == lsx::kDivIEEE ? 0.0 : lsx::dr<D>(1.0, np);
    const double s_xm = uc - lsx::dq<D>(S_L * Gdxm, nm, ynm);
    const double s_xp = uc - lsx::dq<D>(S_L * Gdxp, np, ynp);
    const double alpha_x = dmax(fabs(s_xm), fabs(s_xp));
    const double s_ym = vc - lsx::dq<D>(S_L * Gdym, nm, ynm);
    const double s_yp = vc - lsx::dq<D>(S_L * Gdyp, np, ynp);
    const double alpha_y = dmax(fabs(s_ym), fabs(s_yp));
    const double alpha = dmax(alpha_x, alpha_y);
    const double diss_x = alpha * (Gdxp - Gdxm) * 0.5;
    const double diss_y = alpha * (Gdyp - Gdym) * 0.5;
    const double H = transport - propagation - (diss_x + diss_y);
    return -H;
}

// The same with the faces built for this cell alone.
__device__ __forceinline__ double srhs(const SGeom& g, const double* __restrict__ G, int64_t c, double uc, double vc,
                                       double S_L) {
    const double* P = G + c;
    const int64_t s = g.nxt;
    return srhs_f(g, sface(P[-3], P[-2], P[-1], P[0], P[1]), sface(P[-2], P[-1], P[0], P[1], P[2]),
                  sface(P[-1], P[0], P[1], P[2], P[3]), sface(P[-3 * s], P[-2 * s], P[-s], P[0], P[s]),
                  sface(P[-2 * s], P[-s], P[0], P[s], P[2 * s]), sface(P[-s], P[0], P[s], P[2 * s], P[3 * s]), uc,
                  vc, S_L);
}

// Stage combine at cell c; Gn_c is G_n at c.
template <int STAGE>
__device__ __forceinline__ double scombine(double gn, double gs, double r, double dt) {
    if (STAGE == 1) return gn + dt * r;
    if (STAGE == 2) return 0.75 * gn + 0.25 * (gs + dt * r);
    return (1.0 / 3.0) * gn + (2.0 / 3.0) * (gs + dt * r);
}

// Periodic images of interior index q (padded) along an axis of n points.
__device__ __forceinline__ int s_nimg(int q, int n) { return 1 + (q < 6) + (q >= n); }
__device__ __forceinline__ int s_img(int q, int n, int a) {
    if (a == 0) return q;
    return (a == 1 && q < 6) ? q + n : q - n;
}

// One RK3 stage over the interior plus the anchored rows and periodic ghosts,
// as a y-march of one wavefront.  The 64 lanes of a strip of kSW = 62 cells
// hold the x faces between interior columns xs-1+l and xs+l (l = 0..63), so
// each x face is built once and read by the three cells that use it; each
// lane also carries its column's 7 y-values and the y faces j-1/2, j+1/2 in
// registers and builds one new y face per row.  A cell's arithmetic is
// srhs_f's on the same face values as srhs, so the result is bitwise the
// per-cell one.  Row 3 takes the initial field; row ny+2 (padded) is
// 2 row(ny+1) - row(ny) of this stage's output, which the same wavefront has
// just computed (the last chunk holds at least 3 rows).  out may alias Gn (stage 3):
// a cell reads G_n only at itself, before writing it.  The divisions run as
// the FMA sequences of ls_crdiv.hpp while every G value the wavefront has
// loaded passed lsx::crdiv_safe (a sticky per-lane flag and a wave vote), S_L
// does, and g.crdiv is set; otherwise as divisions.  Both give the same bits.
constexpr int kSW = 62;
#ifndef LS_SAMPLE_ROWS
#define LS_SAMPLE_ROWS 4  // rows per chunk but the last
#endif

struct SLds {
    double g[72];    // G of the current row at padded column xs + p
    double f[64];    // lane l's x face
    double2 c[68];   // FMA rows: ssub of the centre g[p]
};

template <int STAGE>
__global__ __launch_bounds__(64) void k_sample_ymarch(SGeom g, int strips, int nch, const double* __restrict__ Gs,
                                                      const double* Gn, double* out, const double* __restrict__ u,
                                                      const double* __restrict__ v,
                                                      const double* __restrict__ init, double S_L, double dt) {
    __shared__ SLds L;
    const int t = (int)blockIdx.x;
    const int sidx = t % strips, ch = t / strips;
    const int l = (int)threadIdx.x;
    const int xs = sidx * kSW;               // first interior column of the strip
    const int W = min(kSW, g.nx - xs);
    const int X = xs + 2 + l;                // padded column of lane l
    const bool col = X <= g.nx + 5;
    const bool cell = l >= 1 && l <= W;
    // chunks of LS_SAMPLE_ROWS rows; the last one takes the rest (3 rows or more)
    const int r0 = 3 + ch * LS_SAMPLE_ROWS, r1 = ch + 1 == nch ? g.ny + 3 : r0 + LS_SAMPLE_ROWS;
    const int64_t s = g.nxt;
    const bool same = Gs == Gn;
    // LDS columns outside the lanes' own: p = 0, 1 (lanes 0, 1) and 66, 67 (lanes 2, 3)
    const bool xl = l < 4;
    const int xp = l < 2 ? l : 64 + l;
    const bool xok = xs + xp <= g.nx + 5;
    double ring[7];  // G(X, r-3 .. r+3)
#pragma unroll
    for (int q = 0; q < 7; q++) ring[q] = col ? Gs[(int64_t)(r0 - 3 + q) * s + X] : 0.0;
    bool bad = false;  // a G value this lane loaded failed crdiv_safe
#pragma unroll
    for (int q = 0; q < 7; q++) bad |= !lsx::crdiv_safe(ring[q]);
    const bool fd_ok = g.crdiv && lsx::crdiv_safe(S_L);
    double fy0, fy1;  // y faces r-1/2, r+1/2
    // FMA rows: ssub of the centres ring[3], ring[4] (the y face r+3/2 of the
    // first row needs them, and each row passes them on).  The FMA rows are a
    // prefix of the march (bad is sticky), so what a row carries was computed
    // by FMA rows.
    double2 cy0 = make_double2(0.0, 0.0), cy1 = cy0;
    if (fd_ok && __ballot(bad) == 0) {
        fy0 = sface<lsx::kDivFma>(ring[0], ring[1], ring[2], ring[3], ring[4]);
        fy1 = sface<lsx::kDivFma>(ring[1], ring[2], ring[3], ring[4], ring[5]);
        cy0 = ssub(ring[2], ring[3], ring[4]);
        cy1 = ssub(ring[3], ring[4], ring[5]);
    } else {
        fy0 = sface<lsx::kDivIEEE>(ring[0], ring[1], ring[2], ring[3], ring[4]);
        fy1 = sface<lsx::kDivIEEE>(ring[1], ring[2], ring[3], ring[4], ring[5]);
    }
    double cu = 0.0, cv = 0.0, cgn = 0.0, xh = 0.0;
    if (cell) {
        const int64_t c = (int64_t)r0 * s + X;
        cu = u[c];
        cv = v[c];
        if (!same) cgn = Gn[c];
    }
    if (xl && xok) xh = Gs[(int64_t)r0 * s + xs + xp];
    double o1 = 0.0, o0 = 0.0;  // this stage's output at rows r-1, r-2
    for (int r = r0; r < r1; r++) {
        L.g[l + 2] = ring[3];
        if (xl) L.g[xp] = xh;
        bad |= !lsx::crdiv_safe(xh);
        const bool fd = fd_ok && __ballot(bad) == 0;  // this row's quotients by the FMA sequences
        double nr = 0.0, nu = 0.0, nv = 0.0, ngn = 0.0, nxh = 0.0;
        if (r + 1 < r1) {
            if (col) nr = Gs[(int64_t)(r + 4) * s + X];
            if (cell) {
                const int64_t c1 = (int64_t)(r + 1) * s + X;
                nu = u[c1];
                nv = v[c1];
                if (!same) ngn = Gn[c1];
            }
            if (xl && xok) nxh = Gs[(int64_t)(r + 1) * s + xs + xp];
        }
        __syncthreads();  // row in LDS
        double fy2;  // y face r+3/2
        double2 cy2 = make_double2(0.0, 0.0);  // FMA rows: ssub of the centre ring[5]
        if (fd) {  // (wavefront-uniform; the block is one wavefront)
            L.c[l + 2] = ssub(L.g[l + 1], L.g[l + 2], L.g[l + 3]);
            if (l < 2) {  // centres g[1] and g[66]
                const int p = l == 0 ? 1 : 66;
                L.c[p] = ssub(L.g[p - 1], L.g[p], L.g[p + 1]);
            }
            cy2 = ssub(ring[4], ring[5], ring[6]);
            fy2 = sface_t(ring[2], ring[3], ring[4], ring[5], ring[6], cy0, cy1, cy2);
            __syncthreads();  // centre data
            L.f[l] = sface_t(L.g[l], L.g[l + 1], L.g[l + 2], L.g[l + 3], L.g[l + 4], L.c[l + 1], L.c[l + 2], L.c[l + 3]);
        } else {
            L.f[l] = sface<lsx::kDivIEEE>(L.g[l], L.g[l + 1], L.g[l + 2], L.g[l + 3], L.g[l + 4]);
            fy2 = sface<lsx::kDivIEEE>(ring[2], ring[3], ring[4], ring[5], ring[6]);
        }
        __syncthreads();  // x faces
        if (cell) {
            const int64_t c = (int64_t)r * s + X;
            double val;
            if (r == 3) {
                val = init[c];
            } else if (r == g.ny + 2) {
                val = 2.0 * o1 - o0;
            } else {
                const double rr = fd ? srhs_f<lsx::kDivFma>(g, L.f[l - 1], L.f[l], L.f[l + 1], fy0, fy1, fy2, cu, cv, S_L)
                                     : srhs_f<lsx::kDivIEEE>(g, L.f[l - 1], L.f[l], L.f[l + 1], fy0, fy1, fy2, cu, cv, S_L);
                val = scombine<STAGE>(same ? ring[3] : cgn, ring[3], rr, dt);
            }
            const int nj = s_nimg(r, g.ny), ni = s_nimg(X, g.nx);
            for (int b = 0; b < nj; b++)
                for (int a = 0; a < ni; a++) out[(int64_t)s_img(r, g.ny, b) * s + s_img(X, g.nx, a)] = val;
            o0 = o1;
            o1 = val;
        }
#pragma unroll
        for (int q = 0; q < 6; q++) ring[q] = ring[q + 1];
        ring[6] = nr;
        bad |= !lsx::crdiv_safe(nr);
        fy0 = fy1;
        fy1 = fy2;
        cy0 = cy1;
        cy1 = cy2;
        cu = nu;
        cv = nv;
        cgn = ngn;
        xh = nxh;
    }
}

// compute_geqn_rhs into a padded array (zero outside the interior).
__global__ __launch_bounds__(256) void k_sample_rhs(SGeom g, const double* __restrict__ G,
                                                    const double* __restrict__ u, const double* __restrict__ v,
                                                    double S_L, double* __restrict__ rhs) {
    const int i = blockIdx.x * kSX + threadIdx.x;
    const int j = blockIdx.y * kSY + threadIdx.y;
    if (i >= g.nxt || j >= g.nyt) return;
    const int64_t c = (int64_t)j * g.nxt + i;
    const bool in = i >= 3 && i <= g.nx + 2 && j >= 3 && j <= g.ny + 2;
    rhs[c] = in ? srhs(g, G, c, u[c], v[c], S_L) : 0.0;
}

// Field2D::apply_periodic (field.cpp:16-35): ghost (j,i) takes the interior
// cell at its periodic image on both axes.
__global__ __launch_bounds__(256) void k_sample_periodic(SGeom g, double* G) {
    const int i = blockIdx.x * kSX + threadIdx.x;
    const int j = blockIdx.y * kSY + threadIdx.y;
    if (i >= g.nxt || j >= g.nyt) return;
    const bool gi = i < 3 || i > g.nx + 2, gj = j < 3 || j > g.ny + 2;
    if (!gi && !gj) return;
    const int si = i < 3 ? i + g.nx : (i > g.nx + 2 ? i - g.nx : i);
    const int sj = j < 3 ? j + g.ny : (j > g.ny + 2 ? j - g.ny : j);
    G[(int64_t)j * g.nxt + i] = G[(int64_t)sj * g.nxt + si];
}

// Field2D::apply_anchored_bc (field.cpp:37-63), first part: per column, the
// bottom rows (row 3 anchored to the initial field, rows 2..0 extrapolated)
// and the top rows (row ny+2 .. ny+5 extrapolated).
__global__ __launch_bounds__(256) void k_sample_anchor_rows(SGeom g, double* G, const double* __restrict__ init) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nxt) return;
    const int64_t s = g.nxt;
    double* A = G + i;
    A[3 * s] = init[3 * s + i];
    A[2 * s] = 2.0 * A[3 * s] - A[4 * s];
    A[1 * s] = 2.0 * A[2 * s] - A[3 * s];
    A[0] = 2.0 * A[1 * s] - A[2 * s];
    const int ny = g.ny;
    A[(ny + 2) * s] = 2.0 * A[(ny + 1) * s] - A[ny * s];
    A[(ny + 3) * s] = 2.0 * A[(ny + 2) * s] - A[(ny + 1) * s];
    A[(ny + 4) * s] = 2.0 * A[(ny + 3) * s] - A[(ny + 2) * s];
    A[(ny + 5) * s] = 2.0 * A[(ny + 4) * s] - A[(ny + 3) * s];
}

// Second part: per row (all rows), the left and right ghost columns.
__global__ __launch_bounds__(256) void k_sample_anchor_cols(SGeom g, double* G) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.nyt) return;
    double* A = G + (int64_t)j * g.nxt;
    const int nx = g.nx;
    A[2] = 2.0 * A[3] - A[4];
    A[1] = 2.0 * A[2] - A[3];
    A[0] = 2.0 * A[1] - A[2];
    A[nx + 3] = 2.0 * A[nx + 2] - A[nx + 1];
    A[nx + 4] = 2.0 * A[nx + 3] - A[nx + 2];
    A[nx + 5] = 2.0 * A[nx + 4] - A[nx + 3];
}

// One iteration of Field2D::reinitialize (field.cpp:111-150): phi_out =
// phi + dtau * (-s (|grad phi| - 1)) on the interior, s from the initial
// field, the reference's one-sided gradient picks.
__global__ __launch_bounds__(256) void k_sample_reinit(SGeom g, const double* __restrict__ phi,
                                                       const double* __restrict__ init, double dtau,
                                                       double* __restrict__ out) {
    const int i = 3 + blockIdx.x * kSX + threadIdx.x;
    const int j = 3 + blockIdx.y * kSY + threadIdx.y;
    if (i > g.nx + 2 || j > g.ny + 2) return;
    const int64_t c = (int64_t)j * g.nxt + i, s = g.nxt;
    const double i0 = init[c];
    const double sg = i0 / sqrt(i0 * i0 + g.dx * g.dx);
    const double dGdx_pos = (phi[c] - phi[c - 1]) / g.dx;
    const double dGdx_neg = (phi[c + 1] - phi[c]) / g.dx;
    const double dGdy_pos = (phi[c] - phi[c - s]) / g.dy;
    const double dGdy_neg = (phi[c + s] - phi[c]) / g.dy;
    double grad;
    if (sg > 0.0) {
        const double px = dmax(dmax(dGdx_pos, 0.0), -dmin(dGdx_neg, 0.0));
        const double py = dmax(dmax(dGdy_pos, 0.0), -dmin(dGdy_neg, 0.0));
        grad = sqrt(px * px + py * py);
    } else {
        const double qx = dmax(dmax(-dGdx_pos, 0.0), dmin(dGdx_neg, 0.0));
        const double qy = dmax(dmax(-dGdy_pos, 0.0), dmin(dGdy_neg, 0.0));
        grad = sqrt(qx * qx + qy * qy);
    }
    out[c] = phi[c] + dtau * (-sg * (grad - 1.0));
}

// compute_first_order_upwind (first_order_upwind.cpp:26-62) over rows and
// columns 3 .. n+3 (the reference's loop bounds reach the first high-side
// ghost row and column).
__global__ __launch_bounds__(256) void k_sample_first_order(SGeom g, const double* __restrict__ G,
                                                            const double* __restrict__ u,
                                                            const double* __restrict__ v, double dt, double S_L,
                                                            double* __restrict__ out) {
    const int i = 3 + blockIdx.x * kSX + threadIdx.x;
    const int j = 3 + blockIdx.y * kSY + threadIdx.y;
    if (i > g.nx + 3 || j > g.ny + 3) return;
    const int64_t c = (int64_t)j * g.nxt + i, s = g.nxt;
    const double Gx_mns = (G[c] - G[c - 1]) / g.dx;
    const double Gx_pls = (G[c + 1] - G[c]) / g.dx;
    const double Gy_mns = (G[c] - G[c - s]) / g.dy;
    const double Gy_pls = (G[c + s] - G[c]) / g.dy;
    const double dG_dx = 0.5 * (Gx_mns + Gx_pls);
    const double dG_dy = 0.5 * (Gy_mns + Gy_pls);
    const double mag_sq = dG_dx * dG_dx + dG_dy * dG_dy;
    double uf = u[c], vf = v[c];
    if (mag_sq > 1e-10 * 1e-10) {
        const double mag = sqrt(mag_sq);
        uf = u[c] - S_L * dG_dx / mag;
        vf = v[c] - S_L * dG_dy / mag;
    }
    const double a_pls = dmax(uf, 0.0), a_mns = dmin(uf, 0.0);
    const double b_pls = dmax(vf, 0.0), b_mns = dmin(vf, 0.0);
    out[c] = G[c] - dt * (a_pls * Gx_mns + a_mns * Gx_pls + b_pls * Gy_mns + b_mns * Gy_pls);
}

__global__ __launch_bounds__(256) void k_sample_copy_box(SGeom g, const double* __restrict__ src, double* dst,
                                                         int hi) {
    const int i = 3 + blockIdx.x * kSX + threadIdx.x;
    const int j = 3 + blockIdx.y * kSY + threadIdx.y;
    if (i > g.nx + 2 + hi || j > g.ny + 2 + hi) return;
    const int64_t c = (int64_t)j * g.nxt + i;
    dst[c] = src[c];
}

// The Bunsen-flame field of sample/main.cpp:95-116 (slope computed on the host).
__global__ __launch_bounds__(256) void k_sample_bunsen(SGeom g, double slope, double R, double X0, double Y0,
                                                       double* __restrict__ G) {
    const int i = blockIdx.x * kSX + threadIdx.x;
    const int j = blockIdx.y * kSY + threadIdx.y;
    if (i >= g.nxt || j >= g.nyt) return;
    const double x = (i - 3) * g.dx;
    const double y = (j - 3) * g.dy;
    double distance;
    if (x < X0)
        distance = slope * (x - (X0 - R)) - (y - Y0);
    else
        distance = -slope * (x - (X0 + R)) - (y - Y0);
    G[(int64_t)j * g.nxt + i] = -distance;
}

__global__ __launch_bounds__(256) void k_sample_fill2(int64_t n, double a, double b, double* __restrict__ u,
                                                      double* __restrict__ v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    u[t] = a;
    v[t] = b;
}

dim3 interior_grid(const SGeom& g, int extra) {
    return dim3((g.nx + extra + kSX - 1) / kSX, (g.ny + extra + kSY - 1) / kSY, 1);
}
dim3 full_grid(const SGeom& g) { return dim3((g.nxt + kSX - 1) / kSX, (g.nyt + kSY - 1) / kSY, 1); }

}  // namespace

struct ls_sample {
    SGeom g;
    int64_t n;
    double* G1 = nullptr;
    double* G2 = nullptr;
};

namespace {

int launch_ok(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, what);
    return LS_OK;
}

int rk3(ls_sample* s, double* G, const double* u, const double* v, double dt, double S_L, const double* init,
        hipStream_t st) {
    const SGeom& g = s->g;
    const int strips = (g.nx + kSW - 1) / kSW;
    const int nch = (g.ny - 3) / LS_SAMPLE_ROWS + 1;  // the last chunk starts at row ny or below
    const dim3 grid((unsigned)(strips * nch)), blk(64, 1, 1);
    hipLaunchKernelGGL(k_sample_ymarch<1>, grid, blk, 0, st, g, strips, nch, G, G, s->G1, u, v, init, S_L, dt);
    hipLaunchKernelGGL(k_sample_ymarch<2>, grid, blk, 0, st, g, strips, nch, s->G1, G, s->G2, u, v, init, S_L, dt);
    hipLaunchKernelGGL(k_sample_ymarch<3>, grid, blk, 0, st, g, strips, nch, s->G2, G, G, u, v, init, S_L, dt);
    return launch_ok("sample stage launch failed");
}

int anchored(ls_sample* s, double* G, const double* init, hipStream_t st) {
    const SGeom& g = s->g;
    hipLaunchKernelGGL(k_sample_anchor_rows, dim3((g.nxt + 255) / 256), dim3(256), 0, st, g, G, init);
    hipLaunchKernelGGL(k_sample_anchor_cols, dim3((g.nyt + 255) / 256), dim3(256), 0, st, g, G);
    return launch_ok("anchored boundary launch failed");
}

int periodic(ls_sample* s, double* G, hipStream_t st) {
    hipLaunchKernelGGL(k_sample_periodic, full_grid(s->g), dim3(kSX, kSY, 1), 0, st, s->g, G);
    return launch_ok("periodic fill launch failed");
}

}  // namespace

extern "C" {

int64_t ls_sample_field_size(int32_t nx, int32_t ny) {
    if (nx < 6 || ny < 6) return -1;
    return (int64_t)(nx + 6) * (ny + 6);
}

int ls_sample_create(int32_t nx, int32_t ny, double dx, double dy, ls_sample_t** out) {
    if (!out) return lsi::set_error(LS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (nx < 6 || ny < 6) return lsi::set_error(LS_ERR_INVALID, "the sample scheme needs nx, ny >= 6");
    if ((int64_t)(nx + 6) * (ny + 6) >= (int64_t)1 << 31) return lsi::set_error(LS_ERR_INVALID, "grid too large");
    if (!(dx > 0.0) || !(dy > 0.0)) return lsi::set_error(LS_ERR_INVALID, "dx and dy must be > 0");
    ls_sample* s = new ls_sample();
    s->g.nx = nx;
    s->g.ny = ny;
    s->g.nxt = nx + 6;
    s->g.nyt = ny + 6;
    s->g.dx = dx;
    s->g.dy = dy;
    s->g.rdx = 1.0 / dx;
    s->g.rdy = 1.0 / dy;
    s->g.crdiv = lsx::crdiv_geom_ok(dx, dy, 1.0, 1e-14);  // 1e-14: the eps under the |grad G| of the LLF speeds
    s->n = (int64_t)s->g.nxt * s->g.nyt;
    const size_t fb = (size_t)s->n * sizeof(double);
    if (hipMalloc((void**)&s->G1, fb) != hipSuccess || hipMalloc((void**)&s->G2, fb) != hipSuccess) {
        (void)hipFree(s->G1);
        (void)hipFree(s->G2);
        delete s;
        return lsi::set_error(LS_ERR_NOMEM, "sample workspace allocation failed");
    }
    // every stage writes all cells of G1/G2 (interior and ghost images);
    // zero them once so that no call ever reads undefined memory
    if (hipMemset(s->G1, 0, fb) != hipSuccess || hipMemset(s->G2, 0, fb) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(s->G1);
        (void)hipFree(s->G2);
        delete s;
        return lsi::set_error(LS_ERR_DEVICE, "sample workspace initialisation failed");
    }
    *out = s;
    return LS_OK;
}

int ls_sample_destroy(ls_sample_t* s) {
    if (!s) return LS_OK;
    (void)hipFree(s->G1);
    (void)hipFree(s->G2);
    delete s;
    return LS_OK;
}

int ls_sample_rhs(ls_sample_t* s, const double* G, const double* u, const double* v, double S_L, double* rhs,
                  void* stream) {
    if (!s || !G || !u || !v || !rhs) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    hipLaunchKernelGGL(k_sample_rhs, full_grid(s->g), dim3(kSX, kSY, 1), 0, reinterpret_cast<hipStream_t>(stream),
                       s->g, G, u, v, S_L, rhs);
    return launch_ok("sample rhs launch failed");
}

int ls_sample_rk3_step(ls_sample_t* s, double* G, const double* u, const double* v, double dt, double S_L,
                       const double* G_initial, void* stream) {
    if (!s || !G || !u || !v || !G_initial) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    return rk3(s, G, u, v, dt, S_L, G_initial, reinterpret_cast<hipStream_t>(stream));
}

int ls_sample_apply_periodic(ls_sample_t* s, double* G, void* stream) {
    if (!s || !G) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    return periodic(s, G, reinterpret_cast<hipStream_t>(stream));
}

int ls_sample_apply_anchored_bc(ls_sample_t* s, double* G, const double* G_initial, void* stream) {
    if (!s || !G || !G_initial) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    return anchored(s, G, G_initial, reinterpret_cast<hipStream_t>(stream));
}

int ls_sample_reinitialize(ls_sample_t* s, double* G, const double* G_initial, int32_t num_iters, void* stream) {
    if (!s || !G || !G_initial) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    if (num_iters < 0) return lsi::set_error(LS_ERR_INVALID, "num_iters must be >= 0");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const SGeom& g = s->g;
    const double dtau = 0.2 * (g.dy < g.dx ? g.dy : g.dx);
    double* a = G;
    double* b = s->G1;
    for (int it = 0; it < num_iters; it++) {
        // ghost cells of b are all rewritten by the boundary below
        hipLaunchKernelGGL(k_sample_reinit, interior_grid(g, 0), dim3(kSX, kSY, 1), 0, st, g, a, G_initial, dtau, b);
        int rc = anchored(s, b, G_initial, st);
        if (rc) return rc;
        double* t = a;
        a = b;
        b = t;
    }
    if (a != G && hipMemcpyAsync(G, a, (size_t)s->n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
        return lsi::set_error(LS_ERR_DEVICE, "copy failed");
    return launch_ok("reinitialisation launch failed");
}

int ls_sample_first_order_upwind(ls_sample_t* s, double* G, const double* u, const double* v, double dt,
                                 double S_L, void* stream) {
    if (!s || !G || !u || !v) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const SGeom& g = s->g;
    hipLaunchKernelGGL(k_sample_first_order, interior_grid(g, 1), dim3(kSX, kSY, 1), 0, st, g, G, u, v, dt, S_L,
                       s->G2);
    hipLaunchKernelGGL(k_sample_copy_box, interior_grid(g, 1), dim3(kSX, kSY, 1), 0, st, g, s->G2, G, 1);
    return launch_ok("first-order upwind launch failed");
}

int ls_sample_bunsen(ls_sample_t* s, double R, double flame_angle_deg, double X0, double Y0, double* G,
                     void* stream) {
    if (!s || !G) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    const double slope = std::tan((90.0 - flame_angle_deg) * M_PI / 180.0);
    hipLaunchKernelGGL(k_sample_bunsen, full_grid(s->g), dim3(kSX, kSY, 1), 0, reinterpret_cast<hipStream_t>(stream),
                       s->g, slope, R, X0, Y0, G);
    return launch_ok("Bunsen initial field launch failed");
}

int ls_sample_forced_velocity(ls_sample_t* s, double t, int32_t freq, double u0, double v0, double eps_u,
                              double eps_v, double* u, double* v, void* stream) {
    if (!s || !u || !v) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    // sample/main.cpp:131-136, evaluated on the host with the same libm cos
    const double uu = u0 + eps_u * v0 * cos(freq * 2 * M_PI * (t));
    const double vv = v0 + eps_v * v0 * cos(freq * 2 * M_PI * (t));
    hipLaunchKernelGGL(k_sample_fill2, dim3((unsigned)((s->n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), s->n, uu, vv, u, v);
    return launch_ok("velocity fill launch failed");
}

int ls_sample_advance(ls_sample_t* s, double* G, double* u, double* v, const double* G_initial, double S_L,
                      double dt, double t0, int64_t nsteps, int32_t freq, double u0, double v0, double eps_u,
                      double eps_v, double* t_out, void* stream) {
    if (!s || !G || !u || !v || !G_initial || !t_out) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    if (nsteps < 0) return lsi::set_error(LS_ERR_INVALID, "nsteps must be >= 0");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    double t = t0;
    int rc;
    if (nsteps > 0) {
        // loop head of sample/main.cpp:127-128; later iterations find the state
        // the previous stage-3 boundary left, on which both calls are no-ops
        // for the interior and produce the same periodic ghosts
        if ((rc = anchored(s, G, G_initial, st)) || (rc = periodic(s, G, st))) return rc;
    }
    for (int64_t k = 0; k < nsteps; k++) {
        if ((rc = ls_sample_forced_velocity(s, t, freq, u0, v0, eps_u, eps_v, u, v, stream))) return rc;
        if ((rc = rk3(s, G, u, v, dt, S_L, G_initial, st))) return rc;
        t += dt;
    }
    if (nsteps > 0) {
        // loop tail (main.cpp:140-141): periodic then anchored
        if ((rc = periodic(s, G, st)) || (rc = anchored(s, G, G_initial, st))) return rc;
    }
    if (hipStreamSynchronize(st) != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, "sample advance failed");
    *t_out = t;
    return LS_OK;
}

}  // extern "C"
