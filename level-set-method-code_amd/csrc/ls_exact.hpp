// ls_exact.hpp -- device numerics of the exact mode.
//
// Every expression keeps the reference's C++ evaluation order and is compiled
// with -ffp-contract=off, so that each cell reproduces the contraction-free
// CPU result bit for bit (IEEE double, correctly rounded division and sqrt).
// The marching stage kernels may carry out the divisions by 6 and by a
// divisor shared between several quotients (the weight sum, the grid
// spacing, |grad G|) with the FMA sequences of ls_crdiv.hpp, which return the
// same correctly rounded quotients; the per-cell kernels here divide.
//
//   smoothness indicators  Local_code/level-set_MPI_3D/include/weno5.h:28-37
//   weno5_left / right     Local_code/level-set_MPI_3D/include/weno5.h:39-81
//   upwind axis derivative Local_code/level-set_MPI_3D/include/weno5.h:90-196
//   u_eff RHS              Local_code/level-set_MPI_3D/include/rk3.h:31-81
//                          Local_code/level-set_MPI_2D/include/rk3.h:54-98
//   Godunov |grad phi|     Local_code/level-set_MPI_3D/include/weno5.h:202-248
//   HCR-2 step             Local_code/level-set_MPI_3D/include/reinitialization.h:46-193
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ls_crdiv.hpp"

namespace lsx {

constexpr double kWenoEps = 1.0e-6;

// Kernel-side geometry, passed by value.
struct Geom {
    int32_t dim, nx, ny, nz, ng;
    int32_t nxt, nyt, nzt;
    int32_t bc[3];
    int64_t sy, sz;      // strides of y and z in doubles
    double dx, dy, dz;
    double rdx, rdy, rdz;  // RN(1/dx), RN(1/dy), RN(1/dz) for the marching kernels' lsx::qdiv_u
    double s_l, eps;
    double dtau, beta;
};

// beta_k of the three 3-point sub-stencils of (a,b,c,d,e).
__device__ __forceinline__ double beta3(double p, double q, double r) {
    double d = r - 2.0 * q + p;
    return (13.0 / 12.0) * d * d + 0.25 * (r - p) * (r - p);
}

// weno5_left on (v0..v4): reconstruction at the right face of v2.  D selects
// how the divisions are carried out (ls_crdiv.hpp); every choice returns the
// same correctly rounded quotients.
template <int D = kDivIEEE>
__device__ __forceinline__ double weno_l(double v0, double v1, double v2, double v3, double v4) {
    double b0 = beta3(v0, v1, v2);
    double b1 = beta3(v1, v2, v3);
    double b2 = beta3(v2, v3, v4);
    double a0 = dr<D>(0.1, (kWenoEps + b0) * (kWenoEps + b0));
    double a1 = dr<D>(0.6, (kWenoEps + b1) * (kWenoEps + b1));
    double a2 = dr<D>(0.3, (kWenoEps + b2) * (kWenoEps + b2));
    double s = a0 + a1 + a2;
    const double ys = D == kDivIEEE ? 0.0 : dr<D>(1.0, s);
    double w0 = dq<D>(a0, s, ys), w1 = dq<D>(a1, s, ys), w2 = dq<D>(a2, s, ys);
    double p0 = d6<D>(2.0 * v0 - 7.0 * v1 + 11.0 * v2);
    double p1 = d6<D>(-v1 + 5.0 * v2 + 2.0 * v3);
    double p2 = d6<D>(2.0 * v2 + 5.0 * v3 - v4);
    return w0 * p0 + w1 * p1 + w2 * p2;
}

// weno5_right on (v0..v4): the indicators are taken on the mirrored stencil
// (v4,v3,v2,v1,v0), so beta_0 uses (v4,v3,v2), beta_1 (v3,v2,v1), beta_2 (v2,v1,v0).
template <int D = kDivIEEE>
__device__ __forceinline__ double weno_r(double v0, double v1, double v2, double v3, double v4) {
    double b0 = beta3(v4, v3, v2);
    double b1 = beta3(v3, v2, v1);
    double b2 = beta3(v2, v1, v0);
    double a0 = dr<D>(0.3, (kWenoEps + b0) * (kWenoEps + b0));
    double a1 = dr<D>(0.6, (kWenoEps + b1) * (kWenoEps + b1));
    double a2 = dr<D>(0.1, (kWenoEps + b2) * (kWenoEps + b2));
    double s = a0 + a1 + a2;
    const double ys = D == kDivIEEE ? 0.0 : dr<D>(1.0, s);
    double w0 = dq<D>(a0, s, ys), w1 = dq<D>(a1, s, ys), w2 = dq<D>(a2, s, ys);
    double p0 = d6<D>(11.0 * v2 - 7.0 * v3 + 2.0 * v4);
    double p1 = d6<D>(2.0 * v1 + 5.0 * v2 - v3);
    double p2 = d6<D>(-v0 + 5.0 * v1 + 2.0 * v2);
    return w0 * p0 + w1 * p1 + w2 * p2;
}

#if defined(__HIPCC__)
// The smoothness data of the 3-point sub-stencil centred at c in each
// orientation, as weno_l (x: beta3(G(c-1), G(c), G(c+1))) and weno_r
// (y: beta3(G(c+1), G(c), G(c-1))) compute it -- the two add the outer
// values in opposite orders, so their bits may differ -- each as
// (t, rcp_u(t)), t = (eps + beta)^2: the divisor of the weight numerator and
// the reciprocal part of its rdiv_u.  A sub-stencil serves three faces of one
// orientation, so the marching kernels compute this once per centre.
__device__ __forceinline__ double2 sub_l(double gm, double g0, double gp) {
    const double b = beta3(gm, g0, gp);
    const double t = (kWenoEps + b) * (kWenoEps + b);
    return make_double2(t, rcp_u(t));
}
__device__ __forceinline__ double2 sub_r(double gm, double g0, double gp) { return sub_l(gp, g0, gm); }

// weno_l / weno_r with the FMA division sequences (kDivFma), from the
// sub-stencil data (t, rcp_u(t)) of its three centres in the order weno_l /
// weno_r takes its beta_0, beta_1, beta_2: the same operations on the same
// values, so the same bits.
template <bool MIRROR>
__device__ __forceinline__ double weno_t(double v0, double v1, double v2, double v3, double v4, double2 s0,
                                         double2 s1, double2 s2) {
    double a0 = rdiv_y(MIRROR ? 0.3 : 0.1, s0.x, s0.y);
    double a1 = rdiv_y(0.6, s1.x, s1.y);
    double a2 = rdiv_y(MIRROR ? 0.1 : 0.3, s2.x, s2.y);
    double s = a0 + a1 + a2;
    const double ys = rdiv_u(1.0, s);
    double w0 = qdiv_u(a0, s, ys), w1 = qdiv_u(a1, s, ys), w2 = qdiv_u(a2, s, ys);
    double p0, p1, p2;
    if (MIRROR) {
        p0 = div6_u(11.0 * v2 - 7.0 * v3 + 2.0 * v4);
        p1 = div6_u(2.0 * v1 + 5.0 * v2 - v3);
        p2 = div6_u(-v0 + 5.0 * v1 + 2.0 * v2);
    } else {
        p0 = div6_u(2.0 * v0 - 7.0 * v1 + 11.0 * v2);
        p1 = div6_u(-v1 + 5.0 * v2 + 2.0 * v3);
        p2 = div6_u(2.0 * v2 + 5.0 * v3 - v4);
    }
    return w0 * p0 + w1 * p1 + w2 * p2;
}
#endif

// The part of weno_u after the weight numerators a0..a2.
template <int D>
__device__ __forceinline__ double weno_u_tail(double u0, double u1, double u2, double u3, double u4, double a0,
                                              double a1, double a2, bool mirror) {
    double s = a0 + a1 + a2;
    const double ys = D == kDivIEEE ? 0.0 : dr<D>(1.0, s);
    double w0 = dq<D>(a0, s, ys), w1 = dq<D>(a1, s, ys), w2 = dq<D>(a2, s, ys);
    // left: p0 = (2u0 - 7u1) + 11u2, p1 = (-u1 + 5u2) + 2u3, p2 = (2u2 + 5u3) - u4;
    // right: the first and last term of each swapped
    const double e0 = 2.0 * u0, f0 = 11.0 * u2;
    const double e1 = -u1, f1 = 2.0 * u3;
    const double e2 = 2.0 * u2, f2 = -u4;
    double p0 = d6<D>((mirror ? f0 : e0) - 7.0 * u1 + (mirror ? e0 : f0));
    double p1 = d6<D>((mirror ? f1 : e1) + 5.0 * u2 + (mirror ? e1 : f1));
    double p2 = d6<D>((mirror ? f2 : e2) + 5.0 * u3 + (mirror ? e2 : f2));
    return w0 * p0 + w1 * p1 + w2 * p2;
}

// weno_l(u0..u4) when mirror is false; when mirror is true weno_r on the
// stencil (u4, u3, u2, u1, u0), i.e. with its values given in mirrored order.
// weno5_right is weno5_left's arithmetic on the mirrored stencil with the
// linear weights (0.3, 0.6, 0.1), except that each candidate polynomial adds
// its three terms in the opposite order (weno5.h:39-81).  Both orders are
// selected per lane, so one call covers either orientation without a branch
// and returns the bits of weno_l / weno_r.
template <int D = kDivIEEE>
__device__ __forceinline__ double weno_u(double u0, double u1, double u2, double u3, double u4, bool mirror) {
    double b0 = beta3(u0, u1, u2);
    double b1 = beta3(u1, u2, u3);
    double b2 = beta3(u2, u3, u4);
    const double d0 = mirror ? 0.3 : 0.1, d2 = mirror ? 0.1 : 0.3;
    double a0 = dr<D>(d0, (kWenoEps + b0) * (kWenoEps + b0));
    double a1 = dr<D>(0.6, (kWenoEps + b1) * (kWenoEps + b1));
    double a2 = dr<D>(d2, (kWenoEps + b2) * (kWenoEps + b2));
    return weno_u_tail<D>(u0, u1, u2, u3, u4, a0, a1, a2, mirror);
}

#if defined(__HIPCC__)
// weno_u with the FMA sequences from the sub-stencil data (t, rcp_u(t)) of
// beta3(u0, u1, u2), beta3(u1, u2, u3), beta3(u2, u3, u4): sub_l at the
// centres u1, u2, u3 when mirror is false; when it is true the stencil is
// given mirrored, and those are sub_r at the same centres.
__device__ __forceinline__ double weno_tu(double u0, double u1, double u2, double u3, double u4, double2 s0,
                                          double2 s1, double2 s2, bool mirror) {
    const double d0 = mirror ? 0.3 : 0.1, d2 = mirror ? 0.1 : 0.3;
    double a0 = rdiv_y(d0, s0.x, s0.y);
    double a1 = rdiv_y(0.6, s1.x, s1.y);
    double a2 = rdiv_y(d2, s2.x, s2.y);
    return weno_u_tail<kDivFma>(u0, u1, u2, u3, u4, a0, a1, a2, mirror);
}
#endif

// Upwinded WENO-5 derivative from the 7-point axis stencil m3..p3.
__device__ __forceinline__ double dweno(double m3, double m2, double m1, double c0, double p1,
                                        double p2, double p3, double vel, double h) {
    if (vel >= 0.0) {
        double gl = weno_l<kDivIEEE>(m3, m2, m1, c0, p1);
        double gr = weno_l<kDivIEEE>(m2, m1, c0, p1, p2);
        return (gr - gl) / h;
    }
    double gl = weno_r<kDivIEEE>(m2, m1, c0, p1, p2);
    double gr = weno_r<kDivIEEE>(m1, c0, p1, p2, p3);
    return (gr - gl) / h;
}

// RHS -(u_eff . grad_w G) at the cell P points to.
template <int DIM>
__device__ __forceinline__ double rhs_cell(const Geom& g, const double* __restrict__ P, double uc,
                                           double vc, double wc, double s_l) {
    const int64_t sy = g.sy, sz = g.sz;
    double gx = 0.0, gy = 0.0, gz = 0.0, gm = 1.0;
    if (s_l > g.eps) {
        gx = (P[1] - P[-1]) / (2.0 * g.dx);
        gy = (P[sy] - P[-sy]) / (2.0 * g.dy);
        if (DIM == 3) {
            gz = (P[sz] - P[-sz]) / (2.0 * g.dz);
            gm = sqrt(gx * gx + gy * gy + gz * gz + g.eps);
        } else {
            gm = sqrt(gx * gx + gy * gy + g.eps);
        }
    }
    double ue = uc - s_l * gx / gm;
    double ve = vc - s_l * gy / gm;
    double dgx = dweno(P[-3], P[-2], P[-1], P[0], P[1], P[2], P[3], ue, g.dx);
    double dgy = dweno(P[-3 * sy], P[-2 * sy], P[-sy], P[0], P[sy], P[2 * sy], P[3 * sy], ve, g.dy);
    if (DIM == 3) {
        double we = wc - s_l * gz / gm;
        double dgz = dweno(P[-3 * sz], P[-2 * sz], P[-sz], P[0], P[sz], P[2 * sz], P[3 * sz], we, g.dz);
        return -(ue * dgx + ve * dgy + we * dgz);
    }
    return -(ue * dgx + ve * dgy);
}

// Shu-Osher stage combine (rk3.h:103-150).
template <int STAGE>
__device__ __forceinline__ double stage_combine(double gn, double gs, double r, double dt) {
    if (STAGE == 1) return gn + dt * r;
    if (STAGE == 2) return 0.75 * gn + 0.25 * gs + 0.25 * dt * r;
    return (1.0 / 3.0) * gn + (2.0 / 3.0) * gs + (2.0 / 3.0) * dt * r;
}

// One HCR-2 pseudo-time step at a cell from phi_n (P) and phi_0 (Q) at the
// cell and its neighbours in flag order x-, x+, y-, y+, z-, z+ (entries 5 and
// 6 unused in 2D), in the reference's operation order (reinitialization.h,
// 3D :136-193).  Crossing flags and r~ are recomputed from phi_0 with the
// arithmetic of computeInterfaceCrossings, so they equal the reference's
// stored arrays.
template <int DIM>
__device__ __forceinline__ double hcr2_core(const double (&P)[7], const double (&Q)[7], double dx, double dy,
                                            double dz, double dtau, double beta) {
    constexpr int NB = DIM == 3 ? 6 : 4;
    const double pc = Q[0];
    int flag = 0;
#pragma unroll
    for (int q = 0; q < NB; q++)
        if (pc * Q[q + 1] < 0.0) flag |= 1 << q;
    // r~ (the reference sums the crossing neighbours in flag order); only cells
    // beside the interface have a flag, so most wavefronts skip this
    double rt = 0.0;
    if (flag != 0) {
        double sum0 = 0.0;
#pragma unroll
        for (int q = 0; q < NB; q++)
            if ((flag >> q) & 1) sum0 += Q[q + 1];
        if (fabs(sum0) > 1e-15) rt = pc / sum0;
    }

    const double phi = P[0];
    const double S = pc / sqrt(pc * pc + dx * dx);
    // Godunov terms.  The reference takes, per axis, max(a^2, b^2) with
    // a = max(D-, 0), b = min(D+, 0) for S > 0 and the mirrored min/max
    // otherwise, D- = (phi_c - phi_lo) / h, D+ = (phi_hi - phi_c) / h.  With
    // s = +-1 the sign of S that is max(A, B)^2 over A = max(s (phi_c - phi_lo)
    // / h, 0), B = max(s (phi_c - phi_hi) / h, 0): negation is exact, and a
    // rounded square, a rounded difference and a rounded quotient by h > 0
    // are monotone, so max(A, B) = max(RN(RN(s phi_c - min(s phi_lo,
    // s phi_hi)) / h), 0) and the axis costs one division instead of two.
    const int sgn = S > 0.0 ? 0 : (int)0x80000000u;  // the sign bit of -s
    double R[7];
#pragma unroll
    for (int q = 0; q < 7; q++) R[q] = __hiloint2double(__double2hiint(P[q]) ^ sgn, __double2loint(P[q]));
    const double gx = fmax((R[0] - fmin(R[1], R[2])) / dx, 0.0);
    const double gy = fmax((R[0] - fmin(R[3], R[4])) / dy, 0.0);
    double s2 = gx * gx + gy * gy;
    if (DIM == 3) {
        const double gz = fmax((R[0] - fmin(R[5], R[6])) / dz, 0.0);
        s2 = s2 + gz * gz;
    }
    const double term = -S * (sqrt(s2) - 1.0);

    double forcing = 0.0;
    if (flag != 0) {
        bool stable = (P[0] > 0.0) == (pc > 0.0);
#pragma unroll
        for (int q = 0; q < NB; q++)
            if ((flag >> q) & 1) stable = stable && ((P[q + 1] > 0.0) == (Q[q + 1] > 0.0));
        if (stable) {
            double sum = 0.0;
#pragma unroll
            for (int q = 0; q < NB; q++)
                if ((flag >> q) & 1) sum += P[q + 1];
            forcing = beta * (rt * sum - phi) / dx;
        }
    }
    return phi + dtau * (term + forcing);
}

// hcr2_core at the cell pn / p0 point at (padded layout of g).
template <int DIM>
__device__ __forceinline__ double hcr2_cell(const Geom& g, const double* __restrict__ pn,
                                            const double* __restrict__ p0) {
    const int64_t off[6] = {-1, 1, -g.sy, g.sy, -g.sz, g.sz};
    double P[7], Q[7];
    P[0] = pn[0];
    Q[0] = p0[0];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        P[q + 1] = (DIM == 3 || q < 4) ? pn[off[q]] : 0.0;
        Q[q + 1] = (DIM == 3 || q < 4) ? p0[off[q]] : 0.0;
    }
    return hcr2_core<DIM>(P, Q, g.dx, g.dy, g.dz, g.dtau, g.beta);
}

}  // namespace lsx
