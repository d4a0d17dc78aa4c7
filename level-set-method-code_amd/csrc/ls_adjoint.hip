// ls_adjoint.hip -- discrete adjoint of the 3D WENO-5 / TVD-RK3 / HCR-2 solver.
//
// Loss  L = sum over interior cells of (H_e(G_T) - H_e(G_target))^2 with the
// smoothed Heaviside H_e(x) = 0 (x < -e), 1 (x > e),
// (1 + x/e + sin(pi x/e)/pi) / 2 otherwise.  ls_adjoint_run returns L and
// dL/dG0, dL/du, dL/dv, dL/dw and dL/dS_L (per-cell fields) by reverse-mode
// differentiation of exactly the discrete scheme the forward solver runs:
//   * RK3 stage (Local_code/level-set_MPI_3D/include/rk3.h:31-125): the RHS
//     -(u_eff . grad_w G) is differentiated through the WENO-5 faces
//     (beta -> alpha -> omega -> face, weno5.h:28-81), the central gradient
//     and the u_eff normalisation (rk3.h:31-81); the upwind choice, the sign
//     of u_eff (weno5.h:90-196), is the forward one;
//   * ghost fill (boundary.h:250-261): a ghost cell copies an interior cell,
//     so its adjoint is added into that cell ("fold");
//   * HCR-2 step (reinitialization.h:46-193): smoothed sign of phi_0,
//     Godunov |grad phi| (weno5.h:202-248) with the forward's max/min picks,
//     forcing beta (r~ sum phi_n - phi_n) / dx with r~ = phi_0 / sum phi_0; the
//     crossing flags and the stability test are discrete and keep their
//     forward values.
// The forward trajectory is checkpointed every `ckpt` steps; during the
// reverse sweep each segment is recomputed once by the forward kernels, which
// are deterministic, so every branch above is re-derived from a state that is
// bitwise the forward state.
//
// Each cell's partial derivatives P_q[c] = lambda_c * d out_c / d G[c + off_q]
// must be summed at the target cells c + off_q.  The stage adjoint (k_adj_zm)
// and the HCR-2 adjoint (k_adj_hcr2_zm) sum their partials inside the block
// and leave the block-edge spills to a combine kernel.  Fixed summation
// order, no atomics: the gradient is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ls_api.h"
#include "ls_exact.hpp"
#include "ls_fast_math.hpp"
#include "ls_internal.hpp"

namespace {

constexpr int kAX = 64, kAY = 4;
constexpr int kLossBlocks = 1024;
constexpr double kPi = 3.14159265358979323846;

struct AGeom {
    int32_t nx, ny, nz;
    int32_t bc[3];
    int64_t sy, sz, n;
    double dx, dy, dz, eps, dtau, beta;
};

__device__ __forceinline__ int64_t cell(const AGeom& g, int i, int j, int k) {
    return (int64_t)(k + 3) * g.sz + (int64_t)(j + 3) * g.sy + (i + 3);
}

// The face gradients below are derivatives held to the tangent checker's
// tolerance, not to the forward's bits, so they let the compiler contract
// a*b + c into FMAs (the library is built with contraction off).  The upwind
// choice, which must match the forward, is made outside them.

// Smoothness indicator of (a,b,c) and its gradient.
__device__ __forceinline__ double beta_g(double a, double b, double c, double& da, double& db, double& dc) {
#pragma clang fp contract(fast)
    const double D = a - 2.0 * b + c, E = c - a;
    da = (13.0 / 6.0) * D - 0.5 * E;
    db = -(13.0 / 3.0) * D;
    dc = (13.0 / 6.0) * D + 0.5 * E;
    return (13.0 / 12.0) * D * D + 0.25 * E * E;
}

// Left-biased WENO-5 face of (v0..v4) with linear weights (d0,d1,d2), and its
// gradient df[0..4].  The right-biased face of the reference is this on the
// reversed stencil with weights (0.3, 0.6, 0.1) (weno5.h:60-81).  Divisions
// are refined reciprocals (one per smoothness term, one for the weight sum):
// the adjoint is a derivative, held to the tangent checker's tolerance, not to
// the forward's bits.
__device__ __forceinline__ double face_grad(const double* v, double d0, double d1, double d2, double* df) {
#pragma clang fp contract(fast)
    double gb[3][3], b[3];
    b[0] = beta_g(v[0], v[1], v[2], gb[0][0], gb[0][1], gb[0][2]);
    b[1] = beta_g(v[1], v[2], v[3], gb[1][0], gb[1][1], gb[1][2]);
    b[2] = beta_g(v[2], v[3], v[4], gb[2][0], gb[2][1], gb[2][2]);
    const double d[3] = {d0, d1, d2};
    double a[3], re[3], s = 0.0;
    for (int k = 0; k < 3; k++) {
        re[k] = lsf::rcp_nr(lsx::kWenoEps + b[k]);
        a[k] = d[k] * (re[k] * re[k]);
        s += a[k];
    }
    constexpr double s6 = 1.0 / 6.0;
    const double p[3] = {(2.0 * v[0] - 7.0 * v[1] + 11.0 * v[2]) * s6, (-v[1] + 5.0 * v[2] + 2.0 * v[3]) * s6,
                         (2.0 * v[2] + 5.0 * v[3] - v[4]) * s6};
    const double rs = lsf::rcp_nr(s);
    double w[3], f = 0.0;
    for (int k = 0; k < 3; k++) {
        w[k] = a[k] * rs;
        f += w[k] * p[k];
    }
    df[0] = w[0] * (2.0 / 6.0);
    df[1] = w[0] * (-7.0 / 6.0) + w[1] * (-1.0 / 6.0);
    df[2] = w[0] * (11.0 / 6.0) + w[1] * (5.0 / 6.0) + w[2] * (2.0 / 6.0);
    df[3] = w[1] * (2.0 / 6.0) + w[2] * (5.0 / 6.0);
    df[4] = w[2] * (-1.0 / 6.0);
    for (int k = 0; k < 3; k++) {
        // d f / d beta_k = -2 w_k (p_k - f) / (eps + beta_k)
        const double gk = -2.0 * w[k] * (p[k] - f) * re[k];
        df[k] += gk * gb[k][0];
        df[k + 1] += gk * gb[k][1];
        df[k + 2] += gk * gb[k][2];
    }
    return f;
}

// Upwind WENO-5 derivative along one axis from the stencil s7 = (m3..p3), and
// its gradient P[0..6]; ih = 1/h.
__device__ __forceinline__ double dweno_grad(const double* s7, double vel, double ih, double* P) {
#pragma clang fp contract(fast)
    double fa, fb, ga[5], gb[5];
    for (int q = 0; q < 7; q++) P[q] = 0.0;
    if (vel >= 0.0) {
        fa = face_grad(s7 + 0, 0.1, 0.6, 0.3, ga);  // face i-1/2 from m3..p1
        fb = face_grad(s7 + 1, 0.1, 0.6, 0.3, gb);  // face i+1/2 from m2..p2
        for (int q = 0; q < 5; q++) {
            P[q] -= ga[q] * ih;
            P[q + 1] += gb[q] * ih;
        }
    } else {
        const double ra[5] = {s7[5], s7[4], s7[3], s7[2], s7[1]};  // face i-1/2, mirrored m2..p2
        const double rb[5] = {s7[6], s7[5], s7[4], s7[3], s7[2]};  // face i+1/2, mirrored m1..p3
        fa = face_grad(ra, 0.3, 0.6, 0.1, ga);
        fb = face_grad(rb, 0.3, 0.6, 0.1, gb);
        for (int q = 0; q < 5; q++) {
            P[5 - q] -= ga[q] * ih;
            P[6 - q] += gb[q] * ih;
        }
    }
    return (fb - fa) * ih;
}

// Plane offsets of the 7-point stencil: 0 centre, x-, x+, y-, y+, z-, z+.
__device__ __forceinline__ int64_t off7(const AGeom& g, int q) {
    switch (q) {
        case 0: return 0;
        case 1: return -1;
        case 2: return 1;
        case 3: return -g.sy;
        case 4: return g.sy;
        case 5: return -g.sz;
        default: return g.sz;
    }
}

// Blocks of the z-marching stage adjoint: x segments of at most 64 cells
// (one wavefront row each), 4-row y tiles, z chunks; segments and chunks are
// balanced, so every segment and chunk spans at least 6 cells.
struct ZMGeom {
    int nsx, nty, nch;
};

__device__ __forceinline__ int split_at(int n, int parts, int q) { return (int)(((int64_t)n * q) / parts); }

// Stage partials, summed in the block.  A block walks its z chunk one plane
// at a time; each cell c computes its 19 partials lam_c * dRHS_c/dG[c + o]
// (lam_c = cdt * lam_out[c]): the centre and the offsets -3..3 of each axis,
// through the WENO-5 faces (dweno_grad) and the central gradient of u_eff.
//   * x: the 6 x partials of a row go through LDS; every cell of the segment
//     sums the ones that land on it, and lanes 0..5 sum the ones that leave
//     the segment (3 cells each side) into the spill plane EX.
//   * y: likewise through LDS across the tile's rows; the sums for the 3 rows
//     below and the 3 above the tile go to the spill planes EYlo and EYhi.
//     (Those of two neighbouring 4-row tiles overlap, hence two planes.)
//   * z: each cell carries accumulators for the 7 target planes k-3..k+3 of
//     its column; a target plane is complete three planes later and goes to
//     `core`, or, below and above the chunk, to the spill plane EZ.
// A spill position has one writer per pass, because segments and chunks span
// at least 6 cells and tiles at least 3 rows, and every pass writes the same
// positions; positions never written stay zero.  The sums run in a fixed
// order: the result is deterministic.
__global__ __launch_bounds__(kAX * kAY) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_adj_zm(AGeom g, ZMGeom zm, const double* __restrict__ Gs,
                                                     const double* __restrict__ u, const double* __restrict__ v,
                                                     const double* __restrict__ w, const double* __restrict__ sl,
                                                     double sl0, const double* __restrict__ lam_out, double cdt,
                                                     double* __restrict__ core, double* __restrict__ EX,
                                                     double* __restrict__ EYlo, double* __restrict__ EYhi,
                                                     double* __restrict__ EZ,
                                                     double* __restrict__ gu, double* __restrict__ gv,
                                                     double* __restrict__ gw, double* __restrict__ gsl) {
    __shared__ double sX[kAY][6][kAX];
    __shared__ double sY[kAY][6][kAX];
    const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
    int b = (int)blockIdx.x;
    const int sxi = b % zm.nsx;
    b /= zm.nsx;
    const int ty = b % zm.nty, ch = b / zm.nty;
    const int xs = split_at(g.nx, zm.nsx, sxi), L = split_at(g.nx, zm.nsx, sxi + 1) - xs;
    const int kb = split_at(g.nz, zm.nch, ch), ke = split_at(g.nz, zm.nch, ch + 1);
    const int i = xs + lx, j = ty * kAY + ly;
    const bool row = j < g.ny;
    const bool act = row && lx < L;
    double zacc[7];  // targets k-3 .. k+3 of the current source plane k
#pragma unroll
    for (int q = 0; q < 7; q++) zacc[q] = 0.0;
    for (int k = kb; k < ke; k++) {
        if (act) {
            const int64_t c = cell(g, i, j, k);
            // every load of the cell up front: one round trip per plane
            const double lam = cdt * lam_out[c];
            const double uc = u[c], vc = v[c], wc = w[c], s_l = sl ? sl[c] : sl0;
            const double gu0 = gu[c], gv0 = gv[c], gw0 = gw[c], gs0 = gsl ? gsl[c] : 0.0;
            // The partials one axis at a time, each stored as soon as it is
            // known, so that the 19 are never live together: x and y into LDS,
            // z into the accumulators.
            const double* G = Gs + c;
            const double ihx = 1.0 / g.dx, ihy = 1.0 / g.dy, ihz = 1.0 / g.dz;
            double gx = 0.0, gy = 0.0, gz = 0.0, m = 1.0, im = 1.0;
            const bool on = s_l > g.eps;
            if (on) {  // u_eff in the exact forward's arithmetic: its sign picks the upwind side
                gx = (G[1] - G[-1]) / (2.0 * g.dx);
                gy = (G[g.sy] - G[-g.sy]) / (2.0 * g.dy);
                gz = (G[g.sz] - G[-g.sz]) / (2.0 * g.dz);
                m = sqrt(gx * gx + gy * gy + gz * gz + g.eps);
                im = lsf::rcp_nr(m);
            }
            const double ue = uc - s_l * gx / m, ve = vc - s_l * gy / m, we = wc - s_l * gz / m;
            double s7[7], P[7];
#pragma unroll
            for (int q = 0; q < 7; q++) s7[q] = G[q - 3];
            const double dgx = dweno_grad(s7, ue, ihx, P);
            double ctr = ue * P[3];
            double xm1 = -ue * P[2], xp1 = -ue * P[4];  // the offsets -1, +1 get the S_L terms too
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q != 2 && q != 3) sX[ly][q][lx] = (-ue * P[q < 3 ? q : q + 1]) * lam;
#pragma unroll
            for (int q = 0; q < 7; q++) s7[q] = G[(int64_t)(q - 3) * g.sy];
            const double dgy = dweno_grad(s7, ve, ihy, P);
            ctr += ve * P[3];
            double ym1 = -ve * P[2], yp1 = -ve * P[4];
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q != 2 && q != 3) sY[ly][q][lx] = (-ve * P[q < 3 ? q : q + 1]) * lam;
#pragma unroll
            for (int q = 0; q < 7; q++) s7[q] = G[(int64_t)(q - 3) * g.sz];
            const double dgz = dweno_grad(s7, we, ihz, P);
            ctr += we * P[3];
            double zm1 = -we * P[2], zp1 = -we * P[4];
            // z partial to plane k + o, o = -3, -2, +2, +3 (the +-1 ones below)
            zacc[0] += (-we * P[0]) * lam;
            zacc[1] += (-we * P[1]) * lam;
            zacc[5] += (-we * P[5]) * lam;
            zacc[6] += (-we * P[6]) * lam;
            double qim = 0.0;
            if (on) {
                // d RHS / d g_b = s_l (dg_b - g_b Q / m^2) / m with Q = sum_a dg_a g_a
                const double Q = dgx * gx + dgy * gy + dgz * gz;
                const double qm = Q * (im * im);
                const double rx = s_l * (dgx - gx * qm) * im;
                const double ry = s_l * (dgy - gy * qm) * im;
                const double rz = s_l * (dgz - gz * qm) * im;
                xm1 -= rx * (0.5 * ihx);
                xp1 += rx * (0.5 * ihx);
                ym1 -= ry * (0.5 * ihy);
                yp1 += ry * (0.5 * ihy);
                zm1 -= rz * (0.5 * ihz);
                zp1 += rz * (0.5 * ihz);
                qim = Q * im;
            }
            sX[ly][2][lx] = xm1 * lam;
            sX[ly][3][lx] = xp1 * lam;
            sY[ly][2][lx] = ym1 * lam;
            sY[ly][3][lx] = yp1 * lam;
            zacc[2] += zm1 * lam;
            zacc[4] += zp1 * lam;
            zacc[3] += -ctr * lam;
            gu[c] = gu0 - lam * dgx;
            gv[c] = gv0 - lam * dgy;
            gw[c] = gw0 - lam * dgz;
            if (gsl && on) gsl[c] = gs0 + lam * qim;
        } else {
#pragma unroll
            for (int q = 0; q < 6; q++) {
                sX[ly][q][lx] = 0.0;
                sY[ly][q][lx] = 0.0;
            }
        }
        __syncthreads();
        // x (y) partial q of source s lands on s + o, o = -3,-2,-1,1,2,3
        if (act) {
            double xsum = 0.0, ysum = 0.0;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const int o = q < 3 ? q - 3 : q - 2;
                const int s = lx - o, r = ly - o;
                if (s >= 0 && s < L) xsum += sX[ly][q][s];
                if (r >= 0 && r < kAY) ysum += sY[r][q][lx];
            }
            zacc[3] += xsum + ysum;
        } else if (lx < L && j <= g.ny + 2) {
            // a ghost row inside the tile (j >= ny): only the y partials of the
            // tile's own rows land here; no other writer of core at this position
            double ysum = 0.0;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const int r = ly - (q < 3 ? q - 3 : q - 2);
                if (r >= 0 && r < kAY) ysum += sY[r][q][lx];
            }
            core[cell(g, i, j, k)] = ysum;
        }
        // the 3 rows below (EYlo) and above (EYhi) the tile, column lx
        if (lx < L) {
            for (int e = ly; e < 6; e += kAY) {
                const int r = e < 3 ? e - 3 : kAY + e - 3;
                const int jt = ty * kAY + r;
                if (jt <= g.ny + 2) {
                    double ye = 0.0;
#pragma unroll
                    for (int q = 0; q < 6; q++) {
                        const int s = r - (q < 3 ? q - 3 : q - 2);
                        if (s >= 0 && s < kAY) ye += sY[s][q][lx];
                    }
                    (e < 3 ? EYlo : EYhi)[cell(g, i, jt, k)] = ye;
                }
            }
        }
        if (row && lx < 6) {  // the 3 targets on each side of the segment
            const int p = lx < 3 ? lx - 3 : L + lx - 3;
            double e = 0.0;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const int s = p - (q < 3 ? q - 3 : q - 2);
                if (s >= 0 && s < L) e += sX[ly][q][s];
            }
            EX[cell(g, xs + p, j, k)] = e;
        }
        __syncthreads();
        if (act) {  // target plane k-3 has all of this chunk's partials
            const int kt = k - 3;
            (kt >= kb ? core : EZ)[cell(g, i, j, kt)] = zacc[0];
        }
#pragma unroll
        for (int q = 0; q < 6; q++) zacc[q] = zacc[q + 1];
        zacc[6] = 0.0;
    }
    if (act) {  // targets ke-3 .. ke+2
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int kt = ke - 3 + q;
            (q < 3 ? core : EZ)[cell(g, i, j, kt)] = zacc[q];
        }
    }
}

// target[p] = a * self[p] + core[p] + EX[p] + EYlo[p] + EYhi[p] + EZ[p]
// (+ target[p] if accumulate) over every padded cell p: the block sums of
// k_adj_zm and its spills.  Planes are zero where nothing was written.  With
// acc2 given, also acc2[p] += a2 * self[p] (the stage's G_n term), or
// acc2[p] = a2 * self[p] when acc2_init is set (the first term of a step).
__global__ __launch_bounds__(256) void k_adj_combine(AGeom g, const double* __restrict__ core,
                                                     const double* __restrict__ EX, const double* __restrict__ EYlo,
                                                     const double* __restrict__ EYhi, const double* __restrict__ EZ,
                                                     const double* __restrict__ self, double a,
                                                     double* __restrict__ target, int accumulate,
                                                     double* __restrict__ acc2, double a2, int acc2_init) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < g.n; p += (int64_t)gridDim.x * blockDim.x) {
        const double sp = self ? self[p] : 0.0;
        double acc = self ? a * sp : 0.0;
        acc += core[p] + EX[p] + EYlo[p] + EYhi[p] + EZ[p];
        target[p] = accumulate ? target[p] + acc : acc;
        if (acc2) acc2[p] = acc2_init ? a2 * sp : acc2[p] + a2 * sp;
    }
}

// Number of padded positions that hold interior index q along an axis of n
// points (itself plus its ghost images), and the a-th of them.  Periodic:
// padded q+3 and, for q < 3, q+3+n, for q >= n-3, q+3-n.  Zero gradient:
// the edge points also fill their three ghosts.
__device__ __forceinline__ int n_img(int q, int n, int bc) {
    if (bc == LS_BC_PERIODIC) return 1 + (q < 3) + (q >= n - 3);
    return 1 + 3 * (q == 0) + 3 * (q == n - 1);
}
__device__ __forceinline__ int img_at(int q, int n, int bc, int a) {
    if (a == 0) return q + 3;
    if (bc == LS_BC_PERIODIC) return (a == 1 && q < 3) ? q + 3 + n : q + 3 - n;
    return q == 0 ? a - 1 : n + 2 + a;
}

// Cells within 3 of a face of the interior (the only ones with ghost images),
// enumerated as the z slabs (6 planes), then the y slabs of the remaining
// planes, then the x slabs of the remaining rows.  Needs nx, ny, nz >= 6.
__host__ __device__ __forceinline__ int64_t shell_cells(int nx, int ny, int nz) {
    return 6LL * nx * ny + 6LL * nx * (nz - 6) + 6LL * (ny - 6) * (nz - 6);
}
__device__ __forceinline__ bool shell_cell(const AGeom& g, int64_t e, int& i, int& j, int& k) {
    const int64_t a = 6LL * g.nx * g.ny, b = 6LL * g.nx * (g.nz - 6), c = 6LL * (g.ny - 6) * (g.nz - 6);
    if (e < a) {
        const int64_t pl = (int64_t)g.nx * g.ny;
        const int kk = (int)(e / pl), r = (int)(e - kk * pl);
        k = kk < 3 ? kk : g.nz - 6 + kk;
        j = r / g.nx;
        i = r - j * g.nx;
        return true;
    }
    e -= a;
    if (e < b) {
        const int64_t pl = 6LL * g.nx;
        const int kz = (int)(e / pl), r = (int)(e - kz * pl), jj = r / g.nx;
        k = 3 + kz;
        j = jj < 3 ? jj : g.ny - 6 + jj;
        i = r - jj * g.nx;
        return true;
    }
    e -= b;
    if (e < c) {
        const int64_t pl = 6LL * (g.ny - 6);
        const int kz = (int)(e / pl), r = (int)(e - kz * pl), ii = r % 6;
        k = 3 + kz;
        j = 3 + r / 6;
        i = ii < 3 ? ii : g.nx - 6 + ii;
        return true;
    }
    return false;
}

// Adjoint of a ghost fill: each interior cell near a face adds the adjoints of
// the ghost cells that copy it and clears them.  One thread per shell cell.
__global__ __launch_bounds__(256) void k_fold(AGeom g, double* __restrict__ lam) {
    int i, j, k;
    if (!shell_cell(g, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int ni = n_img(i, g.nx, g.bc[0]), nj = n_img(j, g.ny, g.bc[1]), nk = n_img(k, g.nz, g.bc[2]);
    if (ni * nj * nk == 1) return;
    double acc = 0.0;
    for (int c = 0; c < nk; c++)
        for (int b = 0; b < nj; b++)
            for (int a = 0; a < ni; a++) {
                if (!(a | b | c)) continue;
                const int64_t q = (int64_t)img_at(k, g.nz, g.bc[2], c) * g.sz +
                                  (int64_t)img_at(j, g.ny, g.bc[1], b) * g.sy + img_at(i, g.nx, g.bc[0], a);
                acc += lam[q];
                lam[q] = 0.0;
            }
    lam[cell(g, i, j, k)] += acc;
}

// HCR-2 partials of one pseudo-time step phi_new = step(phi_n; phi_0):
// d phi_new_c / d phi_n[c + off7(q)] (dN) and / d phi_0[c + off7(q)] (dZ) of
// one HCR-2 iteration at cell c, from phi_n (N) and phi_0 (Z) on its 7-point
// stencil in off7 order.
__device__ __forceinline__ void hcr2_partials(const AGeom& g, const double (&N)[7], const double (&Z)[7],
                                              double (&dN)[7], double (&dZ)[7]) {
    for (int q = 0; q < 7; q++) {
        dN[q] = 0.0;
        dZ[q] = 0.0;
    }
    const double pc = Z[0];
    double sum0 = 0.0;
    int flag = 0;
    for (int q = 0; q < 6; q++)
        if (pc * Z[q + 1] < 0.0) {
            sum0 += Z[q + 1];
            flag |= 1 << q;
        }
    const bool rt_on = flag != 0 && fabs(sum0) > 1e-15;
    const double rt = rt_on ? pc / sum0 : 0.0;
    const double den = pc * pc + g.dx * g.dx;
    const double S = pc / sqrt(den);
    const double dS = g.dx * g.dx / (den * sqrt(den));
    // Godunov |grad phi_n|: s2 = sum over axes of max(A^2, B^2)
    const double h[3] = {g.dx, g.dy, g.dz};
    double s2 = 0.0, gs2[7];
    for (int q = 0; q < 7; q++) gs2[q] = 0.0;
    for (int a = 0; a < 3; a++) {
        const int qm = 1 + 2 * a, qp = 2 + 2 * a;
        const double Dm = (N[0] - N[qm]) / h[a], Dp = (N[qp] - N[0]) / h[a];
        double A, B;
        bool Aon, Bon;
        if (S > 0.0) {
            A = fmax(Dm, 0.0);
            B = fmin(Dp, 0.0);
            Aon = Dm > 0.0;
            Bon = Dp < 0.0;
        } else {
            A = fmin(Dm, 0.0);
            B = fmax(Dp, 0.0);
            Aon = Dm < 0.0;
            Bon = Dp > 0.0;
        }
        if (A * A >= B * B) {
            s2 += A * A;
            if (Aon) {
                gs2[0] += 2.0 * A / h[a];
                gs2[qm] -= 2.0 * A / h[a];
            }
        } else {
            s2 += B * B;
            if (Bon) {
                gs2[qp] += 2.0 * B / h[a];
                gs2[0] -= 2.0 * B / h[a];
            }
        }
    }
    const double gm = sqrt(s2);
    // phi_new = phi_n + dtau (-S (gm - 1) + forcing)
    dN[0] += 1.0;
    if (gm > 0.0)
        for (int q = 0; q < 7; q++) dN[q] -= g.dtau * S * gs2[q] / (2.0 * gm);
    dZ[0] -= g.dtau * dS * (gm - 1.0);
    if (flag != 0) {
        bool stable = (N[0] > 0.0) == (pc > 0.0);
        for (int q = 0; q < 6; q++)
            if ((flag >> q) & 1) stable = stable && ((N[q + 1] > 0.0) == (Z[q + 1] > 0.0));
        if (stable) {
            double sumn = 0.0;
            for (int q = 0; q < 6; q++)
                if ((flag >> q) & 1) sumn += N[q + 1];
            const double bd = g.beta / g.dx;
            dN[0] -= g.dtau * bd;
            for (int q = 0; q < 6; q++)
                if ((flag >> q) & 1) dN[q + 1] += g.dtau * bd * rt;
            if (rt_on) {
                dZ[0] += g.dtau * bd * sumn / sum0;
                for (int q = 0; q < 6; q++)
                    if ((flag >> q) & 1) dZ[q + 1] -= g.dtau * bd * pc * sumn / (sum0 * sum0);
            }
        }
    }
}

// One HCR-2 iteration's adjoint, summed in the block like k_adj_zm.  The
// partials lam_c * d phi_new_c / d phi_n[c + o] (and / d phi_0) land on the
// cell and its six neighbours: x and y neighbours are summed through LDS, z
// in three accumulators per column (targets k-1, k, k+1).  Spills: the column
// left and right of the segment (EX), the row below and above the tile (EY),
// the plane below and above the chunk (EZ); the spills of two neighbouring
// blocks never coincide, so each plane has one writer per position.  The
// sums for phi_n go to planes[0..3] (core, EX, EY, EZ), those for phi_0 to
// planes[4..7]; positions never written stay zero.
__global__ __launch_bounds__(kAX * kAY) void k_adj_hcr2_zm(AGeom g, ZMGeom zm, const double* __restrict__ pn,
                                                            const double* __restrict__ p0,
                                                            const double* __restrict__ lam_new,
                                                            double* __restrict__ planes) {
    __shared__ double sX[2][kAY][2][kAX];  // [phi_n, phi_0][row][to x-1, to x+1][lane]
    __shared__ double sY[2][kAY][2][kAX];  // [phi_n, phi_0][row][to y-1, to y+1][lane]
    const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
    int b = (int)blockIdx.x;
    const int sxi = b % zm.nsx;
    b /= zm.nsx;
    const int ty = b % zm.nty, ch = b / zm.nty;
    const int xs = split_at(g.nx, zm.nsx, sxi), L = split_at(g.nx, zm.nsx, sxi + 1) - xs;
    const int kb = split_at(g.nz, zm.nch, ch), ke = split_at(g.nz, zm.nch, ch + 1);
    const int y0 = ty * kAY, i = xs + lx, j = y0 + ly;
    const bool row = j < g.ny;
    const bool act = row && lx < L;
    double za[2][3];  // per output: targets k-1, k, k+1 of the current plane k
#pragma unroll
    for (int o = 0; o < 2; o++)
#pragma unroll
        for (int q = 0; q < 3; q++) za[o][q] = 0.0;
    for (int k = kb; k < ke; k++) {
        if (act) {
            const int64_t c = cell(g, i, j, k);
            const double lam = lam_new[c];
            double N[7], Z[7], dN[7], dZ[7];
#pragma unroll
            for (int q = 0; q < 7; q++) {
                N[q] = pn[c + off7(g, q)];
                Z[q] = p0[c + off7(g, q)];
            }
            hcr2_partials(g, N, Z, dN, dZ);
#pragma unroll
            for (int o = 0; o < 2; o++) {
                const double(&D)[7] = o ? dZ : dN;
                sX[o][ly][0][lx] = lam * D[1];
                sX[o][ly][1][lx] = lam * D[2];
                sY[o][ly][0][lx] = lam * D[3];
                sY[o][ly][1][lx] = lam * D[4];
                za[o][0] += lam * D[5];
                za[o][1] += lam * D[0];
                za[o][2] += lam * D[6];
            }
        } else {
#pragma unroll
            for (int o = 0; o < 2; o++)
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    sX[o][ly][e][lx] = 0.0;
                    sY[o][ly][e][lx] = 0.0;
                }
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < 2; o++) {
            double* core = planes + (int64_t)(4 * o) * g.n;
            double* EX = core + g.n;
            double* EY = core + 2 * g.n;
            double ys = 0.0;  // y neighbours' partials landing on this cell
            if (ly + 1 < kAY) ys += sY[o][ly + 1][0][lx];
            if (ly >= 1) ys += sY[o][ly - 1][1][lx];
            if (act) {
                double xs_ = 0.0;
                if (lx + 1 < L) xs_ += sX[o][ly][0][lx + 1];
                if (lx >= 1) xs_ += sX[o][ly][1][lx - 1];
                za[o][1] += xs_ + ys;
            } else if (lx < L && j <= g.ny + 2) {
                core[cell(g, i, j, k)] = ys;  // a ghost row inside the tile: y partials only
            }
            if (row && lx < 2)  // the column left (lane 0) and right (lane 1) of the segment
                EX[cell(g, lx == 0 ? xs - 1 : xs + L, j, k)] = lx == 0 ? sX[o][ly][0][0] : sX[o][ly][1][L - 1];
            if (lx < L) {  // the row below (row 0) and above (row 1) the tile
                if (ly == 0) EY[cell(g, i, y0 - 1, k)] = sY[o][0][0][lx];
                else if (ly == 1 && y0 + kAY <= g.ny + 2) EY[cell(g, i, y0 + kAY, k)] = sY[o][kAY - 1][1][lx];
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int o = 0; o < 2; o++) {
                double* core = planes + (int64_t)(4 * o) * g.n;
                double* EZ = core + 3 * g.n;
                (k - 1 >= kb ? core : EZ)[cell(g, i, j, k - 1)] = za[o][0];
                za[o][0] = za[o][1];
                za[o][1] = za[o][2];
                za[o][2] = 0.0;
            }
        }
    }
    if (act) {  // targets ke-1 (complete) and ke (above the chunk)
#pragma unroll
        for (int o = 0; o < 2; o++) {
            double* core = planes + (int64_t)(4 * o) * g.n;
            core[cell(g, i, j, ke - 1)] = za[o][0];
            (core + 3 * g.n)[cell(g, i, j, ke)] = za[o][1];
        }
    }
}

// lamN[p] = the phi_n sums (planes 0..3) and lamZ[p] += the phi_0 sums
// (planes 4..7; lamZ[p] = them when init is set) over every padded cell p.
__global__ __launch_bounds__(256) void k_hcr2_combine(AGeom g, const double* __restrict__ planes,
                                                      double* __restrict__ lamN, double* __restrict__ lamZ,
                                                      int init) {
    const int64_t n = g.n;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        lamN[p] = ((planes[p] + planes[n + p]) + planes[2 * n + p]) + planes[3 * n + p];
        const double z = ((planes[4 * n + p] + planes[5 * n + p]) + planes[6 * n + p]) + planes[7 * n + p];
        lamZ[p] = init ? z : lamZ[p] + z;
    }
}

__device__ __forceinline__ double heaviside(double x, double e, double& d) {
    if (x < -e) {
        d = 0.0;
        return 0.0;
    }
    if (x > e) {
        d = 0.0;
        return 1.0;
    }
    d = (1.0 + cos(kPi * x / e)) / (2.0 * e);
    return 0.5 * (1.0 + x / e + sin(kPi * x / e) / kPi);
}

// Per-block partial loss sums over interior cells; lam (optional) receives
// dL/dG on interior cells.
__global__ __launch_bounds__(256) void k_loss(AGeom g, const double* __restrict__ G, const double* __restrict__ T,
                                              double e, double* __restrict__ lam, double* __restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    const int64_t nc = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nc; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((int64_t)g.nx * g.ny));
        const int64_t c = cell(g, i, j, k);
        double dG, dT;
        const double r = heaviside(G[c], e, dG) - heaviside(T[c], e, dT);
        acc += r * r;
        if (lam) lam[c] = 2.0 * r * dG;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_loss_final(const double* __restrict__ part, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < kLossBlocks; b++) s += part[b];
        *out = s;
    }
}

__global__ __launch_bounds__(256) void k_axpy(int64_t n, double a, const double* __restrict__ x,
                                              double* __restrict__ y) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        y[p] += a * x[p];
}

AGeom make_ageom(const ls_params_t& p) {
    AGeom g;
    std::memset(&g, 0, sizeof(g));
    g.nx = p.nx;
    g.ny = p.ny;
    g.nz = p.nz;
    for (int a = 0; a < 3; a++) g.bc[a] = p.bc[a];
    g.sy = p.nx + 6;
    g.sz = g.sy * (p.ny + 6);
    g.n = g.sz * (p.nz + 6);
    g.dx = p.dx;
    g.dy = p.dy;
    g.dz = p.dz;
    g.eps = p.eps;
    g.dtau = p.reinit_dtau;
    g.beta = p.reinit_beta;
    return g;
}

}  // namespace

struct ls_adjoint {
    ls_solver_t* s = nullptr;
    ls_params_t p;
    AGeom g;
    int nsteps = 0, ckpt = 0, nseg = 0, iters = 0;
    size_t fb = 0;
    ZMGeom zm;                  // block layout of the z-marching stage adjoint
    double* P = nullptr;        // 19 partial planes, zero where unwritten (0..7 HCR-2's, 14..18 the stage adjoint's)
    double* lam = nullptr;      // adjoint of the current state
    double* lam2 = nullptr;     // scratch adjoints
    double* lam3 = nullptr;
    double* acc = nullptr;      // adjoint of G_n while a step is reversed
    double* G1 = nullptr;       // recomputed stage fields
    double* G2 = nullptr;
    double* part = nullptr;     // loss partials + result
    double* loss_host = nullptr;
    std::vector<double*> ck;    // checkpoints: state at the start of step m*ckpt
    std::vector<double*> seg;   // states of one segment (ckpt + 1)
    std::vector<double*> pre;   // pre-reinitialisation states of one segment
    std::vector<double*> hit;   // HCR-2 iterates 1..iters-1 (hit[0] unused)
    std::vector<double*> g1s;   // stage fields G1, G2 of each step of one segment
    std::vector<double*> g2s;
    std::vector<double*> all;   // every device allocation
};

namespace {

void release(ls_adjoint* a) {
    for (double* q : a->all) (void)hipFree(q);
    if (a->loss_host) (void)hipHostFree(a->loss_host);
    delete a;
}

}  // namespace

extern "C" {

int ls_adjoint_create(ls_solver_t* s, int32_t nsteps, int32_t ckpt, ls_adjoint_t** out) {
    if (!out) return lsi::set_error(LS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!s) return lsi::set_error(LS_ERR_INVALID, "solver is NULL");
    const ls_params_t& p = lsi::solver_params(s);
    if (p.dim != 3) return lsi::set_error(LS_ERR_INVALID, "the adjoint covers the 3D solver");
    for (int a = 0; a < 3; a++)
        if (p.bc[a] != LS_BC_PERIODIC && p.bc[a] != LS_BC_ZERO_GRAD)
            return lsi::set_error(LS_ERR_INVALID, "the adjoint supports periodic and zero-gradient boundaries");
    if (p.nghost != 3) return lsi::set_error(LS_ERR_INVALID, "the adjoint needs 3 ghost layers");
    if (p.nx < 6 || p.ny < 6 || p.nz < 6) return lsi::set_error(LS_ERR_INVALID, "the adjoint needs >= 6 points per axis");
    if (nsteps < 1 || ckpt < 1) return lsi::set_error(LS_ERR_INVALID, "nsteps and ckpt must be >= 1");
    ls_adjoint* a = new ls_adjoint();
    a->s = s;
    a->p = p;
    a->g = make_ageom(p);
    a->nsteps = nsteps;
    a->ckpt = std::min(ckpt, nsteps);
    a->nseg = (nsteps + a->ckpt - 1) / a->ckpt;
    a->iters = (p.reinit_interval > 0) ? p.reinit_iters : 0;
    a->fb = (size_t)a->g.n * sizeof(double);
    // Balanced x segments of at most 64 cells and z chunks of at least 6
    // planes; as many chunks as fill the GPU's resident blocks once (one round,
    // no tail), as the occupancy query reports for this kernel.
    a->zm.nsx = (p.nx + kAX - 1) / kAX;
    a->zm.nty = (p.ny + kAY - 1) / kAY;
    {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adj_zm, kAX * kAY, 0) != hipSuccess) {
            delete a;
            return lsi::set_error(LS_ERR_DEVICE, "device query for the adjoint failed");
        }
        const int64_t cols = (int64_t)a->zm.nsx * a->zm.nty;
        const int64_t fill = std::max<int64_t>(1, ((int64_t)cus * std::max(per_cu, 1) + cols / 2) / cols);
        a->zm.nch = (int)std::max<int64_t>(1, std::min<int64_t>(fill, p.nz / 6));
    }
    auto field = [&](size_t count) -> double* {
        double* d = nullptr;
        if (hipMalloc((void**)&d, count * a->fb) != hipSuccess) return nullptr;
        a->all.push_back(d);
        return d;
    };
    bool ok = (a->P = field(19)) && (a->lam = field(1)) && (a->lam2 = field(1)) && (a->lam3 = field(1)) &&
              (a->acc = field(1)) && (a->G1 = field(1)) && (a->G2 = field(1));
    if (ok) {
        ok = hipMalloc((void**)&a->part, (kLossBlocks + 1) * sizeof(double)) == hipSuccess;
        if (ok) a->all.push_back(a->part);
    }
    for (int q = 0; ok && q < a->nseg; q++) ok = (a->ck.emplace_back(field(1)) != nullptr);
    for (int q = 0; ok && q <= a->ckpt; q++) ok = (a->seg.emplace_back(field(1)) != nullptr);
    for (int q = 0; ok && a->iters > 0 && q < a->ckpt; q++) ok = (a->pre.emplace_back(field(1)) != nullptr);
    if (a->iters > 0) a->hit.push_back(nullptr);
    for (int q = 1; ok && q < a->iters; q++) ok = (a->hit.emplace_back(field(1)) != nullptr);
    for (int q = 0; ok && q < a->ckpt; q++)
        ok = (a->g1s.emplace_back(field(1)) != nullptr) && (a->g2s.emplace_back(field(1)) != nullptr);
    if (ok) ok = hipHostMalloc((void**)&a->loss_host, sizeof(double), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        release(a);
        return lsi::set_error(LS_ERR_NOMEM, "adjoint workspace allocation failed");
    }
    // The planes are written on interior cells only; their ghost regions must
    // read as zero in every gather.
    if (hipMemset(a->P, 0, 19 * a->fb) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        release(a);
        return lsi::set_error(LS_ERR_DEVICE, "adjoint workspace initialisation failed");
    }
    *out = a;
    return LS_OK;
}

int ls_adjoint_destroy(ls_adjoint_t* a) {
    if (a) release(a);
    return LS_OK;
}

int64_t ls_adjoint_workspace_bytes(const ls_adjoint_t* a) {
    if (!a) return -1;
    // every allocation is one field except P (19 fields) and the loss partials
    return (int64_t)(a->all.size() - 1 + 18) * (int64_t)a->fb + (kLossBlocks + 1) * (int64_t)sizeof(double);
}

int ls_adjoint_loss(ls_solver_t* s, const double* G, const double* G_target, double eps_h, double* loss,
                    void* stream) {
    if (!s || !G || !G_target || !loss) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    const ls_params_t& p = lsi::solver_params(s);
    if (p.dim != 3 || p.nghost != 3) return lsi::set_error(LS_ERR_INVALID, "the loss covers 3D fields with 3 ghosts");
    if (!(eps_h > 0.0)) return lsi::set_error(LS_ERR_INVALID, "eps_h must be > 0");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const AGeom g = make_ageom(p);
    // the partials go to the solver's reduction scratch (allocated by ls_create)
    size_t sbytes = 0;
    double* part = static_cast<double*>(lsi::solver_scratch(s, &sbytes));
    static_assert(kLossBlocks > 0, "loss blocks");
    if (sbytes < (kLossBlocks + 1) * sizeof(double)) return lsi::set_error(LS_ERR_INVALID, "solver scratch too small");
    hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(256), 0, st, g, G, G_target, eps_h, (double*)nullptr, part);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, st, part, part + kLossBlocks);
    double h = 0.0;
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(&h, part + kLossBlocks, sizeof(double), hipMemcpyDeviceToHost, st) == hipSuccess &&
                    hipStreamSynchronize(st) == hipSuccess;
    if (!ok) return lsi::set_error(LS_ERR_DEVICE, "loss evaluation failed");
    *loss = h;
    return LS_OK;
}

int ls_adjoint_run(ls_adjoint_t* a, const double* G0, const double* u, const double* v, const double* w,
                   const double* sl_field, const double* G_target, double eps_h, double dt, double* loss,
                   double* gG0, double* gu, double* gv, double* gw, double* gsl, void* stream) {
    if (!a || !G0 || !u || !v || !w || !G_target || !loss) return lsi::set_error(LS_ERR_INVALID, "NULL argument");
    if (!(eps_h > 0.0) || !(dt > 0.0)) return lsi::set_error(LS_ERR_INVALID, "eps_h and dt must be > 0");
    const bool grad = gu || gv || gw || gsl || gG0;
    if (grad && (!gu || !gv || !gw)) return lsi::set_error(LS_ERR_INVALID, "gu, gv and gw are required for a gradient");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    ls_solver_t* s = a->s;
    const AGeom& g = a->g;
    const size_t fb = a->fb;
    const int ri = a->p.reinit_interval;
    const int iters = a->iters;
    const int ckpt = a->ckpt;
    const double sl0 = a->p.s_l;
    const dim3 cblk(kAX, kAY, 1);
    const dim3 lgrid(1024), lblk(256);
    const dim3 fgrid((unsigned)((shell_cells(g.nx, g.ny, g.nz) + 255) / 256));
    // z-marching stage adjoint: its block sums and spills are P[14..18]
    // (planes nothing else writes)
    const ZMGeom zm = a->zm;
    const dim3 zgrid((unsigned)(zm.nsx * zm.nty * zm.nch));
    double* core = a->P + 14 * g.n;
    double* EX = a->P + 15 * g.n;
    double* EZ = a->P + 16 * g.n;
    double* EYlo = a->P + 17 * g.n;
    double* EYhi = a->P + 18 * g.n;
    int rc = LS_OK;
#define ADJ_HIP(call)                                                          \
    do {                                                                       \
        if ((call) != hipSuccess) return lsi::set_error(LS_ERR_DEVICE, #call); \
    } while (0)
#define ADJ_RC(call)                  \
    do {                              \
        if ((rc = (call))) return rc; \
    } while (0)

    auto reinit_after = [&](int n) { return iters > 0 && n > 0 && n % ri == 0; };

    // Forward step n from `from` into `to`, its stage fields into g1, g2; with
    // reinitialisation the RK3 result goes to `prestate` and the HCR-2
    // iterates through hit[].
    auto fwd_step = [&](const double* from, double* to, double* prestate, int n, double* g1, double* g2) -> int {
        ADJ_RC(ls_rk3_stage(s, 1, from, from, g1, u, v, w, sl_field, dt, st));
        ADJ_RC(ls_rk3_stage(s, 2, g1, from, g2, u, v, w, sl_field, dt, st));
        if (!reinit_after(n)) return ls_rk3_stage(s, 3, g2, from, to, u, v, w, sl_field, dt, st);
        ADJ_RC(ls_rk3_stage(s, 3, g2, from, prestate, u, v, w, sl_field, dt, st));
        const double* in = prestate;
        for (int it = 0; it < iters; it++) {
            double* o = it + 1 == iters ? to : a->hit[it + 1];
            ADJ_RC(ls_hcr2_step(s, in, prestate, o, st));
            in = o;
        }
        return LS_OK;
    };

    // Forward sweep: checkpoints and G_T.
    ADJ_HIP(hipMemcpyAsync(a->ck[0], G0, fb, hipMemcpyDeviceToDevice, st));
    ADJ_RC(ls_fill_ghosts(s, a->ck[0], st));
    const double* cur = a->ck[0];
    for (int n = 0; n < a->nsteps; n++) {
        const bool to_ck = (n + 1) % ckpt == 0 && n + 1 < a->nsteps;
        double* nxt = to_ck ? a->ck[(n + 1) / ckpt] : (cur == a->seg[0] ? a->seg[1] : a->seg[0]);
        ADJ_RC(fwd_step(cur, nxt, iters > 0 ? a->pre[0] : nullptr, n, a->G1, a->G2));
        cur = nxt;
    }
    double* lamT = grad ? a->lam : nullptr;
    if (grad) ADJ_HIP(hipMemsetAsync(a->lam, 0, fb, st));
    hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(256), 0, st, g, cur, G_target, eps_h, lamT, a->part);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, st, a->part, a->part + kLossBlocks);
    ADJ_HIP(hipGetLastError());
    ADJ_HIP(hipMemcpyAsync(a->loss_host, a->part + kLossBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
    if (!grad) {
        ADJ_HIP(hipStreamSynchronize(st));
        *loss = *a->loss_host;
        return LS_OK;
    }
    ADJ_HIP(hipMemsetAsync(gu, 0, fb, st));
    ADJ_HIP(hipMemsetAsync(gv, 0, fb, st));
    ADJ_HIP(hipMemsetAsync(gw, 0, fb, st));
    if (gsl) ADJ_HIP(hipMemsetAsync(gsl, 0, fb, st));

    // Reverse of one stage out = A G_n + B G_s + C dt RHS(G_s) followed by a
    // ghost fill.  lam_in: adjoint of out (its ghosts are folded here).  The
    // adjoint of G_s goes to lam_s (overwritten); A lam_in is added to acc
    // (stage 3, the first of a step's reverse, starts acc with it).  Stage 1
    // has G_s = G_n: everything goes to acc.
    auto stage_vjp = [&](int stage, const double* Gs, double* lam_in, double* lam_s) -> int {
        const double A = stage == 1 ? 1.0 : (stage == 2 ? 0.75 : 1.0 / 3.0);
        const double B = stage == 1 ? 0.0 : (stage == 2 ? 0.25 : 2.0 / 3.0);
        const double C = stage == 1 ? 1.0 : (stage == 2 ? 0.25 : 2.0 / 3.0);
        hipLaunchKernelGGL(k_fold, fgrid, lblk, 0, st, g, lam_in);
        hipLaunchKernelGGL(k_adj_zm, zgrid, cblk, 0, st, g, zm, Gs, u, v, w, sl_field, sl0, lam_in, C * dt, core,
                           EX, EYlo, EYhi, EZ, gu, gv, gw, gsl);
        if (stage == 1)
            hipLaunchKernelGGL(k_adj_combine, lgrid, lblk, 0, st, g, core, EX, EYlo, EYhi, EZ, lam_in, A + B, a->acc,
                               1, (double*)nullptr, 0.0, 0);
        else
            hipLaunchKernelGGL(k_adj_combine, lgrid, lblk, 0, st, g, core, EX, EYlo, EYhi, EZ, lam_in, B, lam_s, 0,
                               a->acc, A, stage == 3 ? 1 : 0);
        ADJ_HIP(hipGetLastError());
        return LS_OK;
    };

    // Reverse of a reinitialisation: a->lam holds the adjoint of its output;
    // on return it holds the adjoint of phi_0 (the pre-reinit state).
    auto hcr2_vjp = [&](const double* phi0) -> int {
        // lam2: adjoint of phi_0 through the r~ and sign terms, started by the
        // first (last-iteration) combine
        for (int it = iters - 1; it >= 0; it--) {
            const double* in = it == 0 ? phi0 : a->hit[it];
            hipLaunchKernelGGL(k_fold, fgrid, lblk, 0, st, g, a->lam);
            hipLaunchKernelGGL(k_adj_hcr2_zm, zgrid, cblk, 0, st, g, zm, in, phi0, a->lam, a->P);
            hipLaunchKernelGGL(k_hcr2_combine, lgrid, lblk, 0, st, g, a->P, a->lam3, a->lam2,
                               it == iters - 1 ? 1 : 0);
            std::swap(a->lam, a->lam3);
        }
        // the first iterate is phi_0 itself
        hipLaunchKernelGGL(k_axpy, lgrid, lblk, 0, st, g.n, 1.0, a->lam2, a->lam);
        ADJ_HIP(hipGetLastError());
        return LS_OK;
    };

    // Reverse sweep, one checkpoint segment at a time.
    for (int sg = a->nseg - 1; sg >= 0; sg--) {
        const int n0 = sg * ckpt, n1 = std::min(a->nsteps, n0 + ckpt);
        const double* start = a->ck[sg];
        for (int n = n0; n < n1; n++) {
            const double* from = n == n0 ? start : a->seg[n - n0];
            ADJ_RC(fwd_step(from, a->seg[n - n0 + 1], iters > 0 ? a->pre[n - n0] : nullptr, n, a->g1s[n - n0],
                            a->g2s[n - n0]));
        }
        for (int n = n1 - 1; n >= n0; n--) {
            const int q = n - n0;
            const double* Gn = q == 0 ? start : a->seg[q];
            if (reinit_after(n)) {
                const double* in = a->pre[q];
                for (int it = 1; it < iters; it++) {
                    ADJ_RC(ls_hcr2_step(s, in, a->pre[q], a->hit[it], st));
                    in = a->hit[it];
                }
                ADJ_RC(hcr2_vjp(a->pre[q]));
            }
            ADJ_RC(stage_vjp(3, a->g2s[q], a->lam, a->lam2));   // lam2: adjoint of G2; starts acc
            ADJ_RC(stage_vjp(2, a->g1s[q], a->lam2, a->lam3));  // lam3: adjoint of G1
            ADJ_RC(stage_vjp(1, Gn, a->lam3, nullptr));
            std::swap(a->lam, a->acc);  // adjoint of G_n
        }
    }
    // G0's ghosts were filled from its interior before the first step.
    hipLaunchKernelGGL(k_fold, fgrid, lblk, 0, st, g, a->lam);
    ADJ_HIP(hipGetLastError());
    if (gG0) ADJ_HIP(hipMemcpyAsync(gG0, a->lam, fb, hipMemcpyDeviceToDevice, st));
    ADJ_HIP(hipStreamSynchronize(st));
    *loss = *a->loss_host;
    return LS_OK;
#undef ADJ_HIP
#undef ADJ_RC
}

int ls_adjoint_run_sets(ls_adjoint_t* const* ctx, void* const* streams, int32_t nctx, const ls_adjoint_set_t* sets,
                        int32_t nsets) {
    if (!ctx || !sets || nctx < 1 || nsets < 0) return lsi::set_error(LS_ERR_INVALID, "bad context or set list");
    for (int c = 0; c < nctx; c++) {
        if (!ctx[c]) return lsi::set_error(LS_ERR_INVALID, "NULL adjoint context");
        for (int d = 0; d < c; d++)
            if (ctx[d] == ctx[c] || ctx[d]->s == ctx[c]->s)
                return lsi::set_error(LS_ERR_INVALID, "contexts of one call must not share a context or solver");
    }
    auto run_one = [&](int c, int q) {
        const ls_adjoint_set_t& e = sets[q];
        return ls_adjoint_run(ctx[c], e.G0, e.u, e.v, e.w, e.sl_field, e.G_target, e.eps_h, e.dt, e.loss, e.gG0,
                              e.gu, e.gv, e.gw, e.gsl, streams ? streams[c] : nullptr);
    };
    const int nthr = std::min<int>(nctx, nsets);
    if (nthr <= 1) {
        for (int q = 0; q < nsets; q++) {
            const int rc = run_one(0, q);
            if (rc) return rc;
        }
        return LS_OK;
    }
    // One host thread per context; context c takes sets c, c + nctx, ... in
    // order on its own stream.  The first failure (lowest set index) is
    // reported with its message.
    std::mutex mu;
    int err_q = nsets, err_rc = LS_OK;
    std::string err_msg;
    std::vector<std::thread> pool;
    for (int c = 0; c < nthr; c++) {
        pool.emplace_back([&, c]() {
            if (hipSetDevice(lsi::solver_device(ctx[c]->s)) != hipSuccess) {
                std::lock_guard<std::mutex> lock(mu);
                if (c < err_q) { err_q = c; err_rc = LS_ERR_DEVICE; err_msg = "hipSetDevice failed in a set worker"; }
                return;
            }
            for (int q = c; q < nsets; q += nctx) {
                const int rc = run_one(c, q);
                if (rc) {
                    std::lock_guard<std::mutex> lock(mu);
                    if (q < err_q) { err_q = q; err_rc = rc; err_msg = ls_last_error(); }
                    return;
                }
            }
        });
    }
    for (auto& t : pool) t.join();
    if (err_rc) return lsi::set_error(err_rc, ("parameter set " + std::to_string(err_q) + ": " + err_msg).c_str());
    return LS_OK;
}

}  // extern "C"
