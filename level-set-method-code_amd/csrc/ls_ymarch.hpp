// ls_ymarch.hpp -- the 2D fast-mode RK3 stage as a y-march of one wavefront,
// shared by the single-grid solver and the batched engine.
//
// A wavefront owns a strip of 64 columns (one cell per lane) and walks a
// chunk of rows.  The column's 7 y-values and 5 y smoothness terms sit in
// registers and shift by one row per step, so a G value is loaded once per
// chunk (plus 6 halo rows per chunk and 6 halo columns per row).  The current
// row, with its 3-cell x halo, goes through LDS for the x stencil.  Each WENO
// face is reconstructed once, in the orientations its two cells need, and
// shared between them: x faces through LDS, the y face carried in registers to
// the next row, which rebuilds it only where its orientation differs.  Both
// axes go over one denominator, so a cell costs one refined reciprocal.  The
// arithmetic is the fast mode's (ls_fast_math.hpp), or with EXACT the
// reference's (weno5_left / weno5_right values, its divisions and expression
// order, bitwise equal to the oracle).  Either way the result does not depend
// on the strip or chunk partition, since every face is a pure function of
// its stencil.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ls_api.h"
#include "ls_exact.hpp"
#include "ls_fast_math.hpp"

namespace lsf {

constexpr int kYW = 64;  // strip width: one column per lane of a wavefront

// The a-th padded index that holds interior index q after a ghost fill
// (a = 0 is the cell itself), and how many there are (periodic wrap or
// zero-gradient clamp, boundary.h:98-168).
__device__ __forceinline__ int n_images(int q, int n, int bc) {
    if (bc == LS_BC_PERIODIC) return 1 + (q < 3) + (q >= n - 3);
    if (bc == LS_BC_ZERO_GRAD) return 1 + 3 * (q == 0) + 3 * (q == n - 1);
    return 1;
}
__device__ __forceinline__ int image(int q, int n, int bc, int a) {
    if (a == 0) return q + 3;
    if (bc == LS_BC_PERIODIC) return (a == 1 && q < 3) ? q + 3 + n : q + 3 - n;
    return q == 0 ? a - 1 : n + 2 + a;  // zero-gradient: 0,1,2 or n+3,n+4,n+5
}

// One 2D grid as the y-march sees it.
struct YGrid {
    int32_t nx, ny, bc0, bc1;
    int32_t fuse_ghosts;  // write the ghost images of the output (periodic / zero-gradient axes only)
    int32_t pad;
    int64_t sy, off;      // padded row pitch; element offset of padded (0, 0)
    double s_l, eps, dt;
    double inv2dx, inv2dy, inv6dx, inv6dy;
    double dx, dy;        // exact mode divides by these, as the reference does
    double rdx, rdy;      // RN(1/dx), RN(1/dy) for lsx::qdiv_u
    int32_t crdiv;        // spacings, eps and a scalar S_L admit the FMA division sequences
    int32_t pad2;
};

// Rows the y-march loads ahead of the one it computes (1 or 2).
#ifndef LS_YMARCH_PF
#define LS_YMARCH_PF 1
#endif

// What the y-march loads for one row of its strip.
struct YRow {
    double g;    // the ring's next value G(i, j + 3)
    double u, v, gn, sl;
    double hal;  // this lane's x-halo value (lanes 0..5)
};

// LDS of one wavefront.
template <bool EXACT>
struct YLds {
    double g[kYW + 8];                    // G of the current row at x = x0 - 3 + p
    double t[EXACT ? 1 : kYW + 8];        // fast: smoothness term T at the same positions
    double2 s[EXACT ? 2 : 1][EXACT ? kYW + 8 : 1];  // exact: sub-stencil data (t, 1/t) there,
                                                    // [0] left-biased, [1] mirrored (lsx::sub_l / sub_r)
    double2 f[2][kYW + 1];                // x face x0 - 1/2 + q as (n, d): [0] left-biased, [1] mirrored
    unsigned char ori[kYW];               // upwind orientation along x of each lane's cell
};

// Fast stage combine (rk3.h:103-150 with FMA).
template <int STAGE>
__device__ __forceinline__ double ystage(double gn, double gc, double rr, double dt) {
    if (STAGE == 1) return fma(dt, rr, gn);
    if (STAGE == 2) return fma(0.25 * dt, rr, fma(0.25, gc, 0.75 * gn));
    return fma((2.0 / 3.0) * dt, rr, fma(2.0 / 3.0, gc, (1.0 / 3.0) * gn));
}

// The RHS of one cell from its two faces per axis and the effective velocity:
// -(ue (F(i+1/2)-F(i-1/2))/dx + ve (F(j+1/2)-F(j-1/2))/dy) over one denominator.
__device__ __forceinline__ double yrhs(const YGrid& m, Face xl, Face xh, Face yl, Face yh, double ue, double ve) {
    const double Nx = fma(xh.n, xl.d, -(xl.n * xh.d)), Dx = xh.d * xl.d;
    const double Ny = fma(yh.n, yl.d, -(yl.n * yh.d)), Dy = yh.d * yl.d;
    const double num = fma((ue * m.inv6dx) * Nx, Dy, ((ve * m.inv6dy) * Ny) * Dx);
    return -(num * rcp_nr(Dx * Dy));
}

// Effective velocity u - S_L grad G / sqrt(|grad G|^2 + eps) from the central
// differences (rk3.h:54-98).
__device__ __forceinline__ void yueff(const YGrid& m, double gxm, double gxp, double gym, double gyp, double s_l,
                                      double& ue, double& ve) {
    if (s_l > m.eps) {
        const double gx = (gxp - gxm) * m.inv2dx;
        const double gy = (gyp - gym) * m.inv2dy;
        const double k = s_l * rsqrt_nr(fma(gx, gx, gy * gy) + m.eps);
        ue = fma(-k, gx, ue);
        ve = fma(-k, gy, ve);
    }
}

// Ghost images of an interior output value (periodic / zero-gradient).
__device__ __forceinline__ void ystore_images(const YGrid& m, double* __restrict__ out, int i, int j, double o) {
    if (m.fuse_ghosts && (i < 3 || i >= m.nx - 3 || j < 3 || j >= m.ny - 3)) {
        const int nxi = n_images(i, m.nx, m.bc0), nyi = n_images(j, m.ny, m.bc1);
        for (int b = 0; b < nyi; b++) {
            const int64_t pj = m.off + (int64_t)image(j, m.ny, m.bc1, b) * m.sy;
            for (int e = 0; e < nxi; e++)
                if (b | e) out[pj + image(i, m.nx, m.bc0, e)] = o;
        }
    }
}

// The face c-1/2 and c+1/2 of one axis in one orientation, from the axis
// stencil s[0..6] = G(c-3..c+3) and t[0..4] = T(c-2..c+2).
__device__ __forceinline__ void yfaces(const double (&s)[7], const double (&t)[5], bool L, Face& lo, Face& hi) {
    if (L) {
        lo = face_core<false>(s[0], s[1], s[2], s[3], s[4], t[0], t[1], t[2]);
        hi = face_core<false>(s[1], s[2], s[3], s[4], s[5], t[1], t[2], t[3]);
    } else {
        lo = face_core<true>(s[5], s[4], s[3], s[2], s[1], t[3], t[2], t[1]);
        hi = face_core<true>(s[6], s[5], s[4], s[3], s[2], t[4], t[3], t[2]);
    }
}

// One cell with every face built for it alone: the y-march's arithmetic
// without the sharing (bitwise equal to it), for grids too small to give the
// y-march enough wavefronts.
template <int STAGE>
__device__ __forceinline__ void ycell(const YGrid& m, int i, int j, const double* __restrict__ Gs,
                                      const double* __restrict__ Gn, double* __restrict__ out,
                                      const double* __restrict__ u, const double* __restrict__ v,
                                      const double* __restrict__ sl) {
    const int64_t c = m.off + (int64_t)(j + 3) * m.sy + (i + 3);
    const double* P = Gs + c;
    double sx[7], sy[7], tx[5], ty[5];
#pragma unroll
    for (int q = 0; q < 7; q++) {
        sx[q] = P[q - 3];
        sy[q] = P[(q - 3) * m.sy];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) {
        tx[q] = tsq(sx[q], sx[q + 1], sx[q + 2]);
        ty[q] = tsq(sy[q], sy[q + 1], sy[q + 2]);
    }
    double ue = u[c], ve = v[c];
    yueff(m, sx[2], sx[4], sy[2], sy[4], sl ? sl[c] : m.s_l, ue, ve);
    Face xl, xh, yl, yh;
    yfaces(sx, tx, ue >= 0.0, xl, xh);
    yfaces(sy, ty, ve >= 0.0, yl, yh);
    const double rr = yrhs(m, xl, xh, yl, yh, ue, ve);
    const double o = ystage<STAGE>(Gs == Gn ? sx[3] : Gn[c], sx[3], rr, m.dt);
    out[c] = o;
    ystore_images(m, out, i, j, o);
}

// Rows [y0, y1) of the strip that starts at column x0.  Called by every lane
// of a one-wavefront block (blockDim.x == 64) with the same arguments.  In
// the exact mode the divisions run as the FMA sequences of ls_crdiv.hpp while
// every G and S_L value the wavefront has loaded passed lsx::crdiv_safe (a
// sticky per-lane flag and a wave vote) and m.crdiv is set; otherwise as
// divisions.  Both give the same bits.  SM says whether G_s is G_n: -1 test
// the pointers, 1 known equal (stage 1 of a step: no separate G_n load), 0
// known different.
template <int STAGE, bool EXACT = false, int SM = -1>
__device__ __forceinline__ void ymarch_strip(const YGrid& m, int x0, int y0, int y1,
                                             const double* __restrict__ Gs, const double* __restrict__ Gn,
                                             double* __restrict__ out, const double* __restrict__ u,
                                             const double* __restrict__ v, const double* __restrict__ sl,
                                             YLds<EXACT>& L) {
    const int l = (int)threadIdx.x;
    const int xe = min(x0 + kYW, m.nx);
    const int i = x0 + l;
    const bool act = i < xe;
    const int64_t sy = m.sy;
    const int64_t col = m.off + (act ? i : x0) + 3;  // padded column of this lane's cell
    // halo lanes: 0..2 hold x0-3..x0-1, 3..5 hold xe..xe+2 (ghosts at the grid edges)
    const bool hl = l < 6;
    const int hp = l < 3 ? l : xe - x0 + l;  // LDS position of the halo value
    const int64_t hcol = m.off + (l < 3 ? x0 + l : xe + l);
    const bool same = SM < 0 ? Gs == Gn : SM == 1;  // the combine's G_n is the stencil's centre
    auto prow = [&](int j) { return (int64_t)(j + 3) * sy; };

    double r[7];  // G(i, j-3 .. j+3)
#pragma unroll
    for (int q = 0; q < 7; q++) r[q] = act ? Gs[col + prow(y0 - 3 + q)] : 0.0;
    bool bad = false;  // exact mode: a value this lane loaded failed crdiv_safe
    if (EXACT)
#pragma unroll
        for (int q = 0; q < 7; q++) bad |= !lsx::crdiv_safe(r[q]);
    double ty[5];  // T(i, j-2 .. j+2) along y (fast mode)
#pragma unroll
    for (int q = 0; q < 5; q++) ty[q] = EXACT ? 0.0 : tsq(r[q], r[q + 1], r[q + 2]);
    // point data of row jr and the ring's next value G(i, jr + 3).  Fast mode
    // loads them as buffer loads: a descriptor based at the row's padded x = 0
    // (wavefront-uniform) and a 32-bit byte offset per lane fixed for the whole
    // march.  (The exact variant keeps plain loads: its scalar registers are
    // full, and the descriptors spilled them, 1.6% slower.)
    const int64_t base = m.off;  // padded (0, 0) of this grid
    const uint32_t cb = (uint32_t)(col - base) * 8u, hb = (uint32_t)(hcol - base) * 8u;
    auto ld = [](const double* row, uint32_t off) {
        if constexpr (EXACT) {
            return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(row) + off);
        } else {
            const __amdgpu_buffer_rsrc_t r =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(row), (short)0, -1, 0x00020000);
            return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
        }
    };
    auto fetch = [&](int jr, bool with_g) {
        YRow q;
        q.g = q.u = q.v = q.gn = q.hal = 0.0;
        q.sl = m.s_l;
        if (jr < y1) {
            const int64_t rj = base + prow(jr);
            if (act) {
                if (with_g) q.g = ld(Gs + base + prow(jr + 3), cb);
                q.u = ld(u + rj, cb);
                q.v = ld(v + rj, cb);
                if (!same) q.gn = ld(Gn + rj, cb);
                if (sl) q.sl = ld(sl + rj, cb);
            }
            if (hl) q.hal = ld(Gs + rj, hb);
        }
        return q;
    };
    auto st = [&](double* row, uint32_t off, double val) {
        if constexpr (EXACT) {
            *reinterpret_cast<double*>(reinterpret_cast<char*>(row) + off) = val;
        } else {
            const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(row, (short)0, -1, 0x00020000);
            typedef unsigned int u32x2 __attribute__((__vector_size__(8)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, val), r, (int)off, 0, 0);
        }
    };
    // does any lane of this strip own a cell next to an x edge (whose ghost
    // images it writes)?  Rows away from the y edges then skip the images
    // with a scalar test.
    const bool strip_xedge = __ballot(act && (i < 3 || i >= m.nx - 3)) != 0;
    YRow cur = fetch(y0, false);  // G(i, y0 + 3) is r[6]
    YRow ahead = cur;             // row j + 1 when prefetching two rows ahead
    if (LS_YMARCH_PF == 2) ahead = fetch(y0 + 1, true);
    Face ylo;
    ylo.n = ylo.d = 0.0;
    int ylo_L = -1;  // orientation ylo was built in (-1: none yet)

    for (int j = y0; j < y1; j++) {
        // --- the row into LDS; prefetch
        if (act) L.g[3 + l] = r[3];
        if (hl) L.g[hp] = cur.hal;
        bool fd = false;  // exact mode: this row's quotients by the FMA sequences
        if (EXACT) {
            bad |= !lsx::crdiv_safe(cur.hal) || (sl && !lsx::crdiv_safe(cur.sl));
            fd = m.crdiv && __ballot(bad) == 0;
        }
        const YRow fresh = fetch(j + LS_YMARCH_PF, true);
        __syncthreads();  // row
        // --- smoothness terms along x, u_eff, orientations
        if constexpr (!EXACT) {
            if (act) L.t[3 + l] = tsq(L.g[2 + l], L.g[3 + l], L.g[4 + l]);
            if (l < 4) {  // T at x0-2, x0-1, xe, xe+1
                const int p = l < 2 ? 1 + l : xe - x0 + 1 + l;
                L.t[p] = tsq(L.g[p - 1], L.g[p], L.g[p + 1]);
            }
        }
        double ue = cur.u, ve = cur.v;
        if (act && EXACT) {  // computeRHS (rk3.h:54-98), the reference's expression order
            auto ueff = [&](auto dtag) {
                constexpr int D = decltype(dtag)::value;
                double gx = 0.0, gy = 0.0, gm = 1.0;
                if (cur.sl > m.eps) {  // 2h is exact, and RN(1/(2h)) = RN(1/h) / 2
                    gx = lsx::dq<D>(L.g[4 + l] - L.g[2 + l], 2.0 * m.dx, 0.5 * m.rdx);
                    gy = lsx::dq<D>(r[4] - r[2], 2.0 * m.dy, 0.5 * m.rdy);
                    gm = sqrt(gx * gx + gy * gy + m.eps);
                }
                const double ygm = D == lsx::kDivIEEE ? 0.0 : lsx::dr<D>(1.0, gm);
                ue = cur.u - lsx::dq<D>(cur.sl * gx, gm, ygm);
                ve = cur.v - lsx::dq<D>(cur.sl * gy, gm, ygm);
            };
            if (fd) ueff(std::integral_constant<int, lsx::kDivFma>());
            else ueff(std::integral_constant<int, lsx::kDivIEEE>());
        } else if (act) {
            yueff(m, L.g[2 + l], L.g[4 + l], r[2], r[4], cur.sl, ue, ve);
        }
        const bool Lx = ue >= 0.0, Ly = ve >= 0.0;
        L.ori[l] = (unsigned char)Lx;
        if constexpr (EXACT) {
            // the x sub-stencils' (t, 1/t), once per centre for the three faces
            // of an orientation that use them; the strip's x faces take the
            // orientations of its cells, so only the sets those use
            if (fd) {
                const bool wl = __ballot(act && Lx) != 0, wr = __ballot(act && !Lx) != 0;
                auto sub = [&](int p) {  // centre L.g[p]
                    const double gm = L.g[p - 1], g0 = L.g[p], gp = L.g[p + 1];
                    if (wl) L.s[0][p] = lsx::sub_l(gm, g0, gp);
                    if (wr) L.s[1][p] = lsx::sub_r(gm, g0, gp);
                };
                if (act) sub(3 + l);
                if (l < 4) sub(l < 2 ? 1 + l : xe - x0 + 1 + l);  // centres x0-2, x0-1, xe, xe+1
            }
        }
        __syncthreads();  // T, orientations
        // --- faces
        Face yhi;
        yhi.n = yhi.d = 0.0;
        if (act) {
            // x face i+1/2 for this cell and its high neighbour
            bool nL = Lx, nR = !Lx;
            if (i + 1 < xe) {
                const bool o = L.ori[l + 1];
                nL = nL || o;
                nR = nR || !o;
            }
            if constexpr (EXACT) {  // weno5_left / weno5_right values (weno5.h:62-122); .x holds the value
                // orientations are per-lane choices of the stencil order
                // (lsx::weno_u), not branches: every lane builds the first x
                // face it needs, lanes that need both then the mirrored one
                auto faces = [&](auto dtag) {
                    constexpr int D = decltype(dtag)::value;
                    const double* g = &L.g[3 + l];
                    auto xf = [&](bool mir) {  // weno_l(g[-2..2]) or weno_r(g[-1..3])
                        const double* b = mir ? g + 3 : g - 2;
                        const int st = mir ? -1 : 1;
                        if constexpr (D == lsx::kDivFma) {
                            // beta_0..2 at the centres g[-1], g[0], g[1] (left-biased set)
                            // or g[2], g[1], g[0] (mirrored set)
                            const double2* sb = mir ? &L.s[1][3 + l + 2] : &L.s[0][3 + l - 1];
                            return lsx::weno_tu(b[0], b[st], b[2 * st], b[3 * st], b[4 * st], sb[0], sb[st],
                                                sb[2 * st], mir);
                        } else {
                            return lsx::weno_u<D>(b[0], b[st], b[2 * st], b[3 * st], b[4 * st], mir);
                        }
                    };
                    const bool m1 = !nL;
                    L.f[m1 ? 1 : 0][l + 1].x = xf(m1);
                    if (nL && nR) L.f[1][l + 1].x = xf(true);
                    const bool m = !Ly;
                    if (ylo_L != (int)Ly)
                        ylo.n = lsx::weno_u<D>(m ? r[5] : r[0], m ? r[4] : r[1], m ? r[3] : r[2], m ? r[2] : r[3],
                                               m ? r[1] : r[4], m);
                    yhi.n = lsx::weno_u<D>(m ? r[6] : r[1], m ? r[5] : r[2], m ? r[4] : r[3], m ? r[3] : r[4],
                                           m ? r[2] : r[5], m);
                };
                if (fd) faces(std::integral_constant<int, lsx::kDivFma>());
                else faces(std::integral_constant<int, lsx::kDivIEEE>());
            } else {
                face_put<1, 1, kYW + 1>(&L.g[3 + l], &L.t[3 + l], nL, nR, &L.f[0][l + 1]);
                // y face j-1/2 in this cell's orientation (carried unless it differs), and j+1/2
                if (ylo_L != (int)Ly) {
                    if (Ly) ylo = face_core<false>(r[0], r[1], r[2], r[3], r[4], ty[0], ty[1], ty[2]);
                    else ylo = face_core<true>(r[5], r[4], r[3], r[2], r[1], ty[3], ty[2], ty[1]);
                }
                if (Ly) yhi = face_core<false>(r[1], r[2], r[3], r[4], r[5], ty[1], ty[2], ty[3]);
                else yhi = face_core<true>(r[6], r[5], r[4], r[3], r[2], ty[4], ty[3], ty[2]);
            }
        }
        if (l == 0) {  // x face x0-1/2 in the orientation of the strip's first cell
            const bool m = !Lx;  // weno_l(L.g[0..4]) or weno_r(L.g[1..5])
            if constexpr (EXACT) {
                if (fd) {  // centres L.g[1..3] (left-biased set) or L.g[4..2] (mirrored set)
                    const double2* sb = m ? &L.s[1][4] : &L.s[0][1];
                    const int st = m ? -1 : 1;
                    L.f[m ? 1 : 0][0].x = lsx::weno_tu(L.g[m ? 5 : 0], L.g[m ? 4 : 1], L.g[m ? 3 : 2], L.g[m ? 2 : 3],
                                                       L.g[m ? 1 : 4], sb[0], sb[st], sb[2 * st], m);
                } else
                    L.f[m ? 1 : 0][0].x = lsx::weno_u<lsx::kDivIEEE>(L.g[m ? 5 : 0], L.g[m ? 4 : 1], L.g[m ? 3 : 2],
                                                                     L.g[m ? 2 : 3], L.g[m ? 1 : 4], m);
            } else {
                face_put<1, 1, kYW + 1>(&L.g[2], &L.t[2], Lx, !Lx, &L.f[0][0]);
            }
        }
        __syncthreads();  // faces
        // --- derivatives over one denominator, RHS, stage combine
        double o2 = 0.0;
        if (act) {
            const int o = Lx ? 0 : 1;
            const double2 fl = L.f[o][l], fh = L.f[o][l + 1];
            if (EXACT) {  // (F(i+1/2) - F(i-1/2)) / h, rhs and the stage in the reference's order
                const double ax = fh.x - fl.x, ay = yhi.n - ylo.n;
                const double dgx = fd ? lsx::qdiv_u(ax, m.dx, m.rdx) : ax / m.dx;
                const double dgy = fd ? lsx::qdiv_u(ay, m.dy, m.rdy) : ay / m.dy;
                const double rr = -(ue * dgx + ve * dgy);
                o2 = lsx::stage_combine<STAGE>(same ? r[3] : cur.gn, r[3], rr, m.dt);
            } else {  // both axes over one denominator
                Face xl, xh;
                xl.n = fl.x;
                xl.d = fl.y;
                xh.n = fh.x;
                xh.d = fh.y;
                const double rr = yrhs(m, xl, xh, ylo, yhi, ue, ve);
                o2 = ystage<STAGE>(same ? r[3] : cur.gn, r[3], rr, m.dt);
            }
            st(out + base + prow(j), cb, o2);
        }
        // ghost cells of out: written by the interior cell they copy
        if (m.fuse_ghosts && (strip_xedge || j < 3 || j >= m.ny - 3)) {
            if (act) ystore_images(m, out, i, j, o2);
        }
        ylo = yhi;
        ylo_L = (int)Ly;
#pragma unroll
        for (int q = 0; q < 6; q++) r[q] = r[q + 1];
        if (LS_YMARCH_PF == 2) {
            cur = ahead;
            ahead = fresh;
        } else {
            cur = fresh;
        }
        r[6] = cur.g;
        if (EXACT) bad |= !lsx::crdiv_safe(r[6]);
#pragma unroll
        for (int q = 0; q < 4; q++) ty[q] = ty[q + 1];
        ty[4] = EXACT ? 0.0 : tsq(r[4], r[5], r[6]);
    }
}

}  // namespace lsf
