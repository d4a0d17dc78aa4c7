// ls_fast_math.hpp -- device arithmetic of the fast mode, shared by the
// single-grid and batched kernels.  See ls_fast.hip for the derivation.
#pragma once
#include <hip/hip_runtime.h>

#include "ls_exact.hpp"

namespace lsf {

// Reciprocal with two Newton-Raphson refinements: close to 1 ulp.
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// a / b to about 1 ulp: reciprocal, quotient and one residual correction.
__device__ __forceinline__ double div_fast(double a, double b) {
    double y = rcp_nr(b);
    double q = a * y;
    double r = fma(-b, q, a);
    return fma(r, y, q);
}

// 1/sqrt(x) to about 1 ulp (x > 0).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * x;
    double e = fma(-h * y, y, 0.5);
    y = fma(y, e, y);
    e = fma(-h * y, y, 0.5);
    return fma(y, e, y);
}

// t = 16 (eps + beta)^2 of the 3-point sub-stencil (a, b, c) (weno5.h:28-37),
// as (4 eps + 4 beta)^2 with 4 beta = (13/3) d^2 + e^2.  The faces use the
// T's only in ratios, and scaling by a power of two commutes with rounding,
// so the weights are those of (eps + beta)^2 bit for bit, one product fewer.
__device__ __forceinline__ double tsq(double a, double b, double c) {
    const double d = fma(-2.0, b, a + c);
    const double e = c - a;
    const double beta4 = fma((13.0 / 3.0) * d, d, e * e);
    const double s = 4.0 * lsx::kWenoEps + beta4;
    return s * s;
}

// ---------------------------------------------------------------------------
// Shared faces (k_stage_zface and the 2D y-march).  A face is kept as a
// numerator n and denominator d, value n / (6 d).
struct Face {
    double n, d;
};

// Face between cells c and c+1.  v0..v4 run from the far upwind side: G(c-2..c+2)
// left-biased, G(c+3..c-1) mirrored; t0..t2 are T at the centres of v1..v3.
// The linear weights (0.1, 0.6, 0.3) / (0.3, 0.6, 0.1) are scaled by 10,
// a factor common to numerator and denominator.
template <bool MIRROR>
__device__ __forceinline__ Face face_core(double v0, double v1, double v2, double v3, double v4, double t0,
                                          double t1, double t2) {
    const double p0 = fma(2.0, v0, fma(-7.0, v1, 11.0 * v2));
    const double p1 = fma(5.0, v2, fma(2.0, v3, -v1));
    const double p2 = fma(2.0, v2, fma(5.0, v3, -v4));
    const double w0 = MIRROR ? 3.0 * (t1 * t2) : t1 * t2;
    const double w1 = 6.0 * (t0 * t2);
    const double w2 = MIRROR ? t0 * t1 : 3.0 * (t0 * t1);
    Face f;
    f.n = fma(w0, p0, fma(w1, p1, w2 * p2));
    f.d = (w0 + w1) + w2;
    return f;
}

// The face c+1/2 from LDS: g points at G(c), t at T(c); GS and TS are the
// strides of the axis in g and t.
template <bool MIRROR, int GS, int TS>
__device__ __forceinline__ Face face_lds(const double* g, const double* t) {
    if (MIRROR) return face_core<true>(g[3 * GS], g[2 * GS], g[GS], g[0], g[-GS], t[2 * TS], t[TS], t[0]);
    return face_core<false>(g[-2 * GS], g[-GS], g[0], g[GS], g[2 * GS], t[-TS], t[0], t[TS]);
}

// Stores the face c+1/2 in each orientation asked for: slot[0] left-biased,
// slot[OS] mirrored.
template <int GS, int TS, int OS>
__device__ __forceinline__ void face_put(const double* g, const double* t, bool needL, bool needR, double2* slot) {
    if (needL) {
        const Face f = face_lds<false, GS, TS>(g, t);
        slot[0] = make_double2(f.n, f.d);
    }
    if (needR) {
        const Face f = face_lds<true, GS, TS>(g, t);
        slot[OS] = make_double2(f.n, f.d);
    }
}

}  // namespace lsf
