// ls_crdiv.hpp -- correctly rounded FP64 quotients from multiplies and FMAs.
//
// The exact mode must reproduce the reference's IEEE divisions bit for bit.
// On gfx950 a division a/b is an 11-instruction sequence (two v_div_scale,
// the quarter-rate v_rcp_f64, Newton FMAs, v_div_fmas, v_div_fixup).  Two
// cases of the WENO-5 scheme admit cheaper sequences with the same correctly
// rounded result (round to nearest even):
//
//  * div6_u(x) = RN(x/6).  q = RN(x c) with c = RN(1/6), then one correction
//    q' = RN(q + r c) with the exact residual r = x - 6q.  Why it is exact:
//    x/6 = (x/2)/3, and x/2 is a multiple of 2^k for a k at most two below the
//    ulp u of the quotient, so x/6 is a multiple of u/3 and lies at least u/6
//    from every rounding midpoint.  q is within 1.5 u of x/6, so r is a small
//    multiple of u/2 (exact in the FMA), and q + r c differs from x/6 by
//    |r| |c - 1/6| <= 9 u 2^-56, far below u/6: it rounds as x/6 does.
//
//  * qdiv_u(a, b, y) = RN(a/b) given y = RN(1/b), for quotients that share a
//    divisor (the three nonlinear weights a_k / s, the axis derivatives F/h,
//    the effective-velocity terms over |grad G|).  q0 = RN(a y) is within
//    1.5 ulp; one correction q1 = RN(q0 + RN(a - b q0) y) is faithful; the
//    second is Markstein's theorem (P. Markstein, "Computation of elementary
//    functions on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1),
//    1990): for faithful q1 and y = RN(1/b), r1 = a - b q1 is exact and
//    RN(q1 + r1 y) = RN(a/b).
//
// Both arguments need every intermediate in the normal range; the _u forms
// state operand ranges that guarantee it and do not check them.  div6 and
// qdiv check (a sufficient part of) those ranges and take the division
// outside them; the host tests run the sequences through them.  The
// corrections are written q + (b q - a)(-y), not q + (a - b q) y, so that a
// zero numerator keeps its sign as the division does (b q - a is +0 for a
// zero a, and +0 times a negative factor is -0); the divisor must be > 0.
//
// The marching stage kernels do not check per quotient.  They check each G
// and S_L value once, as it is loaded (crdiv_safe), and use the _u forms
// while every value so far passed; the grid spacings and eps are checked on
// the host (crdiv_geom_ok).  For G, S_L in {0} U [2^-200, 2^50] and spacings
// in [2^-60, 2^60], eps in [2^-200, 2^50], every quotient of the stage falls
// in the _u ranges: the WENO numerators are 0 or in [2^-252, 2^55] (integer
// combinations of the stencil, multiples of its smallest quantum); the
// smoothness terms (eps + beta)^2 lie in [2^-40, 2^214], so the weights'
// numerators and sum lie in [2^-218, 2^42] and the weights in [2^-260, 1];
// the face values are 0 or in [2^-567, 2^55] and so are the differences of
// two, whose quotients by h are 0 or in [2^-627, 2^116]; central differences
// are 0 or in [2^-252, 2^51], |grad G| (with its eps) is in [2^-100, 2^117],
// and S_L times a gradient is 0 or in [2^-827, 2^166], over |grad G| at least
// 2^-944.
//
// The functions are host+device so that the CPU tests check them against the
// host's division (tests/test_crdiv.py).
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define LS_CRDIV_FN __host__ __device__ __forceinline__
#else
#define LS_CRDIV_FN static inline
#endif

namespace lsx {

// RN(x/6) for x = 0 or |x| in [2^-960, 2^1000].
LS_CRDIV_FN double div6_u(double x) {
    constexpr double c = 1.0 / 6.0;  // RN(1/6)
    const double q = x * c;
    return fma(fma(q, 6.0, -x), -c, q);
}

// RN(a/b) given y = RN(1/b), for b in [2^-300, 2^300] and a = 0 or
// |a| in [2^-960, 2^960] with |a|/b in [2^-1000, 2^1000].
LS_CRDIV_FN double qdiv_u(double a, double b, double y) {
    const double q0 = a * y;
    const double q1 = fma(fma(q0, b, -a), -y, q0);
    return fma(fma(q1, b, -a), -y, q1);
}

LS_CRDIV_FN double div6(double x) {
    const double m = fabs(x);
    if (!((m >= 0x1p-960 && m <= 0x1p1000) || m == 0.0)) return x / 6.0;
    return div6_u(x);
}

// (checks a sufficient part of qdiv_u's condition: |a| in [2^-600, 2^600])
LS_CRDIV_FN double qdiv(double a, double b, double y) {
    const double m = fabs(a);
    if (!(((m >= 0x1p-600 && m <= 0x1p600) || m == 0.0) && b >= 0x1p-300 && b <= 0x1p300)) return a / b;
    return qdiv_u(a, b, y);
}

// The value class under which the marching kernels use the _u forms.
LS_CRDIV_FN bool crdiv_safe(double v) {
    const double m = fabs(v);
    return (m >= 0x1p-200 && m <= 0x1p50) || m == 0.0;
}

// Host-side part of the same condition: spacings and the gradient epsilon.
LS_CRDIV_FN bool crdiv_geom_ok(double dx, double dy, double dz, double eps) {
    return dx >= 0x1p-60 && dx <= 0x1p60 && dy >= 0x1p-60 && dy <= 0x1p60 && dz >= 0x1p-60 && dz <= 0x1p60 &&
           eps >= 0x1p-200 && eps <= 0x1p50;
}

// Division flavours of the exact-mode helpers: the IEEE division itself, or
// the unchecked sequences (the caller guarantees their ranges).
enum : int { kDivIEEE = 0, kDivFma = 1 };

template <int D>
LS_CRDIV_FN double d6(double x) {
    if constexpr (D == kDivIEEE) return x / 6.0;
    else return div6_u(x);
}

template <int D>
LS_CRDIV_FN double dq(double a, double b, double y) {
    if constexpr (D == kDivIEEE) return a / b;
    else return qdiv_u(a, b, y);
}

#if defined(__HIPCC__)
// RN(a/b) for a quotient with a divisor of its own (the WENO weights'
// numerators d_k / (eps + beta_k)^2, and the reciprocals 1/s and 1/|grad G|
// that qdiv_u takes as y).  It is the sequence the compiler emits for a / b
// on gfx950 -- v_rcp_f64, two Newton steps, the quotient and one residual
// correction -- without the range scaling (two v_div_scale) and the
// special-case fix-up (v_div_fixup) around it.  v_div_scale returns its
// operand unchanged, and leaves VCC clear so that v_div_fmas is a plain FMA,
// unless b or 1/b is subnormal, a/b is subnormal, a is zero or below
// 2^-969, or the exponent of a exceeds that of b by 768 or more;
// v_div_fixup returns the quotient unchanged for finite nonzero operands
// whose quotient is normal.  Inside those ranges both sequences compute the
// same values, so this one returns the IEEE quotient too.  A zero numerator
// (which v_div_scale hands to v_div_fixup) keeps its sign through the
// correction's form (b q - a)(-y), as in qdiv_u.  Range: b in
// [2^-1000, 2^1000], a = 0 or |a| in [2^-960, 2^960], |a|/b in
// [2^-1000, 2^700]; the exact-mode uses stay far inside it (divisors
// (eps + beta)^2 in [2^-40, 2^214], weight sums in [2^-218, 2^42], |grad G|
// in [2^-100, 2^117]; numerators 0.1, 0.3, 0.6 and 1).  Device only.
//
// The sequence splits into a part that reads b alone (rcp_u, the refined
// reciprocal) and the quotient from it (rdiv_y), so that several quotients by
// one divisor can share the first part; rdiv_u is their composition.
__device__ __forceinline__ double rcp_u(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    return fma(y, e, y);
}

__device__ __forceinline__ double rdiv_y(double a, double b, double y) {
    const double q = a * y;
    return fma(fma(b, q, -a), -y, q);
}

__device__ __forceinline__ double rdiv_u(double a, double b) { return rdiv_y(a, b, rcp_u(b)); }

// sqrt(x) for x in [2^-767, 2^1000]: v_rsq_f64 and the Goldschmidt
// iteration with two Newton corrections that the compiler emits for sqrt on
// gfx950, without the scaling of arguments below 2^-767 and the fix-up for 0
// and inf around it.  Inside that range the two run the same instructions, so
// this returns sqrt's bits.  Device only.
__device__ __forceinline__ double sqrt_u(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    double d = fma(-g, g, x);
    h = fma(h, r, h);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// a / b in the flavour D: the IEEE division, or rdiv_u.
template <int D>
__device__ __forceinline__ double dr(double a, double b) {
    if constexpr (D == kDivIEEE) return a / b;
    else return rdiv_u(a, b);
}
#endif

}  // namespace lsx
