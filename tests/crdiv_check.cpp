// Host check of the correctly rounded quotient sequences of ls_crdiv.hpp
// against the IEEE division, over random operands, operands whose quotient
// lies within a few 2^-106 ulp of a rounding midpoint, exact quotients,
// signed zeros, range edges and the WENO-5 weight regime.
//
//   crdiv_check N [weak]   -> prints "div6 <mismatches> qdiv <mismatches> cases <n>"
//   crdiv_check N range    -> prints "range <violations> quotients <n>": WENO-5 faces,
//       derivatives and effective velocities (both schemes) on random stencils whose
//       values pass crdiv_safe, counting quotients outside the _u forms' domains
//
// "weak" drops the corrections (x/6 as x RN(1/6), a/b as a RN(1/b)), which
// are not correctly rounded; the test uses it to show that the cases catch it.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ls_crdiv.hpp"

static uint64_t s_state = 0x9e3779b97f4a7c15ull;
static uint64_t rnd() {  // splitmix64
    uint64_t z = (s_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static double bits(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static uint64_t ubits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

// Random double with exponent in [lo, hi] and a random (or patterned) mantissa.
static double rdouble(int lo, int hi) {
    const uint64_t r = rnd();
    uint64_t man = rnd() & ((1ull << 52) - 1);
    switch (r & 15) {
        case 0: man = 0; break;                              // power of two
        case 1: man = (1ull << 52) - 1; break;               // all ones
        case 2: man &= 0xfffull; break;                      // few low bits
        case 3: man |= (1ull << 52) - (1ull << 40); break;   // high ones
        default: break;
    }
    const int e = lo + (int)((r >> 8) % (uint64_t)(hi - lo + 1));
    const uint64_t sign = (r >> 40) & 1;
    return bits((sign << 63) | ((uint64_t)(e + 1023) << 52) | man);
}

static double next_up(double x) { return nextafter(x, INFINITY); }

static double qdiv_weak(double a, double b, double y) {  // qdiv without its corrections
    const double m = fabs(a);
    if (!(((m >= 0x1p-600 && m <= 0x1p600) || m == 0.0) && b >= 0x1p-300 && b <= 0x1p300)) return a / b;
    return a * y;
}
static double div6_weak(double x) {  // div6 without its correction
    const double m = fabs(x);
    if (!((m >= 0x1p-960 && m <= 0x1p1000) || m == 0.0)) return x / 6.0;
    return x * (1.0 / 6.0);
}

// Domain checks of the unchecked forms.
static long s_viol = 0, s_quot = 0;
static double d6r(double x) {
    const double m = fabs(x);
    s_quot++;
    if (!(m == 0.0 || (m >= 0x1p-960 && m <= 0x1p1000))) s_viol++;
    return x / 6.0;
}
static double dqr(double a, double b) {
    const double m = fabs(a);
    s_quot++;
    if (!(b >= 0x1p-300 && b <= 0x1p300 && (m == 0.0 || (m >= 0x1p-960 && m <= 0x1p960 && m / b >= 0x1p-1000 &&
                                                           m / b <= 0x1p1000))))
        s_viol++;
    return a / b;
}
// A stencil value that passes crdiv_safe: 0, or a magnitude in [2^-200, 2^50].
static double safe_value() {
    const uint64_t r = rnd();
    if ((r & 7) == 0) return 0.0;
    int e;
    switch ((r >> 3) & 3) {
        case 0: e = -200; break;
        case 1: e = 49; break;
        default: e = -200 + (int)((r >> 8) % 250); break;
    }
    return ldexp((r >> 40) & 1 ? -1.0 : 1.0, e) * (1.0 + (rnd() >> 12) * 0x1p-52);
}
static double beta3r(double p, double q, double r) {  // the reference solver's indicator
    double d = r - 2.0 * q + p;
    return (13.0 / 12.0) * d * d + 0.25 * (r - p) * (r - p);
}
static double sbetar(double f0, double f1, double f2) {  // the sample scheme's indicator
    const double p = f0 - 2.0 * f1 + f2, q = f0 - 4.0 * f1 + 3.0 * f2;
    return (13.0 / 12.0) * (p * p) + (1.0 / 4.0) * (q * q);
}
static double weno_lr(const double* v) {  // weno5_left with its quotients checked
    const double eps = 1.0e-6;
    const double b0 = beta3r(v[0], v[1], v[2]), b1 = beta3r(v[1], v[2], v[3]), b2 = beta3r(v[2], v[3], v[4]);
    const double a0 = 0.1 / ((eps + b0) * (eps + b0)), a1 = 0.6 / ((eps + b1) * (eps + b1)),
                 a2 = 0.3 / ((eps + b2) * (eps + b2));  // divisions in every flavour
    const double s = a0 + a1 + a2;
    const double w0 = dqr(a0, s), w1 = dqr(a1, s), w2 = dqr(a2, s);
    return w0 * d6r(2.0 * v[0] - 7.0 * v[1] + 11.0 * v[2]) + w1 * d6r(-v[1] + 5.0 * v[2] + 2.0 * v[3]) +
           w2 * d6r(2.0 * v[2] + 5.0 * v[3] - v[4]);
}
static double sface_r(const double* v) {  // the sample scheme's face with its quotients checked
    const double eps = 1.0e-6;
    const double b0 = sbetar(v[0], v[1], v[2]), b1 = sbetar(v[1], v[2], v[3]), b2 = sbetar(v[2], v[3], v[4]);
    const double a0 = 0.1 / ((eps + b0) * (eps + b0)), a1 = 0.6 / ((eps + b1) * (eps + b1)),
                 a2 = 0.3 / ((eps + b2) * (eps + b2));
    const double sum = a0 + a1 + a2;
    const double w0 = dqr(a0, sum), w1 = dqr(a1, sum), w2 = dqr(a2, sum);
    return d6r(w0 * (2.0 * v[0] - 7.0 * v[1] + 11.0 * v[2])) + d6r(w1 * (-v[1] + 5.0 * v[2] + 2.0 * v[3])) +
           d6r(w2 * (2.0 * v[2] + 5.0 * v[3] - v[4]));
}
static void range_case() {
    double v[7];
    for (double& x : v) x = safe_value();
    const double h = ldexp(1.0 + (rnd() >> 12) * 0x1p-52, (int)(rnd() % 121) - 61);  // in [2^-61, 2^60)
    const double hh = h < 0x1p-60 ? 0x1p-60 : h;
    const double sl = safe_value(), eps = ldexp(1.0, -200 + (int)(rnd() % 251));
    // faces c-1/2, c+1/2 (left-biased) and the derivative over h
    const double fl = weno_lr(v), fh = weno_lr(v + 1);
    const double d = dqr(fh - fl, hh);
    // central difference, |grad G| over three axes, S_L times it over |grad G|
    const double gx = dqr(v[4] - v[2], 2.0 * hh);
    const double gy = dqr(v[5] - v[1], 2.0 * hh);
    const double gz = dqr(v[6] - v[0], 2.0 * hh);
    const double gm = sqrt(gx * gx + gy * gy + gz * gz + eps);
    dqr(sl * gx, gm);
    dqr(sl * gz, gm);
    (void)d;
    // sample scheme: three faces, the one-sided derivatives and the LLF speeds
    const double s0 = sface_r(v), s1 = sface_r(v + 1), s2 = sface_r(v + 2);
    const double gm0 = dqr(s1 - s0, hh), gp0 = dqr(s2 - s1, hh);
    const double nm = sqrt(gm0 * gm0 + gp0 * gp0 + 1e-14);
    dqr(sl * gm0, nm);
    dqr(sl * gp0, nm);
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    if (argc > 2 && !strcmp(argv[2], "range")) {
        for (long it = 0; it < n; it++) range_case();
        printf("range %ld quotients %ld\n", s_viol, s_quot);
        return 0;
    }
    const bool weak = argc > 2 && !strcmp(argv[2], "weak");
    long bad6 = 0, badq = 0, cases = 0;
    auto check6 = [&](double x) {
        cases++;
        const double got = weak ? div6_weak(x) : lsx::div6(x);
        if (ubits(got) != ubits(x / 6.0)) {
            if (bad6 < 5 && !weak) fprintf(stderr, "div6 %a: %a vs %a\n", x, got, x / 6.0);
            bad6++;
        }
    };
    auto checkq = [&](double a, double b) {
        cases++;
        const double y = 1.0 / b;
        const double got = weak ? qdiv_weak(a, b, y) : lsx::qdiv(a, b, y);
        if (ubits(got) != ubits(a / b)) {
            if (badq < 5 && !weak) fprintf(stderr, "qdiv %a / %a: %a vs %a\n", a, b, got, a / b);
            badq++;
        }
    };
    // signed zeros, specials, range edges
    const double sp[] = {0.0, -0.0, 1.0, -1.0, 6.0, 3.0, INFINITY, -INFINITY, NAN, 0x1p-500, 0x1p500,
                         0x1.fffffffffffffp-501, 0x1.0000000000001p500, 0x1p-1074, 0x1p-1022, 0x1.fffffffffffffp1023};
    for (double x : sp) {
        check6(x);
        for (double b : sp) {
            if (isnan(x) || isnan(b)) continue;  // NaN payloads are not compared
            checkq(x, b);
        }
    }
    for (long it = 0; it < n; it++) {
        // div6: random, and numerators of exact or near-exact quotients
        check6(rdouble(-600, 600));
        check6(rdouble(-1074 + 52, -400));   // through the division fallback
        const double q = rdouble(-300, 300);
        check6(6.0 * q);
        check6(next_up(6.0 * q));
        // qdiv: random operands
        checkq(rdouble(-400, 400), rdouble(-400, 400));  // b < 0 takes the division
        checkq(rdouble(-400, 400), fabs(rdouble(-400, 400)));
        // near-midpoint quotients: a = RN(b m) for a midpoint m of two doubles
        const double b = fabs(rdouble(-200, 200));
        const double lo = rdouble(-200, 200);
        const double m = lo + (next_up(lo) - lo) * 0.5;  // RN may land on lo or its neighbour
        const double a = b * m;
        checkq(a, b);
        checkq(next_up(a), b);
        checkq(nextafter(a, -INFINITY), b);
        // exact quotients
        checkq(b * (double)(int64_t)(rnd() >> 40), b);
        // WENO-5 weight regime: alpha_k = d_k/t_k, s = a0 + a1 + a2, w_k = alpha_k/s
        const double t0 = ldexp(1.0 + (rnd() >> 11) * 0x1p-53, -40 + (int)(rnd() % 80));
        const double t1 = ldexp(1.0 + (rnd() >> 11) * 0x1p-53, -40 + (int)(rnd() % 80));
        const double t2 = ldexp(1.0 + (rnd() >> 11) * 0x1p-53, -40 + (int)(rnd() % 80));
        const double a0 = 0.1 / t0, a1 = 0.6 / t1, a2 = 0.3 / t2;
        const double s = a0 + a1 + a2;
        checkq(a0, s);
        checkq(a1, s);
        checkq(a2, s);
        // qdiv_u over its whole stated domain: b in [2^-300, 2^300], |a| in
        // [2^-960, 2^960], |a|/b in [2^-1000, 2^1000] (no fallback)
        {
            const double bu = fabs(rdouble(-300, 299));
            const int eb = ilogb(bu);
            const int lo = eb - 999 > -960 ? eb - 999 : -960, hi = eb + 998 < 959 ? eb + 998 : 959;
            const double au = rdouble(lo, hi);
            cases++;
            const double got = weak ? au * (1.0 / bu) : lsx::qdiv_u(au, bu, 1.0 / bu);
            if (ubits(got) != ubits(au / bu)) {
                if (badq < 5 && !weak) fprintf(stderr, "qdiv_u %a / %a: %a vs %a\n", au, bu, got, au / bu);
                badq++;
            }
        }
        // div6_u at the ends of its domain
        {
            const double xu = rnd() & 1 ? rdouble(-960, -900) : rdouble(940, 999);
            cases++;
            const double got = weak ? xu * (1.0 / 6.0) : lsx::div6_u(xu);
            if (ubits(got) != ubits(xu / 6.0)) {
                if (bad6 < 5 && !weak) fprintf(stderr, "div6_u %a: %a vs %a\n", xu, got, xu / 6.0);
                bad6++;
            }
        }
    }
    printf("div6 %ld qdiv %ld cases %ld\n", bad6, badq, cases);
    return 0;
}
