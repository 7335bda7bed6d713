/*
 * fj_math.h -- portable FP64 elementary functions used by every computation
 * of the SCF path (GPU kernels, and the CPU oracle that checks them).
 *
 * The fJmodels algorithm takes discrete decisions on computed values (root
 * bracketing on the sign of a derivative that vanishes near pericentre,
 * early root returns when |F| < 5e-4, adaptive-quadrature bisection), so two
 * implementations that differ by one ulp in sinh() or cos() take different
 * branches on a few nodes in 10^5 and move those nodes' actions by percents.
 * Both sides therefore use these functions, written only with IEEE +,-,*,/,
 * sqrt, fused multiply-add (C99 fma: one rounding, so the same bits on every
 * conforming implementation), round-to-nearest-even rint and exact scaling,
 * with no other contraction: the same inputs give the same bits on the host
 * and on gfx950.  Accuracy is a few ulp (checked against mpmath in
 * tests/test_fjmath.py).
 *
 * Methods: Cody-Waite argument reduction (ln 2 and pi/2 split into exactly
 * representable parts, as in fdlibm) followed by Taylor polynomials in
 * fused Horner form; expm1 by the same reduction, 2^k (expm1(r) + 1) - 1
 * formed with one fma, so it needs no branch on |x|; sinh and cosh from one
 * expm1 and one division, again without a branch; log via the atanh series
 * of (m-1)/(m+1); asin via its Taylor series on |x| <= 1/2 and the
 * half-angle identity beyond; atan by two half-angle reductions and its Taylor
 * series.
 */
#ifndef FJ_MATH_H
#define FJ_MATH_H

#if defined(__HIPCC__)
#define FJM_FN __host__ __device__ static inline
#else
#define FJM_FN static inline
#endif

#include <math.h>
#include <stdint.h>
#include <string.h>

/* Translation units that include this file are compiled without contraction
 * (-ffp-contract=off); the pragma keeps these functions so regardless. */
#if defined(__HIPCC__)
#pragma clang fp contract(off)
#endif

FJM_FN double fjm_bits2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
FJM_FN uint64_t fjm_d2bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

/* x * 2^k, k in [-2044, 2046] (two exact power-of-two factors) */
FJM_FN double fjm_ldexp(double x, int k) {
    int k1 = k / 2, k2 = k - k / 2;
    return (x * fjm_bits2d((uint64_t)(1023 + k1) << 52)) * fjm_bits2d((uint64_t)(1023 + k2) << 52);
}

#define FJM_FMA(a, b, c) fma((a), (b), (c))

FJM_FN double fjm_rint(double x) {   /* round to nearest, ties to even (IEEE roundToIntegralTiesToEven) */
    return rint(x);
}

/* 2^k for k in [-1022, 1023] */
FJM_FN double fjm_pow2i(int k) { return fjm_bits2d((uint64_t)(1023 + k) << 52); }

/* exp(r) - 1 for |r| <= ln2/2 (+ a hair): Taylor to r^13 in fused Horner form */
FJM_FN double fjm_expm1_red(double r) {
    double p = 1.0 / 6227020800.0;                   /* 1/13! */
    p = FJM_FMA(p, r, 1.0 / 479001600.0);
    p = FJM_FMA(p, r, 1.0 / 39916800.0);
    p = FJM_FMA(p, r, 1.0 / 3628800.0);
    p = FJM_FMA(p, r, 1.0 / 362880.0);
    p = FJM_FMA(p, r, 1.0 / 40320.0);
    p = FJM_FMA(p, r, 1.0 / 5040.0);
    p = FJM_FMA(p, r, 1.0 / 720.0);
    p = FJM_FMA(p, r, 1.0 / 120.0);
    p = FJM_FMA(p, r, 1.0 / 24.0);
    p = FJM_FMA(p, r, 1.0 / 6.0);
    p = FJM_FMA(p, r, 0.5);
    return FJM_FMA(r * r, p, r);
}

/* x = k ln2 + r, |r| <= ln2/2 */
FJM_FN double fjm_reduce_ln2(double x, double *k) {
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    *k = fjm_rint(x * inv_ln2);
    return FJM_FMA(-*k, ln2_lo, FJM_FMA(-*k, ln2_hi, x));
}

/* exp(x) */
FJM_FN double fjm_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return fjm_bits2d(0x7ff0000000000000ull);
    if (x < -745.1332191019412) return 0.0;
    double k;
    const double r = fjm_reduce_ln2(x, &k);
    const double e = 1.0 + fjm_expm1_red(r);
    const int ki = (int)k;
    if (ki < -1021 || ki > 1023) return fjm_ldexp(e, ki);
    return e * fjm_pow2i(ki);
}

/* exp(x) - 1: 2^k expm1(r) + (2^k - 1) with one rounding; 2^k - 1 is exact
 * for |k| <= 53.  Beyond x = 700 the result is exp(x) - 1 (+inf past 709.78);
 * below -40 it is -1. */
FJM_FN double fjm_expm1(double x) {
    if (x != x) return x;
    if (x > 700.0) return fjm_exp(x) - 1.0;
    if (x < -40.0) return -1.0;
    double k;
    const double r = fjm_reduce_ln2(x, &k);
    const double t = fjm_pow2i((int)k);
    return FJM_FMA(t, fjm_expm1_red(r), t - 1.0);
}

/* sinh and cosh of the same argument: with m = expm1(|x|), e = m + 1,
 * sinh|x| = (m + m/e)/2 (no cancellation for small |x|), cosh x = (e + 1/e)/2 */
FJM_FN void fjm_sinhcosh(double x, double *s, double *c) {
    const double ax = x < 0 ? -x : x;
    if (ax > 700.0) {   /* e^-|x| is negligible: both are e^|x| / 2 = (e^(|x|/2) / 2) e^(|x|/2) */
        const double e = fjm_exp(0.5 * ax);
        const double h = (0.5 * e) * e;
        *s = x < 0 ? -h : h;
        *c = h;
        return;
    }
    const double m = fjm_expm1(ax);
    const double e = m + 1.0;
    const double ie = 1.0 / e;
    const double sh = 0.5 * FJM_FMA(m, ie, m);
    *s = x < 0 ? -sh : sh;
    *c = 0.5 * (e + ie);
}
FJM_FN double fjm_sinh(double x) { double s, c; fjm_sinhcosh(x, &s, &c); return s; }
FJM_FN double fjm_cosh(double x) { double s, c; fjm_sinhcosh(x, &s, &c); return c; }

/* sin and cos (|x| < 1e5: three-part Cody-Waite reduction by pi/2, |r| <= pi/4;
 * Taylor to r^19 and r^18, whose next terms are below 2^-60 relative) */
FJM_FN void fjm_sincos(double x, double *sn, double *cs) {
    const double pio2_1 = 1.57079632673412561417e+00, pio2_2 = 6.07710050630396597660e-11,
                 pio2_3 = 2.02226624871116645580e-21, inv_pio2 = 6.36619772367581382433e-01;
    const double k = fjm_rint(x * inv_pio2);
    const double r = FJM_FMA(-k, pio2_3, FJM_FMA(-k, pio2_2, FJM_FMA(-k, pio2_1, x)));
    const double r2 = r * r;
    double p = -1.0 / 121645100408832000.0;     /* -1/19! */
    p = FJM_FMA(p, r2, 1.0 / 355687428096000.0);
    p = FJM_FMA(p, r2, -1.0 / 1307674368000.0);
    p = FJM_FMA(p, r2, 1.0 / 6227020800.0);
    p = FJM_FMA(p, r2, -1.0 / 39916800.0);
    p = FJM_FMA(p, r2, 1.0 / 362880.0);
    p = FJM_FMA(p, r2, -1.0 / 5040.0);
    p = FJM_FMA(p, r2, 1.0 / 120.0);
    p = FJM_FMA(p, r2, -1.0 / 6.0);
    const double s = FJM_FMA(r * r2, p, r);
    double q = -1.0 / 6402373705728000.0;       /* -1/18! */
    q = FJM_FMA(q, r2, 1.0 / 20922789888000.0);
    q = FJM_FMA(q, r2, -1.0 / 87178291200.0);
    q = FJM_FMA(q, r2, 1.0 / 479001600.0);
    q = FJM_FMA(q, r2, -1.0 / 3628800.0);
    q = FJM_FMA(q, r2, 1.0 / 40320.0);
    q = FJM_FMA(q, r2, -1.0 / 720.0);
    q = FJM_FMA(q, r2, 1.0 / 24.0);
    q = FJM_FMA(q, r2, -0.5);
    const double c = FJM_FMA(r2, q, 1.0);
    /* quadrant k mod 4: (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s) */
    const int quad = (int)k & 3;
    const double a = (quad & 1) ? c : s, b = (quad & 1) ? s : c;
    *sn = (quad & 2) ? -a : a;
    *cs = ((quad + 1) & 2) ? -b : b;
}
FJM_FN double fjm_sin(double x) { double s, c; fjm_sincos(x, &s, &c); return s; }
FJM_FN double fjm_cos(double x) { double s, c; fjm_sincos(x, &s, &c); return c; }

/* asin for |x| <= 0.5: Taylor series sum c_n x^(2n+1), c_n = (2n)!/(4^n (n!)^2 (2n+1)) */
#if defined(__HIP_DEVICE_COMPILE__)
__constant__
#endif
static const double fjm_asin_c[28] = {
    1.66666666666666666667e-1, 7.5e-2, 4.46428571428571428571e-2,
    3.03819444444444444444e-2, 2.23721590909090909091e-2, 1.73527644230769230769e-2,
    1.396484375e-2, 1.15518008961397058824e-2, 9.76160952919407894737e-3,
    8.39033580961681547619e-3, 7.3125258735988451087e-3, 6.4472103118896484375e-3,
    5.74003767084192346644e-3, 5.15330968231990419585e-3, 4.6601434869150961599e-3,
    4.24090709367936307734e-3, 3.88096455883766923632e-3, 3.56920539382593454541e-3,
    3.29705950347348474539e-3, 3.05782164925803066935e-3, 2.84617840110894216788e-3,
    2.65787063820728993354e-3, 2.48944867824688349464e-3, 2.33809189211197518693e-3,
    2.20147397371013820542e-3, 2.07766103251816744278e-3, 1.96503361627728361844e-3,
    1.86222640640312748928e-3};

FJM_FN double fjm_asin_small(double x) {
    const double x2 = x * x;
    double p = fjm_asin_c[27];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll 1
#endif
    for (int i = 26; i >= 0; i--) p = FJM_FMA(p, x2, fjm_asin_c[i]);
    return FJM_FMA(x * x2, p, x);
}

FJM_FN double fjm_acos(double x) {
    const double pio2_hi = 1.57079632679489655800e+00, pio2_lo = 6.12323399573676603587e-17;
    if (x != x || x > 1.0 || x < -1.0) return fjm_bits2d(0x7ff8000000000000ull);
    if (x <= 0.5 && x >= -0.5) return pio2_hi - (fjm_asin_small(x) - pio2_lo);
    if (x > 0.5) return 2.0 * fjm_asin_small(sqrt((1.0 - x) * 0.5));
    return (pio2_hi - 2.0 * fjm_asin_small(sqrt((1.0 + x) * 0.5))) + (pio2_hi + 2.0 * pio2_lo);
}

/* log(x), x > 0 */
FJM_FN double fjm_log(double x) {
    if (x != x || x < 0) return fjm_bits2d(0x7ff8000000000000ull);
    if (x == 0) return -fjm_bits2d(0x7ff0000000000000ull);
    if (x == fjm_bits2d(0x7ff0000000000000ull)) return x;
    int e = 0;
    if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; e = -54; }   /* 2^54 */
    uint64_t u = fjm_d2bits(x);
    e += (int)((u >> 52) & 0x7ff) - 1022;
    double m = fjm_bits2d((u & 0x000fffffffffffffull) | 0x3fe0000000000000ull);   /* [0.5, 1) */
    if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 27.0;
    p = FJM_FMA(p, s2, 1.0 / 25.0);
    p = FJM_FMA(p, s2, 1.0 / 23.0);
    p = FJM_FMA(p, s2, 1.0 / 21.0);
    p = FJM_FMA(p, s2, 1.0 / 19.0);
    p = FJM_FMA(p, s2, 1.0 / 17.0);
    p = FJM_FMA(p, s2, 1.0 / 15.0);
    p = FJM_FMA(p, s2, 1.0 / 13.0);
    p = FJM_FMA(p, s2, 1.0 / 11.0);
    p = FJM_FMA(p, s2, 1.0 / 9.0);
    p = FJM_FMA(p, s2, 1.0 / 7.0);
    p = FJM_FMA(p, s2, 1.0 / 5.0);
    p = FJM_FMA(p, s2, 1.0 / 3.0);
    const double lm = 2.0 * FJM_FMA(s * s2, p, s);
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double de = (double)e;
    return FJM_FMA(de, ln2_hi, FJM_FMA(de, ln2_lo, lm));
}

/* atan(x): |x| > 1 by pi/2 - atan(1/x); two halvings t -> t / (1 + sqrt(1 + t^2))
 * (atan t = 2 atan of the result) bring |t| below tan(pi/16) < 0.2, where the Taylor
 * series to t^25 is exact to far below an ulp. */
FJM_FN double fjm_atan(double x) {
    if (x != x) return x;
    const double ax = x < 0 ? -x : x;
    const int inv = ax > 1.0;
    double t = inv ? 1.0 / ax : ax;
    t = t / (1.0 + sqrt(1.0 + t * t));
    t = t / (1.0 + sqrt(1.0 + t * t));
    const double t2 = t * t;
    double p = 1.0 / 25.0;
    p = FJM_FMA(p, t2, -1.0 / 23.0);
    p = FJM_FMA(p, t2, 1.0 / 21.0);
    p = FJM_FMA(p, t2, -1.0 / 19.0);
    p = FJM_FMA(p, t2, 1.0 / 17.0);
    p = FJM_FMA(p, t2, -1.0 / 15.0);
    p = FJM_FMA(p, t2, 1.0 / 13.0);
    p = FJM_FMA(p, t2, -1.0 / 11.0);
    p = FJM_FMA(p, t2, 1.0 / 9.0);
    p = FJM_FMA(p, t2, -1.0 / 7.0);
    p = FJM_FMA(p, t2, 1.0 / 5.0);
    p = FJM_FMA(p, t2, -1.0 / 3.0);
    double r = 4.0 * FJM_FMA(t * t2, p, t);
    if (inv) r = (1.57079632679489655800e+00 - r) + 6.12323399573676603587e-17;
    return x < 0 ? -r : r;
}

/* atan2(y, x) for finite arguments; atan2(0, 0) = 0 */
FJM_FN double fjm_atan2(double y, double x) {
    const double pi_hi = 3.14159265358979311600e+00, pi_lo = 1.22464679914735317720e-16;
    if (x > 0) return fjm_atan(y / x);
    if (x < 0) {
        const double a = fjm_atan(y / x);
        return y >= 0 ? (a + pi_lo) + pi_hi : (a - pi_lo) - pi_hi;
    }
    if (y > 0) return 1.57079632679489655800e+00;
    if (y < 0) return -1.57079632679489655800e+00;
    return 0.0;
}

/* asinh(x) */
FJM_FN double fjm_asinh(double x) {
    const double ax = x < 0 ? -x : x;
    double r;
    if (ax < 0.5) {
        /* Taylor: x + sum_{n>=1} (-1)^n c_n x^(2n+1), c_n as in asin */
        const double x2 = ax * ax;
        const double c[28] = {
            1.66666666666666666667e-1, 7.5e-2, 4.46428571428571428571e-2,
            3.03819444444444444444e-2, 2.23721590909090909091e-2, 1.73527644230769230769e-2,
            1.396484375e-2, 1.15518008961397058824e-2, 9.76160952919407894737e-3,
            8.39033580961681547619e-3, 7.3125258735988451087e-3, 6.4472103118896484375e-3,
            5.74003767084192346644e-3, 5.15330968231990419585e-3, 4.6601434869150961599e-3,
            4.24090709367936307734e-3, 3.88096455883766923632e-3, 3.56920539382593454541e-3,
            3.29705950347348474539e-3, 3.05782164925803066935e-3, 2.84617840110894216788e-3,
            2.65787063820728993354e-3, 2.48944867824688349464e-3, 2.33809189211197518693e-3,
            2.20147397371013820542e-3, 2.07766103251816744278e-3, 1.96503361627728361844e-3,
            1.86222640640312748928e-3};
        double p = c[27];                        /* n = 28: + */
        for (int i = 26; i >= 0; i--) p = FJM_FMA(p, x2, (i & 1) ? c[i] : -c[i]);   /* sign (-1)^(i+1) */
        r = FJM_FMA(ax * x2, p, ax);
    } else if (ax > 1e8) {
        r = fjm_log(ax) + 6.93147180559945309417e-01;
    } else {
        r = fjm_log(ax + sqrt(ax * ax + 1.0));
    }
    return x < 0 ? -r : r;
}

/* tanh(x) */
FJM_FN double fjm_tanh(double x) {
    const double ax = x < 0 ? -x : x;
    double t;
    if (ax > 20.0) t = 1.0;
    else {
        const double e = fjm_expm1(2.0 * ax);
        t = e / (e + 2.0);
    }
    return x < 0 ? -t : t;
}

/* x^n for integer n (binary powering) */
FJM_FN double fjm_powi(double x, int n) {
    int m = n < 0 ? -n : n;
    double r = 1.0, b = x;
    while (m) {
        if (m & 1) r *= b;
        m >>= 1;
        if (m) b *= b;
    }
    return n < 0 ? 1.0 / r : r;
}

/* x^y, x > 0 (exp(y log x)); integer y in [-64, 64] uses binary powering */
FJM_FN double fjm_pow(double x, double y) {
    if (y == (double)(int)y && y >= -64 && y <= 64) return fjm_powi(x, (int)y);
    if (x == 0) return y > 0 ? 0.0 : fjm_bits2d(0x7ff0000000000000ull);
    return fjm_exp(y * fjm_log(x));
}

#endif
