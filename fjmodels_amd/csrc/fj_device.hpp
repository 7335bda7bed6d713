// fj_device.hpp -- device building blocks of the fJmodels SCF path on gfx950.
//
// Potential evaluation (pot/Potential.cpp:206-276, include/UtilsLeg.h:43-158)
// is fused: one radial bracket search and one pass over the Legendre
// coefficients give Phi, dPhi/dR and dPhi/dz together, with the derivative
// recurrences run alongside instead of re-running legend() per order.
// Root finding (include/Utils.h:134-167) and the adaptive quadratures
// (include/Utils.h:173-185, pot/Potential.cpp:161-206) are written as
// steppers so that a kernel can drive many of them in lock step.
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FJ_HD __host__ __device__
#else
// Host build of the same code (test harness only): plain C++.
#include <math.h>
#define __device__
#define __host__
#define __forceinline__ inline
#define __constant__
#define FJ_HD
#endif
#include <stdint.h>

#include "fj_math.h"

namespace fj {

// include/Utils.h:22-33
constexpr double EPSABSROOTS = 1e-8;
constexpr double EPSRELROOTS = 1e-6;
constexpr double EPSABSINTEG = 1e-5;
constexpr double EPSRELINTEG = 1e-3;
constexpr double PI = 3.141592653589793;
constexpr double PIH = 1.570796326794897;
constexpr double TPI = 6.283185307179586;
constexpr double FPI = 12.56637061435917;
constexpr double TINY = 1e-12;
constexpr double DBL_EPS = 2.220446049250313e-16;
constexpr double DBL_MINV = 2.2250738585072014e-308;
constexpr double DBL_MAXV = 1.7976931348623157e+308;

// Machine constants of the QAGS stepper (QUADPACK's d1mach / r1mach): the
// double instance is dqagse, the float instance QUADPACK's single-precision qagse.
template <class T> struct QLim;
template <> struct QLim<double> {
    static constexpr double eps = DBL_EPS, uflow = DBL_MINV, oflow = DBL_MAXV;
};
template <> struct QLim<float> {
    static constexpr float eps = 1.1920928955078125e-07f, uflow = 1.1754943508222875e-38f,
                           oflow = 3.4028234663852886e+38f;
};
constexpr int MAXPOLY = 32;

// MAX/MIN of include/Utils.h:47-48 (a NaN first argument compares false).
FJ_HD __forceinline__ double fj_max(double a, double b) { return a > b ? a : b; }
FJ_HD __forceinline__ double fj_min(double a, double b) { return a < b ? a : b; }

// Integer and quotient constants of the Legendre recurrences, filled by the
// host with the same integer-to-double conversions and divisions as the
// reference expressions (UtilsLeg.h:43-79), so every step rounds as there:
//   evenlegend step np:  p = -pm2*el[np]*elm1[np]/ed[np] + pm1*(c2 - eb[np]);  p *= eg[np]
//   legend step i:       P_i = (l2m1[i]*c*P_{i-1} - lm1[i]*P_{i-2}) / li[i]
// ied and ili are the reciprocals 1/ed and 1/li, for fj_div_rn.
struct LegCoef {
    double el[MAXPOLY], elm1[MAXPOLY], ed[MAXPOLY], eb[MAXPOLY], eg[MAXPOLY], ied[MAXPOLY];
    double l2m1[2 * MAXPOLY], lm1[2 * MAXPOLY], li[2 * MAXPOLY], ili[2 * MAXPOLY];
};

// a / b rounded to nearest, given y = RN(1/b), with three fused
// multiply-adds instead of a division.  q0 = RN(a y) is within 1.5 ulp of
// a/b; one correction q1 = RN(q0 + (a - b q0) y) is within one ulp (its
// error is 2^-52 of q0's plus the final rounding), so the remainder
// a - b q1 is exact and Markstein's theorem makes q2 = RN(q1 + (a - b q1) y)
// the correctly rounded quotient.  For a, b, a/b in the normal range (and
// a == 0, where q0 carries the IEEE sign); tests/test_device_code_cpu.py
// checks it against IEEE division.
FJ_HD __forceinline__ double fj_div_rn(double a, double b, double y) {
    const double q0 = a * y;
    const double q1 = fma(fma(-q0, b, a), y, q0);
    const double q2 = fma(fma(-q1, b, a), y, q1);
    return a == 0.0 ? q0 : q2;
}

FJ_HD __forceinline__ double evenleg_step(const LegCoef *lc, int np, double pm2, double pm1, double c2) {
    double p = fj_div_rn(-pm2 * lc->el[np] * lc->elm1[np], lc->ed[np], lc->ied[np]) + pm1 * (c2 - lc->eb[np]);
    p *= lc->eg[np];
    return p;
}

// Potential tables as the node kernels see them (LDS or global).
// Radial bracket table: cell k of NLUT equal cells in log2(r) over
// [ar[0], ar[nr-1]] holds the bracket index of the cell's left edge; a lookup
// plus a short exact walk gives the same bracket as the bisection of
// topbottom (Utils.h:103-128) for strictly increasing radii.
constexpr int NLUT = 256;

struct PotView {
    int nr, npoly;
    const double *ar;    // [nr]
    const double *phil;  // [nr][npoly]
    const double *Pr;    // [nr][npoly]
    const LegCoef *lc;
    const int *lut;      // [NLUT]
    float lg0, lsc;      // log2(ar[0]), NLUT / (log2(ar[nr-1]) - log2(ar[0]))
    const double *rk;    // [nr][RK]: divisors of radial_interp per cell (fill_radk_cell)
};

// Per radial cell bot (bracket [ar[bot], ar[bot+1]]): the divisor of f1 and
// its reciprocal, then for the quadratic correction with thr = bot-1 and with
// thr = bot+2: the divisor of f2, its reciprocal, and f3.  Each is formed
// with the operations of the quotient it serves in radial_interp, so the
// interpolation weights keep their bits.
constexpr int RK = 8;
FJ_HD __forceinline__ void fill_radk_cell(int nr, const double *ar, int bot, double *k) {
    const int top = bot + 1;
    for (int i = 0; i < RK; i++) k[i] = 0.0;
    if (top >= nr) return;
    const double d = ar[top] - ar[bot];
    k[0] = d;
    k[1] = 1.0 / d;
    if (bot > 0) {
        const int thr = bot - 1;
        const double den = (ar[thr] - ar[top]) * (ar[thr] - ar[bot]);
        k[2] = den; k[3] = 1.0 / den; k[4] = (ar[thr] - ar[bot]) / (ar[top] - ar[bot]);
    }
    if (top + 1 < nr) {
        const int thr = top + 1;
        const double den = (ar[thr] - ar[top]) * (ar[thr] - ar[bot]);
        k[5] = den; k[6] = 1.0 / den; k[7] = (ar[thr] - ar[bot]) / (ar[top] - ar[bot]);
    }
}

FJ_HD __forceinline__ float fj_log2f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __log2f(x);
#else
    return log2f(x);
#endif
}

// Host fill of the bracket table (any float log2 is fine: the walk in
// radial_interp corrects a cell off by one).
inline void make_lut(int nr, const double *ar, int *lut, float *lg0, float *lsc) {
    const double a = log2((double)(float)ar[0]), b = log2((double)(float)ar[nr - 1]);
    *lg0 = (float)a;
    *lsc = b > a ? (float)(NLUT / (b - a)) : 0.f;
    int m = 0;
    for (int k = 0; k < NLUT; k++) {
        const double x = b > a ? exp2(a + k * (b - a) / NLUT) : ar[0];
        while (m < nr - 2 && ar[m + 1] <= x) m++;
        lut[k] = m;
    }
}

// ---------------------------------------------------------------------------
// Radial bracket + interpolation weights (UtilsLeg.h:96-158, Utils.h:130-146).
// r is already clamped to >= ar[0] where the caller requires it.
// ---------------------------------------------------------------------------
struct Interp {
    int bot, top, thr;
    double f1, f2, f3;
    bool quad, kep;
    double q;   // ar[nr-1]/r on the Keplerian branch
};

FJ_HD __forceinline__ Interp radial_interp(const PotView &T, double r) {
    Interp I;
    const int nr = T.nr;
    const double *ar = T.ar;
    I.kep = (r >= ar[nr - 1]);
    I.quad = false;
    I.bot = 0; I.top = 0; I.thr = 0; I.f1 = 0; I.f2 = 0; I.f3 = 0; I.q = 0;
    if (I.kep) {
        I.q = ar[nr - 1] / r;
        return I;
    }
    // (xm[top]-x)*(x-xm[m]) >= 0 is x >= xm[m] for x inside [xm[0], xm[top]]:
    // the bracket is the largest bot <= nr-2 with r >= ar[bot] (0 if none)
    const float t = (fj_log2f((float)r) - T.lg0) * T.lsc;
    int bot = T.lut[t > 0.f ? (t < (float)(NLUT - 1) ? (int)t : NLUT - 1) : 0];
    while (bot < nr - 2 && r >= ar[bot + 1]) bot++;
    while (bot > 0 && !(r >= ar[bot])) bot--;
    const int top = bot + 1;
    I.bot = bot; I.top = top;
    const double db = r - ar[bot];
    const double *k = T.rk + RK * bot;
    I.f1 = fj_div_rn(db, k[0], k[1]);                      // db / (ar[top] - ar[bot])
    if (top < 10) {
        I.quad = true;
        const bool below = I.f1 < 0.5 && bot > 0;
        I.thr = below ? bot - 1 : top + 1;
        const double dt = r - ar[top];
        const double *kq = k + (below ? 2 : 5);
        I.f2 = fj_div_rn(dt * db, kq[0], kq[1]);           // dt db / ((ar[thr]-ar[top]) (ar[thr]-ar[bot]))
        I.f3 = kq[2];                                      // (ar[thr]-ar[bot]) / (ar[top]-ar[bot])
    }
    return I;
}

// Interpolated coefficient np of table tab (interior branch).
FJ_HD __forceinline__ double interp_coef(const Interp &I, const double *tab, int npoly, int np) {
    const double t = tab[I.top * npoly + np], b = tab[I.bot * npoly + np];
    double v = I.f1 * t + (1 - I.f1) * b;
    if (I.quad) {
        const double h = tab[I.thr * npoly + np];
        v += I.f2 * (h - b - I.f3 * (t - b));
    }
    return v;
}

// intpo2's radius: unchanged beyond the grid, else MAX(r, ar[0]).
FJ_HD __forceinline__ double I_clamp(const PotView &T, double r) {
    return r >= T.ar[T.nr - 1] ? r : fj_max(r, T.ar[0]);
}

// Interpolated coefficient np on either branch (intpo2, UtilsLeg.h:96-124):
// beyond r_max phil[nr-1][np] (ar[nr-1]/r)^(2np+1).
FJ_HD __forceinline__ double coef_at(const PotView &T, const Interp &I, const double *tab, int np) {
    if (I.kep) return tab[(T.nr - 1) * T.npoly + np] * fjm_powi(I.q, 2 * np + 1);
    return interp_coef(I, tab, T.npoly, np);
}

// Phi(R,z): Potential::operator() (Potential.cpp:206-219) with evenlegend.
FJ_HD __forceinline__ double pot_phi(const PotView &T, double R, double z) {
    const double c = z / sqrt(R * R + z * z), r = sqrt(R * R + z * z);
    const Interp I = radial_interp(T, I_clamp(T, r));
    const int np_ = T.npoly;
    const double c2 = c * c;
    double phi = coef_at(T, I, T.phil, 0);
    double pm2 = 1.0, pm1 = 1.5 * c2 - .5;
    for (int np = 1; np < np_; np++) {
        double p;
        if (np == 1) p = pm1;
        else { p = evenleg_step(T.lc, np, pm2, pm1, c2); pm2 = pm1; pm1 = p; }
        phi += coef_at(T, I, T.phil, np) * p;
    }
    return phi;
}

// Phi, dPhi/dR, dPhi/dz at one point: Potential::operator(), dR, dz, dtheta,
// dr (Potential.cpp:206-244, 261-276) with dlegend (UtilsLeg.h:51-57) and
// the three-table intpo2 (UtilsLeg.h:127-158).  One radial search serves
// all three; every value rounds as in the separate reference routines.
// With grad false only Phi is formed (the same bits as pot_phi).
FJ_HD __forceinline__ double pot_eval(const PotView &T, double R, double z, bool grad, double *dR, double *dz) {
    const double cphi = z / sqrt(R * R + z * z);
    double r = sqrt(z * z + R * R);
    if (r < T.ar[0]) r = T.ar[0];
    const Interp I = radial_interp(T, r);
    const int np_ = T.npoly;
    if (grad) {
        const double s = R / r;
        double c = z / r;
        if (c == 1) c -= 1e-8;
        // dtheta: sum_j phip[j] dlegend(c, 2j), times -sqrt(1-c^2)
        double dth = 0, Pm2 = 0, Pm1 = 1.0;   // P_{i-2}, P_{i-1}
        for (int j = 0; j < np_; j++) {
            double dpl = 0;
            if (j > 0) {
                const int n = 2 * j;
                double P1 = Pm1, P0 = Pm2;
                for (int i = n - 1; i <= n; i++) {
                    double Pi;
                    if (i == 1) Pi = c;
                    else Pi = fj_div_rn(T.lc->l2m1[i] * c * P1 - T.lc->lm1[i] * P0, T.lc->li[i], T.lc->ili[i]);
                    P0 = P1; P1 = Pi;
                }
                Pm2 = P0; Pm1 = P1;   // P_{n-1}, P_n
                dpl = (T.lc->li[n] * Pm2 - T.lc->li[n] * c * Pm1) / (1 - c * c);
            }
            dth += coef_at(T, I, T.phil, j) * dpl;
        }
        dth *= -sqrt(1 - c * c);
        // dr: sum dphip[np] P_2np(cos(acos(c)))
        const double ct = fjm_cos(fjm_acos(c)), ct2 = ct * ct;
        double g1 = 0, pm2 = 1.0, pm1 = 1.5 * ct2 - .5;
        for (int np = 0; np < np_; np++) {
            double p;
            if (np == 0) p = 1.0;
            else if (np == 1) p = pm1;
            else { p = evenleg_step(T.lc, np, pm2, pm1, ct2); pm2 = pm1; pm1 = p; }
            double dph;
            if (I.kep) dph = -(2 * np + 1) * T.phil[(T.nr - 1) * np_ + np] * fjm_powi(I.q, 2 * np + 1) / r;
            else dph = interp_coef(I, T.Pr, np_, np);
            g1 += dph * p;
        }
        *dR = dth * (c / r) + g1 * s;
        *dz = dth * (-s / r) + g1 * c;
    }
    // Phi at (R, z): c from the unclamped radius
    const double cc2 = cphi * cphi;
    double phi = coef_at(T, I, T.phil, 0), a2 = 1.0, a1 = 1.5 * cc2 - .5;
    for (int np = 1; np < np_; np++) {
        double p;
        if (np == 1) p = a1;
        else { p = evenleg_step(T.lc, np, a2, a1, cc2); a2 = a1; a1 = p; }
        phi += coef_at(T, I, T.phil, np) * p;
    }
    return phi;
}

FJ_HD __forceinline__ double pot_phi_grad(const PotView &T, double R, double z, double *dR, double *dz) {
    return pot_eval(T, R, z, true, dR, dz);
}

// ---------------------------------------------------------------------------
// Brent root finder in GSL iterate form (Brent 1973) driven as
// wrp_GSLroots (Utils.h:134-167), with F(low), F(high) supplied.
// ---------------------------------------------------------------------------
struct Brent {
    double a, b, c, d, e, fa, fb, fc;
    double lo, hi, root;
    int cont;    // 1 while the wrapper loop continues
    int iters;
};

// Returns true if the wrapper answers without iterating (|F| < 5e-4 at an end,
// or a bracket already narrower than the tolerance, which returns 0 as there).
FJ_HD __forceinline__ bool brent_start(Brent &s, double low, double high, double flow, double fhigh) {
    s.iters = 0;
    if (fabs(flow) < 5e-4) { s.root = low; s.cont = 0; return true; }
    if (fabs(fhigh) < 5e-4) { s.root = high; s.cont = 0; return true; }
    s.a = low; s.fa = flow; s.b = high; s.fb = fhigh; s.c = high; s.fc = fhigh;
    s.d = high - low; s.e = high - low;
    s.lo = low; s.hi = high; s.root = 0;
    s.cont = (high - low > EPSABSROOTS) ? 1 : 0;
    return !s.cont;
}

// Computes the next abscissa to evaluate; returns false if this iterate
// terminates without a new evaluation (fb == 0 or |m| <= tol).
FJ_HD __forceinline__ bool brent_propose(Brent &s, double &xnew) {
    double a = s.a, b = s.b, c = s.c, fa = s.fa, fb = s.fb, fc = s.fc, d = s.d, e = s.e;
    int ac_equal = 0;
    if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) { ac_equal = 1; c = a; fc = fa; d = b - a; e = b - a; }
    if (fabs(fc) < fabs(fb)) { ac_equal = 1; a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    const double tol = 0.5 * DBL_EPS * fabs(b);
    const double m = 0.5 * (c - b);
    if (fb == 0) {
        s.a = a; s.b = b; s.c = c; s.fa = fa; s.fb = fb; s.fc = fc; s.d = d; s.e = e;
        s.root = b; s.lo = b; s.hi = b;
        return false;
    }
    if (fabs(m) <= tol) {
        s.a = a; s.b = b; s.c = c; s.fa = fa; s.fb = fb; s.fc = fc; s.d = d; s.e = e;
        s.root = b;
        if (b < c) { s.lo = b; s.hi = c; } else { s.lo = c; s.hi = b; }
        return false;
    }
    if (fabs(e) < tol || fabs(fa) <= fabs(fb)) { d = m; e = m; }
    else {
        double p, q, r, ss = fb / fa;
        if (ac_equal) { p = 2 * m * ss; q = 1 - ss; }
        else {
            q = fa / fc; r = fb / fc;
            p = ss * (2 * m * q * (q - r) - (b - a) * (r - 1));
            q = (q - 1) * (r - 1) * (ss - 1);
        }
        if (p > 0) q = -q; else p = -p;
        if (2 * p < fmin(3 * m * q - fabs(tol * q), fabs(e * q))) { e = d; d = p / q; }
        else { d = m; e = m; }
    }
    a = b; fa = fb;
    if (fabs(d) > tol) b += d; else b += (m > 0 ? +tol : -tol);
    s.a = a; s.b = b; s.c = c; s.fa = fa; s.d = d; s.e = e; s.fc = fc;
    xnew = b;
    return true;
}

FJ_HD __forceinline__ void brent_accept(Brent &s, double fb) {
    s.fb = fb;
    s.root = s.b;
    double c = s.c;
    if ((fb < 0 && s.fc < 0) || (fb > 0 && s.fc > 0)) c = s.a;
    if (s.b < c) { s.lo = s.b; s.hi = c; } else { s.lo = c; s.hi = s.b; }
}

// gsl_root_test_interval + the wrapper's width test: decides whether the
// wrapper loop runs again.
FJ_HD __forceinline__ void brent_after(Brent &s) {
    s.iters++;
    double lo = s.lo, hi = s.hi, min_abs;
    if ((lo > 0.0 && hi > 0.0) || (lo < 0.0 && hi < 0.0)) min_abs = fmin(fabs(lo), fabs(hi));
    else min_abs = 0;
    const double tol = EPSABSROOTS + EPSRELROOTS * min_abs;
    const int cont = fabs(hi - lo) < tol ? 0 : 1;
    s.cont = (cont || hi - lo > EPSABSROOTS) ? 1 : 0;
    if (s.iters > 1000) s.cont = 0;
}

// Full root find for a functor F (one lane).
template <class F>
FJ_HD __forceinline__ double brent_solve(F &f, double low, double high, double flow, double fhigh, int *nit = nullptr) {
    Brent s;
    if (brent_start(s, low, high, flow, fhigh)) { if (nit) *nit = 0; return s.root; }
    while (s.cont) {
        double x;
        if (brent_propose(s, x)) brent_accept(s, f(x));
        brent_after(s);
    }
    if (nit) *nit = s.iters;
    return s.root;
}

// ---------------------------------------------------------------------------
// QUADPACK Gauss-Kronrod rules (dqk21, dqk15)
// ---------------------------------------------------------------------------
__constant__ static const double kXGK21[11] = {
    0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
    0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
    0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
    0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
    0.294392862701460198131126603103866, 0.148874338981631210884826001129720, 0.0};
__constant__ static const double kWGK21[11] = {
    0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
    0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
    0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
    0.123491976262065851077208980178525, 0.134709217311473325928054001771707,
    0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
    0.149445554002916905664936468389821};
__constant__ static const double kWG10[5] = {
    0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
    0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
    0.295524224714752870173892994651338};
__constant__ static const double kXGK15[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.0};
__constant__ static const double kWGK15[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
__constant__ static const double kWG7[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

template <class T>
struct QKOutT { T result, abserr, resabs, resasc; };
using QKOut = QKOutT<double>;

template <class T>
FJ_HD __forceinline__ T qk_rescale(T err, T resabs, T resasc) {
    err = fabs(err);
    if (resasc != 0 && err != 0) {
        const T t = 200 * err / resasc;
        const T scale = t * sqrt(t);   // (200 err/resasc)^1.5
        err = scale < 1 ? resasc * scale : resasc;
    }
    if (resabs > QLim<T>::uflow / (50 * QLim<T>::eps)) {
        const T min_err = 50 * QLim<T>::eps * resabs;
        if (min_err > err) err = min_err;
    }
    return err;
}

// Storage for the 2(NK-1) off-centre function values of one rule.  FvLocal
// keeps them in the caller's frame; FvStrided points into memory shared by a
// block of FV_LANES lanes (LDS on the device), element k at p[k * FV_LANES].
// The stride is a compile-time constant so that every element's address is
// the lane's base plus an immediate offset: with a run-time stride the
// compiler precomputed the 20 addresses per lane and spilled them.
template <int N>
struct FvLocal {
    double v[N];
    FJ_HD double &operator[](int k) { return v[k]; }
};
constexpr int FV_LANES = 256;   // block size of every kernel that stores rule values in LDS
struct FvStrided {
    double *p;
    FJ_HD double &operator[](int k) const { return p[k * FV_LANES]; }
};

// Kronrod rule with NK abscissae >= 0 (11: QK21, 8: QK15).  The integrand is
// evaluated in one rolled loop (one inlined copy of f); the sums are taken in
// QUADPACK's order: the pairs at the Gauss abscissae xgk[1], xgk[3], ...,
// then the Kronrod-only pairs xgk[0], xgk[2], ...  fv[2m], fv[2m+1] hold the
// values left and right of the centre for pair m of that visiting order.
template <int NK, class F, class FV>
FJ_HD __forceinline__ QKOut qk_rule(F &f, double a, double b, FV &&fv) {
    const double *xgk = NK == 11 ? kXGK21 : kXGK15;
    const double *wgk = NK == 11 ? kWGK21 : kWGK15;
    const double *wg = NK == 11 ? kWG10 : kWG7;
    constexpr int NG = (NK - 1) / 2;     // Gauss pairs
    const double center = 0.5 * (a + b), half = 0.5 * (b - a), ahalf = fabs(half);
    const double fc = f(center);
    double rg = 0, rk = fc * wgk[NK - 1], rabs = fabs(rk);
    if (NK % 2 == 0) rg = fc * wg[NK / 2 - 1];
    double f1 = 0;
#pragma unroll 1
    for (int s = 0; s < 2 * (NK - 1); s++) {
        const int m = s >> 1;
        const int jj = m < NG ? 2 * m + 1 : 2 * (m - NG);
        const double absc = half * xgk[jj];
        const double fx = f((s & 1) ? center + absc : center - absc);
        fv[s] = fx;
        if (s & 1) {
            const double fs = f1 + fx;
            if (m < NG) rg += wg[m] * fs;
            rk += wgk[jj] * fs;
            rabs += wgk[jj] * (fabs(f1) + fabs(fx));
        } else {
            f1 = fx;
        }
    }
    const double mean = rk * 0.5;
    double rasc = wgk[NK - 1] * fabs(fc - mean);
#pragma unroll
    for (int j = 0; j < NK - 1; j++) {
        const int m = (j & 1) ? (j >> 1) : NG + (j >> 1);
        rasc += wgk[j] * (fabs(fv[2 * m] - mean) + fabs(fv[2 * m + 1] - mean));
    }
    const double err = (rk - rg) * half;
    QKOut o;
    o.result = rk * half;
    o.resabs = rabs * ahalf;
    o.resasc = rasc * ahalf;
    o.abserr = qk_rescale(err, o.resabs, o.resasc);
    return o;
}

// FP32 pieces may contract to fused multiply-adds (the FP64 code may not).
#pragma clang fp contract(fast)
// The 2(NK-1) off-centre values of one FP32 rule, element k at p[k * FV_LANES] (LDS).
struct FvStrided32 {
    float *p;
    FJ_HD float &operator[](int k) const { return p[k * FV_LANES]; }
};

// qk_rule in FP32 (QUADPACK's single-precision qk21 / qk15): the same
// visiting order and sums, with the abscissae and weights rounded to float.
template <int NK, class F, class FV>
FJ_HD __forceinline__ QKOutT<float> qk_rule32(const F &f, float a, float b, FV &&fv) {
    const double *xgk = NK == 11 ? kXGK21 : kXGK15;
    const double *wgk = NK == 11 ? kWGK21 : kWGK15;
    const double *wg = NK == 11 ? kWG10 : kWG7;
    constexpr int NG = (NK - 1) / 2;
    const float center = 0.5f * (a + b), half = 0.5f * (b - a), ahalf = fabs(half);
    const float fc = f(center);
    float rg = 0.f, rk = fc * (float)wgk[NK - 1], rabs = fabs(rk);
    if (NK % 2 == 0) rg = fc * (float)wg[NK / 2 - 1];
    float f1 = 0.f;
#pragma unroll 1
    for (int s = 0; s < 2 * (NK - 1); s++) {
        const int m = s >> 1;
        const int jj = m < NG ? 2 * m + 1 : 2 * (m - NG);
        const float absc = half * (float)xgk[jj];
        const float fx = f((s & 1) ? center + absc : center - absc);
        fv[s] = fx;
        if (s & 1) {
            const float fs = f1 + fx, wk = (float)wgk[jj];
            if (m < NG) rg += (float)wg[m] * fs;
            rk += wk * fs;
            rabs += wk * (fabs(f1) + fabs(fx));
        } else {
            f1 = fx;
        }
    }
    const float mean = rk * 0.5f;
    float rasc = (float)wgk[NK - 1] * fabs(fc - mean);
#pragma unroll
    for (int j = 0; j < NK - 1; j++) {
        const int m = (j & 1) ? (j >> 1) : NG + (j >> 1);
        rasc += (float)wgk[j] * (fabs(fv[2 * m] - mean) + fabs(fv[2 * m + 1] - mean));
    }
    const float err = (rk - rg) * half;
    QKOutT<float> o;
    o.result = rk * half;
    o.resabs = rabs * ahalf;
    o.resasc = rasc * ahalf;
    o.abserr = qk_rescale(err, o.resabs, o.resasc);
    return o;
}

#pragma clang fp contract(off)

// ---------------------------------------------------------------------------
// QUADPACK dqagse as a stepper: the caller evaluates one Kronrod rule on
// [next_a, next_b] and hands the result to qags_consume().  Lists are
// 1-based arrays of length cap+1 in a per-lane workspace.
// ---------------------------------------------------------------------------
// Workspace layout.  FJ_WS_CHUNK_BYTES 0: lists interleaved element by
// element across lanes (element i of lane L at i * ld + L), so a wave
// touching the same index reads one contiguous run.  FJ_WS_CHUNK_BYTES B (a
// power of two): interleaved by B-byte chunks (element i of lane L at
// (i / CH) * CH * ld + L * CH + i % CH, CH = B / sizeof(T)), so a lane
// walking consecutive entries (the epsilon table in qelg, the error order in
// qpsrt) stays within one chunk for CH steps.
#ifndef FJ_WS_CHUNK_BYTES
#define FJ_WS_CHUNK_BYTES 64
#endif
template <class T>
FJ_HD constexpr int ws_chunk() { return FJ_WS_CHUNK_BYTES ? FJ_WS_CHUNK_BYTES / (int)sizeof(T) : 1; }
FJ_HD constexpr long long ws_round(long long n, int ch) { return (n + ch - 1) / ch * ch; }

// One lane's view of a list: p is the lane's first element (chunk).
template <class T>
struct Lanes {
    T *p;
    long long ld;
    FJ_HD T &operator[](int i) const {
        constexpr unsigned CH = (unsigned)ws_chunk<T>();
        const unsigned u = (unsigned)i;
        return p[(long long)(u / CH) * CH * ld + (u % CH)];
    }
};

// A lane's QAGS workspace inside a buffer shared by nlanes lanes (layout at
// lane_work()).  Addresses are formed at each access from the uniform base
// and stride and the lane index, so they do not occupy registers across the
// integrand evaluations.
template <class T>
struct QagsWorkT {
    T *ws;           // buffer base (uniform)
    long long ld;    // lanes sharing the buffer (uniform)
    int n1;          // cap + 1 (uniform)
    int lane;
    static constexpr int CH = ws_chunk<T>();
    FJ_HD long long l1() const { return ws_round(n1, CH); }   // per-lane length of a list
    FJ_HD Lanes<T> at(long long off) const { return {ws + off * ld + (long long)lane * CH, ld}; }
    FJ_HD Lanes<T> list(int k) const { return at(k * l1()); }
    FJ_HD Lanes<T> alist() const { return list(0); }
    FJ_HD Lanes<T> blist() const { return list(1); }
    FJ_HD Lanes<T> rlist() const { return list(2); }
    FJ_HD Lanes<T> elist() const { return list(3); }
    FJ_HD Lanes<T> rlist2() const { return at(4 * l1()); }
    FJ_HD Lanes<T> res3la() const { return at(4 * l1() + ws_round(56, CH)); }
    FJ_HD Lanes<int> iord() const {
        return {reinterpret_cast<int *>(ws + (4 * l1() + ws_round(56, CH) + ws_round(4, CH)) * ld) +
                    (long long)lane * ws_chunk<int>(),
                ld};
    }
};
using QagsWork = QagsWorkT<double>;

// Interval limit the packed counters hold (QUADPACK's limit; the reference's
// WRKSP is 1000).
constexpr int QAGS_MAX_LIMIT = 2047;

template <class T>
struct QagsT {
    // inputs
    T a0, b0, epsabs, epsrel;
    int limit;
    T next_a, next_b;
    // dqagse locals
    T area, area1, error1, defab1, defabs, dres, erlarg, erlast, errbnd, errmax,
        errsum, ertest, small, correc, result, abserr;
    // dqagse's integers, packed into three words so that a persistent
    // stepper keeps them in three registers: indices below 2048 (limit <=
    // QAGS_MAX_LIMIT), the roundoff counters below 32 (dqagse stops at 10 /
    // 20), nres and ktmin count steps, numrl2 stays below 52 (limexp 50).
    unsigned last : 11, maxerr : 11, state : 2, ier : 3, ierro : 2, extrap : 1, noext : 1, overflow : 1;
    unsigned nrmax : 11, nres : 11, numrl2 : 6;
    unsigned iroff1 : 5, iroff2 : 5, iroff3 : 5, ktmin : 11;
    signed ksgn : 2;
    // state: 0 first rule, 1 left half pending, 2 right half pending, 3 done;
    // overflow: workspace capacity reached before limit
};
using Qags = QagsT<double>;

// T is taken from q only (the bounds and tolerances convert to it).
template <class T> struct QNoDeduce { using type = T; };
template <class T>
FJ_HD __forceinline__ void qags_begin(QagsT<T> &q, typename QNoDeduce<T>::type a, typename QNoDeduce<T>::type b,
                                      typename QNoDeduce<T>::type epsabs, typename QNoDeduce<T>::type epsrel,
                                      int limit);
template <class T>
FJ_HD __forceinline__ void qags_consume(QagsT<T> &q, const QagsWorkT<T> &w, int cap, const QKOutT<T> &r);

template <int NK, class F>
FJ_HD __forceinline__ double qags_solve(F &f, double a, double b, double epsabs, double epsrel, int limit,
                                             const QagsWork &w, int cap, int *ier = nullptr, int *nqk = nullptr) {
    Qags q;
    qags_begin(q, a, b, epsabs, epsrel, limit);
    int n = 0;
    FvLocal<2 * (NK - 1)> fv;
    while (q.state != 3) {
        QKOut r = qk_rule<NK>(f, q.next_a, q.next_b, fv);
        n++;
        qags_consume(q, w, cap, r);
    }
    if (ier) *ier = q.overflow ? -1 : q.ier;
    if (nqk) *nqk = n;
    return q.result;
}

}  // namespace fj

namespace fj {

// QUADPACK dqpsrt (1-based lists).
template <class T>
FJ_HD __forceinline__ void qpsrt(int limit, int last, int &maxerr, T &ermax, const Lanes<T> &elist,
                                      const Lanes<int> &iord, int &nrmax) {
    int i, j, k, ido, ibeg, jbnd, isucc, jupbn;
    T errmax, errmin;
    if (last <= 2) { iord[1] = 1; iord[2] = 2; goto L90; }
    errmax = elist[maxerr];
    if (nrmax != 1) {
        ido = nrmax - 1;
        for (i = 1; i <= ido; i++) {
            isucc = iord[nrmax - 1];
            if (errmax <= elist[isucc]) break;
            iord[nrmax] = isucc;
            nrmax--;
        }
    }
    jupbn = last;
    if (last > (limit / 2 + 2)) jupbn = limit + 3 - last;
    errmin = elist[last];
    jbnd = jupbn - 1;
    ibeg = nrmax + 1;
    if (ibeg <= jbnd) {
        for (i = ibeg; i <= jbnd; i++) {
            isucc = iord[i];
            if (errmax >= elist[isucc]) goto L60;
            iord[i - 1] = isucc;
        }
    }
    iord[jbnd] = maxerr;
    iord[jupbn] = last;
    goto L90;
L60:
    iord[i - 1] = maxerr;
    k = jbnd;
    for (j = i; j <= jbnd; j++) {
        isucc = iord[k];
        if (errmin < elist[isucc]) { iord[k + 1] = last; goto L90; }
        iord[k + 1] = isucc;
        k--;
    }
    iord[i] = last;
L90:
    maxerr = iord[nrmax];
    ermax = elist[maxerr];
}

// QUADPACK dqelg (epsilon algorithm), epstab and res3la 1-based.
template <class T>
FJ_HD __forceinline__ void qelg(int &n, const Lanes<T> &epstab, T &result, T &abserr,
                                     const Lanes<T> &res3la, int &nres) {
    const T epmach = QLim<T>::eps, oflow = QLim<T>::oflow;
    nres++;
    abserr = oflow;
    result = epstab[n];
    if (n >= 3) {
        const int limexp = 50;
        epstab[n + 2] = epstab[n];
        const int newelm = (n - 1) / 2;
        epstab[n] = oflow;
        const int num = n;
        int k1 = n;
        for (int i = 1; i <= newelm; i++) {
            const int k2 = k1 - 1, k3 = k1 - 2;
            T res = epstab[k1 + 2];
            const T e0 = epstab[k3], e1 = epstab[k2], e2 = res;
            const T e1abs = fabs(e1), delta2 = e2 - e1, err2 = fabs(delta2);
            const T tol2 = fmax(fabs(e2), e1abs) * epmach;
            const T delta3 = e1 - e0, err3 = fabs(delta3);
            const T tol3 = fmax(e1abs, fabs(e0)) * epmach;
            if (!(err2 > tol2 || err3 > tol3)) {
                result = res;
                abserr = err2 + err3;
                abserr = fmax(abserr, T(0.5e+01) * epmach * fabs(result));
                return;
            }
            const T e3 = epstab[k1];
            epstab[k1] = e1;
            const T delta1 = e1 - e3, err1 = fabs(delta1);
            const T tol1 = fmax(e1abs, fabs(e3)) * epmach;
            bool shrink = (err1 <= tol1 || err2 <= tol2 || err3 <= tol3);
            T ss = 0;
            if (!shrink) {
                ss = T(1.0) / delta1 + T(1.0) / delta2 - T(1.0) / delta3;
                if (!(fabs(ss * e1) > T(0.1e-03))) shrink = true;
            }
            if (shrink) { n = i + i - 1; break; }
            res = e1 + T(1.0) / ss;
            epstab[k1] = res;
            k1 = k1 - 2;
            const T error = err2 + fabs(res - e2) + err3;
            if (!(error > abserr)) { abserr = error; result = res; }
        }
        if (n == limexp) n = 2 * (limexp / 2) - 1;
        int ib = ((num / 2) * 2 == num) ? 2 : 1;
        const int ie = newelm + 1;
        for (int i = 1; i <= ie; i++) { epstab[ib] = epstab[ib + 2]; ib += 2; }
        if (num != n) {
            int indx = num - n + 1;
            for (int i = 1; i <= n; i++) { epstab[i] = epstab[indx]; indx++; }
        }
        if (nres >= 4) {
            abserr = fabs(result - res3la[3]) + fabs(result - res3la[2]) + fabs(result - res3la[1]);
            res3la[1] = res3la[2]; res3la[2] = res3la[3]; res3la[3] = result;
        } else {
            res3la[nres] = result;
            abserr = oflow;
        }
    }
    abserr = fmax(abserr, T(0.5e+01) * epmach * fabs(result));
}

template <class T>
FJ_HD __forceinline__ void qags_begin(QagsT<T> &q, typename QNoDeduce<T>::type a, typename QNoDeduce<T>::type b,
                                      typename QNoDeduce<T>::type epsabs, typename QNoDeduce<T>::type epsrel,
                                      int limit) {
    q.a0 = a; q.b0 = b; q.epsabs = epsabs; q.epsrel = epsrel; q.limit = limit;
    q.state = 0; q.next_a = a; q.next_b = b;
    q.ier = 0; q.last = 0; q.result = 0; q.abserr = 0; q.overflow = 0;
    if (epsabs <= 0 && epsrel < fmax(50 * QLim<T>::eps, T(0.5e-28))) { q.ier = 6; q.state = 3; }
    if (limit > QAGS_MAX_LIMIT) { q.ier = 6; q.state = 3; }
}

// Sets up the bisection of interval maxerr for loop index q.last: the rule
// runs on the left half [a1, b1] first.  The interval stays in the workspace
// until both halves are done, so its ends are read back there (keeping them
// in the state instead measured 0.5 % slower: more spills).
template <class T>
FJ_HD __forceinline__ void qags_start_iter(QagsT<T> &q, const QagsWorkT<T> &w, int cap) {
    if (q.last > cap) { q.overflow = 1; q.state = 3; return; }
    const T a = w.alist()[q.maxerr], b = w.blist()[q.maxerr];
    q.erlast = q.errmax;
    q.next_a = a; q.next_b = T(0.5) * (a + b);
    q.state = 1;
}

// dqagse labels 100..140
template <class T>
FJ_HD __forceinline__ void qags_finish(QagsT<T> &q, const QagsWorkT<T> &w, int from115) {
    const T oflow = QLim<T>::oflow;
    int go115 = from115, go130 = 0;
    if (!go115) {
        if (q.abserr == oflow) go115 = 1;
        else if (q.ier + q.ierro != 0) {
            if (q.ierro == 3) q.abserr = q.abserr + q.correc;
            if (q.ier == 0) q.ier = 3;
            if (q.result != 0.0 && q.area != 0.0) {
                if (q.abserr / fabs(q.result) > q.errsum / fabs(q.area)) go115 = 1;
            } else if (q.abserr > q.errsum) go115 = 1;
            else if (q.area == 0.0) go130 = 1;
        }
        if (!go115 && !go130) {
            if (q.ksgn == -1 && fmax(fabs(q.result), fabs(q.area)) <= q.defabs * T(0.1e-01)) go130 = 1;
            else if (T(0.1e-01) > (q.result / q.area) || (q.result / q.area) > T(0.1e+03) || q.errsum > fabs(q.area))
                q.ier = 6;
        }
    }
    if (go115) {
        T s = 0;
        for (int k = 1; k <= q.last; k++) s = s + w.rlist()[k];
        q.result = s;
        q.abserr = q.errsum;
    }
    if (q.ier > 2) q.ier = q.ier - 1;
    q.state = 3;
}

template <class T>
FJ_HD __forceinline__ void qags_consume(QagsT<T> &q, const QagsWorkT<T> &w, int cap, const QKOutT<T> &r) {
    const T epmach = QLim<T>::eps, uflow = QLim<T>::uflow, oflow = QLim<T>::oflow;
    if (q.state == 0) {
        q.result = r.result; q.abserr = r.abserr; q.defabs = r.resabs;
        const T resabs = r.resasc;
        q.dres = fabs(q.result);
        q.errbnd = fmax(q.epsabs, q.epsrel * q.dres);
        q.last = 1;
        w.alist()[1] = q.a0; w.blist()[1] = q.b0;
        w.rlist()[1] = q.result; w.elist()[1] = q.abserr; w.iord()[1] = 1;
        if (q.abserr <= T(1.0e+02) * epmach * q.defabs && q.abserr > q.errbnd) q.ier = 2;
        if (q.limit == 1) q.ier = 1;
        if (q.ier != 0 || (q.abserr <= q.errbnd && q.abserr != resabs) || q.abserr == 0.0) { q.state = 3; return; }
        w.rlist2()[1] = q.result;
        q.errmax = q.abserr; q.maxerr = 1; q.area = q.result; q.errsum = q.abserr; q.abserr = oflow;
        q.nrmax = 1; q.nres = 0; q.numrl2 = 2; q.ktmin = 0; q.extrap = 0; q.noext = 0;
        q.ierro = 0; q.iroff1 = 0; q.iroff2 = 0; q.iroff3 = 0;
        q.erlarg = 0; q.ertest = 0; q.small = 0; q.correc = 0;
        q.ksgn = -1;
        if (q.dres >= (T(1.0) - T(0.5e+02) * epmach) * q.defabs) q.ksgn = 1;
        q.last = 2;
        qags_start_iter(q, w, cap);
        return;
    }
    if (q.state == 1) {
        q.area1 = r.result; q.error1 = r.abserr; q.defab1 = r.resasc;
        q.next_a = q.next_b; q.next_b = w.blist()[q.maxerr];   // right half [a2, b2], a2 = b1
        q.state = 2;
        return;
    }
    // state 2: both halves known -- body of the dqagse loop for index last
    const int last = q.last, limit = q.limit;
    const T area2 = r.result, error2 = r.abserr, defab2 = r.resasc;
    const T area1 = q.area1, error1 = q.error1;
    const T a1 = w.alist()[q.maxerr], b2 = w.blist()[q.maxerr], b1 = T(0.5) * (a1 + b2), a2 = b1;
    const T area12 = area1 + area2;
    const T erro12 = error1 + error2;
    q.errsum = q.errsum + erro12 - q.errmax;
    q.area = q.area + area12 - w.rlist()[q.maxerr];
    if (!(q.defab1 == error1 || defab2 == error2)) {
        if (!(fabs(w.rlist()[q.maxerr] - area12) > T(0.1e-04) * fabs(area12) || erro12 < T(0.99) * q.errmax)) {
            if (q.extrap) q.iroff2++;
            else q.iroff1++;
        }
        if (last > 10 && erro12 > q.errmax) q.iroff3++;
    }
    w.rlist()[q.maxerr] = area1;
    w.rlist()[last] = area2;
    q.errbnd = fmax(q.epsabs, q.epsrel * fabs(q.area));
    if (q.iroff1 + q.iroff2 >= 10 || q.iroff3 >= 20) q.ier = 2;
    if (q.iroff2 >= 5) q.ierro = 3;
    if (last == limit) q.ier = 1;
    if (fmax(fabs(a1), fabs(b2)) <= (T(1.0) + T(0.1e+03) * epmach) * (fabs(a2) + T(0.1e+04) * uflow)) q.ier = 4;
    if (error2 > error1) {
        w.alist()[q.maxerr] = a2;
        w.alist()[last] = a1;
        w.blist()[last] = b1;
        w.rlist()[q.maxerr] = area2;
        w.rlist()[last] = area1;
        w.elist()[q.maxerr] = error2;
        w.elist()[last] = error1;
    } else {
        w.alist()[last] = a2;
        w.blist()[q.maxerr] = b1;
        w.blist()[last] = b2;
        w.elist()[q.maxerr] = error1;
        w.elist()[last] = error2;
    }
    {
        int maxerr = q.maxerr, nrmax = q.nrmax;
        qpsrt(limit, last, maxerr, q.errmax, w.elist(), w.iord(), nrmax);
        q.maxerr = maxerr; q.nrmax = nrmax;
    }
    if (q.errsum <= q.errbnd) { qags_finish(q, w, 1); return; }
    if (q.ier != 0) { qags_finish(q, w, 0); return; }
    bool next = false;
    if (last == 2) {
        q.small = fabs(q.b0 - q.a0) * T(0.375);
        q.erlarg = q.errsum;
        q.ertest = q.errbnd;
        w.rlist2()[2] = q.area;
        next = true;
    } else if (q.noext) {
        next = true;
    } else {
        q.erlarg = q.erlarg - q.erlast;
        if (fabs(b1 - a1) > q.small) q.erlarg = q.erlarg + erro12;
        if (!q.extrap) {
            if (fabs(w.blist()[q.maxerr] - w.alist()[q.maxerr]) > q.small) next = true;
            else { q.extrap = 1; q.nrmax = 2; }
        }
        if (!next) {
            if (!(q.ierro == 3 || q.erlarg <= q.ertest)) {
                const int id = q.nrmax;
                int jupbnd = last;
                if (last > (2 + limit / 2)) jupbnd = limit + 3 - last;
                for (int k = id; k <= jupbnd; k++) {
                    q.maxerr = w.iord()[q.nrmax];
                    q.errmax = w.elist()[q.maxerr];
                    if (fabs(w.blist()[q.maxerr] - w.alist()[q.maxerr]) > q.small) { next = true; break; }
                    q.nrmax++;
                }
            }
        }
        if (!next) {
            q.numrl2++;
            w.rlist2()[q.numrl2] = q.area;
            T reseps, abseps;
            {
                int numrl2 = q.numrl2, nres = q.nres;
                qelg(numrl2, w.rlist2(), reseps, abseps, w.res3la(), nres);
                q.numrl2 = numrl2; q.nres = nres;
            }
            q.ktmin++;
            if (q.ktmin > 5 && q.abserr < T(0.1e-02) * q.errsum) q.ier = 5;
            if (abseps < q.abserr) {
                q.ktmin = 0;
                q.abserr = abseps;
                q.result = reseps;
                q.correc = q.erlarg;
                q.ertest = fmax(q.epsabs, q.epsrel * fabs(reseps));
                if (q.abserr <= q.ertest) { qags_finish(q, w, 0); return; }
            }
            if (q.numrl2 == 1) q.noext = 1;
            if (q.ier == 5) { qags_finish(q, w, 0); return; }
            q.maxerr = w.iord()[1];
            q.errmax = w.elist()[q.maxerr];
            q.nrmax = 1;
            q.extrap = 0;
            q.small = q.small * T(0.5);
            q.erlarg = q.errsum;
            next = true;
        }
    }
    q.last = last + 1;
    if (q.last > limit) { q.last = limit; qags_finish(q, w, 0); return; }
    qags_start_iter(q, w, cap);
}

}  // namespace fj

namespace fj {

// QAGS workspace of nlanes lanes, interleaved by lane (element by element
// or by chunks, FJ_WS_CHUNK_BYTES): per lane the doubles alist, blist,
// rlist, elist [cap+1], epsilon table [56], res3la [4], then the ints
// iord [cap+1], each rounded up to whole chunks.  The float instance of the
// same cap takes no more bytes than the double one.
FJ_HD constexpr long long ws_lane_doubles(int cap) {
    constexpr int CH = ws_chunk<double>();
    return 4 * ws_round(cap + 1, CH) + ws_round(56, CH) + ws_round(4, CH);
}
FJ_HD constexpr size_t ws_doubles(int cap, long long nlanes) {
    return (size_t)(ws_lane_doubles(cap) * nlanes + (ws_round(cap + 1, ws_chunk<int>()) * nlanes + 1) / 2);
}

template <class T>
FJ_HD inline QagsWorkT<T> lane_work(T *ws, int cap, long long lane, long long nlanes) {
    QagsWorkT<T> w;
    w.ws = ws; w.ld = nlanes; w.n1 = cap + 1; w.lane = (int)lane;
    return w;
}

// Host-side fill of the recurrence coefficients.
inline LegCoef make_legcoef(int npoly) {
    LegCoef lc{};
    for (int np = 2; np < npoly && np < MAXPOLY; np++) {
        const int l = 2 * (np - 1), l2 = 2 * l;
        lc.el[np] = (double)l;
        lc.elm1[np] = (double)(l - 1);
        lc.ed[np] = (double)((l2 + 1) * (l2 - 1));
        lc.eb[np] = (l2 * l + l2 - 1) / (double)((l2 - 1) * (l2 + 3));
        lc.eg[np] = (l2 + 1) * (l2 + 3) / (double)((l + 1) * (l + 2));
        lc.ied[np] = 1.0 / lc.ed[np];
    }
    for (int i = 0; i < 2 * MAXPOLY; i++) {
        lc.l2m1[i] = (double)(2 * i - 1);
        lc.lm1[i] = (double)(i - 1);
        lc.li[i] = (double)i;
        lc.ili[i] = i > 0 ? 1.0 / (double)i : 0.0;
    }
    return lc;
}

}  // namespace fj
