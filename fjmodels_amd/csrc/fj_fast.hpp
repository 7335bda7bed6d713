// fj_fast.hpp -- per-node work of the fast node phase (FJ_PREC_FAST).
//
// The same Staeckel-fudge actions as uvOrb (uvOrb/uvOrb.cpp:15-262), computed
// another way, for speed:
//   * the reference point of the u side is the node itself instead of umid
//     (Getu0, uvOrb.cpp:109-129): umid only sets the zero of Er and dU, and
//     cancels exactly from uturnfn (uvOrb.cpp:25-31, 61-73), so the bracket
//     walk from u = 58 and its Brent solve are not needed;
//   * the turning points are bracketed by the reference's own walks
//     (GetTurnu from u = 1, GetTurnv from pi/4: uvOrb.cpp:131-224, which
//     decide which allowed interval of the u and v momenta is the orbit's)
//     and, after wrp_GSLroots' endpoint test (|F| < 5e-4), solved by
//     Anderson-Bjorck regula falsi in FP64 to 1e-9 relative;
//   * the action integrals Juint / Jvint (uvOrb.cpp:226-262) are fixed-order
//     Gauss-Legendre rules over psi (the cos-mapping of the reference makes
//     the integrands vanish smoothly at both turning points).
// With energies measured from E, the two momentum functions are
//   F_u(u) = B_u - C / sh^2 u - (sh^2 u + sin^2 v0) (Phi(u, v0) - E)
//   F_v(v) = B_v - C / sin^2 v - (sh^2 u0 + sin^2 v) (Phi(u0, v) - E)
// with C = Lz^2 / (2 Delta^2), K = (sh^2 u0 + sin^2 v0) (Phi(R, z) - E),
// B_u = pu^2 / (2 Delta^2) + C / sh^2 u0 + K, B_v = pv^2 / (2 Delta^2) + C / sin^2 v0 + K,
// so that F_u(u0) = pu^2 / (2 Delta^2) and F_v(v0) = pv^2 / (2 Delta^2).
// Jr = sqrt(2) Delta Du / pi int_0^pi sqrt(F_u(ubar - Du cos psi)) sin psi dpsi and
// Jz = 2 sqrt(2) Delta (pi/2 - vt) / pi int_0^pi/2 sqrt(F_v(pi/2 - (pi/2 - vt) cos psi)) sin psi dpsi.
// Host/device code: the FP64 parts also compile into the host test build.
#pragma once
#include "fj_node.hpp"

namespace fj {

struct FastNode {
    double Delta, E, C, Bu, Bv;
    double shu0, chu0, sv0, cv0;   // node's spheroidal coordinates (v0 folded into [0, pi/2])
    double u0, v0;
    double uin, uout, vt;          // turning points
    int ju, jv;                    // 1: the integral is needed (else the action is 0)
};

// Anderson-Bjorck regula falsi on a bracket [a, b] with fa * fb <= 0; stops
// when the bracket is below rtol relative.  *nev counts evaluations.
template <class F>
FJ_HD inline double fast_root(const F &f, double a, double b, double fa, double fb, double rtol, int *nev) {
    if (fa == 0) return a;
    if (fb == 0) return b;
    for (int it = 0; it < 100; it++) {
        const double x = b - fb * (b - a) / (fb - fa);
        const double fx = f(x);
        (*nev)++;
        if (fx == 0) return x;
        if ((fx > 0) != (fb > 0)) { a = b; fa = fb; }
        else {
            const double m = 1 - fx / fb;
            fa *= m > 0 ? m : 0.5;
        }
        b = x; fb = fx;
        if (fabs(b - a) <= rtol * fabs(b)) break;
    }
    return b;
}

struct FastFu {
    using Tab = PotView;
    PotView T;
    const FastNode *n;
    FJ_HD double operator()(double u) const {
        double shu, chu;
        fjm_sinhcosh(u, &shu, &chu);
        const double sh2 = shu * shu, D = n->Delta;
        const double Phi = pot_phi(T, D * shu * n->sv0, D * chu * n->cv0);
        return n->Bu - n->C / sh2 - (sh2 + n->sv0 * n->sv0) * (Phi - n->E);
    }
};

// dF_u/du (duturnfn, uvOrb.cpp:74-79, with F_u's reference point)
struct FastDFu {
    PotView T;
    const FastNode *n;
    FJ_HD double operator()(double u) const {
        double shu, chu;
        fjm_sinhcosh(u, &shu, &chu);
        const double sh2 = shu * shu, D = n->Delta;
        double dR, dz;
        const double Phi = pot_phi_grad(T, D * shu * n->sv0, D * chu * n->cv0, &dR, &dz);
        const double dPhidu = D * (dR * chu * n->sv0 + dz * shu * n->cv0);
        return 2 * (n->E - Phi) * shu * chu - (sh2 + n->sv0 * n->sv0) * dPhidu + 2 * n->C * chu / (sh2 * shu);
    }
};

struct FastFv {
    PotView T;
    const FastNode *n;
    FJ_HD double operator()(double v) const {
        double s, cs;
        fjm_sincos(v, &s, &cs);
        const double s2 = s * s, D = n->Delta;
        const double Phi = pot_phi(T, D * n->shu0 * s, D * n->chu0 * cs);
        return n->Bv - n->C / s2 - (n->shu0 * n->shu0 + s2) * (Phi - n->E);
    }
};

constexpr double FAST_ROOT_RTOL = 1e-9;

// wrp_GSLroots (include/Utils.h:134-167): an end of the bracket whose |F| is
// below 5e-4 is taken as the root (low first); otherwise the root is solved
// for, on f, to rtol relative.  The early return decides the reference's
// turning points of near-circular and near-planar orbits, so it is kept;
// flo and fhi are FP64 values.
template <class F>
FJ_HD inline double fast_root_ref(const F &f, double lo, double hi, double flo, double fhi, int *nev,
                                  double rtol = FAST_ROOT_RTOL) {
    if (fabs(flo) < 5e-4) return lo;
    if (fabs(fhi) < 5e-4) return hi;
    return fast_root(f, lo, hi, flo, fhi, rtol, nev);
}

// The evaluations of fast_setup: all in FP64 (host test build).  The device
// kernel's policy (fj_mixed.hpp) walks and solves on FP32 evaluations and
// checks each bracket's ends in FP64.
struct FastPolicy64 {
    using Tab = PotView;
    using U = FastFu;
    using V = FastFv;
    using DU = FastDFu;
    static constexpr double rtol = FAST_ROOT_RTOL;
    FJ_HD static DU make_du(const PotView &T, const Tab &, const FastNode *n) { return DU{T, n}; }
};

// Brackets of the two u turning points found by GetTurnu's walk.
struct UBracket {
    double uinner, u1, u2, uouter;
    bool noroot;
};

// GetTurnu(1) (uvOrb.cpp:131-201): the reference's bracket walk from u = 1.
// The walk only takes signs of F_u (and of dF_u in its outer branches), so
// Uw and dU may be lower-precision evaluations.
template <class UF, class DUF>
FJ_HD inline UBracket fast_walk_u(const UF &Uw, const DUF &dU, int *nev) {
    UBracket b;
    double ut_ = 1, ut = Uw(ut_), dutt = 0, eps = .1;
    (*nev)++;
    if (ut < 0) { dutt = dU(ut_); (*nev)++; }   // read only on this side
    double uinner = ut_, uouter = ut_, u1 = ut_, u2 = ut_;
    bool noroot = false;
    int guard = 0;
    if (ut < 0) {
        if (dutt > 0) {
            uinner = ut_;
            while (guard++ < 4000) {
                while (ut < 0 && dutt > 0 && guard++ < 4000) {
                    uinner = ut_; ut_ *= (1 + eps);
                    ut = Uw(ut_); dutt = dU(ut_); *nev += 2;
                }
                if (ut >= 0) { u1 = ut_; break; }
                ut_ /= (1 + eps); eps *= .2;
                if (eps < TINY) { uinner = uouter = ut_; noroot = true; break; }
                ut_ *= (1 + eps);
                ut = Uw(ut_); dutt = dU(ut_); *nev += 2;
            }
            if (!noroot) {
                while (ut > 0 && guard++ < 8000) { u2 = ut_; ut_ *= 1.1; ut = Uw(ut_); (*nev)++; }
                uouter = ut_;
            }
        } else {
            while (guard++ < 4000) {
                while (ut < 0 && dutt < 0 && guard++ < 4000) {
                    uouter = ut_; ut_ /= (1 + eps);
                    ut = Uw(ut_); dutt = dU(ut_); *nev += 2;
                }
                if (ut >= 0) { u2 = ut_; break; }
                ut_ *= (1 + eps); eps *= .2;
                if (eps < TINY) { uinner = uouter = ut_; noroot = true; break; }
                ut_ /= (1 + eps);
                ut = Uw(ut_); dutt = dU(ut_); *nev += 2;
            }
            if (!noroot) {
                while (ut > 0 && guard++ < 8000) { u1 = ut_; ut_ *= .9; ut = Uw(ut_); (*nev)++; }
                uinner = ut_;
            }
        }
    } else {
        const double f_at_1 = ut;
        u1 = u2 = ut_;
        while (ut > 0 && guard++ < 4000) { u2 = ut_; ut_ *= 1.1; ut = Uw(ut_); (*nev)++; }
        uouter = ut_;
        ut = f_at_1;   // the inward walk starts where the outward one stopped (uvOrb.cpp:186-190)
        while (ut > 0 && guard++ < 8000) { u1 = ut_; ut_ *= .9; ut = Uw(ut_); (*nev)++; }
        uinner = ut_;
    }
    b.uinner = uinner; b.u1 = u1; b.u2 = u2; b.uouter = uouter; b.noroot = noroot;
    return b;
}

// Everything of one node before the action integrals: H < 0 test, Delta(E),
// spheroidal coordinates (Uvpt.cpp:11-24), E, the constants above and the
// turning points.  Returns 0 (H >= 0), 1 (bound) or 2 (E >= 0: unbound).
// *nev counts the potential evaluations.  The u and v walks and the root
// solves run on P::U and P::V (built from tw and the node); the ends of each
// bracket are checked in FP64, and the endpoint test takes the FP64 values.
template <class P>
FJ_HD inline int fast_setup(const PotView &T, const typename P::Tab &tw, const double *Eg, const double *Dg, int ng,
                            double R, double z, double Phi_h, const double *V, FastNode &n, int *nev) {
    n.ju = 0; n.jv = 0;
    const double H = .5 * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) + Phi_h;   // DF.cpp:83
    if (!(H < 0)) return 0;
    const double Delta = d_deltafn(Eg, Dg, ng, H);
    const double Delta2 = Delta * Delta, R2 = R * R, z2 = z * z, r2 = R2 + z2;
    const double shu2 = .5 * (r2 - Delta2 + sqrt((Delta2 - r2) * (Delta2 - r2) + 4 * Delta2 * R2)) / Delta2;
    const double shu = sqrt(shu2), chu = sqrt(1 + shu2);
    const double cv2 = .5 * (r2 + Delta2 - sqrt((Delta2 + r2) * (Delta2 + r2) - 4 * Delta2 * z2)) / Delta2;
    const double cva = sqrt(cv2), sv2 = 1 - cv2, sva = sqrt(sv2 > 0 ? sv2 : 0);
    const double cv = z < 0 ? -cva : cva;
    const double pR = V[0], pz = V[2];
    const double pu = Delta * (pR * chu * sva + pz * shu * cv);
    const double pv = Delta * (pR * shu * cv - pz * chu * sva);
    const double Lz = R * V[1];
    n.Delta = Delta;
    n.E = .5 * (pR * pR + pz * pz) + .5 * Lz * Lz / R2 + Phi_h;   // uvOrb.cpp:22
    if (n.E >= 0 || isnan(n.E)) return 2;
    n.C = .5 * Lz * Lz / Delta2;
    const double K = (shu2 + sv2) * (Phi_h - n.E);
    n.Bu = .5 * pu * pu / Delta2 + n.C / shu2 + K;
    n.Bv = .5 * pv * pv / Delta2 + n.C / sv2 + K;
    n.shu0 = shu; n.chu0 = chu; n.sv0 = sva; n.cv0 = cva;   // Phi is even in z
    n.u0 = fjm_asinh(shu);
    n.v0 = fjm_acos(cva);                                   // in [0, pi/2]
    const FastFu U{T, &n};
    const typename P::DU dU = P::make_du(T, tw, &n);
    const typename P::U Uw{tw, &n};
    const UBracket b = fast_walk_u(Uw, dU, nev);
    if (!b.noroot) {
        // FP64 values at the bracket ends (the roots' brackets and wrp_GSLroots'
        // endpoint test); where the low-precision walk took a sign that FP64
        // does not confirm, the end moves on by the walk's own step
        double ui = b.uinner, u1 = b.u1, u2 = b.u2, uo = b.uouter;
        double fi = U(ui), f1 = U(u1), f2 = u2 == u1 ? f1 : U(u2), fo = U(uo);
        *nev += 3 + (u2 != u1);
        int g = 0;
        while (fi > 0 && g++ < 200) { u1 = ui; f1 = fi; ui *= .9; fi = U(ui); (*nev)++; }
        while (!(f1 > 0) && u1 < u2 && g++ < 400) { ui = u1; fi = f1; u1 /= .9; f1 = U(u1); (*nev)++; }
        while (fo > 0 && g++ < 600) { u2 = uo; f2 = fo; uo *= 1.1; fo = U(uo); (*nev)++; }
        while (!(f2 > 0) && u2 > u1 && g++ < 800) { uo = u2; fo = f2; u2 /= 1.1; f2 = U(u2); (*nev)++; }
        n.uin = fast_root_ref(Uw, ui, u1, fi, f1, nev, P::rtol);
        n.uout = fast_root_ref(Uw, u2, uo, f2, fo, nev, P::rtol);
    } else {
        n.uin = b.uinner; n.uout = b.uouter;
    }
    n.ju = !(fabs(n.uout - n.uin) < TINY);
    // GetTurnv(pi/4) (uvOrb.cpp:203-224): halve towards pi/2 (F < 0 at pi/4)
    // or towards 0 (F >= 0) until F changes sign.  The first value, which
    // picks the direction, is FP64; the walk runs on Vw, then the two ends of
    // its last step are checked in FP64 and moved by the walk's own step
    // where the FP64 sign differs.
    const FastFv Vf{T, &n};
    const typename P::V Vw{tw, &n};
    n.vt = 0;
    {
        const double vtry = .5 * PIH;
        const double f0 = Vf(vtry);
        (*nev)++;
        const bool up = f0 < 0;                 // walk vbot towards pi/2
        // a: last point with f0's sign, b: first point past it
        double a = vtry, b = vtry, fb = f0;
        bool done = false;
        int guard = 0;
        while ((up ? fb < 0 : fb > 0) && guard++ < 4000) {
            a = b;
            b = up ? .5 * (PIH + b) : .5 * b;
            if (up ? PIH - b < TINY : b < TINY) { n.vt = up ? PIH : 0; done = true; break; }
            fb = Vw(b); (*nev)++;
        }
        if (!done) {
            double fa = a == vtry ? f0 : Vf(a), fb64 = Vf(b);
            *nev += 1 + (a != vtry);
            int g = 0;
            // FP64 says a is already past the sign change: step back
            while ((up ? !(fa < 0) : !(fa > 0)) && a != vtry && g++ < 200) {
                b = a; fb64 = fa;
                a = up ? 2 * a - PIH : 2 * a;
                if (up ? a < vtry : a > vtry) a = vtry;
                fa = a == vtry ? f0 : Vf(a); (*nev)++;
            }
            // FP64 says b has not reached it: step on
            while ((up ? fb64 < 0 : fb64 > 0) && g++ < 400) {
                a = b; fa = fb64;
                b = up ? .5 * (PIH + b) : .5 * b;
                if (up ? PIH - b < TINY : b < TINY) { n.vt = up ? PIH : 0; done = true; break; }
                fb64 = Vf(b); (*nev)++;
            }
            if (!done) {
                // wrp_GSLroots on the reference's bracket, [vtry, b] or [b, vtry]:
                // its endpoint test (low end first), then the root, solved
                // within the walk's last step [a, b]
                const double lo = up ? vtry : b, hi = up ? b : vtry;
                const double flo = up ? f0 : fb64, fhi = up ? fb64 : f0;
                if (fabs(flo) < 5e-4) n.vt = lo;
                else if (fabs(fhi) < 5e-4) n.vt = hi;
                else if (up) n.vt = fast_root(Vw, a, b, fa, fb64, P::rtol, nev);
                else n.vt = fast_root(Vw, b, a, fb64, fa, P::rtol, nev);
            }
        }
    }
    n.jv = !(fabs(n.vt - PIH) < TINY);
    return 1;
}

// Order of the Gauss-Legendre rule of the v integral for a u-rule of n
// points: n / 2 (at least 4).  The v integrand is far smoother (its median
// error at 16 points is below 1e-14 of the node's total action), so half
// the u rule's order keeps J_z more accurate than J_r.
FJ_HD constexpr int fast_v_order(int n) { return n / 2 < 4 ? 4 : n / 2; }

// The two action integrals with Gauss-Legendre rules on [0, 1]: ngl points
// (x[k], w[k]) for J_r, nglv points (xv[k], wv[k]) for J_z.  FP64 version
// (host test build and reference for the FP32 kernel).
FJ_HD inline void fast_actions64(const PotView &T, const FastNode &n, int ngl, const double *x, const double *w,
                                 int nglv, const double *xv, const double *wv, double *Jr, double *Jz) {
    const double rtwo = 1.4142135623730951;
    *Jr = 0; *Jz = 0;
    const FastFu fu{T, &n};
    const FastFv fv{T, &n};
    if (n.ju) {
        const double ubar = .5 * (n.uin + n.uout), du = .5 * (n.uout - n.uin);
        double s = 0;
        for (int k = 0; k < ngl; k++) {
            double sp, cp;
            fjm_sincos(PI * x[k], &sp, &cp);
            const double F = fu(ubar - du * cp);
            s += w[k] * sqrt(F > 0 ? F : 0) * sp;
        }
        *Jr = rtwo * n.Delta * du * s;
    }
    if (n.jv) {
        const double dv = PIH - n.vt;
        double s = 0;
        for (int k = 0; k < nglv; k++) {
            double sp, cp;
            fjm_sincos(PIH * xv[k], &sp, &cp);
            const double F = fv(PIH - dv * cp);
            s += wv[k] * sqrt(F > 0 ? F : 0) * sp;
        }
        *Jz = rtwo * n.Delta * dv * s;
    }
}

}  // namespace fj
