// fj_node.hpp -- per-velocity-node work of df() up to the action integrals
// (df/DF.cpp:78-89, uvOrb/Uvpt.cpp:11-24, uvOrb/uvOrb.cpp:15-262), the
// orbit functions, and the action integrands.  Host/device code: the same
// source runs in the gfx950 kernels and in a host test build.
#pragma once
#include "fj_device.hpp"

namespace fj {

// include/Utils.h:149-155 linterp (x strictly inside uses the bisection)
FJ_HD inline double d_linterp(const double *xm, const double *ym, int np, double x) {
    if (x <= xm[0]) return ym[0];
    if (x >= xm[np - 1]) return ym[np - 1];
    int top = np - 1, bot = 0;
    while (top - bot > 1) {
        int m = (top + bot) >> 1;
        if (x >= xm[m]) bot = m; else top = m;
    }
    return ym[bot] + (x - xm[bot]) / (xm[top] - xm[bot]) * (ym[top] - ym[bot]);
}

// delta/tabulate.cpp:40-44
FJ_HD inline double d_deltafn(const double *Eg, const double *Dg, int n, double E) {
    if (E < Eg[0]) return sqrt(fj_max(.000001, Dg[0]));
    else if (E < Eg[n - 1]) return sqrt(fj_max(.000001, d_linterp(Eg, Dg, n, E)));
    return sqrt(fj_max(.000001, Dg[n - 1]));
}

// ---------------------------------------------------------------------------
// Orbit functions of uvOrb (uvOrb.cpp:42-83) with per-node constants
// ---------------------------------------------------------------------------
struct NodeConst {
    double Delta, E, Lzsq, sv, cv, sv2, sinv, cosv;  // u-side
    double shuX, chuX, shu2;                         // v-side: sinh/cosh(X.u), X.shu2
    double sh1sq, K1, K2, K3, invsh1sq;              // uturnfn: Er*ch1sq, (sh1sq+sv2)*Phiu1, .5Lzsq/D^2, 1/sh1sq
    double Iv, KV;                                   // vturnfn: Iv, (1+shu2)*Phiv(PIH)
};

struct DUturn {   // duturnfn (uvOrb.cpp:74-79)
    PotView T;        // by value: the table pointers stay visible as LDS addresses
    const NodeConst *c;
    long long *nev;
    FJ_HD __forceinline__ double operator()(double u) const {
        double shu, chu;
        fjm_sinhcosh(u, &shu, &chu);
        const double shu2 = shu * shu;
        const double D = c->Delta;
        double dR, dz;
        const double Phi0 = pot_phi_grad(T, D * shu * c->sinv, D * chu * c->cosv, &dR, &dz);
        (*nev)++;
        const double dPhidu = D * (dR * chu * c->sv + dz * shu * c->cv);
        return 2 * (c->E - Phi0) * shu * chu - (shu2 + c->sv2) * dPhidu + c->Lzsq * chu / (D * D * shu2 * shu);
    }
};

struct Uturn {    // uturnfn (uvOrb.cpp:70-73) with dU (61-63)
    PotView T;        // by value: the table pointers stay visible as LDS addresses
    const NodeConst *c;
    int *nphi;        // potential evaluations, or null
    FJ_HD __forceinline__ double operator()(double u) const {
        double shu, chu;
        fjm_sinhcosh(u, &shu, &chu);
        const double shu2 = shu * shu;
        const double D = c->Delta;
        if (nphi) (*nphi)++;
        const double Phi = pot_phi(T, D * shu * c->sinv, D * chu * c->cosv);
        const double dU = (shu2 + c->sv2) * Phi - c->K2;
        return c->E * (shu2 - c->sh1sq) + c->K1 - dU - c->K3 * (1 / shu2 - c->invsh1sq);
    }
};

struct Vturn {    // vturnfn (uvOrb.cpp:49-52) with dV (46-48)
    PotView T;        // by value: the table pointers stay visible as LDS addresses
    const NodeConst *c;
    int *nphi;        // potential evaluations, or null
    FJ_HD __forceinline__ double operator()(double v) const {
        double s, cs;
        fjm_sincos(v, &s, &cs);
        const double sv2 = s * s;
        const double D = c->Delta;
        if (nphi) (*nphi)++;
        const double Phi = pot_phi(T, D * c->shuX * s, D * c->chuX * cs);
        const double dV = c->KV - (c->shu2 + sv2) * Phi;
        return c->E * sv2 + c->Iv + dV - .5 * c->Lzsq / (D * D * sv2);
    }
};

// Action integrands Juint / Jvint (uvOrb.cpp:53-56, 80-83)
// The node constants one action integrand reads (uvOrb.cpp:61-67, 70-73,
// 84-90, 226-262), packed so that a lane holds only those of its own kind:
//   k[0] Delta, k[1] E, then
//   Ju (type 0): ubar, Deltau, sinv, cosv, sv2, sh1sq, K1, K2, K3, invsh1sq
//   Jv (type 1): vturn, shuX, chuX, shu2, Iv, KV, Lzsq
constexpr int JK = 12;

struct JInt {
    const PotView *T;
    double k[JK];
    int type;   // 0: Ju, 1: Jv
    FJ_HD double operator()(double psi) const {
        double sp, cp;
        fjm_sincos(psi, &sp, &cp);
        double F;
        const double D = k[0], E = k[1];
        if (type == 0) {
            // Juint (uvOrb.cpp:80-83, integrated by GetJu 226-245): u = ubar - Deltau cos(psi)
            const double u = k[2] - k[3] * cp;
            double shu, chu;
            fjm_sinhcosh(u, &shu, &chu);
            const double shu2 = shu * shu;
            const double Phi = pot_phi(*T, D * shu * k[4], D * chu * k[5]);
            const double dU = (shu2 + k[6]) * Phi - k[9];
            F = E * (shu2 - k[7]) + k[8] - dU - k[10] * (1 / shu2 - k[11]);
        } else {
            // Jvint (uvOrb.cpp:53-56, integrated by GetJv 247-262): v = pi/2 - (pi/2 - vturn) cos(psi)
            const double v = PIH - (PIH - k[2]) * cp;
            double s, cs;
            fjm_sincos(v, &s, &cs);
            const double sv2 = s * s;
            const double Phi = pot_phi(*T, D * k[3] * s, D * k[4] * cs);
            const double dV = k[7] - (k[5] + sv2) * Phi;
            F = E * sv2 + k[6] + dV - .5 * k[8] / (D * D * sv2);
        }
        return sqrt(fj_max(0., F)) * sp;
    }
    // Jr (uvOrb.cpp:236) or Jz (uvOrb.cpp:254) from the integral
    FJ_HD double action(double integral) const {
        const double rtwo = 1.4142135623730951;   // sqrt(2) (uvOrb.cpp:13)
        if (type == 0) return rtwo * k[0] * k[3] * integral / PI;
        return 2 * rtwo * k[0] * (PIH - k[2]) * integral / PI;
    }
};

FJ_HD inline JInt make_jint(const PotView *T, const NodeConst &c, double ubar, double Deltau, double vturn,
                            int type) {
    JInt f;
    f.T = T;
    f.type = type;
    f.k[0] = c.Delta; f.k[1] = c.E;
    if (type == 0) {
        f.k[2] = ubar; f.k[3] = Deltau; f.k[4] = c.sinv; f.k[5] = c.cosv; f.k[6] = c.sv2; f.k[7] = c.sh1sq;
        f.k[8] = c.K1; f.k[9] = c.K2; f.k[10] = c.K3; f.k[11] = c.invsh1sq;
    } else {
        f.k[2] = vturn; f.k[3] = c.shuX; f.k[4] = c.chuX; f.k[5] = c.shu2; f.k[6] = c.Iv; f.k[7] = c.KV;
        f.k[8] = c.Lzsq; f.k[9] = 0; f.k[10] = 0; f.k[11] = 0;
    }
    return f;
}


// Everything of df()/uvOrb before the two QAGS integrals for one node.
// Returns 0 (H >= 0: f = 0), 1 (bound; c/ubar/Deltau/vturn valid) or 2
// (unbound: Jr = Jz = 1e31).  want_u / want_v flag the integrals to run;
// otherwise Jr / Jz hold their final value (0 at a degenerate turning point).
// nev counts the potential + gradient evaluations (duturnfn); *nphi, when
// given, the potential evaluations, those with a gradient included.
FJ_HD inline int node_setup(const PotView &T, const double *Eg, const double *Dg, int ng, double R, double z,
                            double Phi_h, const double *V, double Dx, bool check_H, NodeConst &c, double &ubar,
                            double &Deltau, double &vturn, int &want_u, int &want_v, double &Jr, double &Jz,
                            long long &nev, int &brent_it, int *bracket_fail = nullptr, int *nphi = nullptr) {
    int flag = 0;
    want_u = 0; want_v = 0; Jr = 0; Jz = 0; ubar = 0; Deltau = 0; vturn = 0;
    const double H = .5 * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) + Phi_h;   // DF.cpp:83
    if (check_H && !(H < 0)) return 0;
    const double Delta = Dx > 0 ? Dx : d_deltafn(Eg, Dg, ng, H);
    // pt(Delta, x, p): Uvpt.cpp:11-24
    const double Delta2 = Delta * Delta, R2 = R * R, z2 = z * z, r2 = R2 + z2;
    const double shu2 = .5 * (r2 - Delta2 + sqrt((Delta2 - r2) * (Delta2 - r2) + 4 * Delta2 * R2)) / Delta2;
    const double shu = sqrt(shu2), u = fjm_asinh(shu);
    const double chu = sqrt(1 + shu2);
    const double cv2 = .5 * (r2 + Delta2 - sqrt((Delta2 + r2) * (Delta2 + r2) - 4 * Delta2 * z2)) / Delta2;
    const double cv = z < 0 ? -sqrt(cv2) : sqrt(cv2);
    const double v = fjm_acos(cv);
    const double sv = fjm_sin(v), sv2 = 1 - cv2;
    const double pR = V[0], pz = V[2], p2 = pR * pR + pz * pz;
    const double pu = Delta * (pR * chu * sv + pz * shu * cv);
    const double pv = Delta * (pR * shu * cv - pz * chu * sv);
    const double Lz = R * V[1];
    c.Delta = Delta; c.Lzsq = Lz * Lz;
    c.E = .5 * p2 + .5 * c.Lzsq / (R * R) + Phi_h;   // uvOrb.cpp:22
    if (c.E >= 0 || isnan(c.E)) {
        flag = 2; Jr = 1e31; Jz = 1e31;   // uvOrb.cpp:38-40
    } else {
        flag = 1;
        c.sv = sv; c.cv = cv; c.sv2 = sv2; c.sinv = sv; c.cosv = fjm_cos(v); c.shu2 = shu2;
        // Getu0(us = 58): uvOrb.cpp:109-129
        DUturn du{T, &c, &nev};
        double utry = 58, ui = 58, uo = 58, dfui = 0, dfuo = 0;
        double dut = du(utry);
        int guard = 0;
        if (dut > 0) {
            while (dut > 0 && guard++ < 4000) { ui = utry; dfui = dut; utry *= 1.2; dut = du(utry); }
            uo = utry; dfuo = dut;
        } else {
            while (dut < 0 && guard++ < 4000) { uo = utry; dfuo = dut; utry *= .8; dut = du(utry); }
            ui = utry; dfui = dut;
        }
        // a bracket walk that ran into its iteration guard (the reference would
        // loop on): counted, the node continues as the reference would after it
        bool bfail = guard >= 4000;
        int it = 0;
        const double umid = brent_solve(du, ui, uo, dfui, dfuo, &it);
        brent_it += it;
        // second pass at umid: uvOrb.cpp:30-33
        double shm, chm;
    fjm_sinhcosh(umid, &shm, &chm);
        c.sh1sq = shm * shm;
        const double ch1sq = chm * chm;
        if (nphi) *nphi += 3 + (int)nev;   // Phiu1, PhiX, PhivPIH and the duturnfn calls so far
        const double Phiu1 = pot_phi(T, Delta * shm * c.sinv, Delta * chm * c.cosv);
        fjm_sinhcosh(u, &c.shuX, &c.chuX);
        const double PhiX = pot_phi(T, Delta * c.shuX * c.sinv, Delta * c.chuX * c.cosv);
        c.invsh1sq = 1 / c.sh1sq;
        c.K2 = (c.sh1sq + sv2) * Phiu1;
        const double dUX = (c.shuX * c.shuX + sv2) * PhiX - c.K2;
        const double Er = -c.E * (shu2 - c.sh1sq) / ch1sq + .5 * (pu * pu) / (Delta2 * ch1sq) +
                          .5 * c.Lzsq / (Delta2 * ch1sq) * (1 / shu2 - c.invsh1sq) + dUX / ch1sq;
        c.K1 = Er * ch1sq;
        c.K3 = .5 * c.Lzsq / (Delta * Delta);
        double sP, cP;
        fjm_sincos(PIH, &sP, &cP);
        const double PhivPIH = pot_phi(T, Delta * c.shuX * sP, Delta * c.chuX * cP);
        c.KV = (1 + shu2) * PhivPIH;
        const double dVX = c.KV - (shu2 + c.sinv * c.sinv) * PhiX;
        c.Iv = .5 * (pv * pv) / Delta2 - c.E * sv2 + .5 * c.Lzsq / (Delta2 * sv2) - dVX;

        // GetTurnu(1): uvOrb.cpp:131-201
        Uturn U{T, &c, nphi};
        const long long nev0 = nev;
        double ut_ = 1, ut = U(ut_), dutt = du(ut_), eps = .1;
        double uinner = ut_, uouter = ut_, u1 = ut_, u2 = ut_, fu1 = ut, fu2 = ut, fuinner = ut, fuouter = ut;
        bool noroot = false;
        guard = 0;
        if (ut < 0) {
            if (dutt > 0) {
                uinner = ut_; fuinner = ut;
                while (guard++ < 4000) {
                    while (ut < 0 && dutt > 0 && guard++ < 4000) {
                        uinner = ut_; fuinner = ut; ut_ *= (1 + eps);
                        ut = U(ut_); dutt = du(ut_);
                    }
                    if (ut >= 0) { u1 = ut_; fu1 = ut; break; }
                    ut_ /= (1 + eps); eps *= .2;
                    if (eps < TINY) { uinner = uouter = ut_; noroot = true; break; }
                    ut_ *= (1 + eps);
                    ut = U(ut_); dutt = du(ut_);
                }
                if (!noroot) {
                    while (ut > 0 && guard++ < 8000) { u2 = ut_; fu2 = ut; ut_ *= 1.1; ut = U(ut_); }
                    uouter = ut_; fuouter = ut;
                }
            } else {
                while (guard++ < 4000) {
                    while (ut < 0 && dutt < 0 && guard++ < 4000) {
                        uouter = ut_; fuouter = ut; ut_ /= (1 + eps);
                        ut = U(ut_); dutt = du(ut_);
                    }
                    if (ut >= 0) { u2 = ut_; fu2 = ut; break; }
                    ut_ *= (1 + eps); eps *= .2;
                    if (eps < TINY) { uinner = uouter = ut_; noroot = true; break; }
                    ut_ /= (1 + eps);
                    ut = U(ut_); dutt = du(ut_);
                }
                if (!noroot) {
                    while (ut > 0 && guard++ < 8000) { u1 = ut_; fu1 = ut; ut_ *= .9; ut = U(ut_); }
                    uinner = ut_; fuinner = ut;
                }
            }
        } else {
            u1 = u2 = ut_; fu1 = fu2 = ut;
            while (ut > 0 && guard++ < 4000) { u2 = ut_; fu2 = ut; ut_ *= 1.1; ut = U(ut_); }
            uouter = ut_; fuouter = ut;
            // the inward walk starts where the outward one stopped; its first
            // pass pairs u1 = uouter with the stale fu1, whose true F is fuouter
            ut = fu1;
            bool first = true;
            while (ut > 0 && guard++ < 8000) {
                u1 = ut_; fu1 = first ? fuouter : ut; first = false;
                ut_ *= .9; ut = U(ut_);
            }
            uinner = ut_; fuinner = ut;
        }
        bfail = bfail || guard >= 4000;
        if (!noroot) {
            int i1 = 0, i2 = 0;
            const double a = brent_solve(U, uinner, u1, fuinner, fu1, &i1);
            const double b = brent_solve(U, u2, uouter, fu2, fuouter, &i2);
            uinner = a; uouter = b;
            brent_it += i1 + i2;
        }
        if (fabs(uouter - uinner) < TINY) Jr = 0;
        else {
            ubar = .5 * (uinner + uouter);
            Deltau = .5 * (uouter - uinner);
            want_u = 1;
        }
        // GetTurnv(.5*PIH): uvOrb.cpp:203-224
        if (nphi) *nphi += (int)(nev - nev0);   // the duturnfn calls of GetTurnu
        Vturn Vf{T, &c, nphi};
        {
            const double vtry = .5 * PIH;
            double vtop = vtry, vbot = vtry, ft, fb;
            const double f0 = Vf(vtry);
            bool done = false;
            guard = 0;
            if (f0 < 0) {
                fb = f0; ft = f0;
                while (fb < 0 && guard++ < 4000) {
                    vbot = .5 * (PIH + vbot);
                    if (PIH - vbot < TINY) { vturn = PIH; done = true; break; }
                    fb = Vf(vbot);
                }
            } else {
                ft = f0; fb = f0;
                while (ft > 0 && guard++ < 4000) {
                    vtop *= .5;
                    if (vtop < TINY) { vturn = 0; done = true; break; }
                    ft = Vf(vtop);
                }
            }
            if (!done) {
                int i3 = 0;
                if (!(vbot < vtop)) vturn = brent_solve(Vf, vtop, vbot, ft, fb, &i3);
                else vturn = brent_solve(Vf, vbot, vtop, fb, ft, &i3);
                brent_it += i3;
            }
        }
        if (fabs(vturn - PIH) < TINY) Jz = 0;
        else want_v = 1;
        if (bracket_fail) *bracket_fail = bfail || guard >= 4000;
                    }

    return flag;
}

}  // namespace fj
