// Host build of the gfx950 node code (fj_node.hpp / fj_device.hpp) for tests:
// the same source the kernels run, compiled for the CPU, so logic can be
// compared with the oracle node by node without a GPU.
#include <vector>

#include "fj_node.hpp"
#include "fj_fast.hpp"
#include "fj_math.h"

using namespace fj;

namespace {
struct CFn {
    double (*f)(double, void *);
    void *p;
    double operator()(double x) const { return f(x, p); }
};
}  // namespace

extern "C" {

// Elementary functions of fj_math.h on n arguments (which: 0 exp, 1 expm1,
// 2 sinh, 3 cosh, 4 sin, 5 cos, 6 acos, 7 asinh, 8 log, 9 tanh, 10 pow(x, y[i])).
int fjh_math(int which, int n, const double *x, double *y) {
    for (int i = 0; i < n; i++) {
        const double a = x[i];
        double r = 0;
        switch (which) {
        case 0: r = fjm_exp(a); break;
        case 1: r = fjm_expm1(a); break;
        case 2: r = fjm_sinh(a); break;
        case 3: r = fjm_cosh(a); break;
        case 4: r = fjm_sin(a); break;
        case 5: r = fjm_cos(a); break;
        case 6: r = fjm_acos(a); break;
        case 7: r = fjm_asinh(a); break;
        case 8: r = fjm_log(a); break;
        case 9: r = fjm_tanh(a); break;
        case 10: r = fjm_pow(a, y[i]); break;
        case 11: r = fjm_atan(a); break;
        case 12: r = fjm_atan2(a, y[i]); break;
        }
        y[i] = r;
    }
    return 0;
}

// Potential view over host arrays, with the bracket table and the radial
// cell divisors filled as the library fills them.
struct HostPot {
    LegCoef lc;
    int lut[NLUT];
    std::vector<double> rk;
    PotView T;
    HostPot(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int nlc)
        : lc(make_legcoef(nlc)), rk((size_t)RK * nr) {
        for (int b = 0; b < nr; b++) fill_radk_cell(nr, ar, b, &rk[(size_t)RK * b]);
        T = PotView{nr, npoly, ar, phil, Pr, &lc, lut, 0.f, 0.f, rk.data()};
        make_lut(nr, ar, lut, &T.lg0, &T.lsc);
    }
};

int fjh_qags(double (*f)(double, void *), void *p, double a, double b, double epsabs, double epsrel, int limit,
             int cap, int rule15, double *result, int *ier, int *nqk) {
    // lane 2 of 3 lanes sharing an interleaved workspace
    std::vector<double> ws(ws_doubles(cap, 3));
    QagsWork w = lane_work(ws.data(), cap, 2, 3);
    CFn fn{f, p};
    if (rule15) *result = qags_solve<8>(fn, a, b, epsabs, epsrel, limit, w, cap, ier, nqk);
    else *result = qags_solve<11>(fn, a, b, epsabs, epsrel, limit, w, cap, ier, nqk);
    return 0;
}

// The float instance of the stepper (QUADPACK's single-precision qagse with
// the FP32 QK21 rule of the mixed-precision kernels) on a float integrand.
struct CFnF {
    float (*f)(float, void *);
    void *p;
    float operator()(float x) const { return f(x, p); }
};
struct FvF {
    float v[20];
    float &operator[](int k) { return v[k]; }
};
int fjh_qags32(float (*f)(float, void *), void *p, float a, float b, float epsabs, float epsrel, int limit,
               int cap, float *result, int *ier, int *nqk) {
    std::vector<double> wsd(ws_doubles(cap, 3));   // 3 lanes, sized as the kernels size it
    const QagsWorkT<float> w = lane_work(reinterpret_cast<float *>(wsd.data()), cap, 1, 3);
    const CFnF fn{f, p};
    QagsT<float> q;
    qags_begin(q, a, b, epsabs, epsrel, limit);
    FvF fv;
    int n = 0;
    while (q.state != 3) {
        const QKOutT<float> r = qk_rule32<11>(fn, q.next_a, q.next_b, fv);
        n++;
        qags_consume(q, w, cap, r);
    }
    *result = q.result;
    *ier = q.overflow ? -1 : q.ier;
    *nqk = n;
    return 0;
}

// Radial bracket of radial_interp (bracket table + walk) at n radii.
int fjh_bracket(int nr, const double *ar, int n, const double *r, int *bot) {
    HostPot H(nr, ar, 1, ar, ar, 1);
    const PotView &T = H.T;
    for (int i = 0; i < n; i++) {
        const Interp I = radial_interp(T, r[i]);
        bot[i] = I.kep ? -1 : I.bot;
    }
    return 0;
}

double fjh_phi_grad(int nr, const double *ar, int npoly, const double *phil, const double *Pr, double R, double z,
                    double *dR, double *dz) {
    HostPot H(nr, ar, npoly, phil, Pr, npoly);
    const PotView &T = H.T;
    return pot_phi_grad(T, R, z, dR, dz);
}

double fjh_phi(int nr, const double *ar, int npoly, const double *phil, const double *Pr, double R, double z) {
    HostPot H(nr, ar, npoly, phil, Pr, npoly);
    const PotView &T = H.T;
    return pot_phi(T, R, z);
}

// Node setup records (NodeBuf field order) at n points, for diffing with the GPU.
int fjh_nodes(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng, const double *Eg,
              const double *Dg, int n, const double *R, const double *z, const double *Phi, const double *vR,
              const double *vphi, const double *vz, const double *Dx, double *out) {
    HostPot H(nr, ar, npoly, phil, Pr, npoly);
    const PotView &T = H.T;
    for (int i = 0; i < n; i++) {
        NodeConst c{};
        double ubar, Deltau, vturn, jr, jz;
        int wu, wv, bit = 0;
        long long nev = 0;
        const double V[3] = {vR[i], vphi[i], vz[i]};
        node_setup(T, Eg, Dg, ng, R[i], z[i], Phi[i], V, Dx ? Dx[i] : 0.0, false, c, ubar, Deltau, vturn, wu, wv,
                   jr, jz, nev, bit);
        const double f[21] = {c.Delta, c.E, c.Lzsq, c.sv, c.cv, c.sv2, c.sinv, c.cosv, c.shuX, c.chuX, c.shu2,
                              c.sh1sq, c.K1, c.K2, c.K3, c.invsh1sq, c.Iv, c.KV, ubar, Deltau, vturn};
        for (int k = 0; k < 21; k++) out[(size_t)k * n + i] = f[k];
    }
    return 0;
}

// Integrand samples and one QAGS run for one node (mirror of the GPU diagnostic).
int fjh_integral(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng, const double *Eg,
                 const double *Dg, double R, double z, double Phi, const double *V, double Dx, int type, int npsi,
                 const double *psi, double *out) {
    HostPot H(nr, ar, npoly, phil, Pr, npoly);
    const PotView &T = H.T;
    NodeConst c{};
    double ubar, Deltau, vturn, jr, jz;
    int wu, wv, bit = 0;
    long long nev = 0;
    node_setup(T, Eg, Dg, ng, R, z, Phi, V, Dx, false, c, ubar, Deltau, vturn, wu, wv, jr, jz, nev, bit);
    const JInt f = make_jint(&T, c, ubar, Deltau, vturn, type);
    for (int i = 0; i < npsi; i++) out[i] = f(psi[i]);
    std::vector<double> ws(ws_doubles(1000, 1));
    QagsWork w = lane_work(ws.data(), 1000, 0, 1);
    int ier = 0, nqk = 0;
    out[npsi] = qags_solve<11>(f, 0, type == 0 ? PI : PIH, EPSABSINTEG, EPSRELINTEG, 1000, w, 1000, &ier, &nqk);
    out[npsi + 1] = ier;
    out[npsi + 2] = nqk;
    const QKOut r = qk_rule<11>(f, 0, type == 0 ? PI : PIH, FvLocal<20>{});
    out[npsi + 3] = r.result; out[npsi + 4] = r.abserr; out[npsi + 5] = r.resabs; out[npsi + 6] = r.resasc;
    return 0;
}

// rhofDF sums at one point as k_points + k_setup + k_integrals + k_moments
// compute them (same per-node code, same reduction order).
int fjh_rhofDF(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng, const double *Eg,
               const double *Dg, const double *xa, const double *wa, const double *xp, const double *wp, double R,
               double z, double A, double B, double dphi_h, double dz_h, double dphi_g, double dz_g, double mass,
               double J0, double chi, double k, int rot_form, double *out6) {
    HostPot H(nr, ar, npoly, phil, Pr, npoly);
    const PotView &T = H.T;
    std::vector<double> ws(ws_doubles(1000, 1));
    QagsWork w = lane_work(ws.data(), 1000, 0, 1);
    const double far = 10 * ar[nr - 1];
    const double Phi_h = pot_phi(T, R, z);
    const double Ve = sqrt(-2 * (Phi_h - pot_phi(T, far, far)));
    const double jacob = Ve * Ve * Ve;
    double part[256][6] = {};
    for (int node = 0; node < 22 * 22 * 22; node++) {
        const int i = node / 484, j = (node / 22) % 22, kk = node % 22;
        const double V[3] = {1e-6 + Ve * xa[i], 1e-6 + Ve * xp[j], 1e-6 + Ve * xa[kk]};
        const double wt = wa[i] * wp[j] * wa[kk];
        NodeConst c;
        double ubar, Deltau, vturn, jr, jz;
        int wu, wv, bit = 0;
        long long nev = 0;
        const int flag = node_setup(T, Eg, Dg, ng, R, z, Phi_h, V, 0.0, true, c, ubar, Deltau, vturn, wu, wv, jr,
                                    jz, nev, bit);
        if (flag == 0) continue;
        if (wu) {
            const JInt f = make_jint(&T, c, ubar, Deltau, vturn, 0);
            jr = f.action(qags_solve<11>(f, 0, PI, EPSABSINTEG, EPSRELINTEG, 1000, w, 1000));
        }
        if (wv) {
            const JInt f = make_jint(&T, c, ubar, Deltau, vturn, 1);
            jz = f.action(qags_solve<11>(f, 0, PIH, EPSABSINTEG, EPSRELINTEG, 1000, w, 1000));
        }
        const double Jphi = R * V[1];
        const double hh = jr + dphi_h * fabs(Jphi) + dz_h * jz, gg = jr + dphi_g * fabs(Jphi) + dz_g * jz;
        double DF = mass / fjm_powi(J0, 3) * fj_max(0., fjm_pow(1. + J0 / hh, A) / (fjm_pow(1. + gg / J0, B))) /
                    fjm_powi(TPI, 3);
        if (chi != 0.) {
            const double t = fjm_tanh(chi * Jphi / J0);
            DF = rot_form == 1 ? (1. - k) * DF + k * t * DF : DF + k * t * DF;
        }
        const double jd = jacob * DF;
        double *s = part[node & 255];
        s[0] += wt * jd; s[1] += wt * (jd * V[1]); s[2] += wt * (jd * V[0] * V[0]);
        s[3] += wt * (jd * V[1] * V[1]); s[4] += wt * (jd * V[2] * V[2]); s[5] += wt * (jd * V[0] * V[2]);
    }
    for (int st = 128; st > 0; st >>= 1)
        for (int t = 0; t < st; t++)
            for (int m = 0; m < 6; m++) part[t][m] += part[t + st][m];
    for (int m = 0; m < 6; m++) out6[m] = part[0][m];
    return 0;
}

// Actions at n phase-space points exactly as k_setup + k_integrals compute them;
// nqk (optional, 2n) receives the number of QK21 rules of the Ju and Jv integrals.
int fjh_actions(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng,
                const double *Eg, const double *Dg, int n, const double *R, const double *z, const double *Phi,
                const double *vR, const double *vphi, const double *vz, const double *Dx, int cap, double *Jr,
                double *Jz, int *nqk) {
    HostPot H(nr, ar, npoly, phil, Pr, npoly);
    const PotView &T = H.T;
    std::vector<double> ws(ws_doubles(cap, 1));
    QagsWork w = lane_work(ws.data(), cap, 0, 1);
    for (int i = 0; i < n; i++) {
        NodeConst c;
        double ubar, Deltau, vturn, jr, jz;
        int wu, wv, bit = 0;
        long long nev = 0;
        const double V[3] = {vR[i], vphi[i], vz[i]};
        node_setup(T, Eg, Dg, ng, R[i], z[i], Phi[i], V, Dx ? Dx[i] : 0.0, false, c, ubar, Deltau, vturn, wu, wv,
                   jr, jz, nev, bit);
        int q1 = 0, q2 = 0, ier = 0;
        if (wu) {
            const JInt f = make_jint(&T, c, ubar, Deltau, vturn, 0);
            const double res = qags_solve<11>(f, 0, PI, EPSABSINTEG, EPSRELINTEG, 1000, w, cap, &ier, &q1);
            jr = f.action(res);
        }
        if (wv) {
            const JInt f = make_jint(&T, c, ubar, Deltau, vturn, 1);
            const double res = qags_solve<11>(f, 0, PIH, EPSABSINTEG, EPSRELINTEG, 1000, w, cap, &ier, &q2);
            jz = f.action(res);
        }
        Jr[i] = jr;
        Jz[i] = jz;
        if (nqk) { nqk[2 * i] = q1; nqk[2 * i + 1] = q2; }
    }
    return 0;
}

// Fast-mode actions (fj_fast.hpp) at n phase-space points with an ngl-point
// Gauss-Legendre rule (x, w on [0, 1]); nev[i]: potential evaluations of the
// turning-point search of point i; tp (optional, 3n): uin, uout, vturn.
int fjh_fast_actions(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng,
                     const double *Eg, const double *Dg, int n, const double *R, const double *z, const double *Phi,
                     const double *vR, const double *vphi, const double *vz, int ngl, const double *x,
                     const double *w, int nglv, const double *xv, const double *wv, double *Jr, double *Jz, int *nev,
                     double *tp) {
    HostPot H(nr, ar, npoly, phil, Pr, npoly);
    const PotView &T = H.T;
    for (int i = 0; i < n; i++) {
        const double V[3] = {vR[i], vphi[i], vz[i]};
        FastNode fn;
        int ne = 0;
        const int flag = fast_setup<FastPolicy64>(T, T, Eg, Dg, ng, R[i], z[i], Phi[i], V, fn, &ne);
        double jr = 0, jz = 0;
        if (flag == 2) { jr = 1e31; jz = 1e31; }
        else if (flag == 1) fast_actions64(T, fn, ngl, x, w, nglv, xv, wv, &jr, &jz);
        Jr[i] = jr; Jz[i] = jz;
        if (nev) nev[i] = ne;
        if (tp) { tp[3 * i] = fn.uin; tp[3 * i + 1] = fn.uout; tp[3 * i + 2] = fn.vt; }
    }
    return 0;
}

// rhofDF's six sums at one point with the fast-mode actions (FP64 integrands,
// ngl-point Gauss-Legendre rule), summed in node order.
int fjh_fast_rhofDF(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng,
                    const double *Eg, const double *Dg, const double *xa, const double *wa, const double *xp,
                    const double *wp, double R, double z, double A, double B, double dphi_h, double dz_h,
                    double dphi_g, double dz_g, double mass, double J0, double chi, double k, int rot_form, int ngl,
                    const double *gx, const double *gw, int nglv, const double *gxv, const double *gwv,
                    double *out6) {
    HostPot H(nr, ar, npoly, phil, Pr, npoly);
    const PotView &T = H.T;
    const double far = 10 * ar[nr - 1];
    const double Phi_h = pot_phi(T, R, z);
    const double Ve = sqrt(-2 * (Phi_h - pot_phi(T, far, far)));
    const double jacob = Ve * Ve * Ve;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int node = 0; node < 22 * 22 * 22; node++) {
        const int i = node / 484, j = (node / 22) % 22, kk = node % 22;
        const double V[3] = {1e-6 + Ve * xa[i], 1e-6 + Ve * xp[j], 1e-6 + Ve * xa[kk]};
        const double wt = wa[i] * wp[j] * wa[kk];
        FastNode fn;
        int ne = 0;
        const int flag = fast_setup<FastPolicy64>(T, T, Eg, Dg, ng, R, z, Phi_h, V, fn, &ne);
        if (flag == 0) continue;
        double jr = 1e31, jz = 1e31;
        if (flag == 1) fast_actions64(T, fn, ngl, gx, gw, nglv, gxv, gwv, &jr, &jz);
        const double Jphi = R * V[1];
        const double hh = jr + dphi_h * fabs(Jphi) + dz_h * jz, gg = jr + dphi_g * fabs(Jphi) + dz_g * jz;
        double DF = mass / fjm_powi(J0, 3) * fj_max(0., fjm_pow(1. + J0 / hh, A) / (fjm_pow(1. + gg / J0, B))) /
                    fjm_powi(TPI, 3);
        if (chi != 0.) {
            const double t = fjm_tanh(chi * Jphi / J0);
            DF = rot_form == 1 ? (1. - k) * DF + k * t * DF : DF + k * t * DF;
        }
        const double jd = jacob * DF;
        s[0] += wt * jd; s[1] += wt * (jd * V[1]); s[2] += wt * (jd * V[0] * V[0]);
        s[3] += wt * (jd * V[1] * V[1]); s[4] += wt * (jd * V[2] * V[2]); s[5] += wt * (jd * V[0] * V[2]);
    }
    for (int m = 0; m < 6; m++) out6[m] = s[m];
    return 0;
}

// fj_div_rn against IEEE division: n random quotients per divisor class.
// mode 0: a, b with random significands and exponents in [-60, 60];
// mode 1: b from the Legendre constants ed/li and radial-cell divisors of a
// log grid, a random.  Returns the number of results that differ in any bit.
long long fjh_divrn_check(long long n, unsigned long long seed, int mode) {
    unsigned long long x = seed | 1;
    auto next = [&x]() { x ^= x << 13; x ^= x >> 7; x ^= x << 17; return x; };
    auto rnd = [&](int emin, int emax) {
        const unsigned long long m = next() & 0x000fffffffffffffull;
        const int e = emin + (int)(next() % (unsigned long long)(emax - emin + 1));
        uint64_t u = ((uint64_t)(1023 + e) << 52) | m;
        if (next() & 1) u |= 0x8000000000000000ull;
        double d; memcpy(&d, &u, 8); return d;
    };
    std::vector<double> divs;
    if (mode == 1) {
        const LegCoef lc = make_legcoef(MAXPOLY);
        for (int np = 2; np < MAXPOLY; np++) divs.push_back(lc.ed[np]);
        for (int i = 1; i < 2 * MAXPOLY; i++) divs.push_back(lc.li[i]);
        const int nr = 200;
        std::vector<double> ar(nr), rk((size_t)RK * nr);
        for (int i = 0; i < nr; i++) ar[i] = 1e-4 * pow(1e6, i / (double)(nr - 1));
        for (int b = 0; b < nr; b++) fill_radk_cell(nr, ar.data(), b, &rk[(size_t)RK * b]);
        for (int b = 0; b < nr - 1; b++) { divs.push_back(rk[RK * b]); if (b < 12) { if (b) divs.push_back(rk[RK * b + 2]); divs.push_back(rk[RK * b + 5]); } }
    }
    long long bad = 0;
    for (long long i = 0; i < n; i++) {
        const double b = mode == 0 ? rnd(-60, 60) : divs[i % divs.size()];
        const double a = rnd(-60, 60);
        const double y = 1.0 / b;
        const double q = fj_div_rn(a, b, y), r = a / b;
        if (memcmp(&q, &r, 8) != 0) bad++;
    }
    // zeros keep the IEEE sign
    for (double a : {0.0, -0.0})
        for (double b : {3.0, -3.0}) {
            const double q = fj_div_rn(a, b, 1.0 / b), r = a / b;
            if (memcmp(&q, &r, 8) != 0) bad++;
        }
    return bad;
}
}  // extern "C"
