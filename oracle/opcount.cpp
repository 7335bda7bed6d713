// opcount.cpp -- operation count of the action-integral work (TEST AND
// MEASUREMENT INFRASTRUCTURE: only tests/ and bench.py's roofline leg load it).
//
// The reference's integrands Juint / Jvint (uvOrb/uvOrb.cpp:42-83,226-262)
// with Phiu / Phiv through Potential::operator() (pot/Potential.cpp:206-219,
// intpo2 and evenlegend, include/UtilsLeg.h:68-124), and the sums of the QK21
// rule (QUADPACK dqk21, as gsl_integration_qk21), are restated here on a
// counting scalar: every add, sub, mul, div and sqrt counts 1, an FMA would
// count 2 (none is written), every elementary-function call (sin, cos, sinh,
// cosh, pow, powi) counts 1, and comparisons, fabs and max count 0.  Two forms:
// the reference's as written (sinh u is called three times per Ju evaluation,
// Phiv(pi/2) is evaluated again in every Jv evaluation), and the same
// integrand with its per-orbit constants computed once.  The QK21 rules are
// counted over the intervals the oracle's QAGS actually bisected to, which the
// oracle records per node (fjo_orbit_state); QAGS's own bookkeeping (the
// interval list, the epsilon extrapolation) is not counted.
#include <cmath>
#include <cstring>

#include "fjoracle.h"

namespace {

thread_local long long g_ops = 0;

struct Cd {
    double v;
    Cd() : v(0) {}
    Cd(double x) : v(x) {}   // constants enter free
};
inline Cd operator+(Cd a, Cd b) { g_ops++; return Cd(a.v + b.v); }
inline Cd operator-(Cd a, Cd b) { g_ops++; return Cd(a.v - b.v); }
inline Cd operator*(Cd a, Cd b) { g_ops++; return Cd(a.v * b.v); }
inline Cd operator/(Cd a, Cd b) { g_ops++; return Cd(a.v / b.v); }
inline Cd operator-(Cd a) { return Cd(-a.v); }
inline Cd &operator+=(Cd &a, Cd b) { a = a + b; return a; }
inline Cd &operator*=(Cd &a, Cd b) { a = a * b; return a; }
inline bool operator<(Cd a, Cd b) { return a.v < b.v; }
inline bool operator>=(Cd a, Cd b) { return a.v >= b.v; }
inline Cd Sqrt(Cd a) { g_ops++; return Cd(std::sqrt(a.v)); }
inline Cd Sin(Cd a) { g_ops++; return Cd(std::sin(a.v)); }
inline Cd Cos(Cd a) { g_ops++; return Cd(std::cos(a.v)); }
inline Cd Sinh(Cd a) { g_ops++; return Cd(std::sinh(a.v)); }
inline Cd Cosh(Cd a) { g_ops++; return Cd(std::cosh(a.v)); }
inline Cd Pow2(Cd a) { g_ops++; return Cd(a.v * a.v); }                  // pow(x, 2)
inline Cd Pow15(Cd a) { g_ops++; return Cd(std::pow(a.v, 1.5)); }
inline Cd Powi(Cd a, int n) { g_ops++; return Cd(std::pow(a.v, n)); }   // fjm_powi
inline Cd Fabs(Cd a) { return Cd(std::fabs(a.v)); }
inline Cd Max0(Cd a) { return Cd(a.v > 0 ? a.v : 0.0); }

struct Tab {
    int nr, npoly;
    const double *ar, *phil;
};

// topbottom (include/Utils.h:112-128): comparisons only
void topbottom(const double *xm, int n, double x, int *bot, int *top) {
    int b = 0, t = n - 1;
    while (t - b > 1) {
        const int m = (t + b) >> 1;
        if (x >= xm[m]) b = m; else t = m;
    }
    *bot = b; *top = t;
}

// Potential::operator() (Potential.cpp:206-219): evenlegend + intpo2 (one table)
Cd phi(const Tab &T, Cd R, Cd z) {
    const Cd c = z / Sqrt(R * R + z * z), r0 = Sqrt(R * R + z * z);
    const int np_ = T.npoly, NR = T.nr;
    Cd pol[64], php[64];
    // evenlegend (UtilsLeg.h:68-79)
    const Cd c2 = c * c;
    pol[0] = 1.0;
    if (np_ > 1) pol[1] = Cd(1.5) * c2 - Cd(.5);
    for (int np = 2; np < np_; np++) {
        const int l = 2 * (np - 1), l2 = 2 * l;
        pol[np] = -pol[np - 2] * Cd(l) * Cd(l - 1) / Cd((double)((l2 + 1) * (l2 - 1))) +
                  pol[np - 1] * (c2 - Cd(l2 * l + l2 - 1) / Cd((double)((l2 - 1) * (l2 + 3))));
        pol[np] *= Cd((l2 + 1) * (l2 + 3)) / Cd((double)((l + 1) * (l + 2)));
    }
    // intpo2 (UtilsLeg.h:96-124)
    Cd r = r0;
    if (r.v >= T.ar[NR - 1]) {
        for (int k = 0; k < np_; k++) php[k] = Cd(T.phil[(NR - 1) * np_ + k]) * Powi(Cd(T.ar[NR - 1]) / r, 2 * k + 1);
    } else {
        if (r.v < T.ar[0]) r = T.ar[0];
        int bot, top;
        topbottom(T.ar, NR, r.v, &bot, &top);
        const Cd db = r - Cd(T.ar[bot]), f1 = db / (Cd(T.ar[top]) - Cd(T.ar[bot]));
        for (int k = 0; k < np_; k++)
            php[k] = f1 * Cd(T.phil[top * np_ + k]) + (Cd(1.0) - f1) * Cd(T.phil[bot * np_ + k]);
        if (top < 10) {
            const int thr = (f1.v < 0.5 && bot > 0) ? bot - 1 : top + 1;
            const Cd dt = r - Cd(T.ar[top]);
            const Cd f2 = dt * db / ((Cd(T.ar[thr]) - Cd(T.ar[top])) * (Cd(T.ar[thr]) - Cd(T.ar[bot])));
            const Cd f3 = (Cd(T.ar[thr]) - Cd(T.ar[bot])) / (Cd(T.ar[top]) - Cd(T.ar[bot]));
            for (int k = 0; k < np_; k++)
                php[k] += f2 * (Cd(T.phil[thr * np_ + k]) - Cd(T.phil[bot * np_ + k]) -
                                f3 * (Cd(T.phil[top * np_ + k]) - Cd(T.phil[bot * np_ + k])));
        }
    }
    Cd p = php[0];
    for (int np = 1; np < np_; np++) p += php[np] * pol[np];
    return p;
}

// The state of one uvOrb (uvOrb.cpp:15-36) that the integrands read.
struct Orb {
    double Delta, E, Er, ch1sq, sh1sq, Phiu1, Lzsq, u, v, sv2, shu2, Iv, ubar, Deltau, vturn;
};

// Juint (uvOrb.cpp:80-83) with uturnfn (70-73), dU (61-63), Phiu (57-60).
// hoist = 0: as the reference writes it; hoist = 1: with the per-orbit
// constants ((sh1sq + sv2) Phiu1, Er ch1sq, .5 Lzsq / Delta^2, 1 / sh1sq,
// sin v, cos v) computed once and sinh u once per evaluation.
thread_local int g_hoist = 0;

Cd juint(const Tab &T, const Orb &o, Cd psi) {
    const Cd u = Cd(o.ubar) - Cd(o.Deltau) * Cos(psi);
    if (g_hoist) {
        const long long keep = g_ops;
        const Cd K2 = (Cd(o.sh1sq) + Cd(o.sv2)) * Cd(o.Phiu1), K1 = Cd(o.Er) * Cd(o.ch1sq);
        const Cd K3 = Cd(.5) * Cd(o.Lzsq) / (Cd(o.Delta) * Cd(o.Delta)), inv1 = Cd(1.0) / Cd(o.sh1sq);
        const Cd Dsv = Cd(o.Delta) * Sin(Cd(o.v)), Dcv = Cd(o.Delta) * Cos(Cd(o.v));
        g_ops = keep;
        const Cd shu = Sinh(u), chu = Cosh(u), shu2 = shu * shu;
        const Cd dU = (shu2 + Cd(o.sv2)) * phi(T, shu * Dsv, chu * Dcv) - K2;
        const Cd F = Cd(o.E) * (shu2 - Cd(o.sh1sq)) + K1 - dU - K3 * (Cd(1.0) / shu2 - inv1);
        return Sqrt(Max0(F)) * Sin(psi);
    }
    const Cd shu2 = Pow2(Sinh(u));
    const Cd phiu = phi(T, Cd(o.Delta) * Sinh(u) * Sin(Cd(o.v)), Cd(o.Delta) * Cosh(u) * Cos(Cd(o.v)));
    const Cd dU = (Pow2(Sinh(u)) + Cd(o.sv2)) * phiu - (Cd(o.sh1sq) + Cd(o.sv2)) * Cd(o.Phiu1);
    const Cd F = Cd(o.E) * (shu2 - Cd(o.sh1sq)) + Cd(o.Er) * Cd(o.ch1sq) - dU -
                 Cd(.5) * Cd(o.Lzsq) / (Cd(o.Delta) * Cd(o.Delta)) * (Cd(1.0) / shu2 - Cd(1.0) / Cd(o.sh1sq));
    return Sqrt(Max0(F)) * Sin(psi);
}

// Jvint (uvOrb.cpp:53-56) with vturnfn (49-52), dV (46-48), Phiv (42-45)
Cd phiv(const Tab &T, const Orb &o, Cd v) {
    return phi(T, Cd(o.Delta) * Sinh(Cd(o.u)) * Sin(v), Cd(o.Delta) * Cosh(Cd(o.u)) * Cos(v));
}
Cd jvint(const Tab &T, const Orb &o, Cd psi) {
    const double PIH = 1.570796326794897;
    const Cd v = Cd(PIH) - (Cd(PIH) - Cd(o.vturn)) * Cos(psi);
    if (g_hoist) {
        // (1 + shu2) Phiv(PIH), Delta sinh u, Delta cosh u and .5 Lzsq / Delta^2 once per orbit
        const long long keep = g_ops;
        const Cd KV = (Cd(1.0) + Cd(o.shu2)) * phiv(T, o, Cd(PIH));
        const Cd Dsu = Cd(o.Delta) * Sinh(Cd(o.u)), Dcu = Cd(o.Delta) * Cosh(Cd(o.u));
        const Cd K3 = Cd(.5) * Cd(o.Lzsq) / (Cd(o.Delta) * Cd(o.Delta));
        g_ops = keep;
        const Cd s = Sin(v), c = Cos(v), sv2 = s * s;
        const Cd dV = KV - (Cd(o.shu2) + sv2) * phi(T, Dsu * s, Dcu * c);
        const Cd F = Cd(o.E) * sv2 + Cd(o.Iv) + dV - K3 / sv2;
        return Sqrt(Max0(F)) * Sin(psi);
    }
    const Cd sv2 = Pow2(Sin(v));
    const Cd dV = (Cd(1.0) + Cd(o.shu2)) * phiv(T, o, Cd(PIH)) - (Cd(o.shu2) + Pow2(Sin(v))) * phiv(T, o, v);
    const Cd F = Cd(o.E) * sv2 + Cd(o.Iv) + dV - Cd(.5) * Cd(o.Lzsq) / (Cd(o.Delta) * Cd(o.Delta) * sv2);
    return Sqrt(Max0(F)) * Sin(psi);
}

const double XGK[11] = {0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
                        0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
                        0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
                        0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
                        0.294392862701460198131126603103866, 0.148874338981631210884826001129720, 0.0};
const double WGK[11] = {0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
                        0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
                        0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
                        0.123491976262065851077208980178525, 0.134709217311473325928054001771707,
                        0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
                        0.149445554002916905664936468389821};
const double WG[5] = {0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
                      0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
                      0.295524224714752870173892994651338};

// dqk21's sums around 21 integrand values (the rescaled error estimate included)
template <class F>
Cd qk21(F f, Cd a, Cd b, long long *ops_integrand) {
    const Cd center = Cd(.5) * (a + b), half = Cd(.5) * (b - a), ahalf = Fabs(half);
    long long fops = 0;
    auto eval = [&](Cd x) {
        const long long s = g_ops;
        const Cd y = f(x);
        fops += g_ops - s;
        return y;
    };
    const Cd fc = eval(center);
    Cd rg = 0.0, rk = fc * Cd(WGK[10]), rabs = Fabs(rk);
    Cd fv1[10], fv2[10];
    for (int j = 0; j < 5; j++) {
        const int jtw = 2 * j + 1;
        const Cd absc = half * Cd(XGK[jtw]);
        const Cd f1 = eval(center - absc), f2 = eval(center + absc), fs = f1 + f2;
        fv1[jtw] = f1; fv2[jtw] = f2;
        rg += Cd(WG[j]) * fs; rk += Cd(WGK[jtw]) * fs;
        rabs += Cd(WGK[jtw]) * (Fabs(f1) + Fabs(f2));
    }
    for (int j = 0; j < 5; j++) {
        const int jtwm1 = 2 * j;
        const Cd absc = half * Cd(XGK[jtwm1]);
        const Cd f1 = eval(center - absc), f2 = eval(center + absc);
        fv1[jtwm1] = f1; fv2[jtwm1] = f2;
        rk += Cd(WGK[jtwm1]) * (f1 + f2);
        rabs += Cd(WGK[jtwm1]) * (Fabs(f1) + Fabs(f2));
    }
    const Cd mean = rk * Cd(.5);
    Cd rasc = Cd(WGK[10]) * Fabs(fc - mean);
    for (int j = 0; j < 10; j++) rasc += Cd(WGK[j]) * (Fabs(fv1[j] - mean) + Fabs(fv2[j] - mean));
    Cd err = (rk - rg) * half;
    rk *= half; rabs *= ahalf; rasc *= ahalf;
    // rescale_error (QUADPACK dqk21 tail): err = resasc min(1, (200 err / resasc)^1.5),
    // at least 50 eps resabs
    err = Fabs(err);
    if (rasc.v != 0 && err.v != 0) {
        const Cd sc = Pow15(Cd(200.0) * err / rasc);
        err = sc.v < 1 ? rasc * sc : rasc;
    }
    if (rabs.v > 2.2250738585072014e-308 / (50 * 2.220446049250313e-16)) {
        const Cd me = Cd(50 * 2.220446049250313e-16) * rabs;
        if (me.v > err.v) err = me;
    }
    *ops_integrand += fops;
    return rk;
}

}  // namespace

extern "C" {

// Operation counts of the action integrals of one bound node: state holds the
// 15 orbit doubles (Delta E Er ch1sq sh1sq Phiu1 Lzsq u v sv2 shu2 Iv ubar
// Deltau vturn), iv the [a, b] of the nu QK21 rules QAGS applied to Ju then
// the nv applied to Jv.  hoist selects the integrand form (see juint).
// out[0] operations of the Ju rules, out[1] of the Jv rules, out[2] of those
// inside integrand evaluations, out[3] the first Ju rule's result, out[4] the
// first Jv rule's result (NAN when that list is empty).
int fjc_rule_ops(int nr, const double *ar, int npoly, const double *phil, const double *state,
                 const double *iv, int nu, int nv, int hoist, double *out) {
    const Tab T{nr, npoly, ar, phil};
    Orb o;
    std::memcpy(&o, state, sizeof o);
    g_hoist = hoist;
    long long integrand = 0, ops[2] = {0, 0};
    double first[2] = {NAN, NAN};
    for (int i = 0; i < nu + nv; i++) {
        const int type = i < nu ? 0 : 1;
        const long long s = g_ops;
        Cd r;
        if (type == 0) r = qk21([&](Cd x) { return juint(T, o, x); }, Cd(iv[2 * i]), Cd(iv[2 * i + 1]), &integrand);
        else r = qk21([&](Cd x) { return jvint(T, o, x); }, Cd(iv[2 * i]), Cd(iv[2 * i + 1]), &integrand);
        ops[type] += g_ops - s;
        if (i == 0 || i == nu) first[type] = r.v;
    }
    out[0] = (double)ops[0];
    out[1] = (double)ops[1];
    out[2] = (double)integrand;
    out[3] = first[0];
    out[4] = first[1];
    return 0;
}

// One potential evaluation (Potential::operator()) at (R, z): returns its
// operation count and the value in *val.
long long fjc_phi_ops(int nr, const double *ar, int npoly, const double *phil, double R, double z, double *val) {
    const Tab T{nr, npoly, ar, phil};
    const long long s = g_ops;
    *val = phi(T, Cd(R), Cd(z)).v;
    return g_ops - s;
}

// The whole action-integral work of the grid points (R[i], z[i]), i < n, on
// the oracle's potential pot (whose phil table, ar grid and velocity rule are
// passed again here): every bound node of the 22^3 rule of rhofDF
// (integ/Integ.cpp:150-163, vel_3D_rule as the oracle's fjo_rhofDF) runs
// through fjo_orbit_state, and its recorded QK21 rules are counted.
// out[0] Ju-rule operations, out[1] Jv-rule operations, out[2] operations
// inside integrand evaluations, out[3] Ju rules, out[4] Jv rules, out[5]
// bound nodes, out[6] largest relative difference between a counted first
// rule and the oracle's (the restatement's check).  OpenMP over points.
int fjc_grid_ops(const fjo_pot *pot, int nr, const double *ar, int npoly, const double *phil, int quad, int n,
                 const double *R, const double *z, int hoist, double *out) {
    double x11[22], w11[22], x01[22], w01[22];
    fjo_quad_tables(x11, w11, x01, w01);
    const double *xa = quad ? x01 : x11;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma omp parallel
    {
        double loc[7] = {0, 0, 0, 0, 0, 0, 0};
        const int cap = 4096;
        double *iv = new double[2 * cap];
#pragma omp for schedule(dynamic, 1)
        for (int pt = 0; pt < n; pt++) {
            const double Phi_h = fjo_phi(pot, R[pt], z[pt]);
            const double far = 10 * ar[nr - 1];
            const double Ve = std::sqrt(-2 * (Phi_h - fjo_phi(pot, far, far)));
            for (int i = 0; i < 22; i++)
                for (int j = 0; j < 22; j++)
                    for (int k = 0; k < 22; k++) {
                        const double V0 = 1e-6 + Ve * xa[i], V1 = 1e-6 + Ve * x11[j], V2 = 1e-6 + Ve * xa[k];
                        const double H = .5 * (V0 * V0 + V1 * V1 + V2 * V2) + Phi_h;
                        if (H >= 0) continue;
                        double st[15], qk[2];
                        int nrule[2] = {0, 0};
                        if (fjo_orbit_state(pot, fjo_deltafn(pot, H), R[pt], z[pt], Phi_h, V0, V1, V2, st, qk, iv,
                                            cap, nrule))
                            continue;
                        if (nrule[0] + nrule[1] > cap) continue;
                        double o[5];
                        fjc_rule_ops(nr, ar, npoly, phil, st, iv, nrule[0], nrule[1], hoist, o);
                        loc[0] += o[0]; loc[1] += o[1]; loc[2] += o[2];
                        loc[3] += nrule[0]; loc[4] += nrule[1]; loc[5] += 1;
                        for (int t = 0; t < 2; t++)
                            if (nrule[t] > 0) {
                                const double d = std::fabs(o[3 + t] - qk[t]) / std::fmax(std::fabs(qk[t]), 1e-300);
                                if (d > loc[6]) loc[6] = d;
                            }
                    }
        }
        delete[] iv;
#pragma omp critical
        {
            for (int t = 0; t < 6; t++) acc[t] += loc[t];
            if (loc[6] > acc[6]) acc[6] = loc[6];
        }
    }
    for (int t = 0; t < 7; t++) out[t] = acc[t];
    return 0;
}
}
