// Host build of the gfx950 node code (fj_node.hpp / fj_device.hpp) for tests:
// the same source the kernels run, compiled for the CPU, so logic can be
// compared with the oracle node by node without a GPU.
#include <vector>

#include "fj_node.hpp"

using namespace fj;

namespace {
struct CFn {
    double (*f)(double, void *);
    void *p;
    double operator()(double x) const { return f(x, p); }
};
}  // namespace

extern "C" {

int fjh_qags(double (*f)(double, void *), void *p, double a, double b, double epsabs, double epsrel, int limit,
             int cap, int rule15, double *result, int *ier, int *nqk) {
    std::vector<double> ws(lane_ws_doubles(cap));
    QagsWork w = lane_work(ws.data(), cap, 0);
    CFn fn{f, p};
    if (rule15) *result = qags_solve<8>(fn, a, b, epsabs, epsrel, limit, w, cap, ier, nqk);
    else *result = qags_solve<11>(fn, a, b, epsabs, epsrel, limit, w, cap, ier, nqk);
    return 0;
}

double fjh_phi_grad(int nr, const double *ar, int npoly, const double *phil, const double *Pr, double R, double z,
                    double *dR, double *dz) {
    LegCoef lc = make_legcoef(npoly);
    PotView T{nr, npoly, ar, phil, Pr, &lc};
    return pot_phi_grad(T, R, z, dR, dz);
}

double fjh_phi(int nr, const double *ar, int npoly, const double *phil, const double *Pr, double R, double z) {
    LegCoef lc = make_legcoef(npoly);
    PotView T{nr, npoly, ar, phil, Pr, &lc};
    return pot_phi(T, R, z);
}

// Node setup records (NodeBuf field order) at n points, for diffing with the GPU.
int fjh_nodes(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng, const double *Eg,
              const double *Dg, int n, const double *R, const double *z, const double *Phi, const double *vR,
              const double *vphi, const double *vz, const double *Dx, double *out) {
    LegCoef lc = make_legcoef(npoly);
    PotView T{nr, npoly, ar, phil, Pr, &lc};
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
    LegCoef lc = make_legcoef(npoly);
    PotView T{nr, npoly, ar, phil, Pr, &lc};
    NodeConst c{};
    double ubar, Deltau, vturn, jr, jz;
    int wu, wv, bit = 0;
    long long nev = 0;
    node_setup(T, Eg, Dg, ng, R, z, Phi, V, Dx, false, c, ubar, Deltau, vturn, wu, wv, jr, jz, nev, bit);
    JInt f{&T, &c, ubar, Deltau, vturn, type};
    for (int i = 0; i < npsi; i++) out[i] = f(psi[i]);
    std::vector<double> ws(lane_ws_doubles(1000));
    QagsWork w = lane_work(ws.data(), 1000, 0);
    int ier = 0, nqk = 0;
    out[npsi] = qags_solve<11>(f, 0, type == 0 ? PI : PIH, EPSABSINTEG, EPSRELINTEG, 1000, w, 1000, &ier, &nqk);
    out[npsi + 1] = ier;
    out[npsi + 2] = nqk;
    const QKOut r = qk_rule<11>(f, 0, type == 0 ? PI : PIH);
    out[npsi + 3] = r.result; out[npsi + 4] = r.abserr; out[npsi + 5] = r.resabs; out[npsi + 6] = r.resasc;
    return 0;
}

// Actions at n phase-space points exactly as k_setup + k_integrals compute them.
int fjh_actions(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng,
                const double *Eg, const double *Dg, int n, const double *R, const double *z, const double *Phi,
                const double *vR, const double *vphi, const double *vz, const double *Dx, int cap, double *Jr,
                double *Jz, int *nqk) {
    LegCoef lc = make_legcoef(npoly);
    PotView T{nr, npoly, ar, phil, Pr, &lc};
    std::vector<double> ws(lane_ws_doubles(cap));
    QagsWork w = lane_work(ws.data(), cap, 0);
    const double rtwo = 1.4142135623730951;
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
            JInt f{&T, &c, ubar, Deltau, vturn, 0};
            const double res = qags_solve<11>(f, 0, PI, EPSABSINTEG, EPSRELINTEG, 1000, w, cap, &ier, &q1);
            jr = rtwo * c.Delta * Deltau * res / PI;
        }
        if (wv) {
            JInt f{&T, &c, ubar, Deltau, vturn, 1};
            const double res = qags_solve<11>(f, 0, PIH, EPSABSINTEG, EPSRELINTEG, 1000, w, cap, &ier, &q2);
            jz = 2 * rtwo * c.Delta * (PIH - vturn) * res / PI;
        }
        Jr[i] = jr;
        Jz[i] = jz;
        if (nqk) nqk[i] = q1 + q2;
    }
    return 0;
}
}
