// san_main.cpp -- drives the oracle (oracle/fjoracle.c) and the host build of
// the device node code (tests/hostbuild/fjhost.cpp) through a small SCF
// workload, built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/sanitize/Makefile).  Test infrastructure: a run must end with status
// 0 and no sanitizer report.
#include <cmath>
#include <cstdio>
#include <vector>

#include "fjoracle.h"

extern "C" {
int fjh_rhofDF(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng, const double *Eg,
               const double *Dg, const double *xa, const double *wa, const double *xp, const double *wp, double R,
               double z, double A, double B, double dphi_h, double dz_h, double dphi_g, double dz_g, double mass,
               double J0, double chi, double k, int rot_form, double *out6);
int fjh_fast_rhofDF(int nr, const double *ar, int npoly, const double *phil, const double *Pr, int ng,
                    const double *Eg, const double *Dg, const double *xa, const double *wa, const double *xp,
                    const double *wp, double R, double z, double A, double B, double dphi_h, double dz_h,
                    double dphi_g, double dz_g, double mass, double J0, double chi, double k, int rot_form, int ngl,
                    const double *glx, const double *glw, int nglv, const double *glxv, const double *glwv,
                    double *out6);
}

static fjo_pot *g_pot;

static int fail(const char *what) {
    std::fprintf(stderr, "san_main: %s\n", what);
    if (g_pot) fjo_pot_free(g_pot);
    return 1;
}

int main() {
    const int nr = 14, npoly = 4, ng = 2, ngrid = 30;
    std::vector<double> ar(nr);
    fjo_log_grid(nr, 0.01, 100.0, ar.data());
    fjo_pot *p = fjo_pot_new(nr, ar.data(), npoly, ng, ngrid);
    g_pot = p;
    std::vector<double> rhl(nr * npoly), phil(nr * npoly), Pr(nr * npoly), Pr2(nr * npoly);
    fjo_guess_rhl(p, 0, 1.0, 1.0, 1.0, rhl.data());                 // Hernquist
    if (fjo_compute_phil(p, rhl.data(), phil.data(), Pr.data(), Pr2.data())) return fail("compute_phil");
    fjo_pot_set_tables(p, phil.data(), Pr.data(), Pr2.data());
    if (fjo_tabulate_delta(p)) return fail("tabulate_delta");
    std::vector<double> E(ngrid), D(ngrid);
    fjo_pot_get_delta(p, E.data(), D.data());
    fjo_df df{1.667, 5.0, 0.55, 0.55, 1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 1};
    // one SCF iteration on the oracle (four threads)
    std::vector<double> t[9];
    for (auto &v : t) v.assign(nr * npoly, 0.0);
    if (fjo_compute_new_phi(p, &df, 4, t[0].data(), t[1].data(), t[2].data(), t[3].data(), t[4].data(),
                            t[5].data(), t[6].data(), t[7].data(), t[8].data()))
        return fail("compute_new_phi");
    double vir[11];
    fjo_vir2(p, t[0].data(), t[1].data(), t[2].data(), t[3].data(), t[4].data(), vir);
    if (!std::isfinite(vir[4])) return fail("vir2");
    // observables
    const double lp = fjo_line_profile(p, &df, t[0].data(), 1.0, 0.3, 0.2);
    const double vd = fjo_vdist(p, &df, 0, 1.0, 0.3, 0.1);
    const double sd = fjo_sigma_df(p, t[0].data(), 1.0, 0.3);
    if (!(std::isfinite(lp) && std::isfinite(vd) && std::isfinite(sd))) return fail("observables");
    // host build of the device node code at two points, exact and fast paths
    double x11[22], w11[22], x01[22], w01[22];
    fjo_quad_tables(x11, w11, x01, w01);
    std::vector<double> gx(64), gw(64), gxv(32), gwv(32);
    fjo_gauleg(64, 0.0, 1.0, gx.data(), gw.data());
    fjo_gauleg(32, 0.0, 1.0, gxv.data(), gwv.data());
    const double pts[2][2] = {{0.5, 0.2}, {3.0, 1.0}};
    for (const auto &q : pts) {
        double a[6], b[6], o[6], vr, s1, s2, s3, s4;
        fjh_rhofDF(nr, ar.data(), npoly, phil.data(), Pr.data(), ngrid, E.data(), D.data(), x11, w11, x11, w11, q[0],
                   q[1], df.A, df.B, df.dphi_h, df.dz_h, df.dphi_g, df.dz_g, df.mass, df.J0, df.chi, df.k,
                   df.rot_form, a);
        fjh_fast_rhofDF(nr, ar.data(), npoly, phil.data(), Pr.data(), ngrid, E.data(), D.data(), x11, w11, x11, w11,
                        q[0], q[1], df.A, df.B, df.dphi_h, df.dz_h, df.dphi_g, df.dz_g, df.mass, df.J0, df.chi, df.k,
                        df.rot_form, 64, gx.data(), gw.data(), 32, gxv.data(), gwv.data(), b);
        fjo_rhofDF(p, &df, q[0], q[1], o, &vr, &s1, &s2, &s3, &s4);
        for (int m = 0; m < 6; m++)
            if (a[m] != o[m]) return fail("host build differs from the oracle");
        std::printf("rho at (%g, %g): oracle %.10g, fast path %.10g\n", q[0], q[1], o[0], b[0]);
        if (!(std::fabs(b[0] - o[0]) < 1e-2 * std::fabs(o[0]))) return fail("fast path far from the oracle");
    }
    fjo_pot_free(p);
    std::printf("san_main: ok (W/K %.6f)\n", vir[4]);
    return 0;
}
