// fj_facade.cpp -- the reference's C++ call surface over the gfx950 C ABI.
// All node, projection, Poisson, Delta and virial work runs in
// libfjmodels.so; this file holds the host-side state the reference keeps in
// globals, parameter parsing and the models/*.out files.
#include "fjm.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace fjm {

unsigned components = 0;
int NR = 0, NGAUSS = 0, NPOLY = 0, NGRID = 0;
std::vector<double> ar, Egrid, Dgrid;
double **phil = nullptr, **Pr = nullptr, **Pr2 = nullptr, **rhl = nullptr, **vrotl = nullptr, **sigRl = nullptr,
       **sigpl = nullptr, **sigzl = nullptr, **sigRzl = nullptr;

namespace {

fj_ctx *g_ctx = nullptr;
int g_device = -1;
// SetModel state (pot/models.cpp:13-14)
double g_mass = 1, g_r0 = 1, g_q = 1;
int g_poisson = FJ_POISSON_QUADPACK;
bool g_grid_dirty = false;

void check(int rc, const char *what) {
    if (rc) {
        std::string msg = std::string(what) + ": " + (g_ctx ? fj_last_error(g_ctx) : "no device context");
        throw Error(rc, msg);
    }
}


// The device context, created and given the grid at first use, so that the
// host-only parts (parameters, grid, files) work without a GPU.
fj_ctx *ctx() {
    if (NR <= 0) throw Error(FJ_E_STATE, "SetGrid must be called first");
    if (!g_ctx) {
        const int rc = fj_ctx_create(g_device < 0 ? 0 : g_device, &g_ctx);
        if (rc) throw Error(rc, "fj_ctx_create failed (no HIP device?)");
        g_grid_dirty = true;
    }
    if (g_grid_dirty) {
        check(fj_set_grid(g_ctx, NR, ar.data(), NGAUSS, NPOLY, NGRID), "fj_set_grid");
        g_grid_dirty = false;
    }
    return g_ctx;
}

double **or_global(double **m, double **g) { return m ? m : g; }

void bind_grid(int ntheta, int npoly, int ngrid, int device) {
    if (g_ctx && g_device != device) {
        fj_ctx_destroy(g_ctx);
        g_ctx = nullptr;
    }
    g_device = device;
    NR = (int)ar.size();
    NGAUSS = ntheta;
    NPOLY = npoly;
    NGRID = ngrid;
    g_grid_dirty = true;
    double ***gl[] = {&phil, &Pr, &Pr2, &rhl, &vrotl, &sigRl, &sigpl, &sigzl, &sigRzl};
    for (double ***g : gl) {
        free_mat(*g);
        *g = mat(NR, NPOLY);
    }
    Egrid.assign(NGRID, 0.0);
    Dgrid.assign(NGRID, 0.0);
}

std::string trim_name(const std::string &s) { return s.size() > 5 ? s.substr(0, 5) : s; }

template <class T>
std::string toString(const T &t, int eps = 2) {   // include/writeOut.h:31-38
    std::stringstream ss;
    ss << std::fixed << std::setprecision(eps) << t;
    return ss.str();
}

}  // namespace

double **mat(int rows, int cols) {
    double **m = new double *[rows > 0 ? rows : 1];
    double *block = new double[(size_t)(rows > 0 ? rows : 1) * (cols > 0 ? cols : 1)]();
    for (int i = 0; i < rows; i++) m[i] = block + (size_t)i * cols;
    if (rows <= 0) m[0] = block;
    return m;
}

void free_mat(double **m) {
    if (!m) return;
    delete[] m[0];
    delete[] m;
}

// ---------------------------------------------------------------------------
// readParam (io/readParam.cpp:31-111): "key value" lines; lines starting
// with '#', ' ', '\t' or empty are skipped; unknown keys are ignored.
// ---------------------------------------------------------------------------
fJParams readParam(const char *fName) {
    std::ifstream pF(fName);
    if (!pF.is_open()) throw Error(FJ_E_ARG, std::string("ERROR in opening parameter file ") + fName);
    fJParams P;
    std::string line;
    while (std::getline(pF, line)) {
        if (line.empty() || line[0] == '#' || line[0] == ' ' || line[0] == '\t') continue;
        std::stringstream ls(line);
        std::string key;
        ls >> key;
        auto val = [&]() { double v = 0; ls >> v; return v; };
        auto str = [&]() { std::string s; ls >> s; return s; };
        if (key == "itermax") P.itermax = (int)val();
        else if (key == "model") P.modName = str();
        else if (key == "A") P.alpha = val();
        else if (key == "B") P.beta = val();
        else if (key == "outfile") P.outName = str();
        else if (key == "chi") P.chi = val();
        else if (key == "h(J):dphi") P.dphi_h_in = val();
        else if (key == "h(J):dz") P.dz_h_in = val();
        else if (key == "g(J):dphi") P.dphi_g_in = val();
        else if (key == "g(J):dz") P.dz_g_in = val();
        else if (key == "mass") P.mass = val();
        else if (key == "r0") P.r0 = val();
        else if (key == "q") P.q = val();
        else if (key == "2:model") P.modName2 = str();
        else if (key == "2:outfile") P.outName2 = str();
        else if (key == "2:A") P.alpha_2 = val();
        else if (key == "2:B") P.beta_2 = val();
        else if (key == "2:chi") P.chi_2 = val();
        else if (key == "2:h(J):dphi") P.dphi_h_in2 = val();
        else if (key == "2:h(J):dz") P.dz_h_in2 = val();
        else if (key == "2:g(J):dphi") P.dphi_g_in2 = val();
        else if (key == "2:g(J):dz") P.dz_g_in2 = val();
        else if (key == "2:mass") P.mass_2 = val();
        else if (key == "2:r0") P.r0_2 = val();
        else if (key == "2:q") P.q_2 = val();
        // keys of this build
        else if (key == "rot_form") P.rot_form = (int)val();
        else if (key == "rot_k") P.rot_k = val();
        else if (key == "2:rot_form") P.rot_form_2 = (int)val();
        else if (key == "2:rot_k") P.rot_k_2 = val();
        else if (key == "nr") P.nr = (int)val();
        else if (key == "ntheta") P.ntheta = (int)val();
        else if (key == "npoly") P.npoly = (int)val();
        else if (key == "ngrid") P.ngrid = (int)val();
        else if (key == "rmin") P.rmin = val();
        else if (key == "rmax") P.rmax = val();
        else if (key == "tol") P.tol = val();
        else if (key == "merge") P.merge = val();
        else if (key == "delta_refresh") P.delta_refresh = (int)val();
        else if (key == "lineprofile") P.lineprofile = (int)val();
        else if (key == "comp2_potential") {
            const std::string s = str();
            if (s == "updated") P.comp2_potential = 0;
            else if (s == "total") P.comp2_potential = 1;
            else throw Error(FJ_E_ARG, "comp2_potential must be updated or total");
        }
        else if (key == "quad") {
            const std::string s = str();
            if (s == "log22") P.quad = FJ_QUAD_LOG22;
            else if (s == "tanhsinh22") P.quad = FJ_QUAD_TANHSINH22;
            else throw Error(FJ_E_ARG, "quad must be log22 or tanhsinh22");
        } else if (key == "poisson") {
            const std::string s = str();
            if (s == "quadpack") P.poisson = FJ_POISSON_QUADPACK;
            else if (s == "blockscan") P.poisson = FJ_POISSON_BLOCKSCAN;
            else throw Error(FJ_E_ARG, "poisson must be quadpack or blockscan");
        } else if (key == "precision") {
            const std::string s = str();
            if (s == "fp64") P.precision = FJ_PREC_FP64;
            else if (s == "mixed") P.precision = FJ_PREC_MIXED;
            else if (s == "fast") P.precision = FJ_PREC_FAST;
            else throw Error(FJ_E_ARG, "precision must be fp64, mixed or fast");
        }
    }
    checkParameterFile(&P);
    return P;
}

// checkParameterFile (readParam.cpp:114-144)
void checkParameterFile(const fJParams *f) {
    const bool one = (f->modName != "null" || (f->alpha != UNSET && f->beta != UNSET)) && f->chi != UNSET &&
                     f->mass != UNSET && f->r0 != UNSET && f->q != UNSET && f->dphi_h_in != UNSET &&
                     f->dz_h_in != UNSET && f->dphi_g_in != UNSET && f->dz_g_in != UNSET;
    if (!one) throw Error(FJ_E_ARG, "ERROR: parameterfile must contain at least one component!");
    components = 1;
    if (f->modName2 != "null" || (f->alpha_2 != UNSET && f->beta_2 != UNSET)) {
        if (f->chi_2 != UNSET && f->mass_2 != UNSET && f->r0_2 != UNSET && f->q_2 != UNSET &&
            f->dphi_h_in2 != UNSET && f->dz_h_in2 != UNSET && f->dphi_g_in2 != UNSET && f->dz_g_in2 != UNSET)
            components++;
        else
            throw Error(FJ_E_ARG, "ERROR in parameterfile: second components is missing some parameter to initialize!");
    }
}

// printParam (readParam.cpp:153-192)
void printParam(const fJParams *f) {
    std::cout << "\n======================================================================" << std::endl;
    std::cout << "= f(J) Model computation" << std::endl;
    std::cout << "======================================================================\n" << std::endl;
    std::cout << "= Maximum number of iterations: " << f->itermax << std::endl;
    std::cout << "=\n" << std::endl;
    std::cout << "PARAMETERS of the model:\n" << std::endl;
    std::cout << " - Classical model:\t" << f->modName << std::endl;
    std::cout << " - h(J) :\t\tJr + " << f->dphi_h_in << " Jphi + " << f->dz_h_in << " Jz" << std::endl;
    std::cout << " - g(J) :\t\tJr + " << f->dphi_g_in << " Jphi + " << f->dz_g_in << " Jz" << std::endl;
    std::cout << " - Chi :\t\t" << f->chi << std::endl;
    std::cout << " - Mass :\t\t" << f->mass << std::endl;
    std::cout << " - r0   :\t\t" << f->r0 << std::endl;
    std::cout << " - Flattening :\t\t" << f->q << std::endl;
    std::cout << std::endl;
    if (f->outName != "null") std::cout << " the model's output file will be " << f->outName << "_*.out\n" << std::endl;
    if (components > 1) {
        std::cout << "PARAMETERS of the second component model:\n" << std::endl;
        std::cout << " - Classical model:\t" << f->modName2 << std::endl;
        std::cout << " - h(J) :\t\tJr + " << f->dphi_h_in2 << " Jphi + " << f->dz_h_in2 << " Jz" << std::endl;
        std::cout << " - g(J) :\t\tJr + " << f->dphi_g_in2 << " Jphi + " << f->dz_g_in2 << " Jz" << std::endl;
        std::cout << " - Chi :\t\t" << f->chi_2 << std::endl;
        std::cout << " - Mass :\t\t" << f->mass_2 << std::endl;
        std::cout << " - r0   :\t\t" << f->r0_2 << std::endl;
        std::cout << " - Flattening :\t\t" << f->q_2 << std::endl;
        std::cout << std::endl;
        if (f->outName2 != "null")
            std::cout << " the model's output file will be " << f->outName2 << "_*.out\n" << std::endl;
    }
}

// ---------------------------------------------------------------------------
// Grid
// ---------------------------------------------------------------------------
void SetGrid(double rmax, int nr, int ntheta, int npoly, int ngrid, double rmin, int device) {
    if (nr < 2) throw Error(FJ_E_ARG, "SetGrid: nr must be >= 2");
    if (!(rmin > 0)) rmin = 5e-6 * rmax;   // include/Utils.h:79
    std::vector<double> r(nr);
    for (int i = 0; i < nr; i++) r[i] = pow(10., log10(rmin) + (log10(rmax) - log10(rmin)) * i / (nr - 1));
    SetGridRadii(r, ntheta, npoly, ngrid, device);
}

void SetGridRadii(const std::vector<double> &radii, int ntheta, int npoly, int ngrid, int device) {
    ar = radii;
    bind_grid(ntheta, npoly, ngrid, device);
}

fj_ctx *context() { return ctx(); }

void set_quad_rule(int preset) { check(fj_set_quad_rule(ctx(), preset), "fj_set_quad_rule"); }

void set_poisson_mode(int mode) { g_poisson = mode; }

void set_precision(int mode) { check(fj_set_precision(ctx(), mode), "fj_set_precision"); }

void SetModel(const fJParams &f, unsigned comp) {
    if (comp == 1) { g_mass = f.mass; g_r0 = f.r0; g_q = f.q; }
    else { g_mass = f.mass_2; g_r0 = f.r0_2; g_q = f.q_2; }
}

// ---------------------------------------------------------------------------
// Potential
// ---------------------------------------------------------------------------
Potential::Potential(unsigned comp_in) : comp(comp_in) {
    if (NR <= 0) throw Error(FJ_E_STATE, "SetGrid must be called before constructing a Potential");
    philP = mat(NR, NPOLY);
    PrP = mat(NR, NPOLY);
    Pr2P = mat(NR, NPOLY);
    rhlP = mat(NR, NPOLY);
}

Potential::~Potential() {
    free_mat(philP);
    free_mat(PrP);
    free_mat(Pr2P);
    free_mat(rhlP);
}

void Potential::selectGuessRho(const std::string &n) {
    if (n == "Isochrone") guess_ = FJ_GUESS_ISOCHRONE;
    else if (n == "Hernquist") guess_ = FJ_GUESS_HERNQUIST;
    else if (n == "NFW") guess_ = FJ_GUESS_NFW;
    else if (n == "NFWext") guess_ = FJ_GUESS_NFWEXT;
    else throw Error(FJ_E_ARG, "selectGuessRho: unknown model '" + n + "'");
}

void Potential::computeGuessRhl() {
    if (guess_ < 0) throw Error(FJ_E_STATE, "selectGuessRho first");
    check(fj_guess_rhl(ctx(), guess_, g_mass, g_r0, g_q, rhlP[0]), "fj_guess_rhl");
    if (comp == 1) memcpy(rhl[0], rhlP[0], sizeof(double) * NR * NPOLY);
}

void Potential::computePhil(double **rhlH) {
    rhlH = or_global(rhlH, rhl);
    check(fj_poisson(ctx(), g_poisson, rhlH[0], philP[0], PrP[0], Pr2P[0]), "fj_poisson");
    if (comp == 1) {
        const size_t b = sizeof(double) * NR * NPOLY;
        memcpy(phil[0], philP[0], b);
        memcpy(Pr[0], PrP[0], b);
        memcpy(Pr2[0], Pr2P[0], b);
        syncPotential();
    }
}

void Potential::eval(int n, const double *R, const double *z, double *phi, double *dR, double *dz) const {
    check(fj_eval_potential(ctx(), n, R, z, phi, dR, dz), "fj_eval_potential");
}

double Potential::operator()(double R, double z) const {
    double phi;
    eval(1, &R, &z, &phi, nullptr, nullptr);
    return phi;
}

double Potential::dR(double R, double z) const {
    double phi, a, b;
    eval(1, &R, &z, &phi, &a, &b);
    return a;
}

double Potential::dz(double R, double z) const {
    double phi, a, b;
    eval(1, &R, &z, &phi, &a, &b);
    return b;
}

void syncPotential() { check(fj_set_potential(ctx(), phil[0], Pr[0], Pr2[0]), "fj_set_potential"); }

void updatePhil(const Potential *p1, const Potential *p2) {
    for (int n = 0; n < NR; n++)
        for (int np = 0; np < NPOLY; np++) {
            phil[n][np] = p1->philP[n][np] + p2->philP[n][np];
            Pr[n][np] = p1->PrP[n][np] + p2->PrP[n][np];
            Pr2[n][np] = p1->Pr2P[n][np] + p2->Pr2P[n][np];
        }
    syncPotential();
}

// setDF (df/DF.cpp:36-63) and SetModel's J0 = sqrt(mass r0), mass, chi.
void setDF(fJParams *f, Potential *p, unsigned comp) {
    (void)p;
    const bool c1 = comp == 1;
    const std::string &name = c1 ? f->modName : f->modName2;
    double &A = c1 ? f->alpha : f->alpha_2, &B = c1 ? f->beta : f->beta_2;
    if (name == "Hernquist") { A = 1.667; B = 5.; }
    else if (name == "Isochrone") { A = 0.0; B = 5.; }
    else if (name == "NFW") { A = 1.667; B = 3.; }
    else if (!(A != UNSET && B != UNSET))
        throw Error(FJ_E_ARG, c1 ? "either set parameter 'model' or 'A' and 'B'"
                                 : "either set parameter '2:model' or '2:A' and '2:B'");
    fj_df_params d;
    d.A = A; d.B = B;
    d.mass = c1 ? f->mass : f->mass_2;
    d.J0 = sqrt(d.mass * (c1 ? f->r0 : f->r0_2));
    d.chi = c1 ? f->chi : f->chi_2;
    d.dphi_h = c1 ? f->dphi_h_in : f->dphi_h_in2;
    d.dz_h = c1 ? f->dz_h_in : f->dz_h_in2;
    d.dphi_g = c1 ? f->dphi_g_in : f->dphi_g_in2;
    d.dz_g = c1 ? f->dz_g_in : f->dz_g_in2;
    d.rot_form = c1 ? f->rot_form : f->rot_form_2;
    const double k = c1 ? f->rot_k : f->rot_k_2;
    d.k = k != UNSET ? k : (d.rot_form == 1 ? 0.5 : 1.0);
    check(fj_set_df(ctx(), &d), "fj_set_df");
    printf("===========  DF Set as: ==============\n a=%f b=%f \n", A, B);
}

void computeNewPhi(Potential *p, double **rhlH, double **sigRlH, double **sigplH, double **sigzlH,
                   double **sigRzlH, double **vrotlH) {
    rhlH = or_global(rhlH, rhl);
    sigRlH = or_global(sigRlH, sigRl);
    sigplH = or_global(sigplH, sigpl);
    sigzlH = or_global(sigzlH, sigzl);
    sigRzlH = or_global(sigRzlH, sigRzl);
    vrotlH = or_global(vrotlH, vrotl);
    check(fj_compute_new_phi(ctx(), g_poisson, rhlH[0], sigRlH[0], sigplH[0], sigzlH[0], sigRzlH[0],
                             vrotlH[0], p->philP[0], p->PrP[0], p->Pr2P[0]),
          "fj_compute_new_phi");
    if (p->comp == 1) {
        const size_t b = sizeof(double) * NR * NPOLY;
        memcpy(phil[0], p->philP[0], b);
        memcpy(Pr[0], p->PrP[0], b);
        memcpy(Pr2[0], p->Pr2P[0], b);
        syncPotential();
    }
}

double rhofDF(double R, double z, Potential *p, double *vrot, double *sigR, double *sigp, double *sigz,
              double *sigRz) {
    (void)p;
    double rho;
    rhofDF(1, &R, &z, &rho, vrot, sigR, sigp, sigz, sigRz);
    return rho;
}

void rhofDF(int n, const double *R, const double *z, double *rho, double *vrot, double *sigR, double *sigp,
            double *sigz, double *sigRz) {
    check(fj_rhofDF(ctx(), n, R, z, nullptr, rho, vrot, sigR, sigp, sigz, sigRz), "fj_rhofDF");
}

void tabulateDelta(Potential *p) {
    (void)p;
    check(fj_tabulate_delta(ctx(), Egrid.data(), Dgrid.data()), "fj_tabulate_delta");
}

// Deltafn (delta/tabulate.cpp:40-44) with linterp (include/Utils.h:149-155):
// Delta = sqrt(max(1e-6, Delta^2(E))), Delta^2 clamped at the table ends.
double Deltafn(double E) {
    const int n = NGRID;
    if (n <= 0) throw Error(FJ_E_STATE, "no Delta table");
    double D2;
    if (E <= Egrid[0]) D2 = Dgrid[0];
    else if (E < Egrid[n - 1]) {
        int top = n - 1, bot = 0;
        while (top - bot > 1) {
            const int m = (top + bot) >> 1;
            if (E >= Egrid[m]) bot = m; else top = m;
        }
        D2 = Dgrid[bot] + (E - Egrid[bot]) / (Egrid[top] - Egrid[bot]) * (Dgrid[top] - Dgrid[bot]);
    } else D2 = Dgrid[n - 1];
    return sqrt(D2 > .000001 ? D2 : .000001);
}

void actions(int n, const double *R, const double *z, const double *vR, const double *vphi, const double *vz,
             const double *Delta, double *Jr, double *Jz) {
    std::vector<double> D;
    if (!Delta) { D.assign(n, 0.0); Delta = D.data(); }
    check(fj_actions(ctx(), n, R, z, vR, vphi, vz, Delta, Jr, Jz), "fj_actions");
}

std::vector<double> vir2(Potential *p, double **rhlH, double **vrotlH, double **sigRlH, double **sigplH,
                         double **sigzlH, bool print) {
    (void)p;
    std::vector<double> o(11);
    check(fj_virial(ctx(), or_global(rhlH, rhl)[0], or_global(vrotlH, vrotl)[0], or_global(sigRlH, sigRl)[0],
                    or_global(sigplH, sigpl)[0], or_global(sigzlH, sigzl)[0], o.data()),
          "fj_virial");
    if (print) {   // libs/Stats.cpp:37-40
        printf("Mass(%3.0f): %f %f\n", ar[NR - 1], o[0], o[1]);
        printf("KE, PE, W/K = %f %f %f\n", o[2], o[3], o[4]);
        printf("Kxx, Wxx, Wxx/Kxx = %f %f %f\n", o[5], o[6], o[7]);
        printf("Kzz, Wzz, Wzz/Kzz = %f %f %f\n", o[8], o[9], o[10]);
    }
    return o;
}

fj_stats nodeStats() {
    fj_stats st;
    check(fj_get_stats(ctx(), &st), "fj_get_stats");
    return st;
}

std::vector<double> vir2Total(const std::vector<double> &a, const std::vector<double> &b, bool print) {
    if (a.size() != 11 || b.size() != 11) throw Error(FJ_E_ARG, "vir2Total: two outputs of vir2");
    std::vector<double> o(11);
    o[0] = a[0];                 // r_max^2 dPhi_0/dr of the total potential (the same in both)
    for (int i : {1, 2, 3, 5, 6, 8, 9}) o[i] = a[i] + b[i];
    o[4] = o[3] / o[2];
    o[7] = o[6] / o[5];
    o[10] = o[9] / o[8];
    if (print) {
        printf("Mass(%3.0f): %f %f\n", ar[NR - 1], o[0], o[1]);
        printf("KE, PE, W/K = %f %f %f\n", o[2], o[3], o[4]);
        printf("Kxx, Wxx, Wxx/Kxx = %f %f %f\n", o[5], o[6], o[7]);
        printf("Kzz, Wzz, Wzz/Kzz = %f %f %f\n", o[8], o[9], o[10]);
    }
    return o;
}

double potentialDistance(double **o, double **n) {
    double maxd = 0;
    for (int i = 0; i < NR; i++)
        for (int j = 0; j < NPOLY; j++) {
            const double d = fabs((n[i][j] - o[i][j]) / n[i][0]);
            maxd = maxd > d ? maxd : d;
        }
    return maxd;
}

void savePhi(double **philOLD, double **PrOLD, double **Pr2OLD) {
    for (int i = 0; i < NR; i++)
        for (int j = 0; j < NPOLY; j++) {
            philOLD[i][j] = phil[i][j];
            PrOLD[i][j] = Pr[i][j];
            Pr2OLD[i][j] = Pr2[i][j];
        }
}

double mergePhi(double **philOLD, double **PrOLD, double **Pr2OLD, double f) {
    const double fa = 1 + f;
    double maxd = 0;
    for (int i = 0; i < NR; i++)
        for (int j = 0; j < NPOLY; j++) {
            const double d = fabs((phil[i][j] - philOLD[i][j]) / phil[i][0]);
            maxd = maxd > d ? maxd : d;
            phil[i][j] = fa * phil[i][j] - f * philOLD[i][j];
            Pr[i][j] = fa * Pr[i][j] - f * PrOLD[i][j];
            Pr2[i][j] = fa * Pr2[i][j] - f * Pr2OLD[i][j];
        }
    printf("-- Merge Phi: max distance %f \n", maxd);
    syncPotential();
    return maxd;
}

// ---------------------------------------------------------------------------
// Output files
// ---------------------------------------------------------------------------
std::string composeName(const fJParams &f, int iter, unsigned comp) {
    std::string name;
    if (comp == 1) {
        name = trim_name(f.modName);
        name += "_" + toString(f.dphi_h_in) + "_" + toString(f.dz_h_in) + "_" + toString(f.dphi_g_in) + "_" +
                toString(f.dz_g_in);
    } else {
        name = trim_name(f.modName2);
        name += "_" + toString(f.dphi_h_in2) + "_" + toString(f.dz_h_in2) + "_" + toString(f.dphi_g_in2) + "_" +
                toString(f.dz_g_in2);
    }
    name += "_" + toString(iter) + ".out";
    if (comp == 2) name += "2c";
    return name;
}

namespace {

void write_body(std::ostream &o, const fJParams &f, unsigned comp, double **M[9]) {
    const bool c1 = comp == 1;
    o << NR << " " << NPOLY << " " << NGAUSS << std::endl;
    o << (c1 ? f.dphi_h_in : f.dphi_h_in2) << " " << (c1 ? f.dz_h_in : f.dz_h_in2) << " "
      << (c1 ? f.dphi_g_in : f.dphi_g_in2) << " " << (c1 ? f.dz_g_in : f.dz_g_in2) << std::endl;
    o << (c1 ? f.chi : f.chi_2) << " " << (c1 ? f.mass : f.mass_2) << " " << (c1 ? f.r0 : f.r0_2) << std::endl;
    o << (c1 ? f.alpha : f.alpha_2) << " " << (c1 ? f.beta : f.beta_2) << std::endl;
    for (int i = 0; i < NR; i++) o << ar[i] << std::endl;
    for (int m = 0; m < 9; m++)
        for (int i = 0; i < NR; i++) {
            for (int j = 0; j < NPOLY; j++) o << M[m][i][j] << " ";
            o << std::endl;
        }
}

}  // namespace

std::string writeOut(const fJParams &f, int iter, unsigned comp, double **rhlH, double **sigRlH, double **sigplH,
                     double **sigzlH, double **sigRzlH, double **vrotlH, double **philH, double **PrH,
                     double **Pr2H, const std::string &dir, bool full) {
    // order of the matrices in the file: writeOut.cpp:94-104
    double **M[9] = {or_global(rhlH, rhl),     or_global(vrotlH, vrotl), or_global(sigRlH, sigRl),
                     or_global(sigplH, sigpl), or_global(sigzlH, sigzl), or_global(sigRzlH, sigRzl),
                     or_global(philH, phil),   or_global(PrH, Pr),       or_global(Pr2H, Pr2)};
    const std::string &user = comp == 1 ? f.outName : f.outName2;
    const std::string file = user != "null" ? user + "_" + toString(iter) + ".out" : composeName(f, iter, comp);
    const std::string path = dir + "/" + file;
    std::ofstream o(path.c_str(), std::ios::out);
    if (!o.is_open()) throw Error(FJ_E_ARG, "cannot open " + path + " (does the directory exist?)");
    write_body(o, f, comp, M);
    o.close();
    if (full) {
        std::ofstream s((path + ".f64").c_str(), std::ios::out);
        if (!s.is_open()) throw Error(FJ_E_ARG, "cannot open " + path + ".f64");
        s << std::setprecision(17);
        write_body(s, f, comp, M);
    }
    return path;
}

double SigmaDF(double R, double z) {
    double s;
    check(fj_sigma_df(ctx(), rhl[0], 1, &R, &z, &s), "fj_sigma_df");
    return s;
}

void line_profile(double R, double z, int n, const double *vx, double *out) {
    check(fj_line_profile(ctx(), rhl[0], R, z, n, vx, out), "fj_line_profile");
}

double line_profile(double R, double z, double vx, Potential *p) {
    (void)p;
    double o;
    line_profile(R, z, 1, &vx, &o);
    return o;
}

static double vdist(int which, double R, double z, double v) {
    double o;
    check(fj_velocity_distribution(ctx(), which, R, z, 1, &v, &o), "fj_velocity_distribution");
    return o;
}
double P_vp(double R, double z, double vphi, Potential *p) { (void)p; return vdist(0, R, z, vphi); }
double P_vz(double R, double z, double vz, Potential *p) { (void)p; return vdist(1, R, z, vz); }
double P_vr(double R, double z, double vr, Potential *p) { (void)p; return vdist(2, R, z, vr); }

std::string writeLineProfiles(const fJParams &f, int iter, unsigned comp, Potential *p, const std::string &dir) {
    (void)p;
    const std::string &user = comp == 1 ? f.outName : f.outName2;
    std::string name = user != "null" ? user + "_" + toString(iter) + ".out" : composeName(f, iter, comp);
    name.erase(name.end() - 3, name.end());
    name += "lp";
    const std::string path = dir + "/" + name;
    std::ofstream o(path.c_str(), std::ios::out);
    if (!o.is_open()) throw Error(FJ_E_ARG, "cannot open " + path);
    const int size_lp = 150, half = size_lp / 2;
    const int rad[3] = {0, NR * 2 / 3, NR * 5 / 6};
    const double far = 10 * ar[NR - 1];
    double Rs[4] = {ar[rad[0]], ar[rad[1]], ar[rad[2]], far}, zs[4] = {ar[0], ar[0], ar[0], far}, ph[4];
    check(fj_eval_potential(ctx(), 4, Rs, zs, ph, nullptr, nullptr), "fj_eval_potential");
    double ve[3];
    for (int k = 0; k < 3; k++) ve[k] = sqrt(-2 * (ph[k] - ph[3]));
    std::vector<double> v(size_lp), vx(size_lp), lp[3];
    for (int i = 0; i < half; i++) v[i] = -pow(10., -3. + (half - 1 - i) * (3.) / (half - 1));
    for (int i = half; i < size_lp; i++) v[i] = pow(10., -3. + (i - half) * (3.) / (half - 1));
    for (int i = 0; i < size_lp; i++) vx[i] = 3. * v[i];
    for (int k = 0; k < 3; k++) {
        lp[k].resize(size_lp);
        line_profile(ar[rad[k]], ar[0], size_lp, vx.data(), lp[k].data());
    }
    for (int i = 0; i < size_lp; i++)
        o << ve[0] * 3. * v[i] << " " << lp[0][i] << " " << ve[1] * 3. * v[i] << " " << lp[1][i] << " "
          << ve[2] * 3. * v[i] << " " << lp[2][i] << std::endl;
    return path;
}

OutFile readOut(const std::string &path) {
    std::ifstream in(path.c_str());
    if (!in.is_open()) throw Error(FJ_E_ARG, "cannot open " + path);
    OutFile o;
    in >> o.nr >> o.npoly >> o.ngauss;
    in >> o.dphi_h >> o.dz_h >> o.dphi_g >> o.dz_g >> o.chi >> o.mass >> o.r0 >> o.alpha >> o.beta;
    if (!in || o.nr <= 0 || o.npoly <= 0) throw Error(FJ_E_ARG, "bad header in " + path);
    o.ar.resize(o.nr);
    for (double &x : o.ar) in >> x;
    for (auto &m : o.m) {
        m.resize((size_t)o.nr * o.npoly);
        for (double &x : m) in >> x;
    }
    if (!in) throw Error(FJ_E_ARG, "truncated file " + path);
    return o;
}

void loadPotential(Potential *p, const OutFile &o) {
    if (o.nr != NR || o.npoly != NPOLY) throw Error(FJ_E_ARG, "the file's grid differs from the current grid");
    for (int i = 0; i < NR; i++)
        if (fabs(o.ar[i] - ar[i]) > 1e-5 * ar[i]) throw Error(FJ_E_ARG, "the file's radii differ from the current grid");
    const size_t b = sizeof(double) * NR * NPOLY;
    memcpy(p->philP[0], o.m[6].data(), b);
    memcpy(p->PrP[0], o.m[7].data(), b);
    memcpy(p->Pr2P[0], o.m[8].data(), b);
    if (p->comp == 1) {
        memcpy(phil[0], o.m[6].data(), b);
        memcpy(Pr[0], o.m[7].data(), b);
        memcpy(Pr2[0], o.m[8].data(), b);
        syncPotential();
    }
}

}  // namespace fjm
