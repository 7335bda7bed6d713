// fjm.h -- C++ facade of the fJmodels SCF loop over the gfx950 C ABI
// (include/fjmodels.h).
//
// It keeps the reference's call surface so that a driver written against it
// reads like fJmodels.cpp: readParam / printParam (io/readParam.cpp), SetGrid
// (include/Utils.h:78-83), class Potential (pot/Potential.h:15-49),
// computeNewPhi / updatePhil (pot/potLeg.cpp:21-82), setDF (df/DF.cpp:36-63),
// tabulateDelta / Deltafn (delta/tabulate.cpp:13-44), rhofDF
// (integ/Integ.cpp:150-163), writeOut (io/writeOut.cpp:22-147) and vir2
// (libs/Stats.cpp:12-41).  Differences, all deliberate:
//   * grid sizes are run-time values (the reference's NR, NGAUSS, NPOLY,
//     NGRID are #defines, include/Grid.h:11-14);
//   * the state the reference keeps in globals (include/Grid.h:19-22,
//     df/DF.cpp:18-20, pot/models.cpp:13-14) lives in namespace fjm and in
//     one device context;
//   * errors throw fjm::Error instead of printf + exit();
//   * matrices are mat<double>(rows, cols): row pointers into one contiguous
//     row-major block, so they pass to the C ABI without copies.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fjmodels.h"

namespace fjm {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// mat<T>(rows, cols) (include/Utils.h:44-48): row pointers over one block.
double **mat(int rows, int cols);
void free_mat(double **m);

// ---------------------------------------------------------------------------
// Parameters: io/readParam.cpp, include/readParam.h
// ---------------------------------------------------------------------------
constexpr double UNSET = -999.;

struct fJParams {
    int itermax = 5;
    std::string modName = "null", outName = "null";
    double dphi_h_in = UNSET, dz_h_in = UNSET, dphi_g_in = UNSET, dz_g_in = UNSET;
    double chi = UNSET, mass = UNSET, r0 = UNSET, q = UNSET;
    double alpha = UNSET, beta = UNSET;
    std::string modName2 = "null", outName2 = "null";
    double dphi_h_in2 = UNSET, dz_h_in2 = UNSET, dphi_g_in2 = UNSET, dz_g_in2 = UNSET;
    double chi_2 = UNSET, mass_2 = UNSET, r0_2 = UNSET, q_2 = UNSET;
    double alpha_2 = UNSET, beta_2 = UNSET;
    // Keys this build adds (the reference has none of them):
    //   rot_form 0|1 and rot_k: rotation term of df/DF.cpp:116-117
    //   nr, rmin, rmax, ntheta, npoly, ngrid: the grid (Grid.h:11-14 defines)
    //   quad log22|tanhsinh22: velocity rule (include/GaussPts.h:12)
    //   tol: stop when the potential changes by less than tol
    //        (mergePhi's distance, include/UtilsLeg.h:266-278); 0 = run itermax
    //   poisson quadpack|blockscan
    //   precision fp64|mixed|fast: node phase in FP64 (default), with FP32
    //        action integrands and DF, or the fast node phase with
    //        Gauss-Legendre action integrals (fj_set_precision)
    //   merge f: after each computeNewPhi, Phi <- (1+f) Phi - f Phi_old
    //        (mergePhi, include/UtilsLeg.h:266-278; one component); 0 = off
    //   delta_refresh 1: tabulate Delta again from each new potential (the
    //        reference tabulates it once, from the guess: fJmodels.cpp:127)
    //   lineprofile 1: write the .lp file of writeOut.cpp:159-186 (the
    //        reference's LINEPROFILE build) after each iteration
    //   comp2_potential updated|total: the potential component 2's moments
    //        are computed in.  updated (default, the reference's flow):
    //        component 1's computeNewPhi installs component 1's new
    //        potential as the global one (Potential.cpp:182-188), and
    //        component 2 reads that (fJmodels.cpp:69-73).  total: both
    //        components read the total potential of the previous iteration.
    int rot_form = 0, rot_form_2 = 0;
    double rot_k = UNSET, rot_k_2 = UNSET;
    int nr = 60, ntheta = 30, npoly = 22, ngrid = 30;
    double rmin = UNSET, rmax = 200.;
    int quad = FJ_QUAD_LOG22;
    int poisson = FJ_POISSON_QUADPACK;
    int precision = FJ_PREC_FP64;
    double tol = 0;
    double merge = 0;
    int delta_refresh = 0;
    int lineprofile = 0;
    int comp2_potential = 0;   // 0 updated, 1 total
};

// Number of components found by checkParameterFile (readParam.cpp:125-150).
extern unsigned components;

fJParams readParam(const char *fName = "param.txt");
void checkParameterFile(const fJParams *fJP);
void printParam(const fJParams *fJP);

// ---------------------------------------------------------------------------
// Grid and global tables (include/Grid.h:19-22)
// ---------------------------------------------------------------------------
extern int NR, NGAUSS, NPOLY, NGRID;
extern std::vector<double> ar, Egrid, Dgrid;
extern double **phil, **Pr, **Pr2, **rhl, **vrotl, **sigRl, **sigpl, **sigzl, **sigRzl;

// SetGrid(rmax) (include/Utils.h:78-83): nr log-spaced radii in
// [5e-6 rmax, rmax], or in [rmin, rmax] when rmin > 0.  Allocates the global
// tables; the device context (HIP device `device`) is created, and given the
// grid, at the first call that needs the GPU.
void SetGrid(double rmax, int nr, int ntheta, int npoly, int ngrid, double rmin = 0, int device = 0);
void SetGridRadii(const std::vector<double> &radii, int ntheta, int npoly, int ngrid, int device = 0);

// The device context behind the facade.
fj_ctx *context();
void set_quad_rule(int preset);
// Poisson solver of computePhil / computeNewPhi (FJ_POISSON_QUADPACK default).
void set_poisson_mode(int mode);
// Arithmetic of the node phase (FJ_PREC_FP64 default, FJ_PREC_MIXED).
void set_precision(int mode);

// SetModel (pot/models.cpp:16-25): mass, r0, q of component comp.
void SetModel(const fJParams &fJP, unsigned comp = 1);

// ---------------------------------------------------------------------------
// Potential (pot/Potential.h:15-49)
// ---------------------------------------------------------------------------
class Potential {
  public:
    explicit Potential(unsigned comp_in = 1);
    ~Potential();
    Potential(const Potential &) = delete;
    Potential &operator=(const Potential &) = delete;

    unsigned comp;
    double **philP, **PrP, **Pr2P, **rhlP;

    // selectGuessRho (Potential.cpp:44-57): Isochrone, Hernquist, NFW, NFWext
    void selectGuessRho(const std::string &rhoName);
    // computeGuessRhl (Potential.cpp:63-80) into rhlP, using SetModel's mass, r0, q
    void computeGuessRhl();
    // computePhil (Potential.cpp:163-200): philP/PrP/Pr2P from rhlH; the
    // global phil/Pr/Pr2 too when comp == 1
    void computePhil(double **rhlH = nullptr);
    // Phi, dPhi/dR, dPhi/dz of the total potential (Potential.cpp:206-276)
    double operator()(double R, double z) const;
    double dR(double R, double z) const;
    double dz(double R, double z) const;
    // batched form of the three above; dR/dz may be null
    void eval(int n, const double *R, const double *z, double *phi, double *dR, double *dz) const;

  private:
    int guess_ = -1;
};

// Loads the global phil/Pr/Pr2 into the device context.
void syncPotential();

// updatePhil (potLeg.cpp:21-29): global tables = p1 + p2.
void updatePhil(const Potential *p1, const Potential *p2);

// setDF (df/DF.cpp:36-63) for component comp.  Component 2's (A, B) fall
// back to its own 2:A/2:B (the reference tests component 1's, DF.cpp:54).
void setDF(fJParams *fJP, Potential *p, unsigned comp);

// computeNewPhi (potLeg.cpp:76-82): moments at every grid point, rho[0][0]
// extrapolation, Legendre projection, Poisson solve into p->philP/PrP/Pr2P
// (and the global tables when p->comp == 1).  Defaults are the globals.
void computeNewPhi(Potential *p, double **rhlH = nullptr, double **sigRlH = nullptr, double **sigplH = nullptr,
                   double **sigzlH = nullptr, double **sigRzlH = nullptr, double **vrotlH = nullptr);

// rhofDF (Integ.cpp:150-163) at one point, and at n points.
double rhofDF(double R, double z, Potential *p, double *vrot, double *sigR, double *sigp, double *sigz,
              double *sigRz);
void rhofDF(int n, const double *R, const double *z, double *rho, double *vrot, double *sigR, double *sigp,
            double *sigz, double *sigRz);

// tabulateDelta (delta/tabulate.cpp:13-38) on the current total potential;
// fills Egrid/Dgrid.  Deltafn (tabulate.cpp:40-44) reads them.
void tabulateDelta(Potential *p);
double Deltafn(double E);

// Actions (uvOrb(...).Ju(), .Jv(), uvOrb.cpp:15-262) at n phase-space points;
// Delta[i] <= 0 (or Delta null) uses Deltafn(H).
void actions(int n, const double *R, const double *z, const double *vR, const double *vphi, const double *vz,
             const double *Delta, double *Jr, double *Jz);

// vir2 (libs/Stats.cpp:12-41): prints as the reference and returns the 11
// numbers of fj_virial.
std::vector<double> vir2(Potential *p, double **rhlH = nullptr, double **vrotlH = nullptr,
                         double **sigRlH = nullptr, double **sigplH = nullptr, double **sigzlH = nullptr,
                         bool print = true);

// Statistics of the last node phase (fj_get_stats): timings, node counts,
// Kronrod rules, Brent iterations, potential / gradient evaluations.
fj_stats nodeStats();

// Virial sums of a two-component model: the additive outputs of vir2 for
// each component, both evaluated in the total potential (the reference has no
// two-component virial), added; the ratios formed from the sums; Mass(r_max)
// of the total potential.  Prints in vir2's format when print is set.
std::vector<double> vir2Total(const std::vector<double> &c1, const std::vector<double> &c2, bool print = true);

// mergePhi's distance (include/UtilsLeg.h:266-278): max over (n, np) of
// |new - old| / |new[n][0]|.
double potentialDistance(double **philOld, double **philNew);

// ---------------------------------------------------------------------------
// Observables: integ/Integ.cpp (the reference's quadratures, summed in the
// serial order of their loops; f(J) on the GPU)
// ---------------------------------------------------------------------------
// SigmaDF (Integ.cpp:168-172) from the global rhl.
double SigmaDF(double R, double z);
// line_profile (Integ.cpp:177-190); the second form takes n velocities.
double line_profile(double R, double z, double vx, Potential *p);
void line_profile(double R, double z, int n, const double *vx, double *out);
// P_vp, P_vz, P_vr (Integ.cpp:191-220).
double P_vp(double R, double z, double vphi, Potential *p);
double P_vz(double R, double z, double vz, Potential *p);
double P_vr(double R, double z, double vr, Potential *p);
// The .lp file of writeOut.cpp:159-186 (LINEPROFILE): 150 line-of-sight
// velocities, three radii (ar[0], ar[2 nr/3], ar[5 nr/6] at z = ar[0]), next
// to the .out file of (iter, comp).  Returns its path.
std::string writeLineProfiles(const fJParams &fJP, int iter, unsigned comp, Potential *p,
                              const std::string &dir = "models");

// savePhi (include/UtilsLeg.h:256-263): copies of the global phil, Pr, Pr2.
void savePhi(double **philOLD, double **PrOLD, double **Pr2OLD);
// mergePhi (include/UtilsLeg.h:266-278): the distance of the global phil
// from philOLD, then phil, Pr, Pr2 <- (1+f) new - f old; prints the
// distance as the reference does and installs the merged potential.
double mergePhi(double **philOLD, double **PrOLD, double **Pr2OLD, double f);

// Installs the potential tables of a file written by writeOut (readOut) as
// component p's tables and, for component 1, as the global potential: the
// state after the iteration that wrote the file, for resuming a run.
void loadPotential(Potential *p, const struct OutFile &o);

// ---------------------------------------------------------------------------
// Output: io/writeOut.cpp
// ---------------------------------------------------------------------------
// composeName (writeOut.cpp:22-56)
std::string composeName(const fJParams &fJP, int iter, unsigned comp = 1);
// writeOut (writeOut.cpp:74-138): models/<name>, byte-compatible (ostream
// default 6 significant digits).  With full = true a sidecar <name>.f64 in
// the same layout with 17 significant digits is written too.  Returns the
// path of the .out file.
std::string writeOut(const fJParams &fJP, int iter, unsigned comp = 1, double **rhlH = nullptr,
                     double **sigRlH = nullptr, double **sigplH = nullptr, double **sigzlH = nullptr,
                     double **sigRzlH = nullptr, double **vrotlH = nullptr, double **philH = nullptr,
                     double **PrH = nullptr, double **Pr2H = nullptr, const std::string &dir = "models",
                     bool full = true);

// Reads a file written by writeOut (.out or its .f64 sidecar): header values
// and the nine matrices (rhl, vrotl, sigRl, sigpl, sigzl, sigRzl, phil, Pr,
// Pr2), each [nr][npoly] row-major.
struct OutFile {
    int nr = 0, npoly = 0, ngauss = 0;
    double dphi_h = 0, dz_h = 0, dphi_g = 0, dz_g = 0, chi = 0, mass = 0, r0 = 0, alpha = 0, beta = 0;
    std::vector<double> ar;
    std::vector<double> m[9];
};
OutFile readOut(const std::string &path);

}  // namespace fjm
