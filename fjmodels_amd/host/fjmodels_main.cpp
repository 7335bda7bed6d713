// fjmodels -- self-consistent f(J) models on one MI355X.
//
// The driver of fJmodels.cpp:33-186 over the facade: read the parameter
// file, set the grid, build the guess potential, tabulate Delta, iterate
// computeNewPhi, write models/*.out, print the virial diagnostic and write
// one JSON record per iteration to <outdir>/fjmodels_metrics.jsonl.
//
//   fjmodels [param.txt] [--outdir DIR] [--device N]
//            [--resume FILE [--resume2 FILE2]] [--first-iter K]
//
// Besides the reference's keys the parameter file may set nr, rmin, rmax,
// ntheta, npoly, ngrid (grid), quad (log22 | tanhsinh22), rot_form / rot_k
// (rotation term), tol (stop when the potential changes by less), merge
// (mergePhi's over-relaxation factor, one component), delta_refresh
// (tabulate Delta from every new potential), lineprofile (write the
// .lp line-profile file of every iteration) and comp2_potential (the
// potential component 2's moments are computed in: updated, the reference's
// flow, or total).
//
// --resume continues a run from the potential in a file it wrote (the .f64
// sidecar reproduces the uninterrupted run bit for bit; FILE2: component 2),
// after the guess potential and Delta table are rebuilt as at the start;
// --first-iter K numbers the iterations from K (the run ends at itermax).
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <chrono>
#include <string>

#include "fjm.h"

using namespace fjm;

namespace {

struct Resume {
    std::string file1, file2;
    int first = 0;
};

double now() {
    using C = std::chrono::steady_clock;
    return std::chrono::duration<double>(C::now().time_since_epoch()).count();
}

void copy_mat(double **dst, double **src) { memcpy(dst[0], src[0], sizeof(double) * NR * NPOLY); }

// Per-iteration JSON record (one line per iteration in <outdir>/fjmodels_metrics.jsonl):
// wall time, grid points per second, the potential change, the virial sums
// (17 significant digits) and the node-phase counters of each component.
struct Metrics {
    FILE *f = nullptr;
    explicit Metrics(const std::string &dir) {
        const std::string path = dir + "/fjmodels_metrics.jsonl";
        f = fopen(path.c_str(), "w");
        if (!f) throw Error(FJ_E_ARG, "cannot write " + path);
    }
    ~Metrics() { if (f) fclose(f); }
    void write(int iter, double seconds, double change, const std::vector<double> &vir,
               const std::vector<fj_stats> &comp) {
        static const char *vn[11] = {"mass_r2Pr", "mass_sum", "KE", "PE", "W/K", "KRR", "WRR", "WRR/KRR",
                                     "Kzz", "Wzz", "Wzz/Kzz"};
        const double pts = (double)NR * NGAUSS * (double)comp.size();
        fprintf(f, "{\"iter\": %d, \"components\": %d, \"seconds\": %.6f, \"grid_pt_per_s\": %.3f, "
                   "\"potential_change\": %.17g, \"virial\": {", iter, (int)comp.size(), seconds,
                pts / seconds, change);
        for (int i = 0; i < 11 && i < (int)vir.size(); i++)
            fprintf(f, "%s\"%s\": %.17g", i ? ", " : "", vn[i], vir[i]);
        fprintf(f, "}, \"node_phase\": [");
        for (size_t k = 0; k < comp.size(); k++) {
            const fj_stats &s = comp[k];
            fprintf(f, "%s{\"ms_total\": %.3f, \"ms_setup\": %.3f, \"ms_integrals\": %.3f, "
                       "\"ms_overflow\": %.3f, \"ms_poisson\": %.3f, \"bound_nodes\": %lld, \"tasks\": %lld, "
                       "\"qk_calls\": %lld, \"overflow_tasks\": %lld, \"brent_iters\": %lld, "
                       "\"phi_evals\": %lld, \"grad_evals\": %lld, \"unbound_nodes\": %lld, "
                       "\"bracket_failures\": %lld, \"nonfinite_nodes\": %lld}",
                    k ? ", " : "", s.ms_total, s.ms_setup, s.ms_integrals, s.ms_overflow, s.ms_poisson,
                    s.bound_nodes, s.tasks, s.qk_calls, s.overflow_tasks, s.brent_iters, s.phi_evals,
                    s.grad_evals, s.unbound_nodes, s.bracket_failures, s.nonfinite_nodes);
        }
        fprintf(f, "]}\n");
        fflush(f);
    }
};

// twoComp (fJmodels.cpp:38-81).  In the reference's flow component 1's
// computeNewPhi ends in computePhil, which installs component 1's new
// potential as the global one (pot/Potential.cpp:182-188); component 2's
// moments are then computed in that potential, not in the total
// (fJmodels.cpp:69-73).  comp2_potential total keeps the previous total for
// component 2 instead.  After updatePhil the virial sums of both components
// in the total potential are printed (the reference prints none for two
// components).
void twoComp(fJParams P, const std::string &dir, const Resume &R) {
    printf("\nSetting up two-component f(J) model...\n");
    double **rhl2 = mat(NR, NPOLY), **vrotl2 = mat(NR, NPOLY), **sigRl2 = mat(NR, NPOLY),
           **sigpl2 = mat(NR, NPOLY), **sigzl2 = mat(NR, NPOLY), **sigRzl2 = mat(NR, NPOLY);
    SetModel(P, 1);
    Potential p(1);
    p.selectGuessRho(P.modName != "null" ? P.modName : "Hernquist");
    p.computeGuessRhl();
    p.computePhil(p.rhlP);
    SetModel(P, 2);
    Potential p2(2);
    p2.selectGuessRho(P.modName2 != "null" ? P.modName2 : "Hernquist");
    p2.computeGuessRhl();
    p2.computePhil(p2.rhlP);
    updatePhil(&p, &p2);
    tabulateDelta(&p);
    if (!R.file1.empty()) {
        if (R.file2.empty()) throw Error(FJ_E_ARG, "two components: --resume2 names component 2's file");
        loadPotential(&p, readOut(R.file1));
        loadPotential(&p2, readOut(R.file2));
        updatePhil(&p, &p2);
        if (P.delta_refresh) tabulateDelta(&p);
    }
    double **old = mat(NR, NPOLY), **PrTot = mat(NR, NPOLY), **Pr2Tot = mat(NR, NPOLY);
    Metrics met(dir);
    for (int k = R.first; k < P.itermax; k++) {
        printf("\n Iter:%d\n", k);
        const double t0 = now();
        copy_mat(old, phil);
        if (P.comp2_potential) { copy_mat(PrTot, Pr); copy_mat(Pr2Tot, Pr2); }
        setDF(&P, &p, 1);
        computeNewPhi(&p, rhl, sigRl, sigpl, sigzl, sigRzl, vrotl);
        std::vector<fj_stats> st{nodeStats()};
        if (P.comp2_potential) {   // back to the previous total for component 2
            copy_mat(phil, old);
            copy_mat(Pr, PrTot);
            copy_mat(Pr2, Pr2Tot);
            syncPotential();
        }
        setDF(&P, &p, 2);
        computeNewPhi(&p2, rhl2, sigRl2, sigpl2, sigzl2, sigRzl2, vrotl2);
        st.push_back(nodeStats());
        updatePhil(&p, &p2);
        const std::vector<double> vir = vir2Total(vir2(&p, rhl, vrotl, sigRl, sigpl, sigzl, false),
                                                  vir2(&p2, rhl2, vrotl2, sigRl2, sigpl2, sigzl2, false));
        if (P.delta_refresh) tabulateDelta(&p);
        writeOut(P, k, 1, rhl, sigRl, sigpl, sigzl, sigRzl, vrotl, p.philP, p.PrP, p.Pr2P, dir);
        writeOut(P, k, 2, rhl2, sigRl2, sigpl2, sigzl2, sigRzl2, vrotl2, p2.philP, p2.PrP, p2.Pr2P, dir);
        const double d = potentialDistance(old, phil);
        const double sec = now() - t0;
        printf("-- iteration %d: %.3f s, potential change %.3e\n", k, sec, d);
        met.write(k, sec, d, vir, st);
        if (P.tol > 0 && d < P.tol) {
            printf("-- converged (tolerance %g)\n", P.tol);
            break;
        }
    }
    double **M[] = {rhl2, vrotl2, sigRl2, sigpl2, sigzl2, sigRzl2, old, PrTot, Pr2Tot};
    for (double **m : M) free_mat(m);
}

// oneComp (fJmodels.cpp:86-144)
void oneComp(fJParams P, const std::string &dir, const Resume &R) {
    printf("\nSetting up one-component f(J) model...\n");
    SetModel(P);
    Potential p;
    p.selectGuessRho(P.modName != "null" ? P.modName : "Hernquist");
    p.computeGuessRhl();
    p.computePhil();
    tabulateDelta(&p);
    setDF(&P, &p, 1);
    if (!R.file1.empty()) {
        loadPotential(&p, readOut(R.file1));
        // the uninterrupted run tabulates Delta from the potential it just wrote
        if (P.delta_refresh) tabulateDelta(&p);
    }
    double **old = mat(NR, NPOLY), **PrOld = mat(NR, NPOLY), **Pr2Old = mat(NR, NPOLY);
    Metrics met(dir);
    for (int k = R.first; k < P.itermax; k++) {
        printf("\n Iter:%d\n", k);
        const double t0 = now();
        copy_mat(old, phil);
        if (P.merge != 0) savePhi(old, PrOld, Pr2Old);
        computeNewPhi(&p, rhl, sigRl, sigpl, sigzl, sigRzl, vrotl);
        const std::vector<fj_stats> st{nodeStats()};
        if (P.merge != 0) mergePhi(old, PrOld, Pr2Old, P.merge);
        writeOut(P, k, 1, rhl, sigRl, sigpl, sigzl, sigRzl, vrotl, phil, Pr, Pr2, dir);
        if (P.lineprofile) writeLineProfiles(P, k, 1, &p, dir);
        const std::vector<double> vir = vir2(&p);
        if (P.delta_refresh) tabulateDelta(&p);
        const double d = potentialDistance(old, phil);
        const double sec = now() - t0;
        printf("-- iteration %d: %.3f s, potential change %.3e\n", k, sec, d);
        met.write(k, sec, d, vir, st);
        if (P.tol > 0 && d < P.tol) {
            printf("-- converged (tolerance %g)\n", P.tol);
            break;
        }
    }
    free_mat(old);
    free_mat(PrOld);
    free_mat(Pr2Old);
}

}  // namespace

int main(int argc, char **argv) {
    const double start = now();
    const char *param = "param.txt";
    std::string dir = "models";
    int device = 0;
    Resume R;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--outdir") && i + 1 < argc) dir = argv[++i];
        else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--resume") && i + 1 < argc) R.file1 = argv[++i];
        else if (!strcmp(argv[i], "--resume2") && i + 1 < argc) R.file2 = argv[++i];
        else if (!strcmp(argv[i], "--first-iter") && i + 1 < argc) R.first = atoi(argv[++i]);
        else if (argv[i][0] != '-') param = argv[i];
        else {
            fprintf(stderr, "usage: %s [param.txt] [--outdir DIR] [--device N] [--resume FILE [--resume2 FILE2]] "
                            "[--first-iter K]\n", argv[0]);
            return 2;
        }
    }
    try {
        fJParams P = readParam(param);
        printParam(&P);
        SetGrid(P.rmax, P.nr, P.ntheta, P.npoly, P.ngrid, P.rmin > 0 ? P.rmin : 0, device);
        set_quad_rule(P.quad);
        set_poisson_mode(P.poisson);
        set_precision(P.precision);
        mkdir(dir.c_str(), 0755);
        printf("=\n======================================================================\n");
        if (components == 1) oneComp(P, dir, R);
        else twoComp(P, dir, R);
    } catch (const Error &e) {
        fprintf(stderr, "%s\n", e.what());
        return e.code ? e.code : 1;
    }
    printf("\n----> Elapsed time of computation: %7.5f s\n", now() - start);
    return 0;
}
