/*
 * fjmodels.h -- C ABI of the MI355X (gfx950) implementation of the fJmodels
 * self-consistent-field iteration.
 *
 * The reference has no plugin or FFI layer: its boundary is the C++
 * free-function / class surface called by the driver (fJmodels.cpp) and the
 * per-iteration models/<name>_<iter>.out files.  Each entry point below names the
 * reference interface it replaces.  All arrays are caller-owned, contiguous,
 * row-major doubles in host memory ([nr][npoly] for Legendre coefficient
 * tables, [nr][ntheta] for grid values); the library keeps its own device
 * copies.  Every function returns 0 on success and a nonzero FJ_E* code
 * otherwise (the reference prints and exit()s); fj_last_error() gives the text.
 * One context per host thread; calls on a context are stream-ordered.
 */
#ifndef FJMODELS_H
#define FJMODELS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fj_ctx fj_ctx;

enum {
    FJ_OK = 0,
    FJ_E_ARG = 1,        /* bad argument / sizes                          */
    FJ_E_HIP = 2,        /* HIP runtime error                             */
    FJ_E_STATE = 3,      /* call made before the data it needs was set    */
    FJ_E_NUMERIC = 4,    /* NaN in the Poisson external integral
                            (pot/Potential.cpp:144-146 exits with 1)      */
    FJ_E_DELTA = 5       /* Delta tabulation hit a reference exit/failure
                            path (delta/Delta.cpp:121,145-151)            */
};

/* Quadrature presets for the velocity integral (include/GaussPts.h:12-136). */
enum { FJ_QUAD_LOG22 = 0,        /* xpt11/xw11 on all axes: Int3D_111111_vec */
       FJ_QUAD_TANHSINH22 = 1 }; /* xpt01 x xpt11 x xpt01: Int3D_011101_vec  */

/* Guess densities (pot/models.cpp:40-68, Potential::selectGuessRho). */
enum { FJ_GUESS_HERNQUIST = 0, FJ_GUESS_ISOCHRONE = 1, FJ_GUESS_NFW = 2, FJ_GUESS_NFWEXT = 3 };

/* Poisson radial integrals: QUADPACK QAGS/QAGIU as the reference
 * (Potential.cpp:161-206), or exact piecewise-linear integration (block scan). */
enum { FJ_POISSON_QUADPACK = 0, FJ_POISSON_BLOCKSCAN = 1 };

/* Arithmetic of the node phase (DF evaluation and velocity quadrature).
 * FJ_PREC_MIXED: FP32 action integrands, Kronrod rules, QAGS bookkeeping
 * and DF (tasks that reach the FP32 interval cap rerun in FP64); FP64 Delta,
 * spheroidal coordinates, turning-point roots and the Poisson solve.
 * Results are not bitwise those of FJ_PREC_FP64. */
enum { FJ_PREC_FP64 = 0, FJ_PREC_MIXED = 1, FJ_PREC_FAST = 2 };
/* FJ_PREC_FAST: the same Staeckel-fudge actions computed for speed, one
 * lane per velocity node (24 B of records per node; runs above 64 M nodes
 * sum per lane instead): the u-side reference point at the node (Getu0's
 * umid cancels from uturnfn, uvOrb.cpp:25-31,70-73), the reference's
 * turning-point bracket walks (uvOrb.cpp:131-224) on FP32 evaluations with
 * the bracket ends and wrp_GSLroots' endpoint test in FP64, regula falsi
 * roots, and the action integrals (uvOrb.cpp:226-262) as FP32
 * Gauss-Legendre rules instead of QAGS: fj_set_fast_order points (64 by
 * default) for J_r, half as many (at least 4) for the smoother J_z
 * integrand.  The action surrogate does not run in this mode. */

/* Gauss-Legendre order of the FJ_PREC_FAST J_r integral (4..128); the J_z
 * rule takes half of it, at least 4 points. */
int fj_set_fast_order(fj_ctx *ctx, int n);

/* DF parameters: df/DF.cpp:18-20,36-63 (alpha[2], dphi_h, dz_h, dphi_g, dz_g)
 * and pot/models.cpp:13 (mass, J0, chi).  rot_form 0 is the code's
 * f(1 + k tanh(chi Jphi/J0)) (DF.cpp:117, k = 1); 1 is the documented
 * (1-k) f + k tanh(chi Jphi/J0) f (DF.cpp:116, README.md:80-83, k = 0.5). */
typedef struct {
    double A, B;
    double dphi_h, dz_h, dphi_g, dz_g;
    double mass, J0, chi;
    double k;
    int rot_form;
} fj_df_params;

/* Per-call statistics of the last node-phase call (fj_moments / fj_rhofDF /
 * fj_compute_new_phi) and of the last Poisson solve. */
typedef struct {
    double ms_points, ms_setup, ms_integrals, ms_overflow, ms_moments, ms_project, ms_poisson, ms_total;
    long long points, nodes, bound_nodes, tasks, qk_calls, overflow_tasks, brent_iters;
    int integral_lanes, setup_lanes;
    /* node diagnostics (the reference passes these cases on silently):
     * unbound_nodes: H < 0 but E >= 0 in spheroidal coordinates, Jr = Jz = 1e31 (uvOrb.cpp:38-40);
     * bracket_failures: a turning-point bracket walk (uvOrb.cpp:109-224) stopped by its iteration
     *   guard, where the reference would keep walking;
     * nonfinite_nodes: nodes whose f(J) is NaN or infinite (they poison the moment sums). */
    long long unbound_nodes, bracket_failures, nonfinite_nodes;
    /* node-work counters: potential evaluations Phi(R,z) (those with the gradient included) and
     * potential + gradient evaluations (Potential::operator(), dR, dz: Potential.cpp:206-276) */
    long long phi_evals, grad_evals;
} fj_stats;

/* Context on HIP device `device` (no reference counterpart: the reference
 * keeps this state in globals, include/Grid.h:19-24). */
int fj_ctx_create(int device, fj_ctx **out);
int fj_ctx_destroy(fj_ctx *ctx);
/* The stream the context's calls are ordered on: a hipStream_t of the caller
 * (of the HIP runtime this library is bound to, on the context's device), or
 * NULL for the context's own stream.  Work already queued is finished first;
 * the caller keeps ownership of its stream and must keep it alive while the
 * context uses it.  No counterpart in the reference, which has no device
 * streams: a context on a caller's stream takes the place of the per-run
 * globals of include/Grid.h inside a caller's own stream-ordered work. */
int fj_set_stream(fj_ctx *ctx, void *stream);
const char *fj_last_error(const fj_ctx *ctx);

/* Grid sizes and radii: replaces the compile-time NR/NGAUSS/NPOLY/NGRID of
 * include/Grid.h:11-14 and the ar[] fill of SetGrid (include/Utils.h:96-108).
 * The theta nodes are gauleg(0,1) with ntheta points (UtilsLeg.h:18-39). */
int fj_set_grid(fj_ctx *ctx, int nr, const double *ar, int ntheta, int npoly, int ngrid);

/* Velocity quadrature rule (include/GaussPts.h:12; Integ.cpp:157-158). */
int fj_set_quad_rule(fj_ctx *ctx, int preset);

/* Node-phase arithmetic (FJ_PREC_FP64 default). */
int fj_set_precision(fj_ctx *ctx, int mode);

/* Total potential multipole tables phil, Pr, Pr2 [nr][npoly] that every
 * potential evaluation reads (Grid.h:20 globals; Potential::computePhil
 * copies at Potential.cpp:182-188, updatePhil potLeg.cpp:21-29). */
int fj_set_potential(fj_ctx *ctx, const double *phil, const double *Pr, const double *Pr2);
int fj_get_potential(fj_ctx *ctx, double *phil, double *Pr, double *Pr2);

/* Delta^2(E) table Egrid/Dgrid [ngrid] read by Deltafn (delta/tabulate.cpp:40-44). */
int fj_set_delta_table(fj_ctx *ctx, const double *Egrid, const double *Dgrid);
int fj_get_delta_table(fj_ctx *ctx, double *Egrid, double *Dgrid);

/* setDF (df/DF.cpp:22-63): DF parameters of the component being iterated. */
int fj_set_df(fj_ctx *ctx, const fj_df_params *df);

/* Potential::computeGuessRhl (Potential.cpp:63-80) for the guess density
 * `model` with (mass, r0, q) (SetModel, models.cpp:16-25) -> rhl [nr][npoly]. */
int fj_guess_rhl(fj_ctx *ctx, int model, double mass, double r0, double q, double *rhl);

/* Potential::computePhil (Potential.cpp:110-200): rhl [nr][npoly] ->
 * philP, PrP, Pr2P [nr][npoly].  Does not change the context's potential. */
int fj_poisson(fj_ctx *ctx, int mode, const double *rhl, double *philP, double *PrP, double *Pr2P);

/* The grid loop of computeRhl (potLeg.cpp:54-58): rhofDF (8-argument,
 * Integ.cpp:150-163) at every (ar[n] sin(theta_g), ar[n] cos(theta_g)).
 * Outputs [nr][ntheta]; rho before the rho[0][0] extrapolation. */
int fj_moments(fj_ctx *ctx, double *rho, double *vrot, double *sigR, double *sigp,
               double *sigz, double *sigRz);

/* rhofDF (Integ.cpp:150-163) at n arbitrary points; out6 [n][6] are the six
 * raw quadrature sums (GaussQuad.cpp:160-180), rho/vrot/... [n]. */
int fj_rhofDF(fj_ctx *ctx, int n, const double *R, const double *z, double *out6,
              double *rho, double *vrot, double *sigR, double *sigp, double *sigz, double *sigRz);

/* fj_rhofDF without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory in the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every buffer beyond the caller's inputs and outputs (points, node records,
 * task lists, counters, QAGS workspaces) lives in one static device array of
 * the library, at most 64 MiB, which each GPU the library runs on holds until
 * the process ends; its address is taken once per device by fj_ctx_create.
 * n is limited to FJ_NOALLOC_MAX_POINTS (FJ_E_ARG beyond); the context's
 * potential, Delta table, DF, rule and precision apply, the action surrogate
 * must be off.  Results are those of fj_rhofDF.  All calls on one GPU share
 * the array: calls that overlap on one GPU (two contexts or threads at once)
 * race, so serialise them. */
#define FJ_NOALLOC_MAX_POINTS 4
int fj_rhofDF_noalloc(fj_ctx *ctx, int n, const double *R, const double *z, double *out6,
                      double *rho, double *vrot, double *sigR, double *sigp, double *sigz, double *sigRz);

/* The rho[0][0] fix (potLeg.cpp:68) and get_Ylm (UtilsLeg.h:200-253):
 * grid moments [nr][ntheta] -> Legendre coefficients [nr][npoly]. */
int fj_project_legendre(fj_ctx *ctx, const double *rho, const double *vrot, const double *sigR,
                        const double *sigp, const double *sigz, const double *sigRz,
                        double *rhl, double *vrotl, double *sigRl, double *sigpl,
                        double *sigzl, double *sigRzl);

/* computeNewPhi (potLeg.cpp:76-82): moments + projection + Poisson, device
 * resident.  Output tables [nr][npoly]; the context's potential is not
 * changed (the caller installs philP/PrP/Pr2P, or their sum over
 * components, with fj_set_potential as the reference driver does). */
int fj_compute_new_phi(fj_ctx *ctx, int poisson_mode, double *rhl, double *sigRl, double *sigpl,
                       double *sigzl, double *sigRzl, double *vrotl,
                       double *philP, double *PrP, double *Pr2P);

/* tabulateDelta (delta/tabulate.cpp:13-38) on the context's potential;
 * fills the context's table and copies it to Egrid/Dgrid [ngrid]. */
int fj_tabulate_delta(fj_ctx *ctx, double *Egrid, double *Dgrid);

/* vir2 (libs/Stats.cpp:12-41) with ev_dens (UtilsLeg.h:300-354):
 * out11 = {r_max^2 Pr[nr-1][0], mass sum, KE, PE, W/K, KRR, WRR, WRR/KRR,
 *          Kzz, Wzz, Wzz/Kzz}. */
int fj_virial(fj_ctx *ctx, const double *rhl, const double *vrotl, const double *sigRl,
              const double *sigpl, const double *sigzl, double *out11);

/* Potential::operator(), dR, dz (Potential.cpp:206-276) at n points;
 * dR/dz may be NULL. */
int fj_eval_potential(fj_ctx *ctx, int n, const double *R, const double *z,
                      double *phi, double *dR, double *dz);

/* uvOrb(Delta, Lz = R vphi, Phi(R,z), x, p, pot).Ju()/Jv() (uvOrb.cpp:15-262)
 * at n phase-space points; Delta[i] <= 0 selects Deltafn(H) as df() does
 * (DF.cpp:83-89).  Unbound points give 1e31 (uvOrb.cpp:38-40). */
int fj_actions(fj_ctx *ctx, int n, const double *R, const double *z, const double *vR,
               const double *vphi, const double *vz, const double *Delta, double *Jr, double *Jz);

/* df() (DF.cpp:78-124) at n phase-space points: 0 where H = v^2/2 + Phi(R,z)
 * >= 0, else the DF of fj_set_df at the Staeckel actions (Delta from the
 * table, as df() takes it). */
int fj_df_points(fj_ctx *ctx, int n, const double *R, const double *z, const double *vR,
                 const double *vphi, const double *vz, double *f);

/* Observables of integ/Integ.cpp, the reference's quadratures summed in the
 * serial order of their loops (GaussQuad.cpp:21-71, 124-134):
 *   fj_sigma_df   SigmaDF (Integ.cpp:168-172, SigmaInteg 33-43): 2 Int1D_01
 *                 of ev_dens (UtilsLeg.h:323-334, from rhl [nr][npoly]) along
 *                 r from R to ar[nr-1], at n points (R, z);
 *   fj_line_profile  line_profile (Integ.cpp:177-190, lineProfile 65-95) at n
 *                 line-of-sight velocities vx: Int3D_111111 / SigmaDF;
 *   fj_velocity_distribution  P_vp, P_vz, P_vr (Integ.cpp:191-220, Pvphi /
 *                 Pvz / Pvr 97-138; which = 0, 1, 2) at n values v of the
 *                 fixed component. */
int fj_sigma_df(fj_ctx *ctx, const double *rhl, int n, const double *R, const double *z, double *sigma);
int fj_line_profile(fj_ctx *ctx, const double *rhl, double R, double z, int n, const double *vx, double *out);
int fj_velocity_distribution(fj_ctx *ctx, int which, double R, double z, int n, const double *v, double *out);

int fj_get_stats(const fj_ctx *ctx, fj_stats *out);

/* ------------------------------------------------------------------------
 * Action surrogate.  No counterpart in the reference: it stands in for
 * uvOrb (uvOrb.cpp:15-262) at the quadrature nodes of intermediate SCF
 * iterations.  An MLP 5 -> 128 -> 128 -> 128 -> 2 (tanh) maps
 * (log R, z/R, vR/vc, vz/vc, |vphi|/vc), vc^2 = R dPhi/dR + z dPhi/dz at the
 * grid point, to (log(Jr/s + 1e-6), log((Jz + 1e-6 s) / (L - |Lz| + 1e-6 s)))
 * with s = r vc and L = |r x v|: Jz relative to its spherical-limit value;
 * Jphi = R vphi stays exact.  BF16 weights and activations on the matrix
 * cores with FP32 accumulation, trained with Adam on the GPU on actions the Staeckel fudge
 * computed at the nodes of earlier node phases.
 * ------------------------------------------------------------------------ */
typedef struct {
    long long n_train;     /* samples in the training set */
    long long nodes_sur;   /* bound nodes whose actions came from the network (last node phase) */
    long long fallback;    /* bound nodes sent to the Staeckel fudge: outside the training box,
                              |E| < 0.02 |Phi|, or predicted Jz below jz_min */
    long long judged;      /* random sample computed both ways */
    /* errors of the judged nodes' network actions, |J - J_St| / (Jr_St + Jz_St + |Jphi|): median and
     * 95th percentile over the judged nodes */
    double judge_err_r_med, judge_err_r_p95, judge_err_z_med, judge_err_z_p95;
    double ms_surrogate;   /* inputs + inference + decisions of the last node phase */
    double ms_train;       /* last fj_surrogate_train call */
    double last_loss;      /* mean squared error of the last epoch (normalised targets) */
    /* the same errors' median weighted by each judged node's share of the density sum
     * (quadrature weight x Ve^3 x f(J_St)) */
    double judge_werr_r_med, judge_werr_z_med;
} fj_surrogate_stats;

/* Fresh network (Glorot-uniform weights from seed) and a training set of
 * `capacity` samples (ring buffer). */
int fj_surrogate_init(fj_ctx *ctx, unsigned long long seed, long long capacity);
/* Staeckel node phases from now on add each bound node to the training set
 * with probability keep (0 stops); in surrogate mode the judged nodes are
 * added instead. */
int fj_surrogate_collect(fj_ctx *ctx, double keep);
/* `epochs` passes over the training set in batches (Adam, cosine-decayed
 * rate lr); the first call fixes the input/target normalisation and the
 * training box.  *loss: mean squared error of the last epoch. */
int fj_surrogate_train(fj_ctx *ctx, int epochs, int batch, double lr, double *loss);
/* on = 1: node phases take the actions from the network, with the Staeckel
 * fudge for the fallback nodes and for a random `judge_fraction` of nodes. */
int fj_surrogate_mode(fj_ctx *ctx, int on, double judge_fraction, double jz_min);
int fj_surrogate_get_stats(const fj_ctx *ctx, fj_surrogate_stats *out);
/* Network parameters as fp32 (W1[5][128], b1[128], W2[128][128], b2, W3, b3,
 * W4[128][2], b4[2]: 34,050 values); get and/or set may be NULL. */
int fj_surrogate_params(fj_ctx *ctx, float *get, const float *set);
/* Raw network on normalised inputs x [n][5] -> y [n][2], and the gradient of
 * the mean squared error against t [n][2] (n <= 1,048,576). */
int fj_surrogate_forward(fj_ctx *ctx, long long n, const float *x, float *y);
int fj_surrogate_gradient(fj_ctx *ctx, long long n, const float *x, const float *t, float *grad, double *loss);

/* ------------------------------------------------------------------------
 * Batched models.  No counterpart in the reference (one model per run of
 * fJmodels.cpp): many independent single-component models, each with its
 * own grid, guess potential, DF and Delta table, advanced together.  Every
 * phase of an SCF iteration is one launch over all models, the iteration is
 * captured in a HIP graph, and fj_batch_autotune picks the launch shapes.
 * The arithmetic follows the context's precision (fj_set_precision) and
 * velocity rule of each model; the Poisson step is the block scan.  A batch
 * holds at most 64 M velocity nodes, except one created under FJ_PREC_FAST:
 * it keeps one model's node records at a time plus a 4-byte sort key per
 * node (any number of models), runs its models one after another without a
 * HIP graph, and is iterated under FJ_PREC_FAST only (FJ_E_STATE otherwise),
 * as a QAGS batch is under FJ_PREC_FP64 / FJ_PREC_MIXED only.
 * ------------------------------------------------------------------------ */
typedef struct fj_batch fj_batch;

typedef struct {
    int nr;                /* radii ar[nr] (copied), strictly increasing, nr >= 11 */
    const double *ar;
    int ntheta, npoly;     /* cos(theta) Gauss-Legendre nodes; even orders l < 2 npoly */
    int guess;             /* FJ_GUESS_*, with mass, r0, q (SetModel, models.cpp:16-25) */
    double mass, r0, q;
    fj_df_params df;
    int quad;              /* FJ_QUAD_* */
} fj_model_spec;

typedef struct {
    double ms_setup, ms_tune, ms_last_iter;
    long long nodes, points;
    int setup_block, integ_blocks, graph;   /* chosen launch shapes; graph: iterations replay a HIP graph */
} fj_batch_stats;

int fj_batch_create(fj_ctx *ctx, int nmodels, const fj_model_spec *specs, fj_batch **out);
int fj_batch_destroy(fj_batch *b);
/* Guess potentials and Delta tables of every model. */
int fj_batch_setup(fj_batch *b);
/* Times the candidate launch shapes on one iteration (the models' state is
 * restored afterwards) and keeps the fastest. */
int fj_batch_autotune(fj_batch *b);
/* niter SCF iterations of every model (graph replays after the first). */
int fj_batch_iterate(fj_batch *b, int niter);
/* Model m's state: any pointer may be NULL; tables [nr][npoly], Egrid/Dgrid [30]. */
int fj_batch_get(fj_batch *b, int m, double *rhl, double *sigRl, double *sigpl, double *sigzl, double *vrotl,
                 double *phil, double *Pr, double *Pr2, double *Egrid, double *Dgrid);
/* Installs model m's potential tables [nr][npoly] and Delta table [30] (any
 * pointer may be NULL: that table is kept), as updatePhil (potLeg.cpp:21-29)
 * and tabulateDelta (delta/tabulate.cpp:13-44) leave them for the next
 * computeNewPhi: resuming a model, or starting it from given tables. */
int fj_batch_set(fj_batch *b, int m, const double *phil, const double *Pr, const double *Pr2, const double *Egrid,
                 const double *Dgrid);
int fj_batch_get_stats(const fj_batch *b, fj_batch_stats *out);

#ifdef __cplusplus
}
#endif
#endif
