/*
 * fjoracle.c -- CPU restatement of the fJmodels SCF path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; it is the checker, never the product.
 *
 * Every function cites the reference file:line it restates.  Compile-time
 * sizes of the reference (include/Grid.h:11-14: NR, NGRID, NGAUSS, NPOLY;
 * include/GaussPts.h:12 QUADORD) are run-time fields of fjo_pot.
 * Pinning of each part is documented in tests/test_oracle_*.py.
 */
#include "fjoracle.h"

/* Elementary functions shared with the GPU path (see the header of that file:
 * one implementation on both sides makes the reference's discrete decisions
 * reproducible bit for bit).  Only tests and the checker use this library. */
#include "../fjmodels_amd/csrc/fj_math.h"
#include "../fjmodels_amd/csrc/fj_observe.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* include/Utils.h:22-33 */
#define EPSABSROOTS 1e-8
#define EPSRELROOTS 1e-6
#define EPSABSODE 1e-10
#define EPSRELODE 1e-9
#define EPSABSINTEG 1e-5
#define EPSRELINTEG 1e-3
#define WRKSP 1000
#define PI 3.141592653589793
#define PIH 1.570796326794897
#define TPI 6.283185307179586
#define FPI 12.56637061435917
#define TINY 1e-12
#define MAXD(a, b) ((a) > (b) ? (a) : (b))
#define MIND(a, b) ((a) < (b) ? (a) : (b))

/* ------------------------------------------------------------------------ */
/* Quadrature tables: include/GaussPts.h:76-136 (QUAD22), verbatim decimals. */
/* ------------------------------------------------------------------------ */
static const double xpt11[22] = {
    -1., -0.50118723, -0.25118864, -0.12589254, -0.06309573, -0.03162278, -0.01584893,
    -0.00794328, -0.00398107, -0.00199526, -0.001, 0.001, 0.00199526, 0.00398107,
    0.00794328, 0.01584893, 0.03162278, 0.06309573, 0.12589254, 0.25118864, 0.50118723, 1.};
static const double xw11[22] = {
    0.49881277, 0.37440568, 0.18764735, 0.09404645, 0.04713488, 0.0236234, 0.01183975,
    0.00593393, 0.00297401, 0.00149054, 0.00149763, 0.00149763, 0.00149054, 0.00297401,
    0.00593393, 0.01183975, 0.0236234, 0.04713488, 0.09404645, 0.18764735, 0.37440568,
    0.49881277};
static const double xpt01[22] = {
    1.12614037692e-05, 8.09053908976e-05, 0.000425204351858, 0.00172603419131,
    0.0056596260047, 0.015541463911, 0.0367576002085, 0.0764489202523,
    0.141870454058, 0.237263531493, 0.360233468529, 0.5,
    0.639766531471, 0.762736468507, 0.858129545942, 0.923551079748,
    0.963242399792, 0.984458536089, 0.994340373995, 0.998273965809,
    0.999574795648, 0.999919094609};
static const double xw01[22] = {
    2.42000467048e-05, 0.000146089994545, 0.000647128630152, 0.00222235986647,
    0.00618889854184, 0.0144761668119, 0.0291728616842, 0.0514732311361,
    0.0801433368009, 0.110279557896, 0.133823374104, 0.142799666072,
    0.133823374104, 0.110279557896, 0.0801433368009, 0.0514732311361,
    0.0291728616842, 0.0144761668119, 0.00618889854184, 0.00222235986647,
    0.000647128630152, 0.000146089994545};

struct fjo_pot {
    int nr, npoly, ngauss, ngrid, quad;
    double *ar;
    double *phil, *Pr, *Pr2;     /* [nr][npoly] total potential */
    double *Egrid, *Dgrid;       /* [ngrid] */
};

/* ------------------------------------------------------------------------ */
/* counters (per thread)                                                     */
/* ------------------------------------------------------------------------ */
static _Thread_local fjo_counters g_cnt;
void fjo_counters_get(fjo_counters *c) { *c = g_cnt; }
void fjo_counters_reset(void) { memset(&g_cnt, 0, sizeof g_cnt); }

/* ------------------------------------------------------------------------ */
/* grid                                                                      */
/* ------------------------------------------------------------------------ */
fjo_pot *fjo_pot_new(int nr, const double *ar, int npoly, int ngauss, int ngrid) {
    /* nr >= 11 keeps the quadratic-interpolation neighbour (UtilsLeg.h:113-115,
     * used for the first 10 cells) inside the table. */
    if (nr < 11 || npoly < 1 || npoly > 64 || ngauss < 1 || ngrid < 2) return NULL;
    fjo_pot *p = (fjo_pot *)calloc(1, sizeof *p);
    p->nr = nr; p->npoly = npoly; p->ngauss = ngauss; p->ngrid = ngrid; p->quad = 0;
    p->ar = (double *)malloc(sizeof(double) * nr);
    memcpy(p->ar, ar, sizeof(double) * nr);
    p->phil = (double *)calloc((size_t)nr * npoly, sizeof(double));
    p->Pr = (double *)calloc((size_t)nr * npoly, sizeof(double));
    p->Pr2 = (double *)calloc((size_t)nr * npoly, sizeof(double));
    p->Egrid = (double *)calloc(ngrid, sizeof(double));
    p->Dgrid = (double *)calloc(ngrid, sizeof(double));
    return p;
}
void fjo_pot_free(fjo_pot *p) {
    if (!p) return;
    free(p->ar); free(p->phil); free(p->Pr); free(p->Pr2); free(p->Egrid); free(p->Dgrid);
    free(p);
}
void fjo_pot_set_tables(fjo_pot *p, const double *phil, const double *Pr, const double *Pr2) {
    size_t n = (size_t)p->nr * p->npoly;
    memcpy(p->phil, phil, n * sizeof(double));
    memcpy(p->Pr, Pr, n * sizeof(double));
    memcpy(p->Pr2, Pr2, n * sizeof(double));
}
void fjo_pot_get_tables(const fjo_pot *p, double *phil, double *Pr, double *Pr2) {
    size_t n = (size_t)p->nr * p->npoly;
    memcpy(phil, p->phil, n * sizeof(double));
    memcpy(Pr, p->Pr, n * sizeof(double));
    memcpy(Pr2, p->Pr2, n * sizeof(double));
}
void fjo_pot_set_delta(fjo_pot *p, const double *E, const double *D) {
    memcpy(p->Egrid, E, p->ngrid * sizeof(double));
    memcpy(p->Dgrid, D, p->ngrid * sizeof(double));
}
void fjo_pot_get_delta(const fjo_pot *p, double *E, double *D) {
    memcpy(E, p->Egrid, p->ngrid * sizeof(double));
    memcpy(D, p->Dgrid, p->ngrid * sizeof(double));
}
void fjo_pot_set_quad(fjo_pot *p, int preset) { p->quad = preset; }

/* include/Utils.h:103-108 SetGrid(rmax) with rmin made explicit */
void fjo_log_grid(int n, double rmin, double rmax, double *ar) {
    for (int i = 0; i < n; i++)
        ar[i] = pow(10., log10(rmin) + (log10(rmax) - log10(rmin)) * i / (n - 1));
}

/* ------------------------------------------------------------------------ */
/* interpolation: include/Utils.h:130-155                                    */
/* ------------------------------------------------------------------------ */
static int topbottom(const double *xm, int n, double x, int *botx, int *topx) {
    int m, top = n - 1, bot = 0;
    if ((xm[top] - x) * (x - xm[bot]) < 0.) {
        if (fabs(xm[top] - x) < 1e-5 * fabs(x - xm[bot])) {
            *topx = top; *botx = top - 1; return 1;
        } else if (fabs(x - xm[bot]) < 1e-5 * fabs(x - xm[top])) {
            *topx = bot + 1; *botx = bot; return 1;
        } else
            return 0;
    }
    while (top - bot > 1) {
        m = (top + bot) / 2;
        if ((xm[top] - x) * (x - xm[m]) >= 0) bot = m;
        else top = m;
    }
    *topx = top; *botx = bot;
    return 1;
}

double fjo_linterp(const double *xm, const double *ym, int np, double x) {
    int top = 0, bot = 0;
    if (x <= xm[0]) return ym[0];
    if (x >= xm[np - 1]) return ym[np - 1];
    topbottom(xm, np, x, &bot, &top);
    return ym[bot] + (x - xm[bot]) / (xm[top] - xm[bot]) * (ym[top] - ym[bot]);
}

/* ------------------------------------------------------------------------ */
/* Legendre machinery: include/UtilsLeg.h:18-92                              */
/* ------------------------------------------------------------------------ */
void fjo_gauleg(int n, double x1, double x2, double *x, double *w) {
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1), pp = 0, z1;
    for (int i = 0; i < m; i++) {
        double z = cos(3.141592654 * (i + 0.75) / (n + 0.5));
        do {
            double p1 = 1, p2 = 0, p3;
            for (int j = 0; j < n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2 * (j + 1) - 1) * z * p2 - j * p3) / (double)(j + 1);
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > 3.0e-11);
        x[i] = xm - xl * z;
        x[n - i - 1] = xm + xl * z;
        w[i] = 2 * xl / ((1 - z * z) * pp * pp);
        w[n - i - 1] = w[i];
    }
}

void fjo_legend(double *allpol, double c, int npoly) {
    npoly++;
    allpol[0] = 1;
    if (npoly < 2) return;
    allpol[1] = c;
    for (int i = 2; i < npoly; i++)
        allpol[i] = ((2 * i - 1) * c * allpol[i - 1] - (i - 1) * allpol[i - 2]) / (double)i;
}

double fjo_dlegend(double c, int n) {
    if (n == 0) return 0;
    double allpol[n + 1];
    fjo_legend(allpol, c, n);
    return (n * allpol[n - 1] - n * c * allpol[n]) / (1 - c * c);
}

void fjo_evenlegend(int npoly, double c, double *pol) {
    double c2 = c * c;
    pol[0] = 1;
    if (npoly < 2) return;
    pol[1] = 1.5 * c2 - .5;
    for (int np = 2; np < npoly; np++) {
        int l = 2 * (np - 1), l2 = 2 * l;
        pol[np] = -pol[np - 2] * l * (l - 1) / (double)((l2 + 1) * (l2 - 1)) +
                  pol[np - 1] * (c2 - (l2 * l + l2 - 1) / (double)((l2 - 1) * (l2 + 3)));
        pol[np] *= (l2 + 1) * (l2 + 3) / (double)((l + 1) * (l + 2));
    }
}

/* include/UtilsLeg.h:96-124 (one table) */
static void intpo2_1(const fjo_pot *p, double r, const double *philH, double *phip) {
    const int npoly = p->npoly, NR = p->nr;
    const double *ar = p->ar;
    if (r >= ar[NR - 1]) {
        for (int k = 0; k < npoly; k++)
            phip[k] = philH[(NR - 1) * npoly + k] * fjm_powi(ar[NR - 1] / r, 2 * k + 1);
    } else {
        int top = 0, bot = 0;
        r = MAXD(r, ar[0]);
        topbottom(ar, NR, r, &bot, &top);
        double db = r - ar[bot], f1 = db / (ar[top] - ar[bot]);
        for (int k = 0; k < npoly; k++)
            phip[k] = f1 * philH[top * npoly + k] + (1 - f1) * philH[bot * npoly + k];
        if (top < 10) {
            int thr;
            if (f1 < 0.5 && bot > 0) thr = bot - 1; else thr = top + 1;
            double dt = r - ar[top];
            double f2 = dt * db / ((ar[thr] - ar[top]) * (ar[thr] - ar[bot])),
                   f3 = (ar[thr] - ar[bot]) / (ar[top] - ar[bot]);
            for (int k = 0; k < npoly; k++)
                phip[k] += f2 * (philH[thr * npoly + k] - philH[bot * npoly + k] -
                                 f3 * (philH[top * npoly + k] - philH[bot * npoly + k]));
        }
    }
}

/* include/UtilsLeg.h:127-158 (three tables) */
static void intpo2_3(const fjo_pot *p, double r, double *phip, double *dphip, double *d2phip) {
    const int npoly = p->npoly, nr = p->nr;
    const double *ar = p->ar, *philH = p->phil, *PrH = p->Pr, *Pr2H = p->Pr2;
    if (r >= ar[nr - 1]) {
        for (int k = 0; k < npoly; k++) {
            phip[k] = philH[(nr - 1) * npoly + k] * fjm_powi(ar[nr - 1] / r, 2 * k + 1);
            dphip[k] = -(2 * k + 1) * philH[(nr - 1) * npoly + k] * fjm_powi(ar[nr - 1] / r, 2 * k + 1) / r;
            d2phip[k] = (2 * k + 2) * (2 * k + 1) * philH[(nr - 1) * npoly + k] *
                        fjm_powi(ar[nr - 1] / r, 2 * k + 1) / r / r;
        }
    } else {
        int top = 0, bot = 0;
        topbottom(ar, nr, r, &bot, &top);
        double db = r - ar[bot], f1 = db / (ar[top] - ar[bot]);
        for (int k = 0; k < npoly; k++) {
            phip[k] = f1 * philH[top * npoly + k] + (1 - f1) * philH[bot * npoly + k];
            dphip[k] = f1 * PrH[top * npoly + k] + (1 - f1) * PrH[bot * npoly + k];
            d2phip[k] = f1 * Pr2H[top * npoly + k] + (1 - f1) * Pr2H[bot * npoly + k];
        }
        if (top < 10) {
            int thr;
            if (f1 < 0.5 && bot > 0) thr = bot - 1; else thr = top + 1;
            double dt = r - ar[top];
            double f2 = dt * db / ((ar[thr] - ar[top]) * (ar[thr] - ar[bot])),
                   f3 = (ar[thr] - ar[bot]) / (ar[top] - ar[bot]);
            for (int k = 0; k < npoly; k++) {
                int t = top * npoly + k, b = bot * npoly + k, h = thr * npoly + k;
                phip[k] += f2 * (philH[h] - philH[b] - f3 * (philH[t] - philH[b]));
                dphip[k] += f2 * (PrH[h] - PrH[b] - f3 * (PrH[t] - PrH[b]));
                d2phip[k] += f2 * (Pr2H[h] - Pr2H[b] - f3 * (Pr2H[t] - Pr2H[b]));
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Potential evaluation: pot/Potential.cpp:206-276                            */
/* ------------------------------------------------------------------------ */
double fjo_phi(const fjo_pot *p, double R, double z) {
    double pol[64], phip[64];
    double c = z / sqrt(R * R + z * z), r = sqrt(R * R + z * z);
    g_cnt.phi_evals++;
    intpo2_1(p, r, p->phil, phip);
    fjo_evenlegend(p->npoly, c, pol);
    double phi = phip[0];
    for (int np = 1; np < p->npoly; np++) phi += phip[np] * pol[np];
    return phi;
}

static double pot_dr(const fjo_pot *p, double r, double theta, double *d2Phidr2) {
    double pol[64], phip[64], dphip[64], d2phip[64];
    double grad1 = 0, grad2 = 0;
    fjo_evenlegend(p->npoly, fjm_cos(theta), pol);
    intpo2_3(p, r, phip, dphip, d2phip);
    for (int np = 0; np < p->npoly; np++) {
        grad1 += dphip[np] * pol[np];
        grad2 += d2phip[np] * pol[np];
    }
    *d2Phidr2 = grad2;
    return grad1;
}

static double pot_dtheta(const fjo_pot *p, double r, double c) {
    double dphitheta = 0, dpol[64], phinterp[64];
    intpo2_1(p, r, p->phil, phinterp);
    for (int k = 0; k < p->npoly; k++) dpol[k] = fjo_dlegend(c, 2 * k);
    for (int j = 0; j < p->npoly; j++) dphitheta += phinterp[j] * dpol[j];
    dphitheta *= -sqrt(1 - c * c);
    return dphitheta;
}

double fjo_dR(const fjo_pot *p, double R, double z) {
    double r = sqrt(z * z + R * R);
    g_cnt.grad_evals++;
    if (r < p->ar[0]) r = p->ar[0];
    double s = R / r, c = z / r;
    if (c == 1) c -= 1e-8;
    double theta = fjm_acos(c), d2;
    return pot_dtheta(p, r, c) * (c / r) + pot_dr(p, r, theta, &d2) * s;
}

double fjo_dz(const fjo_pot *p, double R, double z) {
    double r = sqrt(z * z + R * R);
    if (r < p->ar[0]) r = p->ar[0];
    double s = R / r, c = z / r;
    if (c == 1) c -= 1e-8;
    double theta = fjm_acos(c), d2;
    return pot_dtheta(p, r, c) * (-s / r) + pot_dr(p, r, theta, &d2) * c;
}

/* ------------------------------------------------------------------------ */
/* Brent root finder: GSL roots/brent.c iterate form of Brent (1973),         */
/* driven by the reference wrapper include/Utils.h:134-167.                   */
/* ------------------------------------------------------------------------ */
typedef struct { double a, b, c, d, e, fa, fb, fc; } brent_state;

static void brent_init(brent_state *s, double xl, double fl, double xu, double fu) {
    s->a = xl; s->fa = fl;
    s->b = xu; s->fb = fu;
    s->c = xu; s->fc = fu;
    s->d = xu - xl; s->e = xu - xl;
}

/* One iteration; returns the new root estimate and bracket. */
static void brent_iterate(brent_state *s, fjo_fn F, void *prm, double *root, double *xlo, double *xup) {
    double tol, m;
    int ac_equal = 0;
    double a = s->a, b = s->b, c = s->c, fa = s->fa, fb = s->fb, fc = s->fc, d = s->d, e = s->e;
    if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
        ac_equal = 1; c = a; fc = fa; d = b - a; e = b - a;
    }
    if (fabs(fc) < fabs(fb)) {
        ac_equal = 1; a = b; b = c; c = a; fa = fb; fb = fc; fc = fa;
    }
    tol = 0.5 * DBL_EPSILON * fabs(b);
    m = 0.5 * (c - b);
    if (fb == 0) { *root = b; *xlo = b; *xup = b; return; }
    if (fabs(m) <= tol) {
        *root = b;
        if (b < c) { *xlo = b; *xup = c; } else { *xlo = c; *xup = b; }
        return;
    }
    if (fabs(e) < tol || fabs(fa) <= fabs(fb)) {
        d = m; e = m;
    } else {
        double p, q, r, ss = fb / fa;
        if (ac_equal) {
            p = 2 * m * ss;
            q = 1 - ss;
        } else {
            q = fa / fc;
            r = fb / fc;
            p = ss * (2 * m * q * (q - r) - (b - a) * (r - 1));
            q = (q - 1) * (r - 1) * (ss - 1);
        }
        if (p > 0) q = -q; else p = -p;
        if (2 * p < MIND(3 * m * q - fabs(tol * q), fabs(e * q))) {
            e = d; d = p / q;
        } else {
            d = m; e = m;
        }
    }
    a = b; fa = fb;
    if (fabs(d) > tol) b += d;
    else b += (m > 0 ? +tol : -tol);
    fb = F(b, prm);
    s->a = a; s->b = b; s->c = c; s->d = d; s->e = e; s->fa = fa; s->fb = fb; s->fc = fc;
    *root = b;
    if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) c = a;
    if (b < c) { *xlo = b; *xup = c; } else { *xlo = c; *xup = b; }
}

/* gsl_root_test_interval as documented: |xu-xl| < epsabs + epsrel*min(|xl|,|xu|)
 * (min taken as 0 when the interval contains the origin). */
static int test_interval(double xl, double xu, double epsabs, double epsrel) {
    double min_abs;
    if ((xl > 0.0 && xu > 0.0) || (xl < 0.0 && xu < 0.0)) min_abs = MIND(fabs(xl), fabs(xu));
    else min_abs = 0;
    double tol = epsabs + epsrel * min_abs;
    return fabs(xu - xl) < tol ? 0 : 1; /* 0 = success, 1 = continue */
}

/* wrp_GSLroots with F(low), F(high) already known (the wrapper evaluates the
 * same pure function at the same arguments again; the values are equal). */
static double roots_known(fjo_fn F, void *prm, double low, double high, double flow, double fhigh, int *iters) {
    if (fabs(flow) < 5e-4) return low;
    if (fabs(fhigh) < 5e-4) return high;
    brent_state s;
    brent_init(&s, low, flow, high, fhigh);
    int cont = 0, it = 0;
    double Rout = 0;
    while (cont || high - low > EPSABSROOTS) {
        it++;
        brent_iterate(&s, F, prm, &Rout, &low, &high);
        cont = test_interval(low, high, EPSABSROOTS, EPSRELROOTS);
        if (it > 10000) break; /* the reference would loop forever here */
    }
    g_cnt.brent_iters += it;
    if (iters) *iters = it;
    return Rout;
}

double fjo_roots(fjo_fn F, void *prm, double low, double high, int *iters) {
    double fl = F(low, prm), fh = F(high, prm);
    if (iters) *iters = 0;
    return roots_known(F, prm, low, high, fl, fh, iters);
}

/* ------------------------------------------------------------------------ */
/* QUADPACK Gauss-Kronrod rules dqk21 / dqk15                                */
/* ------------------------------------------------------------------------ */
static const double xgk21[11] = {
    0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
    0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
    0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
    0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
    0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
    0.000000000000000000000000000000000};
static const double wgk21[11] = {
    0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
    0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
    0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
    0.123491976262065851077208980178525, 0.134709217311473325928054001771707,
    0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
    0.149445554002916905664936468389821};
static const double wg10[5] = {
    0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
    0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
    0.295524224714752870173892994651338};
static const double xgk15[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double wgk15[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double wg7[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

static double rescale_error(double err, double resabs, double resasc) {
    err = fabs(err);
    if (resasc != 0 && err != 0) {
        /* (200 err / resasc)^1.5, evaluated as t sqrt(t) */
        double t = 200 * err / resasc;
        double scale = t * sqrt(t);
        if (scale < 1) err = resasc * scale; else err = resasc;
    }
    if (resabs > DBL_MIN / (50 * DBL_EPSILON)) {
        double min_err = 50 * DBL_EPSILON * resabs;
        if (min_err > err) err = min_err;
    }
    return err;
}

/* n = number of Kronrod abscissae >= 0 (11 for QK21, 8 for QK15) */
static void qk_rule(int n, const double *xgk, const double *wg, const double *wgk,
                    fjo_fn f, void *prm, double a, double b,
                    double *result, double *abserr, double *resabs, double *resasc) {
    double fv1[11], fv2[11];
    const double center = 0.5 * (a + b), half = 0.5 * (b - a), ahalf = fabs(half);
    const double fc = f(center, prm);
    double rg = 0, rk = fc * wgk[n - 1], rabs = fabs(rk), rasc, mean;
    if (n % 2 == 0) rg = fc * wg[n / 2 - 1];
    for (int j = 0; j < (n - 1) / 2; j++) {
        int jtw = j * 2 + 1;
        double absc = half * xgk[jtw];
        double f1 = f(center - absc, prm), f2 = f(center + absc, prm), fs = f1 + f2;
        fv1[jtw] = f1; fv2[jtw] = f2;
        rg += wg[j] * fs;
        rk += wgk[jtw] * fs;
        rabs += wgk[jtw] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < n / 2; j++) {
        int jtwm1 = j * 2;
        double absc = half * xgk[jtwm1];
        double f1 = f(center - absc, prm), f2 = f(center + absc, prm);
        fv1[jtwm1] = f1; fv2[jtwm1] = f2;
        rk += wgk[jtwm1] * (f1 + f2);
        rabs += wgk[jtwm1] * (fabs(f1) + fabs(f2));
    }
    mean = rk * 0.5;
    rasc = wgk[n - 1] * fabs(fc - mean);
    for (int j = 0; j < n - 1; j++) rasc += wgk[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    double err = (rk - rg) * half;
    rk *= half; rabs *= ahalf; rasc *= ahalf;
    *result = rk; *resabs = rabs; *resasc = rasc;
    *abserr = rescale_error(err, rabs, rasc);
}

/* Interval recorder of fjo_orbit_state: the [a, b] of every QK21 rule applied. */
static _Thread_local double *g_rec;
static _Thread_local int g_rec_n, g_rec_cap;

void fjo_qk21(fjo_fn f, void *prm, double a, double b, double *r, double *e, double *ra, double *rs) {
    g_cnt.qk21_calls++;
    if (g_rec) {
        if (g_rec_n < g_rec_cap) { g_rec[2 * g_rec_n] = a; g_rec[2 * g_rec_n + 1] = b; }
        g_rec_n++;
    }
    qk_rule(11, xgk21, wg10, wgk21, f, prm, a, b, r, e, ra, rs);
}
void fjo_qk15(fjo_fn f, void *prm, double a, double b, double *r, double *e, double *ra, double *rs) {
    qk_rule(8, xgk15, wg7, wgk15, f, prm, a, b, r, e, ra, rs);
}

/* QUADPACK dqelg: epsilon algorithm.  epstab is 1-based sized >= 53. */
static void qelg(int *n_io, double *epstab, double *result, double *abserr, double *res3la, int *nres) {
    const double epmach = DBL_EPSILON, oflow = DBL_MAX;
    int n = *n_io;
    (*nres)++;
    *abserr = oflow;
    *result = epstab[n];
    if (n < 3) goto L100;
    {
        const int limexp = 50;
        epstab[n + 2] = epstab[n];
        int newelm = (n - 1) / 2;
        epstab[n] = oflow;
        int num = n, k1 = n;
        for (int i = 1; i <= newelm; i++) {
            int k2 = k1 - 1, k3 = k1 - 2;
            double res = epstab[k1 + 2];
            double e0 = epstab[k3], e1 = epstab[k2], e2 = res;
            double e1abs = fabs(e1), delta2 = e2 - e1, err2 = fabs(delta2);
            double tol2 = MAXD(fabs(e2), e1abs) * epmach;
            double delta3 = e1 - e0, err3 = fabs(delta3);
            double tol3 = MAXD(e1abs, fabs(e0)) * epmach;
            if (!(err2 > tol2 || err3 > tol3)) {
                *result = res;
                *abserr = err2 + err3;
                goto L100;
            }
            double e3 = epstab[k1];
            epstab[k1] = e1;
            double delta1 = e1 - e3, err1 = fabs(delta1);
            double tol1 = MAXD(e1abs, fabs(e3)) * epmach;
            double ss, epsinf;
            if (err1 <= tol1 || err2 <= tol2 || err3 <= tol3) goto L20;
            ss = 0.1e+01 / delta1 + 0.1e+01 / delta2 - 0.1e+01 / delta3;
            epsinf = fabs(ss * e1);
            if (epsinf > 0.1e-03) goto L30;
        L20:
            n = i + i - 1;
            goto L50;
        L30:
            res = e1 + 0.1e+01 / ss;
            epstab[k1] = res;
            k1 = k1 - 2;
            {
                double error = err2 + fabs(res - e2) + err3;
                if (!(error > *abserr)) {
                    *abserr = error;
                    *result = res;
                }
            }
        }
    L50:
        if (n == limexp) n = 2 * (limexp / 2) - 1;
        {
            int ib = ((num / 2) * 2 == num) ? 2 : 1;
            int ie = newelm + 1;
            for (int i = 1; i <= ie; i++) {
                int ib2 = ib + 2;
                epstab[ib] = epstab[ib2];
                ib = ib2;
            }
            if (num != n) {
                int indx = num - n + 1;
                for (int i = 1; i <= n; i++) {
                    epstab[i] = epstab[indx];
                    indx++;
                }
            }
        }
        if (*nres >= 4) {
            *abserr = fabs(*result - res3la[3]) + fabs(*result - res3la[2]) + fabs(*result - res3la[1]);
            res3la[1] = res3la[2];
            res3la[2] = res3la[3];
            res3la[3] = *result;
        } else {
            res3la[*nres] = *result;
            *abserr = oflow;
        }
    }
L100:
    *abserr = MAXD(*abserr, 0.5e+01 * epmach * fabs(*result));
    *n_io = n;
}

/* QUADPACK dqpsrt: maintain descending ordering of error estimates (1-based). */
static void qpsrt(int limit, int last, int *maxerr, double *ermax, const double *elist, int *iord, int *nrmax) {
    int i, j, k, ido, ibeg, jbnd, isucc, jupbn;
    double errmax, errmin;
    if (last > 2) goto L10;
    iord[1] = 1;
    iord[2] = 2;
    goto L90;
L10:
    errmax = elist[*maxerr];
    if (*nrmax == 1) goto L30;
    ido = *nrmax - 1;
    for (i = 1; i <= ido; i++) {
        isucc = iord[*nrmax - 1];
        if (errmax <= elist[isucc]) goto L30;
        iord[*nrmax] = isucc;
        (*nrmax)--;
    }
L30:
    jupbn = last;
    if (last > (limit / 2 + 2)) jupbn = limit + 3 - last;
    errmin = elist[last];
    jbnd = jupbn - 1;
    ibeg = *nrmax + 1;
    if (ibeg > jbnd) goto L50;
    for (i = ibeg; i <= jbnd; i++) {
        isucc = iord[i];
        if (errmax >= elist[isucc]) goto L60;
        iord[i - 1] = isucc;
    }
L50:
    iord[jbnd] = *maxerr;
    iord[jupbn] = last;
    goto L90;
L60:
    iord[i - 1] = *maxerr;
    k = jbnd;
    for (j = i; j <= jbnd; j++) {
        isucc = iord[k];
        if (errmin < elist[isucc]) goto L80;
        iord[k + 1] = isucc;
        k--;
    }
    iord[i] = last;
    goto L90;
L80:
    iord[k + 1] = last;
L90:
    *maxerr = iord[*nrmax];
    *ermax = elist[*maxerr];
}

typedef void (*qk_fn)(fjo_fn, void *, double, double, double *, double *, double *, double *);

/* QUADPACK dqagse (1-based workspace), with the rule as a parameter: QK21
 * for QAGS, QK15 for the transformed QAGIU integrand. Returns ier. */
static int qagse(qk_fn rule, fjo_fn f, void *prm, double a, double b, double epsabs, double epsrel,
                 int limit, double *result, double *abserr, int *neval, int *last_out) {
    const double epmach = DBL_EPSILON, uflow = DBL_MIN, oflow = DBL_MAX;
    double *alist = (double *)malloc(sizeof(double) * (limit + 1));
    double *blist = (double *)malloc(sizeof(double) * (limit + 1));
    double *rlist = (double *)malloc(sizeof(double) * (limit + 1));
    double *elist = (double *)malloc(sizeof(double) * (limit + 1));
    int *iord = (int *)malloc(sizeof(int) * (limit + 1));
    double rlist2[56], res3la[4];
    double area, area1, area2, area12, a1, a2, b1, b2, defabs, defab1, defab2, dres, erlarg = 0,
           erlast, errbnd, errmax, error1, error2, erro12, errsum, ertest = 0, resabs, reseps,
           abseps, small = 0, correc = 0;
    int ier = 0, ierro, iroff1, iroff2, iroff3, k, ksgn, ktmin, last, maxerr, nres, nrmax, numrl2,
        extrap, noext, id, jupbnd;

    *neval = 0; last = 0; *result = 0; *abserr = 0;
    alist[1] = a; blist[1] = b; rlist[1] = 0; elist[1] = 0;
    if (epsabs <= 0 && epsrel < MAXD(50 * epmach, 0.5e-28)) { ier = 6; goto L999; }

    ierro = 0;
    rule(f, prm, a, b, result, abserr, &defabs, &resabs);
    dres = fabs(*result);
    errbnd = MAXD(epsabs, epsrel * dres);
    last = 1;
    rlist[1] = *result; elist[1] = *abserr; iord[1] = 1;
    if (*abserr <= 1.0e+02 * epmach * defabs && *abserr > errbnd) ier = 2;
    if (limit == 1) ier = 1;
    if (ier != 0 || (*abserr <= errbnd && *abserr != resabs) || *abserr == 0.0) goto L140;

    rlist2[1] = *result;
    errmax = *abserr; maxerr = 1; area = *result; errsum = *abserr; *abserr = oflow;
    nrmax = 1; nres = 0; numrl2 = 2; ktmin = 0; extrap = 0; noext = 0;
    iroff1 = 0; iroff2 = 0; iroff3 = 0;
    ksgn = -1;
    if (dres >= (1.0 - 0.5e+02 * epmach) * defabs) ksgn = 1;

    for (last = 2; last <= limit; last++) {
        a1 = alist[maxerr];
        b1 = 0.5 * (alist[maxerr] + blist[maxerr]);
        a2 = b1;
        b2 = blist[maxerr];
        erlast = errmax;
        rule(f, prm, a1, b1, &area1, &error1, &resabs, &defab1);
        rule(f, prm, a2, b2, &area2, &error2, &resabs, &defab2);
        area12 = area1 + area2;
        erro12 = error1 + error2;
        errsum = errsum + erro12 - errmax;
        area = area + area12 - rlist[maxerr];
        if (defab1 == error1 || defab2 == error2) goto L15;
        if (fabs(rlist[maxerr] - area12) > 0.1e-04 * fabs(area12) || erro12 < 0.99 * errmax) goto L10;
        if (extrap) iroff2++;
        if (!extrap) iroff1++;
    L10:
        if (last > 10 && erro12 > errmax) iroff3++;
    L15:
        rlist[maxerr] = area1;
        rlist[last] = area2;
        errbnd = MAXD(epsabs, epsrel * fabs(area));
        if (iroff1 + iroff2 >= 10 || iroff3 >= 20) ier = 2;
        if (iroff2 >= 5) ierro = 3;
        if (last == limit) ier = 1;
        if (MAXD(fabs(a1), fabs(b2)) <= (1.0 + 0.1e+03 * epmach) * (fabs(a2) + 0.1e+04 * uflow)) ier = 4;
        if (error2 > error1) goto L20;
        alist[last] = a2;
        blist[maxerr] = b1;
        blist[last] = b2;
        elist[maxerr] = error1;
        elist[last] = error2;
        goto L30;
    L20:
        alist[maxerr] = a2;
        alist[last] = a1;
        blist[last] = b1;
        rlist[maxerr] = area2;
        rlist[last] = area1;
        elist[maxerr] = error2;
        elist[last] = error1;
    L30:
        qpsrt(limit, last, &maxerr, &errmax, elist, iord, &nrmax);
        if (errsum <= errbnd) goto L115;
        if (ier != 0) goto L100;
        if (last == 2) goto L80;
        if (noext) continue;
        erlarg = erlarg - erlast;
        if (fabs(b1 - a1) > small) erlarg = erlarg + erro12;
        if (extrap) goto L40;
        if (fabs(blist[maxerr] - alist[maxerr]) > small) continue;
        extrap = 1;
        nrmax = 2;
    L40:
        if (ierro == 3 || erlarg <= ertest) goto L60;
        id = nrmax;
        jupbnd = last;
        if (last > (2 + limit / 2)) jupbnd = limit + 3 - last;
        for (k = id; k <= jupbnd; k++) {
            maxerr = iord[nrmax];
            errmax = elist[maxerr];
            if (fabs(blist[maxerr] - alist[maxerr]) > small) goto L90;
            nrmax++;
        }
    L60:
        numrl2++;
        rlist2[numrl2] = area;
        qelg(&numrl2, rlist2, &reseps, &abseps, res3la, &nres);
        ktmin++;
        if (ktmin > 5 && *abserr < 0.1e-02 * errsum) ier = 5;
        if (abseps >= *abserr) goto L70;
        ktmin = 0;
        *abserr = abseps;
        *result = reseps;
        correc = erlarg;
        ertest = MAXD(epsabs, epsrel * fabs(reseps));
        if (*abserr <= ertest) goto L100;
    L70:
        if (numrl2 == 1) noext = 1;
        if (ier == 5) goto L100;
        maxerr = iord[1];
        errmax = elist[maxerr];
        nrmax = 1;
        extrap = 0;
        small = small * 0.5;
        erlarg = errsum;
        continue;
    L80:
        small = fabs(b - a) * 0.375;
        erlarg = errsum;
        ertest = errbnd;
        rlist2[2] = area;
        continue;
    L90:;
    }
    if (last > limit) last = limit;
L100:
    if (*abserr == oflow) goto L115;
    if (ier + ierro == 0) goto L110;
    if (ierro == 3) *abserr = *abserr + correc;
    if (ier == 0) ier = 3;
    if (*result != 0.0 && area != 0.0) goto L105;
    if (*abserr > errsum) goto L115;
    if (area == 0.0) goto L130;
    goto L110;
L105:
    if (*abserr / fabs(*result) > errsum / fabs(area)) goto L115;
L110:
    if (ksgn == -1 && MAXD(fabs(*result), fabs(area)) <= defabs * 0.1e-01) goto L130;
    if (0.1e-01 > (*result / area) || (*result / area) > 0.1e+03 || errsum > fabs(area)) ier = 6;
    goto L130;
L115:
    *result = 0.0;
    for (k = 1; k <= last; k++) *result = *result + rlist[k];
    *abserr = errsum;
L130:
    if (ier > 2) ier = ier - 1;
L140:
    *neval = (rule == fjo_qk15) ? 30 * last - 15 : 42 * last - 21;
L999:
    if (last_out) *last_out = last;
    free(alist); free(blist); free(rlist); free(elist); free(iord);
    return ier;
}

int fjo_qags(fjo_fn f, void *prm, double a, double b, double epsabs, double epsrel,
             int limit, double *result, double *abserr, int *neval, int *last) {
    int ne = 0;
    int ier = qagse(fjo_qk21, f, prm, a, b, epsabs, epsrel, limit, result, abserr, &ne, last);
    if (neval) *neval = ne;
    return ier;
}

/* GSL qagiu: x = a + (1-t)/t on t in (0,1], integrand (f(x)/t)/t, QK15. */
typedef struct { fjo_fn f; void *prm; double a; } iu_par;
static double iu_transform(double t, void *p) {
    iu_par *q = (iu_par *)p;
    double x = q->a + (1 - t) / t;
    double y = q->f(x, q->prm);
    return (y / t) / t;
}
int fjo_qagiu(fjo_fn f, void *prm, double a, double epsabs, double epsrel,
              int limit, double *result, double *abserr, int *neval, int *last) {
    iu_par q = {f, prm, a};
    int ne = 0;
    int ier = qagse(fjo_qk15, iu_transform, &q, 0.0, 1.0, epsabs, epsrel, limit, result, abserr, &ne, last);
    if (neval) *neval = ne;
    return ier;
}

/* ------------------------------------------------------------------------ */
/* RK8PD (Prince & Dormand 1981, RK8(7)13M) and the GSL-documented evolve     */
/* loop with standard y-control: used by delta/Delta.cpp:53-90.               */
/* ------------------------------------------------------------------------ */
static const double rk_c[13] = {0.0, 1.0 / 18.0, 1.0 / 12.0, 1.0 / 8.0, 5.0 / 16.0, 3.0 / 8.0,
                                59.0 / 400.0, 93.0 / 200.0, 5490023248.0 / 9719169821.0,
                                13.0 / 20.0, 1201146811.0 / 1299019798.0, 1.0, 1.0};
static const double rk_a[13][12] = {
    {0},
    {1.0 / 18.0},
    {1.0 / 48.0, 1.0 / 16.0},
    {1.0 / 32.0, 0.0, 3.0 / 32.0},
    {5.0 / 16.0, 0.0, -75.0 / 64.0, 75.0 / 64.0},
    {3.0 / 80.0, 0.0, 0.0, 3.0 / 16.0, 3.0 / 20.0},
    {29443841.0 / 614563906.0, 0.0, 0.0, 77736538.0 / 692538347.0, -28693883.0 / 1125000000.0,
     23124283.0 / 1800000000.0},
    {16016141.0 / 946692911.0, 0.0, 0.0, 61564180.0 / 158732637.0, 22789713.0 / 633445777.0,
     545815736.0 / 2771057229.0, -180193667.0 / 1043307555.0},
    {39632708.0 / 573591083.0, 0.0, 0.0, -433636366.0 / 683701615.0, -421739975.0 / 2616292301.0,
     100302831.0 / 723423059.0, 790204164.0 / 839813087.0, 800635310.0 / 3783071287.0},
    {246121993.0 / 1340847787.0, 0.0, 0.0, -37695042795.0 / 15268766246.0,
     -309121744.0 / 1061227803.0, -12992083.0 / 490766935.0, 6005943493.0 / 2108947869.0,
     393006217.0 / 1396673457.0, 123872331.0 / 1001029789.0},
    {-1028468189.0 / 846180014.0, 0.0, 0.0, 8478235783.0 / 508512852.0,
     1311729495.0 / 1432422823.0, -10304129995.0 / 1701304382.0, -48777925059.0 / 3047939560.0,
     15336726248.0 / 1032824649.0, -45442868181.0 / 3398467696.0, 3065993473.0 / 597172653.0},
    {185892177.0 / 718116043.0, 0.0, 0.0, -3185094517.0 / 667107341.0,
     -477755414.0 / 1098053517.0, -703635378.0 / 230739211.0, 5731566787.0 / 1027545527.0,
     5232866602.0 / 850066563.0, -4093664535.0 / 808688257.0, 3962137247.0 / 1805957418.0,
     65686358.0 / 487910083.0},
    {403863854.0 / 491063109.0, 0.0, 0.0, -5068492393.0 / 434740067.0,
     -411421997.0 / 543043805.0, 652783627.0 / 914296604.0, 11173962825.0 / 925320556.0,
     -13158990841.0 / 6184727034.0, 3936647629.0 / 1978049680.0, -160528059.0 / 685178525.0,
     248638103.0 / 1413531060.0, 0.0}};
static const double rk_b8[13] = {14005451.0 / 335480064.0, 0.0, 0.0, 0.0, 0.0,
                                 -59238493.0 / 1068277825.0, 181606767.0 / 758867731.0,
                                 561292985.0 / 797845732.0, -1041891430.0 / 1371343529.0,
                                 760417239.0 / 1151165299.0, 118820643.0 / 751138087.0,
                                 -528747749.0 / 2220607170.0, 1.0 / 4.0};
static const double rk_b7[13] = {13451932.0 / 455176623.0, 0.0, 0.0, 0.0, 0.0,
                                 -808719846.0 / 976000145.0, 1757004468.0 / 5645159321.0,
                                 656045339.0 / 265891186.0, -3867574721.0 / 1518517206.0,
                                 465885868.0 / 322736535.0, 53011238.0 / 667516719.0,
                                 2.0 / 45.0, 0.0};

#define RK_MAXDIM 8
void fjo_rk8pd_step(fjo_ode_fn fn, void *prm, int dim, double t, double h,
                    double *y, double *yerr, const double *dydt_in) {
    double k[13][RK_MAXDIM], yt[RK_MAXDIM];
    for (int i = 0; i < dim; i++) k[0][i] = dydt_in[i];
    for (int s = 1; s < 13; s++) {
        for (int i = 0; i < dim; i++) {
            double acc = 0;
            for (int j = 0; j < s; j++)
                if (rk_a[s][j] != 0.0) acc += rk_a[s][j] * k[j][i];
            yt[i] = y[i] + h * acc;
        }
        fn(t + rk_c[s] * h, yt, k[s], prm);
    }
    for (int i = 0; i < dim; i++) {
        double s8 = 0, s7 = 0;
        for (int j = 0; j < 13; j++) {
            if (rk_b8[j] != 0.0) s8 += rk_b8[j] * k[j][i];
            if (rk_b7[j] != 0.0) s7 += rk_b7[j] * k[j][i];
        }
        y[i] += h * s8;
        yerr[i] = h * (s7 - s8);
    }
}

int fjo_evolve_apply(fjo_ode_fn fn, void *prm, int dim, double eps_abs, double eps_rel,
                     double *t, double t1, double *h, double *y) {
    const double t0 = *t;
    double h0 = *h, y0[RK_MAXDIM], yerr[RK_MAXDIM], dydt_in[RK_MAXDIM];
    int final_step;
    const double dt = t1 - t0;
    memcpy(y0, y, sizeof(double) * dim);
    fn(t0, y, dydt_in, prm);
    for (int tries = 0; tries < 1000; tries++) {
        if ((dt >= 0.0 && h0 > dt) || (dt < 0.0 && h0 < dt)) { h0 = dt; final_step = 1; }
        else final_step = 0;
        fjo_rk8pd_step(fn, prm, dim, t0, h0, y, yerr, dydt_in);
        *t = final_step ? t1 : t0 + h0;
        /* standard control, a_y = 1, a_dydt = 0, order 8 */
        const double h_old = h0;
        double rmax = DBL_MIN;
        for (int i = 0; i < dim; i++) {
            const double D0 = eps_rel * fabs(y[i]) + eps_abs;
            const double r = fabs(yerr[i]) / fabs(D0);
            rmax = MAXD(r, rmax);
        }
        if (rmax > 1.1) {
            double r = 0.9 / fjm_pow(rmax, 1.0 / 8.0);
            if (r < 0.2) r = 0.2;
            h0 = r * h_old;
            double t_curr = *t, t_next = *t + h0;
            if (fabs(h0) < fabs(h_old) && t_next != t_curr) {
                memcpy(y, y0, sizeof(double) * dim);
                continue;
            }
            h0 = h_old;
        } else if (rmax < 0.5) {
            double r = 0.9 / fjm_pow(rmax, 1.0 / 9.0);
            if (r > 5.0) r = 5.0;
            if (r < 1.0) r = 1.0;
            h0 = r * h_old;
        }
        if (!final_step) *h = h0;
        return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Guess densities: pot/models.cpp:40-68                                    */
/* ------------------------------------------------------------------------ */
typedef struct { double mass, r0, q, q2; } guess_par;
static double rho_guess(int model, const guess_par *g, double R, double z) {
    double m = sqrt(R * R + z * z / g->q2);
    switch (model) {
    case 0: /* Hernquist */
        return g->mass / (m / g->r0 * fjm_powi(g->r0 + m, 3)) / TPI;
    case 1: { /* Isochrone */
        double a = sqrt(g->r0 * g->r0 + m * m);
        return g->mass / FPI / (g->q * fjm_powi((a + g->r0) * a, 2) * a) * g->r0 * (g->r0 + 2 * a);
    }
    case 2: /* NFW */
        return g->mass / (m / g->r0 * fjm_powi(g->r0 + m, 2)) / FPI;
    default: /* NFWext */
        return .01 / (m / 25. * fjm_powi(25. + m, 1)) / FPI;
    }
}

/* pot/Potential.cpp:63-80 computeGuessRhl, with initLeg/fillPoly (UtilsLeg.h:81-92) */
void fjo_guess_rhl(const fjo_pot *p, int model, double mass, double r0, double q, double *rhl) {
    const int NR = p->nr, NP = p->npoly, NG = p->ngauss;
    double *ci = (double *)malloc(NG * sizeof(double)), *wi = (double *)malloc(NG * sizeof(double));
    double *si = (double *)malloc(NG * sizeof(double)), *poly = (double *)malloc((size_t)NP * NG * sizeof(double));
    double pol[64];
    guess_par g = {mass, r0, q, q * q};
    fjo_gauleg(NG, 0, 1, ci, wi);
    for (int i = 0; i < NG; i++) si[i] = sqrt(1 - ci[i] * ci[i]);
    for (int i = 0; i < NG; i++) {
        fjo_evenlegend(NP, ci[i], pol);
        for (int np = 0; np < NP; np++) poly[np * NG + i] = 2 * pol[np] * wi[i];
    }
    for (int nr = 0; nr < NR; nr++)
        for (int np = 0; np < NP; np++) rhl[nr * NP + np] = 0.;
    for (int nr = 0; nr < NR; nr++)
        for (int np = 0; np < NP; np++)
            for (int ng = 0; ng < NG; ng++)
                rhl[nr * NP + np] += poly[np * NG + ng] * rho_guess(model, &g, p->ar[nr] * si[ng], p->ar[nr] * ci[ng]);
    rhl[0] = 2 * rho_guess(model, &g, p->ar[0] * si[0], p->ar[0] * ci[0]);
    free(ci); free(wi); free(si); free(poly);
}

/* pot/Potential.cpp:83-101 */
typedef struct { int l; const double *rhl_h; const double *ar; int nr; } I_par;
static double I_internal(double x, void *params) {
    I_par *p = (I_par *)params;
    return fjm_powi(x, p->l + 2) * fjo_linterp(p->ar, p->rhl_h, p->nr, x);
}
static double I_external(double x, void *params) {
    I_par *p = (I_par *)params;
    const double *ar = p->ar;
    int NR = p->nr;
    double rhol;
    if (x > ar[NR - 1]) rhol = MAXD(0., p->rhl_h[NR - 1] * fjm_powi(ar[NR - 1] / x, 4));
    else rhol = fjo_linterp(ar, p->rhl_h, NR, x);
    return rhol / fjm_powi(x, p->l - 1);
}

/* pot/Potential.cpp:110-155 computeInts + 163-200 computePhil.  Returns 0, or 1
 * where the reference exits (NaN in I_ext). */
int fjo_compute_phil(const fjo_pot *p, const double *rhl, double *philP, double *PrP, double *Pr2P) {
    const int NR = p->nr, NP = p->npoly;
    const double *ar = p->ar;
    double *rhl_h = (double *)malloc(NR * sizeof(double));
    double *Iint = (double *)malloc((size_t)NR * NP * sizeof(double));
    double *Iext = (double *)malloc((size_t)NR * NP * sizeof(double));
    int rc = 0;
    for (int np = 0; np < NP; np++) {
        for (int i = 0; i < NR; i++) rhl_h[i] = rhl[i * NP + np];
        for (int nr = 0; nr < NR; nr++) {
            I_par par = {2 * np, rhl_h, ar, NR};
            double res, err;
            fjo_qags(I_internal, &par, 0., ar[nr], 1e-4, 1e-3, 1000, &res, &err, NULL, NULL);
            Iint[nr * NP + np] = res;
            double epsabs = 1e-4, epsrel = 1e-3;
            int status = 1, guard = 0;
            while (status != 0) {
                status = fjo_qagiu(I_external, &par, ar[nr], epsabs, epsrel, 1000, &res, &err, NULL, NULL);
                epsabs *= 2; epsrel *= 2;
                if (isnan(res) || ++guard > 200) { rc = 1; status = 0; }
            }
            Iext[nr * NP + np] = res;
        }
    }
    for (int np = 0; np < NP; np++) {
        int l = 2 * np;
        double G = 1;
        for (int n = 0; n < NR; n++) {
            double Ii = Iint[n * NP + np], Ie = Iext[n * NP + np];
            philP[n * NP + np] = -TPI * G * (Ii / fjm_powi(ar[n], l + 1) + Ie * fjm_powi(ar[n], l));
            PrP[n * NP + np] = -TPI * G * (-(l + 1) * Ii / fjm_powi(ar[n], l + 2) + l * Ie * fjm_powi(ar[n], l - 1));
            Pr2P[n * NP + np] = -TPI * G * ((l + 2) * (l + 1) * Ii / fjm_powi(ar[n], l + 3) +
                                            (l - 1) * l * Ie * fjm_powi(ar[n], l - 2) - (2 * l + 1) * rhl[n * NP + np]);
        }
    }
    free(rhl_h); free(Iint); free(Iext);
    return rc;
}

/* include/UtilsLeg.h:200-253 (12-argument get_Ylm).  rho row 0 is modified
 * as in the reference (monopole fill). Arrays [nr][ngauss] -> [nr][npoly]. */
void fjo_get_ylm(const fjo_pot *p, double *rho, const double *vrot, const double *sigR,
                 const double *sigp, const double *sigz, const double *sigRz,
                 double *rhL, double *vrotL, double *sigRL, double *sigpL, double *sigzL, double *sigRzL) {
    const int nr = p->nr, ng = p->ngauss, npoly = p->npoly;
    double *ci = (double *)malloc(ng * sizeof(double)), *wi = (double *)malloc(ng * sizeof(double));
    double *poly = (double *)malloc((size_t)npoly * ng * sizeof(double)), pol[64];
    fjo_gauleg(ng, 0, 1, ci, wi);
    for (int g = 0; g < ng; g++) {
        fjo_evenlegend(npoly, ci[g], pol);
        for (int np = 0; np < npoly; np++) poly[np * ng + g] = 2 * pol[np] * wi[g];
    }
    for (int j = 1; j < ng; j++) rho[j] = rho[0];
    for (int n = 0; n < nr; n++)
        for (int np = 0; np < npoly; np++) {
            int o = n * npoly + np;
            rhL[o] = 0; vrotL[o] = 0; sigRL[o] = 0; sigpL[o] = 0; sigzL[o] = 0; sigRzL[o] = 0;
        }
    for (int n = 0; n < nr; n++)
        for (int np = 0; np < npoly; np++)
            for (int g = 0; g < ng; g++) {
                int o = n * npoly + np, i = n * ng + g;
                double w = poly[np * ng + g];
                rhL[o] += w * rho[i];
                vrotL[o] += w * vrot[i];
                sigRL[o] += w * sigR[i];
                sigpL[o] += w * sigp[i];
                sigzL[o] += w * sigz[i];
                sigRzL[o] += w * sigRz[i];
            }
    rhL[0] = 2 * rho[0]; vrotL[0] = 2 * vrot[0]; sigRL[0] = 2 * sigR[0];
    sigpL[0] = 2 * sigp[0]; sigzL[0] = 2 * sigz[0]; sigRzL[0] = 2 * sigRz[0];
    for (int np = 1; np < npoly; np++) {
        rhL[np] = 0; vrotL[np] = 0; sigRL[np] = 0; sigpL[np] = 0; sigzL[np] = 0; sigRzL[np] = 0;
    }
    free(ci); free(wi); free(poly);
}

/* ------------------------------------------------------------------------ */
/* Focal distance Delta(E): delta/tabulate.cpp, delta/Delta.cpp, Utils.h     */
/* ------------------------------------------------------------------------ */
typedef struct { const fjo_pot *p; double E, Lzsq; } pELz;
static double Rcfn(double R, void *params) {   /* Utils.h:218-221 */
    pELz *par = (pELz *)params;
    return par->Lzsq / fjm_powi(R, 3) - fjo_dR(par->p, R, 0);
}
static double GetRc(double Lzsq, double R, const fjo_pot *p) {   /* Utils.h:222-241 */
    pELz par = {p, 0., Lzsq};
    double Ri = R, Ro = R, dfRi, dfRo, dfRt = Rcfn(R, &par);
    if (dfRt > 0) {
        while (dfRt > 0) { Ri = R; dfRi = dfRt; R *= 1.2; dfRt = Rcfn(R, &par); }
        Ro = R; dfRo = dfRt;
    } else {
        while (dfRt < 0) { Ro = R; dfRo = dfRt; R *= .8; dfRt = Rcfn(R, &par); }
        Ri = R; dfRi = dfRt;
    }
    return roots_known(Rcfn, &par, Ri, Ro, dfRi, dfRo, NULL);
}
static double RcEfn(double R, void *params) {  /* Utils.h:243-246 */
    pELz *par = (pELz *)params;
    return .5 * R * fjo_dR(par->p, R, 0) + fjo_phi(par->p, R, 0) - par->E;
}
static double RcE(double E, double Rc, const fjo_pot *p) {   /* Utils.h:247-263 */
    pELz par = {p, E, 0.};
    double R0 = .8 * Rc, f0 = RcEfn(R0, &par);
    while (f0 > 0) { R0 *= .8; f0 = RcEfn(R0, &par); }
    double R1 = 1.2 * Rc, f1 = RcEfn(R1, &par);
    while (f1 < 0) { R1 *= 1.2; f1 = RcEfn(R1, &par); }
    return roots_known(RcEfn, &par, R0, R1, f0, f1, NULL);
}

#define DSIZE 200
typedef struct {
    const fjo_pot *p;
    double E, Lz, R0par;           /* eqsPar */
    double Rc, R0, D2, sqDR, v0, cv, sv, Ri, zi;
    int status;
    double *Rp, *zp; int N;        /* Rzptr */
} delta_t;

static double vsq(const delta_t *d, double x) {  /* Delta.cpp:28-30 */
    return 2 * (d->E - (fjo_phi(d->p, x, 0.) + .5 * fjm_pow(d->Lz / x, 2)));
}
static void eqsMotion(double t, const double *y, double *f, void *params) {  /* Delta.cpp:33-42 */
    (void)t;
    delta_t *d = (delta_t *)params;
    f[0] = y[2];
    f[1] = y[3];
    f[2] = -fjo_dR(d->p, y[0], y[1]) + fjm_pow(d->Lz, 2.0) / fjm_pow(y[0], 3.0);
    f[3] = -fjo_dz(d->p, y[0], y[1]);
}
/* Delta.cpp:53-90; cond 0: z>=0 (Cond_zgt0), 1: pz>=0 (Cond_pzgt0) */
static int evolveOrbit(delta_t *d, double x0, double *R, double *z, int cond) {
    double steph = 2e-3, y[4], t = 0;
    y[0] = x0; y[1] = 0.; y[2] = 0.; y[3] = sqrt(vsq(d, x0));
    int i = 0;
    R[i] = y[0]; z[i] = y[1];
    while (i < DSIZE - 1 && (cond == 0 ? y[1] : y[3]) >= 0.) {
        if (fjo_evolve_apply(eqsMotion, d, 4, EPSABSODE, EPSRELODE, &t, 5 * x0 / sqrt(vsq(d, x0)), &steph, y))
            d->status |= 4;
        i++;
        R[i] = y[0]; z[i] = y[1];
    }
    return i;
}
static double getDR(delta_t *d, double x0) {   /* Delta.cpp:93-104 */
    double R[DSIZE], z[DSIZE];
    int N = evolveOrbit(d, x0, R, z, 0);
    if (fabs(z[N - 1] - z[N]) > 0 && fabs(R[N - 1] - R[N]) != 0)
        return x0 - (R[N] - z[N] * (R[N - 1] - R[N]) / (z[N - 1] - z[N]));
    else
        return x0 - R[N];
}
static double wrp_getDR(double x0, void *params) { return getDR((delta_t *)params, x0); }
static double dPhiEff(const delta_t *d, double x) { return fjo_dR(d->p, x, 0) - fjm_pow(d->Lz, 2) / fjm_powi(x, 3); }
static double ensureVsq(delta_t *d, double x0) {   /* Delta.cpp:116-124 */
    int count = 0;
    while (vsq(d, x0) < 0) {
        if (dPhiEff(d, x0) < 0) x0 *= 1.1;
        else x0 *= .9;
        count++;
        if (count > 50) { d->status |= 1; return x0; }
    }
    return x0;
}
static double getJr0(delta_t *d) {   /* Delta.cpp:129-164 */
    double x0 = d->Rc, x1, dx0, dx1, A = .8;
    int k = 0;
    x0 = ensureVsq(d, x0); dx0 = getDR(d, x0);
    x1 = ensureVsq(d, x0 / A); dx1 = getDR(d, x1);
    while (dx0 * dx1 > 0) {
        if (fabs(dx1) < fabs(dx0)) { x0 = x1; dx0 = dx1; }
        else A = fjm_pow(A, -0.9);
        x1 = ensureVsq(d, x0 * A); dx1 = getDR(d, x1);
        k++;
        if (k == 21) {
            d->status |= 2;
            if (fabs(dx0) < 1.e-7) return -x0;
            return -1;
        }
    }
    double low = MIND(x0, x1), high = MAXD(x0, x1);
    double flow = (low == x0) ? dx0 : dx1, fhigh = (high == x1) ? dx1 : dx0;
    if (x0 == x1) { flow = dx0; fhigh = dx1; }
    return roots_known(wrp_getDR, d, low, high, flow, fhigh, NULL);
}
static double ds2dv(const delta_t *d, double v) {   /* Delta.cpp:170-172 */
    return 2 * (d->zi * d->sqDR * fjm_sin(v) - fjm_cos(v) * (d->Ri * d->R0 + d->D2 * fjm_sin(v)));
}
static double wrp_ds2dv(double v, void *params) { return ds2dv((delta_t *)params, v); }
static double getv(delta_t *d) {   /* Delta.cpp:189-208 */
    double low = 0, high = PIH;
    double fl = ds2dv(d, low), fh = ds2dv(d, high);
    if (fabs(fl) < EPSABSROOTS) d->v0 = low;
    else if (fabs(fh) < EPSABSROOTS) d->v0 = high;
    else d->v0 = roots_known(wrp_ds2dv, d, low, high, fl, fh, NULL);
    d->sv = fjm_sin(d->v0); d->cv = fjm_cos(d->v0);
    return d->v0;
}
static double ds2dD2TOT(delta_t *d, double D2h) {   /* Delta.cpp:210-234 */
    d->D2 = D2h; d->sqDR = sqrt(d->D2 + d->R0 * d->R0);
    double sum = 0.;
    for (int i = 0; i < d->N; i++) {
        d->Ri = d->Rp[i]; d->zi = d->zp[i];
        getv(d);
        double sv = d->sv, cv = d->cv, zi = d->zi, Ri = d->Ri, sqDR = d->sqDR, R0 = d->R0, D2 = d->D2;
        double dvdD2 = .5 * sv * (zi / sqDR - 2 * cv) / (cv * zi * sqDR + sv * Ri * R0 - D2 * (cv * cv - sv * sv));
        sum += (cv - zi / sqDR) * cv + ds2dv(d, d->v0) * dvdD2;
    }
    return sum;
}
static double wrp_ds2dD2TOT(double D2h, void *params) { return ds2dD2TOT((delta_t *)params, D2h); }

/* Delta.cpp:18-25 (ctor) + 248-289 getBestDelta; returns Delta^2. */
double fjo_best_delta(const fjo_pot *p, double E, double Lz, double R0, int *status) {
    delta_t d;
    memset(&d, 0, sizeof d);
    d.p = p; d.E = E; d.Lz = Lz; d.R0par = R0;
    d.Rc = RcE(E, R0, p);
    d.R0 = getJr0(&d);
    double R[DSIZE], z[DSIZE];
    d.N = evolveOrbit(&d, d.R0, R, z, 1);
    d.Rp = R; d.zp = z;
    double low = -0.001, high = 5, D2out = 0;
    double Dmin = ds2dD2TOT(&d, low), Dmax = ds2dD2TOT(&d, high);
    int guard = 0;
    while (Dmin * Dmax > 0 && fabs(Dmax) > 1e-4 && fabs(Dmin) > 1e-4) {
        if (Dmax < 0) { high *= 5; Dmax = ds2dD2TOT(&d, high); }
        else { low -= .1; Dmin = ds2dD2TOT(&d, low); }
        if (++guard > 400) { d.status |= 8; break; }
    }
    if (fabs(Dmin) < EPSABSROOTS) D2out = low;
    else if (fabs(Dmax) < EPSABSROOTS) D2out = high;
    else D2out = roots_known(wrp_ds2dD2TOT, &d, low, high, Dmin, Dmax, NULL);
    if (status) *status = d.status;
    return D2out;
}

/* delta/tabulate.cpp:13-38.  Only Lzgrid[1], Rgrid[1], Ecgrid[1], Vmax[1] are
 * read by the energy loop; the remaining entries of the radius loop are pure
 * and unused, so the loop stops after i = 1. */
int fjo_tabulate_delta(fjo_pot *p) {
    const int NR = p->nr;
    double Lzgrid[2], Vmax[2], Ecgrid[2], Rgrid[2];
    double Rmax = p->ar[NR - 1];
    double Lzmax = Rmax * 0.98 * sqrt(-2 * fjo_phi(p, Rmax, 0)), Rmin = 1e-3 * Rmax;
    double Lzmin = Rmin * sqrt(Rmin * fjo_dR(p, Rmin, 0)), R = .01 * Rmax;
    for (int i = 0; i < 2; i++) {
        Lzgrid[i] = Lzmin + i * (Lzmax - Lzmin) / (double)(NR - 1);
        double Lzsq = Lzgrid[i] * Lzgrid[i];
        R = Rgrid[i] = GetRc(Lzsq, R, p);
        double vc2 = R * fjo_dR(p, R, 0);
        Ecgrid[i] = .5 * vc2 + fjo_phi(p, R, 0);
        Vmax[i] = .98 * sqrt(-2 * Ecgrid[i]);
    }
    int st = 0;
    for (int nE = 0; nE < p->ngrid; nE++) {
        const double tv = Vmax[1] * (nE + 1) / (double)p->ngrid;
        p->Egrid[nE] = Ecgrid[1] + .5 * (tv * tv);
        int s = 0;
        p->Dgrid[nE] = fjo_best_delta(p, p->Egrid[nE], .25 * Lzgrid[1], Rgrid[1], &s);
        st |= s;
    }
    return st;
}

/* delta/tabulate.cpp:40-44 */
double fjo_deltafn(const fjo_pot *p, double E) {
    const int NG = p->ngrid;
    if (E < p->Egrid[0]) return sqrt(MAXD(.000001, p->Dgrid[0]));
    else if (E < p->Egrid[NG - 1]) return sqrt(MAXD(.000001, fjo_linterp(p->Egrid, p->Dgrid, NG, E)));
    else return sqrt(MAXD(.000001, p->Dgrid[NG - 1]));
}

/* ------------------------------------------------------------------------ */
/* Stackel fudge: uvOrb/Uvpt.cpp:11-35, uvOrb/uvOrb.cpp:15-262               */
/* ------------------------------------------------------------------------ */
typedef struct {
    const fjo_pot *pot;
    double Delta, u, v, E, Er, Lz, Lzsq, Iv;
    double uinner, uouter, umid, vturn;
    double shu2, sh1sq, ch1sq, Phiu1, ubar, Deltau;
    double sv, cv, sv2;
} uvorb_t;

static double Phiu(uvorb_t *o, double u) {
    double R = o->Delta * fjm_sinh(u) * fjm_sin(o->v), z = o->Delta * fjm_cosh(u) * fjm_cos(o->v);
    return fjo_phi(o->pot, R, z);
}
static double Phiv(uvorb_t *o, double v) {
    double R = o->Delta * fjm_sinh(o->u) * fjm_sin(v), z = o->Delta * fjm_cosh(o->u) * fjm_cos(v);
    return fjo_phi(o->pot, R, z);
}
static double dV(uvorb_t *o, double v) {   /* uvOrb.cpp:46-48 */
    return (1 + o->shu2) * Phiv(o, PIH) - (o->shu2 + fjm_sin(v) * fjm_sin(v)) * Phiv(o, v);
}
static double vturnfn(uvorb_t *o, double v) {   /* uvOrb.cpp:49-52 */
    double sv2 = fjm_sin(v) * fjm_sin(v);
    return o->E * sv2 + o->Iv + dV(o, v) - .5 * o->Lzsq / (o->Delta * o->Delta * sv2);
}
static double Jvint(double psi, void *prm) {   /* uvOrb.cpp:53-56 */
    uvorb_t *o = (uvorb_t *)prm;
    double v = PIH - (PIH - o->vturn) * fjm_cos(psi);
    return sqrt(MAXD(0., vturnfn(o, v))) * fjm_sin(psi);
}
static double dU(uvorb_t *o, double u) {   /* uvOrb.cpp:61-63 */
    return (fjm_sinh(u) * fjm_sinh(u) + o->sv2) * Phiu(o, u) - (o->sh1sq + o->sv2) * o->Phiu1;
}
static double dPhiu(uvorb_t *o, double u, double *f) {   /* uvOrb.cpp:64-69 */
    double R = o->Delta * fjm_sinh(u) * fjm_sin(o->v), z = o->Delta * fjm_cosh(u) * fjm_cos(o->v);
    f[0] = fjo_dR(o->pot, R, z);
    f[1] = fjo_dz(o->pot, R, z);
    return fjo_phi(o->pot, R, z);
}
static double uturnfn(double u, void *prm) {   /* uvOrb.cpp:70-73 */
    uvorb_t *o = (uvorb_t *)prm;
    double shu2 = fjm_sinh(u) * fjm_sinh(u);
    return o->E * (shu2 - o->sh1sq) + o->Er * o->ch1sq - dU(o, u) -
           .5 * o->Lzsq / (o->Delta * o->Delta) * (1 / shu2 - 1 / o->sh1sq);
}
static double duturnfn(double u, void *prm) {   /* uvOrb.cpp:74-79 */
    uvorb_t *o = (uvorb_t *)prm;
    double shu = fjm_sinh(u), chu = fjm_cosh(u), shu2 = shu * shu, f[2];
    double Phi0 = dPhiu(o, u, f);
    double dPhidu = o->Delta * (f[0] * chu * o->sv + f[1] * shu * o->cv);
    return 2 * (o->E - Phi0) * shu * chu - (shu2 + o->sv2) * dPhidu +
           o->Lzsq * chu / (o->Delta * o->Delta * shu2 * shu);
}
static double Juint(double psi, void *prm) {   /* uvOrb.cpp:80-83 */
    uvorb_t *o = (uvorb_t *)prm;
    double u = o->ubar - o->Deltau * fjm_cos(psi);
    return sqrt(MAXD(0., uturnfn(u, o))) * fjm_sin(psi);
}
static double wrp_vturnfn(double v, void *prm) { return vturnfn((uvorb_t *)prm, v); }

static void Getu0(uvorb_t *o, double utry) {   /* uvOrb.cpp:109-129 */
    double ui = utry, uo = utry, dfui = 0, dfuo = 0, dut = duturnfn(utry, o);
    if (dut > 0) {
        while (dut > 0) { ui = utry; dfui = dut; utry *= 1.2; dut = duturnfn(utry, o); }
        uo = utry; dfuo = dut;
    } else {
        while (dut < 0) { uo = utry; dfuo = dut; utry *= .8; dut = duturnfn(utry, o); }
        ui = utry; dfui = dut;
    }
    o->umid = roots_known(duturnfn, o, ui, uo, dfui, dfuo, NULL);
}

static void GetTurnu(uvorb_t *o, double utry) {   /* uvOrb.cpp:131-201 */
    double u1 = 0, u2 = 0, fu1 = 0, fu2 = 0, fuinner = 0, fuouter = 0;
    double ut = uturnfn(utry, o), dut = duturnfn(utry, o), eps = .1;
    if (ut < 0) {
        if (dut > 0) {
            o->uinner = utry; fuinner = ut;
            while (1) {
                while (ut < 0 && dut > 0) {
                    o->uinner = utry; fuinner = ut; utry *= (1 + eps);
                    ut = uturnfn(utry, o); dut = duturnfn(utry, o);
                }
                if (ut >= 0) { u1 = utry; fu1 = ut; break; }
                else {
                    utry /= (1 + eps); eps *= .2;
                    if (eps < TINY) { o->uinner = o->uouter = utry; return; }
                    utry *= (1 + eps);
                    ut = uturnfn(utry, o); dut = duturnfn(utry, o);
                }
            }
            while (ut > 0) { u2 = utry; fu2 = ut; utry *= 1.1; ut = uturnfn(utry, o); }
            o->uouter = utry; fuouter = ut;
        } else {
            while (1) {
                while (ut < 0 && dut < 0) {
                    o->uouter = utry; fuouter = ut; utry /= (1 + eps);
                    ut = uturnfn(utry, o); dut = duturnfn(utry, o);
                }
                if (ut >= 0) { u2 = utry; fu2 = ut; break; }
                else {
                    utry *= (1 + eps); eps *= .2;
                    if (eps < TINY) { o->uinner = o->uouter = utry; return; }
                    utry /= (1 + eps);
                    ut = uturnfn(utry, o); dut = duturnfn(utry, o);
                }
            }
            while (ut > 0) { u1 = utry; fu1 = ut; utry *= .9; ut = uturnfn(utry, o); }
            o->uinner = utry; fuinner = ut;
        }
    } else {
        u1 = u2 = utry; fu1 = fu2 = ut;
        while (ut > 0) { u2 = utry; fu2 = ut; utry *= 1.1; ut = uturnfn(utry, o); }
        o->uouter = utry; fuouter = ut;
        ut = fu1;
        while (ut > 0) { u1 = utry; fu1 = ut; utry *= .9; ut = uturnfn(utry, o); }
        o->uinner = utry; fuinner = ut;
    }
    /* The wrapper re-evaluates F at both ends.  In the allowed-region branch
     * the first pass of the inward walk pairs u1 = uouter with the stale
     * value fu1, so the end values are evaluated afresh here, as there. */
    (void)fuinner; (void)fuouter; (void)fu1; (void)fu2;
    double uin = o->uinner, uout = o->uouter;
    o->uinner = fjo_roots(uturnfn, o, uin, u1, NULL);
    o->uouter = fjo_roots(uturnfn, o, u2, uout, NULL);
}

static double GetTurnv(uvorb_t *o, double vtry) {   /* uvOrb.cpp:203-224 */
    double vtop = vtry, vbot = vtry, ft = 0, fb = 0;
    double f0 = vturnfn(o, vtry);
    if (f0 < 0) {
        fb = f0; ft = f0;
        while (fb < 0) {
            vbot = .5 * (PIH + vbot);
            if (PIH - vbot < TINY) return PIH;
            fb = vturnfn(o, vbot);
        }
    } else {
        ft = f0; fb = f0;
        while (ft > 0) {
            vtop *= .5;
            if (vtop < TINY) return 0;
            ft = vturnfn(o, vtop);
        }
    }
    if (!(vbot < vtop)) return roots_known(wrp_vturnfn, o, vtop, vbot, ft, fb, NULL);
    return roots_known(wrp_vturnfn, o, vbot, vtop, fb, ft, NULL);
}

/* uvOrb.cpp:15-36 + Ju 226-245 + Jv 247-262 */
/* state != NULL: also returns the orbit state the integrands read (15 doubles:
   Delta E Er ch1sq sh1sq Phiu1 Lzsq u v sv2 shu2 Iv ubar Deltau vturn) and in
   qk[0], qk[1] the first QK21 rule of the Ju and Jv integrals (NAN where that
   action is zero without an integral). */
static int actions_impl(const fjo_pot *p, double Delta, double R, double z, double Phi0,
                        double vR, double vphi, double vz, double *Jr, double *Jz,
                        double *state, double *qk, double *iv, int cap, int *nrule) {
    uvorb_t o;
    memset(&o, 0, sizeof o);
    o.pot = p;
    double Lz = R * vphi;
    o.Lz = Lz; o.Lzsq = Lz * Lz; o.Delta = Delta;
    /* pt(Delta, x, p): Uvpt.cpp:11-24 */
    double Delta2 = Delta * Delta, R2 = R * R, z2 = z * z, r2 = R2 + z2;
    double shu2 = .5 * (r2 - Delta2 + sqrt((Delta2 - r2) * (Delta2 - r2) + 4 * Delta2 * R2)) / Delta2;
    double shu = sqrt(shu2), u = fjm_asinh(shu);
    double chu2 = 1 + shu2, chu = sqrt(chu2);
    double cv2 = .5 * (r2 + Delta2 - sqrt((Delta2 + r2) * (Delta2 + r2) - 4 * Delta2 * z2)) / Delta2;
    double cv = z < 0 ? -sqrt(cv2) : sqrt(cv2);
    double v = fjm_acos(cv), sv = fjm_sin(v), sv2 = 1 - cv2;
    double pR = vR, pz = vz, p2 = pR * pR + pz * pz;
    double pu = Delta * (pR * chu * sv + pz * shu * cv);
    double pv = Delta * (pR * shu * cv - pz * chu * sv);
    (void)chu2;
    o.u = u; o.v = v; o.shu2 = shu2; o.sv = sv; o.cv = cv; o.sv2 = sv2;
    o.E = .5 * p2 + .5 * o.Lzsq / (R * R) + Phi0;
    if (o.E >= 0 || isnan(o.E)) { *Jr = 1e31; *Jz = 1e31; g_cnt.unbound_nodes++; return 1; }
    g_cnt.bound_nodes++;
    Getu0(&o, 58);
    o.sh1sq = fjm_sinh(o.umid) * fjm_sinh(o.umid); o.ch1sq = fjm_cosh(o.umid) * fjm_cosh(o.umid);
    o.Phiu1 = Phiu(&o, o.umid);
    o.Er = -o.E * (shu2 - o.sh1sq) / o.ch1sq + .5 * (pu * pu) / (Delta2 * o.ch1sq) +
           .5 * o.Lzsq / (Delta2 * o.ch1sq) * (1 / shu2 - 1 / o.sh1sq) + dU(&o, u) / o.ch1sq;
    o.Iv = .5 * (pv * pv) / Delta2 - o.E * sv2 + .5 * o.Lzsq / (Delta2 * sv2) - dV(&o, v);
    const double rtwo = sqrt(2);
    if (iv) { g_rec = iv; g_rec_n = 0; g_rec_cap = cap; nrule[0] = nrule[1] = 0; }
    /* GetJu */
    GetTurnu(&o, 1);
    if (fabs(o.uouter - o.uinner) < TINY) *Jr = 0;
    else {
        o.ubar = .5 * (o.uinner + o.uouter);
        o.Deltau = .5 * (o.uouter - o.uinner);
        double res, err;
        fjo_qags(Juint, &o, 0, PI, EPSABSINTEG, EPSRELINTEG, WRKSP, &res, &err, NULL, NULL);
        *Jr = rtwo * Delta * o.Deltau * res / PI;
    }
    if (iv) nrule[0] = g_rec_n;
    /* GetJv */
    o.vturn = GetTurnv(&o, .5 * PIH);
    if (fabs(o.vturn - PIH) < TINY) *Jz = 0;
    else {
        double res, err;
        fjo_qags(Jvint, &o, 0, PIH, EPSABSINTEG, EPSRELINTEG, WRKSP, &res, &err, NULL, NULL);
        *Jz = 2 * rtwo * Delta * (PIH - o.vturn) * res / PI;
    }
    if (iv) { nrule[1] = g_rec_n - nrule[0]; g_rec = NULL; }
    if (state) {
        const double st[15] = {o.Delta, o.E, o.Er, o.ch1sq, o.sh1sq, o.Phiu1, o.Lzsq, o.u, o.v,
                               o.sv2, o.shu2, o.Iv, o.ubar, o.Deltau, o.vturn};
        memcpy(state, st, sizeof st);
        double e, ra, rs;
        qk[0] = qk[1] = NAN;
        if (!(fabs(o.uouter - o.uinner) < TINY)) qk_rule(11, xgk21, wg10, wgk21, Juint, &o, 0, PI, &qk[0], &e, &ra, &rs);
        if (!(fabs(o.vturn - PIH) < TINY)) qk_rule(11, xgk21, wg10, wgk21, Jvint, &o, 0, PIH, &qk[1], &e, &ra, &rs);
    }
    return 0;
}

int fjo_actions(const fjo_pot *p, double Delta, double R, double z, double Phi0,
                double vR, double vphi, double vz, double *Jr, double *Jz) {
    return actions_impl(p, Delta, R, z, Phi0, vR, vphi, vz, Jr, Jz, NULL, NULL, NULL, 0, NULL);
}

int fjo_orbit_state(const fjo_pot *p, double Delta, double R, double z, double Phi0,
                    double vR, double vphi, double vz, double *state, double *qk,
                    double *iv, int cap, int *nrule) {
    double Jr, Jz;
    return actions_impl(p, Delta, R, z, Phi0, vR, vphi, vz, &Jr, &Jz, state, qk, iv, cap, nrule);
}

/* df/DF.cpp:65-76 */
double fjo_df_hg(const fjo_df *d, double Jr, double Jphi, double Jz) {
    double h = Jr + d->dphi_h * fabs(Jphi) + d->dz_h * Jz;
    double g = Jr + d->dphi_g * fabs(Jphi) + d->dz_g * Jz;
    return d->mass / fjm_powi(d->J0, 3) * MAXD(0., fjm_pow(1. + d->J0 / h, d->A) / (fjm_pow(1. + g / d->J0, d->B))) /
           (fjm_powi(TPI, 3));
}

/* df/DF.cpp:78-124 with the double-power-law branch (108-123) live. */
double fjo_df_eval(const fjo_pot *p, const fjo_df *d, const double *x, const double *v) {
    double R = x[0], Lz = R * v[1], Phi_h = x[2];
    double H = .5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + Phi_h;
    g_cnt.nodes++;
    if (H >= 0) return 0.;
    double Jr, Jz, Jphi = Lz;
    fjo_actions(p, fjo_deltafn(p, H), x[0], x[1], Phi_h, v[0], v[1], v[2], &Jr, &Jz);
    double DFeven = fjo_df_hg(d, Jr, Jphi, Jz);
    if (d->chi == 0.) return DFeven;
    double t = fjm_tanh(d->chi * Jphi / d->J0);
    if (d->rot_form == 1) return (1. - d->k) * DFeven + d->k * t * DFeven;
    return DFeven + d->k * t * DFeven;
}

/* integ/Integ.cpp:150-163 (8-argument rhofDF) with Int3D_111111_vec
 * (GaussQuad.cpp:160-180) / Int3D_011101_vec (137-157) and rhoInteg_vec
 * (Integ.cpp:46-62); the 6 sums run in serial (i,j,k) order. */
double fjo_rhofDF(const fjo_pot *p, const fjo_df *d, double R, double z, double *out,
                  double *vrot, double *sigR, double *sigp, double *sigz, double *sigRz) {
    double Phi_h = fjo_phi(p, R, z);
    double far = 10 * p->ar[p->nr - 1];
    double Ve = sqrt(-2 * (Phi_h - fjo_phi(p, far, far)));
    double x[3] = {R, z, Phi_h};
    const double *xa = p->quad ? xpt01 : xpt11, *wa = p->quad ? xw01 : xw11;
    /* Summation order of the GPU reduction: node n goes to partial sum n mod
     * 256 in increasing n, then the 256 partials are added pairwise (stride
     * 128, 64, ..., 1).  The reference leaves the order to the OpenMP
     * reduction (GaussQuad.cpp:164). */
    double part[256][6];
    memset(part, 0, sizeof part);
    for (int i = 0; i < 22; i++)
        for (int j = 0; j < 22; j++)
            for (int k = 0; k < 22; k++) {
                const int node = (i * 22 + j) * 22 + k;
                double V[3] = {1e-6 + Ve * xa[i], 1e-6 + Ve * xpt11[j], 1e-6 + Ve * xa[k]};
                double jacob = Ve * Ve * Ve;
                double DF = fjo_df_eval(p, d, x, V);
                double w = wa[i] * xw11[j] * wa[k];
                double r0 = jacob * DF, r1 = jacob * DF * V[1], r2 = jacob * DF * V[0] * V[0],
                       r3 = jacob * DF * V[1] * V[1], r4 = jacob * DF * V[2] * V[2],
                       r5 = jacob * DF * V[0] * V[2];
                double *s = part[node & 255];
                s[0] += w * r0; s[1] += w * r1; s[2] += w * r2;
                s[3] += w * r3; s[4] += w * r4; s[5] += w * r5;
            }
    for (int st = 128; st > 0; st >>= 1)
        for (int t = 0; t < st; t++)
            for (int m = 0; m < 6; m++) part[t][m] += part[t + st][m];
    double *s = part[0];
    for (int m = 0; m < 6; m++) out[m] = s[m];
    *vrot = out[1] / out[0];
    *sigR = sqrt(out[2] / out[0]); *sigp = sqrt(out[3] / out[0]); *sigz = sqrt(out[4] / out[0]);
    *sigRz = sqrt(fabs(out[5]) / out[0]);
    return p->quad ? 4 * out[0] : out[0];
}

/* rhofDF with the reference's own parallel structure (GaussQuad.cpp:160-180):
 * one OpenMP team over the 22^3 nodes of the point, collapse(3), static
 * schedule, reduction(+) on the six sums (their order left to OpenMP).
 * For the CPU timing baseline only; the sums are not fixed-order. */
static double rhofDF_teams(const fjo_pot *p, const fjo_df *d, double R, double z, int nthreads) {
    double Phi_h = fjo_phi(p, R, z);
    double far = 10 * p->ar[p->nr - 1];
    double Ve = sqrt(-2 * (Phi_h - fjo_phi(p, far, far)));
    double x[3] = {R, z, Phi_h};
    const double *xa = p->quad ? xpt01 : xpt11, *wa = p->quad ? xw01 : xw11;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
#ifdef _OPENMP
#pragma omp parallel for collapse(3) schedule(static) num_threads(nthreads) reduction(+ : s0, s1, s2, s3, s4, s5)
#endif
    for (int i = 0; i < 22; i++)
        for (int j = 0; j < 22; j++)
            for (int k = 0; k < 22; k++) {
                double V[3] = {1e-6 + Ve * xa[i], 1e-6 + Ve * xpt11[j], 1e-6 + Ve * xa[k]};
                double jd = Ve * Ve * Ve * fjo_df_eval(p, d, x, V);
                double w = wa[i] * xw11[j] * wa[k];
                s0 += w * jd; s1 += w * jd * V[1]; s2 += w * jd * V[0] * V[0];
                s3 += w * jd * V[1] * V[1]; s4 += w * jd * V[2] * V[2]; s5 += w * jd * V[0] * V[2];
            }
    (void)nthreads;
    return s0 + s1 + s2 + s3 + s4 + s5;
}

/* CPU timing baseline of the grid loop of computeRhl (potLeg.cpp:54-58) on
 * every stride-th grid point: variant 0 grid-parallel (one point per thread,
 * dynamic schedule), variant 1 reference-structured (serial loop over the
 * points, one OpenMP team over the nodes of each).  Returns the wall time in
 * seconds; *npts the number of points. */
double fjo_time_grid(const fjo_pot *p, const fjo_df *d, int nthreads, int variant, int stride, int *npts) {
    const int NR = p->nr, NG = p->ngauss;
    double *ci = (double *)malloc(NG * sizeof(double)), *wi = (double *)malloc(NG * sizeof(double));
    fjo_gauleg(NG, 0, 1, ci, wi);
    const int n = (NR * NG + stride - 1) / stride;
    double *res = (double *)malloc(n * sizeof(double));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const double t0 = omp_get_wtime();
#endif
    if (variant == 0) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
        for (int k = 0; k < n; k++) {
            const int idx = k * stride, nr = idx / NG, ng = idx % NG;
            double out[6], a, b, c, e, f;
            const double sg = sqrt(1 - ci[ng] * ci[ng]);
            res[k] = fjo_rhofDF(p, d, p->ar[nr] * sg, p->ar[nr] * ci[ng], out, &a, &b, &c, &e, &f);
        }
    } else {
        for (int k = 0; k < n; k++) {
            const int idx = k * stride, nr = idx / NG, ng = idx % NG;
            const double sg = sqrt(1 - ci[ng] * ci[ng]);
            res[k] = rhofDF_teams(p, d, p->ar[nr] * sg, p->ar[nr] * ci[ng], nthreads);
        }
    }
    double dt = 0;
#ifdef _OPENMP
    dt = omp_get_wtime() - t0;
#endif
    *npts = n;
    free(ci); free(wi); free(res);
    return dt;
}

/* pot/potLeg.cpp:35-58: the grid loop, without the rho[0][0] fix. */
void fjo_compute_rhl_grid(const fjo_pot *p, const fjo_df *d, int nthreads,
                          double *rho, double *vrot, double *sigR, double *sigp, double *sigz, double *sigRz) {
    const int NR = p->nr, NG = p->ngauss;
    double *ci = (double *)malloc(NG * sizeof(double)), *wi = (double *)malloc(NG * sizeof(double));
    double *si = (double *)malloc(NG * sizeof(double));
    fjo_gauleg(NG, 0, 1, ci, wi);
    for (int i = 0; i < NG; i++) si[i] = sqrt(1 - ci[i] * ci[i]);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int idx = 0; idx < NR * NG; idx++) {
        int nr = idx / NG, ng = idx % NG;
        double out[6];
        rho[idx] = fjo_rhofDF(p, d, p->ar[nr] * si[ng], p->ar[nr] * ci[ng], out, vrot + idx, sigR + idx,
                              sigp + idx, sigz + idx, sigRz + idx);
    }
    (void)nthreads;
    free(ci); free(wi); free(si);
}

/* pot/potLeg.cpp:76-82 computeNewPhi = computeRhl (35-70) + computePhil.
 * The potential tables of p are not modified; the caller updates them. */
int fjo_compute_new_phi(fjo_pot *p, const fjo_df *d, int nthreads,
                        double *rhl, double *vrotl, double *sigRl, double *sigpl,
                        double *sigzl, double *sigRzl, double *philP, double *PrP, double *Pr2P) {
    const int NR = p->nr, NG = p->ngauss;
    size_t n = (size_t)NR * NG;
    double *rho = (double *)malloc(6 * n * sizeof(double));
    double *vrot = rho + n, *sigR = rho + 2 * n, *sigp = rho + 3 * n, *sigz = rho + 4 * n, *sigRz = rho + 5 * n;
    fjo_compute_rhl_grid(p, d, nthreads, rho, vrot, sigR, sigp, sigz, sigRz);
    const double *ar = p->ar;
    rho[0] = (rho[2 * NG] - rho[1 * NG]) / (ar[2] - ar[1]) * (ar[0] - ar[1]) + rho[1 * NG];
    fjo_get_ylm(p, rho, vrot, sigR, sigp, sigz, sigRz, rhl, vrotl, sigRl, sigpl, sigzl, sigRzl);
    int rc = fjo_compute_phil(p, rhl, philP, PrP, Pr2P);
    free(rho);
    return rc;
}

/* include/UtilsLeg.h:300-321 int_dens (5-table, float weight) */
static void int_dens5(const fjo_pot *p, double r, const double *rhl, const double *sigRl,
                      const double *sigpl, const double *sigzl, const double *vrotl,
                      double *densp, double *Vrotp, double *sigup, double *sigpp, double *sigvp) {
    const int NR = p->nr, NP = p->npoly;
    if (r > p->ar[NR - 1]) {
        for (int i = 0; i < NP; i++) { densp[i] = 0; Vrotp[i] = 0; sigup[i] = 0; sigpp[i] = 0; sigvp[i] = 0; }
    } else {
        int top = 0, bot = 0;
        topbottom(p->ar, NR, r, &bot, &top);
        float f = (r - p->ar[bot]) / (p->ar[top] - p->ar[bot]);
        for (int i = 0; i < NP; i++) {
            int t = top * NP + i, b = bot * NP + i;
            densp[i] = f * rhl[t] + (1 - f) * rhl[b];
            sigup[i] = f * sigRl[t] + (1 - f) * sigRl[b];
            sigpp[i] = f * sigpl[t] + (1 - f) * sigpl[b];
            sigvp[i] = f * sigzl[t] + (1 - f) * sigzl[b];
            Vrotp[i] = f * vrotl[t] + (1 - f) * vrotl[b];
        }
    }
}
/* include/UtilsLeg.h:285-298 int_dens (1 table, double weight) */
static void int_dens1(const fjo_pot *p, double r, const double *rhl, double *densp) {
    const int NR = p->nr, NP = p->npoly;
    if (r > p->ar[NR - 1]) {
        for (int i = 0; i < NP; i++) densp[i] = 0;
    } else {
        int top = 0, bot = 0;
        topbottom(p->ar, NR, r, &bot, &top);
        double f = (r - p->ar[bot]) / (p->ar[top] - p->ar[bot]);
        for (int i = 0; i < NP; i++) densp[i] = f * rhl[top * NP + i] + (1 - f) * rhl[bot * NP + i];
    }
}

/* libs/Stats.cpp:12-41 with ev_dens (UtilsLeg.h:323-354) */
void fjo_vir2(const fjo_pot *p, const double *rhl, const double *vrotl, const double *sigRl,
              const double *sigpl, const double *sigzl, double *out) {
    const int NR = p->nr, NG = p->ngauss, NP = p->npoly;
    const double *ar = p->ar;
    double *ci = (double *)malloc(NG * sizeof(double)), *wi = (double *)malloc(NG * sizeof(double));
    fjo_gauleg(NG, 0, 1, ci, wi);
    double pot = 0, KRR = 0, Kzz = 0, WRR = 0, Wzz = 0;
    double pol[64], densp[64], Vrotp[64], sigup[64], sigpp[64], sigvp[64];
    for (int i = 1; i < NR; i++) {
        double r = .5 * (ar[i - 1] + ar[i]), dr = ar[i] - ar[i - 1];
        for (int j = 0; j < NG; j++) {
            double sij = sqrt(1 - ci[j] * ci[j]);
            double z = r * ci[j], R = r * sij;
            /* ev_dens (7-argument) */
            double c = z / sqrt(R * R + z * z), rr = sqrt(R * R + z * z);
            fjo_evenlegend(NP, c, pol);
            int_dens5(p, rr, rhl, sigRl, sigpl, sigzl, vrotl, densp, Vrotp, sigup, sigpp, sigvp);
            double dens = .5 * densp[0];
            double Vrot = .5 * Vrotp[0], Sigu = .5 * sigup[0], Sigp = .5 * sigpp[0], Sigv = .5 * sigvp[0];
            for (int np = 1; np < NP; np++) {
                double fac = .5 * (4 * np + 1);
                dens += fac * densp[np] * pol[np]; Sigu += fac * sigup[np] * pol[np];
                Sigp += fac * sigpp[np] * pol[np]; Sigv += fac * sigvp[np] * pol[np];
                Vrot += fac * Vrotp[np] * pol[np];
            }
            (void)Vrot;
            pot += .5 * wi[j] * dens * fjo_phi(p, R, z) * r * r * dr;
            KRR += .5 * wi[j] * dens * ((r * Sigu) * (r * Sigu) + (r * Sigp) * (r * Sigp)) * dr;
            Kzz += .5 * wi[j] * dens * ((r * Sigv) * (r * Sigv)) * dr;
            WRR -= wi[j] * dens * R * fjo_dR(p, R, z) * r * r * dr;
            Wzz -= wi[j] * dens * z * fjo_dz(p, R, z) * r * r * dr;
        }
    }
    KRR *= 4 * PI; Kzz *= 4 * PI; WRR *= 4 * PI; Wzz *= 4 * PI; pot *= 4 * PI;
    double sum = 0.;
    for (int i = 0; i < NR; i++) {
        /* ev_dens (1-table) at (ar[i], 0) */
        double R = ar[i], z = 0;
        double c = z / sqrt(R * R + z * z), rr = sqrt(R * R + z * z);
        fjo_evenlegend(NP, c, pol);
        int_dens1(p, rr, rhl, densp);
        double dens = .5 * densp[0];
        for (int np = 1; np < NP; np++) dens += .5 * (4 * np + 1) * densp[np] * pol[np];
        sum += 2 * TPI * ar[i] * ar[i] * dens;
    }
    out[0] = ar[NR - 1] * ar[NR - 1] * p->Pr[(NR - 1) * NP + 0];
    out[1] = sum;
    out[2] = KRR + Kzz; out[3] = pot; out[4] = pot / (KRR + Kzz);
    out[5] = KRR; out[6] = WRR; out[7] = WRR / KRR;
    out[8] = Kzz; out[9] = Wzz; out[10] = Wzz / Kzz;
    free(ci); free(wi);
}

/* ---- observables (integ/Integ.cpp) --------------------------------------- */

/* ev_dens (include/UtilsLeg.h:323-334) with the one-table int_dens. */
static double ev_dens1(const fjo_pot *p, const double *rhl, double R, double z) {
    const int NP = p->npoly;
    double pol[64], densp[64];
    double c = z / sqrt(R * R + z * z), r = sqrt(R * R + z * z);
    fjo_evenlegend(NP, c, pol);
    int_dens1(p, r, rhl, densp);
    double dens = .5 * densp[0];
    for (int np = 1; np < NP; np++) {
        double fac = .5 * (4 * np + 1);
        dens += fac * densp[np] * pol[np];
    }
    return dens;
}

/* SigmaDF (Integ.cpp:168-172) = 2 Int1D_01 (GaussQuad.cpp:31-39) of
 * SigmaInteg (Integ.cpp:33-43): r from R to ar[nr-1]. */
double fjo_sigma_df(const fjo_pot *p, const double *rhl, double R, double z) {
    const double rmax = p->ar[p->nr - 1];
    double sum = 0.;
    for (int i = 0; i < 22; i++) {
        double r = R + xpt01[i] * (rmax - R);
        double jacob = rmax - R;
        sum += xw01[i] * (jacob * ev_dens1(p, rhl, r, z));
    }
    return 2 * sum;
}

static void phi_ve(const fjo_pot *p, double R, double z, double *Phi, double *Ve) {
    double far = 10 * p->ar[p->nr - 1];
    *Phi = fjo_phi(p, R, z);
    *Ve = sqrt(-2 * (*Phi - fjo_phi(p, far, far)));
}

/* line_profile (Integ.cpp:177-190): Int3D_111111 (GaussQuad.cpp:124-134) of
 * lineProfile (Integ.cpp:64-92) in serial (i, j, k) order, over SigmaDF. */
double fjo_line_profile(const fjo_pot *p, const fjo_df *d, const double *rhl, double R, double z, double vx) {
    double Phi, Ve;
    phi_ve(p, R, z, &Phi, &Ve);
    double Sigma = fjo_sigma_df(p, rhl, R, z);
    double x[3] = {R, z, Phi}, sum = 0;
    for (int i = 0; i < 22; i++)
        for (int j = 0; j < 22; j++)
            for (int k = 0; k < 22; k++) {
                double V[3], jacob;
                fjob_lp_node(R, Ve, p->ar[p->nr - 1], vx, xpt11[i], xpt11[j], xpt11[k], V, &jacob);
                sum += xw11[i] * xw11[j] * xw11[k] * (jacob * fjo_df_eval(p, d, x, V));
            }
    return sum / Sigma;
}

/* P_vp / P_vz / P_vr (Integ.cpp:191-220): Int2D_0101 / Int2D_0111
 * (GaussQuad.cpp:52-71) of Pvphi / Pvz / Pvr (Integ.cpp:97-138), serial
 * (i, j) order; which = 0, 1, 2. */
double fjo_vdist(const fjo_pot *p, const fjo_df *d, int which, double R, double z, double v) {
    double Phi, Ve;
    phi_ve(p, R, z, &Phi, &Ve);
    const double *xb = which == 0 ? xpt01 : xpt11, *wb = which == 0 ? xw01 : xw11;
    double x[3] = {R, z, Phi}, sum = 0;
    for (int i = 0; i < 22; i++)
        for (int j = 0; j < 22; j++) {
            double V[3];
            fjob_pv_node(which, Ve, v, xpt01[i], xb[j], V);
            sum += xw01[i] * wb[j] * (Ve * Ve * fjo_df_eval(p, d, x, V));
        }
    return sum;
}

/* The velocity rules of include/GaussPts.h:76-136 (xpt11, xw11, xpt01, xw01). */
void fjo_quad_tables(double *x11, double *w11, double *x01, double *w01) {
    memcpy(x11, xpt11, sizeof xpt11); memcpy(w11, xw11, sizeof xw11);
    memcpy(x01, xpt01, sizeof xpt01); memcpy(w01, xw01, sizeof xw01);
}
