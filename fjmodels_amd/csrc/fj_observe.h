/*
 * fj_observe.h -- quadrature nodes of the observables of integ/Integ.cpp:
 * line profiles (lineProfile, 64-92) and velocity distributions (Pvphi,
 * Pvz, Pvr, 97-138).  The velocity at a node and its Jacobian factor are
 * formed here, in C that the library's host code and the CPU oracle both
 * compile, so that both evaluate f(J) at the same bits.
 */
#ifndef FJ_OBSERVE_H
#define FJ_OBSERVE_H

#include "fj_math.h"

/* lineProfile at node (vy, x, vz) of [-1,1]^3 for the grid point (R, z)
 * with escape speed Ve, line-of-sight velocity vx and outer radius armax:
 * the velocity V (with the 1e-6 offsets of Integ.cpp:83) and the factor
 * Ve^2 |x| armax^2 / R_los.  f is evaluated at the grid point itself, as in
 * the reference (df(Rz, V) with the Rz of line_profile). */
FJM_FN void fjob_lp_node(double R, double Ve, double armax, double vx, double vy, double x, double vz,
                         double *V, double *jacob) {
    const double xl = x * armax;
    const double Rl = sqrt(R * R + xl * xl);
    const double y = R;
    const double v_abs = sqrt(vx * vx + vy * vy);
    const double alpha = fjm_acos(vx / v_abs), theta = fjm_atan2(y, x);
    double s, c;
    fjm_sincos(alpha + theta, &s, &c);
    const double vr = v_abs * c, vphi = v_abs * s;
    V[0] = 1e-6 + Ve * vr;
    V[1] = 1e-6 + Ve * vphi;
    V[2] = 1e-6 + Ve * vz;
    *jacob = Ve * Ve * (fabs(x) * (armax * armax)) / Rl;
}

/* Pvphi (which 0), Pvz (1), Pvr (2) at node (a, b) for the fixed velocity
 * component v: Pvphi (a, b) = (vr, vz) in [0,1]^2; Pvz (vr, vphi) and Pvr
 * (vz, vphi) in [0,1] x [-1,1].  The factor is Ve^2. */
FJM_FN void fjob_pv_node(int which, double Ve, double v, double a, double b, double *V) {
    if (which == 0) {
        V[0] = 1e-6 + Ve * a; V[1] = v; V[2] = 1e-6 + Ve * b;
    } else if (which == 1) {
        V[0] = 1e-6 + Ve * a; V[1] = 1e-6 + Ve * b; V[2] = v;
    } else {
        V[0] = v; V[1] = 1e-6 + Ve * b; V[2] = 1e-6 + Ve * a;
    }
}

#endif
