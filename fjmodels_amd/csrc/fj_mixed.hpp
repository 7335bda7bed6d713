// fj_mixed.hpp -- FP32 pieces of the mixed-precision node phase.
//
// In the mixed mode the action integrands (Juint, Jvint: uvOrb.cpp:53-56, 80-83,
// with the potential of Potential.cpp:206-219), their Kronrod rules and the
// QAGS bookkeeping (QUADPACK's single-precision qk21 and qagse: the float
// instance of the stepper in fj_device.hpp) and the DF (DF.cpp:65-119) run in
// FP32 with the GPU's single-precision elementary functions; the turning
// points, Delta, the action prefactor and everything of the grid phase stay
// FP64, and integrals that pass the FP32 interval cap are redone in FP64.
// These functions are device-only: the FP64 path (fj_device.hpp,
// fj_node.hpp) is the one the oracle reproduces.
#pragma once

#include "fj_device.hpp"
#include "fj_node.hpp"
#include "fj_fast.hpp"

namespace fj {

#pragma clang fp contract(fast)

// Potential tables in FP32 (LDS).  Even Legendre step as
// P_2np = (la[np] P_2np-4 + P_2np-2 (c^2 - lb[np])) lg[np]  (UtilsLeg.h:68-79).
// The monopole column is held as a float pair: phil[n][0] + phil0lo[n] is the
// FP64 value to 2^-48, so that a potential difference Phi - c with c close to
// Phi keeps FP32 relative accuracy.
struct PotView32 {
    int nr, npoly;
    const float *ar, *phil, *phil0lo;
    const float *la, *lb, *lg;
    const int *lut;    // bracket table of PotView (fj_device.hpp)
    float lg0, lsc;
    const float *Pr;   // Phi_l' table (staged for the fast phase only, else null)
};

// FP32 division and square root.  Fast = false: IEEE, correctly rounded (the
// mixed path).  Fast = true: the hardware reciprocal and square root
// (v_rcp_f32, v_sqrt_f32; about 1 ulp), for the fast node phase, whose error
// is the fixed-order action rule's (1e-5 relative), not FP32 rounding.
template <bool Fast>
__device__ __forceinline__ float fdiv32(float a, float b) {
    if constexpr (Fast) return a * __builtin_amdgcn_rcpf(b);
    else return a / b;
}
template <bool Fast>
__device__ __forceinline__ float fsqrt32(float x) {
    if constexpr (Fast) return __builtin_amdgcn_sqrtf(x);
    else return sqrtf(x);
}

// Phi(R, z) - shift.  The shift enters through the monopole, whose
// interpolation runs on differences of the float-pair table, so that at the
// bottom of a deep potential (shift close to Phi, E - Phi << |Phi|) the
// result carries no cancellation.
template <bool Fast = false>
__device__ __forceinline__ float pot_phi32(const PotView32 &T, float R, float z, float shift) {
    const float r2 = R * R + z * z;
    const float rs = fsqrt32<Fast>(r2);
    const float c = fdiv32<Fast>(z, rs);
    const int nr = T.nr, np_ = T.npoly;
    const float *ar = T.ar;
    const float c2 = c * c;
    // the coefficients are formed and summed in one pass (no per-lane array)
    float phi, pm2 = 1.f, pm1 = 1.5f * c2 - .5f;
    float r = rs;
    if (r >= ar[nr - 1]) {
        // Keplerian tail: phil[nr-1][l] (ar[nr-1]/r)^(l+1), l = 2np
        const float q = fdiv32<Fast>(ar[nr - 1], r), q2 = q * q;
        const float *last = T.phil + (nr - 1) * np_;
        phi = fmaf(last[0], q, -shift) + T.phil0lo[nr - 1] * q;
        float qp = q * q2;
        for (int np = 1; np < np_; np++) {
            float p;
            if (np == 1) p = pm1;
            else { p = (T.la[np] * pm2 + pm1 * (c2 - T.lb[np])) * T.lg[np]; pm2 = pm1; pm1 = p; }
            phi += (last[np] * qp) * p;
            qp *= q2;
        }
        return phi;
    }
    if (r < ar[0]) r = ar[0];
    const float t = (__log2f(r) - T.lg0) * T.lsc;
    int bot = T.lut[t > 0.f ? (t < (float)(NLUT - 1) ? (int)t : NLUT - 1) : 0];
    while (bot < nr - 2 && r >= ar[bot + 1]) bot++;
    while (bot > 0 && !(r >= ar[bot])) bot--;
    const int top = bot + 1;
    const float db = r - ar[bot];
    const float f1 = fdiv32<Fast>(db, ar[top] - ar[bot]);
    const bool quad = top < 10;
    float f2 = 0, f3 = 0;
    int thr = 0;
    if (quad) {
        thr = (f1 < 0.5f && bot > 0) ? bot - 1 : top + 1;
        f2 = fdiv32<Fast>((r - ar[top]) * db, (ar[thr] - ar[top]) * (ar[thr] - ar[bot]));
        f3 = fdiv32<Fast>(ar[thr] - ar[bot], ar[top] - ar[bot]);
    }
    const float *pb = T.phil + bot * np_, *pt = T.phil + top * np_, *ph = T.phil + thr * np_;
    {
        // monopole minus shift: b - shift + f1 (t - b) + f2 (h - b - f3 (t - b))
        const float *lo = T.phil0lo;
        const float b = pb[0];
        const float bc = (b - shift) + lo[bot];
        const float tb = (pt[0] - b) + (lo[top] - lo[bot]);
        phi = bc + f1 * tb;
        if (quad) phi += f2 * (((ph[0] - b) + (lo[thr] - lo[bot])) - f3 * tb);
    }
    for (int np = 1; np < np_; np++) {
        const float tv = pt[np], b = pb[np];
        float v = b + f1 * (tv - b);
        if (quad) v += f2 * (ph[np] - b - f3 * (tv - b));
        float p;
        if (np == 1) p = pm1;
        else { p = (T.la[np] * pm2 + pm1 * (c2 - T.lb[np])) * T.lg[np]; pm2 = pm1; pm1 = p; }
        phi += v * p;
    }
    return phi;
}

// Phi - shift and dPhi/dR, dPhi/dz in FP32, for the signs the fast phase's
// u walk takes of dF_u/du.  As Potential::dR / dz (Potential.cpp:221-276)
// the radial derivative comes from the Phi_l' table (three-table intpo2,
// UtilsLeg.h:127-158; beyond r_max the derivative of the Keplerian tail) and
// the angular one from the Legendre derivatives, here by the recurrences
// l P_l = (2l - 1) c P_l-1 - (l - 1) P_l-2 and P'_l = P'_l-2 + (2l - 1) P_l-1;
// s and c from the clamped radius, as there.
template <bool Fast = false>
__device__ __forceinline__ float pot_grad32(const PotView32 &T, float R, float z, float shift, float *dR,
                                            float *dz) {
    const int nr = T.nr, np_ = T.npoly;
    const float *ar = T.ar;
    const float rs = fsqrt32<Fast>(R * R + z * z);
    const bool kep = rs >= ar[nr - 1];
    const float r = rs < ar[0] ? ar[0] : rs;
    const float ir = fdiv32<Fast>(1.f, r);
    const float c = Fast ? z * ir : z / r, s = Fast ? R * ir : R / r;
    float phi, g1, dthc = 0.f;
    int bot = 0, top = 0, thr = 0;
    float f1 = 0, f2 = 0, f3 = 0, q = 0, q2 = 0, qp = 0;
    bool quad = false;
    if (kep) {
        q = Fast ? ar[nr - 1] * ir : ar[nr - 1] / r; q2 = q * q;
        const float *last = T.phil + (nr - 1) * np_;
        phi = fmaf(last[0], q, -shift) + T.phil0lo[nr - 1] * q;
        g1 = Fast ? -last[0] * q * ir : -last[0] * q / r;
        qp = q * q2;
    } else {
        const float t = (__log2f(r) - T.lg0) * T.lsc;
        bot = T.lut[t > 0.f ? (t < (float)(NLUT - 1) ? (int)t : NLUT - 1) : 0];
        while (bot < nr - 2 && r >= ar[bot + 1]) bot++;
        while (bot > 0 && !(r >= ar[bot])) bot--;
        top = bot + 1;
        const float db = r - ar[bot];
        f1 = fdiv32<Fast>(db, ar[top] - ar[bot]);
        quad = top < 10;
        if (quad) {
            thr = (f1 < 0.5f && bot > 0) ? bot - 1 : top + 1;
            f2 = fdiv32<Fast>((r - ar[top]) * db, (ar[thr] - ar[top]) * (ar[thr] - ar[bot]));
            f3 = fdiv32<Fast>(ar[thr] - ar[bot], ar[top] - ar[bot]);
        }
        const float *lo = T.phil0lo;
        const float b = T.phil[bot * np_], tb = (T.phil[top * np_] - b) + (lo[top] - lo[bot]);
        phi = ((b - shift) + lo[bot]) + f1 * tb;
        if (quad) phi += f2 * (((T.phil[thr * np_] - b) + (lo[thr] - lo[bot])) - f3 * tb);
        const float pb = T.Pr[bot * np_], pt = T.Pr[top * np_];
        g1 = pb + f1 * (pt - pb);
        if (quad) g1 += f2 * (T.Pr[thr * np_] - pb - f3 * (pt - pb));
    }
    float Pm2 = 0.f, Pm1 = 1.f, dPm2 = 0.f, dPm1 = 0.f;   // P_l-2, P_l-1 and their derivatives
    for (int np = 1; np < np_; np++) {
        for (int i = 2 * np - 1; i <= 2 * np; i++) {
            const float P = Fast ? ((float)(2 * i - 1) * c * Pm1 - (float)(i - 1) * Pm2) * __builtin_amdgcn_rcpf((float)i)
                                 : ((float)(2 * i - 1) * c * Pm1 - (float)(i - 1) * Pm2) / (float)i;
            const float dP = dPm2 + (float)(2 * i - 1) * Pm1;
            Pm2 = Pm1; Pm1 = P; dPm2 = dPm1; dPm1 = dP;
        }
        float v, d;
        if (kep) {
            const float l = T.phil[(nr - 1) * np_ + np];
            v = l * qp;
            d = Fast ? -(float)(2 * np + 1) * v * ir : -(float)(2 * np + 1) * v / r;
            qp *= q2;
        } else {
            const float tv = T.phil[top * np_ + np], b = T.phil[bot * np_ + np];
            v = b + f1 * (tv - b);
            if (quad) v += f2 * (T.phil[thr * np_ + np] - b - f3 * (tv - b));
            const float pt = T.Pr[top * np_ + np], pb = T.Pr[bot * np_ + np];
            d = pb + f1 * (pt - pb);
            if (quad) d += f2 * (T.Pr[thr * np_ + np] - pb - f3 * (pt - pb));
        }
        phi += v * Pm1;
        g1 += d * Pm1;
        dthc += v * dPm1;
    }
    const float dth = -s * dthc;                          // dPhi/dtheta
    *dR = dth * (Fast ? c * ir : c / r) + g1 * s;
    *dz = dth * (Fast ? -s * ir : -s / r) + g1 * c;
    return phi;
}

// sinh and cosh from one expm1 (accurate for small |u| too).
template <bool Fast = false>
__device__ __forceinline__ void sinhcosh32(float u, float *s, float *c) {
    const float em = expm1f(u), ep = em + 1.f;
    if constexpr (Fast) {
        const float ie = __builtin_amdgcn_rcpf(ep);
        *s = 0.5f * em * (em + 2.f) * ie;
        *c = 0.5f * (ep + ie);
    } else {
        *s = 0.5f * em * (em + 2.f) / ep;
        *c = 0.5f * (ep + 1.f / ep);
    }
}

// Juint / Jvint in FP32 with JInt's packing of the node constants, with the
// energies shifted by c = fl(E): k[1] = E - c and the potential constants
// (K2 of Ju; Iv, KV of Jv) taken at Phi - c, so that every term of the
// integrand is of the size of E - Phi (the integrands are invariant under
// the common shift of E and Phi).
struct JInt32 {
    const PotView32 *T;
    float k[JK];
    float c;
    double pre;   // FP64 prefactor of the action: sqrt(2) Delta Deltau, or 2 sqrt(2) Delta (pi/2 - vturn)
    int type;
    __device__ float operator()(float psi) const {
        float sp, cp;
        sincosf(psi, &sp, &cp);
        float F;
        const float D = k[0], E = k[1];
        if (type == 0) {
            const float u = k[2] - k[3] * cp;
            float shu, chu;
            sinhcosh32(u, &shu, &chu);
            const float shu2 = shu * shu;
            const float Phi = pot_phi32(*T, D * shu * k[4], D * chu * k[5], c);
            const float dU = (shu2 + k[6]) * Phi - k[9];
            F = E * (shu2 - k[7]) + k[8] - dU - k[10] * (1.f / shu2 - k[11]);
        } else {
            const float v = 1.5707963f - (1.5707963f - k[2]) * cp;
            float s, cs;
            sincosf(v, &s, &cs);
            const float sv2 = s * s;
            const float Phi = pot_phi32(*T, D * k[3] * s, D * k[4] * cs, c);
            const float dV = k[7] - (k[5] + sv2) * Phi;
            F = E * sv2 + k[6] + dV - .5f * k[8] / (D * D * sv2);
        }
        return sqrtf(fmaxf(0.f, F)) * sp;
    }
    __device__ double action(double integral) const { return pre * integral / PI; }
};

// DF (df_hg, DF.cpp:65-76, and the rotation term DF.cpp:114-119) in FP32.
struct DF32 {
    float A, B, dphi_h, dz_h, dphi_g, dz_g, norm, J0, invJ0, chi, k;
    int rot_form;
};

__device__ __forceinline__ float df32(const DF32 &d, float Jr, float Jphi, float Jz) {
    const float aJ = fabsf(Jphi);
    const float h = Jr + d.dphi_h * aJ + d.dz_h * Jz;
    const float g = Jr + d.dphi_g * aJ + d.dz_g * Jz;
    const float even = d.norm * fmaxf(0.f, powf(1.f + d.J0 / h, d.A) / powf(1.f + g * d.invJ0, d.B));
    if (d.chi == 0.f) return even;
    const float t = tanhf(d.chi * Jphi * d.invJ0);
    if (d.rot_form == 1) return (1.f - d.k) * even + d.k * t * even;
    return even + d.k * t * even;
}

// Fast mode (fj_fast.hpp): the two action integrals as Gauss-Legendre rules
// in FP32.  Per rule node k: weight w_k of the rule on [0, 1] and the fixed
// cos / sin of psi_k = pi x_k (Ju) and psi_k = pi x_k / 2 (Jv); the Ju rule
// reads cu, su of its table, the Jv rule (of fast_v_order points) cv, sv of
// its own.  Energies are
// measured from c = fl(E): Phi - c comes from the float-pair monopole of
// pot_phi32, and E - c is carried separately.
struct FastGL32 {
    float w, cu, su, cv, sv;
};

// F_u of fj_fast.hpp in FP32, for the signs of the u bracket walk.
struct FastFu32 {
    using Tab = const PotView32 *;
    const PotView32 *T;
    const FastNode *n;
    __device__ double operator()(double u) const {
        float sh, ch;
        sinhcosh32<true>((float)u, &sh, &ch);
        const float sh2 = sh * sh;
        const float c = (float)n->E, dE = (float)(n->E - (double)c);
        const float D = (float)n->Delta, sv0 = (float)n->sv0, cv0 = (float)n->cv0;
        const float Pc = pot_phi32<true>(*T, D * sh * sv0, D * ch * cv0, c) - dE;   // Phi - E
        return (double)((float)n->Bu - (float)n->C * __builtin_amdgcn_rcpf(sh2) - (sh2 + sv0 * sv0) * Pc);
    }
};

// dF_u/du of fj_fast.hpp (FastDFu) in FP32, for the signs of the u walk.
struct FastDFu32 {
    const PotView32 *T;
    const FastNode *n;
    __device__ double operator()(double u) const {
        float sh, ch;
        sinhcosh32<true>((float)u, &sh, &ch);
        const float sh2 = sh * sh;
        const float c = (float)n->E, dE = (float)(n->E - (double)c);
        const float D = (float)n->Delta, sv0 = (float)n->sv0, cv0 = (float)n->cv0;
        float dR, dz;
        const float Pc = pot_grad32<true>(*T, D * sh * sv0, D * ch * cv0, c, &dR, &dz) - dE;   // Phi - E
        const float dPhidu = D * (dR * ch * sv0 + dz * sh * cv0);
        return (double)(-2.f * Pc * sh * ch - (sh2 + sv0 * sv0) * dPhidu +
                        2.f * (float)n->C * ch * __builtin_amdgcn_rcpf(sh2 * sh));
    }
};

// F_v of fj_fast.hpp in FP32.
struct FastFv32 {
    const PotView32 *T;
    const FastNode *n;
    __device__ double operator()(double v) const {
        float sn, cs;
        sincosf((float)v, &sn, &cs);
        const float s2 = sn * sn;
        const float c = (float)n->E, dE = (float)(n->E - (double)c);
        const float D = (float)n->Delta, sh0 = (float)n->shu0, ch0 = (float)n->chu0;
        const float Pc = pot_phi32<true>(*T, D * sh0 * sn, D * ch0 * cs, c) - dE;   // Phi - E
        return (double)((float)n->Bv - (float)n->C * __builtin_amdgcn_rcpf(s2) - (sh0 * sh0 + s2) * Pc);
    }
};

// fast_setup's evaluations in the kernel: the u and v walks (with the signs
// of dF_u/du) and the root solves on FP32 evaluations (turning points to
// 2e-6 relative, below what the FP32 action rules resolve), the bracket
// checks and the endpoint test in FP64.
struct FastPolicy32 {
    using Tab = const PotView32 *;
    using U = FastFu32;
    using V = FastFv32;
    using DU = FastDFu32;
    static constexpr double rtol = 2e-6;
    __device__ static DU make_du(const PotView &, Tab tw, const FastNode *n) { return DU{tw, n}; }
};

__device__ __forceinline__ void fast_actions32(const PotView32 &T, const FastNode &n, const FastGL32 *gl, int ngl,
                                               const FastGL32 *glv, int nglv, double *Jr, double *Jz) {
    const double rtwo = 1.4142135623730951;
    const float c = (float)n.E;
    const float dE = (float)(n.E - (double)c);
    const float D = (float)n.Delta, sv0 = (float)n.sv0, cv0 = (float)n.cv0, C = (float)n.C;
    const float sv0sq = sv0 * sv0;
    *Jr = 0;
    *Jz = 0;
    if (n.ju) {
        const double du = .5 * (n.uout - n.uin);
        const float ub = (float)(.5 * (n.uin + n.uout)), duf = (float)du, Bu = (float)n.Bu;
        float s = 0.f;
#pragma unroll 2
        for (int k = 0; k < ngl; k++) {
            const float u = ub - duf * gl[k].cu;
            float sh, ch;
            sinhcosh32<true>(u, &sh, &ch);
            const float sh2 = sh * sh;
            const float Pc = pot_phi32<true>(T, D * sh * sv0, D * ch * cv0, c) - dE;   // Phi - E
            const float F = Bu - C * __builtin_amdgcn_rcpf(sh2) - (sh2 + sv0sq) * Pc;
            s += gl[k].w * (fsqrt32<true>(fmaxf(0.f, F)) * gl[k].su);
        }
        *Jr = rtwo * n.Delta * du * (double)s;
    }
    if (n.jv) {
        const double dv = PIH - n.vt;
        const float dvf = (float)dv, Bv = (float)n.Bv;
        const float sh0 = (float)n.shu0, ch0 = (float)n.chu0, sh0sq = sh0 * sh0;
        float s = 0.f;
#pragma unroll 2
        for (int k = 0; k < nglv; k++) {
            const float v = 1.5707963f - dvf * glv[k].cv;
            float sn, cs;
            sincosf(v, &sn, &cs);
            const float s2 = sn * sn;
            const float Pc = pot_phi32<true>(T, D * sh0 * sn, D * ch0 * cs, c) - dE;
            const float F = Bv - C * __builtin_amdgcn_rcpf(s2) - (sh0sq + s2) * Pc;
            s += glv[k].w * (fsqrt32<true>(fmaxf(0.f, F)) * glv[k].sv);
        }
        *Jz = rtwo * n.Delta * dv * (double)s;
    }
}

// Compensated (Kahan) FP32 accumulator.
struct KSum {
    float s = 0.f, c = 0.f;
    __device__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

#pragma clang fp contract(off)

}  // namespace fj
