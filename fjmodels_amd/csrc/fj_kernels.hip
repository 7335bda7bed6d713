// fj_kernels.hip -- gfx950 kernels of the fJmodels SCF iteration and the
// extern "C" layer declared in include/fjmodels.h.
//
// Node phase of rhofDF (integ/Integ.cpp:150-163) for a batch of points:
//   k_points    per point: Phi(R,z) and the escape speed Ve
//   k_setup     per velocity node: H, Deltafn, spheroidal coordinates, Getu0,
//               turning points (uvOrb.cpp:15-36,109-224); appends the two
//               action integrals of every bound node to task lists
//   k_integrals persistent: every lane evaluates one Gauss-Kronrod rule per
//               step for its own QAGS task and takes a new task when done,
//               so the expensive evaluations stay in lock step
//   k_moments   per point: f(J) at every node and the 6 weighted sums
//               (GaussQuad.cpp:160-180) reduced in a fixed order
// Fast node phase (FJ_PREC_FAST, fj_fast.hpp): k_fast fuses node setup,
// Gauss-Legendre action integrals, f(J) and the moment sums (no node
// records); k_fast_finish adds the slices of each point in a fixed order.
// Grid phase: k_project (potLeg.cpp:68 + UtilsLeg.h:200-253), k_poisson
// (Potential.cpp:110-200), k_guess (Potential.cpp:63-80), k_delta
// (delta/tabulate.cpp + delta/Delta.cpp), k_virial (libs/Stats.cpp).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/fjmodels.h"
#include "fj_device.hpp"
#include "fj_observe.h"
#include "fj_node.hpp"
#include "fj_mixed.hpp"
#include "fj_mlp.hpp"

using namespace fj;

namespace {

constexpr int NQ1 = 22;                 // nodes per velocity axis (QUADORD)
constexpr int NQ = NQ1 * NQ1 * NQ1;     // 10,648 nodes per point
constexpr int QCAP = 64;                // QAGS intervals kept per lane before overflow
// Mixed mode: FP32 integrands whose QAGS passes this many intervals are
// limited by their own rounding (near-circular or near-planar orbits, where
// F(u) is a small difference of large terms); they are redone in FP64.
#ifndef FJ_QCAP_MIXED
#define FJ_QCAP_MIXED 24
#endif
constexpr int QCAP_MIXED = FJ_QCAP_MIXED;
constexpr int QCAP_FULL = 1000;         // the reference's workspace limit (Utils.h:55)
// The float workspace of the mixed path lives in the buffer sized for QCAP doubles.
static_assert(QCAP_MIXED <= QCAP, "FP32 interval cap above the workspace's");

// include/GaussPts.h:76-136
__constant__ double c_xpt11[22] = {
    -1., -0.50118723, -0.25118864, -0.12589254, -0.06309573, -0.03162278, -0.01584893,
    -0.00794328, -0.00398107, -0.00199526, -0.001, 0.001, 0.00199526, 0.00398107,
    0.00794328, 0.01584893, 0.03162278, 0.06309573, 0.12589254, 0.25118864, 0.50118723, 1.};
__constant__ double c_xw11[22] = {
    0.49881277, 0.37440568, 0.18764735, 0.09404645, 0.04713488, 0.0236234, 0.01183975,
    0.00593393, 0.00297401, 0.00149054, 0.00149763, 0.00149763, 0.00149054, 0.00297401,
    0.00593393, 0.01183975, 0.0236234, 0.04713488, 0.09404645, 0.18764735, 0.37440568,
    0.49881277};
__constant__ double c_xpt01[22] = {
    1.12614037692e-05, 8.09053908976e-05, 0.000425204351858, 0.00172603419131,
    0.0056596260047, 0.015541463911, 0.0367576002085, 0.0764489202523,
    0.141870454058, 0.237263531493, 0.360233468529, 0.5,
    0.639766531471, 0.762736468507, 0.858129545942, 0.923551079748,
    0.963242399792, 0.984458536089, 0.994340373995, 0.998273965809,
    0.999574795648, 0.999919094609};
__constant__ double c_xw01[22] = {
    2.42000467048e-05, 0.000146089994545, 0.000647128630152, 0.00222235986647,
    0.00618889854184, 0.0144761668119, 0.0291728616842, 0.0514732311361,
    0.0801433368009, 0.110279557896, 0.133823374104, 0.142799666072,
    0.133823374104, 0.110279557896, 0.0801433368009, 0.0514732311361,
    0.0291728616842, 0.0144761668119, 0.00618889854184, 0.00222235986647,
    0.000647128630152, 0.000146089994545};

// ---------------------------------------------------------------------------
// Tables in LDS
// ---------------------------------------------------------------------------
struct TabArgs {
    int nr, npoly, ngrid;
    const double *ar, *phil, *Pr, *E, *D;
    const int *lut;
    float lg0, lsc;
    const double *rk;   // [nr][RK] radial cell divisors (fill_radk_cell)
};

// Legendre recurrence constants (index-only values, the same for every grid)
// in constant memory: wave-uniform reads go through the scalar cache.
__constant__ LegCoef c_legcoef;

__host__ __device__ inline size_t tab_lds_doubles(int nr, int npoly, int ngrid) {
    return (size_t)nr + 2 * (size_t)nr * npoly + 2 * (size_t)ngrid + (size_t)RK * nr + NLUT / 2;
}

struct LdsTabs {
    PotView T;
    const double *E, *D;
    int ngrid;
};

__device__ inline LdsTabs stage_tables(const TabArgs &a, double *smem) {
    const int nr = a.nr, np = a.npoly, ng = a.ngrid;
    double *ar = smem, *phil = ar + nr, *Pr = phil + nr * np, *E = Pr + nr * np, *D = E + ng, *rk = D + ng;
    int *lut = reinterpret_cast<int *>(rk + RK * nr);
    for (int i = threadIdx.x; i < nr; i += blockDim.x) ar[i] = a.ar[i];
    for (int i = threadIdx.x; i < nr * np; i += blockDim.x) { phil[i] = a.phil[i]; Pr[i] = a.Pr[i]; }
    for (int i = threadIdx.x; i < ng; i += blockDim.x) { E[i] = a.E[i]; D[i] = a.D[i]; }
    for (int i = threadIdx.x; i < RK * nr; i += blockDim.x) rk[i] = a.rk[i];
    for (int i = threadIdx.x; i < NLUT; i += blockDim.x) lut[i] = a.lut[i];
    __syncthreads();
    LdsTabs t;
    t.T.nr = nr; t.T.npoly = np; t.T.ar = ar; t.T.phil = phil; t.T.Pr = Pr;
    t.T.lc = &c_legcoef;
    t.T.lut = lut; t.T.lg0 = a.lg0; t.T.lsc = a.lsc; t.T.rk = rk;
    t.E = E; t.D = D; t.ngrid = ng;
    return t;
}

// FP32 copies of the radial grid, Phi_l table and even-Legendre step
// constants for the mixed-precision integrand, after the FP64 tables.
__host__ __device__ inline size_t tab32_lds_floats(int nr, int npoly, bool with_pr = false) {
    return (size_t)2 * nr + (size_t)nr * npoly * (with_pr ? 2 : 1) + 3 * MAXPOLY;
}

// with_pr: also the Phi_l' table (the fast phase's FP32 gradient).
__device__ inline PotView32 stage_tables32(const LdsTabs &L, float *f, bool with_pr = false) {
    const int nr = L.T.nr, np = L.T.npoly;
    float *ar = f, *phil = ar + nr, *la = phil + nr * np, *lb = la + MAXPOLY, *lg = lb + MAXPOLY;
    float *lo = lg + MAXPOLY;
    float *pr = with_pr ? lo + nr : nullptr;
    for (int i = threadIdx.x; i < nr; i += blockDim.x) ar[i] = (float)L.T.ar[i];
    for (int i = threadIdx.x; i < nr * np; i += blockDim.x) phil[i] = (float)L.T.phil[i];
    if (with_pr)
        for (int i = threadIdx.x; i < nr * np; i += blockDim.x) pr[i] = (float)L.T.Pr[i];
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
        const double p0 = L.T.phil[i * np];
        lo[i] = (float)(p0 - (double)(float)p0);
    }
    for (int i = threadIdx.x; i < MAXPOLY; i += blockDim.x) {
        const LegCoef *lc = L.T.lc;
        la[i] = i >= 2 ? (float)(-lc->el[i] * lc->elm1[i] / lc->ed[i]) : 0.f;
        lb[i] = (float)lc->eb[i];
        lg[i] = (float)lc->eg[i];
    }
    __syncthreads();
    PotView32 T;
    T.nr = nr; T.npoly = np; T.ar = ar; T.phil = phil; T.phil0lo = lo; T.la = la; T.lb = lb; T.lg = lg; T.Pr = pr;
    T.lut = L.T.lut; T.lg0 = L.T.lg0; T.lsc = L.T.lsc;
    return T;
}

// Node record between the kernels (structure of arrays).
enum { NF_DELTA, NF_E, NF_LZSQ, NF_SV, NF_CV, NF_SV2, NF_SINV, NF_COSV, NF_SHUX, NF_CHUX, NF_SHU2,
       NF_SH1SQ, NF_K1, NF_K2, NF_K3, NF_INVSH1, NF_IV, NF_KV, NF_UBAR, NF_DELTAU, NF_VTURN, NF_COUNT };

struct NodeBuf {
    double *f;       // [NF_COUNT][cap]
    long long cap;
    double *Jr, *Jz; // [cap]
    int *flag;       // [cap]: 0 zero DF (H >= 0), 1 bound, 2 unbound (E >= 0)
};

__device__ inline void store_const(const NodeBuf &nb, long long i, const NodeConst &c, double ubar, double Deltau,
                                   double vturn) {
    double *f = nb.f;
    const long long s = nb.cap;
    f[NF_DELTA * s + i] = c.Delta; f[NF_E * s + i] = c.E; f[NF_LZSQ * s + i] = c.Lzsq;
    f[NF_SV * s + i] = c.sv; f[NF_CV * s + i] = c.cv; f[NF_SV2 * s + i] = c.sv2;
    f[NF_SINV * s + i] = c.sinv; f[NF_COSV * s + i] = c.cosv; f[NF_SHUX * s + i] = c.shuX;
    f[NF_CHUX * s + i] = c.chuX; f[NF_SHU2 * s + i] = c.shu2; f[NF_SH1SQ * s + i] = c.sh1sq;
    f[NF_K1 * s + i] = c.K1; f[NF_K2 * s + i] = c.K2; f[NF_K3 * s + i] = c.K3;
    f[NF_INVSH1 * s + i] = c.invsh1sq; f[NF_IV * s + i] = c.Iv; f[NF_KV * s + i] = c.KV;
    f[NF_UBAR * s + i] = ubar; f[NF_DELTAU * s + i] = Deltau; f[NF_VTURN * s + i] = vturn;
}

// Node-record fields of the packed integrand constants (JInt::k) per kind;
// -1: unused slot.
__constant__ int c_jfield[2][JK] = {
    {NF_DELTA, NF_E, NF_UBAR, NF_DELTAU, NF_SINV, NF_COSV, NF_SV2, NF_SH1SQ, NF_K1, NF_K2, NF_K3, NF_INVSH1},
    {NF_DELTA, NF_E, NF_VTURN, NF_SHUX, NF_CHUX, NF_SHU2, NF_IV, NF_KV, NF_LZSQ, -1, -1, -1}};

__device__ inline void load_jint(const NodeBuf &nb, long long i, int type, JInt &f) {
    f.type = type;
#pragma unroll
    for (int m = 0; m < JK; m++) {
        const int fld = c_jfield[type][m];
        f.k[m] = fld >= 0 ? nb.f[fld * nb.cap + i] : 0.0;
    }
}

__device__ inline void load_jint(const NodeBuf &nb, long long i, int type, JInt32 &f) {
    JInt g;
    load_jint(nb, i, type, g);
    f.type = type;
#pragma unroll
    for (int m = 0; m < JK; m++) f.k[m] = (float)g.k[m];
    // energy shift (JInt32): c = fl(E)
    const float c = (float)g.k[1];
    const double cd = (double)c;
    f.c = c;
    f.k[1] = (float)(g.k[1] - cd);
    if (type == 0) {
        f.k[9] = (float)(g.k[9] - cd * (g.k[7] + g.k[6]));      // (sh1sq + sv2) (Phi(u1) - c)
    } else {
        f.k[6] = (float)(g.k[6] + cd);                          // Iv at E - c
        f.k[7] = (float)(g.k[7] - cd * (1 + g.k[5]));           // (1 + shu2) (Phi(pi/2) - c)
    }
    const double rtwo = 1.4142135623730951;
    f.pre = type == 0 ? rtwo * g.k[0] * g.k[3] : 2 * rtwo * g.k[0] * (PIH - g.k[2]);
}

// Device counters
struct Counters {
    unsigned long long bound, tasks_u, tasks_v, qk, overflow, brent, grad;
    unsigned long long unbound;        // E >= 0 after the spheroidal transform: Jr = Jz = 1e31
    unsigned long long bracket_fail;   // a turning-point bracket walk stopped by its iteration guard
    unsigned long long nonfinite;      // nodes whose f(J) is NaN or infinite
    unsigned long long phi;            // potential evaluations of k_setup (gradient ones included)
};

// Wave-aggregated append to a task list.
__device__ inline void wave_append(bool pred, unsigned int *counter, int *list, int value) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned int)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) {
        const unsigned long long lower = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
        list[base + __popcll(lower)] = value;
    }
}

struct PointArgs {
    const double *R, *z;   // [npts]
    double *Phi, *Ve;      // [npts]
    double *Vc;            // [npts] or null: circular-speed scale sqrt(R dPhi/dR + z dPhi/dz)
};

// k_points: Phi_h = Phi(R,z); Ve = sqrt(-2(Phi_h - Phi(10 r_max, 10 r_max))) (Integ.cpp:152-153)
__global__ void __launch_bounds__(256) k_points(TabArgs ta, PointArgs pa, int npts) {
    extern __shared__ double smem[];
    const LdsTabs L = stage_tables(ta, smem);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const double far = 10 * L.T.ar[L.T.nr - 1];
    const double Phi_h = pot_phi(L.T, pa.R[i], pa.z[i]);
    pa.Phi[i] = Phi_h;
    pa.Ve[i] = sqrt(-2 * (Phi_h - pot_phi(L.T, far, far)));
    if (pa.Vc) {
        double dR, dz;
        pot_phi_grad(L.T, pa.R[i], pa.z[i], &dR, &dz);
        pa.Vc[i] = sqrt(fj_max(1e-30, pa.R[i] * dR + pa.z[i] * dz));
    }
}

struct SetupArgs {
    const double *R, *z, *Phi, *Ve;
    int p0;             // first point of this chunk
    int npts;           // points in this chunk
    int quad;           // FJ_QUAD_*
    NodeBuf nb;
    unsigned int *cnt_u, *cnt_v;
    int *list_u, *list_v;
    Counters *ctr;
    // explicit phase-space points (fj_actions): node gid is point gid
    const double *vR, *vphi, *vz, *Dx;
    // if set: only the nodes sel[0 .. *nsel) of the chunk
    const int *sel;
    const unsigned int *nsel;
};

__device__ inline void node_velocity(int quad, int node, double Ve, double V[3], double &w) {
    const int i = node / (NQ1 * NQ1), j = (node / NQ1) % NQ1, k = node % NQ1;
    const double *xa = quad ? c_xpt01 : c_xpt11, *wa = quad ? c_xw01 : c_xw11;
    V[0] = 1e-6 + Ve * xa[i];
    V[1] = 1e-6 + Ve * c_xpt11[j];
    V[2] = 1e-6 + Ve * xa[k];
    w = wa[i] * c_xw11[j] * wa[k];
}

// k_setup: everything of df()/uvOrb before the two QAGS integrals, for
// node tid of the launch (every lane of a wave calls it: task appends are
// wave-aggregated).
__device__ __forceinline__ void setup_node(const LdsTabs &L, const SetupArgs &sa, long long tid) {
    const bool explicit_v = sa.vR != nullptr;
    const long long nnodes = sa.sel ? (long long)*sa.nsel : explicit_v ? (long long)sa.npts : (long long)sa.npts * NQ;
    const bool valid = tid < nnodes;
    const long long gid = sa.sel ? (valid ? (long long)sa.sel[tid] : 0) : tid;
    bool want_u = false, want_v = false;
    long long nev = 0;
    int brent_it = 0, flag = 0, bfail = 0, nphi = 0;
    if (valid) {
        const int pl = explicit_v ? (int)gid : (int)(gid / NQ), node = explicit_v ? 0 : (int)(gid % NQ);
        const int p = sa.p0 + pl;
        const double R = sa.R[p], z = sa.z[p], Phi_h = sa.Phi[p], Ve = sa.Ve[p];
        double V[3], w;
        if (explicit_v) { V[0] = sa.vR[p]; V[1] = sa.vphi[p]; V[2] = sa.vz[p]; }
        else node_velocity(sa.quad, node, Ve, V, w);
        double Jr = 0, Jz = 0;
        NodeConst c;
        double ubar = 0, Deltau = 0, vturn = 0;
        int wu = 0, wv = 0;
        const double Dx = (explicit_v && sa.Dx) ? sa.Dx[p] : 0.0;
        flag = node_setup(L.T, L.E, L.D, L.ngrid, R, z, Phi_h, V, Dx, !explicit_v, c, ubar, Deltau, vturn,
                          wu, wv, Jr, Jz, nev, brent_it, &bfail, &nphi);
        want_u = wu; want_v = wv;
        if (flag == 1) store_const(sa.nb, gid, c, ubar, Deltau, vturn);
        sa.nb.flag[gid] = flag;
        sa.nb.Jr[gid] = Jr;
        sa.nb.Jz[gid] = Jz;
    }
    wave_append(want_u, sa.cnt_u, sa.list_u, (int)gid);
    wave_append(want_v, sa.cnt_v, sa.list_v, (int)gid);
    // counters (wave-aggregated)
    const unsigned long long nb_ = __popcll(__ballot(valid && (want_u || want_v || nev > 0)));
    if (__lane_id() == 0 && nb_) atomicAdd(&sa.ctr->bound, nb_);
    const unsigned long long nu_ = __popcll(__ballot(valid && flag == 2));
    if (__lane_id() == 0 && nu_) atomicAdd(&sa.ctr->unbound, nu_);
    if (valid && bfail) atomicAdd(&sa.ctr->bracket_fail, 1ull);
    if (valid && brent_it) atomicAdd(&sa.ctr->brent, (unsigned long long)brent_it);
    if (valid && nev) atomicAdd(&sa.ctr->grad, (unsigned long long)nev);
    if (valid && nphi) atomicAdd(&sa.ctr->phi, (unsigned long long)nphi);
}

#ifndef FJ_SETUP_WAVES
#define FJ_SETUP_WAVES 4   // waves per SIMD the register allocation of k_setup aims at
#endif

// Lane order of k_setup over the npts x NQ nodes of a chunk: tiles of 64
// grid points x NQ nodes, and inside a tile lane r takes velocity node
// r / 64 of point r % 64 (of the tile's points, fewer in the last tile).  A
// wave then follows one velocity node, in units of the local escape speed,
// at 64 neighbouring grid points: orbits of like shape, whose Brent
// iteration counts are alike.  The node records stay point-major (node
// p * NQ + q), so this changes which lane computes a node, never a result.
#ifndef FJ_SETUP_TILED
#define FJ_SETUP_TILED 1
#endif
__device__ inline long long setup_lane_node(long long tid, int npts) {
    if (tid >= (long long)npts * NQ) return tid;          // past the last node: stays invalid
    const long long tile = 64ll * NQ;
    const long long T = tid / tile, r = tid - T * tile;
    const int left = npts - (int)T * 64;
    const int tp = left < 64 ? left : 64;                   // points in this tile, >= 1
    const long long q = r / tp, pt = r - q * tp;            // q < NQ, pt < tp
    return (T * 64 + pt) * NQ + q;
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FJ_SETUP_WAVES)))
k_setup(TabArgs ta, SetupArgs sa) {
    extern __shared__ double smem[];
    const LdsTabs L = stage_tables(ta, smem);
    long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
#if FJ_SETUP_TILED
    if (!sa.sel && !sa.vR) tid = setup_lane_node(tid, sa.npts);
#endif
    setup_node(L, sa, tid);
}

// Sort key of a node for the next node phase's lane order: for a bound node
// (flag 1) its turning points u_in and u_out on the scale of GetTurnu's steps
// (ln 1.1) and its v turning point as the halvings of GetTurnv from pi/4; 0
// for H >= 0, 1 for E >= 0.  Taken in the order of the sorted keys, a wave
// runs nodes whose bracket walks and root solves take like steps.  Any key
// gives the same results: the records stay node-indexed, so only which lane
// computes a node changes.
constexpr int SETUP_KEY_BITS = 25;
__device__ inline unsigned int node_order_key(int flag, double uin, double uout, double vt) {
    if (flag != 1) return flag == 2 ? 1u : 0u;
    // ln u / ln 1.1 + 256, clamped to [0, 511]
    const auto ubin = [](double u) -> unsigned int {
        const float t = u > 0 ? __logf((float)u) * 10.49205f + 256.f : 0.f;
        return t > 0.f ? (t < 511.f ? (unsigned int)t : 511u) : 0u;
    };
    // halvings from pi/4 towards pi/2 (32 + level) or towards 0 (level), level in [0, 31]
    const bool up = vt > 0.25 * PI;
    const double d = up ? PIH - vt : vt;
    const float lv = d > 0 ? __log2f((float)(0.25 * PI / d)) : 31.f;
    const unsigned int lvl = lv > 0.f ? (lv < 31.f ? (unsigned int)lv : 31u) : 0u;
    return (1u << 24) + (ubin(uout) << 15) + (ubin(uin) << 6) + (up ? 32u : 0u) + lvl;
}
// The keys of the nn nodes of the records nb (none if key is null) and the
// identity permutation to sort with them; the node count into *nsel if set.
__global__ void __launch_bounds__(256) k_setup_key(NodeBuf nb, long long nn, unsigned int *key, int *perm,
                                                   unsigned int *nsel) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && nsel) *nsel = (unsigned int)nn;
    if (i >= nn) return;
    perm[i] = (int)i;
    if (!key) return;
    const int flag = nb.flag[i];
    double ub = 0, du = 0, vt = 0;
    if (flag == 1) {
        ub = nb.f[NF_UBAR * nb.cap + i]; du = nb.f[NF_DELTAU * nb.cap + i]; vt = nb.f[NF_VTURN * nb.cap + i];
    }
    key[i] = node_order_key(flag, ub - du, ub + du, vt);
}

// Sort keys of a task list of the FP32 integrator: the intervals the task's
// QAGS took in the previous node phase, most first (0xFFFE - count), so that
// a wave's lanes take tasks of like cost and the costliest start first;
// entries past the list's length (*cnt) sort last.
constexpr int TASK_KEY_BITS = 16;
__global__ void __launch_bounds__(256) k_task_key(const int *list, const unsigned int *cnt, long long n, int type,
                                                  const int *cost, unsigned int *key) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned int k = 0xFFFFu;
    if (i < (long long)*cnt) {
        const int q = cost[2ll * list[i] + type];
        k = 0xFFFEu - (unsigned int)(q < 0 ? 0 : q > 0xFFFE ? 0xFFFE : q);
    }
    key[i] = k;
}

// df_hg (DF.cpp:65-76) and the rotation term (DF.cpp:114-119)
__device__ inline double d_df(const fj_df_params &d, double Jr, double Jphi, double Jz) {
    const double h = Jr + d.dphi_h * fabs(Jphi) + d.dz_h * Jz;
    const double g = Jr + d.dphi_g * fabs(Jphi) + d.dz_g * Jz;
    const double even = d.mass / fjm_powi(d.J0, 3) *
                        fj_max(0., fjm_pow(1. + d.J0 / h, d.A) / (fjm_pow(1. + g / d.J0, d.B))) / fjm_powi(TPI, 3);
    if (d.chi == 0.) return even;
    const double t = fjm_tanh(d.chi * Jphi / d.J0);
    if (d.rot_form == 1) return (1. - d.k) * even + d.k * t * even;
    return even + d.k * t * even;
}

// df() (DF.cpp:78-124) at explicit points whose actions are in Jr, Jz:
// 0 where H = V^2/2 + Phi >= 0, else the double-power-law DF.
__global__ void __launch_bounds__(256) k_df_points(int n, const double *R, const double *Phi, const double *vR,
                                                   const double *vphi, const double *vz, const double *Jr,
                                                   const double *Jz, fj_df_params df, double *f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double H = .5 * (vR[i] * vR[i] + vphi[i] * vphi[i] + vz[i] * vz[i]) + Phi[i];   // DF.cpp:83
    f[i] = H >= 0 ? 0.0 : d_df(df, Jr[i], R[i] * vphi[i], Jz[i]);
}

// ---------------------------------------------------------------------------
// Action surrogate in the node phase.  Inputs of node (R, z, V):
// (log R, z/R, vR/vc, vz/vc, |vphi|/vc) with vc the circular-speed scale of
// the grid point.  Outputs: log(Jr / s + tfloor) with the action scale
// s = r vc of the grid point, and log((Jz + zf s) / (L - |Lz| + zf s)), Jz
// relative to its spherical-limit value (L = |r x V|, Lz = R vphi).
// ---------------------------------------------------------------------------
struct SurNorm {
    float mu[MLP_IN], sig[MLP_IN], tmu[MLP_OUT], tsig[MLP_OUT];
    float lo[MLP_IN], hi[MLP_IN];   // training-set box of the normalised inputs
    float tfloor;                   // Jr target: log(Jr / s + tfloor)
};

__device__ inline void sur_features(double R, double z, double vc, const double V[3], float f[MLP_IN]) {
    f[0] = (float)log(R);
    f[1] = (float)(z / R);
    f[2] = (float)(V[0] / vc);
    f[3] = (float)(V[2] / vc);
    f[4] = (float)(fabs(V[1]) / vc);
}

// Jz of the spherical limit, L - |Lz|: the Jz target is the logarithm of
// (Jz + zf s) / (L - |Lz| + zf s), a correction near zero.
constexpr float SUR_ZFLOOR = 1e-6f;
__device__ inline double sur_jz_sph(double R, double z, const double V[3]) {
    const double Lm = z * V[0] - R * V[2];
    const double L = sqrt((R * R + z * z) * V[1] * V[1] + Lm * Lm);
    return fmax(0.0, L - fabs(R * V[1]));
}

struct SurArgs {
    const double *R, *z, *Phi, *Ve, *Vc;
    int p0, npts, quad;
    NodeBuf nb;
    SurNorm nm;
    bf16_t *x16;            // [nodes][5] normalised inputs
    const float *y;         // [nodes][2] normalised outputs
    double judge_frac, jz_min;
    unsigned int seed;
    unsigned int *nsel;     // fallback + judged nodes -> sel
    int *sel;
    unsigned int *njudge;   // judged nodes -> judge (node, Jr_sur, Jz_sur)
    int *judge_node;
    float *judge_J;         // [2 * judged]
    unsigned int *nfall;
    fj_df_params df;        // weights of the judged nodes' errors
};

__host__ __device__ inline unsigned int hash32(unsigned int x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// Inputs for every node; flag 0 (H >= 0), 1 (bound, surrogate), 3 (bound,
// Staeckel: outside the training box or |H| < 0.02 |Phi|).
__global__ void __launch_bounds__(256) k_sur_prep(SurArgs a) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)a.npts * NQ) return;
    const int p = a.p0 + (int)(gid / NQ), node = (int)(gid % NQ);
    const double R = a.R[p], z = a.z[p], Phi_h = a.Phi[p], Ve = a.Ve[p];
    double V[3], w;
    node_velocity(a.quad, node, Ve, V, w);
    const double H = .5 * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) + Phi_h;   // DF.cpp:83
    int flag = H < 0 ? 1 : 0;
    float f[MLP_IN];
    sur_features(R, z, a.Vc[p], V, f);
    for (int k = 0; k < MLP_IN; k++) {
        const float x = (f[k] - a.nm.mu[k]) / a.nm.sig[k];
        if (flag == 1 && !(x >= a.nm.lo[k] && x <= a.nm.hi[k])) flag = 3;
        a.x16[gid * MLP_IN + k] = f2bf(x);
    }
    if (flag == 1 && fabs(H) < 0.02 * fabs(Phi_h)) flag = 3;
    a.nb.flag[gid] = flag;
    a.nb.Jr[gid] = 0;
    a.nb.Jz[gid] = 0;
}

// Predicted actions; nodes that fall back (flag 3, predicted Jz < jz_min) or
// are judged (hash below judge_frac) go to the Staeckel list.
__global__ void __launch_bounds__(256) k_sur_post(SurArgs a) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = gid < (long long)a.npts * NQ;
    bool fall = false, judge = false;
    float Jr = 0.f, Jz = 0.f;
    if (valid) {
        const int flag = a.nb.flag[gid];
        if (flag == 1) {
            const int p = a.p0 + (int)(gid / NQ), node = (int)(gid % NQ);
            const double R = a.R[p], z = a.z[p];
            const float sc = (float)(sqrt(R * R + z * z) * a.Vc[p]);
            const float lr = a.y[gid * MLP_OUT] * a.nm.tsig[0] + a.nm.tmu[0];
            const float lz = a.y[gid * MLP_OUT + 1] * a.nm.tsig[1] + a.nm.tmu[1];
            double V[3], w;
            node_velocity(a.quad, node, a.Ve[p], V, w);
            const float zs = (float)sur_jz_sph(R, z, V) + sc * SUR_ZFLOOR;
            Jr = fmaxf(0.f, sc * (expf(lr) - a.nm.tfloor));
            Jz = fmaxf(0.f, zs * expf(lz) - sc * SUR_ZFLOOR);
            a.nb.Jr[gid] = Jr;
            a.nb.Jz[gid] = Jz;
            const float aJphi = (float)fabs(R * V[1]);
            if (Jz < (float)a.jz_min * (Jr + Jz + aJphi)) fall = true;
            else if ((double)hash32((unsigned int)gid ^ a.seed) < a.judge_frac * 4294967296.0) judge = true;
        } else if (flag == 3) {
            fall = true;
        }
    }
    wave_append(fall || judge, a.nsel, a.sel, (int)gid);
    const unsigned long long jm = __ballot(judge);
    if (jm) {
        const int lane = __lane_id(), leader = __ffsll((long long)jm) - 1;
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(a.njudge, (unsigned int)__popcll(jm));
        base = __shfl(base, leader);
        if (judge) {
            const unsigned long long lower = lane == 0 ? 0ull : (jm & ((~0ull) >> (64 - lane)));
            const unsigned int q = base + (unsigned int)__popcll(lower);
            a.judge_node[q] = (int)gid;
            a.judge_J[2 * q] = Jr;
            a.judge_J[2 * q + 1] = Jz;
        }
    }
    const unsigned long long fm = __ballot(fall);
    if (__lane_id() == 0 && fm) atomicAdd(a.nfall, (unsigned int)__popcll(fm));
    const unsigned long long sm = __ballot(valid && !fall && a.nb.flag[valid ? gid : 0] == 1);
    if (__lane_id() == 0 && sm) atomicAdd(a.nfall + 1, (unsigned int)__popcll(sm));
}

// Fused surrogate node phase: per 32-node tile and wave, the inputs, the
// network (mlp_tile) and the decisions of k_sur_prep / k_sur_post.
// Dynamic LDS: the packed weights of the two 128 x 128 layers (64 KB).
constexpr int SUR_LDS = 2 * 16 * 64 * 4 * 4 * (int)sizeof(bf16_t);

__global__ void __launch_bounds__(256) k_sur_infer(SurArgs a, const bf16_t *__restrict__ wg,
                                                   const float *__restrict__ bias) {
    extern __shared__ double smem_sur[];
    bf16_t *wl = reinterpret_cast<bf16_t *>(smem_sur);
    {
        const uint2 *src = reinterpret_cast<const uint2 *>(wg + MlpPacked::L2);
        uint2 *dst = reinterpret_cast<uint2 *>(wl);
        for (int i = threadIdx.x; i < SUR_LDS / 8; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const long long nn = (long long)a.npts * NQ, ntiles = (nn + 31) / 32;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    for (long long tile = (long long)blockIdx.x * 4 + wave; tile < ntiles; tile += (long long)gridDim.x * 4) {
        const long long gid = tile * 32 + (lane & 31);
        const bool valid = gid < nn;
        int flag = 0;
        float x[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f};
        int p = 0, node = 0;
        double V[3] = {0, 0, 0};
        if (valid) {
            p = a.p0 + (int)(gid / NQ);
            node = (int)(gid % NQ);
            double w;
            node_velocity(a.quad, node, a.Ve[p], V, w);
            const double H = .5 * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) + a.Phi[p];   // DF.cpp:83
            flag = H < 0 ? 1 : 0;
            float f[MLP_IN];
            sur_features(a.R[p], a.z[p], a.Vc[p], V, f);
            for (int k = 0; k < MLP_IN; k++) {
                x[k] = (f[k] - a.nm.mu[k]) / a.nm.sig[k];
                if (flag == 1 && !(x[k] >= a.nm.lo[k] && x[k] <= a.nm.hi[k])) flag = 3;
            }
            if (flag == 1 && fabs(H) < 0.02 * fabs(a.Phi[p])) flag = 3;
        }
        const v4s x4 = lane < 32 ? pack4(x[0], x[1], x[2], x[3]) : pack4(x[4], 0.f, 0.f, 0.f);
        float y0, y1;
        mlp_tile(wg, wl, bias, x4, y0, y1);
        bool fall = false, judge = false, served = false;
        float Jr = 0.f, Jz = 0.f;
        if (valid && lane < 32) {
            if (flag == 1) {
                const double R = a.R[p], z = a.z[p];
                const float sc = (float)(sqrt(R * R + z * z) * a.Vc[p]);
                const float zs = (float)sur_jz_sph(R, z, V) + sc * SUR_ZFLOOR;
                Jr = fmaxf(0.f, sc * (expf(y0 * a.nm.tsig[0] + a.nm.tmu[0]) - a.nm.tfloor));
                Jz = fmaxf(0.f, zs * expf(y1 * a.nm.tsig[1] + a.nm.tmu[1]) - sc * SUR_ZFLOOR);
                const float aJphi = (float)fabs(R * V[1]);
                if (Jz < (float)a.jz_min * (Jr + Jz + aJphi)) fall = true;
                else {
                    served = true;
                    if ((double)hash32((unsigned int)gid ^ a.seed) < a.judge_frac * 4294967296.0) judge = true;
                }
            } else if (flag == 3) {
                fall = true;
            }
            a.nb.flag[gid] = flag;
            a.nb.Jr[gid] = Jr;
            a.nb.Jz[gid] = Jz;
        }
        wave_append(fall || judge, a.nsel, a.sel, (int)gid);
        const unsigned long long jm = __ballot(judge);
        if (jm) {
            const int leader = __ffsll((long long)jm) - 1;
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(a.njudge, (unsigned int)__popcll(jm));
            base = __shfl(base, leader);
            if (judge) {
                const unsigned long long lower = lane == 0 ? 0ull : (jm & ((~0ull) >> (64 - lane)));
                const unsigned int q = base + (unsigned int)__popcll(lower);
                a.judge_node[q] = (int)gid;
                a.judge_J[2 * q] = Jr;
                a.judge_J[2 * q + 1] = Jz;
            }
        }
        const unsigned long long fm = __ballot(fall), sm = __ballot(served);
        if (lane == 0 && fm) atomicAdd(a.nfall, (unsigned int)__popcll(fm));
        if (lane == 0 && sm) atomicAdd(a.nfall + 1, (unsigned int)__popcll(sm));
    }
}

// Network on given normalised inputs x [n][5] through the fused tile -> y [n][2].
__global__ void __launch_bounds__(256) k_mlp_infer_x(long long n, const float *__restrict__ x,
                                                     const bf16_t *__restrict__ wg, const float *__restrict__ bias,
                                                     float *__restrict__ y) {
    extern __shared__ double smem_sur[];
    bf16_t *wl = reinterpret_cast<bf16_t *>(smem_sur);
    {
        const uint2 *src = reinterpret_cast<const uint2 *>(wg + MlpPacked::L2);
        uint2 *dst = reinterpret_cast<uint2 *>(wl);
        for (int i = threadIdx.x; i < SUR_LDS / 8; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const long long ntiles = (n + 31) / 32;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    for (long long tile = (long long)blockIdx.x * 4 + wave; tile < ntiles; tile += (long long)gridDim.x * 4) {
        const long long r = tile * 32 + (lane & 31);
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        if (r < n) {
            if (lane < 32) for (int k = 0; k < 4; k++) v[k] = x[r * MLP_IN + k];
            else v[0] = x[r * MLP_IN + 4];
        }
        float y0, y1;
        mlp_tile(wg, wl, bias, pack4(v[0], v[1], v[2], v[3]), y0, y1);
        if (r < n && lane < 32) { y[r * MLP_OUT] = y0; y[r * MLP_OUT + 1] = y1; }
    }
}

// Training samples from the last Staeckel node phase: bound nodes with
// finite actions, each kept with probability keep, into a ring buffer.
// Row: inputs [5]; targets log(Jr/s + tf), log(Jz/s + tf); weight
// 0.05 + NQ w f Ve^3 / sum_nodes(w f Ve^3), the node's share of its point's
// density times the node count (mean about 1), so the loss spends its
// effort where the density is made.
constexpr int SUR_TW = 3;   // training-set row: two targets and a weight

struct CollectArgs {
    const double *R, *z, *Phi, *Ve, *Vc;
    int p0, npts, quad;
    NodeBuf nb;
    const int *only;          // if set: only these nodes (judged ones)
    unsigned int n_only;
    double keep;
    unsigned int seed;
    float tfloor;
    fj_df_params df;
    const double *out6;       // [npts][6] raw moment sums of the node phase
    float *X, *T;             // [cap][5], [cap][SUR_TW]
    unsigned long long *pos;  // running write position (mod cap)
    long long cap;
};

__global__ void __launch_bounds__(256) k_sur_collect(CollectArgs a) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = a.only ? (long long)a.n_only : (long long)a.npts * NQ;
    bool take = false;
    long long gid = 0;
    if (tid < n) {
        gid = a.only ? a.only[tid] : tid;
        take = a.nb.flag[gid] == 1 && a.nb.Jr[gid] < 1e30 && a.nb.Jz[gid] < 1e30 &&
               (double)hash32((unsigned int)gid ^ a.seed) < a.keep * 4294967296.0;
    }
    const unsigned long long m = __ballot(take);
    if (!m) return;
    const int lane = __lane_id(), leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(a.pos, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    if (!take) return;
    const unsigned long long lower = lane == 0 ? 0ull : (m & ((~0ull) >> (64 - lane)));
    const long long q = (long long)((base + __popcll(lower)) % (unsigned long long)a.cap);
    const int p = a.p0 + (int)(gid / NQ), node = (int)(gid % NQ);
    double V[3], w;
    node_velocity(a.quad, node, a.Ve[p], V, w);
    float f[MLP_IN];
    sur_features(a.R[p], a.z[p], a.Vc[p], V, f);
    for (int k = 0; k < MLP_IN; k++) a.X[q * MLP_IN + k] = f[k];
    const double sc = sqrt(a.R[p] * a.R[p] + a.z[p] * a.z[p]) * a.Vc[p];
    const double Jr = a.nb.Jr[gid], Jz = a.nb.Jz[gid];
    a.T[q * SUR_TW] = (float)log(Jr / sc + a.tfloor);
    a.T[q * SUR_TW + 1] = (float)log((Jz + sc * SUR_ZFLOOR) / (sur_jz_sph(a.R[p], a.z[p], V) + sc * SUR_ZFLOOR));
    const double Ve = a.Ve[p], o0 = a.out6[(long long)p * 6];
    const double share = o0 > 0 ? w * d_df(a.df, Jr, a.R[p] * V[1], Jz) * Ve * Ve * Ve / o0 : 0.0;
#ifndef FJ_SUR_WEIGHTED
#define FJ_SUR_WEIGHTED 1
#endif
    a.T[q * SUR_TW + 2] = FJ_SUR_WEIGHTED ? (float)(0.05 + NQ * share) : 1.f;
}

// Batch b of an epoch: rows (off + (j B + b) P) mod n of the training set,
// normalised (P odd and coprime with n: every row once per epoch).
__global__ void k_sur_batch(long long B, long long first, long long P, long long n, long long off, const float *X,
                            const float *T, SurNorm nm, bf16_t *x16, float *t, float *wt) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const long long idx = (off + (unsigned long long)(first + b) % (unsigned long long)n * (unsigned long long)P) % n;
    for (int k = 0; k < MLP_IN; k++) x16[b * MLP_IN + k] = f2bf((X[idx * MLP_IN + k] - nm.mu[k]) / nm.sig[k]);
    for (int k = 0; k < MLP_OUT; k++) t[b * MLP_OUT + k] = (T[idx * SUR_TW + k] - nm.tmu[k]) / nm.tsig[k];
    wt[b] = T[idx * SUR_TW + 2];
}

// Per-block min and max of the normalised inputs over the training set.
__global__ void k_sur_box(long long n, const float *X, SurNorm nm, float *part) {
    __shared__ float lo[MLP_IN][256], hi[MLP_IN][256];
    float l[MLP_IN], h[MLP_IN];
    for (int k = 0; k < MLP_IN; k++) { l[k] = 3.4e38f; h[k] = -3.4e38f; }
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        for (int k = 0; k < MLP_IN; k++) {
            const float x = (X[i * MLP_IN + k] - nm.mu[k]) / nm.sig[k];
            l[k] = fminf(l[k], x);
            h[k] = fmaxf(h[k], x);
        }
    for (int k = 0; k < MLP_IN; k++) { lo[k][threadIdx.x] = l[k]; hi[k][threadIdx.x] = h[k]; }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < MLP_IN; k++) {
                lo[k][threadIdx.x] = fminf(lo[k][threadIdx.x], lo[k][threadIdx.x + st]);
                hi[k][threadIdx.x] = fmaxf(hi[k][threadIdx.x], hi[k][threadIdx.x + st]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < MLP_IN; k++) { part[blockIdx.x * 2 * MLP_IN + k] = lo[k][0]; part[blockIdx.x * 2 * MLP_IN + MLP_IN + k] = hi[k][0]; }
}

// Errors of the predicted actions at the judged nodes, after the Staeckel
// pass: |J - J_St| / (Jr_St + Jz_St + |Jphi|), the fraction of the DF's
// action combination h or g that the error moves (-1: unbound node).
// err[2i], err[2i+1]: the network's Jr, Jz errors at judged node i over the
// node's total action; err[2n + i]: the node's weight in the density sum,
// quadrature weight x Ve^3 x f(J) with the Staeckel actions.
__global__ void k_sur_judge_err(SurArgs a, unsigned int n, float *err) {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long gid = a.judge_node[i];
    const int p = a.p0 + (int)(gid / NQ), node = (int)(gid % NQ);
    double V[3], w;
    const double Ve = a.Ve[p];
    node_velocity(a.quad, node, Ve, V, w);
    const double Sr = a.nb.Jr[gid], Sz = a.nb.Jz[gid];
    if (!(Sr < 1e30 && Sz < 1e30)) { err[2 * i] = -1.f; err[2 * i + 1] = -1.f; err[2 * n + i] = 0.f; return; }
    const double Jphi = a.R[p] * V[1];
    const double scale = Sr + Sz + fabs(Jphi);
    err[2 * i] = (float)(fabs(a.judge_J[2 * i] - Sr) / scale);
    err[2 * i + 1] = (float)(fabs(a.judge_J[2 * i + 1] - Sz) / scale);
    err[2 * n + i] = (float)(w * Ve * Ve * Ve * d_df(a.df, Sr, Jphi, Sz));
}

struct IntegArgs {
    NodeBuf nb;
    const int *list_u, *list_v;
    const unsigned int *cnt_u, *cnt_v;
    unsigned int *next;           // task counter
    int *ovf_list;                // overflowed tasks as node * 2 + type
    unsigned int *ovf_cnt;
    double *ws;                   // per-lane workspace
    int cap;                      // intervals per lane
    int ntask_override;           // >= 0: tasks are ovf_list[0..n)
    Counters *ctr;
    const unsigned int *ntask_dev;   // if set: tasks are ovf_list[0 .. *ntask_dev) (count on the device)
    const PotView *pv;               // if set: task tables pv[pmodel[node / NQ]] (batched overflow pass)
    const int *pmodel;
    int *cost;                       // FP32 stepper, if set: [2 node + type] the intervals each finished task took
};


// k_integrals: persistent QAGS over the Ju and Jv task lists (Utils.h:173-185,
// uvOrb.cpp:226-262).  Each loop step every busy lane applies one QK21 rule to
// its own current interval; lanes that finish fetch the next task.
#ifndef FJ_INTEG_WAVES
#define FJ_INTEG_WAVES 3   // waves per SIMD the register allocation of k_integrals aims at
#endif

// PerTaskTables: the tables of each task are ia.pv[ia.pmodel[node / NQ]]
// (batched overflow pass); otherwise f.T stays the kernel's LDS view, which
// keeps the view's fields in registers.
// The real type R of the workspace picks the stepper: double is dqagse with
// the FP64 rule (qk_rule), float QUADPACK's single-precision qagse with the
// FP32 rule (qk_rule32, mixed precision).
template <bool PerTaskTables = false, class FI, class FV, class R>
__device__ __forceinline__ void integrals_loop(const IntegArgs &ia, FI &f, const QagsWorkT<R> &w, const FV &fv) {
    const bool from_list = ia.ntask_override >= 0 || ia.ntask_dev;
    const unsigned int nu = from_list ? 0u : *ia.cnt_u, nv = from_list ? 0u : *ia.cnt_v;
    const unsigned int ntask = ia.ntask_dev ? *ia.ntask_dev
                               : ia.ntask_override >= 0 ? (unsigned int)ia.ntask_override : nu + nv;
    int task = -1;    // -1: needs a task, -2: queue drained
    long long node = 0;
    QagsT<R> q;
    unsigned long long nqk = 0, novf = 0;
    while (true) {
        // fetch (wave-aggregated)
        const bool need = (task == -1);
        const unsigned long long mask = __ballot(need);
        if (mask) {
            const int lane = __lane_id();
            const int leader = __ffsll((long long)mask) - 1;
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(ia.next, (unsigned int)__popcll(mask));
            base = __shfl(base, leader);
            if (need) {
                const unsigned long long lower = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
                const unsigned int t = base + (unsigned int)__popcll(lower);
                if (t < ntask) {
                    int type;
                    if (from_list) {
                        const int code = ia.ovf_list[t];   // node * 2 + type
                        node = code >> 1;
                        type = code & 1;
                    } else {
                        type = t < nu ? 0 : 1;
                        node = type == 0 ? ia.list_u[t] : ia.list_v[t - nu];
                    }
                    task = 0;
                    load_jint(ia.nb, node, type, f);
                    if constexpr (PerTaskTables) f.T = &ia.pv[ia.pmodel[node / NQ]];
                    qags_begin(q, 0, type == 0 ? PI : PIH, EPSABSINTEG, EPSRELINTEG, 1000);
                } else {
                    task = -2;
                }
            }
        }
        if (__ballot(task >= 0) == 0) break;
        if (task >= 0) {
            QKOutT<R> r;
            if constexpr (std::is_same<R, float>::value) r = qk_rule32<11>(f, q.next_a, q.next_b, fv);
            else r = qk_rule<11>(f, q.next_a, q.next_b, fv);
            nqk++;
            qags_consume(q, w, ia.cap, r);
            if (q.state == 3) {
                if constexpr (std::is_same<R, float>::value) {
                    if (ia.cost) ia.cost[2 * node + f.type] = (int)q.last;
                }
                if (q.overflow) {
                    const unsigned int slot = atomicAdd(ia.ovf_cnt, 1u);
                    ia.ovf_list[slot] = (int)(node * 2 + f.type);
                    novf++;
                } else {
                    (f.type == 0 ? ia.nb.Jr : ia.nb.Jz)[node] = f.action(q.result);
                }
                task = -1;
            }
        }
    }
    if (nqk) atomicAdd(&ia.ctr->qk, nqk);
    if (novf) atomicAdd(&ia.ctr->overflow, novf);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FJ_INTEG_WAVES)))
k_integrals(TabArgs ta, IntegArgs ia) {
    extern __shared__ double smem[];
    const LdsTabs L = stage_tables(ta, smem);
    const long long lane_g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const QagsWork w = lane_work(ia.ws, ia.cap, lane_g, (long long)gridDim.x * blockDim.x);
    const FvStrided fv{smem + tab_lds_doubles(ta.nr, ta.npoly, ta.ngrid) + threadIdx.x};
    JInt f;
    f.T = &L.T;
    integrals_loop(ia, f, w, fv);
}

// Mixed precision: FP32 integrand, Kronrod rule and QAGS bookkeeping (the
// float instance of the stepper, QUADPACK's qagse); tasks that reach the
// FP32 interval cap go to the FP64 overflow pass.
#ifndef FJ_INTEG32_WAVES
#define FJ_INTEG32_WAVES 4
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FJ_INTEG32_WAVES)))
k_integrals_f32(TabArgs ta, IntegArgs ia) {
    extern __shared__ double smem[];
    const LdsTabs L = stage_tables(ta, smem);
    float *f32 = reinterpret_cast<float *>(smem + tab_lds_doubles(ta.nr, ta.npoly, ta.ngrid));
    const PotView32 T32 = stage_tables32(L, f32);
    const long long lane_g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const QagsWorkT<float> w =
        lane_work(reinterpret_cast<float *>(ia.ws), ia.cap, lane_g, (long long)gridDim.x * blockDim.x);
    const FvStrided32 fv{f32 + tab32_lds_floats(ta.nr, ta.npoly) + threadIdx.x};
    JInt32 f;
    f.T = &T32;
    integrals_loop(ia, f, w, fv);
}

struct MomArgs {
    const double *Phi, *Ve, *R;
    int p0, npts, quad;
    NodeBuf nb;
    fj_df_params df;
    double *out6;                       // [npts_total][6]
    double *rho, *vrot, *sigR, *sigp, *sigz, *sigRz;   // [npts_total]
    unsigned long long *nonfinite;      // if set: count of nodes with a NaN or infinite f(J)
};

// Adds the block's count of non-finite f(J) values to *ctr (one atomic per block).
__device__ inline void count_nonfinite(unsigned long long *ctr, int bad) {
    if (!ctr) return;
    __shared__ int nb_bad;
    if (threadIdx.x == 0) nb_bad = 0;
    __syncthreads();
    if (bad) atomicAdd(&nb_bad, bad);
    __syncthreads();
    if (threadIdx.x == 0 && nb_bad) atomicAdd(ctr, (unsigned long long)nb_bad);
}


// k_moments: one block per point; fixed-order reduction of the 6 sums.
__device__ inline void moments_point(const MomArgs &ma, int pl, double (*red)[256]);

__global__ void __launch_bounds__(256) k_moments(MomArgs ma) {
    __shared__ double red[6][256];
    moments_point(ma, blockIdx.x, red);
}

// The six sums of point p0 + pl (one block of 256 threads).
__device__ inline void moments_point(const MomArgs &ma, int pl, double (*red)[256]) {
    const int p = ma.p0 + pl;
    const double R = ma.R[p], Ve = ma.Ve[p];
    const double jacob = Ve * Ve * Ve;
    double s[6] = {0, 0, 0, 0, 0, 0};
    int bad = 0;
    for (int node = threadIdx.x; node < NQ; node += blockDim.x) {
        const long long gid = (long long)pl * NQ + node;
        const int flag = ma.nb.flag[gid];
        if (flag == 0) continue;   // H >= 0: df returns 0 (DF.cpp:84)
        double V[3], w;
        node_velocity(ma.quad, node, Ve, V, w);
        const double Jphi = R * V[1];
        const double DF = d_df(ma.df, ma.nb.Jr[gid], Jphi, ma.nb.Jz[gid]);
        bad += !isfinite(DF);
        const double jd = jacob * DF;
        s[0] += w * jd;
        s[1] += w * (jd * V[1]);
        s[2] += w * (jd * V[0] * V[0]);
        s[3] += w * (jd * V[1] * V[1]);
        s[4] += w * (jd * V[2] * V[2]);
        s[5] += w * (jd * V[0] * V[2]);
    }
    count_nonfinite(ma.nonfinite, bad);
    for (int m = 0; m < 6; m++) red[m][threadIdx.x] = s[m];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int m = 0; m < 6; m++) red[m][threadIdx.x] += red[m][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double o[6];
        for (int m = 0; m < 6; m++) { o[m] = red[m][0]; ma.out6[(long long)p * 6 + m] = o[m]; }
        // Integ.cpp:160-161
        ma.vrot[p] = o[1] / o[0];
        ma.sigR[p] = sqrt(o[2] / o[0]);
        ma.sigp[p] = sqrt(o[3] / o[0]);
        ma.sigz[p] = sqrt(o[4] / o[0]);
        ma.sigRz[p] = sqrt(fabs(o[5]) / o[0]);
        ma.rho[p] = ma.quad ? 4 * o[0] : o[0];
    }
}

// Mixed precision: f(J) in FP32 with compensated FP32 sums; the pairwise
// tree adds the (sum, compensation) pairs with an exact two-sum.
__device__ inline void moments_point_f32(const MomArgs &ma, int pl, float (*red)[2][256]);

__global__ void __launch_bounds__(256) k_moments_f32(MomArgs ma) {
    __shared__ float red[6][2][256];
    moments_point_f32(ma, blockIdx.x, red);
}

__device__ inline void moments_point_f32(const MomArgs &ma, int pl, float (*red)[2][256]) {
    const int p = ma.p0 + pl;
    const double R = ma.R[p], Ve = ma.Ve[p];
    const float jacob = (float)(Ve * Ve * Ve);
    DF32 d;
    d.A = (float)ma.df.A; d.B = (float)ma.df.B;
    d.dphi_h = (float)ma.df.dphi_h; d.dz_h = (float)ma.df.dz_h;
    d.dphi_g = (float)ma.df.dphi_g; d.dz_g = (float)ma.df.dz_g;
    d.norm = (float)(ma.df.mass / (ma.df.J0 * ma.df.J0 * ma.df.J0) / (TPI * TPI * TPI));
    d.J0 = (float)ma.df.J0; d.invJ0 = (float)(1.0 / ma.df.J0);
    d.chi = (float)ma.df.chi; d.k = (float)ma.df.k; d.rot_form = ma.df.rot_form;
    KSum s[6];
    int bad = 0;
    for (int node = threadIdx.x; node < NQ; node += blockDim.x) {
        const long long gid = (long long)pl * NQ + node;
        if (ma.nb.flag[gid] == 0) continue;
        double Vd[3], wd;
        node_velocity(ma.quad, node, Ve, Vd, wd);
        const float V0 = (float)Vd[0], V1 = (float)Vd[1], V2 = (float)Vd[2], w = (float)wd;
        const float DF = df32(d, (float)ma.nb.Jr[gid], (float)(R * Vd[1]), (float)ma.nb.Jz[gid]);
        bad += !isfinite(DF);
        const float jd = jacob * DF;
        s[0].add(w * jd);
        s[1].add(w * (jd * V1));
        s[2].add(w * (jd * V0 * V0));
        s[3].add(w * (jd * V1 * V1));
        s[4].add(w * (jd * V2 * V2));
        s[5].add(w * (jd * V0 * V2));
    }
    count_nonfinite(ma.nonfinite, bad);
    for (int m = 0; m < 6; m++) { red[m][0][threadIdx.x] = s[m].s; red[m][1][threadIdx.x] = s[m].c; }
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int m = 0; m < 6; m++) {
                const float a = red[m][0][threadIdx.x], b = red[m][0][threadIdx.x + st];
                const float t = a + b, bb = t - a, e = (a - (t - bb)) + (b - bb);   // a + b = t + e exactly
                red[m][0][threadIdx.x] = t;
                red[m][1][threadIdx.x] = red[m][1][threadIdx.x] + red[m][1][threadIdx.x + st] - e;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double o[6];
        for (int m = 0; m < 6; m++) {
            o[m] = (double)red[m][0][0] - (double)red[m][1][0];
            ma.out6[(long long)p * 6 + m] = o[m];
        }
        ma.vrot[p] = o[1] / o[0];
        ma.sigR[p] = sqrt(o[2] / o[0]);
        ma.sigp[p] = sqrt(o[3] / o[0]);
        ma.sigz[p] = sqrt(o[4] / o[0]);
        ma.sigRz[p] = sqrt(fabs(o[5]) / o[0]);
        ma.rho[p] = ma.quad ? 4 * o[0] : o[0];
    }
}

// ---------------------------------------------------------------------------
// Fast node phase
// ---------------------------------------------------------------------------
constexpr int FAST_GL_MAX = 128;
__constant__ FastGL32 c_fgl[FAST_GL_MAX];    // J_r rule (ngl points)
__constant__ FastGL32 c_fglv[FAST_GL_MAX];   // J_z rule (fast_v_order(ngl) points)

struct FastArgs {
    const double *R, *z, *Phi, *Ve;   // [npts]
    int npts, quad, S, ngl, nglv;     // S slices of the velocity nodes per point; the two rules' orders
    fj_df_params df;
    double *part;                     // [npts][S][6] slice sums
    Counters *ctr;
};

// Block b: tile b % tiles of 64 grid points, node slice b / tiles.  Lane l of
// each wave takes point 64 tile + l, so a wave follows one velocity node at
// 64 neighbouring points (orbits of like shape); the 4 waves take every 4th
// node of the slice.  Per lane, FP64 sums over its nodes; the 4 waves' sums
// are added in wave order and stored as the slice's sums.
#ifndef FJ_FAST_BLOCKS_PER_CU
#define FJ_FAST_BLOCKS_PER_CU 32
#endif
#ifndef FJ_FAST_WALK32
#define FJ_FAST_WALK32 1   // walks and root solves on FP32 evaluations, bracket ends checked in FP64
#endif
#ifndef FJ_FAST_WAVES
#define FJ_FAST_WAVES 4   // waves per SIMD the register allocation of k_fast aims at
#endif
// The work of one k_fast block: tile `tile` (64 points) of the fa.npts points
// that start at global point pbase, node slice sl of fa.S; tables in L / T32,
// red: 4 x 6 x 64 doubles of LDS.
__device__ __forceinline__ void fast_block(const LdsTabs &L, const PotView32 &T32, double *red, const FastArgs &fa,
                                           int tile, int sl, int pbase) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pl = tile * 64 + lane;
    const bool valid = pl < fa.npts;
    const int p = pbase + pl;
    const int QS = (NQ + fa.S - 1) / fa.S;
    const int q0 = sl * QS, q1 = q0 + QS < NQ ? q0 + QS : NQ;
    // the lane's six sums live in its LDS slots of red: fewer registers held
    // across the node loop
    double *acc = red + wave * 6 * 64 + lane;
    for (int m = 0; m < 6; m++) acc[m * 64] = 0;
    int nbound = 0, nunb = 0, nbad = 0;
    long long nev = 0;
    if (valid) {
        const double R = fa.R[p], z = fa.z[p], Phi_h = fa.Phi[p], Ve = fa.Ve[p];
        const double jacob = Ve * Ve * Ve;
        for (int q = q0 + wave; q < q1; q += 4) {
            double V[3], w;
            node_velocity(fa.quad, q, Ve, V, w);
            FastNode fn;
            int ne = 0;
#if FJ_FAST_WALK32
            const int flag = fast_setup<FastPolicy32>(L.T, &T32, L.E, L.D, L.ngrid, R, z, Phi_h, V, fn, &ne);
#else
            const int flag = fast_setup<FastPolicy64>(L.T, L.T, L.E, L.D, L.ngrid, R, z, Phi_h, V, fn, &ne);
#endif
            nev += ne;
            if (flag == 0) continue;   // H >= 0: f = 0 (DF.cpp:84)
            double Jr = 1e31, Jz = 1e31;
            if (flag == 1) {
                fast_actions32(T32, fn, c_fgl, fa.ngl, c_fglv, fa.nglv, &Jr, &Jz);
                nbound++;
                nev += (long long)fn.ju * fa.ngl + (long long)fn.jv * fa.nglv;
            } else {
                nunb++;
            }
            const double DF = d_df(fa.df, Jr, R * V[1], Jz);
            nbad += !isfinite(DF);
            const double jd = jacob * DF;
            acc[0] += w * jd;
            acc[64] += w * (jd * V[1]);
            acc[128] += w * (jd * V[0] * V[0]);
            acc[192] += w * (jd * V[1] * V[1]);
            acc[256] += w * (jd * V[2] * V[2]);
            acc[320] += w * (jd * V[0] * V[2]);
        }
    }
    __syncthreads();
    if (wave == 0 && valid) {
        double *o = fa.part + ((size_t)p * fa.S + sl) * 6;
        for (int m = 0; m < 6; m++) {
            double t = red[m * 64 + lane];
            for (int w2 = 1; w2 < 4; w2++) t += red[(w2 * 6 + m) * 64 + lane];
            o[m] = t;
        }
    }
    // counters: one atomic per wave
    unsigned long long b = nbound, u = nunb, nb = nbad, ev = (unsigned long long)nev;
    for (int off = 32; off > 0; off >>= 1) {
        b += __shfl_down(b, off); u += __shfl_down(u, off);
        nb += __shfl_down(nb, off); ev += __shfl_down(ev, off);
    }
    if (lane == 0) {
        if (b) atomicAdd(&fa.ctr->bound, b);
        if (u) atomicAdd(&fa.ctr->unbound, u);
        if (nb) atomicAdd(&fa.ctr->nonfinite, nb);
        if (ev) atomicAdd(&fa.ctr->phi, ev);
    }
}

// LDS of a fast-phase block: FP64 tables, FP32 tables with Phi_l', the
// per-wave sums.
__host__ __device__ inline size_t fast_lds_bytes(int nr, int npoly, int ngrid) {
    return tab_lds_doubles(nr, npoly, ngrid) * sizeof(double) +
           ((tab32_lds_floats(nr, npoly, true) + 1) & ~(size_t)1) * sizeof(float) + 4 * 6 * 64 * sizeof(double);
}

__device__ __forceinline__ double *fast_stage(const TabArgs &ta, double *smem, LdsTabs &L, PotView32 &T32) {
    L = stage_tables(ta, smem);
    float *f32 = reinterpret_cast<float *>(smem + tab_lds_doubles(ta.nr, ta.npoly, ta.ngrid));
    T32 = stage_tables32(L, f32, true);
    return reinterpret_cast<double *>(f32 + ((tab32_lds_floats(ta.nr, ta.npoly, true) + 1) & ~(size_t)1));
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FJ_FAST_WAVES)))
k_fast(TabArgs ta, FastArgs fa) {
    extern __shared__ double smem[];
    LdsTabs L;
    PotView32 T32;
    double *red = fast_stage(ta, smem, L, T32);
    const int tiles = (fa.npts + 63) / 64;
    fast_block(L, T32, red, fa, blockIdx.x % tiles, blockIdx.x / tiles, 0);
}

// Fast node phase with one lane per velocity node, in the order sel (or the
// tiled order of k_setup): node setup and the FP32 action rules of node
// gid, its flag and actions into the node-indexed records (k_moments adds
// the DF), and its sort key for the next pass's order.  The records do not
// depend on the lane order.
struct FastNodeArgs {
    const double *R, *z, *Phi, *Ve;   // [npts]
    int npts, quad, ngl, nglv;
    NodeBuf nb;
    const int *sel;                   // [npts * NQ] lane order, or null
    unsigned int *key;                // [npts * NQ] keys of this pass
    Counters *ctr;
};
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FJ_FAST_WAVES)))
k_fast_nodes(TabArgs ta, FastNodeArgs fa) {
    extern __shared__ double smem[];
    LdsTabs L;
    PotView32 T32;
    (void)fast_stage(ta, smem, L, T32);
    const long long nn = (long long)fa.npts * NQ;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < nn;
    int flag = 0;
    long long nev = 0;
    if (valid) {
        const long long gid = fa.sel ? (long long)fa.sel[t] : setup_lane_node(t, fa.npts);
        const int p = (int)(gid / NQ), q = (int)(gid % NQ);
        const double R = fa.R[p], z = fa.z[p], Phi_h = fa.Phi[p], Ve = fa.Ve[p];
        double V[3], w;
        node_velocity(fa.quad, q, Ve, V, w);
        FastNode fn;
        int ne = 0;
        flag = fast_setup<FastPolicy32>(L.T, &T32, L.E, L.D, L.ngrid, R, z, Phi_h, V, fn, &ne);
        nev = ne;
        double Jr = 0, Jz = 0;
        if (flag == 1) {
            fast_actions32(T32, fn, c_fgl, fa.ngl, c_fglv, fa.nglv, &Jr, &Jz);
            nev += (long long)fn.ju * fa.ngl + (long long)fn.jv * fa.nglv;
        } else if (flag == 2) {
            Jr = 1e31; Jz = 1e31;
        }
        fa.nb.flag[gid] = flag;
        fa.nb.Jr[gid] = Jr;
        fa.nb.Jz[gid] = Jz;
        fa.key[gid] = flag == 1 ? node_order_key(1, fn.uin, fn.uout, fn.vt) : node_order_key(flag, 0, 0, 0);
    }
    unsigned long long b = valid && flag == 1, u = valid && flag == 2, ev = (unsigned long long)nev;
    for (int off = 32; off > 0; off >>= 1) {
        b += __shfl_down(b, off); u += __shfl_down(u, off); ev += __shfl_down(ev, off);
    }
    if (__lane_id() == 0) {
        if (b) atomicAdd(&fa.ctr->bound, b);
        if (u) atomicAdd(&fa.ctr->unbound, u);
        if (ev) atomicAdd(&fa.ctr->phi, ev);
    }
}

// Sums of the S slices of each point, then the moments of rhofDF
// (Integ.cpp:160-161) as k_moments forms them.  One block per point: lane t
// adds slices t, t + 256, ... in order, then a pairwise tree over the lanes
// (a fixed order).
__global__ void __launch_bounds__(256) k_fast_finish(int npts, int S, int quad, const double *part, double *out6,
                                                     double *mom) {
    __shared__ double red[6][256];
    const int p = blockIdx.x, t = threadIdx.x;
    double o[6] = {0, 0, 0, 0, 0, 0};
    for (int s = t; s < S; s += 256)
        for (int m = 0; m < 6; m++) o[m] += part[((size_t)p * S + s) * 6 + m];
    for (int m = 0; m < 6; m++) red[m][t] = o[m];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st)
            for (int m = 0; m < 6; m++) red[m][t] += red[m][t + st];
        __syncthreads();
    }
    if (t != 0) return;
    for (int m = 0; m < 6; m++) o[m] = red[m][0];
    for (int m = 0; m < 6; m++) out6[(size_t)p * 6 + m] = o[m];
    mom[p] = quad ? 4 * o[0] : o[0];
    mom[npts + p] = o[1] / o[0];
    mom[2 * npts + p] = sqrt(o[2] / o[0]);
    mom[3 * npts + p] = sqrt(o[3] / o[0]);
    mom[4 * npts + p] = sqrt(o[4] / o[0]);
    mom[5 * npts + p] = sqrt(fabs(o[5]) / o[0]);
}

// ---------------------------------------------------------------------------
// Grid phase
// ---------------------------------------------------------------------------
// get_Ylm (UtilsLeg.h:200-253) with the rho[0][0] fix (potLeg.cpp:68).
__global__ void k_project(int nr, int ng, int np_, const double *ar, const double *poly,
                          const double *rho, const double *vrot, const double *sigR, const double *sigp,
                          const double *sigz, const double *sigRz, double *rhL, double *vrotL, double *sigRL,
                          double *sigpL, double *sigzL, double *sigRzL) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nr * np_) return;
    const int n = idx / np_, np = idx % np_;
    if (n == 0) {
        if (np == 0) {
            const double rho00 = (rho[2 * ng] - rho[1 * ng]) / (ar[2] - ar[1]) * (ar[0] - ar[1]) + rho[1 * ng];
            rhL[0] = 2 * rho00; vrotL[0] = 2 * vrot[0]; sigRL[0] = 2 * sigR[0];
            sigpL[0] = 2 * sigp[0]; sigzL[0] = 2 * sigz[0]; sigRzL[0] = 2 * sigRz[0];
        } else {
            rhL[np] = 0; vrotL[np] = 0; sigRL[np] = 0; sigpL[np] = 0; sigzL[np] = 0; sigRzL[np] = 0;
        }
        return;
    }
    double a = 0, b = 0, c = 0, d = 0, e = 0, f = 0;
    for (int g = 0; g < ng; g++) {
        const double wv = poly[np * ng + g];
        const int i = n * ng + g;
        a += wv * rho[i]; b += wv * vrot[i]; c += wv * sigR[i];
        d += wv * sigp[i]; e += wv * sigz[i]; f += wv * sigRz[i];
    }
    rhL[idx] = a; vrotL[idx] = b; sigRL[idx] = c; sigpL[idx] = d; sigzL[idx] = e; sigRzL[idx] = f;
}

// Guess densities (pot/models.cpp:40-68)
__device__ inline double d_rho_guess(int model, double mass, double r0, double q, double q2, double R, double z) {
    const double m = sqrt(R * R + z * z / q2);
    switch (model) {
    case 0: return mass / (m / r0 * fjm_powi(r0 + m, 3)) / TPI;
    case 1: {
        const double a = sqrt(r0 * r0 + m * m);
        return mass / FPI / (q * fjm_powi((a + r0) * a, 2) * a) * r0 * (r0 + 2 * a);
    }
    case 2: return mass / (m / r0 * fjm_powi(r0 + m, 2)) / FPI;
    default: return .01 / (m / 25. * fjm_powi(25. + m, 1)) / FPI;
    }
}

// computeGuessRhl (Potential.cpp:63-80)
__global__ void k_guess(int nr, int ng, int np_, const double *ar, const double *ci, const double *si,
                        const double *poly, int model, double mass, double r0, double q, double *rhl) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nr * np_) return;
    const int n = idx / np_, np = idx % np_;
    const double q2 = q * q;
    double s = 0.;
    for (int g = 0; g < ng; g++) s += poly[np * ng + g] * d_rho_guess(model, mass, r0, q, q2, ar[n] * si[g], ar[n] * ci[g]);
    if (idx == 0) s = 2 * d_rho_guess(model, mass, r0, q, q2, ar[0] * si[0], ar[0] * ci[0]);
    rhl[idx] = s;
}

// I_internal / I_external (Potential.cpp:83-101) on column np of rhl.
struct IInt {
    const double *ar, *rhl; int nr, np_, col, l;
    __device__ double val(double x) const {
        if (x <= ar[0]) return rhl[col];
        if (x >= ar[nr - 1]) return rhl[(nr - 1) * np_ + col];
        int top = nr - 1, bot = 0;
        while (top - bot > 1) { int m = (top + bot) >> 1; if (x >= ar[m]) bot = m; else top = m; }
        const double yb = rhl[bot * np_ + col], yt = rhl[top * np_ + col];
        return yb + (x - ar[bot]) / (ar[top] - ar[bot]) * (yt - yb);
    }
    __device__ double operator()(double x) const { return fjm_powi(x, l + 2) * val(x); }
};
struct IExtT {   // QAGIU transform of I_external: x = a + (1-t)/t, (f(x)/t)/t
    IInt in; double a;
    __device__ double operator()(double t) const {
        const double x = a + (1 - t) / t;
        const int nr = in.nr;
        double rhol;
        if (x > in.ar[nr - 1]) rhol = fj_max(0., in.rhl[(nr - 1) * in.np_ + in.col] * fjm_powi(in.ar[nr - 1] / x, 4));
        else rhol = in.val(x);
        const double y = rhol / fjm_powi(x, in.l - 1);
        return (y / t) / t;
    }
};

// Kronrod rule of qk_rule with the integrand spread over the lanes of one
// wave: lane s < 2(NK-1) evaluates the off-centre abscissa s of qk_rule's
// visiting order, lane 2(NK-1) the centre; the values meet in `sh` (LDS, one
// wave per block) and every lane then takes the sums in qk_rule's order, so
// the result carries qk_rule's bits and is the same in all lanes.
template <int NK, class F>
__device__ __forceinline__ QKOut qk_rule_wave(const F &f, double a, double b, double *sh) {
    const double *xgk = NK == 11 ? kXGK21 : kXGK15;
    const double *wgk = NK == 11 ? kWGK21 : kWGK15;
    const double *wg = NK == 11 ? kWG10 : kWG7;
    constexpr int NG = (NK - 1) / 2;
    constexpr int NV = 2 * (NK - 1);
    const double center = 0.5 * (a + b), half = 0.5 * (b - a), ahalf = fabs(half);
    const int t = threadIdx.x;
    if (t <= NV) {
        double x = center;
        if (t < NV) {
            const int m = t >> 1;
            const int jj = m < NG ? 2 * m + 1 : 2 * (m - NG);
            const double absc = half * xgk[jj];
            x = (t & 1) ? center + absc : center - absc;
        }
        sh[t] = f(x);
    }
    __syncthreads();
    const double fc = sh[NV];
    double rg = 0, rk = fc * wgk[NK - 1], rabs = fabs(rk);
    if (NK % 2 == 0) rg = fc * wg[NK / 2 - 1];
#pragma unroll
    for (int m = 0; m < NK - 1; m++) {
        const int jj = m < NG ? 2 * m + 1 : 2 * (m - NG);
        const double f1 = sh[2 * m], fx = sh[2 * m + 1];
        const double fs = f1 + fx;
        if (m < NG) rg += wg[m] * fs;
        rk += wgk[jj] * fs;
        rabs += wgk[jj] * (fabs(f1) + fabs(fx));
    }
    const double mean = rk * 0.5;
    double rasc = wgk[NK - 1] * fabs(fc - mean);
#pragma unroll
    for (int j = 0; j < NK - 1; j++) {
        const int m = (j & 1) ? (j >> 1) : NG + (j >> 1);
        rasc += wgk[j] * (fabs(sh[2 * m] - mean) + fabs(sh[2 * m + 1] - mean));
    }
    __syncthreads();   // sh is rewritten by the next rule
    const double err = (rk - rg) * half;
    QKOut o;
    o.result = rk * half;
    o.resabs = rabs * ahalf;
    o.resasc = rasc * ahalf;
    o.abserr = qk_rescale(err, o.resabs, o.resasc);
    return o;
}

// qags_solve with qk_rule_wave.  Every lane of the wave runs the same
// (uniform) QAGS state; the workspace stores are the same value at the same
// address from all lanes.
template <int NK, class F>
__device__ __forceinline__ double qags_solve_wave(const F &f, double a, double b, double epsabs, double epsrel,
                                                  int limit, const QagsWork &w, int cap, double *sh,
                                                  int *ier = nullptr) {
    Qags q;
    qags_begin(q, a, b, epsabs, epsrel, limit);
    while (q.state != 3) {
        const QKOut r = qk_rule_wave<NK>(f, q.next_a, q.next_b, sh);
        qags_consume(q, w, cap, r);
    }
    if (ier) *ier = q.overflow ? -1 : q.ier;
    return q.result;
}

// computeInts + computePhil (Potential.cpp:110-200), one wave per (n, np):
// the 21 (QK21, internal integral) or 15 (QK15 of QAGIU's transform,
// external integral) integrand values of every rule are taken by the wave's
// lanes at once; the QAGS bookkeeping is the lane-serial one, so the
// coefficients are those of one lane per integral, bit for bit.
__global__ void __launch_bounds__(64) k_poisson(int nr, int np_, const double *ar, const double *rhl, double *ws,
                                                int cap, double *philP, double *PrP, double *Pr2P, int *status) {
    __shared__ double sh[32];
    const int idx = blockIdx.x;
    if (idx >= nr * np_) return;
    const int n = idx / np_, np = idx % np_;
    const int l = 2 * np;
    const QagsWork w = lane_work(ws, cap, idx, (long long)gridDim.x);
    IInt fi{ar, rhl, nr, np_, np, l};
    const double Iint = qags_solve_wave<11>(fi, 0., ar[n], 1e-4, 1e-3, 1000, w, cap, sh);
    IExtT fe{fi, ar[n]};
    double epsabs = 1e-4, epsrel = 1e-3, Iext = 0;
    int ier = 1, guard = 0;
    while (ier != 0) {
        Iext = qags_solve_wave<8>(fe, 0.0, 1.0, epsabs, epsrel, 1000, w, cap, sh, &ier);
        epsabs *= 2; epsrel *= 2;
        if (isnan(Iext) || ++guard > 200) {
            if (threadIdx.x == 0) atomicOr(status, 1);
            break;
        }
    }
    if (threadIdx.x != 0) return;
    const double G = 1, rn = ar[n];
    philP[idx] = -TPI * G * (Iint / fjm_powi(rn, l + 1) + Iext * fjm_powi(rn, l));
    PrP[idx] = -TPI * G * (-(l + 1) * Iint / fjm_powi(rn, l + 2) + l * Iext * fjm_powi(rn, l - 1));
    Pr2P[idx] = -TPI * G * ((l + 2) * (l + 1) * Iint / fjm_powi(rn, l + 3) + (l - 1) * l * Iext * fjm_powi(rn, l - 2) -
                            (2 * l + 1) * rhl[idx]);
}

// computePhil (Potential.cpp:110-200) with the radial integrals taken
// exactly for the piecewise-linear rho_l that I_internal / I_external
// integrate (linterp, constant below ar[0], max(0, rho_l(r_max)) (r_max/x)^4
// beyond r_max): closed-form segment integrals, then a block-wide inclusive
// scan (I_int, from the centre out) and a reverse scan (I_ext, from the
// outside in).  One block per Legendre order; FP64 throughout.  The result
// differs from the QUADPACK path by that path's quadrature error.
constexpr int PSCAN_THREADS = 256;
constexpr int PSCAN_MAXNR = 4 * PSCAN_THREADS;

__device__ inline double ipow_signed(double x, int k) { return fjm_powi(x, k); }

// int_{x0}^{x1} x^k dx
__device__ inline double seg_pow_int(double x0, double x1, int k) {
    if (k == -1) return fjm_log(x1 / x0);
    return (ipow_signed(x1, k + 1) - ipow_signed(x0, k + 1)) / (k + 1);
}

// inclusive scan of v[0..n) (n <= PSCAN_MAXNR) in place; reverse = suffix sums
__device__ void block_scan(double *v, int n, bool reverse, double *part) {
    const int t = threadIdx.x;
    const int per = (n + PSCAN_THREADS - 1) / PSCAN_THREADS;
    const int i0 = t * per, i1 = min(n, i0 + per);
    // serial scan of this thread's run
    double acc = 0;
    if (!reverse) for (int i = i0; i < i1; i++) { acc += v[i]; v[i] = acc; }
    else for (int i = i1 - 1; i >= i0; i--) { acc += v[i]; v[i] = acc; }
    part[t] = acc;
    __syncthreads();
    // Hillis-Steele over the run totals
    for (int off = 1; off < PSCAN_THREADS; off <<= 1) {
        double add = 0;
        if (!reverse) { if (t >= off) add = part[t - off]; }
        else { if (t + off < PSCAN_THREADS) add = part[t + off]; }
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    // offset of this run: total of the runs before it (after it for reverse)
    double off = 0;
    if (!reverse) { if (t > 0) off = part[t - 1]; }
    else { if (t + 1 < PSCAN_THREADS) off = part[t + 1]; }
    for (int i = i0; i < i1; i++) v[i] += off;
    __syncthreads();
}

__device__ inline void poisson_scan_col(int nr, int np_, const double *ar, const double *rhl, double *philP,
                                        double *PrP, double *Pr2P, int np, double *Iint, double *Iext, double *part);

__global__ void __launch_bounds__(PSCAN_THREADS) k_poisson_scan(int nr, int np_, const double *ar, const double *rhl,
                                                              double *philP, double *PrP, double *Pr2P) {
    __shared__ double Iint[PSCAN_MAXNR], Iext[PSCAN_MAXNR], part[PSCAN_THREADS];
    poisson_scan_col(nr, np_, ar, rhl, philP, PrP, Pr2P, blockIdx.x, Iint, Iext, part);
}

// Column np (order l = 2 np) of the block-scan Poisson solve, one block.
__device__ inline void poisson_scan_col(int nr, int np_, const double *ar, const double *rhl, double *philP,
                                        double *PrP, double *Pr2P, int np, double *Iint, double *Iext, double *part) {
    const int l = 2 * np;
    // segment i = [ar[i], ar[i+1]] stored at i+1 (internal) and i (external)
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
        double si = 0, se = 0;
        if (i == 0) {
            si = rhl[np] * ipow_signed(ar[0], l + 3) / (l + 3);   // constant below ar[0]
        } else {
            const double x0 = ar[i - 1], x1 = ar[i];
            const double y0 = rhl[(i - 1) * np_ + np], y1 = rhl[i * np_ + np];
            const double s = (y1 - y0) / (x1 - x0), c0 = y0 - s * x0;
            si = c0 * seg_pow_int(x0, x1, l + 2) + s * seg_pow_int(x0, x1, l + 3);
        }
        if (i < nr - 1) {
            const double x0 = ar[i], x1 = ar[i + 1];
            const double y0 = rhl[i * np_ + np], y1 = rhl[(i + 1) * np_ + np];
            const double s = (y1 - y0) / (x1 - x0), c0 = y0 - s * x0;
            se = c0 * seg_pow_int(x0, x1, 1 - l) + s * seg_pow_int(x0, x1, 2 - l);
        } else {
            const double rm = ar[nr - 1], C = fj_max(0., rhl[(nr - 1) * np_ + np]);
            se = C * ipow_signed(rm, 2 - l) / (l + 2);            // r^-4 tail beyond r_max
        }
        Iint[i] = si;
        Iext[i] = se;
    }
    __syncthreads();
    block_scan(Iint, nr, false, part);
    block_scan(Iext, nr, true, part);
    const double G = 1;
    for (int n = threadIdx.x; n < nr; n += blockDim.x) {
        const double rn = ar[n], Ii = Iint[n], Ie = Iext[n];
        const int idx = n * np_ + np;
        philP[idx] = -TPI * G * (Ii / fjm_powi(rn, l + 1) + Ie * fjm_powi(rn, l));
        PrP[idx] = -TPI * G * (-(l + 1) * Ii / fjm_powi(rn, l + 2) + l * Ie * fjm_powi(rn, l - 1));
        Pr2P[idx] = -TPI * G * ((l + 2) * (l + 1) * Ii / fjm_powi(rn, l + 3) + (l - 1) * l * Ie * fjm_powi(rn, l - 2) -
                                (2 * l + 1) * rhl[idx]);
    }
}

// Potential evaluation at a batch of points (facade, tests).
__global__ void __launch_bounds__(256) k_eval(TabArgs ta, int n, const double *R, const double *z, double *phi,
                                              double *dR, double *dz) {
    extern __shared__ double smem[];
    const LdsTabs L = stage_tables(ta, smem);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dR) {
        double a, b;
        pot_phi_grad(L.T, R[i], z[i], &a, &b);
        dR[i] = a; dz[i] = b;
    }
    phi[i] = pot_phi(L.T, R[i], z[i]);
}

// ---------------------------------------------------------------------------
// Virial diagnostic: libs/Stats.cpp:12-41 with ev_dens (UtilsLeg.h:300-354).
// One lane per midpoint cell; the sums run in the reference's order in one lane.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_vir_terms(TabArgs ta, int ng, const double *ci, const double *wi,
                                                   const double *rhl, const double *sigRl, const double *sigpl,
                                                   const double *sigzl, const double *vrotl, double *terms) {
    extern __shared__ double smem[];
    const LdsTabs L = stage_tables(ta, smem);
    const int nr = L.T.nr, np_ = L.T.npoly;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncell = (nr - 1) * ng;
    if (idx >= ncell + nr) return;
    const double *ar = L.T.ar;
    if (idx >= ncell) {
        // ev_dens(ar[i], 0) with the double-weight int_dens (UtilsLeg.h:285-298)
        const int i = idx - ncell;
        const double R = ar[i], z = 0;
        const double rr = sqrt(R * R + z * z);
        const double c = z / rr, c2 = c * c;
        int top = nr - 1, bot = 0;
        while (top - bot > 1) { int m = (top + bot) >> 1; if (rr >= ar[m]) bot = m; else top = m; }
        const double f = (rr - ar[bot]) / (ar[top] - ar[bot]);
        double dens = .5 * (f * rhl[top * np_] + (1 - f) * rhl[bot * np_]);
        double pm2 = 1.0, pm1 = 1.5 * c2 - .5;
        for (int np = 1; np < np_; np++) {
            double pe = pm1;
            if (np >= 2) { pe = evenleg_step(L.T.lc, np, pm2, pm1, c2); pm2 = pm1; pm1 = pe; }
            const double dp = f * rhl[top * np_ + np] + (1 - f) * rhl[bot * np_ + np];
            dens += .5 * (4 * np + 1) * dp * pe;
        }
        terms[5 * ncell + i] = 2 * TPI * ar[i] * ar[i] * dens;
        return;
    }
    const int i = idx / ng + 1, j = idx % ng;
    const double r = .5 * (ar[i - 1] + ar[i]), dr = ar[i] - ar[i - 1];
    const double sij = sqrt(1 - ci[j] * ci[j]);
    const double z = r * ci[j], R = r * sij;
    const double rr = sqrt(R * R + z * z);
    const double c = z / rr, c2 = c * c;
    double dens = 0, Sigu = 0, Sigp = 0, Sigv = 0;
    if (!(rr > ar[nr - 1])) {
        int top = nr - 1, bot = 0;
        while (top - bot > 1) { int m = (top + bot) >> 1; if (rr >= ar[m]) bot = m; else top = m; }
        const float f = (rr - ar[bot]) / (ar[top] - ar[bot]);   // float weight as UtilsLeg.h:312
        const float omf = 1 - f;
        double pm2 = 1.0, pm1 = 1.5 * c2 - .5;
        for (int np = 0; np < np_; np++) {
            const int t = top * np_ + np, b = bot * np_ + np;
            const double d0 = f * rhl[t] + omf * rhl[b];
            const double s1 = f * sigRl[t] + omf * sigRl[b];
            const double s2 = f * sigpl[t] + omf * sigpl[b];
            const double s3 = f * sigzl[t] + omf * sigzl[b];
            if (np == 0) { dens = .5 * d0; Sigu = .5 * s1; Sigp = .5 * s2; Sigv = .5 * s3; continue; }
            double pe = pm1;
            if (np >= 2) { pe = evenleg_step(L.T.lc, np, pm2, pm1, c2); pm2 = pm1; pm1 = pe; }
            const double fac = .5 * (4 * np + 1);
            dens += fac * d0 * pe; Sigu += fac * s1 * pe; Sigp += fac * s2 * pe; Sigv += fac * s3 * pe;
        }
    }
    double dRv, dzv;
    pot_phi_grad(L.T, R, z, &dRv, &dzv);
    const double phi = pot_phi(L.T, R, z);
    terms[0 * ncell + idx] = .5 * wi[j] * dens * phi * r * r * dr;
    terms[1 * ncell + idx] = .5 * wi[j] * dens * ((r * Sigu) * (r * Sigu) + (r * Sigp) * (r * Sigp)) * dr;
    terms[2 * ncell + idx] = .5 * wi[j] * dens * ((r * Sigv) * (r * Sigv)) * dr;
    terms[3 * ncell + idx] = wi[j] * dens * R * dRv * r * r * dr;
    terms[4 * ncell + idx] = wi[j] * dens * z * dzv * r * r * dr;
}

__global__ void k_vir_sum(int nr, int ng, const double *ar, const double *Pr, int np_, const double *terms,
                          double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int ncell = (nr - 1) * ng;
    double pot = 0, KRR = 0, Kzz = 0, WRR = 0, Wzz = 0, sum = 0;
    for (int k = 0; k < ncell; k++) {
        pot += terms[k]; KRR += terms[ncell + k]; Kzz += terms[2 * ncell + k];
        WRR -= terms[3 * ncell + k]; Wzz -= terms[4 * ncell + k];
    }
    KRR *= 4 * PI; Kzz *= 4 * PI; WRR *= 4 * PI; Wzz *= 4 * PI; pot *= 4 * PI;
    for (int i = 0; i < nr; i++) sum += terms[5 * ncell + i];
    out[0] = ar[nr - 1] * ar[nr - 1] * Pr[(nr - 1) * np_];
    out[1] = sum; out[2] = KRR + Kzz; out[3] = pot; out[4] = pot / (KRR + Kzz);
    out[5] = KRR; out[6] = WRR; out[7] = WRR / KRR; out[8] = Kzz; out[9] = Wzz; out[10] = Wzz / Kzz;
}

// ---------------------------------------------------------------------------
// Focal distance table: delta/tabulate.cpp:13-38, delta/Delta.cpp:18-294,
// include/Utils.h:217-263.  One lane per energy; RK8PD (Prince & Dormand
// 1981) with the standard y-control step rule; orbit samples in global scratch.
// ---------------------------------------------------------------------------
constexpr double c_rk_c[13] = {0.0, 1.0 / 18.0, 1.0 / 12.0, 1.0 / 8.0, 5.0 / 16.0, 3.0 / 8.0, 59.0 / 400.0,
                                  93.0 / 200.0, 5490023248.0 / 9719169821.0, 13.0 / 20.0,
                                  1201146811.0 / 1299019798.0, 1.0, 1.0};
constexpr double c_rk_a[13][12] = {
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
    {246121993.0 / 1340847787.0, 0.0, 0.0, -37695042795.0 / 15268766246.0, -309121744.0 / 1061227803.0,
     -12992083.0 / 490766935.0, 6005943493.0 / 2108947869.0, 393006217.0 / 1396673457.0,
     123872331.0 / 1001029789.0},
    {-1028468189.0 / 846180014.0, 0.0, 0.0, 8478235783.0 / 508512852.0, 1311729495.0 / 1432422823.0,
     -10304129995.0 / 1701304382.0, -48777925059.0 / 3047939560.0, 15336726248.0 / 1032824649.0,
     -45442868181.0 / 3398467696.0, 3065993473.0 / 597172653.0},
    {185892177.0 / 718116043.0, 0.0, 0.0, -3185094517.0 / 667107341.0, -477755414.0 / 1098053517.0,
     -703635378.0 / 230739211.0, 5731566787.0 / 1027545527.0, 5232866602.0 / 850066563.0,
     -4093664535.0 / 808688257.0, 3962137247.0 / 1805957418.0, 65686358.0 / 487910083.0},
    {403863854.0 / 491063109.0, 0.0, 0.0, -5068492393.0 / 434740067.0, -411421997.0 / 543043805.0,
     652783627.0 / 914296604.0, 11173962825.0 / 925320556.0, -13158990841.0 / 6184727034.0,
     3936647629.0 / 1978049680.0, -160528059.0 / 685178525.0, 248638103.0 / 1413531060.0, 0.0}};
constexpr double c_rk_b8[13] = {14005451.0 / 335480064.0, 0.0, 0.0, 0.0, 0.0, -59238493.0 / 1068277825.0,
                                   181606767.0 / 758867731.0, 561292985.0 / 797845732.0,
                                   -1041891430.0 / 1371343529.0, 760417239.0 / 1151165299.0,
                                   118820643.0 / 751138087.0, -528747749.0 / 2220607170.0, 1.0 / 4.0};
constexpr double c_rk_b7[13] = {13451932.0 / 455176623.0, 0.0, 0.0, 0.0, 0.0, -808719846.0 / 976000145.0,
                                   1757004468.0 / 5645159321.0, 656045339.0 / 265891186.0,
                                   -3867574721.0 / 1518517206.0, 465885868.0 / 322736535.0,
                                   53011238.0 / 667516719.0, 2.0 / 45.0, 0.0};

struct DeltaLane {
    const PotView *T;
    double E, Lz, Rc, R0, D2, sqDR, Ri, zi, sv, cv, v0;
    double Lzp2;       // pow(Lz, 2.0) of the equations of motion, formed once per energy
    double *Rp, *zp;   // [200] orbit samples (scratch)
    int N, status;
};

__device__ inline double dl_dR(const PotView &T, double R, double z) { double a, b; pot_phi_grad(T, R, z, &a, &b); return a; }

__device__ inline double dl_vsq(const DeltaLane &d, double x) {   // Delta.cpp:28-30
    const double t = d.Lz / x;
    return 2 * (d.E - (pot_phi(*d.T, x, 0.) + .5 * fjm_pow(t, 2)));
}

__device__ inline void dl_eqs(const DeltaLane &d, const double *y, double *f) {   // Delta.cpp:33-42
    double dR, dz;
    pot_phi_grad(*d.T, y[0], y[1], &dR, &dz);
    f[0] = y[2];
    f[1] = y[3];
    f[2] = -dR + d.Lzp2 / fjm_pow(y[0], 3.0);
    f[3] = -dz;
}

// one GSL-style evolve_apply step (RK8PD, y-control abs 1e-10 rel 1e-9)
__device__ inline void dl_evolve(const DeltaLane &d, double &t, double t1, double &h, double *y) {
    const double t0 = t, dt = t1 - t0;
    double h0 = h, y0[4], yerr[4], k[13][4], yt[4];
    for (int i = 0; i < 4; i++) y0[i] = y[i];
    dl_eqs(d, y, k[0]);
    for (int tries = 0; tries < 200; tries++) {
        int final_step;
        if ((dt >= 0.0 && h0 > dt) || (dt < 0.0 && h0 < dt)) { h0 = dt; final_step = 1; }
        else final_step = 0;
        // stages unrolled over compile-time coefficients: k stays in registers
        // and the zero coefficients drop out (the sums skip them, as written)
#pragma unroll
        for (int s = 1; s < 13; s++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                double acc = 0;
#pragma unroll
                for (int j = 0; j < s; j++)
                    if (c_rk_a[s][j] != 0.0) acc += c_rk_a[s][j] * k[j][i];
                yt[i] = y0[i] + h0 * acc;
            }
            dl_eqs(d, yt, k[s]);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double s8 = 0, s7 = 0;
#pragma unroll
            for (int j = 0; j < 13; j++) {
                if (c_rk_b8[j] != 0.0) s8 += c_rk_b8[j] * k[j][i];
                if (c_rk_b7[j] != 0.0) s7 += c_rk_b7[j] * k[j][i];
            }
            y[i] = y0[i] + h0 * s8;
            yerr[i] = h0 * (s7 - s8);
        }
        t = final_step ? t1 : t0 + h0;
        const double h_old = h0;
        double rmax = DBL_MINV;
        for (int i = 0; i < 4; i++) {
            const double D0 = 1e-9 * fabs(y[i]) + 1e-10;
            rmax = fj_max(fabs(yerr[i]) / fabs(D0), rmax);
        }
        if (rmax > 1.1) {
            double r = 0.9 / fjm_pow(rmax, 1.0 / 8.0);
            if (r < 0.2) r = 0.2;
            h0 = r * h_old;
            if (fabs(h0) < fabs(h_old) && (t + h0) != t) continue;
            h0 = h_old;
        } else if (rmax < 0.5) {
            double r = 0.9 / fjm_pow(rmax, 1.0 / 9.0);
            if (r > 5.0) r = 5.0;
            if (r < 1.0) r = 1.0;
            h0 = r * h_old;
        }
        if (!final_step) h = h0;
        return;
    }
}

// Delta.cpp:53-90; cond 0: z >= 0, 1: pz >= 0
__device__ inline int dl_orbit(DeltaLane &d, double x0, int cond) {
    double steph = 2e-3, y[4], t = 0;
    y[0] = x0; y[1] = 0.; y[2] = 0.; y[3] = sqrt(dl_vsq(d, x0));
    int i = 0;
    d.Rp[i] = y[0]; d.zp[i] = y[1];
    while (i < 199 && (cond == 0 ? y[1] : y[3]) >= 0.) {
        dl_evolve(d, t, 5 * x0 / sqrt(dl_vsq(d, x0)), steph, y);
        i++;
        d.Rp[i] = y[0]; d.zp[i] = y[1];
    }
    return i;
}

struct DLgetDR {   // Delta.cpp:93-104
    DeltaLane *d;
    __device__ double operator()(double x0) const {
        const int N = dl_orbit(*d, x0, 0);
        const double *R = d->Rp, *z = d->zp;
        if (fabs(z[N - 1] - z[N]) > 0 && fabs(R[N - 1] - R[N]) != 0)
            return x0 - (R[N] - z[N] * (R[N - 1] - R[N]) / (z[N - 1] - z[N]));
        return x0 - R[N];
    }
};

__device__ inline double dl_ensure(DeltaLane &d, double x0) {   // Delta.cpp:116-124
    int count = 0;
    while (dl_vsq(d, x0) < 0) {
        const double dpe = dl_dR(*d.T, x0, 0) - fjm_pow(d.Lz, 2) / fjm_powi(x0, 3);
        if (dpe < 0) x0 *= 1.1; else x0 *= .9;
        if (++count > 50) { d.status |= 1; return x0; }
    }
    return x0;
}

struct DLds2dv {   // Delta.cpp:170-172
    const DeltaLane *d;
    __device__ double operator()(double v) const {
        return 2 * (d->zi * d->sqDR * fjm_sin(v) - fjm_cos(v) * (d->Ri * d->R0 + d->D2 * fjm_sin(v)));
    }
};

struct DLds2dD2 {  // Delta.cpp:210-234 with getv 189-208
    DeltaLane *d;
    __device__ double operator()(double D2h) const {
        DeltaLane &D = *d;
        D.D2 = D2h; D.sqDR = sqrt(D.D2 + D.R0 * D.R0);
        double sum = 0.;
        DLds2dv g{d};
        for (int i = 0; i < D.N; i++) {
            D.Ri = D.Rp[i]; D.zi = D.zp[i];
            const double fl = g(0), fh = g(PIH);
            if (fabs(fl) < EPSABSROOTS) D.v0 = 0;
            else if (fabs(fh) < EPSABSROOTS) D.v0 = PIH;
            else D.v0 = brent_solve(g, 0, PIH, fl, fh);
            D.sv = fjm_sin(D.v0); D.cv = fjm_cos(D.v0);
            const double sv = D.sv, cv = D.cv, zi = D.zi, Ri = D.Ri, sqDR = D.sqDR, R0 = D.R0, D2 = D.D2;
            const double dvdD2 = .5 * sv * (zi / sqDR - 2 * cv) / (cv * zi * sqDR + sv * Ri * R0 - D2 * (cv * cv - sv * sv));
            sum += (cv - zi / sqDR) * cv + g(D.v0) * dvdD2;
        }
        return sum;
    }
};

struct DLRcE {   // Utils.h:243-246
    const PotView *T; double E;
    __device__ __forceinline__ double operator()(double R) const {
        double dR, dz;
        const double phi = pot_phi_grad(*T, R, 0, &dR, &dz);
        (void)phi;
        return .5 * R * dR + pot_phi(*T, R, 0) - E;
    }
};
struct DLRc {    // Utils.h:218-221
    const PotView *T; double Lzsq;
    __device__ __forceinline__ double operator()(double R) const { return Lzsq / fjm_powi(R, 3) - dl_dR(*T, R, 0); }
};

// Utils.h:222-241 GetRc
__device__ inline double dl_getrc(const PotView &T, double Lzsq, double R) {
    DLRc F{&T, Lzsq};
    double Ri = R, Ro = R, fi = 0, fo = 0, f = F(R);
    int guard = 0;
    if (f > 0) { while (f > 0 && guard++ < 4000) { Ri = R; fi = f; R *= 1.2; f = F(R); } Ro = R; fo = f; }
    else { while (f < 0 && guard++ < 4000) { Ro = R; fo = f; R *= .8; f = F(R); } Ri = R; fi = f; }
    return brent_solve(F, Ri, Ro, fi, fo);
}

// tabulateDelta prologue (tabulate.cpp:14-25), entries 0 and 1 of the radius loop.
__device__ inline void delta_prep(const PotView &T, double *prep);

__global__ void k_delta_prep(TabArgs ta, double *prep) {
    extern __shared__ double smem[];
    const LdsTabs L = stage_tables(ta, smem);
    if (threadIdx.x != 0) return;
    delta_prep(L.T, prep);
}

__device__ inline void delta_prep(const PotView &T, double *prep) {
    const int NR = T.nr;
    const double Rmax = T.ar[NR - 1];
    const double Lzmax = Rmax * 0.98 * sqrt(-2 * pot_phi(T, Rmax, 0)), Rmin = 1e-3 * Rmax;
    const double Lzmin = Rmin * sqrt(Rmin * dl_dR(T, Rmin, 0));
    double R = .01 * Rmax, Lz[2], Rg[2], Ec[2], Vm[2];
    for (int i = 0; i < 2; i++) {
        Lz[i] = Lzmin + i * (Lzmax - Lzmin) / (double)(NR - 1);
        R = Rg[i] = dl_getrc(T, Lz[i] * Lz[i], R);
        const double vc2 = R * dl_dR(T, R, 0);
        Ec[i] = .5 * vc2 + pot_phi(T, R, 0);
        Vm[i] = .98 * sqrt(-2 * Ec[i]);
    }
    prep[0] = Lz[1]; prep[1] = Rg[1]; prep[2] = Ec[1]; prep[3] = Vm[1];
}

// tabulate.cpp:27-35 + Delta::getBestDelta (Delta.cpp:248-289), one lane per energy.
__device__ inline void delta_energy(const PotView &T, int NG, const double *prep, double *scratch, int nE,
                                    double *Eout, double *Dout, int *status);

// One energy per wave (block of 64, lane 0 works): the energies' nested
// Brent / RK8PD loops take different trip counts, and lanes of one wave
// would wait for each other at every loop level.
__global__ void __launch_bounds__(64) k_delta(TabArgs ta, const double *prep, double *scratch, double *Eout, double *Dout, int *status) {
    extern __shared__ double smem[];
    const LdsTabs L = stage_tables(ta, smem);
    const int nE = blockIdx.x;
    if (threadIdx.x != 0 || nE >= L.ngrid) return;
    delta_energy(L.T, L.ngrid, prep, scratch + (size_t)nE * 400, nE, Eout, Dout, status);
}

// scratch: 400 doubles of this lane
__device__ inline void delta_energy(const PotView &T, int NG, const double *prep, double *scratch, int nE,
                                    double *Eout, double *Dout, int *status) {
    const double Lz1 = prep[0], Rg1 = prep[1], Ec1 = prep[2], Vm1 = prep[3];
    const double t = Vm1 * (nE + 1) / (double)NG;
    const double E = Ec1 + .5 * (t * t);
    DeltaLane d;
    d.T = &T; d.E = E; d.Lz = .25 * Lz1; d.status = 0;
    d.Lzp2 = fjm_pow(d.Lz, 2.0);
    d.Rp = scratch; d.zp = d.Rp + 200;
    d.D2 = 0; d.sqDR = 0; d.Ri = 0; d.zi = 0; d.sv = 0; d.cv = 0; d.v0 = 0; d.N = 0;
    // Delta ctor: Rc = RcE(E, R0) (Utils.h:247-263)
    {
        DLRcE F{&T, E};
        double R0 = .8 * Rg1, f0 = F(R0);
        int guard = 0;
        while (f0 > 0 && guard++ < 4000) { R0 *= .8; f0 = F(R0); }
        double R1 = 1.2 * Rg1, f1 = F(R1);
        while (f1 < 0 && guard++ < 8000) { R1 *= 1.2; f1 = F(R1); }
        d.Rc = brent_solve(F, R0, R1, f0, f1);
    }
    // getJr0 (Delta.cpp:129-164)
    DLgetDR gdr{&d};
    double x0 = d.Rc, x1, dx0, dx1, A = .8;
    int k = 0;
    double R0res;
    x0 = dl_ensure(d, x0); dx0 = gdr(x0);
    x1 = dl_ensure(d, x0 / A); dx1 = gdr(x1);
    bool early = false;
    while (dx0 * dx1 > 0) {
        if (fabs(dx1) < fabs(dx0)) { x0 = x1; dx0 = dx1; }
        else A = fjm_pow(A, -0.9);
        x1 = dl_ensure(d, x0 * A); dx1 = gdr(x1);
        k++;
        if (k == 21) {
            d.status |= 2;
            R0res = fabs(dx0) < 1.e-7 ? -x0 : -1;
            early = true;
            break;
        }
    }
    if (!early) {
        const double low = fj_min(x0, x1), high = fj_max(x0, x1);
        const double flow = (low == x0) ? dx0 : dx1, fhigh = (high == x1) ? dx1 : dx0;
        R0res = brent_solve(gdr, low, high, x0 == x1 ? dx0 : flow, x0 == x1 ? dx1 : fhigh);
    }
    d.R0 = R0res;
    d.N = dl_orbit(d, d.R0, 1);
    DLds2dD2 F{&d};
    double low = -0.001, high = 5, D2out = 0;
    double Dmin = F(low), Dmax = F(high);
    int guard = 0;
    while (Dmin * Dmax > 0 && fabs(Dmax) > 1e-4 && fabs(Dmin) > 1e-4) {
        if (Dmax < 0) { high *= 5; Dmax = F(high); }
        else { low -= .1; Dmin = F(low); }
        if (++guard > 400) { d.status |= 8; break; }
    }
    if (fabs(Dmin) < EPSABSROOTS) D2out = low;
    else if (fabs(Dmax) < EPSABSROOTS) D2out = high;
    else D2out = brent_solve(F, low, high, Dmin, Dmax);
    Eout[nE] = E;
    Dout[nE] = D2out;
    if (d.status) atomicOr(status, d.status);
}


// ===========================================================================
// Batched models: every SCF phase of a batch of independent single-component
// models (own grid, DF, Delta table) is one launch.  Blocks map to models
// through prefix tables, so each block still stages its own model's tables in
// LDS; the persistent integrator gives every model blocks in proportion to
// its task count.  No host synchronisation inside an iteration (graph capture).
// ===========================================================================
struct ModelDev {
    TabArgs ta;              // phil/Pr/E/D point at this model's installed tables
    double *Pr2;
    fj_df_params df;
    int quad;
    int po, np;              // first point (in the batch), points
    int nr, ntheta, npoly;
    int guess;               // FJ_GUESS_*
    double mass, r0, q;
    const double *ci, *si, *poly;   // theta nodes and projection weights [npoly][ntheta]
    double *rhl, *vrotl, *sigRl, *sigpl, *sigzl, *sigRzl;   // projected [nr][npoly]
    double *prep;            // 4 doubles: tabulateDelta prologue
};

__device__ inline int find_seg(const int *start, int M, int x) {   // largest m with start[m] <= x
    int lo = 0, hi = M - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ inline PotView global_potview(const TabArgs &ta) {
    PotView T;
    T.nr = ta.nr; T.npoly = ta.npoly; T.ar = ta.ar; T.phil = ta.phil; T.Pr = ta.Pr; T.lc = &c_legcoef;
    T.lut = ta.lut; T.lg0 = ta.lg0; T.lsc = ta.lsc; T.rk = ta.rk;
    return T;
}

__global__ void kb_potviews(const ModelDev *md, int M, PotView *pv) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < M) pv[m] = global_potview(md[m].ta);
}

// counters of model m: cnt[8 m + 0] u tasks, 1 v tasks, 2 next task; cnt[8 M + 0]
// overflowed tasks, 1 next overflow task, 2 second overflow count
__global__ void kb_zero(unsigned int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void kb_zero64(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void __launch_bounds__(256) kb_points(const ModelDev *md, const int *pmodel, PointArgs pa, int npts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const PotView T = global_potview(md[pmodel[i]].ta);
    const double far = 10 * T.ar[T.nr - 1];
    const double Phi_h = pot_phi(T, pa.R[i], pa.z[i]);
    pa.Phi[i] = Phi_h;
    pa.Ve[i] = sqrt(-2 * (Phi_h - pot_phi(T, far, far)));
    if (pa.Vc) {
        double dR, dz;
        pot_phi_grad(T, pa.R[i], pa.z[i], &dR, &dz);
        pa.Vc[i] = sqrt(fj_max(1e-30, pa.R[i] * dR + pa.z[i] * dz));
    }
}

// Node setup of every model; model m owns blocks [bstart[m], bstart[m+1]).
// perm: the batch's lane order (each model's nodes permuted within the
// model's own range of node indices).
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(FJ_SETUP_WAVES))) kb_setup(const ModelDev *md, const int *bstart, int M, SetupArgs base,
                                               unsigned int *cnt, const int *perm) {
    extern __shared__ double smem[];
    const int m = find_seg(bstart, M, blockIdx.x);
    const ModelDev &d = md[m];
    const LdsTabs L = stage_tables(d.ta, smem);
    SetupArgs sa = base;
    sa.p0 = 0;
    sa.npts = d.po + d.np;
    sa.quad = d.quad;
    const long long no = (long long)d.po * NQ;
    sa.list_u = base.list_u + no;
    sa.list_v = base.list_v + no;
    sa.cnt_u = cnt + 8 * m;
    sa.cnt_v = cnt + 8 * m + 1;
    const long long local = (long long)(blockIdx.x - bstart[m]) * BS + threadIdx.x;
    setup_node(L, sa, local < (long long)d.np * NQ ? (long long)perm[no + local] : (long long)sa.npts * NQ);
}

// Sort keys of the batch's lane order: node_order_key of the previous
// iteration's records, the model index above it so that every model's nodes
// keep their own range of positions.  Writes the identity into perm.
__global__ void __launch_bounds__(256) kb_order_key(NodeBuf nb, long long nn, const int *pmodel,
                                                    unsigned long long *key, int *perm) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    perm[i] = (int)i;
    const int flag = nb.flag[i];
    double ub = 0, du = 0, vt = 0;
    if (flag == 1) {
        ub = nb.f[NF_UBAR * nb.cap + i]; du = nb.f[NF_DELTAU * nb.cap + i]; vt = nb.f[NF_VTURN * nb.cap + i];
    }
    key[i] = ((unsigned long long)pmodel[i / NQ] << SETUP_KEY_BITS) | node_order_key(flag, ub - du, ub + du, vt);
}

// Blocks of the persistent integrator per model: one for every model with
// tasks, the rest in proportion to the task counts (one thread, fixed order).
__global__ void kb_assign(int M, const unsigned int *cnt, int B, int *block_model) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long tot = 0;
    int with = 0;
    for (int m = 0; m < M; m++) {
        const unsigned int t = cnt[8 * m] + cnt[8 * m + 1];
        tot += t;
        with += t > 0;
    }
    int b = 0;
    const int spare = B > with ? B - with : 0;
    unsigned long long acc = 0;
    int given = 0;
    for (int m = 0; m < M && b < B; m++) {
        const unsigned int t = cnt[8 * m] + cnt[8 * m + 1];
        if (!t) continue;
        acc += t;
        const int upto = tot ? (int)((unsigned long long)spare * acc / tot) : 0;
        int nb = 1 + (upto - given);
        given = upto;
        for (; nb > 0 && b < B; nb--) block_model[b++] = m;
    }
    for (; b < B; b++) block_model[b] = -1;
}

template <class FI>
__device__ __forceinline__ IntegArgs batch_integ_args(const IntegArgs &base, const ModelDev &d, unsigned int *cnt,
                                                      int m) {
    IntegArgs ia = base;
    const long long no = (long long)d.po * NQ;
    ia.list_u = base.list_u + no;
    ia.list_v = base.list_v + no;
    ia.cnt_u = cnt + 8 * m;
    ia.cnt_v = cnt + 8 * m + 1;
    ia.next = cnt + 8 * m + 2;
    return ia;
}

// The next model after m (cyclically) whose task queue is not drained, or -1.
__device__ inline int kb_next_model(unsigned int *cnt, int M, int m) {
    for (int k = 1; k <= M; k++) {
        const int c = (m + k) % M;
        const unsigned int tot = cnt[8 * c] + cnt[8 * c + 1];
        if (atomicAdd(&cnt[8 * c + 2], 0u) < tot) return c;
    }
    return -1;
}

// Integrals of every model: a block starts on the model kb_assign gave it
// and, once that model's queue is drained, moves on to the next model with
// tasks left (re-staging its tables), so that no model's tail idles the
// device.  The hop bound only guards the loop: every visit either takes
// tasks or finds the queue drained for good.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FJ_INTEG_WAVES)))
kb_integrals(const ModelDev *md, const int *block_model, int M, IntegArgs base, unsigned int *cnt) {
    extern __shared__ double smem[];
    __shared__ int s_m;
    int m = block_model[blockIdx.x];
    const long long lane_g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int hop = 0; hop < 4 * M + 8 && m >= 0; hop++) {
        const ModelDev &d = md[m];
        const LdsTabs L = stage_tables(d.ta, smem);
        const IntegArgs ia = batch_integ_args<JInt>(base, d, cnt, m);
        const QagsWork w = lane_work(ia.ws, ia.cap, lane_g, (long long)gridDim.x * blockDim.x);
        const FvStrided fv{smem + tab_lds_doubles(d.ta.nr, d.ta.npoly, d.ta.ngrid) + threadIdx.x};
        JInt f;
        f.T = &L.T;
        integrals_loop(ia, f, w, fv);
        __syncthreads();
        if (threadIdx.x == 0) s_m = kb_next_model(cnt, M, m);
        __syncthreads();
        m = s_m;
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FJ_INTEG32_WAVES)))
kb_integrals_f32(const ModelDev *md, const int *block_model, int M, IntegArgs base, unsigned int *cnt) {
    extern __shared__ double smem[];
    __shared__ int s_m;
    int m = block_model[blockIdx.x];
    const long long lane_g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int hop = 0; hop < 4 * M + 8 && m >= 0; hop++) {
        const ModelDev &d = md[m];
        const LdsTabs L = stage_tables(d.ta, smem);
        float *f32 = reinterpret_cast<float *>(smem + tab_lds_doubles(d.ta.nr, d.ta.npoly, d.ta.ngrid));
        const PotView32 T32 = stage_tables32(L, f32);
        const IntegArgs ia = batch_integ_args<JInt32>(base, d, cnt, m);
        const QagsWorkT<float> w =
            lane_work(reinterpret_cast<float *>(ia.ws), ia.cap, lane_g, (long long)gridDim.x * blockDim.x);
        const FvStrided32 fv{f32 + tab32_lds_floats(d.ta.nr, d.ta.npoly) + threadIdx.x};
        JInt32 f;
        f.T = &T32;
        integrals_loop(ia, f, w, fv);
        __syncthreads();
        if (threadIdx.x == 0) s_m = kb_next_model(cnt, M, m);
        __syncthreads();
        m = s_m;
    }
}

// Overflowed tasks of every model, FP64, full interval limit, tables read
// from global memory per task.
__global__ void __launch_bounds__(256) kb_integrals_ovf(IntegArgs ia) {
    extern __shared__ double smem[];
    const long long lane_g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const QagsWork w = lane_work(ia.ws, ia.cap, lane_g, (long long)gridDim.x * blockDim.x);
    const FvStrided fv{smem + threadIdx.x};
    JInt f;
    f.T = ia.pv;
    integrals_loop<true>(ia, f, w, fv);
}

__global__ void __launch_bounds__(256) kb_moments(const ModelDev *md, const int *pmodel, MomArgs base, int mixed) {
    __shared__ double red[6][256];
    const int p = blockIdx.x;
    const ModelDev &d = md[pmodel[p]];
    MomArgs ma = base;
    ma.p0 = 0;
    ma.df = d.df;
    ma.quad = d.quad;
    if (mixed) moments_point_f32(ma, p, reinterpret_cast<float (*)[2][256]>(&red[0][0]));
    else moments_point(ma, p, red);
}

// rho[0][0] fix + get_Ylm for every model; entries [rstart[m], rstart[m+1]).
__global__ void kb_project(const ModelDev *md, const int *rstart, int M, const double *mom, int npts_total) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rstart[M]) return;
    const int m = find_seg(rstart, M, idx);
    const ModelDev &d = md[m];
    const int loc = idx - rstart[m], np_ = d.npoly, ng = d.ntheta;
    const int n = loc / np_, np = loc % np_;
    const double *rho = mom + d.po, *vrot = mom + npts_total + d.po, *sigR = mom + 2 * npts_total + d.po,
                 *sigp = mom + 3 * npts_total + d.po, *sigz = mom + 4 * npts_total + d.po,
                 *sigRz = mom + 5 * npts_total + d.po;
    const double *ar = d.ta.ar;
    if (n == 0) {
        if (np == 0) {
            const double rho00 = (rho[2 * ng] - rho[1 * ng]) / (ar[2] - ar[1]) * (ar[0] - ar[1]) + rho[1 * ng];
            d.rhl[0] = 2 * rho00; d.vrotl[0] = 2 * vrot[0]; d.sigRl[0] = 2 * sigR[0];
            d.sigpl[0] = 2 * sigp[0]; d.sigzl[0] = 2 * sigz[0]; d.sigRzl[0] = 2 * sigRz[0];
        } else {
            d.rhl[np] = 0; d.vrotl[np] = 0; d.sigRl[np] = 0; d.sigpl[np] = 0; d.sigzl[np] = 0; d.sigRzl[np] = 0;
        }
        return;
    }
    double a = 0, b = 0, c = 0, e = 0, f = 0, g2 = 0;
    for (int g = 0; g < ng; g++) {
        const double wv = d.poly[np * ng + g];
        const int i = n * ng + g;
        a += wv * rho[i]; b += wv * vrot[i]; c += wv * sigR[i];
        e += wv * sigp[i]; f += wv * sigz[i]; g2 += wv * sigRz[i];
    }
    d.rhl[loc] = a; d.vrotl[loc] = b; d.sigRl[loc] = c; d.sigpl[loc] = e; d.sigzl[loc] = f; d.sigRzl[loc] = g2;
}

// Guess density coefficients of every model (Potential::computeGuessRhl).
__global__ void kb_guess(const ModelDev *md, const int *rstart, int M) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rstart[M]) return;
    const int m = find_seg(rstart, M, idx);
    const ModelDev &d = md[m];
    const int loc = idx - rstart[m], np_ = d.npoly, ng = d.ntheta;
    const int n = loc / np_, np = loc % np_;
    const double q2 = d.q * d.q;
    const double *ar = d.ta.ar;
    double s = 0.;
    for (int g = 0; g < ng; g++)
        s += d.poly[np * ng + g] * d_rho_guess(d.guess, d.mass, d.r0, d.q, q2, ar[n] * d.si[g], ar[n] * d.ci[g]);
    if (loc == 0) s = 2 * d_rho_guess(d.guess, d.mass, d.r0, d.q, q2, ar[0] * d.si[0], ar[0] * d.ci[0]);
    d.rhl[loc] = s;
}

// Block-scan Poisson of every model into its installed tables; block (np, m).
__global__ void __launch_bounds__(PSCAN_THREADS) kb_poisson(const ModelDev *md) {
    __shared__ double Iint[PSCAN_MAXNR], Iext[PSCAN_MAXNR], part[PSCAN_THREADS];
    const ModelDev &d = md[blockIdx.y];
    if ((int)blockIdx.x >= d.npoly) return;
    poisson_scan_col(d.nr, d.npoly, d.ta.ar, d.rhl, const_cast<double *>(d.ta.phil), const_cast<double *>(d.ta.Pr),
                     d.Pr2, blockIdx.x, Iint, Iext, part);
}

__global__ void kb_delta_prep(const ModelDev *md, const PotView *pv, int M) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < M) delta_prep(pv[m], md[m].prep);
}

// One (model, energy) per wave, as k_delta.
__global__ void __launch_bounds__(64) kb_delta(const ModelDev *md, const PotView *pv, int M, double *scratch, int *status) {
    const int i = blockIdx.x;
    const int NG = md[0].ta.ngrid;
    if (threadIdx.x != 0 || i >= M * NG) return;
    const int m = i / NG, nE = i % NG;
    delta_energy(pv[m], NG, md[m].prep, scratch + (size_t)i * 400, nE, const_cast<double *>(md[m].ta.E),
                 const_cast<double *>(md[m].ta.D), status);
}

// ---------------------------------------------------------------------------
// Scratch of fj_rhofDF_noalloc: one statically sized device array holds every
// buffer the node phase needs beyond the caller's inputs and outputs (points,
// node records, task lists, counters, QAGS workspaces), so that the call
// allocates nothing.  Each GPU the library runs on holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_LANES = 8192;        // persistent integrator lanes (32 blocks of 256)
constexpr int NOALLOC_OVF_LANES = 256;     // lanes of the full-limit overflow pass
constexpr long long NOALLOC_NODES = (long long)FJ_NOALLOC_MAX_POINTS * NQ;

struct NoallocLayout {
    size_t R, z, Phi, Ve, out6, mom, nodef, Jr, Jz, flag, list_u, list_v, ovf, cnt, ctr, ws, wsfull, total;
};
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr NoallocLayout noalloc_layout() {
    NoallocLayout L{};
    size_t o = 0;
    const size_t P = FJ_NOALLOC_MAX_POINTS, N = (size_t)NOALLOC_NODES, D = sizeof(double), I = sizeof(int);
    L.R = o; o = na_align(o + P * D);
    L.z = o; o = na_align(o + P * D);
    L.Phi = o; o = na_align(o + P * D);
    L.Ve = o; o = na_align(o + P * D);
    L.out6 = o; o = na_align(o + 6 * P * D);
    L.mom = o; o = na_align(o + 6 * P * D);
    L.nodef = o; o = na_align(o + (size_t)NF_COUNT * N * D);
    L.Jr = o; o = na_align(o + N * D);
    L.Jz = o; o = na_align(o + N * D);
    L.flag = o; o = na_align(o + N * I);
    L.list_u = o; o = na_align(o + N * I);
    L.list_v = o; o = na_align(o + N * I);
    L.ovf = o; o = na_align(o + 2 * N * I);
    L.cnt = o; o = na_align(o + 16 * I);
    L.ctr = o; o = na_align(o + sizeof(Counters));
    L.ws = o; o = na_align(o + ws_doubles(QCAP, NOALLOC_LANES) * D);
    L.wsfull = o; o = na_align(o + ws_doubles(QCAP_FULL, NOALLOC_OVF_LANES) * D);
    L.total = o;
    return L;
}
constexpr NoallocLayout NA = noalloc_layout();
static_assert(NA.total <= (size_t)64 << 20, "no-allocation scratch above its 64 MiB budget");

static __device__ __align__(256) unsigned char scratch[NA.total];

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
struct fj_ctx {
    int device = 0;
    hipStream_t stream = nullptr;       // the stream every call is ordered on
    hipStream_t own_stream = nullptr;   // the context's own (stream unless fj_set_stream gave another)
    int nr = 0, ntheta = 0, npoly = 0, ngrid = 0, quad = 0, prec = 0;
    bool have_pot = false, have_delta = false, have_df = false;
    std::vector<double> ar, ci, wi, si, poly;
    fj_df_params df{};
    double *d_ar = nullptr, *d_phil = nullptr, *d_Pr = nullptr, *d_Pr2 = nullptr, *d_E = nullptr, *d_D = nullptr;
    double *d_ci = nullptr, *d_si = nullptr, *d_poly = nullptr;
    int *d_lut = nullptr;
    float lut_lg0 = 0, lut_lsc = 0;
    double *d_rk = nullptr;   // [nr][RK] radial cell divisors
    // node phase scratch
    long long node_cap = 0;
    double *d_nodef = nullptr, *d_Jr = nullptr, *d_Jz = nullptr;
    int *d_flag = nullptr, *d_list_u = nullptr, *d_list_v = nullptr, *d_ovf = nullptr;
    unsigned int *d_cnt = nullptr;   // [0] u, [1] v, [2] next, [3] ovf, [4] next2, [6..8] surrogate, [10] k_setup nodes
    Counters *d_ctr = nullptr;
    double *d_ws = nullptr; long long ws_lanes = 0; int ws_cap = 0;
    double *d_wsfull = nullptr; long long wsfull_lanes = 0;
    // lane order of k_setup (k_setup_key): keys and node permutation, in and
    // out of the radix sort [2][order_cap] each, the sort's scratch
    unsigned int *d_okey = nullptr; int *d_operm = nullptr; long long order_cap = 0;
    unsigned char *d_otemp = nullptr; size_t otemp_bytes = 0;
    long long rec_nodes = -1;   // nodes of the point-major records the last node phase left (-1: none)
    int *d_tcost = nullptr;     // [2 node + type] intervals of each FP32 task's QAGS in the last node phase
    // fast node phase, one lane per node: flag, actions and next-pass sort key per node
    int *d_fflag = nullptr; double *d_fJr = nullptr, *d_fJz = nullptr; unsigned int *d_fkey = nullptr;
    long long fnode_cap = 0, fast_keys = -1;   // fast_keys: nodes whose keys the last fast pass wrote
    // point buffers
    int pt_cap = 0;
    double *d_eval = nullptr; int eval_cap = 0;   // fj_eval_potential: R, z, Phi, dR, dz [eval_cap] each
    unsigned char *noalloc = nullptr;   // this device's copy of the fj_rhofDF_noalloc scratch
    // fast node phase: slice sums [npts][S][6], Gauss-Legendre order of the action integrals
    double *d_fpart = nullptr; size_t fpart_cap = 0;
    int fast_ngl = 64;
    bool fgl_ready = false;
    int fast_cus = 0;                 // compute units of the device (fast phase's launch shape)
    double *d_pR = nullptr, *d_pz = nullptr, *d_pPhi = nullptr, *d_pVe = nullptr, *d_out6 = nullptr, *d_mom = nullptr;
    // grid buffers
    double *d_rhl = nullptr, *d_gl = nullptr, *d_pw = nullptr, *d_pws = nullptr; int *d_status = nullptr;
    int integ_lanes = 0, integ_lanes32 = 0;
    long long last_nodes = 0;
    fj_stats st{};
    // action surrogate
    double *d_pVc = nullptr;
    bool sur_init = false, sur_norm = false;
    int sur_on = 0;
    double sur_judge = 0.01, sur_jzmin = 1e-3, sur_keep = 0;
    unsigned int sur_seed = 1, sur_iter = 0;
    Mlp mlp;
    SurNorm nm{};
    float *d_X = nullptr, *d_T = nullptr, *d_bt = nullptr, *d_bw = nullptr;   // set [cap][5], [cap][3]; batch
    long long train_cap = 0;
    unsigned long long *d_pos = nullptr;
    double *d_loss = nullptr;
    long long sur_nodes_cap = 0;
    bf16_t *d_sx16 = nullptr;
    float *d_sy = nullptr, *d_judge_J = nullptr;
    int *d_sel = nullptr, *d_judge_node = nullptr;
    double *d_judge_S = nullptr;
    bf16_t *d_wpk = nullptr;   // packed weights of the fused inference (MlpPacked)
    float *d_wbias = nullptr;
    int sur_blocks = 0;
    fj_surrogate_stats sst{};
    char err[512] = {0};
};

namespace {

int set_err(fj_ctx *c, int code, const char *fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(ctx, x)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            return set_err(ctx, FJ_E_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                                \
    } while (0)

template <class T>
int dalloc(fj_ctx *c, T **p, size_t n) {
    if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (n == 0) return 0;
    const hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(c, FJ_E_HIP, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    }
    return 0;
}

// include/UtilsLeg.h:18-39 gauleg (constant tables, built once per grid)
void host_gauleg(int n, double x1, double x2, double *x, double *w) {
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1), pp = 0, z1;
    for (int i = 0; i < m; i++) {
        double z = cos(3.141592654 * (i + 0.75) / (n + 0.5));
        do {
            double p1 = 1, p2 = 0, p3;
            for (int j = 0; j < n; j++) {
                p3 = p2; p2 = p1;
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

// evenlegend (UtilsLeg.h:68-79), host side for the projection weights
void host_evenlegend(int npoly, double c, double *pol) {
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


TabArgs tab_args(const fj_ctx *c) {
    TabArgs t;
    t.nr = c->nr; t.npoly = c->npoly; t.ngrid = c->ngrid;
    t.ar = c->d_ar; t.phil = c->d_phil; t.Pr = c->d_Pr; t.E = c->d_E; t.D = c->d_D;
    t.lut = c->d_lut; t.lg0 = c->lut_lg0; t.lsc = c->lut_lsc; t.rk = c->d_rk;
    return t;
}
size_t tab_bytes(const fj_ctx *c) { return tab_lds_doubles(c->nr, c->npoly, c->ngrid) * sizeof(double); }
// k_integrals: tables + the 20 off-centre QK21 values of each of its 256 lanes;
// k_integrals_f32 adds the FP32 tables and keeps the 20 values as floats.
size_t integ_bytes(const fj_ctx *c, bool mixed) {
    if (mixed) return tab_bytes(c) + (tab32_lds_floats(c->nr, c->npoly) + 20 * 256) * sizeof(float);
    return tab_bytes(c) + 20 * 256 * sizeof(double);
}

int ensure_points(fj_ctx *c, int n) {
    if (n <= c->pt_cap) return 0;
    int r = 0;
    if ((r = dalloc(c, &c->d_pR, n)) || (r = dalloc(c, &c->d_pz, n)) || (r = dalloc(c, &c->d_pPhi, n)) ||
        (r = dalloc(c, &c->d_pVe, n)) || (r = dalloc(c, &c->d_out6, (size_t)n * 6)) ||
        (r = dalloc(c, &c->d_mom, (size_t)n * 6)) || (r = dalloc(c, &c->d_pVc, n)))
        return r;
    c->pt_cap = n;
    return 0;
}

// Persistent-kernel lane count: fill the device at the kernel's occupancy.
int integ_lanes(fj_ctx *c, bool mixed) {
    int &cached = mixed ? c->integ_lanes32 : c->integ_lanes;
    if (cached) return cached;
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
    int per_cu = 0;
    if (mixed)
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_integrals_f32, 256, integ_bytes(c, true)));
    else
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_integrals, 256, integ_bytes(c, false)));
    if (per_cu < 1) per_cu = 1;
    cached = prop.multiProcessorCount * per_cu * 256;
    return cached;
}


int ensure_nodes(fj_ctx *c, long long nodes) {
    int r = 0;
    if (nodes > c->node_cap) {
        if ((r = dalloc(c, &c->d_nodef, (size_t)nodes * NF_COUNT)) || (r = dalloc(c, &c->d_Jr, (size_t)nodes)) ||
            (r = dalloc(c, &c->d_Jz, (size_t)nodes)) || (r = dalloc(c, &c->d_flag, (size_t)nodes)) ||
            (r = dalloc(c, &c->d_list_u, (size_t)nodes)) || (r = dalloc(c, &c->d_list_v, (size_t)nodes)) ||
            (r = dalloc(c, &c->d_ovf, (size_t)2 * nodes)) || (r = dalloc(c, &c->d_tcost, (size_t)2 * nodes)))
            return r;
        c->node_cap = nodes;
        c->rec_nodes = -1;
    }
    if (!c->d_cnt) {
        if ((r = dalloc(c, &c->d_cnt, 16)) || (r = dalloc(c, &c->d_ctr, 1))) return r;
    }
    const int l64 = integ_lanes(c, false), l32 = integ_lanes(c, true);
    if (l64 == FJ_E_HIP || l32 == FJ_E_HIP) return FJ_E_HIP;
    const int lanes = l64 > l32 ? l64 : l32;
    if (lanes > c->ws_lanes || c->ws_cap != QCAP) {
        if ((r = dalloc(c, &c->d_ws, ws_doubles(QCAP, lanes)))) return r;
        c->ws_lanes = lanes; c->ws_cap = QCAP;
    }
    return 0;
}

long long chunk_points(const fj_ctx *c, int npts) {
    // one launch per phase up to 32M nodes (node records ~6 GB of HBM)
    long long per = (32ll << 20) / NQ;
    if (per < 1) per = 1;
    return npts < per ? npts : per;
    (void)c;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

void launch_integrals(fj_ctx *c, const TabArgs &ta, const IntegArgs &ia, int lanes, bool mixed) {
    if (mixed) k_integrals_f32<<<lanes / 256, 256, integ_bytes(c, true), c->stream>>>(ta, ia);
    else k_integrals<<<lanes / 256, 256, integ_bytes(c, false), c->stream>>>(ta, ia);
}

// Staeckel actions of the tasks k_setup listed (counters d_cnt[0..1]): the
// persistent integrator, then the FP64 full-limit pass for overflowed tasks.
int ensure_order(fj_ctx *c, long long nn);

// order (mixed precision): the task lists sorted by the intervals each task
// took in the previous node phase, whose records are for these nodes.  (In
// FP64 the lists keep k_setup's order: sorting them was slower there.)
int run_integrals(fj_ctx *c, const TabArgs &ta, const NodeBuf &nb, bool mixed, float *t_int, float *t_ovf,
                  unsigned int *novf_out, bool order = false) {
    hipEvent_t e0, e1, e2;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventCreate(&e2));
    const int lanes = integ_lanes(c, mixed);
    IntegArgs ia{nb, c->d_list_u, c->d_list_v, c->d_cnt + 0, c->d_cnt + 1, c->d_cnt + 2,
                 c->d_ovf, c->d_cnt + 3, c->d_ws, mixed ? QCAP_MIXED : QCAP, -1, c->d_ctr};
    if (mixed) ia.cost = c->d_tcost;
    HIPCHK(c, hipEventRecord(e0, c->stream));
    if (mixed && order) {
        int r;
        const long long nn = nb.cap;
        if ((r = ensure_order(c, nn))) return r;
        int *sorted[2] = {c->d_operm, c->d_operm + nn};
        const int *lists[2] = {c->d_list_u, c->d_list_v};
        for (int type = 0; type < 2; type++) {
            k_task_key<<<(unsigned)((nn + 255) / 256), 256, 0, c->stream>>>(lists[type], c->d_cnt + type, nn, type,
                                                                             c->d_tcost, c->d_okey);
            HIPCHK(c, hipGetLastError());
            size_t need = 0;
            HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, need, c->d_okey, c->d_okey + nn, lists[type],
                                                         sorted[type], (int)nn, 0, TASK_KEY_BITS, c->stream));
            if (need > c->otemp_bytes) {
                if ((r = dalloc(c, &c->d_otemp, need))) return r;
                c->otemp_bytes = need;
            }
            HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(c->d_otemp, need, c->d_okey, c->d_okey + nn, lists[type],
                                                         sorted[type], (int)nn, 0, TASK_KEY_BITS, c->stream));
        }
        ia.list_u = sorted[0];
        ia.list_v = sorted[1];
    }
    launch_integrals(c, ta, ia, lanes, mixed);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(e1, c->stream));
    unsigned int novf = 0;
    HIPCHK(c, hipMemcpyAsync(&novf, c->d_cnt + 3, sizeof novf, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (novf) {
        // FP64 integrand, full 1000-interval limit; as many lanes as tasks
        // (in blocks of 256), at most one full device
        const int l64 = integ_lanes(c, false);
        int fl = (int)((novf + 255) / 256 * 256);
        if (fl > l64) fl = l64;
        if (fl > c->wsfull_lanes) {
            int r;
            if ((r = dalloc(c, &c->d_wsfull, ws_doubles(QCAP_FULL, fl)))) return r;
            c->wsfull_lanes = fl;
        }
        IntegArgs io = ia;
        io.next = c->d_cnt + 4; io.ws = c->d_wsfull; io.cap = QCAP_FULL; io.ntask_override = (int)novf;
        io.ovf_cnt = c->d_cnt + 5;
        launch_integrals(c, ta, io, fl, false);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(e2, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *t_int += elapsed(e0, e1);
    *t_ovf += elapsed(e1, e2);
    *novf_out = novf;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2);
    return 0;
}

int sur_ensure_nodes(fj_ctx *c, long long nn) {
    if (nn <= c->sur_nodes_cap) return 0;
    int r;
    if ((r = dalloc(c, &c->d_sx16, (size_t)nn * MLP_IN)) || (r = dalloc(c, &c->d_sy, (size_t)nn * MLP_OUT)) ||
        (r = dalloc(c, &c->d_sel, (size_t)nn)) || (r = dalloc(c, &c->d_judge_node, (size_t)nn)) ||
        (r = dalloc(c, &c->d_judge_J, (size_t)nn * 2)) || (r = dalloc(c, &c->d_judge_S, (size_t)nn * 2)))
        return r;
    c->sur_nodes_cap = nn;
    return 0;
}

// Judge statistics: median and 95th percentile of the judged nodes' errors.
int sur_judge_stats(fj_ctx *c, const SurArgs &a, unsigned int nj) {
    if (!nj) return 0;
    float *d_err = reinterpret_cast<float *>(c->d_judge_S);
    k_sur_judge_err<<<(nj + 255) / 256, 256, 0, c->stream>>>(a, nj, d_err);
    HIPCHK(c, hipGetLastError());
    std::vector<float> E(3 * (size_t)nj);
    HIPCHK(c, hipMemcpyAsync(E.data(), d_err, E.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 2; k++) {
        std::vector<std::pair<float, float>> e;   // (error, density weight)
        e.reserve(nj);
        for (unsigned int i = 0; i < nj; i++)
            if (E[2 * i + k] >= 0) e.push_back({E[2 * i + k], E[2 * (size_t)nj + i]});
        if (e.empty()) continue;
        std::sort(e.begin(), e.end());
        const double med = e[e.size() / 2].first, p95 = e[(size_t)(0.95 * (e.size() - 1))].first;
        // weighted median: the error below which half of the judged density weight lies
        double wtot = 0, acc = 0, wmed = e.back().first;
        for (const auto &x : e) wtot += x.second;
        for (const auto &x : e) {
            acc += x.second;
            if (acc >= 0.5 * wtot) { wmed = x.first; break; }
        }
        if (k == 0) { c->sst.judge_err_r_med = med; c->sst.judge_err_r_p95 = p95; c->sst.judge_werr_r_med = wmed; }
        else { c->sst.judge_err_z_med = med; c->sst.judge_err_z_p95 = p95; c->sst.judge_werr_z_med = wmed; }
    }
    return 0;
}

// Packed weights for the fused inference from the current fp32 parameters.
int sur_pack(fj_ctx *c) {
    k_mlp_pack<<<4 * 16 * 64 * 4 / 256, 256, 0, c->stream>>>(c->mlp.p, c->d_wpk, c->d_wbias);
    HIPCHK(c, hipGetLastError());
    if (!c->sur_blocks) {
        hipDeviceProp_t prop;
        HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
        int per_cu = 0;
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sur_infer, 256, SUR_LDS));
        if (per_cu < 1) return set_err(c, FJ_E_HIP, "fused surrogate kernel does not fit (needs %d B of LDS)", SUR_LDS);
        c->sur_blocks = prop.multiProcessorCount * per_cu;
    }
    return 0;
}

int sur_collect(fj_ctx *c, const NodeBuf &nb, int p0, int np, const int *only, unsigned int n_only, double keep) {
    if (!c->sur_init || keep <= 0) return 0;
    CollectArgs ca{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, c->d_pVc, p0, np, c->quad, nb, only, n_only, keep,
                   hash32(c->sur_seed * 977u + c->sur_iter * 131u + (unsigned)p0), c->nm.tfloor, c->df, c->d_out6,
                   c->d_X, c->d_T, c->d_pos, c->train_cap};
    const long long n = only ? (long long)n_only : (long long)np * NQ;
    if (n > 0) k_sur_collect<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(ca);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// Gauss-Legendre rule on [0, 1] with n nodes (Newton on P_n to full
// precision), as the FP32 constants of the fast action integrals.
void fast_gl_table(int n, FastGL32 *t) {
    for (int i = 0; i < n; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 1;
        for (int it = 0; it < 100; it++) {
            double p1 = 1, p2 = 0;
            for (int j = 0; j < n; j++) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2 * j + 1) * z * p2 - j * p3) / (j + 1);
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            const double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        const double x = 0.5 * (1 - z), w = 1 / ((1 - z * z) * pp * pp);   // node on [0, 1], weight / 2 * 2
        t[i].w = (float)w;
        t[i].cu = (float)cos(M_PI * x); t[i].su = (float)sin(M_PI * x);
        t[i].cv = (float)cos(0.5 * M_PI * x); t[i].sv = (float)sin(0.5 * M_PI * x);
    }
}

// Fast node phase: k_points, k_fast, k_fast_finish; results in d_out6 and d_mom.
// The Gauss-Legendre table of the fast phase's action rules, on the device.
int fast_gl_ensure(fj_ctx *c) {
    if (!c->fgl_ready) {
        FastGL32 t[FAST_GL_MAX];
        fast_gl_table(c->fast_ngl, t);
        HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_fgl), t, sizeof(FastGL32) * c->fast_ngl));
        const int nv = fast_v_order(c->fast_ngl);
        fast_gl_table(nv, t);
        HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_fglv), t, sizeof(FastGL32) * nv));
        c->fgl_ready = true;
    }
    return 0;
}

// Fast node phases up to this many nodes keep node-indexed records (24 B per
// node with the sort key, plus the sort's 16 B); larger ones sum per lane
// (k_fast).
constexpr long long FAST_NODE_RECORDS_MAX = 64ll << 20;

#ifndef FJ_FAST_SORT
#define FJ_FAST_SORT 1   // k_fast_nodes in the order of the previous fast pass's turning points
#endif
int run_fast_nodes(fj_ctx *c, int npts);
int ensure_order(fj_ctx *c, long long nn);
int sort_order(fj_ctx *c, long long nn, const unsigned int *kin);

int run_fast(fj_ctx *c, int npts) {
    int r;
    if ((r = fast_gl_ensure(c))) return r;
    if (!c->d_cnt) {
        if ((r = dalloc(c, &c->d_cnt, 16)) || (r = dalloc(c, &c->d_ctr, 1))) return r;
    }
    if ((long long)npts * NQ <= FAST_NODE_RECORDS_MAX) return run_fast_nodes(c, npts);
    const int tiles = (npts + 63) / 64;
    if (!c->fast_cus) {
        hipDeviceProp_t prop;
        HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
        c->fast_cus = prop.multiProcessorCount;
    }
    // many short blocks (FJ_FAST_BLOCKS_PER_CU per CU) so that the last round of
    // blocks is a small part of the launch; at least 4 nodes per wave
    int S = (FJ_FAST_BLOCKS_PER_CU * c->fast_cus + tiles - 1) / tiles;
    if (S > NQ / 16) S = NQ / 16;
    if (S < 1) S = 1;
    const size_t need = (size_t)npts * S * 6;
    if (need > c->fpart_cap) {
        if ((r = dalloc(c, &c->d_fpart, need))) return r;
        c->fpart_cap = need;
    }
    const TabArgs ta = tab_args(c);
    const size_t lds_pts = tab_bytes(c);
    const size_t lds = fast_lds_bytes(c->nr, c->npoly, c->ngrid);
    hipEvent_t ev[4];
    for (int i = 0; i < 4; i++) HIPCHK(c, hipEventCreate(&ev[i]));
    HIPCHK(c, hipMemsetAsync(c->d_ctr, 0, sizeof(Counters), c->stream));
    memset(&c->st, 0, sizeof c->st);
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    PointArgs pa{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, nullptr};
    k_points<<<(npts + 255) / 256, 256, lds_pts, c->stream>>>(ta, pa, npts);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    FastArgs fa{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, npts, c->quad, S, c->fast_ngl, fast_v_order(c->fast_ngl),
                c->df, c->d_fpart, c->d_ctr};
    k_fast<<<(unsigned)(tiles * S), 256, lds, c->stream>>>(ta, fa);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[2], c->stream));
    k_fast_finish<<<npts, 256, 0, c->stream>>>(npts, S, c->quad, c->d_fpart, c->d_out6, c->d_mom);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[3], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Counters h;
    HIPCHK(c, hipMemcpy(&h, c->d_ctr, sizeof h, hipMemcpyDeviceToHost));
    c->st.ms_points = elapsed(ev[0], ev[1]);
    c->st.ms_setup = elapsed(ev[1], ev[2]);     // the fused node kernel
    c->st.ms_moments = elapsed(ev[2], ev[3]);
    c->st.ms_total = elapsed(ev[0], ev[3]);
    c->st.points = npts;
    c->st.nodes = (long long)npts * NQ;
    c->st.bound_nodes = (long long)h.bound;
    c->st.unbound_nodes = (long long)h.unbound;
    c->st.nonfinite_nodes = (long long)h.nonfinite;
    c->st.phi_evals = (long long)h.phi + 2 * (long long)npts;
    c->st.setup_lanes = tiles * S * 256;
    for (int i = 0; i < 4; i++) (void)hipEventDestroy(ev[i]);
    return 0;
}

// Fast node phase with node-indexed records (k_fast_nodes, then k_moments):
// lanes in the order of the previous fast pass's turning points when it left
// keys for these nodes, else in k_setup's tiled order.
int run_fast_nodes(fj_ctx *c, int npts) {
    int r;
    const long long nn = (long long)npts * NQ;
    if (nn > c->fnode_cap) {
        if ((r = dalloc(c, &c->d_fflag, (size_t)nn)) || (r = dalloc(c, &c->d_fJr, (size_t)nn)) ||
            (r = dalloc(c, &c->d_fJz, (size_t)nn)) || (r = dalloc(c, &c->d_fkey, (size_t)nn)))
            return r;
        c->fnode_cap = nn;
        c->fast_keys = -1;
    }
    if ((r = ensure_order(c, nn))) return r;
    const TabArgs ta = tab_args(c);
    const size_t lds_pts = tab_bytes(c);
    const size_t lds = fast_lds_bytes(c->nr, c->npoly, c->ngrid);
    struct Events {   // destroyed on every return
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events() { for (hipEvent_t x : e) if (x) (void)hipEventDestroy(x); }
    } evs;
    hipEvent_t *ev = evs.e;
    for (int i = 0; i < 4; i++) HIPCHK(c, hipEventCreate(&ev[i]));
    HIPCHK(c, hipMemsetAsync(c->d_ctr, 0, sizeof(Counters), c->stream));
    memset(&c->st, 0, sizeof c->st);
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    PointArgs pa{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, nullptr};
    k_points<<<(npts + 255) / 256, 256, lds_pts, c->stream>>>(ta, pa, npts);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    const int *sel = nullptr;
    if (FJ_FAST_SORT && c->fast_keys == nn) {
        NodeBuf none{nullptr, nn, nullptr, nullptr, nullptr};
        k_setup_key<<<(unsigned)((nn + 255) / 256), 256, 0, c->stream>>>(none, nn, nullptr, c->d_operm, nullptr);
        HIPCHK(c, hipGetLastError());
        if ((r = sort_order(c, nn, c->d_fkey))) return r;
        sel = c->d_operm + nn;
    }
    NodeBuf nb{nullptr, nn, c->d_fJr, c->d_fJz, c->d_fflag};
    FastNodeArgs fa{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, npts, c->quad, c->fast_ngl, fast_v_order(c->fast_ngl),
                    nb, sel, c->d_fkey, c->d_ctr};
    k_fast_nodes<<<(unsigned)((nn + 255) / 256), 256, lds, c->stream>>>(ta, fa);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[2], c->stream));
    MomArgs ma{c->d_pPhi, c->d_pVe, c->d_pR, 0, npts, c->quad, nb, c->df, c->d_out6,
               c->d_mom, c->d_mom + npts, c->d_mom + 2 * npts, c->d_mom + 3 * npts, c->d_mom + 4 * npts,
               c->d_mom + 5 * npts, &c->d_ctr->nonfinite};
    k_moments<<<npts, 256, 0, c->stream>>>(ma);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[3], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->fast_keys = nn;
    Counters h;
    HIPCHK(c, hipMemcpy(&h, c->d_ctr, sizeof h, hipMemcpyDeviceToHost));
    c->st.ms_points = elapsed(ev[0], ev[1]);
    c->st.ms_setup = elapsed(ev[1], ev[2]);     // the node kernel (and the lane order's sort)
    c->st.ms_moments = elapsed(ev[2], ev[3]);
    c->st.ms_total = elapsed(ev[0], ev[3]);
    c->st.points = npts;
    c->st.nodes = nn;
    c->st.bound_nodes = (long long)h.bound;
    c->st.unbound_nodes = (long long)h.unbound;
    c->st.nonfinite_nodes = (long long)h.nonfinite;
    c->st.phi_evals = (long long)h.phi + 2 * (long long)npts;
    c->st.setup_lanes = (nn + 255) / 256 * 256;
    return 0;
}

// k_setup's lane order: nodes sorted by k_setup_key over the records of the
// previous node phase (radix sort on the stream); sets sa->sel / sa->nsel.
// Sorts the keys kin[0..nn) with the node permutation k_setup_key wrote into
// d_operm[0..nn); the sorted node order lands in d_operm[nn..2nn).
int sort_order(fj_ctx *c, long long nn, const unsigned int *kin) {
    int r = 0;
    size_t need = 0;
    unsigned int *kout = c->d_okey + nn;
    int *pin = c->d_operm, *pout = c->d_operm + nn;
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, pin, pout, (int)nn, 0, SETUP_KEY_BITS,
                                                 c->stream));
    if (need > c->otemp_bytes) {
        if ((r = dalloc(c, &c->d_otemp, need))) return r;
        c->otemp_bytes = need;
    }
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(c->d_otemp, need, kin, kout, pin, pout, (int)nn, 0, SETUP_KEY_BITS,
                                                 c->stream));
    return 0;
}

int ensure_order(fj_ctx *c, long long nn) {
    int r = 0;
    if (nn > c->order_cap) {
        if ((r = dalloc(c, &c->d_okey, (size_t)2 * nn)) || (r = dalloc(c, &c->d_operm, (size_t)2 * nn))) return r;
        c->order_cap = nn;
    }
    return 0;
}

int setup_order(fj_ctx *c, const NodeBuf &nb, long long nn, SetupArgs *sa) {
    int r = 0;
    if ((r = ensure_order(c, nn))) return r;
    k_setup_key<<<(unsigned)((nn + 255) / 256), 256, 0, c->stream>>>(nb, nn, c->d_okey, c->d_operm, c->d_cnt + 10);
    HIPCHK(c, hipGetLastError());
    if ((r = sort_order(c, nn, c->d_okey))) return r;
    sa->sel = c->d_operm + nn;
    sa->nsel = c->d_cnt + 10;
    return 0;
}

#ifndef FJ_SETUP_SORT
#define FJ_SETUP_SORT 1   // k_setup in the order of the previous node phase's turning points
#endif

// Node phase for npts points already in d_pR/d_pz; results in d_out6 and d_mom.
int run_nodes(fj_ctx *c, int npts) {
    if (!c->have_pot || !c->have_delta || !c->have_df)
        return set_err(c, FJ_E_STATE, "potential, Delta table and DF must be set before the node phase");
    const bool mixed = c->prec == FJ_PREC_MIXED;
    const bool sur = c->sur_on && c->sur_init && c->sur_norm;
    if (c->prec == FJ_PREC_FAST) {
        if (sur) return set_err(c, FJ_E_STATE, "the action surrogate runs with FJ_PREC_FP64 or FJ_PREC_MIXED");
        return run_fast(c, npts);
    }
    const long long chunk = chunk_points(c, npts);
    int r = ensure_nodes(c, chunk * NQ);
    if (r) return r;
    if (sur && (r = sur_ensure_nodes(c, chunk * NQ))) return r;
    const TabArgs ta = tab_args(c);
    const size_t lds = tab_bytes(c);
    hipEvent_t ev[8];
    for (int i = 0; i < 8; i++) HIPCHK(c, hipEventCreate(&ev[i]));
    HIPCHK(c, hipMemsetAsync(c->d_ctr, 0, sizeof(Counters), c->stream));
    memset(&c->st, 0, sizeof c->st);
    if (sur) {
        const long long nt = c->sst.n_train;
        memset(&c->sst, 0, sizeof c->sst);
        c->sst.n_train = nt;
    }
    c->sur_iter++;
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    // the circular-speed scale Vc feeds only the surrogate's inputs
    const bool need_vc = c->sur_init;
    PointArgs pa{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, need_vc ? c->d_pVc : nullptr};
    k_points<<<(npts + 255) / 256, 256, lds, c->stream>>>(ta, pa, npts);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    float t_setup = 0, t_int = 0, t_ovf = 0, t_mom = 0, t_sur = 0;
    c->st.integral_lanes = integ_lanes(c, mixed);
    for (long long p0 = 0; p0 < npts; p0 += chunk) {
        const int np = (int)((npts - p0) < chunk ? (npts - p0) : chunk);
        const long long nn = (long long)np * NQ;
        NodeBuf nb{c->d_nodef, nn, c->d_Jr, c->d_Jz, c->d_flag};
        HIPCHK(c, hipMemsetAsync(c->d_cnt, 0, 16 * sizeof(unsigned int), c->stream));
        SetupArgs sa{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, (int)p0, np, c->quad, nb,
                     c->d_cnt + 0, c->d_cnt + 1, c->d_list_u, c->d_list_v, c->d_ctr,
                     nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        unsigned int nsel = 0, njudge = 0, nfall = 0;
        SurArgs a{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, c->d_pVc, (int)p0, np, c->quad, nb, c->nm,
                  c->d_sx16, c->d_sy, c->sur_judge, c->sur_jzmin, hash32(c->sur_seed + 7919u * c->sur_iter),
                  c->d_cnt + 6, c->d_sel, c->d_cnt + 7, c->d_judge_node, c->d_judge_J, c->d_cnt + 8, c->df};
        if (sur) {
            // surrogate for every bound node, Staeckel for fallbacks and the judged sample
            HIPCHK(c, hipEventRecord(ev[2], c->stream));
            k_sur_infer<<<c->sur_blocks, 256, SUR_LDS, c->stream>>>(a, c->d_wpk, c->d_wbias);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipEventRecord(ev[3], c->stream));
            unsigned int cn[4];
            HIPCHK(c, hipMemcpyAsync(cn, c->d_cnt + 6, sizeof cn, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            nsel = cn[0]; njudge = cn[1]; nfall = cn[2];
            c->sst.nodes_sur += cn[3];
            t_sur += elapsed(ev[2], ev[3]);
            sa.sel = c->d_sel;
            sa.nsel = c->d_cnt + 6;
        }
        HIPCHK(c, hipEventRecord(ev[2], c->stream));
        // one chunk whose records the previous node phase left: lanes in the order of its turning points
        // (and, in mixed precision, tasks in the order of its interval counts)
        const bool order = FJ_SETUP_SORT && !sur && np == npts && c->rec_nodes == nn;
        if (order && (r = setup_order(c, nb, nn, &sa))) return r;
        const long long nsetup = sur ? (long long)nsel : nn;
        if (nsetup > 0) k_setup<<<(unsigned)((nsetup + 255) / 256), 256, lds, c->stream>>>(ta, sa);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(ev[3], c->stream));
        unsigned int novf = 0;
        if ((r = run_integrals(c, ta, nb, mixed, &t_int, &t_ovf, &novf, order))) return r;
        HIPCHK(c, hipEventRecord(ev[5], c->stream));
        MomArgs ma{c->d_pPhi, c->d_pVe, c->d_pR, (int)p0, np, c->quad, nb, c->df, c->d_out6,
                   c->d_mom, c->d_mom + npts, c->d_mom + 2 * npts, c->d_mom + 3 * npts, c->d_mom + 4 * npts,
                   c->d_mom + 5 * npts, &c->d_ctr->nonfinite};
        if (mixed) k_moments_f32<<<np, 256, 0, c->stream>>>(ma);
        else k_moments<<<np, 256, 0, c->stream>>>(ma);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(ev[6], c->stream));
        // training samples: every Staeckel node phase (probability sur_keep), or the judged nodes
        if (!sur) { if ((r = sur_collect(c, nb, (int)p0, np, nullptr, 0, c->sur_keep))) return r; }
        else if (c->sur_keep > 0 && njudge) {
            if ((r = sur_collect(c, nb, (int)p0, np, c->d_judge_node, njudge, 1.0))) return r;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        t_setup += elapsed(ev[2], ev[3]);
        t_mom += elapsed(ev[5], ev[6]);
        unsigned int cnts[2];
        HIPCHK(c, hipMemcpy(cnts, c->d_cnt, sizeof cnts, hipMemcpyDeviceToHost));
        c->st.tasks += (long long)cnts[0] + cnts[1];
        c->st.overflow_tasks += novf;
        if (sur) {
            c->sst.fallback += nfall;
            c->sst.judged += njudge;
            if ((r = sur_judge_stats(c, a, njudge))) return r;
        }
    }
    HIPCHK(c, hipEventRecord(ev[7], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the records hold every node's turning points when one chunk ran k_setup on all of them
    c->rec_nodes = (!sur && chunk >= npts) ? (long long)npts * NQ : -1;
    Counters h;
    HIPCHK(c, hipMemcpy(&h, c->d_ctr, sizeof h, hipMemcpyDeviceToHost));
    c->st.ms_points = elapsed(ev[0], ev[1]);
    c->st.ms_setup = t_setup; c->st.ms_integrals = t_int; c->st.ms_overflow = t_ovf; c->st.ms_moments = t_mom;
    c->st.ms_total = elapsed(ev[0], ev[7]);
    c->st.points = npts;
    c->st.nodes = (long long)npts * NQ;
    c->st.bound_nodes = (long long)h.bound;
    c->st.qk_calls = (long long)h.qk;
    c->st.brent_iters = (long long)h.brent;
    c->st.unbound_nodes = (long long)h.unbound;
    c->st.bracket_failures = (long long)h.bracket_fail;
    c->st.nonfinite_nodes = (long long)h.nonfinite;
    // k_points: Phi(R,z) and Phi(10 r_max, 10 r_max), plus the gradient at (R,z) for the surrogate's Vc
    const long long pts_grad = need_vc ? npts : 0;
    c->st.grad_evals = (long long)h.grad + pts_grad;
    c->st.phi_evals = (long long)h.phi + 21 * (long long)h.qk + 2 * (long long)npts + pts_grad;
    if (sur) c->sst.ms_surrogate = t_sur;
    if (c->sur_init) {
        unsigned long long pos = 0;
        HIPCHK(c, hipMemcpy(&pos, c->d_pos, sizeof pos, hipMemcpyDeviceToHost));
        c->sst.n_train = (long long)(pos < (unsigned long long)c->train_cap ? pos : (unsigned long long)c->train_cap);
    }
    for (int i = 0; i < 8; i++) (void)hipEventDestroy(ev[i]);
    return 0;
}

int ensure_grid_bufs(fj_ctx *c) {
    int r = 0;
    const size_t nl = (size_t)c->nr * c->npoly;
    if (!c->d_rhl) {
        if ((r = dalloc(c, &c->d_rhl, nl)) || (r = dalloc(c, &c->d_gl, nl * 9)) || (r = dalloc(c, &c->d_status, 1)))
            return r;
        // one lane per (radius, order), launched in blocks of 64
        if ((r = dalloc(c, &c->d_pws, ws_doubles(QCAP_FULL, (long long)((nl + 63) / 64 * 64))))) return r;
    }
    return 0;
}

int grid_points(fj_ctx *c) {
    const int npts = c->nr * c->ntheta;
    int r = ensure_points(c, npts);
    if (r) return r;
    std::vector<double> R(npts), z(npts);
    for (int n = 0; n < c->nr; n++)
        for (int g = 0; g < c->ntheta; g++) {
            R[n * c->ntheta + g] = c->ar[n] * c->si[g];
            z[n * c->ntheta + g] = c->ar[n] * c->ci[g];
        }
    // synchronous: R and z are released when this function returns
    HIPCHK(c, hipMemcpy(c->d_pR, R.data(), npts * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_pz, z.data(), npts * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int run_poisson(fj_ctx *c, int mode, const double *d_rhl, double *d_phil, double *d_Pr, double *d_Pr2) {
    if (mode != FJ_POISSON_QUADPACK && mode != FJ_POISSON_BLOCKSCAN)
        return set_err(c, FJ_E_ARG, "unknown Poisson mode %d", mode);
    int r = ensure_grid_bufs(c);
    if (r) return r;
    const int n = c->nr * c->npoly;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    if (mode == FJ_POISSON_BLOCKSCAN) {
        if (c->nr > PSCAN_MAXNR) return set_err(c, FJ_E_ARG, "block-scan Poisson supports nr <= %d", PSCAN_MAXNR);
        HIPCHK(c, hipEventRecord(a, c->stream));
        k_poisson_scan<<<c->npoly, PSCAN_THREADS, 0, c->stream>>>(c->nr, c->npoly, c->d_ar, d_rhl, d_phil, d_Pr, d_Pr2);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(b, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->st.ms_poisson = elapsed(a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return 0;
    }
    HIPCHK(c, hipMemsetAsync(c->d_status, 0, sizeof(int), c->stream));
    HIPCHK(c, hipEventRecord(a, c->stream));
    k_poisson<<<n, 64, 0, c->stream>>>(c->nr, c->npoly, c->d_ar, d_rhl, c->d_pws, QCAP_FULL, d_phil, d_Pr, d_Pr2,
                                       c->d_status);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(b, c->stream));
    int status = 0;
    HIPCHK(c, hipMemcpyAsync(&status, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->st.ms_poisson = elapsed(a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (status) return set_err(c, FJ_E_NUMERIC, "NaN in the external Poisson integral (Potential.cpp:144-146)");
    return 0;
}


// Actions at explicit phase-space points (d_pR, d_pz filled; velocities given).
int run_actions(fj_ctx *c, int n, const double *vR, const double *vphi, const double *vz, const double *Dx) {
    if (!c->have_pot || !c->have_delta) return set_err(c, FJ_E_STATE, "potential and Delta table must be set");
    int r = ensure_nodes(c, n);
    if (r) return r;
    double *dv = nullptr;
    HIPCHK(c, hipMalloc((void **)&dv, (size_t)n * 4 * sizeof(double)));
    HIPCHK(c, hipMemcpy(dv, vR, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dv + n, vphi, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dv + 2 * n, vz, n * sizeof(double), hipMemcpyHostToDevice));
    if (Dx) HIPCHK(c, hipMemcpy(dv + 3 * n, Dx, n * sizeof(double), hipMemcpyHostToDevice));
    const TabArgs ta = tab_args(c);
    const size_t lds = tab_bytes(c);
    PointArgs pa{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, c->d_pVc};
    k_points<<<(n + 255) / 256, 256, lds, c->stream>>>(ta, pa, n);
    NodeBuf nb{c->d_nodef, n, c->d_Jr, c->d_Jz, c->d_flag};
    c->last_nodes = n;
    c->rec_nodes = -1;
    HIPCHK(c, hipMemsetAsync(c->d_cnt, 0, 16 * sizeof(unsigned int), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_ctr, 0, sizeof(Counters), c->stream));
    SetupArgs sa{c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, 0, n, c->quad, nb, c->d_cnt + 0, c->d_cnt + 1,
                 c->d_list_u, c->d_list_v, c->d_ctr, dv, dv + n, dv + 2 * n, Dx ? dv + 3 * n : nullptr,
                 nullptr, nullptr};
    k_setup<<<(n + 255) / 256, 256, lds, c->stream>>>(ta, sa);
    HIPCHK(c, hipGetLastError());
    IntegArgs ia{nb, c->d_list_u, c->d_list_v, c->d_cnt + 0, c->d_cnt + 1, c->d_cnt + 2,
                 c->d_ovf, c->d_cnt + 3, c->d_ws, QCAP, -1, c->d_ctr};
    launch_integrals(c, ta, ia, integ_lanes(c, false), false);
    HIPCHK(c, hipGetLastError());
    unsigned int novf = 0;
    HIPCHK(c, hipMemcpyAsync(&novf, c->d_cnt + 3, sizeof novf, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (novf) {
        const int fl = 256 * 64;
        if (fl > c->wsfull_lanes) {
            if ((r = dalloc(c, &c->d_wsfull, ws_doubles(QCAP_FULL, fl)))) return r;
            c->wsfull_lanes = fl;
        }
        IntegArgs io = ia;
        io.next = c->d_cnt + 4; io.ws = c->d_wsfull; io.cap = QCAP_FULL; io.ntask_override = (int)novf;
        io.ovf_cnt = c->d_cnt + 5;
        launch_integrals(c, ta, io, fl, false);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    (void)hipFree(dv);
    return 0;
}
}  // namespace

// ===========================================================================
// extern "C"
// ===========================================================================
extern "C" {

int fj_ctx_create(int device, fj_ctx **out) {
    if (!out) return FJ_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FJ_E_HIP;
    if (device < 0 || device >= ndev) return FJ_E_ARG;
    fj_ctx *c = new fj_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return FJ_E_HIP;
    }
    c->own_stream = c->stream;
    // address of this device's scratch array (set-up, once per device)
    static std::mutex mu;
    static unsigned char *addr[64] = {};
    {
        std::lock_guard<std::mutex> lk(mu);
        if (device < 64 && !addr[device]) {
            void *ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
                (void)hipStreamDestroy(c->own_stream);
                delete c;
                return FJ_E_HIP;
            }
            addr[device] = static_cast<unsigned char *>(ptr);
        }
        c->noalloc = device < 64 ? addr[device] : nullptr;
    }
    *out = c;
    return 0;
}

int fj_set_stream(fj_ctx *c, void *stream) {
    if (!c) return FJ_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));   // work already queued stays ordered before the switch
    c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
    return 0;
}

int fj_ctx_destroy(fj_ctx *c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    void *ptrs[] = {c->d_ar, c->d_phil, c->d_Pr, c->d_Pr2, c->d_E, c->d_D, c->d_ci, c->d_si, c->d_poly, c->d_lut,
                    c->d_rk, c->d_nodef, c->d_Jr, c->d_Jz, c->d_flag, c->d_list_u, c->d_list_v, c->d_ovf, c->d_cnt,
                    c->d_ctr, c->d_ws, c->d_wsfull, c->d_pR, c->d_pz, c->d_pPhi, c->d_pVe, c->d_out6,
                    c->d_mom, c->d_rhl, c->d_gl, c->d_pw, c->d_pws, c->d_status, c->d_eval, c->d_fpart,
                    c->d_okey, c->d_operm, c->d_otemp, c->d_fflag, c->d_fJr, c->d_fJz, c->d_fkey,
                    c->d_tcost};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    void *sp[] = {c->d_pVc, c->d_X, c->d_T, c->d_bt, c->d_bw, c->d_pos, c->d_loss, c->d_sx16, c->d_sy, c->d_judge_J,
                  c->d_sel, c->d_judge_node, c->d_judge_S, c->mlp.p, c->mlp.g, c->mlp.m, c->mlp.v, c->mlp.p16,
                  c->mlp.x16, c->mlp.d16, c->mlp.dy16, c->mlp.pre, c->mlp.y, c->mlp.dh, c->mlp.dp32,
                  c->mlp.loss_part, c->mlp.h16[0], c->mlp.h16[1], c->mlp.h16[2], c->mlp.h32[0], c->mlp.h32[1],
                  c->mlp.h32[2], c->mlp.sk.part, c->d_wpk, c->d_wbias};
    for (void *p : sp)
        if (p) (void)hipFree(p);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return 0;
}

const char *fj_last_error(const fj_ctx *c) { return c ? c->err : "null context"; }

int fj_set_grid(fj_ctx *c, int nr, const double *ar, int ntheta, int npoly, int ngrid) {
    if (!c) return FJ_E_ARG;
    // nr >= 11: the quadratic interpolation of the first 10 cells uses the
    // point above the cell (UtilsLeg.h:113-115), which lies past the table
    // when the last cell is one of them.
    if (nr < 11 || !ar || ntheta < 1 || npoly < 1 || npoly > MAXPOLY || ngrid < 2)
        return set_err(c, FJ_E_ARG, "bad grid sizes nr=%d (>= 11) ntheta=%d npoly=%d (<= %d) ngrid=%d", nr, ntheta,
                       npoly, MAXPOLY, ngrid);
    if (!(ar[0] > 0)) return set_err(c, FJ_E_ARG, "radii must be positive");
    for (int i = 1; i < nr; i++)
        if (!(ar[i] > ar[i - 1])) return set_err(c, FJ_E_ARG, "radii must increase");
    HIPCHK(c, hipSetDevice(c->device));
    c->nr = nr; c->ntheta = ntheta; c->npoly = npoly; c->ngrid = ngrid;
    c->ar.assign(ar, ar + nr);
    c->ci.assign(ntheta, 0); c->wi.assign(ntheta, 0); c->si.assign(ntheta, 0);
    host_gauleg(ntheta, 0, 1, c->ci.data(), c->wi.data());
    for (int i = 0; i < ntheta; i++) c->si[i] = sqrt(1 - c->ci[i] * c->ci[i]);
    c->poly.assign((size_t)npoly * ntheta, 0);
    std::vector<double> pol(npoly);
    for (int g = 0; g < ntheta; g++) {
        host_evenlegend(npoly, c->ci[g], pol.data());
        for (int np = 0; np < npoly; np++) c->poly[np * ntheta + g] = 2 * pol[np] * c->wi[g];
    }
    const size_t nl = (size_t)nr * npoly;
    int r = 0;
    if ((r = dalloc(c, &c->d_ar, nr)) || (r = dalloc(c, &c->d_phil, nl)) || (r = dalloc(c, &c->d_Pr, nl)) ||
        (r = dalloc(c, &c->d_Pr2, nl)) || (r = dalloc(c, &c->d_E, ngrid)) || (r = dalloc(c, &c->d_D, ngrid)) ||
        (r = dalloc(c, &c->d_ci, ntheta)) || (r = dalloc(c, &c->d_si, ntheta)) ||
        (r = dalloc(c, &c->d_poly, (size_t)npoly * ntheta)) || (r = dalloc(c, &c->d_lut, NLUT)) ||
        (r = dalloc(c, &c->d_rk, (size_t)RK * nr)))
        return r;
    // grid-size dependent scratch is re-made on demand
    (void)dalloc(c, &c->d_rhl, 0); (void)dalloc(c, &c->d_gl, 0); (void)dalloc(c, &c->d_pws, 0);
    (void)dalloc(c, &c->d_status, 0);
    c->integ_lanes = 0;
    c->integ_lanes32 = 0;
    const LegCoef lc = make_legcoef(MAXPOLY);
    HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_legcoef), &lc, sizeof lc));
    int lut[NLUT];
    make_lut(nr, ar, lut, &c->lut_lg0, &c->lut_lsc);
    HIPCHK(c, hipMemcpy(c->d_lut, lut, sizeof lut, hipMemcpyHostToDevice));
    std::vector<double> rk((size_t)RK * nr);
    for (int b = 0; b < nr; b++) fill_radk_cell(nr, ar, b, &rk[(size_t)RK * b]);
    HIPCHK(c, hipMemcpy(c->d_rk, rk.data(), rk.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_ar, ar, nr * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_ci, c->ci.data(), ntheta * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_si, c->si.data(), ntheta * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_poly, c->poly.data(), c->poly.size() * sizeof(double), hipMemcpyHostToDevice));
    c->have_pot = false; c->have_delta = false;
    return 0;
}

int fj_set_quad_rule(fj_ctx *c, int preset) {
    if (!c) return FJ_E_ARG;
    if (preset != FJ_QUAD_LOG22 && preset != FJ_QUAD_TANHSINH22) return set_err(c, FJ_E_ARG, "unknown rule %d", preset);
    c->quad = preset;
    return 0;
}

int fj_set_precision(fj_ctx *c, int mode) {
    if (!c) return FJ_E_ARG;
    if (mode != FJ_PREC_FP64 && mode != FJ_PREC_MIXED && mode != FJ_PREC_FAST)
        return set_err(c, FJ_E_ARG, "unknown precision %d", mode);
    c->prec = mode;
    return 0;
}

int fj_set_fast_order(fj_ctx *c, int n) {
    if (!c) return FJ_E_ARG;
    if (n < 4 || n > FAST_GL_MAX) return set_err(c, FJ_E_ARG, "Gauss-Legendre order %d outside 4..%d", n, FAST_GL_MAX);
    c->fast_ngl = n;
    c->fgl_ready = false;
    return 0;
}

int fj_set_potential(fj_ctx *c, const double *phil, const double *Pr, const double *Pr2) {
    if (!c || !phil || !Pr || !Pr2) return FJ_E_ARG;
    if (!c->nr) return set_err(c, FJ_E_STATE, "fj_set_grid first");
    const size_t nb = (size_t)c->nr * c->npoly * sizeof(double);
    HIPCHK(c, hipMemcpy(c->d_phil, phil, nb, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_Pr, Pr, nb, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_Pr2, Pr2, nb, hipMemcpyHostToDevice));
    c->have_pot = true;
    return 0;
}

int fj_get_potential(fj_ctx *c, double *phil, double *Pr, double *Pr2) {
    if (!c || !phil || !Pr || !Pr2) return FJ_E_ARG;
    if (!c->have_pot) return set_err(c, FJ_E_STATE, "no potential set");
    const size_t nb = (size_t)c->nr * c->npoly * sizeof(double);
    HIPCHK(c, hipMemcpy(phil, c->d_phil, nb, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(Pr, c->d_Pr, nb, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(Pr2, c->d_Pr2, nb, hipMemcpyDeviceToHost));
    return 0;
}

int fj_set_delta_table(fj_ctx *c, const double *E, const double *D) {
    if (!c || !E || !D) return FJ_E_ARG;
    if (!c->nr) return set_err(c, FJ_E_STATE, "fj_set_grid first");
    HIPCHK(c, hipMemcpy(c->d_E, E, c->ngrid * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_D, D, c->ngrid * sizeof(double), hipMemcpyHostToDevice));
    c->have_delta = true;
    return 0;
}

int fj_get_delta_table(fj_ctx *c, double *E, double *D) {
    if (!c || !E || !D) return FJ_E_ARG;
    if (!c->have_delta) return set_err(c, FJ_E_STATE, "no Delta table");
    HIPCHK(c, hipMemcpy(E, c->d_E, c->ngrid * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(D, c->d_D, c->ngrid * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int fj_set_df(fj_ctx *c, const fj_df_params *df) {
    if (!c || !df) return FJ_E_ARG;
    if (df->chi < 0) return set_err(c, FJ_E_ARG, "chi must be >= 0 (DF.cpp:120-122)");
    if (!(df->J0 > 0)) return set_err(c, FJ_E_ARG, "J0 must be > 0");
    c->df = *df;
    c->have_df = true;
    return 0;
}

int fj_guess_rhl(fj_ctx *c, int model, double mass, double r0, double q, double *rhl) {
    if (!c || !rhl || model < 0 || model > 3) return FJ_E_ARG;
    if (!c->nr) return set_err(c, FJ_E_STATE, "fj_set_grid first");
    int r = ensure_grid_bufs(c);
    if (r) return r;
    const int n = c->nr * c->npoly;
    k_guess<<<(n + 63) / 64, 64, 0, c->stream>>>(c->nr, c->ntheta, c->npoly, c->d_ar, c->d_ci, c->d_si, c->d_poly,
                                                 model, mass, r0, q, c->d_rhl);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(rhl, c->d_rhl, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int fj_poisson(fj_ctx *c, int mode, const double *rhl, double *philP, double *PrP, double *Pr2P) {
    if (!c || !rhl || !philP || !PrP || !Pr2P) return FJ_E_ARG;
    if (!c->nr) return set_err(c, FJ_E_STATE, "fj_set_grid first");
    int r = ensure_grid_bufs(c);
    if (r) return r;
    const size_t n = (size_t)c->nr * c->npoly;
    HIPCHK(c, hipMemcpyAsync(c->d_rhl, rhl, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    double *o = c->d_gl + 6 * n;
    if ((r = run_poisson(c, mode, c->d_rhl, o, o + n, o + 2 * n))) return r;
    HIPCHK(c, hipMemcpy(philP, o, n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(PrP, o + n, n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(Pr2P, o + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int fj_moments(fj_ctx *c, double *rho, double *vrot, double *sigR, double *sigp, double *sigz, double *sigRz) {
    if (!c || !rho || !vrot || !sigR || !sigp || !sigz || !sigRz) return FJ_E_ARG;
    if (!c->nr) return set_err(c, FJ_E_STATE, "fj_set_grid first");
    int r = grid_points(c);
    if (r) return r;
    const int npts = c->nr * c->ntheta;
    if ((r = run_nodes(c, npts))) return r;
    double *outs[6] = {rho, vrot, sigR, sigp, sigz, sigRz};
    for (int m = 0; m < 6; m++)
        HIPCHK(c, hipMemcpy(outs[m], c->d_mom + (size_t)m * npts, npts * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int fj_rhofDF(fj_ctx *c, int n, const double *R, const double *z, double *out6, double *rho, double *vrot,
              double *sigR, double *sigp, double *sigz, double *sigRz) {
    if (!c || n < 1 || !R || !z) return FJ_E_ARG;
    int r = ensure_points(c, n);
    if (r) return r;
    HIPCHK(c, hipMemcpy(c->d_pR, R, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_pz, z, n * sizeof(double), hipMemcpyHostToDevice));
    if ((r = run_nodes(c, n))) return r;
    if (out6) HIPCHK(c, hipMemcpy(out6, c->d_out6, (size_t)n * 6 * sizeof(double), hipMemcpyDeviceToHost));
    double *outs[6] = {rho, vrot, sigR, sigp, sigz, sigRz};
    for (int m = 0; m < 6; m++)
        if (outs[m]) HIPCHK(c, hipMemcpy(outs[m], c->d_mom + (size_t)m * n, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// fj_rhofDF with every device buffer in the static scratch array: the same
// kernels and launch order as run_nodes (FP64 or mixed precision, no
// surrogate), with fewer integrator lanes; counters cleared by a kernel.
int fj_rhofDF_noalloc(fj_ctx *c, int n, const double *R, const double *z, double *out6, double *rho, double *vrot,
                      double *sigR, double *sigp, double *sigz, double *sigRz) {
    if (!c || !R || !z) return FJ_E_ARG;
    if (n < 1 || n > FJ_NOALLOC_MAX_POINTS)
        return set_err(c, FJ_E_ARG, "fj_rhofDF_noalloc: n = %d points, allowed 1..%d", n, FJ_NOALLOC_MAX_POINTS);
    if (!c->noalloc) return set_err(c, FJ_E_STATE, "fj_rhofDF_noalloc: no scratch address for device %d", c->device);
    if (!c->have_pot || !c->have_delta || !c->have_df)
        return set_err(c, FJ_E_STATE, "potential, Delta table and DF must be set before the node phase");
    if (c->sur_on) return set_err(c, FJ_E_STATE, "fj_rhofDF_noalloc: switch the action surrogate off first");
    if (c->prec == FJ_PREC_FAST) return set_err(c, FJ_E_STATE, "fj_rhofDF_noalloc: FJ_PREC_FP64 or FJ_PREC_MIXED only");
    HIPCHK(c, hipSetDevice(c->device));
    unsigned char *b = c->noalloc;
    auto D = [b](size_t off) { return reinterpret_cast<double *>(b + off); };
    auto I = [b](size_t off) { return reinterpret_cast<int *>(b + off); };
    unsigned int *cnt = reinterpret_cast<unsigned int *>(b + NA.cnt);
    Counters *ctr = reinterpret_cast<Counters *>(b + NA.ctr);
    const bool mixed = c->prec == FJ_PREC_MIXED;
    const long long nn = (long long)n * NQ;
    const TabArgs ta = tab_args(c);
    const size_t lds = tab_bytes(c);
    HIPCHK(c, hipMemcpyAsync(D(NA.R), R, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D(NA.z), z, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // the array keeps the previous call's contents: clear the counters
    kb_zero<<<1, 256, 0, c->stream>>>(cnt, 16);
    kb_zero<<<1, 256, 0, c->stream>>>(reinterpret_cast<unsigned int *>(ctr), (int)(sizeof(Counters) / 4));
    PointArgs pa{D(NA.R), D(NA.z), D(NA.Phi), D(NA.Ve), nullptr};
    k_points<<<(n + 255) / 256, 256, lds, c->stream>>>(ta, pa, n);
    HIPCHK(c, hipGetLastError());
    NodeBuf nb{D(NA.nodef), nn, D(NA.Jr), D(NA.Jz), I(NA.flag)};
    SetupArgs sa{D(NA.R), D(NA.z), D(NA.Phi), D(NA.Ve), 0, n, c->quad, nb, cnt + 0, cnt + 1, I(NA.list_u),
                 I(NA.list_v), ctr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    k_setup<<<(unsigned)((nn + 255) / 256), 256, lds, c->stream>>>(ta, sa);
    HIPCHK(c, hipGetLastError());
    IntegArgs ia{nb, I(NA.list_u), I(NA.list_v), cnt + 0, cnt + 1, cnt + 2, I(NA.ovf), cnt + 3, D(NA.ws),
                 mixed ? QCAP_MIXED : QCAP, -1, ctr};
    launch_integrals(c, ta, ia, NOALLOC_LANES, mixed);
    HIPCHK(c, hipGetLastError());
    unsigned int novf = 0;
    HIPCHK(c, hipMemcpyAsync(&novf, cnt + 3, sizeof novf, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (novf) {
        int fl = (int)((novf + 255) / 256 * 256);
        if (fl > NOALLOC_OVF_LANES) fl = NOALLOC_OVF_LANES;
        IntegArgs io = ia;
        io.next = cnt + 4; io.ws = D(NA.wsfull); io.cap = QCAP_FULL; io.ntask_override = (int)novf;
        io.ovf_cnt = cnt + 5;
        launch_integrals(c, ta, io, fl, false);
        HIPCHK(c, hipGetLastError());
    }
    MomArgs ma{D(NA.Phi), D(NA.Ve), D(NA.R), 0, n, c->quad, nb, c->df, D(NA.out6), D(NA.mom), D(NA.mom) + n,
               D(NA.mom) + 2 * n, D(NA.mom) + 3 * n, D(NA.mom) + 4 * n, D(NA.mom) + 5 * n, &ctr->nonfinite};
    if (mixed) k_moments_f32<<<n, 256, 0, c->stream>>>(ma);
    else k_moments<<<n, 256, 0, c->stream>>>(ma);
    HIPCHK(c, hipGetLastError());
    if (out6) HIPCHK(c, hipMemcpyAsync(out6, D(NA.out6), (size_t)n * 6 * sizeof(double), hipMemcpyDeviceToHost,
                                       c->stream));
    double *outs[6] = {rho, vrot, sigR, sigp, sigz, sigRz};
    for (int m = 0; m < 6; m++)
        if (outs[m]) HIPCHK(c, hipMemcpyAsync(outs[m], D(NA.mom) + (size_t)m * n, n * sizeof(double),
                                              hipMemcpyDeviceToHost, c->stream));
    Counters h;
    HIPCHK(c, hipMemcpyAsync(&h, ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    memset(&c->st, 0, sizeof c->st);
    c->st.points = n;
    c->st.nodes = nn;
    c->st.bound_nodes = (long long)h.bound;
    c->st.qk_calls = (long long)h.qk;
    c->st.brent_iters = (long long)h.brent;
    c->st.overflow_tasks = novf;
    c->st.integral_lanes = NOALLOC_LANES;
    c->st.unbound_nodes = (long long)h.unbound;
    c->st.bracket_failures = (long long)h.bracket_fail;
    c->st.nonfinite_nodes = (long long)h.nonfinite;
    c->st.grad_evals = (long long)h.grad;
    c->st.phi_evals = (long long)h.phi + 21 * (long long)h.qk + 2 * (long long)n;
    return 0;
}

int fj_project_legendre(fj_ctx *c, const double *rho, const double *vrot, const double *sigR, const double *sigp,
                        const double *sigz, const double *sigRz, double *rhl, double *vrotl, double *sigRl,
                        double *sigpl, double *sigzl, double *sigRzl) {
    if (!c || !rho || !vrot || !sigR || !sigp || !sigz || !sigRz || !rhl || !vrotl || !sigRl || !sigpl || !sigzl ||
        !sigRzl)
        return FJ_E_ARG;
    if (!c->nr) return set_err(c, FJ_E_STATE, "fj_set_grid first");
    const int npts = c->nr * c->ntheta;
    int r = ensure_points(c, npts);
    if (r || (r = ensure_grid_bufs(c))) return r;
    const double *in[6] = {rho, vrot, sigR, sigp, sigz, sigRz};
    for (int m = 0; m < 6; m++)
        HIPCHK(c, hipMemcpy(c->d_mom + (size_t)m * npts, in[m], npts * sizeof(double), hipMemcpyHostToDevice));
    const size_t nl = (size_t)c->nr * c->npoly;
    double *g = c->d_gl;
    k_project<<<(unsigned)((nl + 63) / 64), 64, 0, c->stream>>>(
        c->nr, c->ntheta, c->npoly, c->d_ar, c->d_poly, c->d_mom, c->d_mom + npts, c->d_mom + 2 * npts,
        c->d_mom + 3 * npts, c->d_mom + 4 * npts, c->d_mom + 5 * npts, g, g + nl, g + 2 * nl, g + 3 * nl, g + 4 * nl,
        g + 5 * nl);
    HIPCHK(c, hipGetLastError());
    double *outs[6] = {rhl, vrotl, sigRl, sigpl, sigzl, sigRzl};
    for (int m = 0; m < 6; m++) HIPCHK(c, hipMemcpy(outs[m], g + m * nl, nl * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int fj_compute_new_phi(fj_ctx *c, int mode, double *rhl, double *sigRl, double *sigpl, double *sigzl,
                       double *sigRzl, double *vrotl, double *philP, double *PrP, double *Pr2P) {
    if (!c || !rhl || !sigRl || !sigpl || !sigzl || !sigRzl || !vrotl || !philP || !PrP || !Pr2P) return FJ_E_ARG;
    if (!c->nr) return set_err(c, FJ_E_STATE, "fj_set_grid first");
    int r = grid_points(c);
    if (r || (r = ensure_grid_bufs(c))) return r;
    const int npts = c->nr * c->ntheta;
    if ((r = run_nodes(c, npts))) return r;
    const size_t nl = (size_t)c->nr * c->npoly;
    double *g = c->d_gl;   // rhl, vrotl, sigRl, sigpl, sigzl, sigRzl, phil, Pr, Pr2
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    HIPCHK(c, hipEventRecord(a, c->stream));
    k_project<<<(unsigned)((nl + 63) / 64), 64, 0, c->stream>>>(
        c->nr, c->ntheta, c->npoly, c->d_ar, c->d_poly, c->d_mom, c->d_mom + npts, c->d_mom + 2 * npts,
        c->d_mom + 3 * npts, c->d_mom + 4 * npts, c->d_mom + 5 * npts, g, g + nl, g + 2 * nl, g + 3 * nl, g + 4 * nl,
        g + 5 * nl);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const float tproj = elapsed(a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if ((r = run_poisson(c, mode, g, g + 6 * nl, g + 7 * nl, g + 8 * nl))) return r;
    c->st.ms_project = tproj;
    double *outs[9] = {rhl, vrotl, sigRl, sigpl, sigzl, sigRzl, philP, PrP, Pr2P};
    for (int m = 0; m < 9; m++) HIPCHK(c, hipMemcpy(outs[m], g + m * nl, nl * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int fj_eval_potential(fj_ctx *c, int n, const double *R, const double *z, double *phi, double *dR, double *dz) {
    if (!c || n < 1 || !R || !z || !phi || (dR && !dz) || (!dR && dz)) return FJ_E_ARG;
    if (!c->have_pot) return set_err(c, FJ_E_STATE, "no potential set");
    // context-owned buffer, grown on demand: repeated calls allocate nothing
    if (n > c->eval_cap) {
        const int cap = n > 1024 ? n : 1024;
        int r = dalloc(c, &c->d_eval, (size_t)cap * 5);
        if (r) { c->eval_cap = 0; return r; }
        c->eval_cap = cap;
    }
    double *d = c->d_eval;
    const size_t cap = (size_t)c->eval_cap, b = (size_t)n * sizeof(double);
    HIPCHK(c, hipMemcpyAsync(d, R, b, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d + cap, z, b, hipMemcpyHostToDevice, c->stream));
    const TabArgs ta = tab_args(c);
    k_eval<<<(n + 255) / 256, 256, tab_bytes(c), c->stream>>>(ta, n, d, d + cap, d + 2 * cap,
                                                              dR ? d + 3 * cap : nullptr, dR ? d + 4 * cap : nullptr);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(phi, d + 2 * cap, b, hipMemcpyDeviceToHost, c->stream));
    if (dR) {
        HIPCHK(c, hipMemcpyAsync(dR, d + 3 * cap, b, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(dz, d + 4 * cap, b, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}


int fj_actions(fj_ctx *c, int n, const double *R, const double *z, const double *vR, const double *vphi,
               const double *vz, const double *Delta, double *Jr, double *Jz) {
    if (!c || n < 1 || !R || !z || !vR || !vphi || !vz || !Jr || !Jz) return FJ_E_ARG;
    int r = ensure_points(c, n);
    if (r) return r;
    HIPCHK(c, hipMemcpy(c->d_pR, R, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_pz, z, n * sizeof(double), hipMemcpyHostToDevice));
    if ((r = run_actions(c, n, vR, vphi, vz, Delta))) return r;
    HIPCHK(c, hipMemcpy(Jr, c->d_Jr, n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(Jz, c->d_Jz, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------------------
// Observables (integ/Integ.cpp:33-43, 63-138, 165-236): f(J) at explicit
// phase-space points on the device, the quadrature sums on the host in the
// serial order of the reference's loops (GaussQuad.cpp:21-71, 124-134).
// ---------------------------------------------------------------------------
int fj_df_points(fj_ctx *c, int n, const double *R, const double *z, const double *vR, const double *vphi,
                 const double *vz, double *f) {
    if (!c || n < 1 || !R || !z || !vR || !vphi || !vz || !f) return FJ_E_ARG;
    if (!c->have_df) return set_err(c, FJ_E_STATE, "no DF set");
    int r = ensure_points(c, n);
    if (r) return r;
    HIPCHK(c, hipMemcpy(c->d_pR, R, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_pz, z, n * sizeof(double), hipMemcpyHostToDevice));
    if ((r = run_actions(c, n, vR, vphi, vz, nullptr))) return r;
    double *d = nullptr;
    HIPCHK(c, hipMalloc((void **)&d, (size_t)4 * n * sizeof(double)));
    hipError_t e = hipMemcpy(d, vR, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, vphi, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 2 * n, vz, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_df_points<<<(n + 255) / 256, 256, 0, c->stream>>>(n, c->d_pR, c->d_pPhi, d, d + n, d + 2 * n, c->d_Jr,
                                                           c->d_Jz, c->df, d + 3 * n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(f, d + 3 * n, n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(c, FJ_E_HIP, "fj_df_points: %s", hipGetErrorString(e));
    return 0;
}

namespace {

struct QuadTables {
    double xpt11[22], xw11[22], xpt01[22], xw01[22];
};

int quad_tables(fj_ctx *c, QuadTables &q) {
    HIPCHK(c, hipMemcpyFromSymbol(q.xpt11, HIP_SYMBOL(c_xpt11), sizeof q.xpt11));
    HIPCHK(c, hipMemcpyFromSymbol(q.xw11, HIP_SYMBOL(c_xw11), sizeof q.xw11));
    HIPCHK(c, hipMemcpyFromSymbol(q.xpt01, HIP_SYMBOL(c_xpt01), sizeof q.xpt01));
    HIPCHK(c, hipMemcpyFromSymbol(q.xw01, HIP_SYMBOL(c_xw01), sizeof q.xw01));
    return 0;
}

// topbottom (include/Utils.h:103-128): 0 if x is outside the table beyond
// the 1e-5 tolerance (bot and top then keep their values).
int host_topbottom(const double *xm, int n, double x, int *botx, int *topx) {
    int top = n - 1, bot = 0;
    if ((xm[top] - x) * (x - xm[bot]) < 0.) {
        if (fabs(xm[top] - x) < 1e-5 * fabs(x - xm[bot])) { *topx = top; *botx = top - 1; return 1; }
        if (fabs(x - xm[bot]) < 1e-5 * fabs(x - xm[top])) { *topx = bot + 1; *botx = bot; return 1; }
        return 0;
    }
    while (top - bot > 1) {
        const int m = (top + bot) / 2;
        if ((xm[top] - x) * (x - xm[m]) >= 0) bot = m;
        else top = m;
    }
    *topx = top; *botx = bot;
    return 1;
}

// ev_dens (include/UtilsLeg.h:323-334) with the one-table int_dens (285-298).
double host_ev_dens(const fj_ctx *c, const double *rhl, double R, double z) {
    const int NR = c->nr, NP = c->npoly;
    double pol[MAXPOLY], densp[MAXPOLY];
    const double cc = z / sqrt(R * R + z * z), r = sqrt(R * R + z * z);
    host_evenlegend(NP, cc, pol);
    if (r > c->ar[NR - 1]) {
        for (int i = 0; i < NP; i++) densp[i] = 0;
    } else {
        int top = 0, bot = 0;
        host_topbottom(c->ar.data(), NR, r, &bot, &top);
        const double f = (r - c->ar[bot]) / (c->ar[top] - c->ar[bot]);
        for (int i = 0; i < NP; i++) densp[i] = f * rhl[top * NP + i] + (1 - f) * rhl[bot * NP + i];
    }
    double dens = .5 * densp[0];
    for (int np = 1; np < NP; np++) {
        const double fac = .5 * (4 * np + 1);
        dens += fac * densp[np] * pol[np];
    }
    return dens;
}

// Phi at (R, z) and the escape speed of rhofDF (Integ.cpp:165-166).
int host_phi_ve(fj_ctx *c, double R, double z, double *Phi, double *Ve) {
    const double far = 10 * c->ar[c->nr - 1];
    const double Rs[2] = {R, far}, zs[2] = {z, far};
    double ph[2];
    int r;
    if ((r = fj_eval_potential(c, 2, Rs, zs, ph, nullptr, nullptr))) return r;
    *Phi = ph[0];
    *Ve = sqrt(-2 * (ph[0] - ph[1]));
    return 0;
}

}  // namespace

int fj_sigma_df(fj_ctx *c, const double *rhl, int n, const double *R, const double *z, double *sigma) {
    if (!c || !rhl || n < 1 || !R || !z || !sigma) return FJ_E_ARG;
    if (c->nr <= 0) return set_err(c, FJ_E_STATE, "no grid set");
    QuadTables q;
    int r;
    if ((r = quad_tables(c, q))) return r;
    const double rmax = c->ar[c->nr - 1];
    for (int m = 0; m < n; m++) {
        double sum = 0.;
        for (int i = 0; i < 22; i++) {
            const double rr = R[m] + q.xpt01[i] * (rmax - R[m]);
            const double jacob = rmax - R[m];
            sum += q.xw01[i] * (jacob * host_ev_dens(c, rhl, rr, z[m]));
        }
        sigma[m] = 2 * sum;
    }
    return 0;
}

int fj_line_profile(fj_ctx *c, const double *rhl, double R, double z, int n, const double *vx, double *out) {
    if (!c || !rhl || n < 1 || !vx || !out) return FJ_E_ARG;
    if (!c->have_pot || !c->have_delta || !c->have_df)
        return set_err(c, FJ_E_STATE, "potential, Delta table and DF must be set");
    QuadTables q;
    int r;
    double Phi, Ve, Sigma;
    if ((r = quad_tables(c, q)) || (r = host_phi_ve(c, R, z, &Phi, &Ve)) || (r = fj_sigma_df(c, rhl, 1, &R, &z, &Sigma)))
        return r;
    const double armax = c->ar[c->nr - 1];
    const long long nn = 22 * 22 * 22;
    const long long np = nn * n;
    if (np > (1ll << 30)) return set_err(c, FJ_E_ARG, "too many velocities in one call");
    std::vector<double> Rv(np, R), zv(np, z), v0(np), v1(np), v2(np), jac(np), f(np);
    for (int m = 0; m < n; m++)
        for (int i = 0; i < 22; i++)
            for (int j = 0; j < 22; j++)
                for (int k = 0; k < 22; k++) {
                    const long long t = m * nn + (i * 22 + j) * 22 + k;
                    double V[3];
                    fjob_lp_node(R, Ve, armax, vx[m], q.xpt11[i], q.xpt11[j], q.xpt11[k], V, &jac[t]);
                    v0[t] = V[0]; v1[t] = V[1]; v2[t] = V[2];
                }
    if ((r = fj_df_points(c, (int)np, Rv.data(), zv.data(), v0.data(), v1.data(), v2.data(), f.data()))) return r;
    for (int m = 0; m < n; m++) {
        double sum = 0;
        for (int i = 0; i < 22; i++)
            for (int j = 0; j < 22; j++)
                for (int k = 0; k < 22; k++) {
                    const long long t = m * nn + (i * 22 + j) * 22 + k;
                    sum += q.xw11[i] * q.xw11[j] * q.xw11[k] * (jac[t] * f[t]);
                }
        out[m] = sum / Sigma;
    }
    return 0;
}

int fj_velocity_distribution(fj_ctx *c, int which, double R, double z, int n, const double *v, double *out) {
    if (!c || which < 0 || which > 2 || n < 1 || !v || !out) return FJ_E_ARG;
    if (!c->have_pot || !c->have_delta || !c->have_df)
        return set_err(c, FJ_E_STATE, "potential, Delta table and DF must be set");
    QuadTables q;
    int r;
    double Phi, Ve;
    if ((r = quad_tables(c, q)) || (r = host_phi_ve(c, R, z, &Phi, &Ve))) return r;
    // Int2D_0101 (Pvphi: both axes [0,1]) or Int2D_0111 (second axis [-1,1])
    const double *xb = which == 0 ? q.xpt01 : q.xpt11, *wb = which == 0 ? q.xw01 : q.xw11;
    const long long nn = 22 * 22, np = nn * n;
    std::vector<double> Rv(np, R), zv(np, z), v0(np), v1(np), v2(np), f(np);
    for (int m = 0; m < n; m++)
        for (int i = 0; i < 22; i++)
            for (int j = 0; j < 22; j++) {
                const long long t = m * nn + i * 22 + j;
                double V[3];
                fjob_pv_node(which, Ve, v[m], q.xpt01[i], xb[j], V);
                v0[t] = V[0]; v1[t] = V[1]; v2[t] = V[2];
            }
    if ((r = fj_df_points(c, (int)np, Rv.data(), zv.data(), v0.data(), v1.data(), v2.data(), f.data()))) return r;
    const double jacob = Ve * Ve;
    for (int m = 0; m < n; m++) {
        double sum = 0;
        for (int i = 0; i < 22; i++)
            for (int j = 0; j < 22; j++) {
                const long long t = m * nn + i * 22 + j;
                sum += q.xw01[i] * wb[j] * (jacob * f[t]);
            }
        out[m] = sum;
    }
    return 0;
}

int fj_virial(fj_ctx *c, const double *rhl, const double *vrotl, const double *sigRl, const double *sigpl,
              const double *sigzl, double *out11) {
    if (!c || !rhl || !vrotl || !sigRl || !sigpl || !sigzl || !out11) return FJ_E_ARG;
    if (!c->have_pot) return set_err(c, FJ_E_STATE, "no potential set");
    const size_t nl = (size_t)c->nr * c->npoly;
    const int ncell = (c->nr - 1) * c->ntheta;
    double *d = nullptr;
    const size_t tot = 5 * nl + 6 * (size_t)(ncell + c->nr) + 16 + c->ntheta;
    HIPCHK(c, hipMalloc((void **)&d, tot * sizeof(double)));
    const double *in[5] = {rhl, vrotl, sigRl, sigpl, sigzl};
    for (int m = 0; m < 5; m++) HIPCHK(c, hipMemcpy(d + m * nl, in[m], nl * sizeof(double), hipMemcpyHostToDevice));
    double *terms = d + 5 * nl, *out = terms + 6 * (size_t)(ncell + c->nr), *wi = out + 16;
    HIPCHK(c, hipMemcpy(wi, c->wi.data(), c->ntheta * sizeof(double), hipMemcpyHostToDevice));
    const TabArgs ta = tab_args(c);
    const int nth = ncell + c->nr;
    k_vir_terms<<<(nth + 255) / 256, 256, tab_bytes(c), c->stream>>>(ta, c->ntheta, c->d_ci, wi, d, d + 2 * nl,
                                                                    d + 3 * nl, d + 4 * nl, d + nl, terms);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        k_vir_sum<<<1, 1, 0, c->stream>>>(c->nr, c->ntheta, c->d_ar, c->d_Pr, c->npoly, terms, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(out11, out, 11 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return set_err(c, FJ_E_HIP, "fj_virial: %s", hipGetErrorString(e));
    return 0;
}

int fj_tabulate_delta(fj_ctx *c, double *Egrid, double *Dgrid) {
    if (!c) return FJ_E_ARG;
    if (!c->have_pot) return set_err(c, FJ_E_STATE, "no potential set");
    double *d = nullptr;
    int *dst = nullptr;
    const int NG = c->ngrid;
    HIPCHK(c, hipMalloc((void **)&d, (8 + (size_t)NG * 402) * sizeof(double)));
    HIPCHK(c, hipMalloc((void **)&dst, sizeof(int)));
    HIPCHK(c, hipMemsetAsync(dst, 0, sizeof(int), c->stream));
    const TabArgs ta = tab_args(c);
    double *prep = d, *E = d + 8, *D = E + NG, *scr = D + NG;
    k_delta_prep<<<1, 64, tab_bytes(c), c->stream>>>(ta, prep);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        k_delta<<<NG, 64, tab_bytes(c), c->stream>>>(ta, prep, scr, E, D, dst);
        e = hipGetLastError();
    }
    int status = 0;
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(&status, dst, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(c->d_E, E, NG * sizeof(double), hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_D, D, NG * sizeof(double), hipMemcpyDeviceToDevice);
    if (e == hipSuccess && Egrid) e = hipMemcpy(Egrid, E, NG * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && Dgrid) e = hipMemcpy(Dgrid, D, NG * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    (void)hipFree(dst);
    if (e != hipSuccess) return set_err(c, FJ_E_HIP, "fj_tabulate_delta: %s", hipGetErrorString(e));
    c->have_delta = true;
    if (status & 1) return set_err(c, FJ_E_DELTA, "ensureVsq gave up after 50 tries (Delta.cpp:121)");
    if (status & 2) return set_err(c, FJ_E_DELTA, "no Jr=0 orbit bracket after 21 tries (Delta.cpp:145-151)");
    if (status & 8) return set_err(c, FJ_E_DELTA, "no Delta^2 bracket");
    return 0;
}


// ---------------------------------------------------------------------------
// Action surrogate
// ---------------------------------------------------------------------------
int fj_surrogate_init(fj_ctx *c, unsigned long long seed, long long capacity) {
    if (!c || capacity < 1024) return FJ_E_ARG;
    int r;
    Mlp &m = c->mlp;
    const long long cap = 1 << 20;   // rows per forward / training batch
    if (!c->sur_init) {
        const size_t N = MlpLayout::N, ch = (size_t)cap * MLP_H;
        if ((r = dalloc(c, &m.p, N)) || (r = dalloc(c, &m.g, N)) || (r = dalloc(c, &m.m, N)) ||
            (r = dalloc(c, &m.v, N)) || (r = dalloc(c, &m.p16, N)) || (r = dalloc(c, &m.x16, (size_t)cap * MLP_IN)) ||
            (r = dalloc(c, &m.d16, ch)) || (r = dalloc(c, &m.dy16, (size_t)cap * MLP_OUT)) ||
            (r = dalloc(c, &m.pre, ch)) || (r = dalloc(c, &m.y, (size_t)cap * MLP_OUT)) || (r = dalloc(c, &m.dh, ch)) ||
            (r = dalloc(c, &m.dp32, ch)) || (r = dalloc(c, &m.loss_part, (size_t)cap * MLP_OUT / 256 + 1)) ||
            (r = dalloc(c, &c->d_bt, (size_t)cap * MLP_OUT)) || (r = dalloc(c, &c->d_bw, (size_t)cap)) || (r = dalloc(c, &c->d_pos, 1)) ||
            (r = dalloc(c, &c->d_loss, 1)) || (r = dalloc(c, &m.sk.part, (size_t)256 * MLP_H * MLP_H)))
            return r;
        m.sk.cap = (long long)256 * MLP_H * MLP_H;
        for (int l = 0; l < 3; l++)
            if ((r = dalloc(c, &m.h16[l], ch)) || (r = dalloc(c, &m.h32[l], ch))) return r;
        if ((r = dalloc(c, &c->d_wpk, (size_t)MlpPacked::N16)) || (r = dalloc(c, &c->d_wbias, (size_t)3 * MLP_H + 2)))
            return r;
        m.cap = cap;
    }
    if (capacity != c->train_cap) {
        if ((r = dalloc(c, &c->d_X, (size_t)capacity * MLP_IN)) || (r = dalloc(c, &c->d_T, (size_t)capacity * SUR_TW)))
            return r;
        c->train_cap = capacity;
    }
    std::vector<float> p;
    mlp_init_params(p, seed);
    HIPCHK(c, hipMemcpy(m.p, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(m.m, 0, p.size() * sizeof(float)));
    HIPCHK(c, hipMemset(m.v, 0, p.size() * sizeof(float)));
    HIPCHK(c, hipMemset(c->d_pos, 0, sizeof(unsigned long long)));
    k_to_bf16<<<mlp_blocks(MlpLayout::N), 256, 0, c->stream>>>(MlpLayout::N, m.p, m.p16);
    HIPCHK(c, hipGetLastError());
    if ((r = sur_pack(c))) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    m.step = 0;
    c->sur_seed = (unsigned int)(seed ^ (seed >> 32));
    c->sur_init = true;
    c->sur_norm = false;
    c->sur_on = 0;
    memset(&c->sst, 0, sizeof c->sst);
#ifndef FJ_SUR_TFLOOR
#define FJ_SUR_TFLOOR 1e-6f
#endif
    c->nm.tfloor = FJ_SUR_TFLOOR;
    return 0;
}

int fj_surrogate_collect(fj_ctx *c, double keep) {
    if (!c || keep < 0 || keep > 1) return FJ_E_ARG;
    if (!c->sur_init) return set_err(c, FJ_E_STATE, "fj_surrogate_init first");
    c->sur_keep = keep;
    return 0;
}

int fj_surrogate_mode(fj_ctx *c, int on, double judge_fraction, double jz_min) {
    if (!c || judge_fraction < 0 || judge_fraction > 1) return FJ_E_ARG;
    if (on && !(c->sur_init && c->sur_norm)) return set_err(c, FJ_E_STATE, "train the surrogate first");
    c->sur_on = on;
    c->sur_judge = judge_fraction;
    c->sur_jzmin = jz_min;
    return 0;
}

namespace {
// Input / target normalisation and the training-set box, from the current set.
int sur_normalise(fj_ctx *c, long long n) {
    std::vector<float> X((size_t)n * MLP_IN), T((size_t)n * SUR_TW);
    HIPCHK(c, hipMemcpy(X.data(), c->d_X, X.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(T.data(), c->d_T, T.size() * sizeof(float), hipMemcpyDeviceToHost));
    auto stats = [&](const std::vector<float> &A, int w, int k, float &mu, float &sg) {
        double s = 0, s2 = 0;
        for (long long i = 0; i < n; i++) { const double x = A[(size_t)i * w + k]; s += x; s2 += x * x; }
        const double m = s / n, var = s2 / n - m * m;
        mu = (float)m;
        sg = (float)sqrt(var > 1e-12 ? var : 1e-12);
    };
    for (int k = 0; k < MLP_IN; k++) stats(X, MLP_IN, k, c->nm.mu[k], c->nm.sig[k]);
    for (int k = 0; k < MLP_OUT; k++) stats(T, SUR_TW, k, c->nm.tmu[k], c->nm.tsig[k]);
    const int nb = 256;
    float *part = nullptr;
    HIPCHK(c, hipMalloc((void **)&part, nb * 2 * MLP_IN * sizeof(float)));
    k_sur_box<<<nb, 256, 0, c->stream>>>(n, c->d_X, c->nm, part);
    std::vector<float> hp(nb * 2 * MLP_IN);
    HIPCHK(c, hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(part);
    for (int k = 0; k < MLP_IN; k++) {
        float lo = 3.4e38f, hi = -3.4e38f;
        for (int b = 0; b < nb; b++) {
            lo = fminf(lo, hp[b * 2 * MLP_IN + k]);
            hi = fmaxf(hi, hp[b * 2 * MLP_IN + MLP_IN + k]);
        }
        c->nm.lo[k] = lo;
        c->nm.hi[k] = hi;
    }
    c->sur_norm = true;
    return 0;
}
}  // namespace

int fj_surrogate_train(fj_ctx *c, int epochs, int batch, double lr, double *loss) {
    if (!c || epochs < 0 || batch < 64 || !(lr > 0)) return FJ_E_ARG;
    if (!c->sur_init) return set_err(c, FJ_E_STATE, "fj_surrogate_init first");
    const long long n = c->sst.n_train;
    if (n < 1024) return set_err(c, FJ_E_STATE, "only %lld training samples", n);
    if (batch > c->mlp.cap) batch = (int)c->mlp.cap;
    int r;
    if (!c->sur_norm && (r = sur_normalise(c, n))) return r;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    HIPCHK(c, hipEventRecord(a, c->stream));
    const long long P = (long long)(2147483647ll % n);   // prime > n: coprime with n
    double last = 0;
    for (int e = 0; e < epochs; e++) {
        const float lre = (float)(lr * 0.5 * (1 + cos(3.141592653589793 * e / (epochs > 1 ? epochs : 1))));
        const long long off = (long long)(hash32(c->sur_seed + 104729u * (unsigned)(c->mlp.step + e)) % (unsigned)n);
        HIPCHK(c, hipMemsetAsync(c->d_loss, 0, sizeof(double), c->stream));
        for (long long j = 0; j < n; j += batch) {
            const long long B = n - j < batch ? n - j : batch;
            k_sur_batch<<<mlp_blocks(B), 256, 0, c->stream>>>(B, j, P, n, off, c->d_X, c->d_T, c->nm, c->mlp.x16,
                                                             c->d_bt, c->d_bw);
            mlp_forward(c->stream, c->mlp, B, c->mlp.x16, c->mlp.y);
            mlp_backward(c->stream, c->mlp, B, c->d_bt, c->d_bw);
            k_sum_into<<<1, 256, 0, c->stream>>>(mlp_blocks(B * MLP_OUT), c->mlp.loss_part, c->d_loss);
            mlp_adam(c->stream, c->mlp, lre);
        }
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(&last, c->d_loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        last /= (double)n * MLP_OUT;
    }
    if ((r = sur_pack(c))) return r;
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->sst.ms_train = elapsed(a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    c->sst.last_loss = last;
    if (loss) *loss = last;
    return 0;
}

int fj_surrogate_get_stats(const fj_ctx *c, fj_surrogate_stats *out) {
    if (!c || !out) return FJ_E_ARG;
    *out = c->sst;
    return 0;
}

int fj_surrogate_params(fj_ctx *c, float *get, const float *set) {
    if (!c) return FJ_E_ARG;
    if (!c->sur_init) return set_err(c, FJ_E_STATE, "fj_surrogate_init first");
    const size_t N = MlpLayout::N;
    if (set) {
        HIPCHK(c, hipMemcpy(c->mlp.p, set, N * sizeof(float), hipMemcpyHostToDevice));
        k_to_bf16<<<mlp_blocks((long long)N), 256, 0, c->stream>>>((long long)N, c->mlp.p, c->mlp.p16);
        HIPCHK(c, hipGetLastError());
        int r;
        if ((r = sur_pack(c))) return r;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (get) HIPCHK(c, hipMemcpy(get, c->mlp.p, N * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

namespace {
int sur_upload_inputs(fj_ctx *c, long long n, const float *x) {
    float *dx = nullptr;
    HIPCHK(c, hipMalloc((void **)&dx, (size_t)n * MLP_IN * sizeof(float)));
    HIPCHK(c, hipMemcpy(dx, x, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
    k_to_bf16<<<mlp_blocks(n * MLP_IN), 256, 0, c->stream>>>(n * MLP_IN, dx, c->mlp.x16);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(dx);
    return 0;
}
}  // namespace

int fj_surrogate_forward(fj_ctx *c, long long n, const float *x, float *y) {
    if (!c || n < 1 || !x || !y) return FJ_E_ARG;
    if (!c->sur_init) return set_err(c, FJ_E_STATE, "fj_surrogate_init first");
    if (n > c->mlp.cap) return set_err(c, FJ_E_ARG, "at most %lld rows", c->mlp.cap);
    float *dx = nullptr;
    HIPCHK(c, hipMalloc((void **)&dx, (size_t)n * MLP_IN * sizeof(float)));
    HIPCHK(c, hipMemcpy(dx, x, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
    k_mlp_infer_x<<<c->sur_blocks, 256, SUR_LDS, c->stream>>>(n, dx, c->d_wpk, c->d_wbias, c->mlp.y);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(y, c->mlp.y, (size_t)n * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(dx);
    return 0;
}

int fj_surrogate_gradient(fj_ctx *c, long long n, const float *x, const float *t, float *grad, double *loss) {
    if (!c || n < 1 || !x || !t || !grad) return FJ_E_ARG;
    if (!c->sur_init) return set_err(c, FJ_E_STATE, "fj_surrogate_init first");
    if (n > c->mlp.cap) return set_err(c, FJ_E_ARG, "at most %lld rows", c->mlp.cap);
    int r;
    if ((r = sur_upload_inputs(c, n, x))) return r;
    HIPCHK(c, hipMemcpy(c->d_bt, t, (size_t)n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemsetAsync(c->d_loss, 0, sizeof(double), c->stream));
    mlp_forward(c->stream, c->mlp, n, c->mlp.x16, c->mlp.y);
    mlp_backward(c->stream, c->mlp, n, c->d_bt);
    k_sum_into<<<1, 256, 0, c->stream>>>(mlp_blocks(n * MLP_OUT), c->mlp.loss_part, c->d_loss);
    HIPCHK(c, hipGetLastError());
    double l = 0;
    HIPCHK(c, hipMemcpyAsync(grad, c->mlp.g, MlpLayout::N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&l, c->d_loss, sizeof l, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = l / ((double)n * MLP_OUT);
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched models (host side)
// ---------------------------------------------------------------------------
struct fj_batch {
    fj_ctx *c = nullptr;
    int M = 0, npts = 0, rows = 0, npoly_max = 0, cus = 0;
    long long nodes = 0;
    std::vector<ModelDev> hmd;
    std::vector<void *> allocs;
    ModelDev *d_md = nullptr;
    PotView *d_pv = nullptr;
    int *d_pmodel = nullptr, *d_bstart[2] = {nullptr, nullptr}, *d_rstart = nullptr, *d_block_model = nullptr;
    int nblk[2] = {0, 0};                      // setup blocks for block sizes 128 and 256
    double *d_R = nullptr, *d_z = nullptr, *d_Phi = nullptr, *d_Ve = nullptr, *d_Vc = nullptr, *d_out6 = nullptr,
           *d_mom = nullptr;
    double *d_nodef = nullptr, *d_Jr = nullptr, *d_Jz = nullptr;
    int *d_flag = nullptr, *d_list_u = nullptr, *d_list_v = nullptr, *d_ovf = nullptr;
    unsigned int *d_cnt = nullptr;
    Counters *d_ctr = nullptr;
    double *d_ws = nullptr, *d_wsfull = nullptr, *d_backup = nullptr;
    long long ws_lanes = 0, wsfull_lanes = 0;
    size_t lds_tab = 0, lds_integ = 0, lds_integ32 = 0;   // LDS of the FP64 / FP32 persistent integrators
    int per_cu64 = 1, per_cu32 = 1;                         // their resident blocks per CU
    int setup_bs = 256, integ_blocks = 0, integ_max_blocks = 0, ovf_blocks = 64;
    bool have_setup = false;
    // fast node phase (a batch created under FJ_PREC_FAST): the models one
    // after another through k_fast_nodes and k_moments, with one model's node
    // records at a time (fmax_nodes: the largest model's nodes) and every
    // node's sort key for the next iteration's lane order
    bool fast = false;
    int fast_ngl = 0;
    long long fmax_nodes = 0;
    int *d_bfflag = nullptr, *d_bfperm = nullptr;   // perm: [2][fmax_nodes], in and out of the sort
    double *d_bfJr = nullptr, *d_bfJz = nullptr;
    unsigned int *d_bfkey = nullptr, *d_bfkout = nullptr;   // key: [nodes] (all models), kout: [fmax_nodes]
    unsigned char *d_bftemp = nullptr;
    size_t bftemp_bytes = 0;
    bool fkeys_ok = false;
    size_t lds_fast = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // lane order of kb_setup (FP64 / mixed): keys and permutation, in and out of
    // the radix sort [2][nodes] each; kb_setup reads the sorted permutation
    unsigned long long *d_bkey = nullptr;
    int *d_bperm = nullptr;
    unsigned char *d_btemp = nullptr;
    size_t btemp_bytes = 0;
    bool rec_ok = false;   // the records hold an iteration's turning points
    int order_bits = SETUP_KEY_BITS;
    fj_batch_stats st{};
};

namespace {

template <class T>
int balloc(fj_batch *b, T **p, size_t n) {
    *p = nullptr;
    const hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(b->c, FJ_E_HIP, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    }
    b->allocs.push_back(*p);
    return 0;
}

void batch_drop_graph(fj_batch *b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

// kb_setup's lane order for the next iteration: every model's nodes sorted
// by the turning points the previous iteration left (on the stream, outside
// any captured graph).  Nothing to do before the first iteration.
int batch_order(fj_batch *b) {
    fj_ctx *c = b->c;
    if (b->fast || !FJ_SETUP_SORT || !b->rec_ok) return 0;
    const long long nn = b->nodes;
    NodeBuf nb{b->d_nodef, nn, b->d_Jr, b->d_Jz, b->d_flag};
    kb_order_key<<<(unsigned)((nn + 255) / 256), 256, 0, c->stream>>>(nb, nn, b->d_pmodel, b->d_bkey, b->d_bperm);
    HIPCHK(c, hipGetLastError());
    size_t bytes = b->btemp_bytes;
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(b->d_btemp, bytes, b->d_bkey, b->d_bkey + nn, b->d_bperm,
                                                 b->d_bperm + nn, (int)nn, 0, b->order_bits, c->stream));
    return 0;
}

// One SCF iteration of every model, stream-ordered, no host synchronisation.
int batch_iteration(fj_batch *b) {
    fj_ctx *c = b->c;
    hipStream_t s = c->stream;
    const int M = b->M;
    const bool mixed = c->prec == FJ_PREC_MIXED;
    if (b->fast) {
        kb_zero64<<<1, 64, 0, s>>>(reinterpret_cast<unsigned long long *>(b->d_ctr), (int)(sizeof(Counters) / 8));
        PointArgs pa{b->d_R, b->d_z, b->d_Phi, b->d_Ve, b->d_Vc};
        kb_points<<<(b->npts + 255) / 256, 256, 0, s>>>(b->d_md, b->d_pmodel, pa, b->npts);
        // each model: its nodes in the order of the previous iteration's turning points (model-local
        // indices), the node kernel, then its points' moments; one model's records at a time
        const long long fm = b->fmax_nodes;
        for (int m = 0; m < M; m++) {
            const ModelDev &d = b->hmd[m];
            const long long n = (long long)d.np * NQ, off = (long long)d.po * NQ;
            const int *sel = nullptr;
            if (FJ_FAST_SORT && b->fkeys_ok) {
                NodeBuf none{nullptr, n, nullptr, nullptr, nullptr};
                k_setup_key<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(none, n, nullptr, b->d_bfperm, nullptr);
                size_t bytes = b->bftemp_bytes;
                HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(b->d_bftemp, bytes, b->d_bfkey + off, b->d_bfkout,
                                                             b->d_bfperm, b->d_bfperm + fm, (int)n, 0,
                                                             SETUP_KEY_BITS, s));
                sel = b->d_bfperm + fm;
            }
            NodeBuf nb{nullptr, n, b->d_bfJr, b->d_bfJz, b->d_bfflag};
            FastNodeArgs fa{b->d_R + d.po, b->d_z + d.po, b->d_Phi + d.po, b->d_Ve + d.po, d.np, d.quad,
                            c->fast_ngl, fast_v_order(c->fast_ngl), nb, sel, b->d_bfkey + off, b->d_ctr};
            k_fast_nodes<<<(unsigned)((n + 255) / 256), 256, b->lds_fast, s>>>(d.ta, fa);
            MomArgs ma{b->d_Phi, b->d_Ve, b->d_R, d.po, d.np, d.quad, nb, d.df, b->d_out6,
                       b->d_mom, b->d_mom + b->npts, b->d_mom + 2 * b->npts, b->d_mom + 3 * b->npts,
                       b->d_mom + 4 * b->npts, b->d_mom + 5 * b->npts, &b->d_ctr->nonfinite};
            k_moments<<<d.np, 256, 0, s>>>(ma);
        }
        kb_project<<<(b->rows + 63) / 64, 64, 0, s>>>(b->d_md, b->d_rstart, M, b->d_mom, b->npts);
        kb_poisson<<<dim3(b->npoly_max, M), PSCAN_THREADS, 0, s>>>(b->d_md);
        HIPCHK(c, hipGetLastError());
        return 0;
    }
    kb_zero<<<(8 * M + 8 + 255) / 256, 256, 0, s>>>(b->d_cnt, 8 * M + 8);
    kb_zero64<<<1, 64, 0, s>>>(reinterpret_cast<unsigned long long *>(b->d_ctr), (int)(sizeof(Counters) / 8));
    PointArgs pa{b->d_R, b->d_z, b->d_Phi, b->d_Ve, b->d_Vc};
    kb_points<<<(b->npts + 255) / 256, 256, 0, s>>>(b->d_md, b->d_pmodel, pa, b->npts);
    NodeBuf nb{b->d_nodef, b->nodes, b->d_Jr, b->d_Jz, b->d_flag};
    SetupArgs sa{b->d_R, b->d_z, b->d_Phi, b->d_Ve, 0, 0, 0, nb, nullptr, nullptr, b->d_list_u, b->d_list_v,
                 b->d_ctr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const int *perm = b->d_bperm + b->nodes;
    if (b->setup_bs == 128)
        kb_setup<128><<<b->nblk[0], 128, b->lds_tab, s>>>(b->d_md, b->d_bstart[0], M, sa, b->d_cnt, perm);
    else
        kb_setup<256><<<b->nblk[1], 256, b->lds_tab, s>>>(b->d_md, b->d_bstart[1], M, sa, b->d_cnt, perm);
    // integrator blocks: the chosen count, at most what is resident at this precision
    int ib = b->cus * (mixed ? b->per_cu32 : b->per_cu64);
    if (b->integ_blocks < ib) ib = b->integ_blocks;
    if (ib < M) ib = M;
    kb_assign<<<1, 64, 0, s>>>(M, b->d_cnt, ib, b->d_block_model);
    unsigned int *g = b->d_cnt + 8 * M;   // [0] overflowed, [1] next overflow task, [2] second overflow
    IntegArgs ia{nb, b->d_list_u, b->d_list_v, nullptr, nullptr, nullptr, b->d_ovf, g,
                 b->d_ws, mixed ? QCAP_MIXED : QCAP, -1, b->d_ctr};
    if (mixed)
        kb_integrals_f32<<<ib, 256, b->lds_integ32, s>>>(b->d_md, b->d_block_model, b->M, ia, b->d_cnt);
    else
        kb_integrals<<<ib, 256, b->lds_integ, s>>>(b->d_md, b->d_block_model, b->M, ia, b->d_cnt);
    IntegArgs io{nb, nullptr, nullptr, nullptr, nullptr, g + 1, b->d_ovf, g + 2, b->d_wsfull, QCAP_FULL, -1,
                 b->d_ctr, g, b->d_pv, b->d_pmodel};
    kb_integrals_ovf<<<b->ovf_blocks, 256, 20 * 256 * sizeof(double), s>>>(io);
    MomArgs ma{b->d_Phi, b->d_Ve, b->d_R, 0, b->npts, 0, nb, fj_df_params{}, b->d_out6, b->d_mom,
               b->d_mom + b->npts, b->d_mom + 2 * b->npts, b->d_mom + 3 * b->npts, b->d_mom + 4 * b->npts,
               b->d_mom + 5 * b->npts, &b->d_ctr->nonfinite};
    kb_moments<<<b->npts, 256, 0, s>>>(b->d_md, b->d_pmodel, ma, mixed ? 1 : 0);
    kb_project<<<(b->rows + 63) / 64, 64, 0, s>>>(b->d_md, b->d_rstart, M, b->d_mom, b->npts);
    kb_poisson<<<dim3(b->npoly_max, M), PSCAN_THREADS, 0, s>>>(b->d_md);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// Grid points of every model (R = r sin theta, z = r cos theta at the Gauss-
// Legendre cos theta nodes), the model table, point -> model map, row
// prefixes and the Legendre constants, copied to the device (the fast-phase
// batch; fj_batch_create does the same inline).
int batch_points(fj_batch *b, const fj_model_spec *spec, const std::vector<int> &pmodel,
                 const std::vector<int> &rstart) {
    fj_ctx *c = b->c;
    const int M = b->M;
    std::vector<double> R(b->npts), z(b->npts);
    for (int m = 0; m < M; m++) {
        const ModelDev &d = b->hmd[m];
        std::vector<double> hc(d.ntheta), hw(d.ntheta);
        host_gauleg(d.ntheta, 0, 1, hc.data(), hw.data());
        for (int n = 0; n < d.nr; n++)
            for (int g = 0; g < d.ntheta; g++) {
                R[d.po + n * d.ntheta + g] = spec[m].ar[n] * sqrt(1 - hc[g] * hc[g]);
                z[d.po + n * d.ntheta + g] = spec[m].ar[n] * hc[g];
            }
    }
    HIPCHK(c, hipMemcpy(b->d_md, b->hmd.data(), M * sizeof(ModelDev), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_pmodel, pmodel.data(), pmodel.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_rstart, rstart.data(), (M + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_R, R.data(), R.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_z, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice));
    const LegCoef lc = make_legcoef(MAXPOLY);
    HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_legcoef), &lc, sizeof lc));
    return 0;
}

// The rest of fj_batch_create for a fast-phase batch: no node records, no
// integrator workspaces; tiles of 64 points per model and the slice sums.
int batch_create_fast(fj_batch *b, const fj_model_spec *spec, const std::vector<int> &pmodel,
                      std::vector<int> &rstart, size_t lds_max, fj_batch **out) {
    fj_ctx *c = b->c;
    const int M = b->M;
    int r;
    rstart.push_back(b->rows);
    b->lds_tab = lds_max;
    long long fm = 0;
    for (int m = 0; m < M; m++) {
        const long long n = (long long)b->hmd[m].np * NQ;
        if (n > fm) fm = n;
    }
    b->fmax_nodes = fm;
    if (fm > 0x7fffffffll) {
        const long long n = fm;
        fj_batch_destroy(b);
        return set_err(c, FJ_E_ARG, "fast batch: a model of %lld nodes; at most %lld per model", n, 0x7fffffffll);
    }
    if ((r = balloc(b, &b->d_md, M)) || (r = balloc(b, &b->d_pv, M)) || (r = balloc(b, &b->d_pmodel, b->npts)) ||
        (r = balloc(b, &b->d_rstart, M + 1)) || (r = balloc(b, &b->d_R, b->npts)) || (r = balloc(b, &b->d_z, b->npts)) ||
        (r = balloc(b, &b->d_Phi, b->npts)) || (r = balloc(b, &b->d_Ve, b->npts)) || (r = balloc(b, &b->d_Vc, b->npts)) ||
        (r = balloc(b, &b->d_out6, (size_t)6 * b->npts)) || (r = balloc(b, &b->d_mom, (size_t)6 * b->npts)) ||
        (r = balloc(b, &b->d_cnt, (size_t)8 * M + 8)) || (r = balloc(b, &b->d_ctr, 1)) ||
        (r = balloc(b, &b->d_backup, (size_t)3 * b->rows)) || (r = balloc(b, &b->d_bfflag, (size_t)fm)) ||
        (r = balloc(b, &b->d_bfJr, (size_t)fm)) || (r = balloc(b, &b->d_bfJz, (size_t)fm)) ||
        (r = balloc(b, &b->d_bfkout, (size_t)fm)) || (r = balloc(b, &b->d_bfperm, (size_t)2 * fm)) ||
        (r = balloc(b, &b->d_bfkey, (size_t)b->nodes))) {
        fj_batch_destroy(b);
        return r;
    }
    // the sort's scratch for the largest model (the others need no more)
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, b->bftemp_bytes, b->d_bfkey, b->d_bfkout, b->d_bfperm,
                                                 b->d_bfperm + fm, (int)fm, 0, SETUP_KEY_BITS, c->stream));
    if ((r = balloc(b, &b->d_bftemp, b->bftemp_bytes))) { fj_batch_destroy(b); return r; }
    if ((r = batch_points(b, spec, pmodel, rstart))) { fj_batch_destroy(b); return r; }
    b->st.nodes = b->nodes;
    b->st.points = b->npts;
    // each model's table view (the Delta fit of fj_batch_setup reads them)
    kb_potviews<<<(M + 63) / 64, 64, 0, c->stream>>>(b->d_md, M, b->d_pv);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = b;
    return 0;
}

// Installed tables of every model <-> backup (autotuning restores them).
int batch_save(fj_batch *b, bool restore) {
    size_t off = 0;
    for (int m = 0; m < b->M; m++) {
        const ModelDev &d = b->hmd[m];
        const size_t n = (size_t)d.nr * d.npoly;
        double *tab[3] = {const_cast<double *>(d.ta.phil), const_cast<double *>(d.ta.Pr), d.Pr2};
        for (int k = 0; k < 3; k++) {
            double *bk = b->d_backup + off;
            HIPCHK(b->c, hipMemcpyAsync(restore ? tab[k] : bk, restore ? bk : tab[k], n * sizeof(double),
                                        hipMemcpyDeviceToDevice, b->c->stream));
            off += n;
        }
    }
    return 0;
}

}  // namespace

extern "C" {

int fj_batch_create(fj_ctx *c, int M, const fj_model_spec *spec, fj_batch **out) {
    if (!c || M < 1 || !spec || !out) return FJ_E_ARG;
    *out = nullptr;
    HIPCHK(c, hipSetDevice(c->device));
    fj_batch *b = new fj_batch();
    b->c = c;
    b->M = M;
    b->fast = c->prec == FJ_PREC_FAST;
    int r = 0;
    const int NG = 30;
    std::vector<int> pmodel, bstart[2], rstart;
    size_t lds_max = 0;
    for (int m = 0; m < M && !r; m++) {
        const fj_model_spec &sp = spec[m];
        if (sp.nr < 11 || !sp.ar || sp.ntheta < 1 || sp.npoly < 1 || sp.npoly > MAXPOLY || sp.nr > PSCAN_MAXNR) {
            r = set_err(c, FJ_E_ARG, "model %d: need 11 <= nr <= %d, ntheta >= 1, 1 <= npoly <= %d", m, PSCAN_MAXNR,
                        MAXPOLY);
            break;
        }
        for (int i = 0; i < sp.nr; i++)
            if (!(sp.ar[i] > 0) || (i && !(sp.ar[i] > sp.ar[i - 1]))) {
                r = set_err(c, FJ_E_ARG, "model %d: radii must be positive and increasing", m);
                break;
            }
        if (r) break;
        ModelDev d{};
        const int nr = sp.nr, np_ = sp.npoly, nt = sp.ntheta, nl = nr * np_;
        double *ar, *phil, *Pr, *Pr2, *E, *D, *ci, *si, *poly, *co, *prep, *rk;
        int *lut;
        if ((r = balloc(b, &ar, nr)) || (r = balloc(b, &phil, nl)) || (r = balloc(b, &Pr, nl)) ||
            (r = balloc(b, &Pr2, nl)) || (r = balloc(b, &E, NG)) || (r = balloc(b, &D, NG)) ||
            (r = balloc(b, &ci, nt)) || (r = balloc(b, &si, nt)) || (r = balloc(b, &poly, (size_t)np_ * nt)) ||
            (r = balloc(b, &co, (size_t)6 * nl)) || (r = balloc(b, &prep, 4)) || (r = balloc(b, &lut, NLUT)) ||
            (r = balloc(b, &rk, (size_t)RK * nr)))
            break;
        std::vector<double> hc(nt), hw(nt), hs(nt), hp((size_t)np_ * nt), pol(np_);
        host_gauleg(nt, 0, 1, hc.data(), hw.data());
        for (int g = 0; g < nt; g++) {
            hs[g] = sqrt(1 - hc[g] * hc[g]);
            host_evenlegend(np_, hc[g], pol.data());
            for (int k = 0; k < np_; k++) hp[(size_t)k * nt + g] = 2 * pol[k] * hw[g];
        }
        int hl[NLUT];
        float lg0, lsc;
        make_lut(nr, sp.ar, hl, &lg0, &lsc);
        std::vector<double> hrk((size_t)RK * nr);
        for (int k = 0; k < nr; k++) fill_radk_cell(nr, sp.ar, k, &hrk[(size_t)RK * k]);
        HIPCHK(c, hipMemcpy(rk, hrk.data(), hrk.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(ar, sp.ar, nr * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(ci, hc.data(), nt * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(si, hs.data(), nt * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(poly, hp.data(), hp.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(lut, hl, sizeof hl, hipMemcpyHostToDevice));
        d.ta.nr = nr; d.ta.npoly = np_; d.ta.ngrid = NG; d.ta.ar = ar; d.ta.phil = phil; d.ta.Pr = Pr;
        d.ta.E = E; d.ta.D = D; d.ta.lut = lut; d.ta.lg0 = lg0; d.ta.lsc = lsc; d.ta.rk = rk;
        d.Pr2 = Pr2; d.df = sp.df; d.quad = sp.quad;
        d.po = b->npts; d.np = nr * nt; d.nr = nr; d.ntheta = nt; d.npoly = np_;
        d.guess = sp.guess; d.mass = sp.mass; d.r0 = sp.r0; d.q = sp.q;
        d.ci = ci; d.si = si; d.poly = poly;
        d.rhl = co; d.vrotl = co + nl; d.sigRl = co + 2 * nl; d.sigpl = co + 3 * nl; d.sigzl = co + 4 * nl;
        d.sigRzl = co + 5 * nl;
        d.prep = prep;
        b->hmd.push_back(d);
        // points and their coordinates
        for (int n = 0; n < nr; n++)
            for (int g = 0; g < nt; g++) pmodel.push_back(m);
        bstart[0].push_back((int)0); bstart[1].push_back(0);
        rstart.push_back(b->rows);
        b->rows += nl;
        b->npts += nr * nt;
        b->npoly_max = np_ > b->npoly_max ? np_ : b->npoly_max;
        const size_t lt = tab_lds_doubles(nr, np_, NG) * sizeof(double);
        lds_max = lt > lds_max ? lt : lds_max;
        // persistent integrator: FP64 (tables + 20 doubles per lane) or FP32 (+ FP32 tables, 20 floats per lane)
        const size_t l64 = lt + 20 * 256 * sizeof(double);
        const size_t l32 = lt + (tab32_lds_floats(nr, np_) + 20 * 256) * sizeof(float);
        b->lds_integ = l64 > b->lds_integ ? l64 : b->lds_integ;
        b->lds_integ32 = l32 > b->lds_integ32 ? l32 : b->lds_integ32;
        const size_t lf = fast_lds_bytes(nr, np_, NG);
        b->lds_fast = lf > b->lds_fast ? lf : b->lds_fast;
    }
    if (r) { fj_batch_destroy(b); return r; }
    b->nodes = (long long)b->npts * NQ;
    if (b->fast) return batch_create_fast(b, spec, pmodel, rstart, lds_max, out);
    if (b->nodes > (64ll << 20)) {
        const long long n = b->nodes;
        fj_batch_destroy(b);
        return set_err(c, FJ_E_ARG, "batch of %lld nodes: at most %lld per batch (split the models)", n, 64ll << 20);
    }
    // block prefixes of the setup launch for block sizes 128 and 256
    for (int k = 0; k < 2; k++) {
        const int bs = k ? 256 : 128;
        int acc = 0;
        for (int m = 0; m < M; m++) {
            bstart[k][m] = acc;
            acc += (int)(((long long)b->hmd[m].np * NQ + bs - 1) / bs);
        }
        b->nblk[k] = acc;
    }
    rstart.push_back(b->rows);
    std::vector<double> R(b->npts), z(b->npts);
    for (int m = 0; m < M; m++) {
        const ModelDev &d = b->hmd[m];
        std::vector<double> hc(d.ntheta), hw(d.ntheta);
        host_gauleg(d.ntheta, 0, 1, hc.data(), hw.data());
        for (int n = 0; n < d.nr; n++)
            for (int g = 0; g < d.ntheta; g++) {
                R[d.po + n * d.ntheta + g] = spec[m].ar[n] * sqrt(1 - hc[g] * hc[g]);
                z[d.po + n * d.ntheta + g] = spec[m].ar[n] * hc[g];
            }
    }
    const long long nn = b->nodes;
    b->lds_tab = lds_max;
    if ((r = balloc(b, &b->d_md, M)) || (r = balloc(b, &b->d_pv, M)) || (r = balloc(b, &b->d_pmodel, b->npts)) ||
        (r = balloc(b, &b->d_bstart[0], M)) || (r = balloc(b, &b->d_bstart[1], M)) ||
        (r = balloc(b, &b->d_rstart, M + 1)) || (r = balloc(b, &b->d_R, b->npts)) || (r = balloc(b, &b->d_z, b->npts)) ||
        (r = balloc(b, &b->d_Phi, b->npts)) || (r = balloc(b, &b->d_Ve, b->npts)) || (r = balloc(b, &b->d_Vc, b->npts)) ||
        (r = balloc(b, &b->d_out6, (size_t)6 * b->npts)) || (r = balloc(b, &b->d_mom, (size_t)6 * b->npts)) ||
        (r = balloc(b, &b->d_nodef, (size_t)NF_COUNT * nn)) || (r = balloc(b, &b->d_Jr, (size_t)nn)) ||
        (r = balloc(b, &b->d_Jz, (size_t)nn)) || (r = balloc(b, &b->d_flag, (size_t)nn)) ||
        (r = balloc(b, &b->d_list_u, (size_t)nn)) || (r = balloc(b, &b->d_list_v, (size_t)nn)) ||
        (r = balloc(b, &b->d_ovf, (size_t)2 * nn)) || (r = balloc(b, &b->d_cnt, (size_t)8 * M + 8)) ||
        (r = balloc(b, &b->d_ctr, 1)) || (r = balloc(b, &b->d_backup, (size_t)3 * b->rows)) ||
        (r = balloc(b, &b->d_bkey, (size_t)2 * nn)) || (r = balloc(b, &b->d_bperm, (size_t)2 * nn))) {
        fj_batch_destroy(b);
        return r;
    }
    {
        // the sort's scratch; until an iteration has left records, kb_setup takes the identity order
        int mb = 0;
        while ((1 << mb) < M) mb++;
        b->order_bits = SETUP_KEY_BITS + mb;
        HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, b->btemp_bytes, b->d_bkey, b->d_bkey + nn, b->d_bperm,
                                                     b->d_bperm + nn, (int)nn, 0, b->order_bits, c->stream));
        if ((r = balloc(b, &b->d_btemp, b->btemp_bytes))) {
            fj_batch_destroy(b);
            return r;
        }
        NodeBuf none{nullptr, nn, nullptr, nullptr, nullptr};
        k_setup_key<<<(unsigned)((nn + 255) / 256), 256, 0, c->stream>>>(none, nn, nullptr, b->d_bperm + nn, nullptr);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipMemcpy(b->d_md, b->hmd.data(), M * sizeof(ModelDev), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_pmodel, pmodel.data(), pmodel.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_bstart[0], bstart[0].data(), M * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_bstart[1], bstart[1].data(), M * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_rstart, rstart.data(), (M + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_R, R.data(), R.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_z, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice));
    // legendre constants (index-only values) for this device
    const LegCoef lc = make_legcoef(MAXPOLY);
    HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_legcoef), &lc, sizeof lc));
    // persistent integrator: as many blocks as fit, at least one per model
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
    int pc64 = 0, pc32 = 0;
    HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc64, kb_integrals, 256, b->lds_integ));
    HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc32, kb_integrals_f32, 256, b->lds_integ32));
    b->per_cu64 = pc64 < 1 ? 1 : pc64;
    b->per_cu32 = pc32 < 1 ? 1 : pc32;
    const int per_cu = b->per_cu64 > b->per_cu32 ? b->per_cu64 : b->per_cu32;
    b->cus = prop.multiProcessorCount;
    b->integ_max_blocks = prop.multiProcessorCount * per_cu;
    if (b->integ_max_blocks < M) b->integ_max_blocks = M;
    b->integ_blocks = b->integ_max_blocks;
    const long long lanes = (long long)b->integ_max_blocks * 256, flanes = (long long)b->ovf_blocks * 256;
    if ((r = balloc(b, &b->d_block_model, (size_t)b->integ_max_blocks)) ||
        (r = balloc(b, &b->d_ws, ws_doubles(QCAP, lanes))) || (r = balloc(b, &b->d_wsfull, ws_doubles(QCAP_FULL, flanes)))) {
        fj_batch_destroy(b);
        return r;
    }
    b->ws_lanes = lanes;
    b->wsfull_lanes = flanes;
    b->st.nodes = nn;
    b->st.points = b->npts;
    b->st.setup_block = b->setup_bs;
    b->st.integ_blocks = b->integ_blocks;
    kb_potviews<<<(M + 63) / 64, 64, 0, c->stream>>>(b->d_md, M, b->d_pv);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = b;
    return 0;
}

int fj_batch_destroy(fj_batch *b) {
    if (!b) return 0;
    batch_drop_graph(b);
    for (void *p : b->allocs)
        if (p) (void)hipFree(p);
    delete b;
    return 0;
}

int fj_batch_setup(fj_batch *b) {
    if (!b) return FJ_E_ARG;
    fj_ctx *c = b->c;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->stream));
    kb_guess<<<(b->rows + 63) / 64, 64, 0, c->stream>>>(b->d_md, b->d_rstart, b->M);
    kb_poisson<<<dim3(b->npoly_max, b->M), PSCAN_THREADS, 0, c->stream>>>(b->d_md);
    kb_delta_prep<<<(b->M + 63) / 64, 64, 0, c->stream>>>(b->d_md, b->d_pv, b->M);
    const int NG = 30, lanes = b->M * NG;
    double *scr = nullptr;
    int *dst = nullptr;
    HIPCHK(c, hipMalloc((void **)&scr, (size_t)lanes * 400 * sizeof(double)));
    HIPCHK(c, hipMalloc((void **)&dst, sizeof(int)));
    HIPCHK(c, hipMemsetAsync(dst, 0, sizeof(int), c->stream));
    kb_delta<<<lanes, 64, 0, c->stream>>>(b->d_md, b->d_pv, b->M, scr, dst);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(e1, c->stream));
    int status = 0;
    HIPCHK(c, hipMemcpyAsync(&status, dst, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    b->st.ms_setup = elapsed(e0, e1);
    (void)hipFree(scr);
    (void)hipFree(dst);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (status & 8) return set_err(c, FJ_E_DELTA, "Delta fit: no bracket for ds2dD2 (delta/Delta.cpp:276-281)");
    b->have_setup = true;
    batch_drop_graph(b);
    return 0;
}

int fj_batch_autotune(fj_batch *b) {
    if (!b) return FJ_E_ARG;
    if (!b->have_setup) return set_err(b->c, FJ_E_STATE, "fj_batch_setup first");
    fj_ctx *c = b->c;
    if (b->fast) {   // one launch shape: nothing to tune
        b->st.ms_tune = 0;
        return 0;
    }
    int r;
    hipEvent_t e0, e1, t0, t1;
    HIPCHK(c, hipEventCreate(&e0)); HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventCreate(&t0)); HIPCHK(c, hipEventCreate(&t1));
    HIPCHK(c, hipEventRecord(t0, c->stream));
    if ((r = batch_save(b, false))) return r;
    const int per_cu_max = b->integ_max_blocks / (b->cus > 0 ? b->cus : 1);
    float best = 1e30f;
    int best_bs = b->setup_bs, best_ib = b->integ_blocks;
    for (int bs : {128, 256})
        for (int k = 1; k <= (per_cu_max > 0 ? per_cu_max : 1); k++) {
            int ib = k * b->cus;
            if (ib < b->M) ib = b->M;
            if (ib > b->integ_max_blocks) ib = b->integ_max_blocks;
            b->setup_bs = bs;
            b->integ_blocks = ib;
            if ((r = batch_save(b, true))) return r;
            HIPCHK(c, hipEventRecord(e0, c->stream));
            if ((r = batch_order(b))) return r;
            if ((r = batch_iteration(b))) return r;
            b->rec_ok = true;
            HIPCHK(c, hipEventRecord(e1, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            const float ms = elapsed(e0, e1);
            if (ms < best) { best = ms; best_bs = bs; best_ib = ib; }
        }
    b->setup_bs = best_bs;
    b->integ_blocks = best_ib;
    if ((r = batch_save(b, true))) return r;
    HIPCHK(c, hipEventRecord(t1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    b->st.ms_tune = elapsed(t0, t1);
    b->st.setup_block = best_bs;
    b->st.integ_blocks = best_ib;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(t0); (void)hipEventDestroy(t1);
    batch_drop_graph(b);
    return 0;
}

int fj_batch_iterate(fj_batch *b, int niter) {
    if (!b || niter < 0) return FJ_E_ARG;
    if (!b->have_setup) return set_err(b->c, FJ_E_STATE, "fj_batch_setup first");
    fj_ctx *c = b->c;
    int r;
    if (b->fast != (c->prec == FJ_PREC_FAST))
        return set_err(c, FJ_E_STATE, b->fast ? "batch created under FJ_PREC_FAST: iterate it under FJ_PREC_FAST"
                                              : "batch created under FJ_PREC_FP64 / FJ_PREC_MIXED: "
                                                "FJ_PREC_FAST needs a batch created under it");
    if (b->fast) {
        // launched directly, without a graph: the lane order's sort runs between the models' launches
        if ((r = fast_gl_ensure(c))) return r;
        b->fast_ngl = c->fast_ngl;
        b->st.graph = 0;
        hipEvent_t e0, e1;
        HIPCHK(c, hipEventCreate(&e0));
        HIPCHK(c, hipEventCreate(&e1));
        for (int it = 0; it < niter; it++) {
            HIPCHK(c, hipEventRecord(e0, c->stream));
            if ((r = batch_iteration(b))) return r;
            HIPCHK(c, hipEventRecord(e1, c->stream));
            b->fkeys_ok = true;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (niter) b->st.ms_last_iter = elapsed(e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return 0;
    }
    if (!b->gexec) {
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        r = batch_iteration(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r) { if (g) (void)hipGraphDestroy(g); return r; }
        if (e != hipSuccess) return set_err(c, FJ_E_HIP, "graph capture: %s", hipGetErrorString(e));
        b->graph = g;
        HIPCHK(c, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->st.graph = 1;
    }
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    for (int it = 0; it < niter; it++) {
        HIPCHK(c, hipEventRecord(e0, c->stream));
        if ((r = batch_order(b))) return r;
        HIPCHK(c, hipGraphLaunch(b->gexec, c->stream));
        HIPCHK(c, hipEventRecord(e1, c->stream));
        b->rec_ok = true;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (niter) b->st.ms_last_iter = elapsed(e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int fj_batch_get(fj_batch *b, int m, double *rhl, double *sigRl, double *sigpl, double *sigzl, double *vrotl,
                 double *phil, double *Pr, double *Pr2, double *Egrid, double *Dgrid) {
    if (!b || m < 0 || m >= b->M) return FJ_E_ARG;
    fj_ctx *c = b->c;
    const ModelDev &d = b->hmd[m];
    const size_t n = (size_t)d.nr * d.npoly * sizeof(double);
    const double *src[10] = {d.rhl, d.sigRl, d.sigpl, d.sigzl, d.vrotl, d.ta.phil, d.ta.Pr, d.Pr2, d.ta.E, d.ta.D};
    double *dst[10] = {rhl, sigRl, sigpl, sigzl, vrotl, phil, Pr, Pr2, Egrid, Dgrid};
    for (int k = 0; k < 10; k++)
        if (dst[k]) HIPCHK(c, hipMemcpy(dst[k], src[k], k < 8 ? n : 30 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int fj_batch_set(fj_batch *b, int m, const double *phil, const double *Pr, const double *Pr2, const double *Egrid,
                 const double *Dgrid) {
    if (!b || m < 0 || m >= b->M) return FJ_E_ARG;
    fj_ctx *c = b->c;
    const ModelDev &d = b->hmd[m];
    const size_t n = (size_t)d.nr * d.npoly * sizeof(double);
    // the model's own device tables (TabArgs holds them read-only for the kernels)
    double *dst[5] = {const_cast<double *>(d.ta.phil), const_cast<double *>(d.ta.Pr), d.Pr2,
                      const_cast<double *>(d.ta.E), const_cast<double *>(d.ta.D)};
    const double *src[5] = {phil, Pr, Pr2, Egrid, Dgrid};
    for (int k = 0; k < 5; k++)
        if (src[k]) HIPCHK(c, hipMemcpy(dst[k], src[k], k < 3 ? n : 30 * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int fj_batch_get_stats(const fj_batch *b, fj_batch_stats *out) {
    if (!b || !out) return FJ_E_ARG;
    *out = b->st;
    return 0;
}

int fj_get_stats(const fj_ctx *c, fj_stats *out) {
    if (!c || !out) return FJ_E_ARG;
    *out = c->st;
    return 0;
}

}  // extern "C"
