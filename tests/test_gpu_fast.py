"""The fast node phase (FJ_PREC_FAST) through the C ABI.

Its actions are the Staeckel fudge of the reference computed another way
(fj_fast.hpp: u-side reference point at the node, the reference's bracket
walks with Anderson-Bjorck roots, Gauss-Legendre action integrals in FP32),
so the checks are:
* per-point moments against the host build of the same algorithm with FP64
  integrands (tests/hostbuild, fjh_fast_rhofDF): the FP32 rounding alone,
  relative 2e-5 on the density and dispersions;
* against the oracle (QAGS, the reference algorithm): within the 64-point
  rule's quadrature error (and QAGS's own, at epsrel 1e-3), relative 5e-4
  on the density;
* the bound / unbound node split equals the FP64 path's;
* a full node phase and Legendre projection on the 40 x 20 grid against the
  FP64 path, each coefficient's difference divided by the monopole at its
  radius below 3e-4.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HB = os.path.join(ROOT, "tests", "hostbuild")
dp = C.POINTER(C.c_double)
X11, W11, _, _ = O.quad_tables()


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)


@pytest.fixture(scope="module")
def model():
    return O.guess_model(40, 0.01, 100.0, 5, 20)[0]


@pytest.fixture(scope="module")
def fctx(ctx, model):
    import fjmodels_amd as fj
    p = model
    ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
    ctx.set_quad(fj.FJ_QUAD_LOG22)
    ctx.set_potential(*p.tables())
    ctx.set_delta(*p.delta())
    ctx.set_df(fj.df_params("Hernquist"))
    yield ctx
    ctx.set_precision(fj.FJ_PREC_FP64)


PTS_R = np.array([0.011, 0.5, 1.3, 4.0, 20.0, 70.0])
PTS_Z = np.array([0.004, 0.9, 0.2, 3.0, 1.0, 50.0])


def host_fast(p, R, z, ngl=64):
    subprocess.run(["make", "-s", "-C", HB], check=True)
    lib = C.CDLL(os.path.join(HB, "libfjhost.so"))
    lib.fjh_fast_rhofDF.argtypes = [C.c_int, dp, C.c_int, dp, dp, C.c_int, dp, dp, dp, dp, dp, dp, C.c_double,
                                    C.c_double] + [C.c_double] * 9 + [C.c_double, C.c_int, C.c_int, dp, dp,
                                                                      C.c_int, dp, dp, dp]
    phil, Pr, _ = p.tables()
    E, D = p.delta()
    gx, gw = np.polynomial.legendre.leggauss(ngl)
    gx, gw = 0.5 * (gx + 1), 0.5 * gw
    nv = max(4, ngl // 2)   # the J_z rule of the fast phase: half the order
    vx, vw = np.polynomial.legendre.leggauss(nv)
    vx, vw = 0.5 * (vx + 1), 0.5 * vw
    d = O.df_params("Hernquist")
    out = np.zeros((len(R), 6))
    for i in range(len(R)):
        o = np.zeros(6)
        lib.fjh_fast_rhofDF(p.nr, P(p.ar), p.npoly, P(phil), P(Pr), p.ngrid, P(E), P(D), P(X11), P(W11), P(X11),
                            P(W11), R[i], z[i], d.A, d.B, d.dphi_h, d.dz_h, d.dphi_g, d.dz_g, d.mass, d.J0, d.chi,
                            d.k, d.rot_form, ngl, P(gx), P(gw), nv, P(vx), P(vw), P(o))
        out[i] = o
    return out


def test_fast_points_match_fp64_integrands(fctx, model):
    import fjmodels_amd as fj
    fctx.set_precision(fj.FJ_PREC_FAST)
    g6, _ = fctx.rhofDF(PTS_R, PTS_Z)
    h6 = host_fast(model, PTS_R, PTS_Z)
    rel = np.abs(g6 - h6) / np.abs(h6[:, :1])
    print("GPU FP32 vs host FP64 integrands, per sum / density:", np.max(rel, axis=0))
    for k in (0, 2, 3, 4):
        assert np.max(rel[:, k] / np.abs(h6[:, k] / h6[:, 0])) < 2e-5, k


def test_fast_points_near_oracle(fctx, model):
    import fjmodels_amd as fj
    fctx.set_precision(fj.FJ_PREC_FAST)
    g6, gm = fctx.rhofDF(PTS_R, PTS_Z)
    st = fctx.stats()
    dfo = O.df_params("Hernquist")
    for i in range(len(PTS_R)):
        _, ref, _ = model.rhofDF(dfo, PTS_R[i], PTS_Z[i])
        rel = np.abs(g6[i] - ref) / ref[0]
        print(i, "rel to density:", rel)
        assert rel[0] < 5e-4 and np.max(rel[[2, 3, 4]] / (ref[[2, 3, 4]] / ref[0])) < 1e-3
    fctx.set_precision(fj.FJ_PREC_FP64)
    fctx.rhofDF(PTS_R, PTS_Z)
    s64 = fctx.stats()
    assert st["bound_nodes"] == s64["bound_nodes"] and st["unbound_nodes"] == s64["unbound_nodes"]
    assert st["nonfinite_nodes"] == 0


def test_fast_grid_phase_against_fp64(fctx):
    import fjmodels_amd as fj
    fctx.set_precision(fj.FJ_PREC_FP64)
    ref = fctx.compute_new_phi(fj.FJ_POISSON_QUADPACK)
    fctx.set_precision(fj.FJ_PREC_FAST)
    got = fctx.compute_new_phi(fj.FJ_POISSON_QUADPACK)
    st = fctx.stats()
    print("fast node phase ms", st["ms_total"])
    for k in ("rhl", "sigRl", "sigpl", "sigzl"):
        err = np.max(np.abs(got[k] - ref[k]) / np.abs(ref[k][:, :1]))
        print(k, "max |diff| / monopole", err)
        assert err < 3e-4, k


def test_lane_order_keeps_results(ctx):
    """From the second node phase on, the node kernels take their nodes in
    the order of the previous phase's turning points (k_setup, and
    k_fast_nodes under FJ_PREC_FAST).  Each node's arithmetic is its own and
    the moments are summed per point in node order, so two phases from the
    same installed potential, the first on a grid of a new size (the tiled
    order: no turning points of these nodes yet) and the second in the
    sorted order, give the same tables bit for bit, in FP64 and in the fast
    phase."""
    import fjmodels_amd as fj
    p = O.guess_model(26, 0.01, 100.0, 5, 9)[0]
    ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
    ctx.set_quad(fj.FJ_QUAD_LOG22)
    ctx.set_potential(*p.tables())
    ctx.set_delta(*p.delta())
    ctx.set_df(fj.df_params("Hernquist"))
    try:
        for prec in (fj.FJ_PREC_FP64, fj.FJ_PREC_FAST):
            ctx.set_precision(prec)
            first = ctx.compute_new_phi(fj.FJ_POISSON_BLOCKSCAN)
            second = ctx.compute_new_phi(fj.FJ_POISSON_BLOCKSCAN)
            print("precision", prec, "setup ms", ctx.stats()["ms_setup"])
            for k in first:
                assert np.array_equal(first[k], second[k]), (prec, k)
    finally:
        ctx.set_precision(fj.FJ_PREC_FP64)
