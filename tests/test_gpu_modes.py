"""The two non-reference modes of the library against the FP64 oracle:

* FJ_POISSON_BLOCKSCAN: the radial integrals of computeInts
  (pot/Potential.cpp:83-155) taken exactly for the piecewise-linear density
  they integrate; checked against the closed-form segment integrals to
  1e-12, and its deviation from the QUADPACK path (epsrel 1e-3) reported.
* FJ_PREC_MIXED: FP32 action integrands and DF, FP64 roots, Delta and
  Poisson; each output coefficient's error against the FP64 oracle, scaled by
  the monopole at that radius (as mergePhi scales, UtilsLeg.h:266-278),
  must stay below 1e-4 for the density and 1e-5 for the potential.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def exact_tables(ar, rhl):
    """phil, Pr, Pr2 from exact integrals of the piecewise-linear rho_l."""
    nr, npoly = rhl.shape
    out = [np.zeros((nr, npoly)) for _ in range(3)]
    for np_ in range(npoly):
        l = 2 * np_
        y = rhl[:, np_]
        Iint = np.zeros(nr)
        Iext = np.zeros(nr)
        acc = y[0] * ar[0] ** (l + 3) / (l + 3)
        Iint[0] = acc
        for i in range(1, nr):
            x0, x1, y0, y1 = ar[i - 1], ar[i], y[i - 1], y[i]
            s = (y1 - y0) / (x1 - x0)
            c0 = y0 - s * x0
            acc += c0 * (x1 ** (l + 3) - x0 ** (l + 3)) / (l + 3) + s * (x1 ** (l + 4) - x0 ** (l + 4)) / (l + 4)
            Iint[i] = acc
        acc = max(0.0, y[-1]) * ar[-1] ** (2 - l) / (2 + l)
        Iext[-1] = acc
        for i in range(nr - 2, -1, -1):
            x0, x1, y0, y1 = ar[i], ar[i + 1], y[i], y[i + 1]
            s = (y1 - y0) / (x1 - x0)
            c0 = y0 - s * x0
            e = 2 - l
            if e == 0:
                acc += c0 * np.log(x1 / x0) + s * (x1 - x0)
            else:
                acc += c0 * (x1 ** e - x0 ** e) / e + s * (x1 ** (e + 1) - x0 ** (e + 1)) / (e + 1)
            Iext[i] = acc
        r = ar
        out[0][:, np_] = -2 * np.pi * (Iint / r ** (l + 1) + Iext * r ** l)
        out[1][:, np_] = -2 * np.pi * (-(l + 1) * Iint / r ** (l + 2) + l * Iext * r ** (l - 1.0))
        out[2][:, np_] = -2 * np.pi * ((l + 2) * (l + 1) * Iint / r ** (l + 3) + (l - 1) * l * Iext * r ** (l - 2.0)
                                       - (2 * l + 1) * y)
    return out


def scaled_err(a, b):
    """max over (n, l) of |a - b| / |b[n][0]|"""
    return float(np.max(np.abs(a - b) / np.abs(b[:, :1])))


@pytest.mark.parametrize("nr,npoly,seed", [(20, 3, 3), (96, 7, 5), (300, 12, 7)])
def test_poisson_blockscan_exact(ctx, nr, npoly, seed):
    import fjmodels_amd as fj
    rng = np.random.default_rng(seed)
    ar = O.log_grid(nr, 1e-3 if nr > 20 else 0.05, 1e3 if nr > 20 else 30.0)
    rhl = np.abs(rng.normal(size=(nr, npoly))) * (1 + 1 / ar[:, None]) + 0.1
    ctx.set_grid(ar, 8, npoly)
    got = ctx.poisson(rhl, fj.FJ_POISSON_BLOCKSCAN)
    ref = exact_tables(ar, rhl)
    # both sides round; the derivative tables cancel terms of size ~l and ~l^2
    # times the result, so their bounds grow with the highest order L = 2 lmax + 1
    L = 2 * (2 * (npoly - 1)) + 1
    tol = [1e-13, 1e-13 * L ** 2, 1e-13 * L ** 3]
    for k, (g, r) in enumerate(zip(got, ref)):
        err = scaled_err(g, r)
        print(nr, npoly, ["phil", "Pr", "Pr2"][k], f"{err:.2e}", "tol", f"{tol[k]:.1e}")
        assert err < tol[k], (k, err)


def test_poisson_blockscan_vs_quadpack(ctx):
    """Deviation of the exact scan from QUADPACK at epsrel 1e-3 on the
    Hernquist guess density of the benchmark grid."""
    import fjmodels_amd as fj
    p, rhl = O.guess_model(40, 0.01, 100.0, 5, 20, delta=False)
    ctx.set_grid(p.ar, 20, 5)
    q = ctx.poisson(rhl, fj.FJ_POISSON_QUADPACK)
    b = ctx.poisson(rhl, fj.FJ_POISSON_BLOCKSCAN)
    errs = [scaled_err(x, y) for x, y in zip(b, q)]
    print("blockscan vs quadpack (phil, Pr, Pr2):", ["%.2e" % e for e in errs])
    assert errs[0] < 5e-3 and errs[1] < 5e-2


def test_mixed_precision_coefficients(ctx):
    """One computeNewPhi in FJ_PREC_MIXED vs the oracle's FP64 one, from the
    same potential and Delta table, on a 16 x 6 grid (l <= 4)."""
    import fjmodels_amd as fj
    p, _ = O.guess_model(16, 0.01, 100.0, 3, 6)
    ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
    ctx.set_potential(*p.tables())
    ctx.set_delta(*p.delta())
    ctx.set_df(fj.df_params("Hernquist"))
    ctx.set_quad(fj.FJ_QUAD_LOG22)
    ctx.set_precision(fj.FJ_PREC_MIXED)
    try:
        g = ctx.compute_new_phi()
    finally:
        ctx.set_precision(fj.FJ_PREC_FP64)
    r = p.compute_new_phi(O.df_params("Hernquist"), nthreads=16)
    errs = {k: scaled_err(g[k], r[k]) for k in ("rhl", "vrotl", "sigRl", "sigpl", "sigzl", "phil", "Pr", "Pr2")}
    print("mixed vs FP64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert errs["rhl"] < 1e-4 and errs["phil"] < 1e-5 and errs["Pr"] < 1e-4
    assert errs["sigRl"] < 1e-4 and errs["sigzl"] < 1e-4
