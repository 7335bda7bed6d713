"""The rotating-NFW and two-component model configurations at their full
sizes, where the CPU oracle cannot run a whole iteration in seconds:

* rotating NFW f(J) (chi = 1, README form (1-k) f + k tanh(chi Jphi/J0) f,
  k = 0.5; df/DF.cpp:116, README.md:80-83), h(J) deltas 0.55, on 96 log radii
  in [1e-3, 1e3] x 32 Gauss-Legendre cos(theta) nodes, l <= 12 (NPOLY 7),
  22^3 velocity nodes per point, NFW guess potential and its Delta table;
* Hernquist bulge (M = 1, J0 = 1) + NFW halo (M = 10, J0 = 5) on 128 x 32,
  NPOLY 7, the eight h/g deltas drawn from numpy's default_rng(0) in
  [0.4, 1.2] in the order (hphi, hz, gphi, gz) per component; 43.6 M nodes
  per component, more than one launch holds (32 M), so the grid runs in two
  chunks.

Checks, each independent of the grid size:
* per-point moments at sampled grid points (both chunks) equal the oracle's
  rhofDF bit for bit;
* a full iteration is deterministic (two runs, same bits);
* the rotation term is odd in Jphi and the velocity rule is symmetric in
  vphi up to the rule's 1e-6 offset, so the README form at k = 0.5 halves
  rho_l and leaves the dispersions unchanged, to ~1e-6;
* the mixed-precision node phase stays within a stated bound of the FP64
  path, per output coefficient.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEYS = ("rhl", "vrotl", "sigRl", "sigpl", "sigzl", "sigRzl", "phil", "Pr", "Pr2")


def node_coords(p, idx):
    """(R, z) of grid point idx = n * ntheta + g, with the oracle's cos(theta)
    nodes (gauleg on [0, 1])."""
    ci, _ = O.gauleg(p.ngauss, 0.0, 1.0)
    n, g = divmod(idx, p.ngauss)
    s = np.sqrt(1 - ci[g] * ci[g])
    return p.ar[n] * s, p.ar[n] * ci[g]


def install(ctx, p, tables, df):
    import fjmodels_amd as fj
    ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
    ctx.set_quad(fj.FJ_QUAD_LOG22)
    ctx.set_precision(fj.FJ_PREC_FP64)
    ctx.set_potential(*tables)
    ctx.set_delta(*p.delta())
    ctx.set_df(df)


def check_points(ctx, p, dfo, idxs):
    rho, vrot, sigR, sigp, sigz, sigRz = [x.ravel() for x in ctx.moments()]
    for i in idxs:
        R, z = node_coords(p, i)
        r_rho, _, r_rest = p.rhofDF(dfo, R, z)
        got = (rho[i], vrot[i], sigR[i], sigp[i], sigz[i], sigRz[i])
        ref = (r_rho,) + tuple(r_rest)
        print("point", i, "rho", rho[i], r_rho)
        assert all(a == b for a, b in zip(got, ref)), (i, got, ref)


@pytest.fixture(scope="module")
def nfw96():
    p, rhl = O.guess_model(96, 1e-3, 1e3, 7, 32, model="NFW")
    return p


def test_rotating_nfw_full_grid(ctx, nfw96):
    import fjmodels_amd as fj
    p = nfw96
    dfo = O.df_params("NFW", chi=1.0, rot_form=1)
    install(ctx, p, p.tables(), fj.df_params("NFW", chi=1.0, rot_form=1))
    npts = p.nr * p.ngauss
    check_points(ctx, p, dfo, [0, 5 * 32 + 7, npts // 2 + 3, npts - 33, npts - 1])


def test_rotating_nfw_full_iteration(ctx, nfw96):
    import fjmodels_amd as fj
    p = nfw96
    install(ctx, p, p.tables(), fj.df_params("NFW", chi=1.0, rot_form=1))
    a = ctx.compute_new_phi()
    b = ctx.compute_new_phi()
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    st = ctx.stats()
    print("nodes", st["nodes"], "bound", st["bound_nodes"], "ms", st["ms_total"])
    assert np.all(np.isfinite(a["phil"])) and np.max(a["vrotl"][:, 0]) > 0

    # chi = 0: the even part alone
    ctx.set_df(fj.df_params("NFW", chi=0.0))
    e = ctx.compute_new_phi()
    mono = np.abs(e["rhl"][:, :1])
    d_rho = np.max(np.abs(a["rhl"] - 0.5 * e["rhl"]) / (0.5 * mono))
    print("rho_l(rot) vs rho_l(even)/2:", d_rho)
    assert d_rho < 1e-5
    for k in ("sigRl", "sigpl", "sigzl"):
        d = np.max(np.abs(a[k] - e[k]) / np.abs(e[k][:, :1]))
        print(k, "rot vs even:", d)
        assert d < 1e-5, k
    assert np.max(np.abs(e["vrotl"])) < 1e-5 * np.max(np.abs(a["vrotl"]))


def test_rotating_nfw_mixed_precision(ctx, nfw96):
    """Per-coefficient error of the mixed node phase (FP32 integrands and DF,
    FP64 roots and Delta) against the FP64 path, both with the QUADPACK
    Poisson solve; errors scaled by the monopole at each radius.

    The mean rotation is the small odd part of the DF near the centre
    (tanh(chi R vphi / J0) ~ 1e-4 at r = 1e-3), a difference of the prograde
    and retrograde halves, so its FP32 error there is the largest: about
    1e-3 of its monopole, against 3e-6 for rho_l and 1e-6 for the
    dispersions."""
    import fjmodels_amd as fj
    p = nfw96
    install(ctx, p, p.tables(), fj.df_params("NFW", chi=1.0, rot_form=1))
    ref = ctx.compute_new_phi()
    ctx.set_precision(fj.FJ_PREC_MIXED)
    try:
        got = ctx.compute_new_phi()
    finally:
        ctx.set_precision(fj.FJ_PREC_FP64)
    bounds = {"rhl": 1e-5, "sigRl": 1e-5, "sigpl": 1e-5, "sigzl": 1e-5, "vrotl": 2e-3, "phil": 1e-6}
    errs = {}
    for k in bounds:
        er = np.max(np.abs(got[k] - ref[k]) / np.abs(ref[k][:, :1]), axis=1)
        worst = np.argsort(er)[::-1][:4]
        print("mixed vs fp64", k, er.max(), "worst radii", [(float("%.3g" % p.ar[i]), float("%.3g" % er[i]))
                                                             for i in worst])
        errs[k] = er.max()
    for k, bound in bounds.items():
        assert errs[k] < bound, (k, errs[k])


def test_two_components_full_grid(ctx):
    import fjmodels_amd as fj
    rng = np.random.default_rng(0)
    d = rng.uniform(0.4, 1.2, 8)
    ar = O.log_grid(128, 1e-3, 1e3)
    p = O.Pot(ar, 7, 32)
    r1 = p.guess_rhl("Hernquist", 1.0, 1.0, 1.0)
    r2 = p.guess_rhl("NFW", 10.0, 2.5, 1.0)
    tot = [a + b for a, b in zip(p.compute_phil(r1), p.compute_phil(r2))]
    p.set_tables(*tot)
    assert p.tabulate_delta() == 0
    comps = [("Hernquist", dict(mass=1.0, r0=1.0, dphi_h=d[0], dz_h=d[1], dphi_g=d[2], dz_g=d[3])),
             ("NFW", dict(mass=10.0, r0=2.5, dphi_h=d[4], dz_h=d[5], dphi_g=d[6], dz_g=d[7]))]
    npts = p.nr * p.ngauss
    per_launch = (32 << 20) // (22 ** 3)
    assert npts > per_launch            # the grid takes two chunks
    for name, kw in comps:
        install(ctx, p, tot, fj.df_params(name, **kw))
        check_points(ctx, p, O.df_params(name, **kw),
                     [1, per_launch - 1, per_launch, per_launch + 1, npts - 1])


def test_two_component_driver_converges(tmp_path):
    """The fjmodels driver on the two-component model at full size, with the
    parameter file shipped beside the driver (mixed node phase, block-scan
    Poisson): the potential change (mergePhi's distance) contracts by about
    0.45 per iteration and falls below the file's tolerance, 1e-5, within its
    15 iterations.  With the reference's QUADPACK Poisson integrals (relative
    tolerance 1e-3) the same run stalls at a change of about 5e-4."""
    import os
    import re
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "fjmodels_amd", "fjmodels")
    par = os.path.join(root, "fjmodels_amd", "params", "two_component_128x32.txt")
    r = subprocess.run([exe, par, "--outdir", str(tmp_path)], capture_output=True, text=True, timeout=200)
    print(r.stdout[-1500:], r.stderr[-500:])
    assert r.returncode == 0
    d = [float(x) for x in re.findall(r"potential change ([0-9.eE+-]+)", r.stdout)]
    assert len(d) <= 15 and "converged (tolerance 1e-05)" in r.stdout
    assert d[-1] < 1e-5
    # geometric contraction after the first iterations
    assert all(b < 0.7 * a for a, b in zip(d[3:], d[4:]))
    assert (tmp_path / ("bulge_%d.out" % (len(d) - 1))).exists()
    assert (tmp_path / ("halo_%d.out" % (len(d) - 1))).exists()
