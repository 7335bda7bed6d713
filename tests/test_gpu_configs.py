"""The other model configurations against the oracle, at sizes the CPU
restatement runs in seconds (FP64, bit for bit):

* rotating NFW f(J) with the documented (1-k) f + k tanh(chi Jphi/J0) f term
  (df/DF.cpp:116, README.md:80-83; k = 0.5, chi = 1), and the code's
  f (1 + k tanh) form, on an NFW guess potential, tanh-sinh and log rules;
* two components through the C ABI in the total-potential flow (the driver's
  comp2_potential total): both components' computeNewPhi read the total
  potential of the previous iteration, and updatePhil (potLeg.cpp:21-29) sums
  them.  The reference's own flow, in which component 2 reads component 1's
  new potential (pot/Potential.cpp:182-188, fJmodels.cpp:69-73), is checked
  through the driver in test_gpu_trajectory.py.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def install(ctx, p, df, quad):
    ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
    ctx.set_quad(quad)
    ctx.set_potential(*p.tables())
    ctx.set_delta(*p.delta())
    ctx.set_df(df)


@pytest.mark.parametrize("rot_form,quad", [(1, 0), (0, 0), (1, 1)])
def test_rotating_nfw(ctx, rot_form, quad):
    import fjmodels_amd as fj
    p = O.guess_model(14, 0.001, 1000.0, 4, 5, model="NFW")[0]
    p.set_quad(quad)
    dfo = O.df_params("NFW", chi=1.0, rot_form=rot_form)
    install(ctx, p, fj.df_params("NFW", chi=1.0, rot_form=rot_form), quad)
    g = ctx.compute_new_phi()
    r = p.compute_new_phi(dfo, nthreads=16)
    for k in ("rhl", "vrotl", "sigRl", "sigpl", "sigzl", "sigRzl", "phil", "Pr", "Pr2"):
        assert np.array_equal(g[k], r[k]), (k, np.max(np.abs(g[k] - r[k])))
    assert np.max(np.abs(g["vrotl"][:, 0])) > 0      # the rotation term is live


def test_two_components(ctx):
    import fjmodels_amd as fj
    ar = O.log_grid(13, 0.01, 100.0)
    npoly, nth = 3, 4
    p = O.Pot(ar, npoly, nth)
    # guess potentials of both components and their total (fJmodels.cpp:47-58)
    r1 = p.guess_rhl("Hernquist", 1.0, 1.0, 1.0)
    r2 = p.guess_rhl("NFW", 10.0, 2.5, 1.0)
    t1, t2 = p.compute_phil(r1), p.compute_phil(r2)
    tot = [a + b for a, b in zip(t1, t2)]
    p.set_tables(*tot)
    assert p.tabulate_delta() == 0
    dfs = [("Hernquist", dict(mass=1.0, r0=1.0, dphi_h=0.5, dz_h=0.9)),
           ("NFW", dict(mass=10.0, r0=2.5, dphi_h=1.1, dz_h=0.6, dphi_g=0.7, dz_g=1.2))]
    ctx.set_grid(ar, nth, npoly, 30)
    ctx.set_quad(fj.FJ_QUAD_LOG22)
    ctx.set_delta(*p.delta())
    for it in range(2):
        ctx.set_potential(*tot)
        outs_g, outs_o = [], []
        for name, kw in dfs:
            ctx.set_df(fj.df_params(name, **kw))
            outs_g.append(ctx.compute_new_phi())
            outs_o.append(p.compute_new_phi(O.df_params(name, **kw), nthreads=16))
        for g, o in zip(outs_g, outs_o):
            for k in ("rhl", "sigRl", "sigzl", "phil", "Pr", "Pr2"):
                assert np.array_equal(g[k], o[k]), (it, k)
        tot = [outs_o[0][k] + outs_o[1][k] for k in ("phil", "Pr", "Pr2")]
        p.set_tables(*tot)
