"""Models of the seeded 256-model sweep draw (fjmodels_amd.draw_models(256,
seed=0): Hernquist / isochrone / NFW guess potentials and DPL f(J) with the
README rotation form, drawn h(J) deltas and chi, l_max 8-12) through the
batched path, against the oracle's FP64 computeNewPhi (pot/potLeg.cpp:76-82)
from the same potential and Delta table.

The drawn grids (24-96 radii x 12-32 cos(theta)) are cut to 14 x 4 so that
the oracle runs in seconds; radii span, l_max, model and DF are the drawn
ones.  Every model starts from the oracle's guess potential (QUADPACK
Poisson, pot/Potential.cpp:110-200) and Delta table, installed with
fj_batch_set; one batched iteration then gives the density and dispersion
coefficients of that potential:
* FP64 node phase: bit for bit the oracle's;
* mixed precision (FP32 integrands, Kronrod rules, QAGS bookkeeping and DF):
  every coefficient within 1e-4 of the monopole at its radius (the mean
  rotation's within 1e-4 of the azimuthal dispersion's);
* the batch's potential update is the block-scan Poisson solve (exact
  integrals of the piecewise-linear rho_l), the reference's is QUADPACK
  (epsrel 1e-3) on its own interpolation of rho_l: on 14 radii they differ
  by up to 6.4e-3 of the monopole (isochrone), so Phi_l is held to 1e-2.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

MOMENTS = ("rhl", "sigRl", "sigpl", "sigzl", "vrotl")


def cut_models():
    import fjmodels_amd as fj
    draw = fj.draw_models(256, seed=0)
    picked, seen = [], set()
    for k, m in enumerate(draw):           # the first model of each kind in the draw
        if m["model"] not in seen:
            seen.add(m["model"])
            picked.append(k)
    out = []
    for k in picked:
        m = dict(draw[k])
        m["ar"] = fj.log_grid(14, float(m["ar"][0]), float(m["ar"][-1]))
        m["ntheta"] = 4
        m["draw_index"] = k
        out.append(m)
    return out


@pytest.fixture(scope="module")
def cfg4_oracle():
    ms = cut_models()
    refs = []
    for m in ms:
        p = O.Pot(m["ar"], m["npoly"], m["ntheta"])
        rhl = p.guess_rhl(m["model"], 1.0, 1.0, 1.0)
        tabs = p.compute_phil(rhl)
        p.set_tables(*tabs)
        assert p.tabulate_delta() == 0
        df = O.DFParams(*[getattr(m["df"], f) for f, _ in O.DFParams._fields_])
        ref = p.compute_new_phi(df, nthreads=16)
        refs.append((tabs, p.delta(), ref))
        print("oracle model", m["draw_index"], m["model"], m["npoly"], round(m["df"].chi, 3), flush=True)
    return ms, refs


def run_batch(ctx, ms, refs, precision):
    import fjmodels_amd as fj
    ctx.set_precision(precision)
    try:
        b = fj.Batch(ctx, ms)
        b.setup()
        for k, (tabs, (E, D), _) in enumerate(refs):
            b.set(k, *tabs, E, D)
        b.iterate(1)
        out = [b.get(k) for k in range(len(ms))]
        b.close()
    finally:
        ctx.set_precision(fj.FJ_PREC_FP64)
    return out


def test_cfg4_fp64_batch_matches_oracle(ctx, cfg4_oracle):
    import fjmodels_amd as fj
    ms, refs = cfg4_oracle
    got = run_batch(ctx, ms, refs, fj.FJ_PREC_FP64)
    for k, (g, (_, _, ref)) in enumerate(zip(got, refs)):
        for n in MOMENTS:
            assert np.array_equal(g[n], ref[n]), (ms[k]["draw_index"], n, np.max(np.abs(g[n] - ref[n])))
        dphi = np.max(np.abs(g["phil"] - ref["phil"]) / np.abs(ref["phil"][:, :1]))
        print("model", ms[k]["draw_index"], "block-scan vs QUADPACK Phi_l:", dphi)
        assert dphi < 1e-2


def test_cfg4_mixed_batch_near_oracle(ctx, cfg4_oracle):
    import fjmodels_amd as fj
    ms, refs = cfg4_oracle
    got = run_batch(ctx, ms, refs, fj.FJ_PREC_MIXED)
    for k, (g, (_, _, ref)) in enumerate(zip(got, refs)):
        for n in MOMENTS:
            # each coefficient over its own monopole; the mean rotation over the
            # azimuthal dispersion's (the rotation's monopole vanishes with chi)
            scale = np.abs(ref["sigpl" if n == "vrotl" else n][:, :1])
            err = np.max(np.abs(g[n] - ref[n]) / scale)
            print("model", ms[k]["draw_index"], n, "mixed vs FP64 oracle:", err)
            assert err < 1e-4, (ms[k]["draw_index"], n, err)
