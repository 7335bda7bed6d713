"""Action surrogate on the GPU.

* the BF16-MFMA network's forward pass and its gradient against the numpy
  reference of tests/mlp_ref.py (itself pinned by finite differences), with
  bf16 rounding emulated at the same places;
* training on Staeckel actions of a small model, and a node phase served by
  the network (fallback and judge sample through the Staeckel fudge): what it
  gains (node-phase time) and what it costs (judge errors, density
  coefficients against the Staeckel node phase from the same potential).
  After 61 epochs on ~0.9 M nodes the median action error is a few per cent
  of Jr + Jz + |Jphi| and the density coefficients move by up to ~10 %: far
  from the 1e-4 the plain path meets, which is why the final iteration of a
  surrogate run always uses the Staeckel fudge.
"""
import time

import numpy as np
import pytest

import mlp_ref as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sctx(ctx):
    ctx.surrogate_init(seed=5, capacity=1 << 20)
    return ctx


def test_forward_matches_reference(sctx):
    p = R.glorot(3).astype(np.float32)
    sctx.surrogate_params(p)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(3001, R.IN)).astype(np.float32)
    y = sctx.surrogate_forward(x)
    ref, _, _ = R.forward(p.astype(np.float64), x.astype(np.float64), emulate=True)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    print("forward max |dy| / max |y|:", err)
    assert err < 1e-2


def test_gradient_matches_reference(sctx):
    p = R.glorot(6).astype(np.float32)
    sctx.surrogate_params(p)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(2500, R.IN)).astype(np.float32)
    t = rng.normal(size=(2500, R.OUT)).astype(np.float32)
    g, loss = sctx.surrogate_gradient(x, t)
    gr, lr = R.grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), emulate=True)
    assert abs(loss - lr) <= 1e-3 * lr, (loss, lr)
    o = 0
    for name, shp in R.SIZES:
        n = int(np.prod(shp))
        e = np.linalg.norm(g[o:o + n] - gr[o:o + n]) / np.linalg.norm(gr[o:o + n])
        print(name, f"rel err {e:.2e}")
        assert e < 3e-2, (name, e)
        o += n


def test_trained_surrogate_node_phase(sctx):
    import fjmodels_amd as fj
    ctx = sctx
    p, _ = O.guess_model(16, 0.01, 100.0, 3, 6)
    ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
    ctx.set_potential(*p.tables())
    ctx.set_delta(*p.delta())
    ctx.set_df(fj.df_params("Hernquist"))
    ctx.surrogate_init(seed=11, capacity=1 << 20)
    ctx.surrogate_collect(1.0)
    ref = ctx.compute_new_phi()                       # Staeckel node phase, collects every bound node
    t_st = ctx.stats()["ms_total"]
    ctx.surrogate_collect(0.0)
    n = ctx.surrogate_stats()["n_train"]
    l0 = ctx.surrogate_train(1, batch=16384, lr=3e-3)
    loss = ctx.surrogate_train(60, batch=16384, lr=3e-3)
    st = ctx.surrogate_stats()
    print(f"samples {n}, loss after 1 epoch {l0:.3e}, after 61 {loss:.3e}, training {st['ms_train']:.0f} ms")
    assert n > 100000 and loss < 0.2 * l0
    ctx.surrogate_mode(True, judge=0.02, jz_min=1e-3)
    try:
        got = ctx.compute_new_phi()
        t_sur = ctx.stats()["ms_total"]
        s = ctx.surrogate_stats()
    finally:
        ctx.surrogate_mode(False)
    err = {k: float(np.max(np.abs(got[k] - ref[k]) / np.abs(ref[k][:, :1]))) for k in ("rhl", "sigRl", "sigzl")}
    print("node phase ms: staeckel", t_st, "surrogate", t_sur, "stats", s, "coefficient errors", err)
    # what the network buys (node-phase time) and what it costs (action and
    # density errors); the bounds are the levels this training reaches
    assert s["nodes_sur"] > 0.8 * (s["nodes_sur"] + s["fallback"])
    assert s["judge_err_r_med"] < 0.06 and s["judge_err_z_med"] < 0.06
    assert err["rhl"] < 0.3
    assert t_sur < 0.8 * t_st
