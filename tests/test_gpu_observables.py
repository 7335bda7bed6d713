"""Observables of integ/Integ.cpp through the C ABI against the oracle, bit
for bit: f(J) at explicit points (fj_df_points), the projected density
SigmaDF, line profiles (line_profile) and the velocity distributions P_vp,
P_vz, P_vr, on a small Hernquist model after one SCF iteration (rotating DF,
so the odd part is exercised too)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model(ctx):
    import fjmodels_amd as fj
    p, _ = O.guess_model(12, 0.05, 50.0, 3, 4)
    df = fj.df_params("Hernquist", chi=1.0, rot_form=1)
    dfo = O.df_params("Hernquist", chi=1.0, rot_form=1)
    ref = p.compute_new_phi(dfo, nthreads=16)
    p.set_tables(ref["phil"], ref["Pr"], ref["Pr2"])
    ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
    ctx.set_potential(*p.tables())
    ctx.set_delta(*p.delta())
    ctx.set_df(df)
    return ctx, p, dfo, ref["rhl"]


def test_df_points(model):
    ctx, p, dfo, _ = model
    rng = np.random.default_rng(3)
    n = 300
    R, z = rng.uniform(0.1, 3, n), rng.uniform(-1, 1, n)
    V = rng.normal(0, 0.25, (3, n))
    got = ctx.df_points(R, z, V[0], V[1], V[2])
    ref = np.array([p.df_eval(dfo, [R[i], z[i], p.phi(R[i], z[i])], V[:, i]) for i in range(n)])
    assert np.count_nonzero(ref) > 100
    assert np.array_equal(got, ref), np.max(np.abs(got - ref))


def test_sigma_df(model):
    ctx, p, _, rhl = model
    R = np.array([p.ar[0], 0.7, 3.0, 20.0])
    z = np.array([p.ar[0], 0.0, 0.5, 2.0])
    got = ctx.sigma_df(rhl, R, z)
    ref = np.array([p.sigma_df(rhl, a, b) for a, b in zip(R, z)])
    assert np.array_equal(got, ref), (got, ref)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_velocity_distributions(model, which):
    ctx, p, dfo, _ = model
    v = np.array([-0.4, -0.05, 0.02, 0.3])
    got = ctx.velocity_distribution(which, 0.8, 0.1, v)
    ref = np.array([p.velocity_distribution(dfo, which, 0.8, 0.1, x) for x in v])
    assert np.all(ref > 0)
    assert np.array_equal(got, ref), (which, got, ref)


def test_line_profile(model):
    ctx, p, dfo, rhl = model
    vx = np.array([-0.3, 0.01, 0.25])
    R, z = float(p.ar[4]), float(p.ar[0])
    got = ctx.line_profile(rhl, R, z, vx)
    ref = np.array([p.line_profile(dfo, rhl, R, z, x) for x in vx])
    assert np.all(ref > 0)
    assert np.array_equal(got, ref), (got, ref)
