"""Observables of integ/Integ.cpp in the oracle (CPU): the projected density
SigmaDF, line profiles and the velocity distributions P_vp, P_vz, P_vr.
Pins: the closed form of SigmaDF for a constant density table, and the
symmetry of P_vp under vphi -> -vphi for a non-rotating DF (f depends on
Lz only through Lz^2 and |Jphi|)."""
import numpy as np

from oracle import oracle as O


def test_sigma_df_constant_density_closed_form():
    p, _ = O.guess_model(12, 0.05, 50.0, 3, 4, delta=False)
    rho0 = 0.37
    rhl = np.zeros((12, 3))
    rhl[:, 0] = 2 * rho0          # ev_dens = rhl[0] / 2 on the midplane
    rmax = p.ar[-1]
    for R in (p.ar[0], 1.0, 20.0):
        got = p.sigma_df(rhl, R, 0.0)
        ref = 2 * sum(w * ((rmax - R) * rho0) for w in O.quad_tables()[3])
        assert abs(got - ref) <= 1e-15 * abs(ref), (R, got, ref)


def test_vphi_distribution_symmetric_without_rotation():
    p, _ = O.guess_model(12, 0.05, 50.0, 3, 4)
    df = O.df_params("Hernquist")
    for v in (0.05, 0.3):
        a = p.velocity_distribution(df, 0, 1.0, 0.2, v)
        b = p.velocity_distribution(df, 0, 1.0, 0.2, -v)
        assert a > 0 and a == b, (v, a, b)
    dfr = O.df_params("Hernquist", chi=1.0)
    a = p.velocity_distribution(dfr, 0, 1.0, 0.2, 0.3)
    b = p.velocity_distribution(dfr, 0, 1.0, 0.2, -0.3)
    assert a > b > 0                # prograde excess of the odd rotation term


def test_line_profile_positive_and_finite():
    p, rhl = O.guess_model(12, 0.05, 50.0, 3, 4)
    df = O.df_params("Hernquist")
    lp = [p.line_profile(df, rhl, 1.0, 0.05, vx) for vx in (-0.2, 0.01, 0.2)]
    assert all(np.isfinite(x) and x > 0 for x in lp), lp
