"""Observables of integ/Integ.cpp in the oracle (CPU): the projected density
SigmaDF, line profiles and the velocity distributions P_vp, P_vz, P_vr.
Pins: the closed form of SigmaDF for a constant density table; the
symmetry of P_vp under vphi -> -vphi for a non-rotating DF (f depends on
Lz only through Lz^2 and |Jphi|); P_vp against an independent fine
integral of f over the same (vR, vz) square; and the evenness of P_vz and
P_vr in the equatorial plane."""
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


def test_vphi_distribution_matches_fine_integral():
    """P_vp (Integ.cpp:196-203, 107-118) integrates Ve^2 f(Ve vR, vphi, Ve vz)
    over (vR, vz) in [0, 1]^2 with the 22-point tanh-sinh rule, i.e. f over
    the square [0, Ve]^2 of physical (vR, vz) at the physical vphi.  A
    48 x 48 Gauss-Legendre integral of df() over that square gives the same
    marginal to the rule's accuracy (0.4 % at these points), at both signs of
    vphi of a rotating DF: a dropped Jacobian factor Ve, a [-1, 1] range or
    vphi in the wrong velocity slot would be off by tens of per cent."""
    p, _ = O.guess_model(12, 0.05, 50.0, 3, 4)
    df = O.df_params("Hernquist", chi=1.0)
    R, z = 1.0, 0.2
    Phi = p.phi(R, z)
    Ve = np.sqrt(-2 * (Phi - p.phi(10 * p.ar[-1], 10 * p.ar[-1])))
    g, w = np.polynomial.legendre.leggauss(48)
    u, wu = 0.5 * Ve * (g + 1), 0.5 * Ve * w
    x = np.array([R, z, Phi])
    for vphi in (-0.3, 0.1, 0.3, 0.8):
        fine = sum(wu[i] * wu[j] * p.df_eval(df, x, np.array([1e-6 + u[i], vphi, 1e-6 + u[j]]))
                   for i in range(48) for j in range(48))
        got = p.velocity_distribution(df, 0, R, z, vphi)
        assert abs(got / fine - 1) < 1e-2, (vphi, got, fine)


def test_vz_vr_distributions_even_in_the_plane():
    """In the equatorial plane the prolate coordinate v is pi/2, so the
    spheroidal momenta separate into one that holds vR and one that holds vz:
    f, and hence P_vz(vz) and P_vr(vR) (Integ.cpp:205-236), are even.  The
    DF rotates (chi = 1), so the distribution's velocity placed in the vphi
    slot would break the symmetry; off the plane the cross terms break it
    too (by 2-10 % at z = 0.2), which the check confirms."""
    p, _ = O.guess_model(12, 0.05, 50.0, 3, 4)
    df = O.df_params("Hernquist", chi=1.0)
    for which in (1, 2):
        for v in (0.1, 0.4):
            a = p.velocity_distribution(df, which, 1.0, 0.0, v)
            b = p.velocity_distribution(df, which, 1.0, 0.0, -v)
            assert a > 0 and abs(a - b) <= 1e-12 * a, (which, v, a, b)
            c = p.velocity_distribution(df, which, 1.0, 0.2, v)
            d = p.velocity_distribution(df, which, 1.0, 0.2, -v)
            assert abs(c / d - 1) > 1e-3, (which, v, c, d)
