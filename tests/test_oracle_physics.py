"""Pins of the oracle's physics (CPU only): closed forms and invariants.

  * potential evaluation (Potential.cpp:206-276): tables linear in r make the
    radial interpolation exact, so Phi, dPhi/dR, dPhi/dz must equal the
    analytic sum of (a_k + b_k r) P_2k(cos theta) and its derivatives;
  * guess densities (models.cpp:40-68): textbook Hernquist, isochrone, NFW;
  * Poisson (Potential.cpp:110-200): QAGS vs the exact integral of the
    piecewise-linear rho_l, and the Hernquist monopole -M/(r+a) on a fine grid;
  * Staeckel-fudge actions (uvOrb.cpp): spherical isochrone with a tiny focal
    distance reproduces the analytic Jr and Jz = L - |Lz|; planar orbits have
    Jz = 0; circular orbits have Jr ~ 0;
  * DF (DF.cpp:65-76,108-119): the documented double-power-law form;
  * virial sums (Stats.cpp:12-41): W = -M^2/(6a) for the Hernquist guess and
    W_RR + W_zz = W (the virial identity for the same density/potential).
"""
import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O


def linear_tables(ar, npoly, rng):
    a = rng.normal(size=npoly) * np.array([1.0] + [0.3] * (npoly - 1))
    b = rng.normal(size=npoly) * 0.05
    phil = a[None, :] + b[None, :] * ar[:, None]
    Pr = np.repeat(b[None, :], len(ar), axis=0)
    return a, b, phil, Pr


def test_potential_eval_exact_on_linear_tables():
    rng = np.random.default_rng(0)
    ar = O.log_grid(40, 0.01, 100.0)
    npoly = 5
    a, b, phil, Pr = linear_tables(ar, npoly, rng)
    p = O.Pot(ar, npoly, 20)
    p.set_tables(phil, Pr, np.zeros_like(phil))
    for _ in range(200):
        r = 10 ** rng.uniform(-1.9, 1.99)
        th = rng.uniform(0.01, np.pi - 0.01)
        R, z = r * np.sin(th), r * np.cos(th)
        c = np.cos(th)
        P = np.array([sp.eval_legendre(2 * k, c) for k in range(npoly)])
        dP = np.array([np.polynomial.legendre.Legendre.basis(2 * k).deriv()(c) for k in range(npoly)])
        phi = np.sum((a + b * r) * P)
        dphidr = np.sum(b * P)
        dphidc = np.sum((a + b * r) * dP)
        # chain rule: c = z/r; dc/dR = -z R / r^3, dc/dz = R^2 / r^3
        dR = dphidr * R / r + dphidc * (-z * R / r ** 3)
        dz = dphidr * z / r + dphidc * (R * R / r ** 3)
        assert abs(p.phi(R, z) - phi) < 1e-12 * max(1, abs(phi))
        assert abs(p.dR(R, z) - dR) < 1e-9 * max(1, abs(dR))
        assert abs(p.dz(R, z) - dz) < 1e-9 * max(1, abs(dz))


def test_potential_keplerian_tail():
    """Beyond r_max each multipole falls as r^-(2k+1) (UtilsLeg.h:98-100)."""
    rng = np.random.default_rng(1)
    ar = O.log_grid(30, 0.01, 50.0)
    _, _, phil, Pr = linear_tables(ar, 4, rng)
    p = O.Pot(ar, 4, 10)
    p.set_tables(phil, Pr, np.zeros_like(phil))
    for r in (60.0, 300.0, 5e3):
        for th in (0.3, 1.1, 2.0):
            c = np.cos(th)
            P = np.array([sp.eval_legendre(2 * k, c) for k in range(4)])
            ref = np.sum(phil[-1] * (ar[-1] / r) ** (2 * np.arange(4) + 1) * P)
            assert abs(p.phi(r * np.sin(th), r * np.cos(th)) - ref) < 1e-13 * abs(ref) + 1e-300


def test_guess_densities_closed_forms():
    ar = O.log_grid(12, 0.05, 20.0)
    p = O.Pot(ar, 1, 8)     # npoly = 1: rhl[:,0] = sum_g 2 w_g rho = int_{-1}^{1} rho dc
    for model, rho in {
        "Hernquist": lambda r, M, a: M * a / (2 * np.pi * r * (r + a) ** 3),
        "NFW": lambda r, M, a: M / (4 * np.pi * r * (r + a) ** 2) * a,
        "Isochrone": lambda r, M, b: M * (3 * (b + np.sqrt(b * b + r * r)) * (b * b + r * r) -
                                          r * r * (b + 3 * np.sqrt(b * b + r * r))) /
                                     (4 * np.pi * (b + np.sqrt(b * b + r * r)) ** 3 * (b * b + r * r) ** 1.5),
    }.items():
        M, a = 2.0, 1.3
        rhl = p.guess_rhl(model, M, a, 1.0)
        exact = 2 * rho(ar, M, a)
        # row 0 is 2 rho(ar[0] si[0], ...), identical for a spherical model
        assert np.max(np.abs(rhl[:, 0] / exact - 1)) < 1e-10, model


def test_poisson_vs_exact_piecewise_integrals():
    """I_int = int_0^r x^(l+2) rho_l and I_ext = int_r^inf rho_l x^(1-l) with
    rho_l piecewise linear (constant below ar[0], r^-4 beyond r_max) have exact
    values; QAGS/QAGIU at epsabs 1e-4, epsrel 1e-3 must land within tolerance,
    and the phil/Pr/Pr2 formulas (Potential.cpp:171-178) follow from them."""
    rng = np.random.default_rng(3)
    ar = O.log_grid(20, 0.05, 30.0)
    npoly = 3
    rhl = np.abs(rng.normal(size=(20, npoly))) * (1 + 1 / ar[:, None]) + 0.1
    p = O.Pot(ar, npoly, 8)
    phil, Pr, Pr2 = p.compute_phil(rhl)

    def exact_int(l, n):
        y = rhl[:, l // 2]
        r = ar[n]
        tot = y[0] * ar[0] ** (l + 3) / (l + 3) if r >= ar[0] else y[0] * r ** (l + 3) / (l + 3)
        for i in range(n):
            x0, x1, y0, y1 = ar[i], ar[i + 1], y[i], y[i + 1]
            s = (y1 - y0) / (x1 - x0)
            c0 = y0 - s * x0
            tot += c0 * (x1 ** (l + 3) - x0 ** (l + 3)) / (l + 3) + s * (x1 ** (l + 4) - x0 ** (l + 4)) / (l + 4)
        return tot

    def exact_ext(l, n):
        y = rhl[:, l // 2]
        tot = 0.0
        for i in range(n, len(ar) - 1):
            x0, x1, y0, y1 = ar[i], ar[i + 1], y[i], y[i + 1]
            s = (y1 - y0) / (x1 - x0)
            c0 = y0 - s * x0
            e = 2 - l   # x^(1-l) integrates to x^(2-l)/(2-l), l = 2 gives log
            if e == 0:
                tot += c0 * np.log(x1 / x0) + s * (x1 - x0)
            else:
                tot += c0 * (x1 ** e - x0 ** e) / e + s * (x1 ** (e + 1) - x0 ** (e + 1)) / (e + 1)
        rm = ar[-1]
        tot += max(0.0, y[-1]) * rm ** 4 * rm ** (-2 - l) / (2 + l)
        return tot

    for np_ in range(npoly):
        l = 2 * np_
        for n in range(len(ar)):
            r = ar[n]
            Ii, Ie = exact_int(l, n), exact_ext(l, n)
            ref_phi = -2 * np.pi * (Ii / r ** (l + 1) + Ie * r ** l)
            ref_pr = -2 * np.pi * (-(l + 1) * Ii / r ** (l + 2) + l * Ie * r ** (l - 1))
            scale = abs(Ii / r ** (l + 1)) + abs(Ie * r ** l)
            assert abs(phil[n, np_] - ref_phi) <= 2.5e-3 * 2 * np.pi * scale + 1e-3, (l, n)
            assert abs(Pr[n, np_] - ref_pr) <= 2.5e-3 * 2 * np.pi * (abs((l + 1) * Ii / r ** (l + 2)) +
                                                                  abs(l * Ie * r ** max(l - 1, 0) / max(r, 1e-300) * r)) + 1e-3


def test_hernquist_monopole_fine_grid():
    p, rhl = O.guess_model(160, 1e-3, 1e3, 2, 8, delta=False)
    phil, Pr, _ = p.tables()
    r = p.ar
    sel = (r > 0.02) & (r < 200)
    # QAGS at epsabs 1e-4 / epsrel 1e-3 on linearly interpolated rho_0 bounds the agreement
    assert np.max(np.abs(phil[sel, 0] / (-1 / (r[sel] + 1)) - 1)) < 1e-2
    # mass inside r_max from the potential derivative (Stats.cpp:37): r^2 Phi_0' ~ M(<r)
    m = r[-1] ** 2 * Pr[-1, 0]
    assert abs(m - r[-1] ** 2 / (r[-1] + 1) ** 2) < 1e-2


@pytest.fixture(scope="module")
def isochrone():
    """Spherical isochrone (M = 1, b = 1) tabulated as a pure monopole on a fine grid."""
    ar = O.log_grid(400, 1e-3, 1e3)
    s = np.sqrt(1 + ar ** 2)
    phil = np.zeros((400, 2))
    Pr = np.zeros((400, 2))
    phil[:, 0] = -1 / (1 + s)
    Pr[:, 0] = ar / (s * (1 + s) ** 2)
    p = O.Pot(ar, 2, 8)
    p.set_tables(phil, Pr, np.zeros_like(phil))
    return p


def test_isochrone_actions_spherical_limit(isochrone):
    p = isochrone
    rng = np.random.default_rng(4)
    n_ok = 0
    for _ in range(60):
        r = 10 ** rng.uniform(-0.7, 0.8)
        th = rng.uniform(0.2, np.pi / 2 - 0.05)
        R, z = r * np.sin(th), r * np.cos(th)
        Phi = p.phi(R, z)
        v = np.sqrt(-2 * Phi) * rng.uniform(0.2, 0.8)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        vR, vphi, vz = v * d
        E = 0.5 * v * v + Phi
        Lvec = np.cross([R, 0, z], [vR, vphi, vz])
        L = np.linalg.norm(Lvec)
        Jr_ex = 1 / np.sqrt(-2 * E) - 0.5 * (L + np.sqrt(L * L + 4))
        Jz_ex = L - abs(R * vphi)
        Jr, Jz = p.actions(1e-3, R, z, vR, vphi, vz, Phi)
        assert abs(Jr - Jr_ex) < 3e-3 * (Jr_ex + L) + 1e-5
        assert abs(Jz - Jz_ex) < 3e-3 * (Jz_ex + L) + 1e-5
        n_ok += 1
    assert n_ok == 60


def test_planar_orbits_have_zero_Jz(isochrone):
    p = isochrone
    for R, vR, vphi in [(0.7, 0.1, 0.5), (2.0, -0.2, 0.3), (0.3, 0.0, 0.9)]:
        Jr, Jz = p.actions(0.05, R, 0.0, vR, vphi, 0.0)
        assert Jz == 0.0
        assert Jr > 0


def test_circular_orbit_has_small_Jr(isochrone):
    p = isochrone
    for R in (0.5, 1.0, 3.0):
        vc = np.sqrt(R * p.dR(R, 0.0))
        Jr, Jz = p.actions(0.05, R, 0.0, 0.0, vc, 0.0)
        assert Jr < 1e-3 * R * vc
        assert Jz < 1e-12 * R * vc


def test_double_power_law_form():
    """f = M/J0^3 (1+J0/h)^A / (1+g/J0)^B / (2 pi)^3, h = Jr + dh_phi|Jphi| + dh_z Jz,
    g = Jr + dg_phi|Jphi| + dg_z Jz (README DF.png); odd part k tanh(chi Jphi/J0) f."""
    import ctypes as C
    d = O.df_params("NFW", dphi_h=0.6, dz_h=0.7, dphi_g=1.1, dz_g=0.9, mass=3.0, r0=2.0)
    rng = np.random.default_rng(5)
    for _ in range(50):
        Jr, Jz = 10 ** rng.uniform(-3, 1, 2)
        Jp = rng.uniform(-3, 3)
        h = Jr + 0.6 * abs(Jp) + 0.7 * Jz
        g = Jr + 1.1 * abs(Jp) + 0.9 * Jz
        J0 = np.sqrt(6.0)
        ref = 3.0 / J0 ** 3 * (1 + J0 / h) ** 1.667 / (1 + g / J0) ** 3.0 / (2 * np.pi) ** 3
        assert abs(O.lib.fjo_df_hg(C.byref(d), Jr, Jp, Jz) / ref - 1) < 1e-13


def test_virial_identity_hernquist_guess():
    p, rhl = O.guess_model(160, 1e-3, 1e3, 3, 10, delta=False)
    z = np.zeros_like(rhl)
    out = p.vir2(rhl, z, z, z, z)
    PE, WRR, Wzz = out[3], out[6], out[9]
    assert abs(PE - (-1 / 6)) < 5e-3          # W = -G M^2 / (6 a)
    assert abs((WRR + Wzz) - PE) < 5e-3       # -int rho x.grad(Phi) = W
    assert abs(out[0] - 1.0) < 1e-2           # r^2 dPhi0/dr at r_max ~ M(<r_max)
