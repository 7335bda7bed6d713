"""Pins of oracle functions whose result no library routine gives directly
(CPU only), each chosen so that a dropped term, a wrong sign or index or a
transposed operand fails:

* get_Ylm (include/UtilsLeg.h:200-253) recovers the coefficients of a known
  even Legendre series up to l = 12 on every row but the first, and applies
  the first row's monopole rule; ev_dens's reconstruction (UtilsLeg.h:336-354)
  returns the series;
* rhofDF's six sums and its vrot / sigma formulas (integ/Integ.cpp:46-62,
  150-163; GaussQuad.cpp:160-180) against the same 22^3-node rule summed
  independently in numpy over the DF at the nodes, for both velocity rules
  and with rotation;
* vir2's kinetic terms (libs/Stats.cpp:12-41) for the Hernquist density with
  the isotropic dispersion of the Jeans equation, computed here by direct
  quadrature: W/K = -2 (virial theorem), K_RR = 2 K_zz exactly (isotropy:
  sigma_R = sigma_phi = sigma_z), W_RR = 2 W_zz (spherical potential) and
  W_RR + W_zz = PE (Clausius);
* the nodes of lineProfile and Pvphi / Pvz / Pvr (Integ.cpp:64-138), which
  the library and the oracle build with the same header, against numpy
  restatements of the reference formulas with the DF at those nodes.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O

X11, W11, X01, W01 = O.quad_tables()


def legendre(l, x):
    return np.polynomial.legendre.Legendre.basis(l)(x)


def test_get_ylm_recovers_legendre_series():
    nr, ng, npoly = 12, 16, 7          # l <= 12; 16 nodes integrate degree 24 exactly
    p = O.Pot(O.log_grid(nr, 0.1, 10.0), npoly, ng)
    ci, wi = O.gauleg(ng)
    rng = np.random.default_rng(5)
    q = rng.normal(size=(6, nr, npoly))                       # series coefficients per field and row
    P = np.array([legendre(2 * k, ci) for k in range(npoly)])   # [npoly][ng]
    fields = [np.ascontiguousarray(q[f] @ P) for f in range(6)]  # [nr][ng]: F(c) = sum q_np P_2np(c)
    rho00 = fields[0][0, 0]
    outs = [np.zeros((nr, npoly)) for _ in range(6)]
    O.lib.fjo_get_ylm(p.h, *[O.ptr(f) for f in fields], *[O.ptr(o) for o in outs])
    expect = q * (2.0 / (4 * np.arange(npoly) + 1))             # int_{-1}^{1} P_2np F dc
    # gauleg stops its Newton iteration at 3e-11 (UtilsLeg.h:17), so the rule is exact to ~1e-10
    for f in range(6):
        assert np.allclose(outs[f][1:], expect[f][1:], rtol=1e-9, atol=1e-9 * np.abs(expect[f]).max()), f
        # first row: monopole from the first node only, the rest zero (UtilsLeg.h:243-249)
        first = fields[0][0, 0] if f == 0 else fields[f][0, 0]
        assert outs[f][0, 0] == 2 * first and np.all(outs[f][0, 1:] == 0), f
    assert np.all(fields[0][0] == rho00)                       # rho's first row set to rho[0][0] in place
    # ev_dens's reconstruction (UtilsLeg.h:336-354): .5 L_0 + sum (4np+1)/2 L_np P_2np = F
    c = np.linspace(-1, 1, 7)
    fac = 0.5 * (4 * np.arange(npoly) + 1)
    rec = (fac * outs[2][5]) @ np.array([legendre(2 * k, c) for k in range(npoly)])
    assert np.allclose(rec, q[2][5] @ np.array([legendre(2 * k, c) for k in range(npoly)]), rtol=1e-9)


@pytest.fixture(scope="module")
def small():
    return O.guess_model(12, 0.05, 50.0, 3, 4)[0]


def numpy_rhofDF(p, dfo, R, z, quad):
    """The 22^3 rule of Int3D_111111_vec (quad 0) or the tanh-sinh variant of
    Int3D_011101_vec (quad 1, density x4), the integrand of rhoInteg_vec."""
    Phi = p.phi(R, z)
    far = 10 * p.ar[-1]
    Ve = np.sqrt(-2 * (Phi - p.phi(far, far)))
    xa, wa = (X01, W01) if quad else (X11, W11)
    x = np.array([R, z, Phi])
    s = np.zeros(6)
    for i in range(22):
        for j in range(22):
            for k in range(22):
                V = 1e-6 + Ve * np.array([xa[i], X11[j], xa[k]])
                f = p.df_eval(dfo, x, V) * Ve ** 3
                w = wa[i] * W11[j] * wa[k]
                s += w * f * np.array([1, V[1], V[0] ** 2, V[1] ** 2, V[2] ** 2, V[0] * V[2]])
    return s


@pytest.mark.parametrize("quad,chi,rot_form", [(0, 0.0, 0), (1, 0.0, 0), (0, 1.0, 1)])
def test_rhofDF_sums_and_dispersions(small, quad, chi, rot_form):
    p = O.Pot(small.ar, small.npoly, small.ngauss, small.ngrid)
    p.set_tables(*small.tables())
    p.set_delta(*small.delta())
    p.set_quad(quad)
    dfo = O.df_params("Hernquist", chi=chi, rot_form=rot_form)
    R, z = 1.3, 0.4
    s = numpy_rhofDF(p, dfo, R, z, quad)
    dens, out6, (vrot, sR, sp, sz, sRz) = p.rhofDF(dfo, R, z)
    assert np.allclose(out6, s, rtol=1e-11, atol=1e-13 * abs(s[0])), (out6, s)
    assert dens == pytest.approx((4 if quad else 1) * s[0], rel=1e-11)
    assert vrot == pytest.approx(s[1] / s[0], rel=1e-9, abs=1e-13)
    assert sR == pytest.approx(np.sqrt(s[2] / s[0]), rel=1e-11)
    assert sp == pytest.approx(np.sqrt(s[3] / s[0]), rel=1e-11)
    assert sz == pytest.approx(np.sqrt(s[4] / s[0]), rel=1e-11)
    assert sRz == pytest.approx(np.sqrt(abs(s[5]) / s[0]), rel=1e-9)
    if chi:
        assert abs(vrot) > 1e-3     # the rotation term is live


def hernquist_sigma2(r):
    """Isotropic Jeans equation for the Hernquist model (G = M = a = 1):
    rho sigma^2 = int_r^inf rho(s) M(s) / s^2 ds, by direct quadrature."""
    rho = lambda s: 1 / (2 * np.pi * s * (1 + s) ** 3)
    M = lambda s: s * s / (1 + s) ** 2
    val, _ = integrate.quad(lambda s: rho(s) * M(s) / (s * s), r, np.inf, epsabs=0, epsrel=1e-12, limit=400)
    return val / rho(r)


def test_vir2_kinetic_terms_jeans_hernquist():
    p, rhl = O.guess_model(160, 1e-3, 1e3, 3, 10, delta=False)
    sig = np.array([np.sqrt(hernquist_sigma2(r)) for r in p.ar])
    s = np.zeros_like(rhl)
    s[:, 0] = 2 * sig                                 # ev_dens: sigma = .5 L_0 for a monopole table
    z = np.zeros_like(rhl)
    out = p.vir2(rhl, z, s, s, s)
    KE, PE, WK, KRR, WRR, Kzz, Wzz = out[2], out[3], out[4], out[5], out[6], out[8], out[9]
    print("KE", KE, "PE", PE, "W/K", WK, "KRR", KRR, "Kzz", Kzz, "WRR", WRR, "Wzz", Wzz)
    assert KRR == 2 * Kzz                              # sigma_R = sigma_phi = sigma_z: exactly twice
    assert KE == KRR + Kzz and WK == PE / KE
    # virial theorem 2K + W = 0, to the grid's discretisation (midpoint radii, densities
    # interpolated linearly between log-spaced radii: both K and PE come out ~1 % high)
    assert abs(WK + 2) < 2e-2
    assert abs(out[7] + 2) < 2e-2 and abs(out[10] + 2) < 2e-2
    assert abs(WRR / Wzz - 2) < 1e-6                   # spherical: <R dPhi/dR> = 2 <z dPhi/dz>
    assert abs((WRR + Wzz) / PE - 1) < 5e-3            # Clausius virial of a self-gravitating system
    assert abs(KE * 12 - 1) < 2e-2                     # K = -W/2 = G M^2 / (12 a)


def numpy_lp_node(R, Ve, armax, vx, vy, x, vz):
    """lineProfile (Integ.cpp:64-92), from the reference's expressions."""
    Rl = np.sqrt(R ** 2 + (x * armax) ** 2)
    y = R
    v_abs = np.sqrt(vx * vx + vy * vy)
    alpha, theta = np.arccos(vx / v_abs), np.arctan2(y, x)
    vr, vphi = v_abs * np.cos(alpha + theta), v_abs * np.sin(alpha + theta)
    V = np.array([1e-6 + Ve * vr, 1e-6 + Ve * vphi, 1e-6 + Ve * vz])
    return V, Ve * Ve * (abs(x) * armax ** 2) / Rl


def test_line_profile_nodes(small):
    p = O.Pot(small.ar, small.npoly, small.ngauss, small.ngrid)
    p.set_tables(*small.tables())
    p.set_delta(*small.delta())
    dfo = O.df_params("Hernquist")
    rhl = O.Pot(small.ar, small.npoly, small.ngauss).guess_rhl("Hernquist")
    R, z = 1.1, 0.3
    Phi = p.phi(R, z)
    far = 10 * p.ar[-1]
    Ve = np.sqrt(-2 * (Phi - p.phi(far, far)))
    Sigma = p.sigma_df(rhl, R, z)
    for vx in (0.35, -0.6):
        s = 0.0
        for i in range(22):
            for j in range(22):
                for k in range(22):
                    V, jac = numpy_lp_node(R, Ve, p.ar[-1], vx, X11[i], X11[j], X11[k])
                    s += W11[i] * W11[j] * W11[k] * (jac * p.df_eval(dfo, np.array([R, z, Phi]), V))
        got = p.line_profile(dfo, rhl, R, z, vx)
        print(vx, got, s / Sigma)
        assert got == pytest.approx(s / Sigma, rel=1e-6)   # node velocities agree to rounding
        assert got > 0


@pytest.mark.parametrize("which", [0, 1, 2])
def test_velocity_distribution_nodes(small, which):
    p = O.Pot(small.ar, small.npoly, small.ngauss, small.ngrid)
    p.set_tables(*small.tables())
    p.set_delta(*small.delta())
    dfo = O.df_params("Hernquist", chi=1.0, rot_form=1)
    R, z, v = 0.9, 0.2, 0.15
    Phi = p.phi(R, z)
    far = 10 * p.ar[-1]
    Ve = np.sqrt(-2 * (Phi - p.phi(far, far)))
    xb, wb = (X01, W01) if which == 0 else (X11, W11)
    s = 0.0
    for i in range(22):
        for j in range(22):
            a, b = X01[i], xb[j]
            if which == 0:     # Pvphi (Integ.cpp:97-109): (vr, vz) in [0,1]^2
                V = [1e-6 + Ve * a, v, 1e-6 + Ve * b]
            elif which == 1:   # Pvz (111-123): (vr, vphi) in [0,1] x [-1,1]
                V = [1e-6 + Ve * a, 1e-6 + Ve * b, v]
            else:              # Pvr (125-138): (vz, vphi) in [0,1] x [-1,1]
                V = [v, 1e-6 + Ve * b, 1e-6 + Ve * a]
            s += W01[i] * wb[j] * (Ve * Ve * p.df_eval(dfo, np.array([R, z, Phi]), np.array(V)))
    got = p.velocity_distribution(dfo, which, R, z, v)
    assert got == pytest.approx(s, rel=1e-12)
    assert got > 0
