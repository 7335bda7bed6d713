"""The gfx950 node code (fj_device.hpp / fj_node.hpp), compiled for the host,
reproduces the oracle bit for bit (CPU only).

Both sides take every discrete decision of the reference algorithm on the
same bits: the potential and its gradient, the spheroidal coordinates,
Getu0 / GetTurnu / GetTurnv with their Brent calls, QAGS on the action
integrands, the DF and the six sums in one reduction order.  The GPU tests
then check that the kernels give the same bits as this host build.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O

HB = os.path.join(ROOT, "tests", "hostbuild")
dp = C.POINTER(C.c_double)

X11 = np.array([-1., -0.50118723, -0.25118864, -0.12589254, -0.06309573, -0.03162278, -0.01584893,
                -0.00794328, -0.00398107, -0.00199526, -0.001, 0.001, 0.00199526, 0.00398107,
                0.00794328, 0.01584893, 0.03162278, 0.06309573, 0.12589254, 0.25118864, 0.50118723, 1.])
W11 = np.array([0.49881277, 0.37440568, 0.18764735, 0.09404645, 0.04713488, 0.0236234, 0.01183975,
                0.00593393, 0.00297401, 0.00149054, 0.00149763, 0.00149763, 0.00149054, 0.00297401,
                0.00593393, 0.01183975, 0.0236234, 0.04713488, 0.09404645, 0.18764735, 0.37440568, 0.49881277])


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)


FNF = C.CFUNCTYPE(C.c_float, C.c_float, C.c_void_p)


@pytest.fixture(scope="module")
def H():
    import subprocess
    subprocess.run(["make", "-s", "-C", HB], check=True)
    lib = C.CDLL(os.path.join(HB, "libfjhost.so"))
    lib.fjh_phi_grad.restype = C.c_double
    lib.fjh_phi_grad.argtypes = [C.c_int, dp, C.c_int, dp, dp, C.c_double, C.c_double, dp, dp]
    lib.fjh_phi.restype = C.c_double
    lib.fjh_phi.argtypes = [C.c_int, dp, C.c_int, dp, dp, C.c_double, C.c_double]
    lib.fjh_actions.argtypes = [C.c_int, dp, C.c_int, dp, dp, C.c_int, dp, dp, C.c_int] + [dp] * 7 + \
        [C.c_int, dp, dp, C.POINTER(C.c_int)]
    lib.fjh_rhofDF.argtypes = [C.c_int, dp, C.c_int, dp, dp, C.c_int, dp, dp, dp, dp, dp, dp, C.c_double,
                               C.c_double] + [C.c_double] * 9 + [C.c_double, C.c_int, dp]
    lib.fjh_bracket.argtypes = [C.c_int, dp, C.c_int, dp, C.POINTER(C.c_int)]
    lib.fjh_divrn_check.restype = C.c_longlong
    lib.fjh_divrn_check.argtypes = [C.c_longlong, C.c_ulonglong, C.c_int]
    lib.fjh_qags.argtypes = [O.FN, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                             C.c_int, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.fjh_qags32.argtypes = [FNF, C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                               C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    return lib


@pytest.fixture(scope="module")
def models():
    out = {}
    out["hq40"] = O.guess_model(40, 0.01, 100.0, 5, 20)[0]
    out["small"] = O.guess_model(12, 0.05, 50.0, 3, 4)[0]
    return out


def tabs(p):
    phil, Pr, Pr2 = p.tables()
    E, D = p.delta()
    return phil, Pr, E, D


@pytest.mark.parametrize("name", ["hq40", "small"])
def test_potential_bitwise(H, models, name):
    p = models[name]
    phil, Pr, _, _ = tabs(p)
    rng = np.random.default_rng(11)
    r = 10 ** rng.uniform(-3.5, 3.5, 3000)
    th = rng.uniform(0, np.pi, 3000)
    for R, z in zip(r * np.sin(th), r * np.cos(th)):
        a, b = C.c_double(), C.c_double()
        phi = H.fjh_phi_grad(p.nr, P(p.ar), p.npoly, P(phil), P(Pr), R, z, C.byref(a), C.byref(b))
        assert phi == p.phi(R, z) and a.value == p.dR(R, z) and b.value == p.dz(R, z), (R, z)
        assert H.fjh_phi(p.nr, P(p.ar), p.npoly, P(phil), P(Pr), R, z) == phi


def point_nodes(p, R, z):
    Phi = p.phi(R, z)
    far = 10 * p.ar[-1]
    Ve = np.sqrt(-2 * (Phi - p.phi(far, far)))
    V = 1e-6 + Ve * np.stack(np.meshgrid(X11, X11, X11, indexing="ij"), -1).reshape(-1, 3)
    H_ = 0.5 * np.sum(V ** 2, axis=1) + Phi
    b = H_ < 0
    return Phi, V[b].copy(), np.array([p.deltafn(h) for h in H_[b]])


@pytest.mark.parametrize("name,pts", [("hq40", [(0.866, 0.5), (60.0, 40.0), (0.011, 0.004)]),
                                      ("small", [(47.198880926701165, 16.50047391037859), (0.9, 2.0)])])
def test_actions_bitwise(H, models, name, pts):
    p = models[name]
    phil, Pr, E, D = tabs(p)
    for R, z in pts:
        Phi, V, Dl = point_nodes(p, R, z)
        if name == "hq40":   # a sample keeps the CPU time down
            V, Dl = V[::7].copy(), Dl[::7].copy()
        n = len(V)
        Jr, Jz = np.zeros(n), np.zeros(n)
        H.fjh_actions(p.nr, P(p.ar), p.npoly, P(phil), P(Pr), p.ngrid, P(E), P(D), n, P(np.full(n, R)),
                      P(np.full(n, z)), P(np.full(n, Phi)), P(V[:, 0]), P(V[:, 1]), P(V[:, 2]), P(Dl), 48,
                      P(Jr), P(Jz), None)
        ref = np.array([p.actions(Dl[i], R, z, V[i, 0], V[i, 1], V[i, 2], Phi) for i in range(n)])
        bad = np.where((Jr != ref[:, 0]) | (Jz != ref[:, 1]))[0]
        assert len(bad) == 0, (name, R, z, len(bad), n, bad[:5])


@pytest.mark.parametrize("name,R,z", [("small", 47.198880926701165, 16.50047391037859), ("small", 2.0, 1.0),
                                      ("hq40", 0.866, 0.5)])
def test_rhofDF_sums_bitwise(H, models, name, R, z):
    p = models[name]
    phil, Pr, E, D = tabs(p)
    d = O.df_params("Hernquist")
    out = np.zeros(6)
    H.fjh_rhofDF(p.nr, P(p.ar), p.npoly, P(phil), P(Pr), p.ngrid, P(E), P(D), P(X11), P(W11), P(X11), P(W11),
                 R, z, d.A, d.B, d.dphi_h, d.dz_h, d.dphi_g, d.dz_g, d.mass, d.J0, d.chi, d.k, d.rot_form, P(out))
    _, ref, _ = p.rhofDF(d, R, z)
    assert np.array_equal(out, ref), (out, ref)


def test_qags_stepper_bitwise(H):
    """The device QAGS stepper against the oracle's dqagse on integrands that
    subdivide and extrapolate (piecewise-linear Poisson-type, endpoint
    singularities, kinks), including a 48-interval workspace that overflows."""
    rng = np.random.default_rng(5)
    xs = np.sort(np.concatenate([[0.0], 10 ** rng.uniform(-2, 1, 30)]))
    ys = np.abs(rng.normal(size=31)) + 0.1
    fs = [lambda x: x ** 2 * np.interp(x, xs, ys),
          lambda x: np.log(x) / np.sqrt(x) if x > 0 else 0.0,
          lambda x: np.sqrt(abs(x - 0.3137)),
          lambda x: np.sqrt(max(0.0, np.cos(3 * x) + 0.2)) * np.sin(x)]
    for f in fs:
        cb = O.FN(lambda x, prm: f(x))
        for a, b in ((0.0, 1.0), (0.0, 3.0)):
            for cap in (1000, 48):
                r, ier, nqk = C.c_double(), C.c_int(), C.c_int()
                H.fjh_qags(cb, None, a, b, 1e-6, 1e-6, 1000, cap, 0, C.byref(r), C.byref(ier), C.byref(nqk))
                rr, ee = C.c_double(), C.c_double()
                ne, last = C.c_int(), C.c_int()
                ier_o = O.lib.fjo_qags(cb, None, a, b, 1e-6, 1e-6, 1000, C.byref(rr), C.byref(ee), C.byref(ne),
                                       C.byref(last))
                if ier.value == -1:        # workspace overflow: the kernel reruns the task at full capacity
                    assert cap == 48 and last.value > 48
                    continue
                assert r.value == rr.value and ier.value == ier_o, (f, a, b, cap)


@pytest.mark.parametrize("kind", ["log", "uneven", "dense_inner"])
def test_radial_bracket_matches_bisection(H, kind):
    """The bracket table + walk of radial_interp gives topbottom's bisection
    bracket (Utils.h:103-128): the largest m <= nr-2 with r >= ar[m], else 0;
    radii on, just off and between the nodes, below ar[0], NaN."""
    rng = np.random.default_rng(1)
    if kind == "log":
        ar = O.log_grid(40, 0.01, 100.0)
    elif kind == "uneven":
        ar = np.sort(10 ** rng.uniform(-3, 3, 97))
    else:
        ar = np.concatenate([np.linspace(1e-3, 1e-2, 60), O.log_grid(30, 0.02, 50.0)])
    nr = len(ar)
    r = np.concatenate([ar[:-1], np.nextafter(ar[:-1], 0), np.nextafter(ar[:-1], np.inf),
                        10 ** rng.uniform(np.log10(ar[0]) - 1, np.log10(ar[-1]), 5000), [np.nan, 1e-300]])
    r = r[~(r >= ar[-1])]
    bot = np.zeros(len(r), dtype=np.int32)
    H.fjh_bracket(nr, P(ar), len(r), P(r), bot.ctypes.data_as(C.POINTER(C.c_int)))
    ref = np.zeros(len(r), dtype=np.int32)
    for i, x in enumerate(r):
        top, b = nr - 1, 0
        while top - b > 1:
            m = (top + b) >> 1
            if x >= ar[m]:
                b = m
            else:
                top = m
        ref[i] = b
    assert np.array_equal(bot, ref), np.where(bot != ref)[0][:5]


@pytest.mark.parametrize("mode", [0, 1], ids=["random", "path_divisors"])
def test_reciprocal_division_is_ieee_division(H, mode):
    """fj_div_rn(a, b, RN(1/b)) (three fused multiply-adds) gives the bits of
    IEEE a/b: on random operands over 2^-60..2^60 and on every divisor the
    potential evaluation divides by (the evenlegend and legend constants, the
    radial cell widths and quadratic-correction products of a 200-radius log
    grid); zeros keep the quotient's sign."""
    assert H.fjh_divrn_check(4_000_000, 12345 + mode, mode) == 0


def test_qags_float_instance(H):
    """The float instance of the QAGS stepper with the FP32 QK21 rule (the
    mixed-precision kernels' bookkeeping: QUADPACK's single-precision qagse)
    against scipy's QUADPACK in double precision on integrands that subdivide
    and extrapolate: the result meets the requested tolerance (plus FP32
    rounding of the integrand), or the stepper reports roundoff (ier 2), the
    workspace cap (-1, the kernels' FP64 fallback) or ier 5 (divergent
    extrapolation), as dqagse would."""
    from scipy.integrate import quad
    import math
    fs = [lambda x: x * x * math.exp(-x),
          lambda x: math.log(x) / math.sqrt(x) if x > 0 else 0.0,
          lambda x: math.sqrt(abs(x - 0.3137)),
          lambda x: math.sqrt(max(0.0, math.cos(3 * x) + 0.2)) * math.sin(x),
          lambda x: 1.0 / (1e-3 + (x - 0.5) ** 2)]
    stats = []
    for f in fs:
        cb = FNF(lambda x, prm: float(f(x)))
        for a, b in ((0.0, 1.0), (0.0, 3.0)):
            ref = quad(f, a, b, epsabs=1e-12, epsrel=1e-12, limit=1000)[0]
            for epsrel in (1e-3, 1e-5):
                r, ier, nqk = C.c_float(), C.c_int(), C.c_int()
                H.fjh_qags32(cb, None, a, b, 1e-5, epsrel, 1000, 64, C.byref(r), C.byref(ier), C.byref(nqk))
                err = abs(r.value - ref)
                tol = max(1e-5, epsrel * abs(ref)) + 1e-6 * abs(ref) + 1e-6
                stats.append((ier.value, nqk.value, err / tol))
                assert ier.value in (0, 2, 5, -1), (f, a, b, ier.value)
                if ier.value == 0:
                    assert err <= 3 * tol, (f, a, b, epsrel, r.value, ref)
                else:
                    assert err <= 100 * tol, (f, a, b, epsrel, ier.value, r.value, ref)
    print(stats)
    assert sum(1 for s in stats if s[0] == 0) >= len(stats) // 2


def test_fast_mode_actions_near_reference(H, models):
    """The fast mode's actions (fj_fast.hpp: the reference's walks from the
    node's own constants, regula falsi roots, 64-point Gauss-Legendre
    integrals), host build with FP64 integrands, against the oracle's
    (Getu0 + Brent + QAGS): the same turning-point brackets and endpoint
    returns, so the differences are quadrature-level (the QAGS tolerance is
    1e-3 relative)."""
    H.fjh_fast_actions.argtypes = [C.c_int, dp, C.c_int, dp, dp, C.c_int, dp, dp, C.c_int] + [dp] * 6 + \
        [C.c_int, dp, dp, C.c_int, dp, dp, dp, dp, C.POINTER(C.c_int), dp]
    p = models["hq40"]
    phil, Pr, E, D = tabs(p)
    gx, gw = np.polynomial.legendre.leggauss(64)
    gx, gw = 0.5 * (gx + 1), 0.5 * gw
    vx, vw = np.polynomial.legendre.leggauss(32)   # the J_z rule: half the order
    vx, vw = 0.5 * (vx + 1), 0.5 * vw
    er, ez = [], []
    for R, z in [(0.011, 0.004), (0.866, 0.5), (8.0, 3.0), (60.0, 40.0)]:
        Phi, V, Dl = point_nodes(p, R, z)
        V = V[::23].copy()
        n = len(V)
        Fr, Fz = np.zeros(n), np.zeros(n)
        H.fjh_fast_actions(p.nr, P(p.ar), p.npoly, P(phil), P(Pr), p.ngrid, P(E), P(D), n, P(np.full(n, R)),
                           P(np.full(n, z)), P(np.full(n, Phi)), P(V[:, 0]), P(V[:, 1]), P(V[:, 2]), 64, P(gx),
                           P(gw), 32, P(vx), P(vw), P(Fr), P(Fz), None, None)
        for i in range(n):
            Jr, Jz = p.actions(p.deltafn(0.5 * np.sum(V[i] ** 2) + Phi), R, z, V[i, 0], V[i, 1], V[i, 2], Phi)
            if Jr > 1e30:
                assert Fr[i] > 1e30
                continue
            s = Jr + Jz + abs(R * V[i, 1])
            er.append(abs(Fr[i] - Jr) / s)
            ez.append(abs(Fz[i] - Jz) / s)
    er, ez = np.array(er), np.array(ez)
    print("Jr: median %.2e max %.2e; Jz: median %.2e max %.2e" % (np.median(er), er.max(), np.median(ez), ez.max()))
    assert np.median(er) < 1e-4 and np.median(ez) < 1e-5
    assert er.max() < 5e-3 and ez.max() < 2e-3
