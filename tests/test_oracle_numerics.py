"""Pins of the oracle's numerical building blocks (CPU only).

Each check is chosen so that a plausible slip (a wrong table entry, a dropped
term, a wrong sign or index) fails it:
  * quadrature tables vs their closed forms (include/GaussPts.h:76-136);
  * gauleg vs numpy.polynomial.legendre.leggauss (UtilsLeg.h:18-39);
  * Legendre recurrences vs scipy.special (UtilsLeg.h:43-79);
  * Gauss-Kronrod rules: exactness on monomials (QUADPACK dqk21/dqk15);
  * QAGS vs scipy.integrate.quad, which wraps QUADPACK dqagse: same result,
    same number of subintervals and evaluations;
  * QAGIU vs scipy's dqagie (different transform arithmetic, same algorithm);
  * Brent (GSL iterate form, Utils.h:134-167) vs scipy.optimize.brentq and the
    wrapper's |F| < 5e-4 early return;
  * RK8PD coefficients: row sums, quadrature order conditions, and 8th-order
    convergence on a linear oscillator.
"""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest
import scipy.integrate as si
import scipy.optimize as so
import scipy.special as sp

from oracle import oracle as O

L = O.lib

XPT11 = [-1., -0.50118723, -0.25118864, -0.12589254, -0.06309573, -0.03162278, -0.01584893,
         -0.00794328, -0.00398107, -0.00199526, -0.001, 0.001, 0.00199526, 0.00398107,
         0.00794328, 0.01584893, 0.03162278, 0.06309573, 0.12589254, 0.25118864, 0.50118723, 1.]
XW11 = [0.49881277, 0.37440568, 0.18764735, 0.09404645, 0.04713488, 0.0236234, 0.01183975,
        0.00593393, 0.00297401, 0.00149054, 0.00149763, 0.00149763, 0.00149054, 0.00297401,
        0.00593393, 0.01183975, 0.0236234, 0.04713488, 0.09404645, 0.18764735, 0.37440568, 0.49881277]


def test_log22_rule_closed_form():
    """xpt11 = +-10^(-3+0.3k), k = 0..10 (8 decimals); weights are the trapezoid
    half-gaps with the +-1 endpoints carrying the full last gap."""
    k = np.arange(11)
    pos = 10.0 ** (-3 + 0.3 * k)
    x = np.concatenate([-pos[::-1], pos])
    assert np.max(np.abs(np.array(XPT11) - x)) < 5e-9
    xs = np.array(XPT11)
    w = np.zeros(22)
    for i in range(1, 21):
        w[i] = 0.5 * (xs[i + 1] - xs[i - 1])
    w[0] = xs[1] - xs[0]
    w[21] = xs[21] - xs[20]
    assert np.max(np.abs(np.array(XW11) - w)) < 1e-8
    assert abs(sum(XW11) - 2.4988) < 1e-3


def test_tanh_sinh_rule_closed_form():
    """xpt01/xw01: x = (1 + tanh(pi/2 sinh t))/2, t = k h, h = 2/11, k = -11..10."""
    src = open(O._HERE + "/fjoracle.c").read()
    start = src.index("static const double xpt01[22]")
    block = src[start:src.index("};", src.index("static const double xw01[22]")) + 2]
    nums = [float(v) for v in __import__("re").findall(r"[0-9.]+e?-?[0-9]*", block.split("=", 1)[1]) if v != "22"]
    xs, ws = np.array(nums[:22]), np.array(nums[-22:])
    h = 2.0 / 11.0
    t = h * np.arange(-11, 11)
    s = np.pi / 2 * np.sinh(t)
    x = 0.5 * (1 + np.tanh(s))
    w = 0.5 * h * (np.pi / 2) * np.cosh(t) / np.cosh(s) ** 2
    assert np.max(np.abs(xs - x)) < 1e-12
    assert np.max(np.abs(ws - w)) < 5e-12


@pytest.mark.parametrize("n", [4, 20, 30, 32])
def test_gauleg_matches_leggauss(n):
    x, w = O.gauleg(n, 0.0, 1.0)
    gx, gw = np.polynomial.legendre.leggauss(n)
    # gauleg fills x ascending on [0,1]
    assert np.max(np.abs(x - 0.5 * (gx + 1))) < 1e-14
    # weights use the derivative from the last Newton iterate (stop at |dz| <= 3e-11)
    assert np.max(np.abs(w - 0.5 * gw)) < 5e-11


def test_legendre_recurrences():
    c = np.linspace(-0.999, 0.999, 41)
    pol = np.zeros(12)
    for cc in c:
        L.fjo_evenlegend(12, cc, O.ptr(pol))
        ref = np.array([sp.eval_legendre(2 * k, cc) for k in range(12)])
        assert np.max(np.abs(pol - ref)) < 1e-13
        allp = np.zeros(23)
        L.fjo_legend(O.ptr(allp), cc, 22)
        assert np.max(np.abs(allp - [sp.eval_legendre(i, cc) for i in range(23)])) < 1e-13
        for n in (0, 2, 8, 22):
            der = np.polynomial.legendre.Legendre.basis(n).deriv()(cc)
            assert abs(L.fjo_dlegend(cc, n) - der) < 1e-10 * max(1, abs(der))


def _qk(which, f, a, b):
    out = [C.c_double() for _ in range(4)]
    fn = O.FN(lambda x, p: f(x))
    getattr(L, which)(fn, None, a, b, *[C.byref(o) for o in out])
    return [o.value for o in out]


@pytest.mark.parametrize("which,deg", [("fjo_qk21", 31), ("fjo_qk15", 22)])
def test_kronrod_exactness(which, deg):
    a, b = -0.3, 1.7
    for k in range(deg + 1):
        r, err, _, _ = _qk(which, lambda x: x ** k, a, b)
        exact = (b ** (k + 1) - a ** (k + 1)) / (k + 1)
        assert abs(r - exact) <= 1e-13 * max(1.0, abs(exact)), (k, r, exact)
    # one degree beyond exactness is not integrated exactly
    k = deg + 2
    r, *_ = _qk(which, lambda x: x ** k, a, b)
    assert abs(r - (b ** (k + 1) - a ** (k + 1)) / (k + 1)) > 1e-12


def _qags(f, a, b, epsabs, epsrel, limit=1000):
    res, err, nev, last = C.c_double(), C.c_double(), C.c_int(), C.c_int()
    fn = O.FN(lambda x, p: f(x))
    ier = L.fjo_qags(fn, None, a, b, epsabs, epsrel, limit, C.byref(res), C.byref(err), C.byref(nev), C.byref(last))
    return res.value, err.value, ier, nev.value, last.value


INTEGRANDS = {
    "smooth": (lambda x: np.exp(-x) * np.cos(3 * x), 0.0, 2.0),
    "sqrt_endpoint": (lambda x: np.sqrt(max(0.0, 1.0 - x * x)) * x, 0.0, 1.0),
    "log_singular": (lambda x: np.log(x) / np.sqrt(x) if x > 0 else 0.0, 0.0, 1.0),
    "kink": (lambda x: abs(x - 0.3137) ** 0.5, 0.0, 1.0),
    "piecewise_linear": (lambda x: x ** 2 * np.interp(x, [0, .1, .25, .6, 1.3, 2.0], [5, 3, 2.5, .4, .1, 0]), 0.0, 2.0),
}


@pytest.mark.parametrize("name", sorted(INTEGRANDS))
@pytest.mark.parametrize("tol", [(1e-5, 1e-3), (1e-10, 1e-10)])
def test_qags_matches_quadpack(name, tol):
    f, a, b = INTEGRANDS[name]
    r, e, ier, nev, last = _qags(f, a, b, *tol)
    ref = si.quad(f, a, b, epsabs=tol[0], epsrel=tol[1], limit=1000, full_output=1)
    rr, re, info = ref[0], ref[1], ref[2]
    assert nev == info["neval"], (nev, info["neval"])
    assert last == info["last"]
    assert abs(r - rr) <= 1e-15 * max(1.0, abs(rr)) * 4
    assert abs(e - re) <= 1e-12 * max(abs(re), 1e-300) + 1e-300


def test_qagiu_matches_quadpack():
    def f(x):
        return 1.0 / (1 + x) ** 2.5 * (1 + 0.3 * np.sin(x))
    for a in (0.0, 0.5, 7.0):
        for tol in ((1e-4, 1e-3), (1e-8, 1e-8)):
            res, err, nev, last = C.c_double(), C.c_double(), C.c_int(), C.c_int()
            fn = O.FN(lambda x, p: f(x))
            ier = L.fjo_qagiu(fn, None, a, tol[0], tol[1], 1000, C.byref(res), C.byref(err), C.byref(nev), C.byref(last))
            ref = si.quad(f, a, np.inf, epsabs=tol[0], epsrel=tol[1], limit=1000, full_output=1)
            assert abs(res.value - ref[0]) <= 1e-12 * abs(ref[0])
            assert (nev.value, last.value) == (ref[2]["neval"], ref[2]["last"])
            assert (ier == 0) == (len(ref) == 3)   # scipy appends a message when ier > 0


@pytest.mark.parametrize("f,lo,hi", [
    (lambda x: np.cos(x) - x, 0.0, 1.5),
    (lambda x: x ** 3 - 2 * x - 5, 1.0, 3.0),
    (lambda x: np.exp(x) - 10, -1.0, 7.0),
    (lambda x: (x - 0.3) * (1 + x * x) * 1e-3 + 0.0, -2.0, 3.0),
])
def test_brent_matches_brentq(f, lo, hi):
    it = C.c_int()
    r = L.fjo_roots(O.FN(lambda x, p: f(x)), None, lo, hi, C.byref(it))
    if abs(f(lo)) < 5e-4 or abs(f(hi)) < 5e-4:
        pytest.skip("endpoint return")
    ref = so.brentq(f, lo, hi, xtol=1e-14, rtol=1e-15)
    assert abs(r - ref) < 1e-8
    assert it.value > 0


def test_brent_wrapper_endpoint_return():
    """wrp_GSLroots returns an endpoint when |F| < 5e-4 there (Utils.h:163-164)."""
    f = lambda x: x - 0.0004
    assert L.fjo_roots(O.FN(lambda x, p: f(x)), None, 0.0, 1.0, None) == 0.0
    g = lambda x: x - 0.9997
    assert L.fjo_roots(O.FN(lambda x, p: g(x)), None, 0.0, 1.0, None) == 1.0


def _rk_tables():
    src = open(O._HERE + "/fjoracle.c").read()

    def grab(name):
        s = src[src.index(name):]
        s = s[s.index("{") + 1:]
        depth, out, cur = 1, [], ""
        for ch in s:
            if ch == "{":
                depth += 1
            elif ch == "}":
                depth -= 1
                if depth == 0:
                    break
            cur += ch
        return cur

    def frac(tok):
        tok = tok.strip()
        if not tok:
            return None
        if "/" in tok:
            a, b = tok.split("/")
            return Fraction(a.strip().rstrip(".0") or "0") / Fraction(b.strip().rstrip(".0") or "1") if False else \
                Fraction(int(float(a)), 1) / Fraction(int(float(b)), 1)
        return Fraction(tok)

    c = [frac(t) for t in grab("rk_c[13] =").split(",") if t.strip()]
    rows = []
    for r in grab("rk_a[13][12] =").split("}")[:-1]:
        r = r.replace("{", "")
        rows.append([frac(t) for t in r.split(",") if t.strip()])
    b8 = [frac(t) for t in grab("rk_b8[13] =").split(",") if t.strip()]
    b7 = [frac(t) for t in grab("rk_b7[13] =").split(",") if t.strip()]
    return c, rows, b8, b7


def test_rk8pd_coefficients():
    c, A, b8, b7 = _rk_tables()
    assert len(c) == 13 and len(A) == 13 and len(b8) == 13 and len(b7) == 13
    for i in range(1, 13):
        assert abs(float(sum(A[i]) - c[i])) < 1e-9, i
    assert abs(float(sum(b8)) - 1) < 1e-12 and abs(float(sum(b7)) - 1) < 1e-12
    for k in range(8):   # bushy-tree conditions sum b c^k = 1/(k+1)
        assert abs(float(sum(b * ci ** k for b, ci in zip(b8, c))) - 1.0 / (k + 1)) < 1e-9
    for k in range(7):
        assert abs(float(sum(b * ci ** k for b, ci in zip(b7, c))) - 1.0 / (k + 1)) < 1e-9


def test_rk8pd_convergence_order():
    """Fixed-step RK8PD on the nonlinear logistic equation y' = y(1-y):
    halving h divides the global error by ~2^8 (a slipped coefficient drops
    the order to a few and the ratio far below 100)."""
    def rhs(t, y, f, p):
        f[0] = y[0] * (1 - y[0])
    fn = O.ODE(rhs)
    T, y0 = 6.0, 0.05

    def run(nsteps):
        y = np.array([y0])
        yerr = np.zeros(1)
        h = T / nsteps
        d = np.zeros(1)
        for s in range(nsteps):
            d[0] = y[0] * (1 - y[0])
            L.fjo_rk8pd_step(fn, None, 1, s * h, h, O.ptr(y), O.ptr(yerr), O.ptr(d))
        exact = 1.0 / (1.0 + np.exp(-T) * (1.0 / y0 - 1.0))
        return abs(y[0] - exact)
    errs = [run(n) for n in (3, 6, 12)]
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(150 < r < 1200 for r in ratios), ratios


def test_evolve_reaches_tolerance():
    """The adaptive evolve loop with y-control 1e-10/1e-9 tracks the oscillator."""
    def rhs(t, y, f, p):
        f[0] = y[1]
        f[1] = -y[0]
    fn = O.ODE(rhs)
    y = np.array([1.0, 0.0])
    t, h = C.c_double(0.0), C.c_double(2e-3)
    steps = 0
    while t.value < 5.0 and steps < 1000:
        L.fjo_evolve_apply(fn, None, 2, 1e-10, 1e-9, C.byref(t), 5.0, C.byref(h), O.ptr(y))
        steps += 1
    assert t.value == 5.0
    assert abs(y[0] - np.cos(5.0)) < 1e-8 and abs(y[1] + np.sin(5.0)) < 1e-8
    assert steps < 60   # step size grows from 2e-3
