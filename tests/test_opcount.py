"""The operation-counting restatement (oracle/opcount.cpp) that prices the
QAGS kernel's roofline.

* A potential evaluation counts what its branch does: interior radius
  without the quadratic correction 15 P - 5 operations (P = npoly), inside
  the first ten cells 22 P + 4, and its value is the oracle's fjo_phi.
* The counted first QK21 rule of each action integral is the oracle's own
  rule (a dropped or mistyped term in the restated integrand changes it).
* The counts over a few grid points of the Hernquist workload: the hoisted
  form is cheaper, the integrand evaluations dominate, and the figure per
  rule falls between the cheapest and the dearest branch.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def counting():
    import importlib.util
    spec = importlib.util.spec_from_file_location("count_ops", os.path.join(ROOT, "oracle", "count_ops.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    lib = m.load()
    lib.fjc_phi_ops.restype = C.c_longlong
    lib.fjc_phi_ops.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double), C.c_double,
                                C.c_double, C.POINTER(C.c_double)]
    return m, lib


@pytest.fixture(scope="module")
def pot():
    p = O.guess_model(40, 0.01, 100.0, 5, 20)[0]
    p.set_quad(0)
    return p


def test_potential_evaluation_counts_and_value(counting, pot):
    _, lib = counting
    P = pot.npoly
    ar = O.arr(pot.ar)
    phil = O.arr(pot.tables()[0])
    v = C.c_double()
    # interior radius beyond the first ten cells: no quadratic correction
    r = 0.5 * (ar[20] + ar[21])
    n = lib.fjc_phi_ops(pot.nr, O.ptr(ar), P, O.ptr(phil), 0.6 * r, 0.8 * r, C.byref(v))
    assert n == 15 * P - 5
    assert v.value == pytest.approx(pot.phi(0.6 * r, 0.8 * r), rel=1e-14)
    # inside the first ten cells: linear interpolation plus the quadratic correction
    r = 0.5 * (ar[3] + ar[4])
    n = lib.fjc_phi_ops(pot.nr, O.ptr(ar), P, O.ptr(phil), 0.6 * r, 0.8 * r, C.byref(v))
    assert n == 21 * P + 4
    assert v.value == pytest.approx(pot.phi(0.6 * r, 0.8 * r), rel=1e-14)
    # beyond the last radius: the Keplerian-power tail
    r = 3 * ar[-1]
    n = lib.fjc_phi_ops(pot.nr, O.ptr(ar), P, O.ptr(phil), 0.6 * r, 0.8 * r, C.byref(v))
    assert n == 12 + 9 * (P - 2) + 3 * P + 2 * (P - 1)
    assert v.value == pytest.approx(pot.phi(0.6 * r, 0.8 * r), rel=1e-13)


def test_counted_rules_are_the_oracle_rules(counting, pot):
    m, lib = counting
    R, z = m.grid_points(pot, 397)          # points 0 and 397: r = 0.01 and r ~ 0.9
    ref = m.count(pot, R, z, 0, lib)
    hoi = m.count(pot, R, z, 1, lib)
    assert ref["bound_nodes"] > 10000
    assert ref["first_rule_rel_diff"] < 1e-6
    assert hoi["first_rule_rel_diff"] < 1e-6
    assert ref["rules_ju"] == hoi["rules_ju"] and ref["rules_jv"] == hoi["rules_jv"]
    assert hoi["ops_per_rule"] < ref["ops_per_rule"]
    assert ref["ops_integrand"] > 0.9 * (ref["ops_ju"] + ref["ops_jv"])
    P = pot.npoly
    lo = 21 * (20 + 15 * P - 5) + 150           # hoisted Ju evaluation at the cheapest branch
    hi = 21 * (36 + 2 * (21 * P + 4) + 40) + 200  # Jv as written, two corrected evaluations
    assert lo < hoi["ops_per_rule"] < ref["ops_per_rule"] < hi
