"""ctypes binding of the CPU restatement (libfjoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (fjmodels_amd)
never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libfjoracle.so")

_dp = C.POINTER(C.c_double)
FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
ODE = C.CFUNCTYPE(None, C.c_double, _dp, _dp, C.c_void_p)


class DFParams(C.Structure):
    _fields_ = [("A", C.c_double), ("B", C.c_double),
                ("dphi_h", C.c_double), ("dz_h", C.c_double),
                ("dphi_g", C.c_double), ("dz_g", C.c_double),
                ("mass", C.c_double), ("J0", C.c_double), ("chi", C.c_double),
                ("k", C.c_double), ("rot_form", C.c_int)]


class Counters(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in
                ("nodes", "bound_nodes", "phi_evals", "grad_evals", "qk21_calls", "brent_iters", "unbound_nodes")]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    v = C.c_void_p
    i = C.c_int
    d = C.c_double
    sig = {
        "fjo_pot_new": (v, [i, _dp, i, i, i]),
        "fjo_pot_free": (None, [v]),
        "fjo_pot_set_tables": (None, [v, _dp, _dp, _dp]),
        "fjo_pot_get_tables": (None, [v, _dp, _dp, _dp]),
        "fjo_pot_set_delta": (None, [v, _dp, _dp]),
        "fjo_pot_get_delta": (None, [v, _dp, _dp]),
        "fjo_pot_set_quad": (None, [v, i]),
        "fjo_log_grid": (None, [i, d, d, _dp]),
        "fjo_gauleg": (None, [i, d, d, _dp, _dp]),
        "fjo_evenlegend": (None, [i, d, _dp]),
        "fjo_legend": (None, [_dp, d, i]),
        "fjo_dlegend": (d, [d, i]),
        "fjo_linterp": (d, [_dp, _dp, i, d]),
        "fjo_roots": (d, [FN, v, d, d, C.POINTER(i)]),
        "fjo_qk21": (None, [FN, v, d, d, _dp, _dp, _dp, _dp]),
        "fjo_qk15": (None, [FN, v, d, d, _dp, _dp, _dp, _dp]),
        "fjo_qags": (i, [FN, v, d, d, d, d, i, _dp, _dp, C.POINTER(i), C.POINTER(i)]),
        "fjo_qagiu": (i, [FN, v, d, d, d, i, _dp, _dp, C.POINTER(i), C.POINTER(i)]),
        "fjo_evolve_apply": (i, [ODE, v, i, d, d, _dp, d, _dp, _dp]),
        "fjo_rk8pd_step": (None, [ODE, v, i, d, d, _dp, _dp, _dp]),
        "fjo_phi": (d, [v, d, d]),
        "fjo_dR": (d, [v, d, d]),
        "fjo_dz": (d, [v, d, d]),
        "fjo_guess_rhl": (None, [v, i, d, d, d, _dp]),
        "fjo_compute_phil": (i, [v, _dp, _dp, _dp, _dp]),
        "fjo_get_ylm": (None, [v] + [_dp] * 12),
        "fjo_deltafn": (d, [v, d]),
        "fjo_tabulate_delta": (i, [v]),
        "fjo_best_delta": (d, [v, d, d, d, C.POINTER(i)]),
        "fjo_actions": (i, [v, d, d, d, d, d, d, d, _dp, _dp]),
        "fjo_df_eval": (d, [v, C.POINTER(DFParams), _dp, _dp]),
        "fjo_df_hg": (d, [C.POINTER(DFParams), d, d, d]),
        "fjo_rhofDF": (d, [v, C.POINTER(DFParams), d, d, _dp, _dp, _dp, _dp, _dp, _dp]),
        "fjo_sigma_df": (d, [v, _dp, d, d]),
        "fjo_quad_tables": (None, [_dp, _dp, _dp, _dp]),
        "fjo_line_profile": (d, [v, C.POINTER(DFParams), _dp, d, d, d]),
        "fjo_vdist": (d, [v, C.POINTER(DFParams), C.c_int, d, d, d]),
        "fjo_compute_rhl_grid": (None, [v, C.POINTER(DFParams), i] + [_dp] * 6),
        "fjo_compute_new_phi": (i, [v, C.POINTER(DFParams), i] + [_dp] * 9),
        "fjo_vir2": (None, [v] + [_dp] * 6),
        "fjo_counters_get": (None, [C.POINTER(Counters)]),
        "fjo_counters_reset": (None, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def ptr(a):
    return a.ctypes.data_as(_dp)


def arr(a):
    return np.ascontiguousarray(a, dtype=np.float64)


MODELS = {"Hernquist": 0, "Isochrone": 1, "NFW": 2, "NFWext": 3}
# DPL slopes by model type (df/DF.cpp:40-42)
SLOPES = {"Hernquist": (1.667, 5.0), "Isochrone": (0.0, 5.0), "NFW": (1.667, 3.0)}


def df_params(model="Hernquist", dphi_h=0.55, dz_h=0.55, dphi_g=1.0, dz_g=1.0,
              mass=1.0, r0=1.0, chi=0.0, k=None, rot_form=0, A=None, B=None):
    a, b = SLOPES.get(model, (A, B))
    if A is not None:
        a = A
    if B is not None:
        b = B
    if k is None:
        k = 0.5 if rot_form == 1 else 1.0
    return DFParams(a, b, dphi_h, dz_h, dphi_g, dz_g, mass, float(np.sqrt(mass * r0)), chi, k, rot_form)


class Pot:
    """Grid + potential tables + Delta table (the reference's Grid.h globals)."""

    def __init__(self, ar, npoly, ngauss, ngrid=30):
        self.ar = arr(ar)
        self.nr, self.npoly, self.ngauss, self.ngrid = len(self.ar), npoly, ngauss, ngrid
        self.h = lib.fjo_pot_new(self.nr, ptr(self.ar), npoly, ngauss, ngrid)
        if not self.h:
            raise ValueError("fjo_pot_new: need nr >= 11, 1 <= npoly <= 64")

    def __del__(self):
        try:
            lib.fjo_pot_free(self.h)
        except Exception:
            pass

    def set_tables(self, phil, Pr, Pr2):
        phil, Pr, Pr2 = arr(phil), arr(Pr), arr(Pr2)
        lib.fjo_pot_set_tables(self.h, ptr(phil), ptr(Pr), ptr(Pr2))

    def tables(self):
        s = (self.nr, self.npoly)
        a, b, c = np.zeros(s), np.zeros(s), np.zeros(s)
        lib.fjo_pot_get_tables(self.h, ptr(a), ptr(b), ptr(c))
        return a, b, c

    def set_delta(self, E, D):
        E, D = arr(E), arr(D)
        lib.fjo_pot_set_delta(self.h, ptr(E), ptr(D))

    def delta(self):
        E, D = np.zeros(self.ngrid), np.zeros(self.ngrid)
        lib.fjo_pot_get_delta(self.h, ptr(E), ptr(D))
        return E, D

    def set_quad(self, preset):
        lib.fjo_pot_set_quad(self.h, preset)

    def phi(self, R, z):
        return lib.fjo_phi(self.h, R, z)

    def dR(self, R, z):
        return lib.fjo_dR(self.h, R, z)

    def dz(self, R, z):
        return lib.fjo_dz(self.h, R, z)

    def guess_rhl(self, model="Hernquist", mass=1.0, r0=1.0, q=1.0):
        out = np.zeros((self.nr, self.npoly))
        lib.fjo_guess_rhl(self.h, MODELS[model], mass, r0, q, ptr(out))
        return out

    def compute_phil(self, rhl):
        rhl = arr(rhl)
        s = (self.nr, self.npoly)
        a, b, c = np.zeros(s), np.zeros(s), np.zeros(s)
        rc = lib.fjo_compute_phil(self.h, ptr(rhl), ptr(a), ptr(b), ptr(c))
        if rc:
            raise RuntimeError("compute_phil: NaN in external integral")
        return a, b, c

    def tabulate_delta(self):
        return lib.fjo_tabulate_delta(self.h)

    def deltafn(self, E):
        return lib.fjo_deltafn(self.h, E)

    def actions(self, Delta, R, z, vR, vphi, vz, Phi0=None):
        if Phi0 is None:
            Phi0 = self.phi(R, z)
        Jr, Jz = C.c_double(), C.c_double()
        lib.fjo_actions(self.h, Delta, R, z, Phi0, vR, vphi, vz, C.byref(Jr), C.byref(Jz))
        return Jr.value, Jz.value

    def df_eval(self, df, x, v):
        """df() (DF.cpp:78-124) at x = (R, z, Phi), velocity v."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        return lib.fjo_df_eval(self.h, C.byref(df), ptr(x), ptr(v))

    def sigma_df(self, rhl, R, z):
        rhl = np.ascontiguousarray(rhl, dtype=np.float64)
        return lib.fjo_sigma_df(self.h, ptr(rhl), R, z)

    def line_profile(self, df, rhl, R, z, vx):
        rhl = np.ascontiguousarray(rhl, dtype=np.float64)
        return lib.fjo_line_profile(self.h, C.byref(df), ptr(rhl), R, z, vx)

    def velocity_distribution(self, df, which, R, z, v):
        return lib.fjo_vdist(self.h, C.byref(df), which, R, z, v)

    def rhofDF(self, df, R, z):
        out = np.zeros(6)
        v = [C.c_double() for _ in range(5)]
        dens = lib.fjo_rhofDF(self.h, C.byref(df), R, z, ptr(out), *[C.byref(x) for x in v])
        return dens, out, [x.value for x in v]

    def grid_moments(self, df, nthreads=0):
        s = (self.nr, self.ngauss)
        outs = [np.zeros(s) for _ in range(6)]
        lib.fjo_compute_rhl_grid(self.h, C.byref(df), nthreads, *[ptr(o) for o in outs])
        return outs

    def compute_new_phi(self, df, nthreads=0):
        s = (self.nr, self.npoly)
        outs = [np.zeros(s) for _ in range(9)]
        rc = lib.fjo_compute_new_phi(self.h, C.byref(df), nthreads, *[ptr(o) for o in outs])
        if rc:
            raise RuntimeError("compute_new_phi failed")
        return dict(zip(["rhl", "vrotl", "sigRl", "sigpl", "sigzl", "sigRzl", "phil", "Pr", "Pr2"], outs))

    def vir2(self, rhl, vrotl, sigRl, sigpl, sigzl):
        out = np.zeros(11)
        a = [arr(x) for x in (rhl, vrotl, sigRl, sigpl, sigzl)]
        lib.fjo_vir2(self.h, *[ptr(x) for x in a], ptr(out))
        return out


def log_grid(n, rmin, rmax):
    a = np.zeros(n)
    lib.fjo_log_grid(n, rmin, rmax, ptr(a))
    return a


def gauleg(n, x1=0.0, x2=1.0):
    x, w = np.zeros(n), np.zeros(n)
    lib.fjo_gauleg(n, x1, x2, ptr(x), ptr(w))
    return x, w


def counters():
    c = Counters()
    lib.fjo_counters_get(C.byref(c))
    return {n: getattr(c, n) for n, _ in Counters._fields_}


def reset_counters():
    lib.fjo_counters_reset()


def quad_tables():
    """(xpt11, xw11, xpt01, xw01) of include/GaussPts.h:76-136."""
    t = [np.zeros(22) for _ in range(4)]
    lib.fjo_quad_tables(*[ptr(x) for x in t])
    return t


def guess_model(nr, rmin, rmax, npoly, ngauss, model="Hernquist", mass=1.0, r0=1.0, q=1.0,
                ngrid=30, delta=True):
    """Grid + guess potential (+ Delta table), as fJmodels.cpp:86-127 sets up."""
    p = Pot(log_grid(nr, rmin, rmax), npoly, ngauss, ngrid)
    rhl = p.guess_rhl(model, mass, r0, q)
    phil, Pr, Pr2 = p.compute_phil(rhl)
    p.set_tables(phil, Pr, Pr2)
    if delta:
        p.tabulate_delta()
    return p, rhl
