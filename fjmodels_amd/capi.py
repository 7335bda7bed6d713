"""ctypes binding of libfjmodels.so (include/fjmodels.h).

Plumbing for tests and bench.py: every call goes through the C ABI into the
gfx950 kernels.  There is no CPU fallback; a missing library or device raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfjmodels.so")

_dp = C.POINTER(C.c_double)

FJ_QUAD_LOG22, FJ_QUAD_TANHSINH22 = 0, 1
FJ_POISSON_QUADPACK, FJ_POISSON_BLOCKSCAN = 0, 1
FJ_PREC_FP64, FJ_PREC_MIXED, FJ_PREC_FAST = 0, 1, 2
NOALLOC_MAX_POINTS = 4   # FJ_NOALLOC_MAX_POINTS of include/fjmodels.h
GUESS = {"Hernquist": 0, "Isochrone": 1, "NFW": 2, "NFWext": 3}
SLOPES = {"Hernquist": (1.667, 5.0), "Isochrone": (0.0, 5.0), "NFW": (1.667, 3.0)}

# Every symbol include/fjmodels.h declares.
EXPORTS = [
    "fj_ctx_create", "fj_ctx_destroy", "fj_last_error", "fj_set_grid", "fj_set_quad_rule",
    "fj_set_precision", "fj_set_fast_order", "fj_set_stream", "fj_set_potential", "fj_get_potential", "fj_set_delta_table",
    "fj_get_delta_table", "fj_set_df", "fj_guess_rhl", "fj_poisson", "fj_moments", "fj_rhofDF",
    "fj_rhofDF_noalloc",
    "fj_project_legendre", "fj_compute_new_phi", "fj_tabulate_delta", "fj_virial",
    "fj_eval_potential", "fj_actions", "fj_df_points", "fj_sigma_df", "fj_line_profile",
    "fj_velocity_distribution", "fj_get_stats",
    "fj_surrogate_init", "fj_surrogate_collect", "fj_surrogate_train", "fj_surrogate_mode",
    "fj_surrogate_get_stats", "fj_surrogate_params", "fj_surrogate_forward", "fj_surrogate_gradient",
    "fj_batch_create", "fj_batch_destroy", "fj_batch_setup", "fj_batch_autotune", "fj_batch_iterate",
    "fj_batch_get", "fj_batch_set", "fj_batch_get_stats",
]

MLP_NPARAM = 5 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 2 + 2   # 34,050


class DFParams(C.Structure):
    _fields_ = [("A", C.c_double), ("B", C.c_double),
                ("dphi_h", C.c_double), ("dz_h", C.c_double),
                ("dphi_g", C.c_double), ("dz_g", C.c_double),
                ("mass", C.c_double), ("J0", C.c_double), ("chi", C.c_double),
                ("k", C.c_double), ("rot_form", C.c_int)]


class Stats(C.Structure):
    _fields_ = ([(n, C.c_double) for n in ("ms_points", "ms_setup", "ms_integrals", "ms_overflow",
                                           "ms_moments", "ms_project", "ms_poisson", "ms_total")] +
                [(n, C.c_longlong) for n in ("points", "nodes", "bound_nodes", "tasks", "qk_calls",
                                             "overflow_tasks", "brent_iters")] +
                [("integral_lanes", C.c_int), ("setup_lanes", C.c_int)] +
                [(n, C.c_longlong) for n in ("unbound_nodes", "bracket_failures", "nonfinite_nodes",
                                             "phi_evals", "grad_evals")])


class SurrogateStats(C.Structure):
    _fields_ = ([(n, C.c_longlong) for n in ("n_train", "nodes_sur", "fallback", "judged")] +
                [(n, C.c_double) for n in ("judge_err_r_med", "judge_err_r_p95", "judge_err_z_med",
                                           "judge_err_z_p95", "ms_surrogate", "ms_train", "last_loss",
                                           "judge_werr_r_med", "judge_werr_z_med")])


class ModelSpec(C.Structure):
    _fields_ = [("nr", C.c_int), ("ar", _dp), ("ntheta", C.c_int), ("npoly", C.c_int), ("guess", C.c_int),
                ("mass", C.c_double), ("r0", C.c_double), ("q", C.c_double), ("df", DFParams), ("quad", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("ms_setup", C.c_double), ("ms_tune", C.c_double), ("ms_last_iter", C.c_double),
                ("nodes", C.c_longlong), ("points", C.c_longlong), ("setup_block", C.c_int),
                ("integ_blocks", C.c_int), ("graph", C.c_int)]


def df_params(model="Hernquist", dphi_h=0.55, dz_h=0.55, dphi_g=1.0, dz_g=1.0,
              mass=1.0, r0=1.0, chi=0.0, k=None, rot_form=0, A=None, B=None):
    """setDF (df/DF.cpp:36-63): slopes from the model type, J0 = sqrt(M r0)."""
    a, b = SLOPES.get(model, (A, B))
    if A is not None:
        a = A
    if B is not None:
        b = B
    if k is None:
        k = 0.5 if rot_form == 1 else 1.0
    return DFParams(a, b, dphi_h, dz_h, dphi_g, dz_g, mass, float(np.sqrt(mass * r0)), chi, k, rot_form)


def load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    v, i, d = C.c_void_p, C.c_int, C.c_double
    sig = {
        "fj_ctx_create": [i, C.POINTER(v)],
        "fj_ctx_destroy": [v],
        "fj_set_grid": [v, i, _dp, i, i, i],
        "fj_set_quad_rule": [v, i],
        "fj_set_precision": [v, i],
        "fj_set_stream": [v, v],
        "fj_set_fast_order": [v, i],
        "fj_set_potential": [v, _dp, _dp, _dp],
        "fj_get_potential": [v, _dp, _dp, _dp],
        "fj_set_delta_table": [v, _dp, _dp],
        "fj_get_delta_table": [v, _dp, _dp],
        "fj_set_df": [v, C.POINTER(DFParams)],
        "fj_guess_rhl": [v, i, d, d, d, _dp],
        "fj_poisson": [v, i, _dp, _dp, _dp, _dp],
        "fj_moments": [v] + [_dp] * 6,
        "fj_rhofDF": [v, i, _dp, _dp, _dp] + [_dp] * 6,
        "fj_rhofDF_noalloc": [v, i, _dp, _dp, _dp] + [_dp] * 6,
        "fj_project_legendre": [v] + [_dp] * 12,
        "fj_compute_new_phi": [v, i] + [_dp] * 9,
        "fj_tabulate_delta": [v, _dp, _dp],
        "fj_virial": [v] + [_dp] * 6,
        "fj_eval_potential": [v, i, _dp, _dp, _dp, _dp, _dp],
        "fj_actions": [v, i] + [_dp] * 8,
        "fj_df_points": [v, i] + [_dp] * 6,
        "fj_sigma_df": [v, _dp, i, _dp, _dp, _dp],
        "fj_line_profile": [v, _dp, d, d, i, _dp, _dp],
        "fj_velocity_distribution": [v, i, d, d, i, _dp, _dp],
        "fj_get_stats": [v, C.POINTER(Stats)],
        "fj_surrogate_init": [v, C.c_ulonglong, C.c_longlong],
        "fj_surrogate_collect": [v, d],
        "fj_surrogate_train": [v, i, i, d, C.POINTER(d)],
        "fj_surrogate_mode": [v, i, d, d],
        "fj_surrogate_get_stats": [v, C.POINTER(SurrogateStats)],
        "fj_surrogate_params": [v, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "fj_surrogate_forward": [v, C.c_longlong, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "fj_surrogate_gradient": [v, C.c_longlong, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.POINTER(d)],
        "fj_batch_create": [v, i, C.POINTER(ModelSpec), C.POINTER(v)],
        "fj_batch_destroy": [v],
        "fj_batch_setup": [v],
        "fj_batch_autotune": [v],
        "fj_batch_iterate": [v, i],
        "fj_batch_get": [v, i] + [_dp] * 10,
        "fj_batch_set": [v, i] + [_dp] * 5,
        "fj_batch_get_stats": [v, C.POINTER(BatchStats)],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.restype = i
        f.argtypes = args
    lib.fj_last_error.restype = C.c_char_p
    lib.fj_last_error.argtypes = [v]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _a(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class FJError(RuntimeError):
    pass


class Context:
    """One fj_ctx on a HIP device (the reference's Grid.h/DF.cpp globals)."""

    def __init__(self, device=0):
        self.lib = lib()
        h = C.c_void_p()
        rc = self.lib.fj_ctx_create(device, C.byref(h))
        if rc:
            raise FJError(f"fj_ctx_create failed ({rc}): no HIP device {device}?")
        self.h = h
        self.nr = self.ntheta = self.npoly = self.ngrid = 0
        self.precision = FJ_PREC_FP64

    def close(self):
        if self.h:
            self.lib.fj_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise FJError(f"{what}: {self.lib.fj_last_error(self.h).decode()} (code {rc})")

    def set_grid(self, ar, ntheta, npoly, ngrid=30):
        self.ar = _a(ar)
        self.nr, self.ntheta, self.npoly, self.ngrid = len(self.ar), ntheta, npoly, ngrid
        self._chk(self.lib.fj_set_grid(self.h, self.nr, _p(self.ar), ntheta, npoly, ngrid), "fj_set_grid")

    def set_quad(self, preset):
        self._chk(self.lib.fj_set_quad_rule(self.h, preset), "fj_set_quad_rule")

    def set_precision(self, mode):
        self._chk(self.lib.fj_set_precision(self.h, mode), "fj_set_precision")
        self.precision = mode

    def set_stream(self, stream):
        """Order this context's calls on a caller's hipStream_t (an integer
        handle of the HIP runtime the library is bound to), or on its own
        stream (None)."""
        self._chk(self.lib.fj_set_stream(self.h, C.c_void_p(stream) if stream else None), "fj_set_stream")

    def set_fast_order(self, n):
        """Gauss-Legendre order of the FJ_PREC_FAST action integrals."""
        self._chk(self.lib.fj_set_fast_order(self.h, n), "fj_set_fast_order")

    def set_potential(self, phil, Pr, Pr2):
        a, b, c = _a(phil), _a(Pr), _a(Pr2)
        self._chk(self.lib.fj_set_potential(self.h, _p(a), _p(b), _p(c)), "fj_set_potential")

    def potential(self):
        s = (self.nr, self.npoly)
        a, b, c = np.zeros(s), np.zeros(s), np.zeros(s)
        self._chk(self.lib.fj_get_potential(self.h, _p(a), _p(b), _p(c)), "fj_get_potential")
        return a, b, c

    def set_delta(self, E, D):
        E, D = _a(E), _a(D)
        self._chk(self.lib.fj_set_delta_table(self.h, _p(E), _p(D)), "fj_set_delta_table")

    def delta(self):
        E, D = np.zeros(self.ngrid), np.zeros(self.ngrid)
        self._chk(self.lib.fj_get_delta_table(self.h, _p(E), _p(D)), "fj_get_delta_table")
        return E, D

    def set_df(self, df):
        self._chk(self.lib.fj_set_df(self.h, C.byref(df)), "fj_set_df")

    def guess_rhl(self, model="Hernquist", mass=1.0, r0=1.0, q=1.0):
        out = np.zeros((self.nr, self.npoly))
        self._chk(self.lib.fj_guess_rhl(self.h, GUESS[model], mass, r0, q, _p(out)), "fj_guess_rhl")
        return out

    def poisson(self, rhl, mode=FJ_POISSON_QUADPACK):
        rhl = _a(rhl)
        s = (self.nr, self.npoly)
        a, b, c = np.zeros(s), np.zeros(s), np.zeros(s)
        self._chk(self.lib.fj_poisson(self.h, mode, _p(rhl), _p(a), _p(b), _p(c)), "fj_poisson")
        return a, b, c

    def moments(self):
        s = (self.nr, self.ntheta)
        o = [np.zeros(s) for _ in range(6)]
        self._chk(self.lib.fj_moments(self.h, *[_p(x) for x in o]), "fj_moments")
        return o

    def rhofDF(self, R, z):
        R, z = _a(R), _a(z)
        n = len(R)
        out6 = np.zeros((n, 6))
        o = [np.zeros(n) for _ in range(6)]
        self._chk(self.lib.fj_rhofDF(self.h, n, _p(R), _p(z), _p(out6), *[_p(x) for x in o]), "fj_rhofDF")
        return out6, o

    def rhofDF_noalloc(self, R, z):
        """fj_rhofDF_noalloc: rhofDF at n <= NOALLOC_MAX_POINTS points without device allocation."""
        R, z = _a(R), _a(z)
        n = len(R)
        out6 = np.zeros((n, 6))
        o = [np.zeros(n) for _ in range(6)]
        self._chk(self.lib.fj_rhofDF_noalloc(self.h, n, _p(R), _p(z), _p(out6), *[_p(x) for x in o]),
                  "fj_rhofDF_noalloc")
        return out6, o

    def project(self, rho, vrot, sigR, sigp, sigz, sigRz):
        ins = [_a(x) for x in (rho, vrot, sigR, sigp, sigz, sigRz)]
        s = (self.nr, self.npoly)
        o = [np.zeros(s) for _ in range(6)]
        self._chk(self.lib.fj_project_legendre(self.h, *[_p(x) for x in ins], *[_p(x) for x in o]),
                  "fj_project_legendre")
        return dict(zip(["rhl", "vrotl", "sigRl", "sigpl", "sigzl", "sigRzl"], o))

    def compute_new_phi(self, mode=FJ_POISSON_QUADPACK):
        s = (self.nr, self.npoly)
        o = [np.zeros(s) for _ in range(9)]
        # C order: rhl, sigRl, sigpl, sigzl, sigRzl, vrotl, philP, PrP, Pr2P
        self._chk(self.lib.fj_compute_new_phi(self.h, mode, *[_p(x) for x in o]), "fj_compute_new_phi")
        return dict(zip(["rhl", "sigRl", "sigpl", "sigzl", "sigRzl", "vrotl", "phil", "Pr", "Pr2"], o))

    def tabulate_delta(self):
        E, D = np.zeros(self.ngrid), np.zeros(self.ngrid)
        self._chk(self.lib.fj_tabulate_delta(self.h, _p(E), _p(D)), "fj_tabulate_delta")
        return E, D

    def virial(self, rhl, vrotl, sigRl, sigpl, sigzl):
        a = [_a(x) for x in (rhl, vrotl, sigRl, sigpl, sigzl)]
        out = np.zeros(11)
        self._chk(self.lib.fj_virial(self.h, *[_p(x) for x in a], _p(out)), "fj_virial")
        return out

    def eval_potential(self, R, z, grad=True):
        R, z = _a(R), _a(z)
        n = len(R)
        phi = np.zeros(n)
        dR = np.zeros(n) if grad else None
        dz = np.zeros(n) if grad else None
        self._chk(self.lib.fj_eval_potential(self.h, n, _p(R), _p(z), _p(phi), _p(dR), _p(dz)),
                  "fj_eval_potential")
        return (phi, dR, dz) if grad else phi

    def actions(self, R, z, vR, vphi, vz, Delta=None):
        R, z, vR, vphi, vz = (_a(x) for x in (R, z, vR, vphi, vz))
        n = len(R)
        D = None if Delta is None else _a(Delta)
        Jr, Jz = np.zeros(n), np.zeros(n)
        self._chk(self.lib.fj_actions(self.h, n, _p(R), _p(z), _p(vR), _p(vphi), _p(vz), _p(D),
                                      _p(Jr), _p(Jz)), "fj_actions")
        return Jr, Jz

    def df_points(self, R, z, vR, vphi, vz):
        R, z, vR, vphi, vz = (_a(x) for x in (R, z, vR, vphi, vz))
        f = np.zeros(len(R))
        self._chk(self.lib.fj_df_points(self.h, len(R), _p(R), _p(z), _p(vR), _p(vphi), _p(vz), _p(f)),
                  "fj_df_points")
        return f

    def sigma_df(self, rhl, R, z):
        rhl, R, z = _a(rhl), _a(np.atleast_1d(R)), _a(np.atleast_1d(z))
        out = np.zeros(len(R))
        self._chk(self.lib.fj_sigma_df(self.h, _p(rhl), len(R), _p(R), _p(z), _p(out)), "fj_sigma_df")
        return out

    def line_profile(self, rhl, R, z, vx):
        rhl, vx = _a(rhl), _a(np.atleast_1d(vx))
        out = np.zeros(len(vx))
        self._chk(self.lib.fj_line_profile(self.h, _p(rhl), R, z, len(vx), _p(vx), _p(out)), "fj_line_profile")
        return out

    def velocity_distribution(self, which, R, z, v):
        """which: 0 P(vphi), 1 P(vz), 2 P(vR)."""
        v = _a(np.atleast_1d(v))
        out = np.zeros(len(v))
        self._chk(self.lib.fj_velocity_distribution(self.h, which, R, z, len(v), _p(v), _p(out)),
                  "fj_velocity_distribution")
        return out

    def stats(self):
        s = Stats()
        self._chk(self.lib.fj_get_stats(self.h, C.byref(s)), "fj_get_stats")
        return {n: getattr(s, n) for n, _ in Stats._fields_}

    # ---- action surrogate -------------------------------------------------
    def surrogate_init(self, seed=1, capacity=4 << 20):
        self._chk(self.lib.fj_surrogate_init(self.h, seed, capacity), "fj_surrogate_init")

    def surrogate_collect(self, keep):
        self._chk(self.lib.fj_surrogate_collect(self.h, keep), "fj_surrogate_collect")

    def surrogate_train(self, epochs, batch=65536, lr=2e-3):
        loss = C.c_double()
        self._chk(self.lib.fj_surrogate_train(self.h, epochs, batch, lr, C.byref(loss)), "fj_surrogate_train")
        return loss.value

    def surrogate_mode(self, on, judge=0.01, jz_min=1e-3):
        self._chk(self.lib.fj_surrogate_mode(self.h, int(on), judge, jz_min), "fj_surrogate_mode")

    def surrogate_stats(self):
        s = SurrogateStats()
        self._chk(self.lib.fj_surrogate_get_stats(self.h, C.byref(s)), "fj_surrogate_get_stats")
        return {n: getattr(s, n) for n, _ in SurrogateStats._fields_}

    def surrogate_params(self, set_=None):
        fp = C.POINTER(C.c_float)
        out = np.zeros(MLP_NPARAM, dtype=np.float32)
        src = None if set_ is None else np.ascontiguousarray(set_, dtype=np.float32)
        self._chk(self.lib.fj_surrogate_params(self.h, out.ctypes.data_as(fp),
                                               None if src is None else src.ctypes.data_as(fp)),
                  "fj_surrogate_params")
        return out

    def surrogate_forward(self, x):
        fp = C.POINTER(C.c_float)
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), 2), dtype=np.float32)
        self._chk(self.lib.fj_surrogate_forward(self.h, len(x), x.ctypes.data_as(fp), y.ctypes.data_as(fp)),
                  "fj_surrogate_forward")
        return y

    def surrogate_gradient(self, x, t):
        fp = C.POINTER(C.c_float)
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._chk(self.lib.fj_surrogate_gradient(self.h, len(x), x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                                 g.ctypes.data_as(fp), C.byref(loss)), "fj_surrogate_gradient")
        return g, loss.value


class Batch:
    """Many single-component models advanced together (fj_batch_*).  models:
    dicts with ar, ntheta, npoly, model (guess), mass, r0, q, df, quad."""

    def __init__(self, ctx, models):
        self.ctx, self.lib = ctx, ctx.lib
        self._ar = [_a(m["ar"]) for m in models]
        specs = (ModelSpec * len(models))()
        for k, m in enumerate(models):
            specs[k] = ModelSpec(len(self._ar[k]), _p(self._ar[k]), m["ntheta"], m["npoly"], GUESS[m["model"]],
                                 m.get("mass", 1.0), m.get("r0", 1.0), m.get("q", 1.0), m["df"], m.get("quad", 0))
        self.shapes = [(len(self._ar[k]), m["npoly"]) for k, m in enumerate(models)]
        h = C.c_void_p()
        ctx._chk(self.lib.fj_batch_create(ctx.h, len(models), specs, C.byref(h)), "fj_batch_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.fj_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self):
        self.ctx._chk(self.lib.fj_batch_setup(self.h), "fj_batch_setup")

    def autotune(self):
        self.ctx._chk(self.lib.fj_batch_autotune(self.h), "fj_batch_autotune")

    def iterate(self, n=1):
        self.ctx._chk(self.lib.fj_batch_iterate(self.h, n), "fj_batch_iterate")

    def get(self, m):
        nr, npoly = self.shapes[m]
        t = [np.zeros((nr, npoly)) for _ in range(8)]
        E, D = np.zeros(30), np.zeros(30)
        self.ctx._chk(self.lib.fj_batch_get(self.h, m, *[_p(x) for x in t], _p(E), _p(D)), "fj_batch_get")
        out = dict(zip(["rhl", "sigRl", "sigpl", "sigzl", "vrotl", "phil", "Pr", "Pr2"], t))
        out["Egrid"], out["Dgrid"] = E, D
        return out

    def set(self, m, phil=None, Pr=None, Pr2=None, Egrid=None, Dgrid=None):
        """Installs model m's potential tables and / or Delta table."""
        nr, npoly = self.shapes[m]
        arrs = []
        for x, n in ((phil, nr * npoly), (Pr, nr * npoly), (Pr2, nr * npoly), (Egrid, 30), (Dgrid, 30)):
            if x is None:
                arrs.append(None)
                continue
            a = _a(x)
            if a.size != n:
                raise ValueError("fj_batch_set: table of %d values where %d are needed" % (a.size, n))
            arrs.append(a)
        self.ctx._chk(self.lib.fj_batch_set(self.h, m, *[None if a is None else _p(a) for a in arrs]),
                      "fj_batch_set")

    def stats(self):
        s = BatchStats()
        self.ctx._chk(self.lib.fj_batch_get_stats(self.h, C.byref(s)), "fj_batch_get_stats")
        return {n: getattr(s, n) for n, _ in BatchStats._fields_}
