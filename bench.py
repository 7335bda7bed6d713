"""Benchmark: grid-point DF moments per second per SCF iteration (plain FP64).

Workload (one GPU): single-component Hernquist double-power-law f(J)
(A = 1.667, B = 5; h: dphi = dz = 0.55; g: dphi = dz = 1.0; chi = 0; M = r0 = 1),
40 log-spaced radii in [0.01, 100] x 20 Gauss-Legendre cos(theta) nodes,
l <= 8 (npoly = 5), 22^3 velocity nodes per point, guess Hernquist potential,
Delta table from the guess potential.  A step is one SCF iteration
(computeNewPhi: node phase + projection + Poisson) followed by installing the
new potential.  Inputs are device resident; the timed loop runs K iterations.

Prints one JSON line.  The roofline entry is for the dominant kernel
(k_integrals, the persistent QAGS over the action integrals): its flops per
launch are the QK21 rules the launch applied (counted by the kernel) times
the operations of one rule, counted by an op-counting build of the reference
integrands and rule (oracle/opcount.cpp, run over every bound node of
iteration 0 of this workload by oracle/count_ops.py: 2,755 per rule in the
reference's form, transcendentals counting 1); its traffic is the HBM bytes
per launch of the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs.  Both
figures are read from the k_integrals profile summary under profiles/.
The CPU baseline times the oracle (the C restatement of the reference
algorithm) with OpenMP on a bounded sample of the same grid points.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6   # TFLOP/s, MI355X
PEAK_FP32 = 157.3  # TFLOP/s, MI355X vector FP32
K_INTEGRALS_PROFILE = os.path.join(ROOT, "profiles", "k_integrals_hernquist_40x20_fp64.md")


def profile_figures(path=K_INTEGRALS_PROFILE):
    """The figures a profile summary states as `key: value` lines: the
    counted operations per QK21 rule (oracle/count_ops.py, op-counting build
    of the action integrals on this workload) and the HBM bytes per launch
    (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, calibrated on a device copy)."""
    out = {}
    try:
        for line in open(path):
            k, sep, v = line.partition(":")
            if sep and k.strip() in ("ops_per_rule_reference_form", "ops_per_rule_hoisted",
                                     "traffic_bytes_per_launch", "traffic_commit"):
                v = v.strip()
                out[k.strip()] = v if k.strip() == "traffic_commit" else float(v)
    except OSError:
        pass
    return out


def log(msg):
    """One progress line on stderr (the JSON line alone goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def line_tasks_bytes(stats):
    """Algorithmic bytes of k_integrals: per action integral the 12 node
    constants read and the action written (doubles)."""
    import numpy as np
    return float(np.mean([s["tasks"] for s in stats])) * 13 * 8


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(npoly, ntheta, nr, ar, tables, delta, threads=None):
    """The oracle (the C restatement of the reference algorithm, built -O3 for
    x86-64-v3, libfjoracle_o3.so) timed on this host's cores on a bounded
    sample of the same iteration: rhofDF on every 4th grid point of the
    40 x 20 grid, grid-parallel (OpenMP dynamic over points, the stronger CPU
    baseline), and on every 16th point reference-structured (serial over
    points, one OpenMP team over the 22^3 nodes of each, as
    integ/GaussQuad.cpp:160-180 and pot/potLeg.cpp:54-58 run it)."""
    import ctypes as C
    from oracle import oracle as O
    threads = threads or len(os.sched_getaffinity(0))   # every core this process may run on
    lib = C.CDLL(os.path.join(ROOT, "oracle", "libfjoracle_o3.so"))
    dp = C.POINTER(C.c_double)
    lib.fjo_pot_new.restype = C.c_void_p
    lib.fjo_pot_new.argtypes = [C.c_int, dp, C.c_int, C.c_int, C.c_int]
    lib.fjo_pot_set_tables.argtypes = [C.c_void_p, dp, dp, dp]
    lib.fjo_pot_set_delta.argtypes = [C.c_void_p, dp, dp]
    lib.fjo_time_grid.restype = C.c_double
    lib.fjo_time_grid.argtypes = [C.c_void_p, C.POINTER(O.DFParams), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.fjo_pot_free.argtypes = [C.c_void_p]
    a = O.arr(ar)
    h = lib.fjo_pot_new(nr, O.ptr(a), npoly, ntheta, len(delta[0]))
    t = [O.arr(x) for x in tables]
    lib.fjo_pot_set_tables(h, *[O.ptr(x) for x in t])
    dl = [O.arr(x) for x in delta]
    lib.fjo_pot_set_delta(h, *[O.ptr(x) for x in dl])
    df = O.df_params("Hernquist")
    out = {}
    for variant, name, stride in ((0, "grid_parallel", 4), (1, "reference_structured", 16)):
        n = C.c_int(0)
        dt = lib.fjo_time_grid(h, C.byref(df), threads, variant, stride, C.byref(n))
        out[name] = {"value": n.value / dt, "points": n.value, "seconds": dt}
    lib.fjo_pot_free(h)
    best = out["grid_parallel"]
    return {"value": best["value"], "unit": "grid-pt DF moments/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "cpus_on_host": os.cpu_count(),
            "sample": "rhofDF at %d of the %d grid points (every 4th) of iteration 0, grid-parallel OpenMP "
                      "(dynamic over points), %d threads, %.1f s; oracle C restatement built -O3 x86-64-v3"
                      % (best["points"], nr * ntheta, threads, best["seconds"]),
            "variants": out}


def mixed_leg(ctx, model, prev, last, steps, pts, npoly):
    """Mixed precision (FP32 action integrands, Kronrod rules, QAGS
    bookkeeping and DF; FP64 roots, Delta and Poisson) with the exact
    block-scan Poisson solve.  Accuracy: one iteration from the same
    potential as the last FP64 iteration, with the same (QUADPACK) Poisson
    solve, every coefficient's difference scaled by the monopole at its
    radius; speed: `steps` iterations continuing from there with the
    block-scan Poisson solve.  Roofline: the counted QK21 flops of all rules of
    the iteration over the FP32 integrator's launch time, against the FP32
    vector peak; the counters do not split off the rules of the FP64
    overflow pass (0.9 % of the tasks), so this bounds the FP32 pass from
    above."""
    import numpy as np
    import torch
    import fjmodels_amd as fj
    ref = last[0]
    ctx.set_potential(*prev)
    model.total = prev
    ctx.set_precision(fj.FJ_PREC_MIXED)
    got = model.iterate()[0]            # same (QUADPACK) Poisson solve as the FP64 reference iteration
    err = {k: float(np.max(np.abs(got[k] - ref[k]) / np.abs(ref[k][:, :1]))) for k in ("rhl", "phil", "sigRl", "sigzl")}
    model.mode = fj.FJ_POISSON_BLOCKSCAN
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = []
    for _ in range(steps):
        model.iterate()
        st.append(ctx.stats())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ctx.set_precision(fj.FJ_PREC_FP64)
    model.mode = fj.FJ_POISSON_QUADPACK
    qk = float(np.mean([x["qk_calls"] for x in st]))
    t_int = float(np.mean([x["ms_integrals"] for x in st])) * 1e-3
    ach = qk * profile_figures().get("ops_per_rule_reference_form", float("nan")) / t_int / 1e12
    return {"value": pts / dt, "ms_per_step": 1e3 * dt, "poisson": "blockscan",
            "err_vs_fp64": err, "err_note": "one iteration from the same potential as the last FP64 one, both with "
                                            "the QUADPACK Poisson solve; |diff| / monopole at the radius",
            "roofline": {"bound": "valu", "pipe": "VALU FP32", "kernel": "k_integrals_f32",
                         "achieved": ach, "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": ach / PEAK_FP32,
                         "ms_per_launch": t_int * 1e3},
            "phases_ms": {k: float(np.mean([x[k] for x in st])) for k in
                          ("ms_setup", "ms_integrals", "ms_overflow", "ms_moments", "ms_poisson")},
            "counters": {k: int(np.mean([x[k] for x in st])) for k in ("qk_calls", "overflow_tasks", "integral_lanes")}}


def fast_leg(ctx, model, prev, last, steps, pts):
    """Fast node phase (FJ_PREC_FAST: one lane per velocity node in
    k_fast_nodes, then k_moments; the Staeckel fudge with the reference's
    turning-point walks, FP32 roots and FP32 Gauss-Legendre action integrals
    of 64 / 32 points instead of QAGS; from the second pass on the lanes
    follow the previous pass's turning points).  Accuracy as in the mixed
    leg (one iteration from the same potential as the last FP64 one,
    QUADPACK Poisson on both sides); speed: `steps` iterations continuing
    from there, QUADPACK Poisson.  k_fast_nodes: its HIP-event time (with
    the lane order's sort) and the potential / integrand evaluations it
    counts."""
    import numpy as np
    import torch
    import fjmodels_amd as fj
    ref = last[0]
    ctx.set_potential(*prev)
    model.total = prev
    ctx.set_precision(fj.FJ_PREC_FAST)
    got = model.iterate()[0]
    err = {k: float(np.max(np.abs(got[k] - ref[k]) / np.abs(ref[k][:, :1])))
           for k in ("rhl", "phil", "sigRl", "sigpl", "sigzl")}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = []
    for _ in range(steps):
        model.iterate()
        st.append(ctx.stats())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ctx.set_precision(fj.FJ_PREC_FP64)
    t_k = float(np.mean([x["ms_setup"] for x in st])) * 1e-3
    evals = float(np.mean([x["phi_evals"] for x in st]))
    return {"value": pts / dt, "ms_per_step": 1e3 * dt, "poisson": "quadpack",
            "err_vs_fp64": err, "err_note": "one iteration from the same potential as the last FP64 one, both with "
                                            "the QUADPACK Poisson solve; |diff| / monopole at the radius",
            "k_fast_nodes": {"ms_per_launch": 1e3 * t_k, "potential_evaluations": evals,
                             "evaluations_per_s": evals / t_k},
            "phases_ms": {k: float(np.mean([x[k] for x in st])) for k in ("ms_points", "ms_setup", "ms_moments")}}


def sync():
    """torch.cuda.synchronize() when torch drives the device in this process
    (the library's own calls return when their results are on the host);
    torch is not imported here, so that a caller that has not loaded it
    keeps a single HIP runtime in the process."""
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        torch.cuda.synchronize()


def drho(a, b):
    """Largest difference of two rho_l tables, scaled by the monopole at each radius."""
    import numpy as np
    return float(np.max(np.abs(a["rhl"] - b["rhl"]) / np.abs(a["rhl"][:, :1])))


def converge(ctx, model, guess_tables, prec, tol=1e-4, max_iter=30, sur=None):
    """One SCF run of `model` from the guess potential to `tol` (successive
    rho_l differ by less than tol of the monopole, the driver's test), with
    the node phase precision `prec`; `sur`: the surrogate schedule (keys
    keep, epochs, judge, gate, max_sur, sur_floor, collect_from, serve_from,
    capacity, batch, lr, seed), whose Staeckel iterations run at `prec`.
    Returns (seconds, last output, virial, surrogate stats, rho_l change per
    iteration, iteration kinds, training seconds, median ms per kind)."""
    import numpy as np
    import fjmodels_amd as fj
    ctx.set_precision(prec)
    ctx.set_potential(*guess_tables)
    model.total = [x.copy() for x in guess_tables]
    model.philP = [tuple(x.copy() for x in guess_tables)]
    sync()
    t0 = time.perf_counter()
    hist, kinds, t_train, t_kind = [], [], 0.0, {}
    judged = []                  # per served iteration: decision and the judge's weighted medians
    prev, out = None, None
    phase = "collect" if sur else "plain"
    n_sur = 0
    for it in range(max_iter):
        if sur and it == sur["collect_from"]:
            ctx.surrogate_init(seed=sur.get("seed", 7), capacity=sur.get("capacity", 1 << 20))
            ctx.surrogate_collect(sur["keep"])
        if sur and phase == "collect" and it == sur["serve_from"]:
            ctx.surrogate_collect(1.0)            # from now on: the judged nodes
            tt = time.perf_counter()
            ctx.surrogate_train(sur["epochs"], batch=sur.get("batch", 32768), lr=sur.get("lr", 3e-3))
            t_train += time.perf_counter() - tt
            ctx.surrogate_mode(True, judge=sur["judge"], jz_min=1e-3)
            phase = "sur"
        if phase == "sur":
            saved = ([x.copy() for x in model.total], [tuple(x.copy() for x in p) for p in model.philP])
        ti = time.perf_counter()
        out = model.iterate()[0]
        ti = 1e3 * (time.perf_counter() - ti)
        t_kind.setdefault("S" if phase == "sur" else "P", []).append(ti)
        if phase == "sur":
            js = ctx.surrogate_stats()
            rejected = max(js["judge_werr_r_med"], js["judge_werr_z_med"]) > sur["gate"]
            judged.append({"iteration": it, "kind": "R" if rejected else "S",
                           "werr_r_med": js["judge_werr_r_med"], "werr_z_med": js["judge_werr_z_med"]})
            if rejected:
                kinds.append("R")                 # rejected by the judge
                ctx.surrogate_mode(False)
                ctx.surrogate_collect(0.0)
                model.total, model.philP = saved
                ctx.set_potential(*model.total)
                phase = "plain"
                continue
        kinds.append("S" if phase == "sur" else "P")
        d = drho(out, prev) if prev is not None else None
        hist.append(d)
        prev = out
        if phase == "sur":
            n_sur += 1
            if (d is not None and d < sur["sur_floor"]) or n_sur >= sur["max_sur"]:
                ctx.surrogate_mode(False)
                ctx.surrogate_collect(0.0)
                phase = "plain"
            else:
                tt = time.perf_counter()
                ctx.surrogate_train(1, batch=sur.get("batch", 32768), lr=sur.get("lr", 3e-3) / 3)
                t_train += time.perf_counter() - tt
        elif phase == "plain" and d is not None and d < tol:
            break
    sync()
    dt = time.perf_counter() - t0
    vir = ctx.virial(out["rhl"], out["vrotl"], out["sigRl"], out["sigpl"], out["sigzl"])
    st = dict(ctx.surrogate_stats(), judge_history=judged) if sur else None
    if sur:
        ctx.surrogate_mode(False)
        ctx.surrogate_collect(0.0)
    ctx.set_precision(fj.FJ_PREC_FP64)
    return dt, out, vir, st, hist, "".join(kinds), t_train, {k: float(np.median(v)) for k, v in t_kind.items()}


def surrogate_leg(ctx, model, guess_tables, tol=1e-4, max_iter=30, keep=0.15, epochs=4, judge=0.05,
                  max_sur=8, sur_floor=4e-3, gate=0.05, collect_from=3, serve_from=5):
    """The Hernquist model converged four times from the guess potential to
    the same tolerance: successive density coefficients differ by less than
    `tol` of the monopole (the driver's convergence test on rho_l).  All
    four use the QUADPACK Poisson solve.

    plain: FP64 Staeckel actions every iteration.
    fast: the fast node phase (FJ_PREC_FAST) every iteration.
    mixed: the same with the mixed-precision node phase.
    surrogate (with the mixed node phase for its Staeckel iterations):
    Staeckel iterations, keeping from iteration `collect_from` on a random
    `keep` of the bound nodes' actions as training samples; at iteration
    `serve_from` the network is trained (`epochs` passes) and serves
    (Jr, Jz) while a random `judge` fraction of
    the nodes is recomputed with Staeckel, scored, added to the training set
    and used for one fine-tuning pass after each iteration; once the served
    iterations change rho_l by less than `sur_floor` (the network's noise
    level) or after `max_sur` of them, Staeckel iterations run to `tol`.
    The judge gates the network: a served iteration whose judged median
    error of Jr or Jz, relative to the node's total action and weighted by
    the node's share of the density sum, exceeds `gate` is discarded (the
    previous potential is restored) and the run continues with Staeckel.

    Reports the wall time of the three runs (training included), the
    median wall time of one Staeckel and one served iteration (the judge's
    rejected ones included), the iteration counts, the differences of the converged rho_l from the plain
    run's (scaled by the monopole) and the virial ratios."""
    import fjmodels_amd as fj
    sur = dict(keep=keep, epochs=epochs, judge=judge, gate=gate, max_sur=max_sur, sur_floor=sur_floor,
               collect_from=collect_from, serve_from=serve_from)

    def run(with_sur, prec):
        return converge(ctx, model, guess_tables, prec, tol, max_iter, sur if with_sur else None)

    t_p, out_p, vir_p, _, h_p, k_p, _, i_p = run(False, fj.FJ_PREC_FP64)
    log("converged plain FP64: %d iterations, %.2f s" % (len(k_p), t_p))
    t_f, out_f, vir_f, _, h_f, k_f, _, i_f = run(False, fj.FJ_PREC_FAST)
    log("converged fast: %d iterations, %.2f s" % (len(k_f), t_f))
    t_m, out_m, vir_m, _, h_m, k_m, _, i_m = run(False, fj.FJ_PREC_MIXED)
    log("converged mixed: %d iterations, %.2f s" % (len(k_m), t_m))
    t_s, out_s, vir_s, st, h_s, k_s, t_tr, i_s = run(True, fj.FJ_PREC_MIXED)
    log("converged surrogate: %s, %.2f s" % (k_s, t_s))
    return {"tol_rho_l": tol, "ms_plain": 1e3 * t_p, "ms_fast": 1e3 * t_f, "ms_mixed": 1e3 * t_m,
            "ms_surrogate": 1e3 * t_s,
            "speedup_fast": t_p / t_f, "speedup": t_p / t_s, "speedup_mixed_only": t_p / t_m,
            "iterations_plain": len(k_p), "iterations_fast": len(k_f), "iterations_mixed": len(k_m),
            "iterations_surrogate_run": k_s,
            "ms_training": 1e3 * t_tr,
            "ms_median_iteration": {"plain": i_p["P"], "fast": i_f["P"], "mixed": i_m["P"],
                                    "surrogate_run_staeckel": i_s["P"], "surrogate_run_served": i_s.get("S")},
            "rho_l_diff_converged_fast": drho(out_f, out_p),
            "rho_l_diff_converged": drho(out_s, out_p), "rho_l_diff_converged_mixed": drho(out_m, out_p),
            "virial_WK_plain": float(vir_p[4]), "virial_WK_fast": float(vir_f[4]), "virial_WK_mixed": float(vir_m[4]),
            "virial_WK_surrogate": float(vir_s[4]),
            "d_rho_l_plain": [None if x is None else float("%.3g" % x) for x in h_p],
            "d_rho_l_fast": [None if x is None else float("%.3g" % x) for x in h_f],
            "d_rho_l_surrogate_run": [None if x is None else float("%.3g" % x) for x in h_s],
            "judge_last_served_iteration": {k: st[k] for k in ("judge_werr_r_med", "judge_werr_z_med",
                                                             "judge_err_r_med", "judge_err_r_p95", "judge_err_z_med",
                                                             "judge_err_z_p95", "n_train", "nodes_sur", "fallback",
                                                             "judged")},
            "note": "speedup: plain FP64 over surrogate + mixed node phase; P = Staeckel iteration (mixed), "
                    "S = served by the network (%.0f%% of nodes judged), R = served iteration "
                    "rejected by the judge (density-weighted median error above %.0f%%) and redone with Staeckel; "
                    "training: %d epochs "
                    "before iteration %d, then one pass per served iteration" % (100 * judge, 100 * gate, epochs,
                                                                                  serve_from)}


def cfg5_leg(tol=1e-4, max_iter=30, keep=0.065, epochs=16, judge=0.05, gate=0.05):
    """Learned-surrogate configuration: the rotating NFW model, 96 log radii
    in [0.01, 100] x 32 cos(theta) nodes, l <= 12 (npoly 7), NFW DPL f(J)
    with chi = 1 and the README rotation form (k = 0.5), NFW guess potential
    and its Delta table, converged from the guess potential to `tol` three
    ways, all with the block-scan Poisson solve (with the reference's
    QUADPACK integrals, epsrel 1e-3, successive rho_l of this model keep
    differing by 2-10 % of the monopole for 30 iterations: the solve's
    adaptive decisions change from one potential to the next): plain FP64;
    the fast node phase; and
    the surrogate schedule: Staeckel (mixed precision) iterations 0-2,
    keeping a random `keep` of the bound nodes of iterations 1 and 2 as
    training samples (about 4 M), training `epochs` passes before iteration
    3, serving iterations 3.. with the judge recomputing a random `judge`
    fraction of nodes; a served iteration whose density-weighted median
    action error exceeds `gate` is discarded and the run continues with
    Staeckel iterations, as do the last iterations to `tol`.  Reports times,
    iteration kinds, rho_l and Phi_l at convergence against the plain run
    (scaled by the monopole), the virial ratios and the judge's figures."""
    import numpy as np
    import fjmodels_amd as fj
    ctx = fj.Context(0)
    ar = fj.log_grid(96, 0.01, 100.0)
    comp = {"model": "NFW", "mass": 1.0, "r0": 1.0, "q": 1.0, "df": fj.df_params("NFW", chi=1.0, rot_form=1)}
    model = fj.Model(ctx, ar, 32, 7, [comp], poisson_mode=fj.FJ_POISSON_BLOCKSCAN)
    model.setup()
    guess = [x.copy() for x in model.total]
    sur = dict(keep=keep, epochs=epochs, judge=judge, gate=gate, max_sur=8, sur_floor=4e-3, collect_from=1,
               serve_from=3, capacity=1 << 22)

    def dphi(a, b):
        return float(np.max(np.abs(a["phil"] - b["phil"]) / np.abs(a["phil"][:, :1])))

    t_p, out_p, vir_p, _, h_p, k_p, _, i_p = converge(ctx, model, guess, fj.FJ_PREC_FP64, tol, max_iter)
    log("cfg5 plain FP64: %d iterations, %.2f s" % (len(k_p), t_p))
    t_f, out_f, vir_f, _, h_f, k_f, _, i_f = converge(ctx, model, guess, fj.FJ_PREC_FAST, tol, max_iter)
    log("cfg5 fast: %d iterations, %.2f s" % (len(k_f), t_f))
    t_s, out_s, vir_s, st, h_s, k_s, t_tr, i_s = converge(ctx, model, guess, fj.FJ_PREC_MIXED, tol, max_iter, sur)
    log("cfg5 surrogate: %s, %.2f s" % (k_s, t_s))
    ctx.close()
    return {"workload": "NFW DPL f(J), chi 1, k 0.5 (README form), 96 log radii [0.01,100] x 32 GL cos(theta), "
                        "l<=12, 22^3 nodes/point, guess potential + Delta table, block-scan Poisson",
            "tol_rho_l": tol, "ms_plain": 1e3 * t_p, "ms_fast": 1e3 * t_f, "ms_surrogate": 1e3 * t_s,
            "speedup_fast": t_p / t_f, "speedup_surrogate": t_p / t_s,
            "iterations": {"plain": k_p, "fast": k_f, "surrogate": k_s},
            "ms_median_iteration": {"plain": i_p["P"], "fast": i_f["P"], "surrogate_run_staeckel": i_s["P"],
                                    "surrogate_run_served": i_s.get("S")},
            "ms_training": 1e3 * t_tr, "training_samples": st["n_train"],
            "rho_l_diff_fast": drho(out_f, out_p), "phi_l_diff_fast": dphi(out_f, out_p),
            "rho_l_diff_surrogate": drho(out_s, out_p), "phi_l_diff_surrogate": dphi(out_s, out_p),
            "virial_WK": {"plain": float(vir_p[4]), "fast": float(vir_f[4]), "surrogate": float(vir_s[4])},
            "judge_last_served_iteration": {k: st[k] for k in ("judge_werr_r_med", "judge_werr_z_med",
                                                             "judge_err_r_med", "judge_err_r_p95",
                                                             "judge_err_z_med", "judge_err_z_p95", "nodes_sur",
                                                             "fallback", "judged")},
            "judge_history": st["judge_history"],
            "gate": gate,
            "note": "iteration kinds: P Staeckel, S served by the network, R served and rejected by the judge"}


def batch_leg(ctx, n_models=8, iters=2, seed=5):
    """Batched models against the same models one by one: n_models
    single-component models of random type (Hernquist, Isochrone, NFW), DF
    parameters (dphi, dz in [0.3, 1.2], chi in [0, 2], README rotation form)
    and uneven grids (nr in 24..40, ntheta in 12..20, l_max 8..12), mixed
    precision node phase and block-scan Poisson.  Every phase of an iteration
    is one launch over all models, replayed from a HIP graph, with launch
    shapes from the autotuner.  Compared with the same models advanced one
    after another through a single-model context (same precision and
    Poisson mode).  Times cover the SCF iterations only (setup, Delta tables
    and tuning are reported apart)."""
    import numpy as np
    import torch
    import fjmodels_amd as fj
    rng = np.random.default_rng(seed)
    models = []
    for k in range(n_models):
        kind = ("Hernquist", "Isochrone", "NFW")[k % 3]
        nr, nt, npoly = int(rng.integers(24, 41)), int(rng.integers(12, 21)), int(rng.integers(5, 8))
        df = fj.df_params(kind, dphi_h=float(rng.uniform(0.3, 1.2)), dz_h=float(rng.uniform(0.3, 1.2)),
                          chi=float(rng.uniform(0, 2)), rot_form=1)
        models.append({"ar": fj.log_grid(nr, 0.01, 100.0), "ntheta": nt, "npoly": npoly, "model": kind, "df": df})
    pts = sum(len(m["ar"]) * m["ntheta"] for m in models)
    ctx.set_precision(fj.FJ_PREC_MIXED)
    t0 = time.perf_counter()
    b = fj.Batch(ctx, models)
    b.setup()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.autotune()
    t_tune = time.perf_counter() - t0
    b.iterate(1)                       # graph capture + first replay
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.iterate(iters)
    torch.cuda.synchronize()
    t_b = (time.perf_counter() - t0) / iters
    st = b.stats()
    b.close()
    t_seq = 0.0
    for m in models:
        ctx.set_grid(m["ar"], m["ntheta"], m["npoly"])
        rhl = ctx.guess_rhl(m["model"], 1.0, 1.0, 1.0)
        ctx.set_potential(*ctx.poisson(rhl, fj.FJ_POISSON_BLOCKSCAN))
        ctx.tabulate_delta()
        ctx.set_df(m["df"])
        out = ctx.compute_new_phi(fj.FJ_POISSON_BLOCKSCAN)
        ctx.set_potential(out["phil"], out["Pr"], out["Pr2"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            out = ctx.compute_new_phi(fj.FJ_POISSON_BLOCKSCAN)
            ctx.set_potential(out["phil"], out["Pr"], out["Pr2"])
        torch.cuda.synchronize()
        t_seq += (time.perf_counter() - t0) / iters
    ctx.set_precision(fj.FJ_PREC_FP64)
    return {"models": n_models, "grid_points": pts, "ms_per_iter_batched": 1e3 * t_b,
            "ms_per_iter_one_by_one": 1e3 * t_seq, "speedup": t_seq / t_b,
            "value_batched": pts / t_b, "unit": "grid-pt DF moments/s",
            "s_setup": t_setup, "s_autotune": t_tune,
            "tuned": {"setup_block": st["setup_block"], "integ_blocks": st["integ_blocks"], "graph": st["graph"]},
            "precision": "mixed (FP32 integrands/DF, FP64 roots/Delta)", "poisson": "blockscan"}


def sweep_leg(ctx, n_models=32, seed=0, iters=2):
    """Model sweep (a bounded sample of the 256-model sweep): the first
    n_models of the seeded draw of fjmodels_amd.draw_models (types uniform
    over Hernquist/Isochrone/NFW, dphi, dz in [0.3, 1.2], chi in [0, 2],
    nr in 24..96, ntheta in 12..32, l_max 8/10/12), mixed precision node
    phase, block-scan Poisson, packed into batches of at most 64 M nodes.
    The timed iterations are graph replays after the first (capture)."""
    import fjmodels_amd as fj
    from fjmodels_amd import sweep as S
    full = fj.draw_models(256, seed=seed)
    full_pts = sum(len(m["ar"]) * m["ntheta"] for m in full)
    ctx.set_precision(fj.FJ_PREC_MIXED)
    try:
        def progress(done, total, sec):
            print("sweep: batch %d of %d done, %.1f s" % (done, total, sec), file=sys.stderr, flush=True)
        _, st = fj.Sweep(ctx, full[:n_models]).run(iters, progress)
    finally:
        ctx.set_precision(fj.FJ_PREC_FP64)
    st["unit"] = "grid-pt DF moments/s"
    st["full_sweep"] = {"models": 256, "grid_points": full_pts, "nodes": full_pts * S.NODES_PER_POINT,
                        "batches": len(S.partition(full))}
    st["precision"] = "mixed (FP32 integrands/DF, FP64 roots/Delta)"
    st["poisson"] = "blockscan"
    return st


def sweep_fast_leg(ctx, n_models=256, seed=0, iters=2):
    """The model sweep under the fast node phase: the first n_models of the
    seeded draw (all 256 by default) in one batch, since fast-phase batches
    hold no node records; every phase one launch over all models (the node
    phase one fused kernel), block-scan Poisson, graph replays after the
    first iteration."""
    import fjmodels_amd as fj
    from fjmodels_amd import sweep as S
    models = fj.draw_models(256, seed=seed)[:n_models]
    ctx.set_precision(fj.FJ_PREC_FAST)
    try:
        sw = fj.Sweep(ctx, models)
        _, st = sw.run(iters)
    finally:
        ctx.set_precision(fj.FJ_PREC_FP64)
    st["unit"] = "grid-pt DF moments/s"
    st["precision"] = "fast node phase (reference walks, FP32 roots, 64-point Gauss-Legendre actions)"
    st["poisson"] = "blockscan"
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision / block-scan leg")
    ap.add_argument("--no-fast", action="store_true", help="skip the fast node phase leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the plain vs surrogate convergence leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-models leg")
    ap.add_argument("--no-cfg5", action="store_true", help="skip the NFW surrogate configuration leg")
    ap.add_argument("--sweep", type=int, default=32,
                    help="models of the seeded 256-model sweep draw to run (0: skip the sweep leg)")
    ap.add_argument("--sweep-fast", type=int, default=256,
                    help="models of the sweep draw to run in one batch under the fast node phase (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device-to-device copy (reference for FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(sharding the grid over GPUs is not built)"}))
        sys.exit(2)

    import numpy as np
    import torch
    import fjmodels_amd as fj

    nr, ntheta, npoly = 40, 20, 5
    ar = fj.log_grid(nr, 0.01, 100.0)
    ctx = fj.Context(0)
    comp = {"model": "Hernquist", "mass": 1.0, "r0": 1.0, "q": 1.0, "df": fj.df_params("Hernquist")}
    model = fj.Model(ctx, ar, ntheta, npoly, [comp])
    model.setup()
    guess_tables = [x.copy() for x in model.total]
    delta = (model.Egrid.copy(), model.Dgrid.copy())

    for _ in range(args.warmup):
        model.iterate()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    last = None
    for _ in range(args.steps):
        prev = [x.copy() for x in model.total]
        last = model.iterate()
        stats.append(ctx.stats())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / args.steps
    pts = nr * ntheta * len(model.components)
    value = pts / (dt / args.steps)

    qk = np.mean([s["qk_calls"] for s in stats])
    t_int = np.mean([s["ms_integrals"] for s in stats]) * 1e-3
    fig = profile_figures()
    ops_rule = fig.get("ops_per_rule_reference_form")
    flops = qk * ops_rule if ops_rule else None
    achieved = flops / t_int / 1e12 if flops else None
    line = {
        "metric": "grid-pt DF moments/s per SCF iter, plain FP64",
        "value": value, "unit": "grid-pt DF moments/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (Hernquist guess potential, 40 x 20 grid)",
        "config": {"workload": "single-component Hernquist DPL f(J), 40 log radii [0.01,100] x 20 GL cos(theta), "
                               "l<=8, 22^3 velocity nodes/point, guess potential + Delta table, SCF iterations",
                   "nr": nr, "ntheta": ntheta, "npoly": npoly, "nodes_per_point": 10648},
        "roofline": {"bound": "valu", "pipe": "VALU FP64 (vector FP64 peak)",
                     "kernel": "k_integrals", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64 if achieved else None,
                     "traffic": fig.get("traffic_bytes_per_launch"),
                     "traffic_note": "HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs "
                                     "(calibrated on a device copy; profiles/%s, build %s); algorithmic bytes "
                                     "(node records) %.3g" % (os.path.basename(K_INTEGRALS_PROFILE),
                                                              fig.get("traffic_commit", "?"),
                                                              line_tasks_bytes(stats)),
                     "flops_per_launch": flops, "ms_per_launch": t_int * 1e3,
                     "ops_per_rule": ops_rule,
                     "ops_note": "operations per QK21 rule from the op-counting build of the reference integrands "
                                 "(oracle/opcount.cpp over iteration 0 of this workload); with the per-orbit "
                                 "constants hoisted: %s" % fig.get("ops_per_rule_hoisted")},
        "phases_ms": {k: float(np.mean([s[k] for s in stats])) for k in
                      ("ms_points", "ms_setup", "ms_integrals", "ms_overflow", "ms_moments", "ms_project",
                       "ms_poisson")},
        "counters": {k: int(np.mean([s[k] for s in stats])) for k in
                     ("bound_nodes", "tasks", "qk_calls", "overflow_tasks", "brent_iters", "integral_lanes",
                      "unbound_nodes", "bracket_failures", "nonfinite_nodes")},
    }
    log("FP64 leg: %.2f ms per iteration" % ms)
    if not args.no_fast:
        line["fast"] = fast_leg(ctx, model, prev, last, args.steps, pts)
        log("fast leg: %.2f ms per iteration" % line["fast"]["ms_per_step"])
    if not args.no_mixed:
        line["mixed"] = mixed_leg(ctx, model, prev, last, args.steps, pts, npoly)
        log("mixed leg: %.2f ms per iteration" % line["mixed"]["ms_per_step"])
    if not args.no_surrogate:
        line["surrogate"] = surrogate_leg(ctx, model, guess_tables)
        log("convergence leg done")
    if not args.no_cfg5:
        line["cfg5"] = cfg5_leg()
        log("cfg5 leg done")
    if not args.no_batch:
        line["batch"] = batch_leg(ctx)
        log("batch leg done")
    if args.sweep > 0:
        line["sweep"] = sweep_leg(ctx, args.sweep)
        log("sweep leg done")
    if args.sweep_fast > 0:
        line["sweep_fast"] = sweep_fast_leg(ctx, args.sweep_fast)
        log("fast sweep leg done")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(npoly, ntheta, nr, ar, guess_tables, delta)
        log("CPU baseline done")
    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = a.numel() * 8
    print(json.dumps(line))


if __name__ == "__main__":
    main()
