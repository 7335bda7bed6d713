"""Operation count of the action integrals of the Hernquist 40 x 20 workload
(MEASUREMENT INFRASTRUCTURE: runs on the CPU, writes the figures the roofline
of the QAGS kernel uses).

Iteration 0: Hernquist guess potential on 40 log radii in [0.01, 100],
npoly = 5, 20 Gauss-Legendre cos(theta) nodes, Delta table from the guess
potential, log-spaced velocity rule (quad = 0, 22^3 nodes).  Every bound node
of every grid point (or every stride-th point) runs through the oracle's
QAGS, and the QK21 rules it applied are counted by libfjopcount.so (the
counting restatement in opcount.cpp), in the reference's integrand form and
with the per-orbit constants hoisted.

  python oracle/count_ops.py [--stride S] [--threads T]
"""
import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:] = [q for q in sys.path if os.path.abspath(q or '.') != HERE]
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402

_dp = C.POINTER(C.c_double)


def load():
    O.build()
    lib = C.CDLL(os.path.join(HERE, "libfjopcount.so"))
    lib.fjc_grid_ops.restype = C.c_int
    lib.fjc_grid_ops.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp]
    lib.fjc_rule_ops.restype = C.c_int
    lib.fjc_rule_ops.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp]
    return lib


def grid_points(p, stride=1):
    ci, _ = O.gauleg(p.ngauss, 0.0, 1.0)
    R, z = [], []
    for idx in range(0, p.nr * p.ngauss, stride):
        i, g = divmod(idx, p.ngauss)
        sg = np.sqrt(1 - ci[g] * ci[g])
        R.append(p.ar[i] * sg)
        z.append(p.ar[i] * ci[g])
    return np.array(R), np.array(z)


def count(p, R, z, hoist, lib=None, quad=0):
    """Counted operations of the action integrals at the points (R, z) on the
    oracle potential p: dict with the Ju / Jv rule operations and counts."""
    lib = lib or load()
    phil = O.arr(p.tables()[0])
    ar = O.arr(p.ar)
    R, z = O.arr(R), O.arr(z)
    out = np.zeros(7)
    lib.fjc_grid_ops(p.h, p.nr, O.ptr(ar), p.npoly, O.ptr(phil), quad, len(R), O.ptr(R), O.ptr(z),
                     int(hoist), O.ptr(out))
    ops_u, ops_v, ops_f, nu, nv, nodes, check = out
    return {"ops_ju": ops_u, "ops_jv": ops_v, "ops_integrand": ops_f, "rules_ju": int(nu),
            "rules_jv": int(nv), "bound_nodes": int(nodes), "first_rule_rel_diff": check,
            "ops_per_rule": (ops_u + ops_v) / max(nu + nv, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stride", type=int, default=1)
    args = ap.parse_args()
    p = O.guess_model(40, 0.01, 100.0, 5, 20)[0]
    p.set_quad(0)
    R, z = grid_points(p, args.stride)
    lib = load()
    print(f"workload: Hernquist 40x20 npoly=5 iteration 0, {len(R)} grid points (stride {args.stride})")
    for hoist, name in ((0, "reference form"), (1, "hoisted constants")):
        t0 = time.time()
        c = count(p, R, z, hoist, lib)
        print(f"{name}: ops_per_rule {c['ops_per_rule']:.2f}  Ju rules {c['rules_ju']}  Jv rules {c['rules_jv']}  "
              f"ops Ju {c['ops_ju']:.6e}  ops Jv {c['ops_jv']:.6e}  integrand share "
              f"{c['ops_integrand'] / (c['ops_ju'] + c['ops_jv']):.4f}  bound nodes {c['bound_nodes']}  "
              f"first-rule rel diff {c['first_rule_rel_diff']:.2e}  ({time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main()
