"""Writes tests/golden/cfg1_hernquist_40x20.npz: the oracle's full-precision
trajectory of the single-component Hernquist model (double-power-law f(J),
A = 1.667, B = 5, h(J) deltas 0.55, g(J) deltas 1.0, chi = 0, M = r0 = J0 = 1)
on 40 log radii in [0.01, 100] x 20 Gauss-Legendre cos(theta) nodes, l <= 8,
log22 velocity rule: the guess potential, its Delta table, then five SCF
iterations (computeNewPhi, install the new potential, vir2), as the
reference's oneComp runs them (fJmodels.cpp:86-144).

Variant choices recorded in the file: DPL f(J) live (df/DF.cpp:94-106
removed), rotation form 0 (chi = 0, unused), log22 rule (Int3D_111111_vec),
run-time grid sizes.

The oracle is the checker (test infrastructure); this script only writes
fixtures.  Run: python tests/golden/make_cfg1_trajectory.py [nthreads]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

NAMES = ["rhl", "vrotl", "sigRl", "sigpl", "sigzl", "sigRzl", "phil", "Pr", "Pr2"]


def main():
    nthreads = int(sys.argv[1]) if len(sys.argv) > 1 else 0
    p, rhl0 = O.guess_model(40, 0.01, 100.0, 5, 20)
    phil0, Pr0, Pr20 = p.tables()
    E, D = p.delta()
    dfo = O.df_params("Hernquist")
    traj = {n: [] for n in NAMES}
    vir = []
    for it in range(5):
        t0 = time.time()
        r = p.compute_new_phi(dfo, nthreads=nthreads)
        p.set_tables(r["phil"], r["Pr"], r["Pr2"])
        vir.append(p.vir2(r["rhl"], r["vrotl"], r["sigRl"], r["sigpl"], r["sigzl"]))
        for n in NAMES:
            traj[n].append(r[n])
        print(f"iteration {it}: {time.time() - t0:.1f} s, W/K {vir[-1][4]:.9f}", flush=True)
    out = os.path.join(HERE, "cfg1_hernquist_40x20.npz")
    np.savez_compressed(out, ar=p.ar, guess_rhl=rhl0, guess_phil=phil0, guess_Pr=Pr0, guess_Pr2=Pr20,
                        Egrid=E, Dgrid=D, vir2=np.array(vir), **{n: np.array(v) for n, v in traj.items()},
                        variant=np.array(["D1:dpl-live", "D2:rot-form-0", "D3:log22", "D4:runtime-grid"]))
    print("wrote", out)


if __name__ == "__main__":
    main()
