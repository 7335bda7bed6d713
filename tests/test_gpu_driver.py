"""End to end through the C++ facade: the `fjmodels` driver (the flow of
fJmodels.cpp:86-144 oneComp) on a small Hernquist model, its full-precision
models/*.out.f64 files against the oracle run on the same parameters
(guess potential, Delta table, two computeNewPhi iterations), bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PARAM = """itermax 2
model Hernquist
outfile e2e
h(J):dphi 0.55
h(J):dz 0.55
g(J):dphi 1.
g(J):dz 1.
chi 0.
mass 1.
r0 1.
q 1.
nr 12
ntheta 4
npoly 3
rmin 0.05
rmax 50
"""

NAMES = ["rhl", "vrotl", "sigRl", "sigpl", "sigzl", "sigRzl", "phil", "Pr", "Pr2"]


def read_f64(path, nr, npoly):
    tok = open(path).read().split()
    v = np.array([float(x) for x in tok[12:]])   # header: 3 + 4 + 3 + 2 values
    ar, rest = v[:nr], v[nr:]
    return ar, {n: rest[i * nr * npoly:(i + 1) * nr * npoly].reshape(nr, npoly) for i, n in enumerate(NAMES)}


def test_driver_matches_oracle(tmp_path):
    exe = os.path.join(ROOT, "fjmodels_amd", "fjmodels")
    assert os.path.exists(exe), "build the package first (make -C fjmodels_amd all)"
    pf = tmp_path / "param.txt"
    pf.write_text(PARAM)
    print("running", exe, flush=True)
    r = subprocess.run([exe, str(pf), "--outdir", str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    nr, npoly, nth = 12, 3, 4
    p = O.Pot(O.log_grid(nr, 0.05, 50.0), npoly, nth)
    rhl = p.guess_rhl("Hernquist")
    p.set_tables(*p.compute_phil(rhl))
    assert p.tabulate_delta() == 0
    dfo = O.df_params("Hernquist")
    for it in range(2):
        ref = p.compute_new_phi(dfo, nthreads=16)
        ar, got = read_f64(tmp_path / f"e2e_{it}.out.f64", nr, npoly)
        assert np.array_equal(ar, p.ar)
        for n in NAMES:
            assert np.array_equal(got[n], ref[n]), (it, n, np.max(np.abs(got[n] - ref[n])))
        assert os.path.exists(tmp_path / f"e2e_{it}.out")
        p.set_tables(ref["phil"], ref["Pr"], ref["Pr2"])
