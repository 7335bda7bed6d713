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


def run_driver(pf, outdir, *extra):
    exe = os.path.join(ROOT, "fjmodels_amd", "fjmodels")
    print("running", exe, *extra, flush=True)
    r = subprocess.run([exe, str(pf), "--outdir", str(outdir), *extra], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    return r.stdout


@pytest.mark.parametrize("refresh", [0, 1])
def test_driver_resume_reproduces_uninterrupted_run(tmp_path, refresh):
    """--resume from the .f64 sidecar of iteration 0, numbered from 1, gives
    the bits of iteration 1 of the uninterrupted run (guess potential and
    Delta table rebuilt as at the start, potential tables from the file;
    with delta_refresh 1 the Delta table is tabulated again from the loaded
    potential, as the uninterrupted run does after writing it)."""
    pf = tmp_path / "param.txt"
    pf.write_text(PARAM + ("delta_refresh 1\n" if refresh else ""))
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    run_driver(pf, a)
    run_driver(pf, b, "--resume", str(a / "e2e_0.out.f64"), "--first-iter", "1")
    assert not (b / "e2e_0.out.f64").exists()
    assert (a / "e2e_1.out.f64").read_bytes() == (b / "e2e_1.out.f64").read_bytes()
    assert (a / "e2e_1.out").read_bytes() == (b / "e2e_1.out").read_bytes()


def test_driver_merge_matches_oracle(tmp_path):
    """merge 0.25: after each computeNewPhi the potential tables become
    1.25 new - 0.25 old (mergePhi, include/UtilsLeg.h:266-278); the written
    tables and the next iteration use them.  Against the oracle with the
    same update, bit for bit, and the printed distance."""
    pf = tmp_path / "param.txt"
    pf.write_text(PARAM + "merge 0.25\n")
    out = run_driver(pf, tmp_path)
    nr, npoly, nth = 12, 3, 4
    p = O.Pot(O.log_grid(nr, 0.05, 50.0), npoly, nth)
    rhl = p.guess_rhl("Hernquist")
    old = p.compute_phil(rhl)
    p.set_tables(*old)
    assert p.tabulate_delta() == 0
    dfo = O.df_params("Hernquist")
    dists = []
    for it in range(2):
        ref = p.compute_new_phi(dfo, nthreads=16)
        dists.append(np.max(np.abs((ref["phil"] - old[0]) / ref["phil"][:, :1])))
        merged = [1.25 * ref[n] - 0.25 * o for n, o in zip(("phil", "Pr", "Pr2"), old)]
        ar, got = read_f64(tmp_path / f"e2e_{it}.out.f64", nr, npoly)
        for n in ("rhl", "sigRl", "sigzl"):
            assert np.array_equal(got[n], ref[n]), (it, n)
        for n, m in zip(("phil", "Pr", "Pr2"), merged):
            assert np.array_equal(got[n], m), (it, n, np.max(np.abs(got[n] - m)))
        p.set_tables(*merged)
        old = merged
    printed = [float(l.split()[-1]) for l in out.splitlines() if l.startswith("-- Merge Phi: max distance")]
    assert len(printed) == 2 and np.allclose(printed, dists, rtol=1e-5, atol=1e-6), (printed, dists)


def test_driver_delta_refresh_matches_oracle(tmp_path):
    """delta_refresh 1: the Delta table is tabulated again from every new
    potential (the reference keeps the guess potential's table); against the
    oracle doing the same, bit for bit."""
    pf = tmp_path / "param.txt"
    pf.write_text(PARAM + "delta_refresh 1\n")
    run_driver(pf, tmp_path)
    nr, npoly, nth = 12, 3, 4
    p = O.Pot(O.log_grid(nr, 0.05, 50.0), npoly, nth)
    p.set_tables(*p.compute_phil(p.guess_rhl("Hernquist")))
    assert p.tabulate_delta() == 0
    dfo = O.df_params("Hernquist")
    for it in range(2):
        ref = p.compute_new_phi(dfo, nthreads=16)
        _, got = read_f64(tmp_path / f"e2e_{it}.out.f64", nr, npoly)
        for n in NAMES:
            assert np.array_equal(got[n], ref[n]), (it, n, np.max(np.abs(got[n] - ref[n])))
        p.set_tables(ref["phil"], ref["Pr"], ref["Pr2"])
        assert p.tabulate_delta() == 0


def test_driver_line_profiles(tmp_path):
    """lineprofile 1: the .lp file of writeOut.cpp:159-186 (150 velocities,
    radii ar[0], ar[2 nr/3], ar[5 nr/6] at z = ar[0]) after iteration 0; its
    values (6 digits) against the oracle's line_profile with the density and
    potential of that iteration."""
    pf = tmp_path / "param.txt"
    pf.write_text(PARAM.replace("itermax 2", "itermax 1") + "lineprofile 1\n")
    run_driver(pf, tmp_path)
    rows = [list(map(float, l.split())) for l in (tmp_path / "e2e_0.lp").read_text().splitlines()]
    assert len(rows) == 150 and all(len(r) == 6 for r in rows)
    nr, npoly, nth = 12, 3, 4
    p = O.Pot(O.log_grid(nr, 0.05, 50.0), npoly, nth)
    p.set_tables(*p.compute_phil(p.guess_rhl("Hernquist")))
    assert p.tabulate_delta() == 0
    dfo = O.df_params("Hernquist")
    ref = p.compute_new_phi(dfo, nthreads=16)
    p.set_tables(ref["phil"], ref["Pr"], ref["Pr2"])
    half = 75
    v = [-10 ** (-3 + (half - 1 - i) * 3 / (half - 1)) for i in range(half)] + \
        [10 ** (-3 + (i - half) * 3 / (half - 1)) for i in range(half, 150)]
    for i in (0, 74, 75, 149):
        for col, k in ((1, 0), (3, nr * 2 // 3), (5, nr * 5 // 6)):
            lp = p.line_profile(dfo, ref["rhl"], p.ar[k], p.ar[0], 3 * v[i])
            assert abs(rows[i][col] - lp) <= 5e-6 * abs(lp), (i, col, rows[i][col], lp)


def test_driver_fast_precision_near_fp64(tmp_path):
    """precision fast: the driver's fast node phase (the reference's turning
    point walks, FP32 roots, 64-point Gauss-Legendre action integrals) on the
    benchmark's radial grid (40 log radii in [0.01, 100], 4 cos(theta)
    nodes) gives each iteration's coefficients within 2e-4 of the FP64 run's
    monopole at the radius.  The rule's error grows with the radial cell
    width (the potential is interpolated piecewise in r, so the integrands
    have kinks): on 12 radii over the same decades it is 6e-3."""
    param = PARAM.replace("nr 12", "nr 40").replace("rmin 0.05", "rmin 0.01").replace("rmax 50", "rmax 100")
    runs = {}
    for prec in ("fp64", "fast"):
        d = tmp_path / prec
        d.mkdir()
        pf = tmp_path / f"param_{prec}.txt"
        pf.write_text(param + f"precision {prec}\n")
        run_driver(pf, d)
        runs[prec] = [read_f64(d / f"e2e_{it}.out.f64", 40, 3)[1] for it in range(2)]
    for it in range(2):
        a, b = runs["fp64"][it], runs["fast"][it]
        for n in ("rhl", "sigRl", "sigzl", "phil"):
            err = np.max(np.abs(b[n] - a[n]) / np.abs(a[n][:, :1]))
            print("iteration", it, n, "fast vs FP64:", err)
            assert err < 2e-4, (it, n, err)
