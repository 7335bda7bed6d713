"""Full SCF trajectories against the oracle, bit for bit.

* The single-component Hernquist model (40 log radii in [0.01, 100] x 20
  cos(theta) nodes, l <= 8, log22 rule) from its guess potential through
  five iterations, as the reference's oneComp runs it (fJmodels.cpp:86-144):
  the Delta table, every coefficient table after every iteration and the
  vir2 numbers against the oracle's trajectory committed in
  tests/golden/cfg1_hernquist_40x20.npz (written by make_cfg1_trajectory.py
  next to it), once through the C ABI and once through the fjmodels driver.
* Two components through the driver (fJmodels.cpp:38-81 twoComp), in the
  reference's flow, where component 1's computePhil installs its new
  potential as the global one (pot/Potential.cpp:182-188) and component 2's
  moments are computed in it (fJmodels.cpp:69-73), and in the flow where both
  read the previous total (comp2_potential total); the summed virial of both
  components in the total potential against the oracle's.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ["rhl", "vrotl", "sigRl", "sigpl", "sigzl", "sigRzl", "phil", "Pr", "Pr2"]
FIX = os.path.join(ROOT, "tests", "golden", "cfg1_hernquist_40x20.npz")


def fixture():
    return np.load(FIX, allow_pickle=False)


def read_f64(path, nr, npoly):
    tok = open(path).read().split()
    v = np.array([float(x) for x in tok[12:]])   # header: 3 + 4 + 3 + 2 values
    ar, rest = v[:nr], v[nr:]
    return ar, {n: rest[i * nr * npoly:(i + 1) * nr * npoly].reshape(nr, npoly) for i, n in enumerate(NAMES)}


def run_driver(param_text, outdir):
    exe = os.path.join(ROOT, "fjmodels_amd", "fjmodels")
    assert os.path.exists(exe), "build the package first (make -C fjmodels_amd all)"
    pf = outdir / "param.txt"
    pf.write_text(param_text)
    print("running", exe, pf, flush=True)
    r = subprocess.run([exe, str(pf), "--outdir", str(outdir)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    with open(outdir / "fjmodels_metrics.jsonl") as f:
        return [json.loads(line) for line in f]


def vir_list(rec):
    keys = ["mass_r2Pr", "mass_sum", "KE", "PE", "W/K", "KRR", "WRR", "WRR/KRR", "Kzz", "Wzz", "Wzz/Kzz"]
    return np.array([rec["virial"][k] for k in keys])


def test_cfg1_trajectory_capi(ctx):
    import fjmodels_amd as fj
    F = fixture()
    ar = F["ar"]
    ctx.set_grid(ar, 20, 5, 30)
    ctx.set_quad(fj.FJ_QUAD_LOG22)
    ctx.set_precision(fj.FJ_PREC_FP64)
    # guess potential (Potential::computeGuessRhl + computePhil) and the Delta table on the GPU
    rhl0 = ctx.guess_rhl("Hernquist", 1.0, 1.0, 1.0)
    assert np.array_equal(rhl0, F["guess_rhl"])
    tabs = ctx.poisson(rhl0)
    for t, n in zip(tabs, ("guess_phil", "guess_Pr", "guess_Pr2")):
        assert np.array_equal(t, F[n]), n
    ctx.set_potential(*tabs)
    E, D = ctx.tabulate_delta()
    assert np.array_equal(E, F["Egrid"]) and np.array_equal(D, F["Dgrid"])
    ctx.set_df(fj.df_params("Hernquist"))
    for it in range(5):
        g = ctx.compute_new_phi()
        bad = [n for n in NAMES if not np.array_equal(g[n], F[n][it])]
        print("iteration", it, "differing tables", bad, flush=True)
        assert not bad, (it, {n: float(np.max(np.abs(g[n] - F[n][it]) / np.maximum(np.abs(F[n][it]), 1e-300)))
                              for n in bad})
        ctx.set_potential(g["phil"], g["Pr"], g["Pr2"])
        v = ctx.virial(g["rhl"], g["vrotl"], g["sigRl"], g["sigpl"], g["sigzl"])
        assert np.array_equal(v, F["vir2"][it]), (it, v, F["vir2"][it])


CFG1 = """itermax 5
model Hernquist
outfile cfg1
h(J):dphi 0.55
h(J):dz 0.55
g(J):dphi 1.
g(J):dz 1.
chi 0.
mass 1.
r0 1.
q 1.
nr 40
ntheta 20
npoly 5
rmin 0.01
rmax 100
"""


def test_cfg1_trajectory_driver(tmp_path):
    F = fixture()
    recs = run_driver(CFG1, tmp_path)
    assert [r["iter"] for r in recs] == list(range(5))
    for it in range(5):
        ar, got = read_f64(tmp_path / f"cfg1_{it}.out.f64", 40, 5)
        assert np.array_equal(ar, F["ar"])
        for n in NAMES:
            assert np.array_equal(got[n], F[n][it]), (it, n)
        assert np.array_equal(vir_list(recs[it]), F["vir2"][it]), it
        st = recs[it]["node_phase"][0]
        assert st["bracket_failures"] == 0 and st["nonfinite_nodes"] == 0 and st["unbound_nodes"] == 0
        assert st["phi_evals"] > 21 * st["qk_calls"] and st["grad_evals"] > 0


TWO = """itermax 2
model Hernquist
outfile c1
h(J):dphi 0.5
h(J):dz 0.9
g(J):dphi 1.
g(J):dz 1.
chi 0.
mass 1.
r0 1.
q 1.
2:model NFW
2:outfile c2
2:h(J):dphi 1.1
2:h(J):dz 0.6
2:g(J):dphi 0.7
2:g(J):dz 1.2
2:chi 0.
2:mass 10.
2:r0 2.5
2:q 1.
nr 13
ntheta 4
npoly 3
rmin 0.01
rmax 100
"""


@pytest.mark.parametrize("flow", ["updated", "total"])
def test_two_component_driver_matches_oracle(tmp_path, flow):
    recs = run_driver(TWO + f"comp2_potential {flow}\n", tmp_path)
    nr, npoly, nth = 13, 3, 4
    p = O.Pot(O.log_grid(nr, 0.01, 100.0), npoly, nth)
    t1 = p.compute_phil(p.guess_rhl("Hernquist", 1.0, 1.0, 1.0))
    t2 = p.compute_phil(p.guess_rhl("NFW", 10.0, 2.5, 1.0))
    tot = [a + b for a, b in zip(t1, t2)]      # updatePhil (potLeg.cpp:21-29)
    p.set_tables(*tot)
    assert p.tabulate_delta() == 0
    df1 = O.df_params("Hernquist", mass=1.0, r0=1.0, dphi_h=0.5, dz_h=0.9)
    df2 = O.df_params("NFW", mass=10.0, r0=2.5, dphi_h=1.1, dz_h=0.6, dphi_g=0.7, dz_g=1.2)
    for it in range(2):
        o1 = p.compute_new_phi(df1, nthreads=16)
        if flow == "updated":   # computePhil of component 1 installs its potential (Potential.cpp:182-188)
            p.set_tables(o1["phil"], o1["Pr"], o1["Pr2"])
        o2 = p.compute_new_phi(df2, nthreads=16)
        tot = [o1[k] + o2[k] for k in ("phil", "Pr", "Pr2")]
        p.set_tables(*tot)
        for name, ref in (("c1", o1), ("c2", o2)):
            ar, got = read_f64(tmp_path / f"{name}_{it}.out.f64", nr, npoly)
            for n in NAMES:
                assert np.array_equal(got[n], ref[n]), (flow, it, name, n)
        v1 = p.vir2(o1["rhl"], o1["vrotl"], o1["sigRl"], o1["sigpl"], o1["sigzl"])
        v2 = p.vir2(o2["rhl"], o2["vrotl"], o2["sigRl"], o2["sigpl"], o2["sigzl"])
        v = v1 + v2
        v[0] = v1[0]
        v[4], v[7], v[10] = v[3] / v[2], v[6] / v[5], v[9] / v[8]
        got = vir_list(recs[it])
        print(flow, it, "W/K", got[4], "oracle", v[4], flush=True)
        assert np.array_equal(got, v), (flow, it, got, v)


def test_two_component_flows_differ(tmp_path):
    """The two flows are different computations: component 2's moments in
    component 1's new potential versus in the previous total."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    run_driver(TWO, a)
    run_driver(TWO + "comp2_potential total\n", b)
    _, ga = read_f64(a / "c2_0.out.f64", 13, 3)
    _, gb = read_f64(b / "c2_0.out.f64", 13, 3)
    _, ha = read_f64(a / "c1_0.out.f64", 13, 3)
    _, hb = read_f64(b / "c1_0.out.f64", 13, 3)
    assert np.array_equal(ha["rhl"], hb["rhl"])            # component 1 reads the same total in both
    assert not np.array_equal(ga["rhl"], gb["rhl"])
