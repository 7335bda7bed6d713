"""C++ facade (fjmodels_amd/host): parameter files, output names and the
models/*.out format, on the CPU.  The reference's behaviour these follow:
io/readParam.cpp:31-144 (keys, skipped lines, component count),
io/writeOut.cpp:22-104 (composeName with 2-decimal fields and the 5-char model
prefix; header lines, ar, nine [NR][NPOLY] matrices with a trailing space,
ostream's default 6 significant digits)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FAC = os.path.join(ROOT, "tests", "facade")

PARAM = """#
# comment lines
#

itermax\t\t7
\t
model\t\tHernquist
h(J):dphi\t0.55
h(J):dz\t\t0.55
g(J):dphi\t1.
g(J):dz\t\t1.
chi\t\t0.
mass\t\t1.
r0\t\t1.
q\t\t1.
 mass 99
nr 12
ntheta 4
npoly 3
rmin 0.05
rmax 50
rot_form 1
tol 1e-5
poisson blockscan
precision mixed
quad tanhsinh22

#2:model NFW
2:A\t\t  0.
2:B\t\t  2.
2:h(J):dphi       0.55
2:h(J):dz         2.
2:g(J):dphi       1.
2:g(J):dz         4.
2:chi\t\t  10.
2:mass            2.
2:r0              2.
2:q               .6
"""


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "fjmodels_amd"), "all"], check=True)
    subprocess.run(["make", "-s", "-C", FAC], check=True)
    d = tmp_path_factory.mktemp("fac")
    pf = d / "param.txt"
    pf.write_text(PARAM)
    out = subprocess.run([os.path.join(FAC, "facade_check"), str(pf), str(d)], capture_output=True, text=True,
                         check=True).stdout
    kv = {}
    for line in out.splitlines():
        k, _, v = line.partition(" ")
        kv[k] = v
    return kv, d


def test_param_file(run):
    kv, _ = run
    assert kv["components"] == "2"          # 2:A/2:B set, 2:model commented out
    assert kv["itermax"] == "7"
    assert kv["model"] == "Hernquist" and kv["outfile"] == "null"
    assert kv["chi"].split()[2] == "1"       # " mass 99" starts with a space: skipped
    f = kv["dphi_h"].split()        # "v1 dz_h v2 dphi_g v3 dz_g v4"
    assert [float(x) for x in f[0::2]] == [0.55, 0.55, 1.0, 1.0]
    assert kv["A2"].split() == ["0", "B2", "2", "mass2", "2"]
    assert kv["grid"].split() == ["12", "4", "3", "30", "0.050000000000000003", "50"]
    assert kv["rot"].split()[0] == "1" and float(kv["rot"].split()[3]) == 1e-5
    assert kv["modes"].split() == ["poisson", "1", "precision", "1"]     # FJ_POISSON_BLOCKSCAN, FJ_PREC_MIXED


def test_bad_mode_keys_are_rejected(tmp_path):
    subprocess.run(["make", "-s", "-C", FAC], check=True)
    for line in ("precision half", "poisson simpson"):
        pf = tmp_path / "p.txt"
        pf.write_text("model Hernquist\nchi 0\n" + line + "\n")
        r = subprocess.run([os.path.join(FAC, "facade_check"), str(pf), str(tmp_path)], capture_output=True,
                           text=True)
        assert r.returncode != 0 and "must be" in (r.stdout + r.stderr), line


def test_compose_name(run):
    kv, _ = run
    assert kv["name1"] == "Hernq_0.55_0.55_1.00_1.00_3.out"
    assert kv["name2"] == "null_0.55_2.00_1.00_4.00_3.out2c"


def test_grid_is_log_spaced(run):
    kv, _ = run
    ar = np.array([float(x) for x in kv["ar"].split()])
    ref = 10 ** (np.log10(0.05) + (np.log10(50) - np.log10(0.05)) * np.arange(12) / 11)
    assert len(ar) == 12 and np.allclose(ar, ref, rtol=1e-15, atol=0)


def test_out_file_layout_and_roundtrip(run):
    kv, d = run
    assert kv["f64_roundtrip"] == "1"
    assert float(kv["out_maxrel"]) < 5e-6      # 6 significant digits
    assert kv["header"].split() == ["12", "3", "4", "-999", "-999"]   # A, B unset until setDF
    path = kv["path"]
    assert os.path.basename(path) == "Hernq_0.55_0.55_1.00_1.00_7.out"
    lines = open(path).read().split("\n")
    assert lines[0] == "12 3 4"
    assert lines[1] == "0.55 0.55 1 1"
    assert lines[2] == "0 1 1"
    assert lines[3] == "-999 -999"
    assert len(lines) == 4 + 12 + 9 * 12 + 1 and lines[-1] == ""
    row = lines[4 + 12]
    assert row.endswith(" ") and len(row.split()) == 3
    # ostream default: 6 significant digits, shortest form
    assert row.split()[0] == "0.333333"
