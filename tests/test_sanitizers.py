"""AddressSanitizer + UndefinedBehaviorSanitizer build of the oracle and of
the host build of the device node code (tests/sanitize): a guess potential,
its Poisson solve and Delta table, one SCF iteration, the virial and the
observables on the oracle, and the exact and fast per-point moments of the
host build at two points (the exact ones equal to the oracle's bit for bit).
The run must end cleanly with no sanitizer report (leaks included)."""
import os
import subprocess

from conftest import ROOT

SAN = os.path.join(ROOT, "tests", "sanitize")


def test_oracle_and_host_build_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", SAN], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(SAN, "_build", "san_main")], capture_output=True, text=True, timeout=600,
                       env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "san_main: ok" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
