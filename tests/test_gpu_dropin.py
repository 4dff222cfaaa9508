"""The reference-style binaries (solver() of solver.h:47 over the C ABI,
built from rk_amd/csrc/dropin with the reference's build-time defines) print
the reference's output contract, and their Result line equals the
reference's known answer for the same defines, digit for digit (%.20e, so
bitwise), with the same step counts.  Each binary runs as a child process.

The reference-driver binaries (oracle/_ref/refmain_*, built by
oracle/build_ref.py) go further: the reference's own src/main.c and problem
object (src/problems/bruss2d-mix.c, lv.c), compiled from its sources, linked
with this build's solver(), method functions and device descriptor over
librkx.so.  Their output is the reference's, line for line where the lines
do not carry timings."""
import json
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "rk_amd", "bin")
KAT = {c["name"]: c for c in json.load(open(os.path.join(HERE, "golden", "kat.json")))["cases"]}

CASES = [("rk_bruss2d_c1", "c1_bruss128_dopri54_tol1e-6"),
         ("rk_bruss2d_defaults", "bruss128_defaults"),
         ("rk_bruss2d_verner65_d", "bruss128_verner65_tol1e-8"),
         ("rk_lv_dopri54", "lv10000_dopri54")]

# the lines log2tab reads (log2tab:75-293), in its order
F = r"\d+(\.\d+)?e[+-]\d+"
LOG2TAB = [r"^ODE: \S+", r"^(Grid size|System size): \d+", r"^Integration interval: \[%s,%s\]" % (F, F),
           r"^Initial value: ", r"^Tolerance: %s" % F, r"^Lock type: \S+", r"^Pad size: \d+",
           r"^Solver type: \S+ embedded Runge-Kutta method", r"^Implementation variant: \S+",
           r"^Method: \S+.*\(order \d+, \d+ stages\)", r"^Initial step size: \d+e[+-]\d+",
           r"^Number of steps: \d+ \(\d+ accepted, \d+ rejected\)",
           r"^Kernel time: %s s \(%s s per step\)" % (F, F), r"^Total time:\s+%s s" % F,
           r"^Result: y\[0\] = %s, \.\.\., y\[\d+\] = %s;" % (F, F)]


@pytest.mark.parametrize("binary,case", CASES)
def test_dropin_binary_matches_reference(binary, case):
    path = os.path.join(BIN, binary)
    assert os.path.exists(path), "drop-in binaries are built by rk_amd/build.py"
    print("running", binary, flush=True)
    r = subprocess.run([path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    k = 0
    for ln in lines:                      # every log2tab line, in order
        if k < len(LOG2TAB) and re.match(LOG2TAB[k], ln):
            k += 1
    assert k == len(LOG2TAB), (LOG2TAB[k] if k < len(LOG2TAB) else "", r.stdout)
    c = KAT[case]
    m = re.search(r"^Number of steps: (\d+) \((\d+) accepted, (\d+) rejected\)", r.stdout, re.M)
    assert (int(m.group(2)), int(m.group(3))) == (c["acc"], c["rej"])
    m = re.search(r"^Result: y\[0\] = (\S+), \.\.\., y\[\d+\] = (\S+); min = (\S+), max = (\S+)$",
                  r.stdout, re.M)
    assert list(m.groups()) == c["footprint"]


# (drop-in binary, known answer, the reference's own seq D binary of the same defines)
REF_DRIVER = [("refmain_bruss_c1", "c1_bruss128_dopri54_tol1e-6", "seq_bruss_D_n128_dopri54_tol1e-6"),
              ("refmain_bruss_defaults", "bruss128_defaults", "seq_bruss_D_n128_defaults"),
              ("refmain_bruss_verner65_tol1e-8", "bruss128_verner65_tol1e-8",
               "seq_bruss_D_n128_verner65_tol1e-8"),
              ("refmain_lv_dopri54", "lv10000_dopri54", "seq_lv_D_10000_dopri54"),
              ("refmain_lv_verner65_t20", "lv10000_verner65_t20", "seq_lv_D_10000_verner65_t20"),
              ("refmain_bruss_c1_F", "c1_bruss128_dopri54_tol1e-6", "seq_bruss_D_n128_dopri54_tol1e-6"),
              ("refmain_lv_dopri54_A", "lv10000_dopri54", "seq_lv_D_10000_dopri54")]
REF_BIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref")
# lines that name the solver or carry a measured time
VARIES = re.compile(r"^(Solver type|Implementation variant|Kernel time|Total time):")


@pytest.mark.parametrize("binary,case,seq", REF_DRIVER)
def test_reference_driver_with_gpu_solver(binary, case, seq):
    path = os.path.join(REF_BIN, binary)
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built here (needs the reference sources at build time)")
    print("running", binary, flush=True)
    r = subprocess.run([path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Solver type: device embedded Runge-Kutta method" in r.stdout
    print("running the reference's", seq, flush=True)
    q = subprocess.run([os.path.join(REF_BIN, seq)], capture_output=True, text=True, timeout=300)
    assert q.returncode == 0
    keep = lambda out: [ln for ln in out.splitlines() if not VARIES.match(ln)]
    assert keep(r.stdout) == keep(q.stdout)
    c = KAT[case]
    m = re.search(r"^Number of steps: (\d+) \((\d+) accepted, (\d+) rejected\)", r.stdout, re.M)
    assert (int(m.group(2)), int(m.group(3))) == (c["acc"], c["rej"])
    m = re.search(r"^Result: y\[0\] = (\S+), \.\.\., y\[\d+\] = (\S+); min = (\S+), max = (\S+)$",
                  r.stdout, re.M)
    assert list(m.groups()) == c["footprint"]


DUMP_CASES = [("rk_bruss2d_dump", "dump_bruss_n4_dopri54"), ("rk_lv_dump", "dump_lv_n3_dopri54")]


@pytest.mark.parametrize("binary,case", DUMP_CASES)
def test_dropin_dump_file_matches_reference(binary, case, tmp_path):
    """Built with -DDUMP, the binary writes ./DUMP as the reference's
    step_control does (include/tools.h:239-256): every attempted step's t, h
    and new y, equal to the reference's file value for value."""
    import numpy as np
    path = os.path.join(BIN, binary)
    assert os.path.exists(path), "drop-in binaries are built by rk_amd/build.py"
    print("running", binary, flush=True)
    r = subprocess.run([path], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    z = np.load(os.path.join(HERE, "golden", case + ".npz"))
    rows = [[float(x) for x in ln.split("\t")] for ln in open(tmp_path / "DUMP")]
    got = np.array(rows)
    assert got.shape == (len(z["t"]), 2 + z["Y"].shape[1])
    assert np.array_equal(got[:, 0], z["t"]) and np.array_equal(got[:, 1], z["h"])
    assert np.array_equal(got[:, 2:], z["Y"])
