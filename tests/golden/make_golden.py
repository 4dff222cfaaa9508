#!/usr/bin/env python3
"""Write golden fixtures from the reference binaries in oracle/_ref/.

Each dump_* binary is the reference's seq D built with -DDUMP
(include/tools.h:239-256): for every attempted step it appends the step's
t, h and the full tentative y vector to ./DUMP at %.20e (which round-trips
exactly to binary64).  This script runs each binary in a scratch directory
and stores, per case, the arrays t[k], h[k], Y[k, :] plus the printed step
counts and Result footprint in tests/golden/<case>.npz (numpy, no pickle).

The reference's other loop orders A, E and F (src/impl/seq/A.c, E.c, F.c)
are run on a subset of the same cases and on three known answers; they must
reproduce D's trajectory and result bitwise, and the index / kat.json record
where they were checked ("same_in", "ref_impls").

Run after `python oracle/build_ref.py`.  The fixtures are data (inputs and
expected outputs); the script that made them is this file.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import build_ref  # noqa: E402  (config table only)

STEPS_RE = re.compile(r"Number of steps: (\d+) \((\d+) accepted, (\d+) rejected\)")
RESULT_RE = re.compile(r"Result: y\[0\] = (\S+), \.\.\., y\[(\d+)\] = (\S+); min = (\S+), max = (\S+)")


def run_case(name):
    kind, prob, impl, defs = build_ref.CONFIGS[name]
    exe = os.path.join(build_ref.OUT, name)
    with tempfile.TemporaryDirectory() as td:
        out = subprocess.run([exe], cwd=td, capture_output=True, text=True, check=True).stdout
        rows = []
        with open(os.path.join(td, "DUMP")) as f:
            for line in f:
                rows.append([float(x) for x in line.split("\t")])
    arr = np.array(rows, dtype=np.float64)
    m1 = STEPS_RE.search(out)
    m2 = RESULT_RE.search(out)
    meta = dict(name=name, problem=prob, defines={k: str(v) for k, v in defs.items()},
                steps=int(m1.group(1)), acc=int(m1.group(2)), rej=int(m1.group(3)),
                footprint=[float(m2.group(1)), float(m2.group(3)), float(m2.group(4)),
                           float(m2.group(5))])
    return arr, meta


def run_result(name):
    """Step counts and Result footprint printed by a reference binary."""
    exe = os.path.join(build_ref.OUT, name)
    with tempfile.TemporaryDirectory() as td:
        out = subprocess.run([exe], cwd=td, capture_output=True, text=True, check=True).stdout
    m1 = STEPS_RE.search(out)
    m2 = RESULT_RE.search(out)
    return (int(m1.group(2)), int(m1.group(3)),
            [float(m2.group(i)) for i in (1, 3, 4, 5)])


def gather_kat():
    """The reference's A, E and F on known-answer cases: each must print the
    case's step counts and footprint; the loop orders that do are recorded in
    the case's "ref_impls"."""
    path = os.path.join(HERE, "kat.json")
    kat = json.load(open(path))
    by_name = {c["name"]: c for c in kat["cases"]}
    for case, (prob, tag, defs) in build_ref.GATHER_KAT.items():
        c = by_name[case]
        impls = ["D"]
        for impl in build_ref.GATHER_IMPLS:
            acc, rej, fp = run_result("seq_%s_%s_%s" % (tag, impl, case))
            if (acc, rej) != (c["acc"], c["rej"]) or fp != [float(x) for x in c["footprint"]]:
                raise SystemExit("reference %s differs from D on %s" % (impl, case))
            impls.append(impl)
        c["ref_impls"] = impls
        print("%-30s reference %s: same steps and footprint" % (case, "/".join(impls)))
    with open(path, "w") as f:       # one case a line
        f.write('{\n "_source": %s,\n "cases": [\n' % json.dumps(kat["_source"]))
        f.write(",\n".join("  " + json.dumps(c) for c in kat["cases"]))
        f.write("\n ]\n}\n")


def main():
    names = [n for n in build_ref.CONFIGS if n.startswith("dump_")
             and build_ref.CONFIGS[n][2] == "D"]
    index = {}
    for n in names:
        arr, meta = run_case(n)
        np.savez_compressed(os.path.join(HERE, n + ".npz"), t=arr[:, 0], h=arr[:, 1], Y=arr[:, 2:])
        # the reference's other loop orders on the same case: the trajectory
        # must be the same, bitwise (no fixture of their own)
        same = []
        for impl in build_ref.GATHER_IMPLS:
            if n + "_" + impl in build_ref.CONFIGS:
                other, m2 = run_case(n + "_" + impl)
                if not np.array_equal(other, arr) or (m2["acc"], m2["rej"]) != (meta["acc"],
                                                                                meta["rej"]):
                    raise SystemExit("reference %s trajectory differs from D on %s" % (impl, n))
                same.append(impl)
        if same:
            meta["same_in"] = same
        index[n] = meta
        print("%-28s attempts=%d n=%d acc=%d rej=%d%s" % (
            n, arr.shape[0], arr.shape[1] - 2, meta["acc"], meta["rej"],
            ("  (= reference " + "/".join(same) + ")") if same else ""))
    with open(os.path.join(HERE, "dump_index.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)
    gather_kat()
    return 0


if __name__ == "__main__":
    sys.exit(main())
