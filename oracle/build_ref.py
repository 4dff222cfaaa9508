#!/usr/bin/env python3
"""Build reference binaries from the reference's own C sources into oracle/_ref/.

Test infrastructure only.  Each binary is one gcc invocation over the
reference's few source files (src/main.c, src/tools.c, src/methods.c,
src/threads.c, one problem and one implementation), with the reference's own
flags (Makefile:46: -O3 -g -D_GNU_SOURCE -D_REENTRANT, -lpthread -lm) and the
compile-time configuration as -D defines (include/config.h).  Nothing is
copied out of the reference tree; outputs go to oracle/_ref/ only.

The drop-in proof (REF_DROPIN) links the reference's own driver
(src/main.c) and problem object (src/problems/<problem>.c), compiled the same
way, with this build's solver(), method functions and device problem
descriptor (rk_amd/csrc/dropin: solver_hip.c, methods_dropin.c,
desc_<problem>.c, all with the same -D defines) over rk_amd/librkx.so -- the
rule INTEGRATION.md gives a maintainer of the reference.  Those binaries are
built only when librkx.so exists.

Usage: python oracle/build_ref.py [--ref /root/reference]
Does nothing (exit 0) when the reference tree is absent.
"""
import argparse
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")

# name -> (kind, problem, impl, defines)
CONFIGS = {
    # known answers (footprints + step counts)
    "seq_bruss_D_n128_dopri54_tol1e-6": ("seq", "bruss2d-mix", "D",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
    "seq_bruss_PipeDls_n128_dopri54_tol1e-6": ("seq", "bruss2d-mix", "PipeDls",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
    "seq_bruss_D_n128_verner65_tol1e-6": ("seq", "bruss2d-mix", "D",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="VERNER65")),
    "seq_bruss_D_n128_verner65_tol1e-8": ("seq", "bruss2d-mix", "D",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-8", H_INIT="1E-3", STEP_LIMIT=0, METHOD="VERNER65")),
    "seq_bruss_D_n128_defaults": ("seq", "bruss2d-mix", "D", dict(BRUSS_GRID_SIZE=128)),
    "seq_lv_D_10000_dopri54": ("seq", "lv", "D",
        dict(LV_COUNT=10000, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
    "seq_lv_D_10000_verner65_t20": ("seq", "lv", "D",
        dict(LV_COUNT=10000, T_END="20.0", TOL="1E-8", STEP_LIMIT=0, METHOD="VERNER65")),
    # CPU baseline: pthreads PipeD4 on the 2000x2000 grid, 200-step cap
    "pthreads_bruss_PipeD4_n2000_dopri54_sl200": ("pthreads", "bruss2d-mix", "PipeD4",
        dict(BRUSS_GRID_SIZE=2000, T_END="1.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=200, METHOD="DOPRI54")),
    "pthreads_bruss_D_n2000_dopri54_sl200": ("pthreads", "bruss2d-mix", "D",
        dict(BRUSS_GRID_SIZE=2000, T_END="1.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=200, METHOD="DOPRI54")),
    # C1 in full on the host's cores (CPU baseline beside the GPU's C1 solve)
    "pthreads_bruss_PipeD4_n128_dopri54_tol1e-6": ("pthreads", "bruss2d-mix", "PipeD4",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
    "pthreads_bruss_D_n128_dopri54_tol1e-6": ("pthreads", "bruss2d-mix", "D",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
    # single-thread baseline: seq D, 20-step cap (about 10 s)
    "seq_bruss_D_n2000_dopri54_sl20": ("seq", "bruss2d-mix", "D",
        dict(BRUSS_GRID_SIZE=2000, T_END="1.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=20, METHOD="DOPRI54")),
}

# the reference's other loop orders (src/impl/seq/A.c, E.c, F.c) on known-answer
# cases: KAT name -> the defines of that case; binaries seq_<tag>_<impl>_<case>
GATHER_IMPLS = ("A", "E", "F")
GATHER_KAT = {
    "c1_bruss128_dopri54_tol1e-6": ("bruss2d-mix", "bruss",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
    "bruss128_verner65_tol1e-8": ("bruss2d-mix", "bruss",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-8", H_INIT="1E-3", STEP_LIMIT=0, METHOD="VERNER65")),
    "lv10000_dopri54": ("lv", "lv",
        dict(LV_COUNT=10000, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
}
for kat, (prob, tag, defs) in GATHER_KAT.items():
    for impl in GATHER_IMPLS:
        CONFIGS["seq_%s_%s_%s" % (tag, impl, kat)] = ("seq", prob, impl, defs)

# DUMP trajectories (every attempted step, full vector) on tiny grids
DUMP_CASES = [
    ("bruss2d-mix", dict(BRUSS_GRID_SIZE=4), "DOPRI54", "4.0", "1E-6", "1E-3"),
    ("bruss2d-mix", dict(BRUSS_GRID_SIZE=5), "VERNER65", "4.0", "1E-6", "1E-3"),
    ("bruss2d-mix", dict(BRUSS_GRID_SIZE=7), "DOPRI54", "1.0", "1E-8", "1E-2"),
    ("bruss2d-mix", dict(BRUSS_GRID_SIZE=6), "RKF23", "0.5", "1E-5", "1E-3"),
    ("bruss2d-mix", dict(BRUSS_GRID_SIZE=3), "DOPRI87", "2.0", "1E-9", "1E-3"),
    ("bruss2d-mix", dict(BRUSS_GRID_SIZE=4), "RK4", "0.2", "0.0", "1E-2"),
    ("bruss2d-mix", dict(BRUSS_GRID_SIZE=2), "HEUN2", "0.1", "0.0", "1E-2"),
    ("bruss2d-mix", dict(BRUSS_GRID_SIZE=3), "SSPRK3", "0.1", "0.0", "1E-2"),
    ("lv", dict(LV_COUNT=3), "DOPRI54", "20.0", "1E-8", "1E-3"),
    ("lv", dict(LV_COUNT=2), "VERNER65", "10.0", "1E-6", "1E-2"),
]
# the same trajectories from the reference's A, E and F (one case per method family)
GATHER_DUMPS = ("dump_bruss_n4_dopri54", "dump_bruss_n5_verner65", "dump_bruss_n3_dopri87",
                "dump_bruss_n6_rkf23", "dump_lv_n3_dopri54")
for prob, d, meth, te, tol, h0 in DUMP_CASES:
    tag = "n%d" % d.get("BRUSS_GRID_SIZE", d.get("LV_COUNT"))
    name = "dump_%s_%s_%s" % ("bruss" if prob.startswith("bruss") else "lv", tag, meth.lower())
    defs = dict(d, T_END=te, TOL=tol, H_INIT=h0, STEP_LIMIT=0, METHOD=meth, DUMP=1)
    CONFIGS[name] = ("seq", prob, "D", defs)
    if name in GATHER_DUMPS:
        for impl in GATHER_IMPLS:
            CONFIGS[name + "_" + impl] = ("seq", prob, impl, defs)


# name -> (reference problem source, drop-in descriptor, defines); each
# mirrors a known answer of tests/golden/kat.json
REF_DROPIN = {
    "refmain_bruss_c1": ("bruss2d-mix", "bruss2d",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
    "refmain_bruss_defaults": ("bruss2d-mix", "bruss2d", dict(BRUSS_GRID_SIZE=128)),
    "refmain_bruss_verner65_tol1e-8": ("bruss2d-mix", "bruss2d",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-8", H_INIT="1E-3", STEP_LIMIT=0, METHOD="VERNER65")),
    "refmain_lv_dopri54": ("lv", "lv",
        dict(LV_COUNT=10000, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54")),
    "refmain_lv_verner65_t20": ("lv", "lv",
        dict(LV_COUNT=10000, T_END="20.0", TOL="1E-8", STEP_LIMIT=0, METHOD="VERNER65")),
    # the same driver with this build's gather loop orders behind solver()
    "refmain_bruss_c1_F": ("bruss2d-mix", "bruss2d",
        dict(BRUSS_GRID_SIZE=128, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54",
             RKX_DROPIN_IMPL="RKX_IMPL_F")),
    "refmain_lv_dopri54_A": ("lv", "lv",
        dict(LV_COUNT=10000, T_END="4.0", TOL="1E-6", H_INIT="1E-3", STEP_LIMIT=0, METHOD="DOPRI54",
             RKX_DROPIN_IMPL="RKX_IMPL_A")),
}
REPO = os.path.dirname(HERE)
DROPIN_DIR = os.path.join(REPO, "rk_amd", "csrc", "dropin")
LIBRKX = os.path.join(REPO, "rk_amd", "librkx.so")


def build_ref_dropin(ref, name, prob, desc, defs):
    """The reference's main.o and problem object + this build's solver,
    methods and descriptor objects + librkx.so (found through an rpath
    relative to the binary, so the tree can move)."""
    with tempfile.TemporaryDirectory() as objdir:
        return _link_ref_dropin(ref, name, prob, desc, defs, objdir)


def _link_ref_dropin(ref, name, prob, desc, defs, objdir):
    out = os.path.join(OUT, name)
    dflags = ["-D%s=%s" % (k, v) for k, v in defs.items()]
    base = ["gcc", "-I" + os.path.join(ref, "include"), "-D_REENTRANT", "-O3", "-g",
            "-D_GNU_SOURCE"] + dflags
    objs = []
    for src in (os.path.join(ref, "src", "main.c"),
                os.path.join(ref, "src", "problems", prob + ".c")):
        o = os.path.join(objdir, os.path.basename(src)[:-2] + ".o")
        r = subprocess.run(base + ["-c", src, "-o", o], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile of %s for %s failed:\n%s" % (src, name, r.stderr[-2000:]))
        objs.append(o)
    r = subprocess.run(["make", "-s", "-C", DROPIN_DIR, "objs", "PROBLEM=" + desc,
                        "DEFINES=" + " ".join("%s=%s" % kv for kv in defs.items()),
                        "OBJDIR=" + objdir], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in objects for %s failed:\n%s" % (name, r.stdout + r.stderr))
    objs += [os.path.join(objdir, f) for f in ("solver_hip.o", "methods_dropin.o",
                                               "desc_%s.o" % desc)]
    cmd = ["gcc", "-o", out] + objs + ["-L" + os.path.dirname(LIBRKX), "-lrkx",
                                       "-Wl,-rpath,$ORIGIN/../../rk_amd", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link of %s failed:\n%s" % (name, r.stderr[-2000:]))
    return name


def build_one(ref, name, kind, prob, impl, defs):
    out = os.path.join(OUT, name)
    srcs = [os.path.join(ref, "src", f) for f in ("main.c", "tools.c", "methods.c", "threads.c")]
    srcs += [os.path.join(ref, "src", "problems", prob + ".c"),
             os.path.join(ref, "src", "impl", kind, impl + ".c")]
    cmd = ["gcc", "-I" + os.path.join(ref, "include"), "-D_REENTRANT", "-O3", "-g", "-D_GNU_SOURCE"]
    cmd += ["-D%s=%s" % (k, v) for k, v in defs.items()]
    cmd += ["-o", out] + srcs + ["-lpthread", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build of %s failed:\n%s" % (name, r.stderr[-2000:]))
    return name


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", default=None, help="comma-separated config names")
    a = ap.parse_args()
    if not os.path.isdir(os.path.join(a.ref, "src")):
        print("reference tree not present; oracle/_ref not built")
        return 0
    os.makedirs(OUT, exist_ok=True)
    names = list(CONFIGS) if not a.only else a.only.split(",")
    names = [n for n in names if n in CONFIGS]
    dropins = [n for n in (list(REF_DROPIN) if not a.only else a.only.split(","))
               if n in REF_DROPIN]
    if dropins and not os.path.exists(LIBRKX):
        print("rk_amd/librkx.so not built: skipping the reference-driver drop-in binaries")
        dropins = []
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = [ex.submit(build_one, a.ref, n, *CONFIGS[n]) for n in names]
        futs += [ex.submit(build_ref_dropin, a.ref, n, *REF_DROPIN[n]) for n in dropins]
        for f in futs:
            f.result()
    print("built %d reference binaries (%d of them reference-driver drop-ins) into %s"
          % (len(names) + len(dropins), len(dropins), OUT))
    return 0


if __name__ == "__main__":
    sys.exit(main())
