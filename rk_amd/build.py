#!/usr/bin/env python3
"""Build rk_amd/librkx.so (HIP kernels + C ABI) for gfx950 with hipcc.

Every translation unit is compiled with -ffp-contract=off (no implicit fused
multiply-add: results must equal the reference's binary64 arithmetic bitwise)
and with packed FP32 disabled.  Objects go to rk_amd/_build/, the library to
rk_amd/librkx.so.  Usage: python rk_amd/build.py [--jobs N] [--verbose]
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librkx.so")

SOURCES = [
    "rkx_host.cpp",
    "rkx_batch.cpp",
    "rkx_methods.cpp",
    "rkx_kernels_d.hip",
    "rkx_kernels_gather.hip",
    "rkx_kernels_lvpairs.hip",
    "rkx_kernels_sweep.hip",
    "rkx_kernels_sweep_dopri87.hip",
    "rkx_kernels_lv.hip",
    "rkx_kernels_chain.hip",
    "rkx_pow.hip",
    "rkx_noalloc.cpp",
]

# Reference-style solver binaries over the C ABI (rk_amd/csrc/dropin): name ->
# (problem object, build-time defines as in the reference's Makefile)
DROPIN = {
    "rk_bruss2d_c1": ("bruss2d", "BRUSS_GRID_SIZE=128 T_END=4.0 TOL=1E-6 H_INIT=1E-3 "
                                 "METHOD=DOPRI54 STEP_LIMIT=0"),
    "rk_bruss2d_defaults": ("bruss2d", "BRUSS_GRID_SIZE=128"),
    "rk_bruss2d_verner65_d": ("bruss2d", "BRUSS_GRID_SIZE=128 T_END=4.0 TOL=1E-8 H_INIT=1E-3 "
                                         "METHOD=VERNER65 STEP_LIMIT=0 "
                                         "RKX_DROPIN_IMPL=RKX_IMPL_D"),
    "rk_lv_dopri54": ("lv", "LV_COUNT=10000 T_END=4.0 TOL=1E-6 H_INIT=1E-3 METHOD=DOPRI54 "
                            "STEP_LIMIT=0"),
    # -DDUMP: the per-attempt trajectory file of include/tools.h:239-256
    "rk_bruss2d_dump": ("bruss2d", "BRUSS_GRID_SIZE=4 T_END=4.0 TOL=1E-6 H_INIT=1E-3 "
                                   "METHOD=DOPRI54 STEP_LIMIT=0 DUMP"),
    "rk_lv_dump": ("lv", "LV_COUNT=3 T_END=20.0 TOL=1E-8 H_INIT=1E-3 METHOD=DOPRI54 "
                         "STEP_LIMIT=0 DUMP"),
}
BIN = os.path.join(HERE, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RKX_ARCH", "gfx950")

CFLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(REPO, "include"),
]


def compile_one(src, verbose):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + CFLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr).strip()
    # hipcc's host pass does not know the packed-fp32 feature; that note is harmless
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (src, "\n".join(lines[-60:])))
    if verbose and lines:
        print("\n".join(lines))
    return obj


def build(jobs=4, verbose=False):
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at " + HIPCC)
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    shutil.move(tmp, LIB)
    build_dropin()
    return LIB


def build_dropin():
    """The reference-style driver binaries of DROPIN, into rk_amd/bin/."""
    for name, (problem, defines) in DROPIN.items():
        cmd = ["make", "-s", "-C", os.path.join(CSRC, "dropin"), "PROBLEM=" + problem,
               "DEFINES=" + defines, "OUT=" + os.path.join(BIN, name)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("drop-in build failed: %s\n%s" % (name, r.stdout + r.stderr))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
    return 0


if __name__ == "__main__":
    sys.exit(main())
