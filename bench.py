#!/usr/bin/env python3
"""Benchmark: BRUSS2D N=2000 full solves on one MI355X.

Workload (one bench "step"): integrate BRUSS2D-MIX on the 2000 x 2000 grid
(n = 8e6 doubles, alpha = 2e-3) over t in [0, 1] with DOPRI5(4) at tol 1e-6,
h0 = 1e-3, with the stage-fused low-storage kernel (PipeDls form) and device
step control: 19743 attempted RK steps per solve.  Inputs are resident in
HBM when the timed region starts.

value  = attempted RK steps per second over the K timed solves.
roofline: the dominant kernel is the fused step kernel (one launch per RK
step).  Algorithmic bytes per launch = 16 n (read y once, write y_new once);
achieved = 16 n / (timed GPU time / attempted steps), timed with HIP events
on the stream the kernels run on.
cpu_baseline: the reference's pthreads PipeD4 (built from the reference's C
sources into oracle/_ref) on the host's cores for a 200-step sample of the
same workload, or the oracle's pthreads D restatement when that binary is
absent.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
N_GRID = 2000
METRIC = "BRUSS2D N=2000 steps/s & HBM GB/s; batched solves/s + p99 latency vs plain"


def traffic_from_profile():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(REPO, "profiles", "sweep_n2000_dopri54_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return float(d["bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(threads):
    ref = os.path.join(REPO, "oracle", "_ref", "pthreads_bruss_PipeD4_n2000_dopri54_sl200")
    if os.path.exists(ref):
        env = dict(os.environ, NUM_THREADS=str(threads))
        out = subprocess.run([ref], capture_output=True, text=True, env=env, timeout=600,
                             cwd="/tmp").stdout
        m = re.search(r"Kernel time: (\S+) s \((\S+) s per step\)", out)
        steps = re.search(r"Number of steps: (\d+)", out)
        per_step = float(m.group(2))
        return {"value": 1.0 / per_step, "unit": "steps/s", "cores": threads,
                "kind": "reference",
                "sample": "pthreads PipeD4 (reference C sources), BRUSS2D N=2000 DOPRI54 tol 1e-6 "
                          "h0 1e-3, first %s attempted steps (STEP_LIMIT=200)" % steps.group(1)}
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    p = pyoracle.make_ode("bruss2d", N_GRID)
    y0 = pyoracle.ode_start(p)
    _, st, _ = pyoracle.solve(p, "DOPRI54", 0.0, 1.0, y0, 1e-6, h0=1e-3, step_limit=60,
                              impl="pthreads_D", threads=threads)
    return {"value": st["steps"] / st["kernel_time"], "unit": "steps/s", "cores": threads,
            "kind": "port",
            "sample": "oracle pthreads D, BRUSS2D N=2000 DOPRI54 tol 1e-6 h0 1e-3, first 60 "
                      "attempted steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--impl", default="PipeDls")
    ap.add_argument("--method", default="DOPRI54")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--tile-waves", type=int, default=0)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--step-limit", type=int, default=0,
                    help="cap each solve at this many attempted steps (profiling runs)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run 5 plain device copies of n doubles, the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(see DESIGN.md for how it would shard); use --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import rk_amd

    assert torch.cuda.is_available()
    stream = torch.cuda.Stream()
    p = rk_amd.Bruss2D(N_GRID, 2e-3)
    y0 = torch.from_numpy(p.ode_start()).cuda()
    y = torch.empty_like(y0)
    torch.cuda.synchronize()
    kw = dict(h0=1e-3, impl=a.impl, tile_waves=a.tile_waves, tile_rows=a.tile_rows,
              stream=stream.cuda_stream, step_limit=a.step_limit)
    if a.calib_copy:
        for _ in range(5):
            y.copy_(y0)                  # reads 8n bytes, writes 8n bytes
        torch.cuda.synchronize()

    for _ in range(a.warmup):
        st, _ = rk_amd.solve(p, a.method, 0.0, 1.0, y0, y, a.tol, **kw)
        print("warmup: %d steps (%d acc, %d rej) in %.3f s" % (st.steps, st.steps_acc,
                                                              st.steps_rej, st.kernel_time),
              file=sys.stderr, flush=True)

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    ev0.record(stream)
    total_steps = 0
    stats = []
    for k in range(a.steps):
        st, _ = rk_amd.solve(p, a.method, 0.0, 1.0, y0, y, a.tol, **kw)
        total_steps += st.steps
        stats.append(st.as_dict())
        print("solve %d: %d steps (%d acc, %d rej) in %.3f s" % (k, st.steps, st.steps_acc,
                                                                st.steps_rej, st.kernel_time),
              file=sys.stderr, flush=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    gpu_s = ev0.elapsed_time(ev1) / 1e3

    n = p.n
    steps_per_s = total_steps / wall
    per_launch_s = gpu_s / total_steps
    achieved = 16.0 * n / per_launch_s / 1e9
    yv = y.cpu().numpy()
    res = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (BRUSS2D start vector of bruss2d-mix.c:51-59)",
        "config": {"workload": "BRUSS2D N=2000 full solve t in [0,1], %s tol %g h0 1e-3, "
                               "%s fused step kernel, device step control" % (a.method, a.tol,
                                                                              a.impl),
                   "grid": N_GRID, "n": n, "method": a.method, "tol": a.tol,
                   "attempted_steps_per_solve": stats[-1]["steps"],
                   "accepted": stats[-1]["acc"], "rejected": stats[-1]["rej"],
                   "hbm_gbs_algorithmic": achieved,
                   "resident_bytes": stats[-1]["resident_bytes"],
                   "footprint": [float(yv[0]), float(yv[-1]), float(yv.min()), float(yv.max())]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile()},
    }
    if not a.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        res["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
