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
sources into oracle/_ref) on 16 host threads for a 200-step sample of the
same workload, or the oracle's pthreads D restatement when that binary is
absent.  Beside it (cpu_reference): the reference's pthreads D (200 steps)
and seq D (one thread, 20 steps) on the same workload, config 1 (N=128,
t in [0, 4]) in full with seq D, pthreads D and PipeD4, and the GPU's time
for config 1 (c1_gpu); host: CPU model, logical CPUs and physical cores.
batched / ensemble / serving: the batched solves/s against plain per-instance
launches on the 512-instance mixed BRUSS2D batch, the LV ensembles, and the
continuous-batching server's p50/p99 latency and solves/s against the plain
solver on a Poisson request stream (see the leg functions for the workloads).
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


def _run_ref(name, threads=1):
    """Run one reference binary of oracle/_ref (built from the reference's C
    sources) and parse its own timer and step lines (src/tools.c:62-79)."""
    ref = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(ref):
        return None
    env = dict(os.environ, NUM_THREADS=str(threads))
    out = subprocess.run([ref], capture_output=True, text=True, env=env, timeout=900,
                         cwd="/tmp").stdout
    m = re.search(r"Kernel time: (\S+) s \((\S+) s per step\)", out)
    steps = re.search(r"Number of steps: (\d+)", out)
    return {"kernel_s": float(m.group(1)), "s_per_step": float(m.group(2)),
            "steps": int(steps.group(1)), "threads": threads}


def host_info():
    """CPU model, logical CPUs and physical cores (lscpu) of the host the CPU
    baselines run on."""
    info = {"cpu_model": "", "logical_cpus": os.cpu_count(), "physical_cores": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        f = dict(re.findall(r"^([^:\n]+):\s*(.+)$", out, re.M))
        info["cpu_model"] = f.get("Model name", "").strip()
        info["physical_cores"] = int(f["Core(s) per socket"]) * int(f["Socket(s)"])
        info["threads_per_core"] = int(f.get("Thread(s) per core", "1"))
    except (OSError, subprocess.SubprocessError, KeyError, ValueError):
        pass
    try:
        info["cpus_allowed"] = len(os.sched_getaffinity(0))
    except OSError:
        pass
    return info


def cpu_leg(threads):
    """The reference's CPU solvers on this host, built from its C sources:
    the 2000x2000 workload (pthreads PipeD4 and pthreads D on `threads`
    threads, first 200 attempted steps; seq D on one thread, first 20) and
    config 1 (N=128, t in [0, 4]) in full, seq D / pthreads D / PipeD4."""
    c2 = {"pthreads_PipeD4": _run_ref("pthreads_bruss_PipeD4_n2000_dopri54_sl200", threads),
          "pthreads_D": _run_ref("pthreads_bruss_D_n2000_dopri54_sl200", threads),
          "seq_D": _run_ref("seq_bruss_D_n2000_dopri54_sl20", 1)}
    c1 = {"seq_D": _run_ref("seq_bruss_D_n128_dopri54_tol1e-6", 1),
          "pthreads_D": _run_ref("pthreads_bruss_D_n128_dopri54_tol1e-6", threads),
          "pthreads_PipeD4": _run_ref("pthreads_bruss_PipeD4_n128_dopri54_tol1e-6", threads)}
    return {"c2_sample": c2, "c1_full": c1}


def cpu_baseline(leg, threads):
    """The headline CPU baseline: pthreads PipeD4 (the reference's fastest
    CPU form) on the 2000x2000 workload, or the oracle's pthreads D
    restatement when the reference binaries are absent."""
    r = leg["c2_sample"]["pthreads_PipeD4"]
    if r is not None:
        return {"value": 1.0 / r["s_per_step"], "unit": "steps/s", "cores": threads,
                "kind": "reference",
                "sample": "pthreads PipeD4 (reference C sources), BRUSS2D N=2000 DOPRI54 tol "
                          "1e-6 h0 1e-3, first %d attempted steps (STEP_LIMIT=200), %d threads"
                          % (r["steps"], threads)}
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


def gpu_c1():
    """Config 1 on the GPU for the CPU comparison: N=128, t in [0, 4], DOPRI54
    tol 1e-6, h0 1e-3 (336 attempted steps), fused kernel, inputs resident."""
    import torch
    import rk_amd
    p = rk_amd.Bruss2D(128, 2e-3)
    y0 = torch.from_numpy(p.ode_start()).cuda()
    y = torch.empty_like(y0)
    torch.cuda.synchronize()
    rk_amd.solve(p, "DOPRI54", 0.0, 4.0, y0, y, 1e-6, h0=1e-3)
    torch.cuda.synchronize()
    t = time.perf_counter()
    st, _ = rk_amd.solve(p, "DOPRI54", 0.0, 4.0, y0, y, 1e-6, h0=1e-3)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    return {"steps": st.steps, "kernel_s": st.kernel_time, "wall_s": wall}


def d_form_passes(s):
    """Vector passes (reads + writes of n doubles) of one step of the stage
    kernels of the D form: first stage reads y and writes w_1..w_{s-1}, dy,
    err; intermediate stage i reads w_i and updates w_{i+1..s-1}, dy, err;
    the last reads w_{s-1}, dy, err, y and writes y_new."""
    return (s + 2) + sum(5 + 2 * (s - 1 - i) for i in range(1, s - 1)) + 5


def gather_form_passes(meth, impl):
    """Vector passes of one step of the gather forms.  A and E: stage 0 reads
    y and writes v_0; stage l gathers (reads y and the v_i it adds, writes w)
    then evaluates (reads w, writes v_l); the output reads the v_i it adds and
    y, and writes y_new.  A adds only the non-zero coefficients of the table.
    F: stage l < s-1 reads w_l (stage 0: y), y and v_0..v_{l-1}, and writes
    v_l and w_{l+1}; the last stage reads w_{s-1} and writes v_{s-1}."""
    from rk_amd.methods import method
    m = method(meth)
    s = m.s
    A = [[m.A[l][i] for i in range(s)] for l in range(s)]
    if impl == "F":
        return 3 + sum(l + 4 for l in range(1, s - 1)) + 2 + (s + 2)
    skip = impl == "A"
    nnz = [sum(1 for i in range(l) if A[l][i] != 0.0 or not skip) for l in range(s)]
    out = 1 + sum(1 for i in range(1, s)
                  if not skip or m.b[i] != 0.0 or m.b[i] - m.b_hat[i] != 0.0)
    return 2 + sum(2 + nnz[l] + 2 for l in range(1, s)) + out + 2


def form_passes(meth, impl):
    from rk_amd.methods import method
    if impl == "D":
        return d_form_passes(method(meth).s)
    if impl in ("A", "E", "F"):
        return gather_form_passes(meth, impl)
    return 2


def c2_matrix_leg(window=(300, 1300), impls=("D", "A", "E", "F", "PipeDls"),
                  methods=("DOPRI54", "VERNER65", "DOPRI87"), tols=(1e-6, 1e-8)):
    """BRUSS2D N=2000, t in [0, 1], h0 1e-3: {DOPRI54, VERNER65, and the 13-stage
    DOPRI87} x tol {1e-6, 1e-8} x {full-storage loop orders D, A, E, F,
    low-storage fused PipeDls}; each timed over attempted steps
    window[0]..window[1] of the solve (differenced, so set-up is excluded),
    with the algorithmic HBM bytes of the form and the device bytes the solve
    holds."""
    import torch
    import rk_amd
    p = rk_amd.Bruss2D(N_GRID, 2e-3)
    y0 = torch.from_numpy(p.ode_start()).cuda()
    y = torch.empty_like(y0)
    torch.cuda.synchronize()
    rows = []
    for meth in methods:
        for tol in tols:
            for impl in impls:
                kw = dict(h0=1e-3, impl=impl)
                a, _ = rk_amd.solve(p, meth, 0.0, 1.0, y0, y, tol, step_limit=window[0], **kw)
                b, _ = rk_amd.solve(p, meth, 0.0, 1.0, y0, y, tol, step_limit=window[1], **kw)
                us = (b.kernel_time - a.kernel_time) / (b.steps - a.steps) * 1e6
                passes = form_passes(meth, impl)
                rows.append({"method": meth, "tol": tol, "impl": impl, "us_per_step": us,
                             "hbm_gbs_algorithmic": passes * 8.0 * p.n / (us * 1e-6) / 1e9,
                             "vector_passes_per_step": passes,
                             "resident_bytes": b.resident_bytes})
    return rows


def loop_order_leg():
    """The step-form autotuner (rk_amd.autotune.solve_form) on three problems:
    BRUSS2D N=2000 (fused sweep), an LV chain of 5e5 cells (fused chain
    kernel) and 4e6 independent LV pairs (no fused kernel: D, A, E, F only;
    "PipeDls" runs D there); DOPRI54, per-step times over attempted steps
    20..120 and the form it picks."""
    import torch
    import rk_amd
    from rk_amd.autotune import solve_form
    cases = [("BRUSS2D N=2000", rk_amd.Bruss2D(N_GRID, 2e-3), 1e-6, 1.0, 1e-3),
             ("LV chain M=5e5", rk_amd.LVChain(500_000, diff=1e-3), 1e-10, 40.0, 1e-4),
             ("LV pairs 4e6", rk_amd.LV(4_000_000), 1e-10, 40.0, 1e-4)]
    rows = []
    for name, p, tol, te, h0 in cases:
        y0 = torch.from_numpy(p.ode_start()).cuda()
        best, times = solve_form(p, "DOPRI54", y0, tol, te=te, h0=h0)
        rows.append({"problem": name, "best": best,
                     "us_per_step": {k: v * 1e6 for k, v in times.items()}})
        del y0
    return rows


def batched_leg(sample_every=1):
    """Mixed batch of 512 BRUSS2D instances (N uniform on 32..512 from seed 0,
    alpha log-uniform on [1e-3, 5e-3], noisy u0; DOPRI54 tol 1e-6, t in
    [0, 4], h0 1e-3): all of them through the batch engine vs the plain
    solver launched per instance, every instance (sample_every > 1 times
    every sample_every-th one and scales).  Both clocks start with the start
    vectors in device memory; the batch's tile shape comes from the autotuner
    (its search time is reported separately, and in a second ratio)."""
    import torch
    import rk_amd
    from rk_amd.workloads import bruss_instance, mixed_bruss
    from rk_amd.autotune import batch_tiles
    inst = [bruss_instance(N, a, k) for k, (N, a) in enumerate(mixed_bruss(512, seed=0))]
    torch.cuda.synchronize()
    t = time.perf_counter()
    (tw, tr), cand = batch_tiles(inst)
    t_tune = time.perf_counter() - t
    with rk_amd.Batch(tile_waves=tw, tile_rows=tr) as b:
        ids = [b.add(p, "DOPRI54", 0.0, 4.0, y0, 1e-6, h0=1e-3) for p, y0 in inst]
        torch.cuda.synchronize()          # start vectors resident in HBM from here
        t = time.perf_counter()
        left = b.run(10 ** 9)
        t_b = time.perf_counter() - t
        sts = [b.status(i)[0] for i in ids]
    assert left == 0
    steps = sum(s.steps for s in sts)
    comp_steps = sum(p.n * s.steps for (p, _), s in zip(inst, sts))
    t_p = 0.0
    for k in range(0, len(inst), sample_every):
        p, y0 = inst[k]
        y0d = torch.from_numpy(y0).cuda()
        yd = torch.empty_like(y0d)
        torch.cuda.synchronize()
        t = time.perf_counter()
        rk_amd.solve(p, "DOPRI54", 0.0, 4.0, y0d, yd, 1e-6, h0=1e-3)
        torch.cuda.synchronize()
        t_p += time.perf_counter() - t
    t_p *= sample_every
    return {"workload": "512 BRUSS2D instances, N uniform 32..512 (seed 0), alpha log-uniform "
                        "[1e-3,5e-3], DOPRI54 tol 1e-6, t in [0,4]",
            "batched_s": t_b, "batched_solves_per_s": len(inst) / t_b,
            "plain_s": t_p, "plain_solves_per_s": len(inst) / t_p,
            "speedup_vs_plain": t_p / t_b,
            "speedup_vs_plain_incl_autotune": t_p / (t_b + t_tune), "attempted_steps": steps,
            "tiles": {"waves": tw, "rows": tr, "autotune_s": t_tune,
                      "candidates_s": {"%dx%d" % k: v for k, v in cand.items()}},
            "hbm_gbs_algorithmic": 16.0 * comp_steps / t_b / 1e9,
            "plain_sample": ("all %d instances timed" % len(inst)) if sample_every == 1 else
                            ("plain solver timed on every %d-th instance (%d of %d), scaled by %d"
                             % (sample_every, len(range(0, len(inst), sample_every)), len(inst),
                                sample_every))}


def ensemble_leg():
    """LV batch: (A) 262144 independent pairs, VERNER65 tol 1e-8, t in [0, 20];
    (B) 64 periodic diffusive chains, M uniform on 1e4..1e6 (seed 1), DOPRI54
    tol 1e-6, t in [0, 4], through the batch engine vs plain per chain."""
    import torch
    import rk_amd
    from rk_amd.ensemble import lv_ensemble, lv_ensemble_inputs
    from rk_amd.workloads import chain_instance, lv_chains
    params, y0 = lv_ensemble_inputs(262144, seed=1)
    P = torch.from_numpy(params).cuda()
    Y0 = torch.from_numpy(y0).cuda()
    lv_ensemble(P, Y0, 0.0, 20.0, "VERNER65", 1e-8)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, acc, rej = lv_ensemble(P, Y0, 0.0, 20.0, "VERNER65", 1e-8)
    torch.cuda.synchronize()
    t_a = time.perf_counter() - t
    att = (acc + rej).double()
    chains = [chain_instance(M, k) for k, M in enumerate(lv_chains(64, seed=1))]
    with rk_amd.Batch() as b:
        for p, c0 in chains:
            b.add(p, "DOPRI54", 0.0, 4.0, c0, 1e-6, h0=1e-3)
        torch.cuda.synchronize()          # start vectors resident in HBM from here
        t = time.perf_counter()
        b.run(10 ** 9)
        t_b = time.perf_counter() - t
    t_p = 0.0
    for p, c0 in chains:
        y0d = torch.from_numpy(c0).cuda()
        yd = torch.empty_like(y0d)
        torch.cuda.synchronize()
        t = time.perf_counter()
        rk_amd.solve(p, "DOPRI54", 0.0, 4.0, y0d, yd, 1e-6, h0=1e-3)
        torch.cuda.synchronize()
        t_p += time.perf_counter() - t
    return {"pairs": {"systems": 262144, "seconds": t_a, "solves_per_s": 262144 / t_a,
                      "mean_attempted_steps": float(att.mean()),
                      "max_attempted_steps": int(att.max())},
            "chains": {"instances": len(chains), "batched_s": t_b, "plain_s": t_p,
                       "batched_solves_per_s": len(chains) / t_b,
                       "plain_solves_per_s": len(chains) / t_p}}


def serving_leg(rate_hz=4.0, duration_s=60.0, hbm_budget=4 << 30, sweep=(2.0, 4.0)):
    """Poisson stream (seed 2) at rate_hz over duration_s: 70 % BRUSS2D N
    uniform 64..1024, 30 % LV chains M uniform 1e4..5e5, tol from {1e-4,
    1e-6, 1e-8}, t_end uniform [0.5, 4], DOPRI54, h0 1e-3.  Served in real
    time by the continuous-batching server under an HBM budget, and by the
    plain solver one request at a time (each solve timed, queueing from the
    arrival times).  The sweep replays the same requests arriving `factor`
    times as fast (8 and 16 req/s at the default rate): the batched server
    runs again in real time; the plain server's latencies follow from its
    measured solve times, which do not depend on arrivals when requests run
    one at a time."""
    from rk_amd.problems import Bruss2D, LVChain
    from rk_amd.serve import Request, replay_plain, serve_batched, serve_plain
    from rk_amd.workloads import serving_stream
    stream = serving_stream(rate_hz, duration_s, seed=2)

    def requests(scale=1.0):
        out = []
        for k, (t, kind, size, tol, te) in enumerate(stream):
            p = Bruss2D(size, 2e-3) if kind == "bruss2d" else LVChain(size, diff=1e-3)
            y0 = p.ode_start_seeded(2, k, 0.01 if kind == "bruss2d" else 0.0)
            out.append(Request(k, t / scale, p, "DOPRI54", 0.0, te, y0, tol, h0=1e-3, kind=kind))
        return out
    _, sb = serve_batched(requests(), hbm_budget=hbm_budget)
    pdone, sp = serve_plain(requests())
    rates = [{"rate_hz": rate_hz, "batched": sb, "plain": sp}]
    for f in sweep:
        _, sbf = serve_batched(requests(f), hbm_budget=hbm_budget)
        rates.append({"rate_hz": rate_hz * f, "batched": sbf, "plain": replay_plain(pdone, f)})
        print("serving at %.0f req/s: batched %.2f solves/s p99 %.2f s; plain %.2f solves/s "
              "p99 %.2f s" % (rate_hz * f, sbf["solves_per_s"], sbf["p99_latency_s"],
                              rates[-1]["plain"]["solves_per_s"],
                              rates[-1]["plain"]["p99_latency_s"]), file=sys.stderr, flush=True)
    return {"workload": "Poisson %.1f req/s over %.0f s (seed 2): 70%% BRUSS2D N 64..1024, "
                        "30%% LV chain M 1e4..5e5, tol {1e-4,1e-6,1e-8}, t_end U[0.5,4]"
                        % (rate_hz, duration_s),
            "hbm_budget_bytes": hbm_budget, "batched": sb, "plain": sp,
            "p99_latency_ratio_plain_over_batched": sp["p99_latency_s"] / sb["p99_latency_s"],
            "rate_sweep": rates}


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
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched, ensemble and serving legs")
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--step-limit", type=int, default=0,
                    help="cap each solve at this many attempted steps (profiling runs)")
    ap.add_argument("--leg", choices=("c2_matrix", "loop_order", "batched", "ensemble", "serving"),
                    help="run only this leg and print its JSON (for profiling runs)")
    ap.add_argument("--grid", type=int, default=N_GRID,
                    help="BRUSS2D grid size of the main leg (profiling runs; the metric is "
                         "quoted at 2000)")
    ap.add_argument("--matrix-methods", default="DOPRI54,VERNER65,DOPRI87",
                    help="tables of the C2 matrix leg (profiling runs)")
    ap.add_argument("--matrix-tols", default="1e-6,1e-8",
                    help="tolerances of the C2 matrix leg (profiling runs)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run 5 plain device copies of n doubles, the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (a single "
              "grid would shard by row bands with one halo exchange and one max all-reduce "
              "per step; batched instances would shard whole); use --gpus 1", file=sys.stderr)
        return 2

    if a.leg:
        mm = tuple(a.matrix_methods.split(","))
        mt = tuple(float(x) for x in a.matrix_tols.split(","))
        legs = {"c2_matrix": lambda: c2_matrix_leg(methods=mm, tols=mt),
                "loop_order": loop_order_leg,
                "batched": batched_leg, "ensemble": ensemble_leg,
                "serving": lambda: serving_leg(a.serve_rate, a.serve_seconds)}
        print(json.dumps({"leg": a.leg, "result": legs[a.leg]()}))
        return 0

    import numpy as np
    import torch
    import rk_amd

    assert torch.cuda.is_available()
    stream = torch.cuda.Stream()
    p = rk_amd.Bruss2D(a.grid, 2e-3)
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
        "config": {"workload": "BRUSS2D N=%d full solve t in [0,1], %s tol %g h0 1e-3, "
                               "%s fused step kernel, device step control" % (a.grid, a.method,
                                                                              a.tol, a.impl),
                   "grid": a.grid, "n": n, "method": a.method, "tol": a.tol,
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
        res["host"] = host_info()
        leg = cpu_leg(threads)
        res["cpu_baseline"] = cpu_baseline(leg, threads)
        res["cpu_reference"] = leg
        res["c1_gpu"] = gpu_c1()
        print("cpu legs done", file=sys.stderr, flush=True)
    if not a.no_batch and a.step_limit == 0:
        res["c2_matrix"] = c2_matrix_leg()
        print("c2 matrix done", file=sys.stderr, flush=True)
        res["loop_order"] = loop_order_leg()
        res["batched"] = batched_leg()
        print("batched: %.2f s vs plain %.2f s" % (res["batched"]["batched_s"],
                                                   res["batched"]["plain_s"]),
              file=sys.stderr, flush=True)
        res["ensemble"] = ensemble_leg()
        print("ensemble done", file=sys.stderr, flush=True)
        res["serving"] = serving_leg(a.serve_rate, a.serve_seconds)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
