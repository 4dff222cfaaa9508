"""Serving loop: continuous batching of solve requests on one GPU.

Requests arrive over time (each: a problem, a method, [t0, te], y0, tol).
The server admits a waiting request as soon as the batch engine's HBM budget
has room for it (low storage -- the fused kernel's two vectors -- when the
problem has a fused kernel, else full storage), keeps every admitted
instance's state resident on the device until its t_end, advances all
admitted instances together (one attempted step each per batched step, each
with its own h and accept/reject), retires finished instances and hands
their results back.  Requests that arrive while others are on the GPU join
the running batch between polls, so their work overlaps on the device.

Latency = completion time - arrival time (queueing included).  The
comparator `serve_plain` runs the same requests through the plain
single-instance solver, one at a time in arrival order, which is what a
caller without this server would do.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from .batch import Batch


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the stream
    problem: object           # rk_amd.Problem
    method: str
    t0: float
    te: float
    y0: np.ndarray
    tol: float
    h0: float = 0.0
    kind: str = ""
    # filled by the server
    admitted: float = float("nan")
    finished: float = float("nan")
    steps: int = 0
    result: np.ndarray = field(default=None, repr=False)

    @property
    def latency(self):
        return self.finished - self.arrival


def percentile(v, q):
    v = np.sort(np.asarray(v, dtype=float))
    return float(np.percentile(v, q)) if len(v) else float("nan")


def summarize(reqs, makespan):
    lat = [r.latency for r in reqs]
    return {"requests": len(reqs), "makespan_s": makespan,
            "solves_per_s": len(reqs) / makespan if makespan > 0 else float("nan"),
            "p50_latency_s": percentile(lat, 50), "p99_latency_s": percentile(lat, 99),
            "mean_latency_s": float(np.mean(lat)) if lat else float("nan")}


def serve_batched(requests, hbm_budget=64 << 30, poll_steps=64, realtime=True, ctrl="device",
                  keep_results=False, log=None):
    """Run the stream through the continuous-batching server.  With realtime,
    a request is only admitted once the wall clock has passed its arrival
    time; otherwise all are available at once (offline batch).  Returns
    (requests, summary)."""
    pending = sorted(requests, key=lambda r: r.arrival)
    active = {}                # instance id -> request
    done = []
    nxt = 0
    peak = 0
    with Batch(hbm_budget=hbm_budget, ctrl=ctrl) as b:
        t_start = time.perf_counter()
        now = lambda: time.perf_counter() - t_start
        while nxt < len(pending) or active:
            # admit every arrived request that fits in the budget (FIFO)
            while nxt < len(pending) and (not realtime or pending[nxt].arrival <= now()):
                r = pending[nxt]
                i = b.add(r.problem, r.method, r.t0, r.te, r.y0, r.tol, h0=r.h0)
                if i is None:
                    if not active:
                        raise RuntimeError("request %d needs %d bytes, more than the HBM "
                                           "budget of %d" % (r.rid, Batch.footprint(
                                               r.problem, r.method), hbm_budget))
                    break                      # budget full: wait for retirements
                r.admitted = now()
                active[i] = r
                nxt += 1
                peak = max(peak, len(active))
            if not active:
                # idle: sleep until the next arrival
                if nxt < len(pending):
                    time.sleep(max(0.0, min(0.01, pending[nxt].arrival - now())))
                continue
            b.run(poll_steps)
            for i in list(active):
                st, fin = b.status(i)
                if fin:
                    r = active.pop(i)
                    r.finished = now()
                    r.steps = st.steps
                    if keep_results:
                        r.result = b.result(i)
                    b.remove(i)
                    done.append(r)
                    if log:
                        log(r)
        makespan = now()
    summ = summarize(done, makespan)
    summ["peak_concurrent"] = peak
    return done, summ


def replay_plain(reqs, scale=1.0):
    """The plain server's latencies for the same requests arriving `scale`
    times as fast, from the solve durations serve_plain measured: the plain
    server runs one request at a time, so a request's solve time does not
    depend on when it arrived, only its start does (max of its arrival and the
    previous completion)."""
    clock = 0.0
    out = []
    for r in sorted(reqs, key=lambda q: q.arrival):
        a = r.arrival / scale
        start = max(clock, a)
        clock = start + (r.finished - r.admitted)
        out.append(clock - a)
    s = {"requests": len(out), "makespan_s": clock,
         "solves_per_s": len(out) / clock if clock > 0 else float("nan"),
         "p50_latency_s": percentile(out, 50), "p99_latency_s": percentile(out, 99),
         "mean_latency_s": float(np.mean(out)) if out else float("nan")}
    return s


def serve_plain(requests, realtime_model=True, keep_results=False):
    """The comparator: each request through the plain solver (rk_amd.solve),
    one at a time in arrival order.  A request starts at max(its arrival,
    the previous request's completion); the solve itself is timed for real,
    the arrival waits are accounted without sleeping."""
    import torch
    from .solver import solve
    clock = 0.0
    done = []
    for r in sorted(requests, key=lambda q: q.arrival):
        y0 = torch.from_numpy(np.ascontiguousarray(r.y0)).cuda()
        y = torch.empty_like(y0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        st, _ = solve(r.problem, r.method, r.t0, r.te, y0, y, r.tol, h0=r.h0)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        start = max(clock, r.arrival) if realtime_model else clock
        r.admitted = start
        r.finished = start + dt
        r.steps = st.steps
        if keep_results:
            r.result = y.cpu().numpy()
        clock = r.finished
        done.append(r)
    return done, summarize(done, clock)
