"""The continuous-batching server: requests that arrive while others run are
admitted into the running batch (HBM budget permitting), and every request's
result equals the plain single-instance solve of the same request bitwise
(same device step control), whatever the mix it ran in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rk():
    import torch
    assert torch.cuda.is_available()
    import rk_amd
    rk_amd.lib()
    return rk_amd


def stream(rk):
    from rk_amd.serve import Request
    items = [(0.00, "bruss2d", 40, 1e-6, 0.6), (0.01, "lvchain", 5000, 1e-8, 2.0),
             (0.02, "bruss2d", 97, 1e-4, 0.5), (0.05, "bruss2d", 24, 1e-8, 1.0),
             (0.08, "lvchain", 12000, 1e-6, 1.5), (0.10, "bruss2d", 130, 1e-6, 0.4),
             (0.12, "bruss2d", 64, 1e-6, 0.8), (0.15, "lvchain", 3001, 1e-4, 3.0)]
    out = []
    for k, (t, kind, size, tol, te) in enumerate(items):
        p = rk.Bruss2D(size, 2e-3) if kind == "bruss2d" else rk.LVChain(size, diff=1e-3)
        y0 = p.ode_start_seeded(2, k, 0.01 if kind == "bruss2d" else 0.0)
        out.append(Request(k, t, p, "DOPRI54", 0.0, te, y0, tol, h0=1e-3, kind=kind))
    return out


def plain_results(rk, reqs):
    from rk_amd.serve import serve_plain
    done, _ = serve_plain(reqs, keep_results=True)
    return {r.rid: (r.result, r.steps) for r in done}


@pytest.mark.parametrize("realtime,budget", [(True, 64 << 30), (False, 64 << 30),
                                             (False, 800_000)])
def test_served_results_equal_plain(rk, realtime, budget):
    from rk_amd.serve import serve_batched
    want = plain_results(rk, stream(rk))
    done, summ = serve_batched(stream(rk), hbm_budget=budget, poll_steps=16,
                               realtime=realtime, keep_results=True)
    assert summ["requests"] == len(want)
    if not realtime:
        assert summ["peak_concurrent"] >= 2
    for r in done:
        y, steps = want[r.rid]
        assert r.steps == steps, r.rid
        assert np.array_equal(r.result, y), r.rid
        assert r.admitted >= (r.arrival if realtime else 0.0) and r.finished >= r.admitted


def test_served_results_equal_oracle(rk, orc):
    """Served requests against the CPU oracle (seq D restated), not only
    against the plain GPU solver: the requests of the stream above that the
    oracle finishes in seconds, each bitwise with its step counts."""
    from rk_amd.serve import serve_batched
    done, _ = serve_batched(stream(rk), hbm_budget=64 << 30, poll_steps=16, realtime=False,
                            keep_results=True)
    checked = 0
    for r in done:
        if r.problem.n > 40_000:
            continue
        if r.kind == "bruss2d":
            po = orc.make_ode("bruss2d", r.problem.N, r.problem.alpha)
        else:
            po = orc.make_ode("lvchain", r.problem.cells, diff=r.problem.diff)
        yo, so, _ = orc.solve(po, r.method, r.t0, r.te, r.y0, r.tol, h0=r.h0)
        assert r.steps == so["steps"], r.rid
        assert np.array_equal(r.result, yo), r.rid
        checked += 1
    assert checked >= 4


def test_budget_limits_resident_set(rk):
    """An 800 kB budget holds only a few of the requests at a time (about 1.7
    MB in all): the server queues the rest and admits them as earlier ones
    retire."""
    from rk_amd.batch import Batch
    reqs = stream(rk)
    need = [Batch.footprint(r.problem, r.method) for r in reqs]
    assert sum(need) > 800_000 and max(need) <= 800_000
    from rk_amd.serve import serve_batched
    done, _ = serve_batched(reqs, hbm_budget=800_000, realtime=False)
    assert len(done) == len(reqs)
    # some request had to wait for a retirement
    assert max(r.admitted for r in done) > min(r.finished for r in done)


def test_requests_join_a_running_batch(rk):
    """A long request starts alone; short ones arriving while it runs are
    admitted into the running batch (continuous batching, not a queue) and
    finish before it, with results equal to their plain solves."""
    from rk_amd.serve import Request, serve_batched, serve_plain
    def reqs():
        out = []
        p = rk.Bruss2D(400, 5e-3)          # about 8000 steps: runs for tenths of a second
        out.append(Request(0, 0.0, p, "DOPRI54", 0.0, 4.0, p.ode_start_seeded(5, 0, 0.01), 1e-6,
                           h0=1e-3))
        for k in range(1, 6):
            q = rk.Bruss2D(24 + 8 * k, 2e-3)
            out.append(Request(k, 0.02 * k, q, "DOPRI54", 0.0, 0.3, q.ode_start_seeded(5, k, 0.01),
                               1e-6, h0=1e-3))
        return out
    want, _ = serve_plain(reqs(), keep_results=True)
    want = {r.rid: r.result for r in want}
    done, summ = serve_batched(reqs(), poll_steps=16, realtime=True, keep_results=True)
    by = {r.rid: r for r in done}
    assert summ["peak_concurrent"] >= 2
    assert all(by[k].finished < by[0].finished for k in range(1, 6))
    assert all(by[k].admitted < by[0].finished for k in range(1, 6))
    for k, r in by.items():
        assert np.array_equal(r.result, want[k]), k
