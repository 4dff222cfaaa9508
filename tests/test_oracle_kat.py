"""The oracle against the reference's known answers (tests/golden/kat.json)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))["cases"]


def _case_params(c):
    marks = []
    if c.get("slow"):
        marks.append(pytest.mark.slow)
    return pytest.param(c, id=c["name"], marks=marks)


CPU_CASES = [_case_params(c) for c in KAT if not c.get("gpu_only")]


@pytest.mark.parametrize("case", CPU_CASES)
@pytest.mark.parametrize("impl", ["D", "PipeDls"])
def test_kat_footprint(orc, case, impl):
    p = orc.make_ode(case["problem"], case["grid"], case.get("alpha", 2e-3))
    y0 = orc.ode_start(p)
    y, st, _ = orc.solve(p, case["method"], case["t0"], case["te"], y0, case["tol"],
                         h0=case["h0"], step_limit=case["step_limit"], impl=impl)
    assert (st["acc"], st["rej"]) == (case["acc"], case["rej"])
    want = [float(x) for x in case["footprint"]]
    assert orc.footprint(y) == want          # bitwise


GATHER_CASES = [pytest.param(c, impl, id="%s-%s" % (c["name"], impl))
                for c in KAT for impl in c.get("ref_impls", []) if impl != "D"]


@pytest.mark.parametrize("case,impl", GATHER_CASES)
def test_kat_gather_forms(orc, case, impl):
    """The oracle's A, E and F on the known answers the reference's own A, E
    and F binaries printed (the same as D's)."""
    p = orc.make_ode(case["problem"], case["grid"], case.get("alpha", 2e-3))
    y0 = orc.ode_start(p)
    y, st, _ = orc.solve(p, case["method"], case["t0"], case["te"], y0, case["tol"],
                         h0=case["h0"], step_limit=case["step_limit"], impl=impl)
    assert (st["acc"], st["rej"]) == (case["acc"], case["rej"])
    assert orc.footprint(y) == [float(x) for x in case["footprint"]]


@pytest.mark.parametrize("meth,tol", [("DOPRI54", 1e-7), ("VERNER65", 1e-9), ("DOPRI87", 1e-10),
                                      ("RKF23", 1e-5), ("RK4", 0.0), ("HEUN2", 0.0),
                                      ("SSPRK3", 0.0)])
@pytest.mark.parametrize("prob,grid", [("bruss2d", 13), ("lv", 40), ("lvchain", 57)])
def test_gather_forms_equal_d(orc, meth, tol, prob, grid):
    """Every method table (zero entries in A, b and b - b_hat differ between
    them, so A's skips differ) on every problem: A, E and F give D's attempted
    steps and final vector bitwise."""
    p = orc.make_ode(prob, grid)
    y0 = orc.ode_start(p)
    te = 0.3 if tol > 0 else 0.05
    ref = orc.solve(p, meth, 0.0, te, y0, tol, h0=1e-3, trace_cap=5000)
    for impl in ("A", "E", "F"):
        y, st, tr = orc.solve(p, meth, 0.0, te, y0, tol, h0=1e-3, impl=impl, trace_cap=5000)
        assert np.array_equal(tr, ref[2]), impl
        assert np.array_equal(y, ref[0]), impl


def test_schedules_agree_bitwise(orc):
    """D, PipeD, PipeDls, A, E, F and pthreads D are the same arithmetic in another
    order of blocks; the reference's own cross-variant check (Makefile:154-173)."""
    p = orc.make_ode("bruss2d", 24)
    y0 = orc.ode_start(p)
    outs = []
    for impl, thr in (("D", 1), ("PipeD", 1), ("PipeDls", 1), ("A", 1), ("E", 1), ("F", 1),
                      ("pthreads_D", 3)):
        y, st, tr = orc.solve(p, "DOPRI54", 0.0, 2.0, y0, 1e-7, h0=1e-3, impl=impl,
                              threads=thr, trace_cap=2000 if impl != "pthreads_D" else 0)
        outs.append((impl, y, st))
    for impl, y, st in outs[1:]:
        assert np.array_equal(y, outs[0][1]), impl
        assert (st["acc"], st["rej"]) == (outs[0][2]["acc"], outs[0][2]["rej"])


def test_pipe_rejects_bad_block(orc):
    p = orc.make_ode("bruss2d", 8)
    y0 = orc.ode_start(p)
    with pytest.raises(RuntimeError):
        orc.solve(p, "DOPRI54", 0.0, 0.1, y0, 1e-6, h0=1e-3, impl="PipeD", block=8)  # B < d = 16
