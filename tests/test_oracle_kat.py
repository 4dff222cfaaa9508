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


def test_schedules_agree_bitwise(orc):
    """D, PipeD, PipeDls and pthreads D are the same arithmetic in another order of
    blocks; the reference's own cross-variant check (Makefile:154-173)."""
    p = orc.make_ode("bruss2d", 24)
    y0 = orc.ode_start(p)
    outs = []
    for impl, thr in (("D", 1), ("PipeD", 1), ("PipeDls", 1), ("pthreads_D", 3)):
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
