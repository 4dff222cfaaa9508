"""Parity at the sizes the benchmark runs, through the code paths only those
sizes reach:

* the mixed batch of 512 BRUSS2D instances (N uniform on 32..512, seed 0)
  with the autotuned tile and the band heights that shrink as the batch
  drains, every instance bitwise equal to its plain solve, and the stiffest
  instances bitwise equal to the oracle;
* LV chains of 10^4 and 6*10^5 cells: the second has more segments (595)
  than an instance's workgroup cap (256), so workgroups loop over segments;
  fused kernel, D form and oracle bitwise equal;
* BRUSS2D N=2000 VERNER65 at tol 1e-8 over the whole interval through the D
  form, against the reference's known answer.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
KAT = {c["name"]: c for c in json.load(open(os.path.join(HERE, "golden", "kat.json")))["cases"]}


@pytest.fixture(scope="module")
def rk():
    import torch
    assert torch.cuda.is_available()
    import rk_amd
    rk_amd.lib()
    return rk_amd


def plain(rk, p, y0, meth, te, tol, h0, **kw):
    import torch
    y0d = torch.from_numpy(np.ascontiguousarray(y0)).cuda()
    yd = torch.empty_like(y0d)
    torch.cuda.synchronize()
    st, _ = rk.solve(p, meth, 0.0, te, y0d, yd, tol, h0=h0, **kw)
    torch.cuda.synchronize()
    return yd.cpu().numpy(), st


def test_c3_mixed_batch_512(rk, orc):
    from rk_amd.autotune import batch_tiles
    from rk_amd.workloads import bruss_instance, mixed_bruss
    te = 0.1            # the stiffest instances take about 300 steps, most a few
    mix = mixed_bruss(512, seed=0)
    inst = [bruss_instance(N, a, k) for k, (N, a) in enumerate(mix)]
    (tw, tr), cand = batch_tiles(inst, te=te, steps=16)
    print("autotuned tile: %d waves x %d rows" % (tw, tr), flush=True)
    # the instance with the most steps (explicit diffusion: h ~ 1/(alpha (N-1)^2))
    stiff = sorted(range(len(mix)), key=lambda k: -mix[k][1] * (mix[k][0] - 1) ** 2)
    last = stiff[0]
    with rk.Batch(tile_waves=tw, tile_rows=tr) as b:
        ids = [b.add(p, "DOPRI54", 0.0, te, y0, 1e-6, h0=1e-3) for p, y0 in inst]
        rows_seen = [b.tile(ids[last])[1]]
        while b.run(64) > 0:
            rows_seen.append(b.tile(ids[last])[1])
        got = [(b.result(i), b.status(i)[0]) for i in ids]
    print("band rows of the last instance as the batch drained:", sorted(set(rows_seen)),
          flush=True)
    assert min(rows_seen) < rows_seen[0], "bands did not shrink as the batch drained"
    for k, (p, y0) in enumerate(inst):
        ya, sa = plain(rk, p, y0, "DOPRI54", te, 1e-6, 1e-3)
        y, st = got[k]
        assert (st.steps_acc, st.steps_rej) == (sa.steps_acc, sa.steps_rej), k
        assert np.array_equal(y, ya), k
    print("512 instances equal their plain solves", flush=True)
    for k in stiff[:2]:
        N, a = mix[k]
        po = orc.make_ode("bruss2d", N, a)
        yo, so, _ = orc.solve(po, "DOPRI54", 0.0, te, inst[k][1], 1e-6, h0=1e-3)
        y, st = got[k]
        print("N=%d alpha=%.3g: %d steps, oracle %d" % (N, a, st.steps, so["steps"]), flush=True)
        assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"])
        assert np.array_equal(y, yo)


@pytest.mark.parametrize("M", [10_000, 600_000])
def test_c4b_chain_sizes(rk, orc, M):
    from rk_amd.workloads import chain_instance
    p, y0 = chain_instance(M, 3)
    po = orc.make_ode("lvchain", M, diff=1e-3)
    yo, so, _ = orc.solve(po, "DOPRI54", 0.0, 4.0, y0, 1e-6, h0=1e-3, step_limit=25)
    for impl in ("PipeDls", "D"):
        y, st = plain(rk, p, y0, "DOPRI54", 4.0, 1e-6, 1e-3, impl=impl, step_limit=25)
        assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"]), impl
        assert np.array_equal(y, yo), impl
        # low storage: y twice; D form: y twice, w_1..w_6, dy, err (plus the state)
        assert st.resident_bytes // (p.n * 8) == (2 if impl == "PipeDls" else 10), impl


def test_c2_verner65_tol1e8_d_form(rk):
    """BRUSS2D N=2000, VERNER65, tol 1e-8, t in [0, 1] through the stage
    kernels of the D form (15 749 steps, about 20 s): the reference's step
    counts and footprint digit for digit."""
    c = KAT["c2_bruss2000_verner65_tol1e-8"]
    p = rk.Bruss2D(c["grid"], c["alpha"])
    y, st = plain(rk, p, p.ode_start(), c["method"], c["te"], c["tol"], c["h0"], impl="D")
    assert (st.steps_acc, st.steps_rej) == (c["acc"], c["rej"])
    assert [y[0], y[-1], y.min(), y.max()] == [float(x) for x in c["footprint"]]
