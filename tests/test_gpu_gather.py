"""GPU parity of the gather loop orders A, E and F (src/impl/seq/A.c, E.c,
F.c; RKX_IMPL_A/E/F): every attempted step and the final vector equal the
oracle's and the reference's known answers bitwise -- the same as D's, since
every stage vector, dy and err receives the same terms in the same order.
The reference's own A, E and F binaries reproduce D's DUMP trajectories and
known answers ("same_in" in dump_index.json, "ref_impls" in kat.json)."""
import json
import os

import numpy as np
import pytest

from test_gpu_parity import dev, gpu_problem, run_gpu

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
KAT = json.load(open(os.path.join(GOLD, "kat.json")))["cases"]
DUMPS = json.load(open(os.path.join(GOLD, "dump_index.json")))
FORMS = ["A", "E", "F"]


@pytest.fixture(scope="module")
def rk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rk_amd
    rk_amd.lib()
    return rk_amd


@pytest.mark.parametrize("name", sorted(DUMPS))
@pytest.mark.parametrize("impl", FORMS)
def test_golden_dump(rk, name, impl, tmp_path):
    """Host step control with the DUMP file: every attempted step's t, h and
    tentative y equal the reference's trajectory."""
    meta = DUMPS[name]
    d = meta["defines"]
    z = np.load(os.path.join(GOLD, name + ".npz"))
    prob = "bruss2d" if meta["problem"].startswith("bruss") else "lv"
    grid = int(d.get("BRUSS_GRID_SIZE", d.get("LV_COUNT", 0)))
    p = gpu_problem(rk, prob, grid)
    m = rk.method(d["METHOD"])
    tol = float(d["TOL"]) if m.has_emb else 0.0
    path = tmp_path / "DUMP"
    print("dump %s impl %s" % (name, impl), flush=True)
    y, st, _ = run_gpu(rk, p, m, 0.0, float(d["T_END"]), p.ode_start(), tol,
                       h0=float(d["H_INIT"]), impl=impl, dump_path=str(path))
    rows = np.array([[float(x) for x in ln.split("\t")] for ln in open(path)])
    assert rows.shape == (len(z["t"]), 2 + z["Y"].shape[1])
    assert np.array_equal(rows[:, 0], z["t"]) and np.array_equal(rows[:, 1], z["h"])
    assert np.array_equal(rows[:, 2:], z["Y"])
    assert (st.steps_acc, st.steps_rej) == (meta["acc"], meta["rej"])
    assert np.array_equal(y, z["Y"][-1])


KAT_CPU = [pytest.param(c, id=c["name"]) for c in KAT if not c.get("gpu_only")]


@pytest.mark.parametrize("case", KAT_CPU)
@pytest.mark.parametrize("impl", FORMS)
def test_kat_device_control(rk, orc, case, impl):
    p = gpu_problem(rk, case["problem"], case["grid"], case.get("alpha", 2e-3))
    y0 = p.ode_start()
    print("kat %s impl %s" % (case["name"], impl), flush=True)
    y, st, tr = run_gpu(rk, p, case["method"], case["t0"], case["te"], y0, case["tol"],
                        h0=case["h0"], step_limit=case["step_limit"], impl=impl, trace_cap=100000)
    assert (st.steps_acc, st.steps_rej) == (case["acc"], case["rej"])
    want = [float(x) for x in case["footprint"]]
    assert [float(y[0]), float(y[-1]), float(y.min()), float(y.max())] == want
    if case["grid"] <= 128:
        po = orc.make_ode(case["problem"], case["grid"], case.get("alpha", 2e-3))
        yo, so, tro = orc.solve(po, case["method"], case["t0"], case["te"], y0, case["tol"],
                                h0=case["h0"], step_limit=case["step_limit"], impl=impl,
                                trace_cap=100000)
        assert np.array_equal(y, yo)
        assert np.array_equal(tr, tro)


CHAIN = [(3, "DOPRI54"), (1011, "VERNER65"), (4100, "RK4"), (3000, "DOPRI87")]


@pytest.mark.parametrize("M,meth", CHAIN)
@pytest.mark.parametrize("impl", FORMS)
def test_chain_host_control(rk, orc, M, meth, impl):
    diff = 0.37 if M < 10 else 1e-3
    p = rk.LVChain(M, diff=diff)
    po = orc.make_ode("lvchain", M, diff=diff)
    y0 = p.ode_start_seeded(3, M, 0.0)
    m = rk.method(meth)
    tol = 1e-7 if m.has_emb else 0.0
    te = 0.5 if m.has_emb else 0.05
    print("chain M=%d %s impl %s" % (M, meth, impl), flush=True)
    yo, so, tro = orc.solve(po, meth, 0.0, te, y0, tol, h0=1e-3, impl=impl, trace_cap=20000)
    y, st, tr = run_gpu(rk, p, meth, 0.0, te, y0, tol, h0=1e-3, impl=impl, ctrl="host",
                        trace_cap=20000)
    assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"])
    assert np.array_equal(tr, tro)
    assert np.array_equal(y, yo)


@pytest.mark.parametrize("impl", FORMS)
def test_c2_full_interval(rk, impl):
    """Config 2 (BRUSS2D N=2000, DOPRI54 tol 1e-6, t in [0,1]): 19 743 attempted
    steps through the gather form, against the reference's footprint."""
    case = [c for c in KAT if c["name"] == "c2_bruss2000_dopri54_tol1e-6"][0]
    p = rk.Bruss2D(case["grid"], case["alpha"])
    print("c2 impl %s" % impl, flush=True)
    y, st, _ = run_gpu(rk, p, case["method"], case["t0"], case["te"], p.ode_start(), case["tol"],
                       h0=case["h0"], impl=impl)
    assert (st.steps_acc, st.steps_rej) == (case["acc"], case["rej"])
    assert [float(y[0]), float(y[-1]), float(y.min()), float(y.max())] == \
        [float(x) for x in case["footprint"]]


def test_forms_equal_on_large_grid(rk):
    """N=4000 (3.2e7 components, beyond every cache), three steps: A, E and F
    leave the same vector as D."""
    import torch
    p = rk.Bruss2D(4000, 2e-3)
    y0 = dev(p.ode_start_seeded(5, 0, 0.01))
    out = {}
    for impl in ["D"] + FORMS:
        print("N=4000 impl %s" % impl, flush=True)
        y = torch.empty_like(y0)
        st, _ = rk.solve(p, "DOPRI54", 0.0, 1.0, y0, y, 1e-6, h0=1e-7, step_limit=3, impl=impl)
        torch.cuda.synchronize()
        out[impl] = (y.cpu().numpy(), (st.steps_acc, st.steps_rej), st.resident_bytes)
        del y
    n = p.n
    for impl in FORMS:
        assert out[impl][1] == out["D"][1]
        assert np.array_equal(out[impl][0], out["D"][0]), impl
    s = rk.method("DOPRI54").s
    # vectors of n held: y x2, v_0..v_{s-1}, w (F: two w); the rest is a few KB
    assert out["A"][2] // (8 * n) == 2 + s + 1 == out["E"][2] // (8 * n)
    assert out["F"][2] // (8 * n) == 2 + s + 2


def test_solve_form_autotuner(rk):
    """The loop-order autotuner times every form on an LV-pairs problem (no
    fused kernel) and returns one of them."""
    from rk_amd.autotune import FORMS as ALL, solve_form
    p = rk.LV(200000)
    y0 = dev(p.ode_start())
    best, times = solve_form(p, "DOPRI54", y0, 1e-10, te=40.0, h0=1e-4, window=(5, 25))
    assert set(times) == set(ALL) and best in ALL
    assert all(t > 0 for t in times.values())
