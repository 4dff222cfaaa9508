"""GPU parity of the HIP path against the CPU oracle (seq D restated) and the
reference's known answers.

Bar: bitwise.  With host step control (the C library's pow, as in
include/tools.h:226) and with device step control (the C library's pow
restated on the device) every attempted step's (t, h, err) and the final
vector equal the oracle's; known answers of the compiled reference match
digit for digit.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))["cases"]
GOLD = os.path.join(HERE, "golden")
DUMPS = json.load(open(os.path.join(GOLD, "dump_index.json")))

REL_TOL = 1e-12


@pytest.fixture(scope="module")
def rk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import rk_amd
    rk_amd.lib()
    return rk_amd


def dev(a):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    torch.cuda.synchronize()
    return t


def run_gpu(rk, prob, meth, t0, te, y0, tol, **kw):
    import torch
    y0d = dev(y0)
    yd = torch.empty_like(y0d)
    st, tr = rk.solve(prob, meth, t0, te, y0d, yd, tol, **kw)
    torch.cuda.synchronize()
    return yd.cpu().numpy(), st, tr


def max_rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def gpu_problem(rk, name, grid, alpha=2e-3):
    if name == "bruss2d":
        return rk.Bruss2D(grid, alpha)
    if name == "lv":
        return rk.LV(grid)
    return rk.LVChain(grid)


# ---- right-hand sides ---------------------------------------------------------

@pytest.mark.parametrize("N", [2, 3, 5, 8, 33, 130])
def test_bruss_rhs_bitwise(rk, orc, N):
    import torch
    rng = np.random.default_rng(N)
    p = rk.Bruss2D(N, 2.7e-3)
    po = orc.make_ode("bruss2d", N, 2.7e-3)
    y = rng.uniform(0.1, 5.0, p.n)
    yd, fd = dev(y), torch.zeros(p.n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    p.eval_rng_device(yd.data_ptr(), fd.data_ptr())
    assert np.array_equal(fd.cpu().numpy(), orc.eval_rng(po, y))
    # sub-range with absolute indexing (ode.h:35)
    fd.zero_()
    torch.cuda.synchronize()          # the library runs on its own stream
    i, j = 3, p.n - 4
    p.eval_rng_device(yd.data_ptr(), fd.data_ptr(), i, j)
    got = fd.cpu().numpy()
    ref = orc.eval_rng(po, y, i, j)
    assert np.array_equal(got[i:j + 1], ref[i:j + 1]) and not got[:i].any()


@pytest.mark.parametrize("kind,grid", [("bruss2d", 2), ("bruss2d", 17), ("lv", 9),
                                       ("lvchain", 11)])
def test_host_callbacks_bitwise(rk, orc, kind, grid):
    """ode_eval_comp / ode_eval_rng on host buffers (rkx_eval_comp_host,
    rkx_eval_rng_host: only the stencil window is staged) equal the oracle's
    bitwise for every component and for ranges at both ends and inside."""
    import ctypes as C
    from rk_amd._lib import check, lib
    rng = np.random.default_rng(grid)
    p = gpu_problem(rk, kind, grid) if kind != "lvchain" else rk.LVChain(grid, diff=0.02)
    po = orc.make_ode(kind, grid, diff=0.02)
    y = rng.uniform(0.1, 3.0, p.n)
    dp = C.POINTER(C.c_double)
    for i in range(p.n):
        v = C.c_double()
        check(lib().rkx_eval_comp_host(C.byref(p.desc), i, 0.0, y.ctypes.data_as(dp), C.byref(v)))
        assert v.value == orc.eval_comp(po, y, i), i
    for i, j in ((0, 0), (0, p.n - 1), (1, 4), (p.n - 3, p.n - 1), (p.n // 2 - 1, p.n // 2 + 2)):
        f = np.full(p.n, -7.0)
        check(lib().rkx_eval_rng_host(C.byref(p.desc), i, j, 0.0, y.ctypes.data_as(dp),
                                      f.ctypes.data_as(dp)))
        ref = orc.eval_rng(po, y, i, j)
        assert np.array_equal(f[i:j + 1], ref[i:j + 1]), (i, j)
        assert (f[:i] == -7.0).all() and (f[j + 1:] == -7.0).all()


@pytest.mark.parametrize("kind", ["lv", "lvchain"])
def test_lv_rhs_bitwise(rk, orc, kind):
    import torch
    rng = np.random.default_rng(5)
    p = rk.LV(37) if kind == "lv" else rk.LVChain(37, diff=0.013)
    po = orc.make_ode(kind, 37, diff=0.013)
    y = rng.uniform(0.1, 2.0, p.n)
    yd, fd = dev(y), torch.zeros(p.n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    p.eval_rng_device(yd.data_ptr(), fd.data_ptr())
    assert np.array_equal(fd.cpu().numpy(), orc.eval_rng(po, y))


# ---- full solves, bitwise with host step control ------------------------------

CFGS = [("D", 0, 0), ("PipeDls", 1, 16), ("PipeDls", 2, 7), ("PipeDls", 3, 20),
        ("PipeDls", 4, 32), ("PipeD", 4, 100)]


@pytest.mark.parametrize("impl,waves,rows", CFGS)
def test_c1_host_control_bitwise(rk, orc, impl, waves, rows):
    """Config 1 (BRUSS2D N=128, DOPRI54, tol 1e-6, t in [0,4], h0 1e-3)."""
    N = 128
    po = orc.make_ode("bruss2d", N)
    y0 = orc.ode_start(po)
    yo, so, tro = orc.solve(po, "DOPRI54", 0.0, 4.0, y0, 1e-6, h0=1e-3, trace_cap=1000)
    p = rk.Bruss2D(N)
    y, st, tr = run_gpu(rk, p, "DOPRI54", 0.0, 4.0, y0, 1e-6, h0=1e-3, impl=impl, ctrl="host",
                        tile_waves=waves, tile_rows=rows, trace_cap=1000)
    assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"]) == (320, 16)
    assert np.array_equal(tr, tro)
    assert np.array_equal(y, yo)


# Grids whose columns past the full-width strips fit a strip of half the lanes:
# the fused kernel pairs two bands of that narrow strip in one workgroup.
# (N, waves, rows, table, alpha): odd band counts (one half without a band),
# ragged last bands, no full-width strip at all (N = 60), whole-trip tables.
PAIRED = [(130, 2, 9, "DOPRI54", 2e-3), (260, 4, 37, "DOPRI54", 2e-3),
          (280, 4, 50, "VERNER65", 2e-3), (60, 4, 11, "RKF23", 2e-3),
          (130, 2, 20, "DOPRI87", 3e-3)]


@pytest.mark.parametrize("N,waves,rows,meth,alpha", PAIRED)
def test_paired_narrow_strip_bitwise(rk, orc, N, waves, rows, meth, alpha):
    po = orc.make_ode("bruss2d", N, alpha)
    y0 = orc.ode_start(po)
    tol = 1e-9 if meth == "DOPRI87" else 1e-6
    yo, so, tro = orc.solve(po, meth, 0.0, 0.3, y0, tol, h0=1e-4, trace_cap=2000)
    p = rk.Bruss2D(N, alpha)
    y, st, tr = run_gpu(rk, p, meth, 0.0, 0.3, y0, tol, h0=1e-4, impl="PipeDls", ctrl="host",
                        tile_waves=waves, tile_rows=rows, trace_cap=2000)
    assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"])
    assert st.steps_acc > 5
    assert np.array_equal(tr, tro)
    assert np.array_equal(y, yo)


@pytest.mark.parametrize("name", sorted(DUMPS))
@pytest.mark.parametrize("impl", ["D", "PipeDls"])
def test_golden_dump_host_control(rk, name, impl):
    """Every attempted step's t and h equal the reference's DUMP output bitwise."""
    meta = DUMPS[name]
    d = meta["defines"]
    z = np.load(os.path.join(GOLD, name + ".npz"))
    prob = "bruss2d" if meta["problem"].startswith("bruss") else "lv"
    grid = int(d.get("BRUSS_GRID_SIZE", d.get("LV_COUNT", 0)))
    p = gpu_problem(rk, prob, grid)
    m = rk.method(d["METHOD"])
    tol = float(d["TOL"]) if m.has_emb else 0.0
    y, st, tr = run_gpu(rk, p, m, 0.0, float(d["T_END"]), p.ode_start(), tol,
                        h0=float(d["H_INIT"]), impl=impl, ctrl="host",
                        trace_cap=len(z["t"]) + 5)
    assert (st.steps_acc, st.steps_rej) == (meta["acc"], meta["rej"])
    assert np.array_equal(tr[:, 0], z["t"]) and np.array_equal(tr[:, 1], z["h"])
    assert [float(y[0]), float(y[-1]), float(y.min()), float(y.max())] == meta["footprint"]
    assert np.array_equal(y, z["Y"][-1])


@pytest.mark.parametrize("name", sorted(DUMPS))
@pytest.mark.parametrize("impl", ["D", "PipeDls"])
def test_golden_dump_file(rk, name, impl, tmp_path):
    """rkx_opts.dump_path writes the reference's DUMP file (include/tools.h:239-256):
    every attempted step's t, h and new y (rejected steps included) equal the
    reference's, value for value."""
    meta = DUMPS[name]
    d = meta["defines"]
    z = np.load(os.path.join(GOLD, name + ".npz"))
    prob = "bruss2d" if meta["problem"].startswith("bruss") else "lv"
    grid = int(d.get("BRUSS_GRID_SIZE", d.get("LV_COUNT", 0)))
    p = gpu_problem(rk, prob, grid)
    m = rk.method(d["METHOD"])
    tol = float(d["TOL"]) if m.has_emb else 0.0
    path = tmp_path / "DUMP"
    y, st, _ = run_gpu(rk, p, m, 0.0, float(d["T_END"]), p.ode_start(), tol,
                       h0=float(d["H_INIT"]), impl=impl, dump_path=str(path))
    rows = np.array([[float(x) for x in ln.split("\t")] for ln in open(path)])
    assert rows.shape == (len(z["t"]), 2 + z["Y"].shape[1])
    assert np.array_equal(rows[:, 0], z["t"]) and np.array_equal(rows[:, 1], z["h"])
    assert np.array_equal(rows[:, 2:], z["Y"])
    assert (st.steps_acc, st.steps_rej) == (meta["acc"], meta["rej"])


@pytest.mark.parametrize("N,waves,rows", [(20, 0, 0), (40, 1, 9), (40, 2, 40), (130, 0, 0),
                                          (130, 2, 17)])
def test_dopri87_fused_equals_d_form(rk, N, waves, rows):
    """DOPRI87 (13 stages) through its fused kernel (one wave per SIMD, whole
    ring trips, band heights rounded up to fill them) gives the D form's
    steps and vector bitwise, for short grids (general ghost rows), both
    strip widths and several band heights; two vectors resident."""
    p = rk.Bruss2D(N, 3e-3)
    y0 = p.ode_start()
    yd, std, _ = run_gpu(rk, p, "DOPRI87", 0.0, 0.5, y0, 1e-9, h0=1e-4, impl="D")
    yf, stf, _ = run_gpu(rk, p, "DOPRI87", 0.0, 0.5, y0, 1e-9, h0=1e-4, impl="PipeDls",
                         tile_waves=waves, tile_rows=rows)
    assert (stf.steps_acc, stf.steps_rej) == (std.steps_acc, std.steps_rej)
    assert np.array_equal(yf, yd)
    assert stf.resident_bytes < std.resident_bytes
    assert stf.resident_bytes <= 2 * p.n * 8 + 65536


def test_dopri87_strip_too_wide_is_an_error(rk):
    p = rk.Bruss2D(300)
    with pytest.raises(rk.RkxError):
        run_gpu(rk, p, "DOPRI87", 0.0, 0.01, p.ode_start(), 1e-9, h0=1e-4, impl="PipeDls",
                tile_waves=4)


CHAIN = [(3, "DOPRI54"), (7, "RKF23"), (1010, "DOPRI54"), (1011, "DOPRI54"),
         (2500, "VERNER65"), (4100, "RK4"), (3000, "SSPRK3"), (999, "HEUN2")]


@pytest.mark.parametrize("M,meth", CHAIN)
@pytest.mark.parametrize("impl", ["D", "PipeDls"])
def test_chain_host_control_bitwise(rk, orc, M, meth, impl):
    """The LV chain through the stage kernels and the fused segment kernel:
    segments of 1024 cells minus an s-cell halo on each side, so the sizes
    cover a chain shorter than the halo, exactly one segment, one cell more,
    and several segments; every attempted step equals the oracle's bitwise."""
    diff = 0.37 if M < 10 else 1e-3
    p = rk.LVChain(M, diff=diff)
    po = orc.make_ode("lvchain", M, diff=diff)
    y0 = p.ode_start_seeded(3, M, 0.0)
    m = rk.method(meth)
    tol = 1e-7 if m.has_emb else 0.0
    te = 0.5 if m.has_emb else 0.05
    yo, so, tro = orc.solve(po, meth, 0.0, te, y0, tol, h0=1e-3, trace_cap=20000)
    y, st, tr = run_gpu(rk, p, meth, 0.0, te, y0, tol, h0=1e-3, impl=impl, ctrl="host",
                        trace_cap=20000)
    assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"])
    assert np.array_equal(tr, tro)
    assert np.array_equal(y, yo)


@pytest.mark.parametrize("meth", ["HEUN2", "SSPRK3", "RKF23", "RK4", "DOPRI54", "VERNER65",
                                  "DOPRI87"])
def test_lv_pairs_fused_bitwise(rk, orc, meth):
    """LV pairs through the per-pair fused kernel (low storage): every attempted
    step (t, h, err) and the final vector equal the oracle's seq D, with
    constants where d0 != c1 and start values off the reference's."""
    consts = (1.3, 0.7, 0.45, 0.9)
    p = rk.LV(4099, *consts)
    po = orc.make_ode("lv", 4099, lv=consts)
    y0 = 0.5 + np.random.default_rng(3).random(p.n)
    m = rk.method(meth)
    tol = 1e-9 if m.has_emb else 0.0
    te = 6.0 if m.has_emb else 0.3
    yo, so, tro = orc.solve(po, meth, 0.0, te, y0, tol, h0=1e-3, trace_cap=20000)
    for ctrl in ("host", "device"):
        print("lv pairs %s %s" % (meth, ctrl), flush=True)
        y, st, tr = run_gpu(rk, p, meth, 0.0, te, y0, tol, h0=1e-3, impl="PipeDls", ctrl=ctrl,
                            trace_cap=20000)
        assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"])
        assert np.array_equal(tr, tro)
        assert np.array_equal(y, yo)
        assert st.resident_bytes < 3 * p.n * 8       # two vectors: the low-storage form


def test_lv_pairs_fused_equals_d_form_large(rk):
    """4e6 pairs (n = 8e6, as at C2), a few steps: fused ≡ D bitwise."""
    import torch
    p = rk.LV(4_000_000)
    y0 = dev(0.5 + np.random.default_rng(5).random(p.n))
    out = {}
    for impl in ("PipeDls", "D"):
        print("lv pairs 4e6 %s" % impl, flush=True)
        y = torch.empty_like(y0)
        st, _ = rk.solve(p, "VERNER65", 0.0, 10.0, y0, y, 1e-10, h0=1e-4, step_limit=25,
                         impl=impl)
        torch.cuda.synchronize()
        out[impl] = (y.cpu().numpy(), (st.steps_acc, st.steps_rej))
        del y
    assert out["PipeDls"][1] == out["D"][1]
    assert np.array_equal(out["PipeDls"][0], out["D"][0])


def test_chain_device_control_bitwise(rk, orc):
    M = 5000
    p = rk.LVChain(M)
    po = orc.make_ode("lvchain", M)
    y0 = p.ode_start_seeded(1, 0, 0.0)
    yo, so, _ = orc.solve(po, "DOPRI54", 0.0, 2.0, y0, 1e-8, h0=1e-4)
    y, st, _ = run_gpu(rk, p, "DOPRI54", 0.0, 2.0, y0, 1e-8, h0=1e-4, impl="PipeDls")
    assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"])
    assert np.array_equal(y, yo)
    assert st.resident_bytes < 3 * p.n * 8         # two vectors: the low-storage form


# ---- device step control ----------------------------------------------------------

KAT_CPU = [pytest.param(c, id=c["name"]) for c in KAT if not c.get("gpu_only")]


@pytest.mark.parametrize("case", KAT_CPU)
@pytest.mark.parametrize("impl", ["D", "PipeDls"])
def test_kat_device_control(rk, orc, case, impl):
    p = gpu_problem(rk, case["problem"], case["grid"], case.get("alpha", 2e-3))
    y0 = p.ode_start()
    y, st, tr = run_gpu(rk, p, case["method"], case["t0"], case["te"], y0, case["tol"],
                        h0=case["h0"], step_limit=case["step_limit"], impl=impl, trace_cap=100000)
    assert (st.steps_acc, st.steps_rej) == (case["acc"], case["rej"])
    want = [float(x) for x in case["footprint"]]
    got = [float(y[0]), float(y[-1]), float(y.min()), float(y.max())]
    assert got == want                     # bitwise: the pow service is the C library's
    if case["grid"] <= 128:
        po = orc.make_ode(case["problem"], case["grid"], case.get("alpha", 2e-3))
        yo, so, tro = orc.solve(po, case["method"], case["t0"], case["te"], y0, case["tol"],
                                h0=case["h0"], step_limit=case["step_limit"], trace_cap=100000)
        assert np.array_equal(y, yo)
        assert np.array_equal(tr, tro)


@pytest.mark.parametrize("case", [c for c in KAT if c.get("gpu_only")], ids=lambda c: c["name"])
def test_c2_full_interval(rk, case):
    """Config 2: BRUSS2D N=2000 over t in [0,1] (19.7 k / 15.7 k steps), against the
    reference's full-run footprints and step counts."""
    p = rk.Bruss2D(case["grid"], case["alpha"])
    y, st, _ = run_gpu(rk, p, case["method"], case["t0"], case["te"], p.ode_start(), case["tol"],
                       h0=case["h0"], impl="PipeDls")
    assert (st.steps_acc, st.steps_rej) == (case["acc"], case["rej"])
    want = np.array([float(x) for x in case["footprint"]])
    got = np.array([y[0], y[-1], y.min(), y.max()])
    assert max_rel(got, want) <= REL_TOL
    assert list(got) == list(want)          # and in fact bitwise


@pytest.mark.parametrize("meth,tol", [("DOPRI54", 1e-6), ("VERNER65", 1e-8)])
def test_c1_accuracy_against_long_double(rk, orc, meth, tol):
    """Config 1 on the device against the long-double reference at tol 1e-13
    (the FP64 path's global error, not parity): within 10 tol."""
    p = rk.Bruss2D(128)
    y0 = p.ode_start()
    po = orc.make_ode("bruss2d", 128)
    yr, _ = orc.solve_ld(po, "DOPRI54", 0.0, 4.0, y0, 1e-13, h0=1e-3)
    y, st, _ = run_gpu(rk, p, meth, 0.0, 4.0, y0, tol, h0=1e-3, impl="PipeDls")
    err = max_rel(y, yr)
    print("C1 %s tol %g: max rel error vs long double %.3e" % (meth, tol, err))
    assert err <= 10 * tol


def test_c2_full_interval_d_form(rk):
    """Config 2 (DOPRI54, tol 1e-6) through the full-storage stage kernels of
    the D form over the whole interval: the reference's step counts and
    footprint bitwise (19 743 steps, about 19 s)."""
    case = [c for c in KAT if c["name"] == "c2_bruss2000_dopri54_tol1e-6"][0]
    p = rk.Bruss2D(case["grid"], case["alpha"])
    y, st, _ = run_gpu(rk, p, case["method"], case["t0"], case["te"], p.ode_start(), case["tol"],
                       h0=case["h0"], impl="D")
    assert (st.steps_acc, st.steps_rej) == (case["acc"], case["rej"])
    assert [y[0], y[-1], y.min(), y.max()] == [float(x) for x in case["footprint"]]
    assert st.resident_bytes > 9 * p.n * 8           # full storage: y x2, w_1..w_6, dy, err


def test_zero_length_interval(rk):
    p = rk.Bruss2D(8)
    y0 = p.ode_start()
    y, st, _ = run_gpu(rk, p, "DOPRI54", 1.0, 1.0, y0, 1e-6, h0=1e-3)
    assert st.steps == 0 and np.array_equal(y, y0)


def test_large_grid_fused_equals_d_form(rk):
    """A 20000 x 20000 grid (n = 8e8 doubles, 6.4 GB per vector; the D form
    holds 64 GB): the first attempted steps through the fused kernel and the
    D form agree bitwise -- row offsets and tile counts near the index limits
    of the two forms' different addressing."""
    import torch
    N = 20000
    p = rk.Bruss2D(N, 2e-3)
    print("N=%d: building the start vector" % N, flush=True)
    y0 = dev(p.ode_start())
    out = {}
    for impl in ("PipeDls", "D"):
        print("N=%d: %s" % (N, impl), flush=True)
        y = torch.empty_like(y0)
        st, _ = rk.solve(p, "DOPRI54", 0.0, 1.0, y0, y, 1e-6, h0=1e-7, step_limit=3, impl=impl)
        torch.cuda.synchronize()
        out[impl] = (y, (st.steps_acc, st.steps_rej))
    assert out["PipeDls"][1] == out["D"][1] and sum(out["D"][1]) == 3
    assert torch.equal(out["PipeDls"][0], out["D"][0])
    assert not torch.equal(out["D"][0], y0)                # the solution moved


def test_step_limit(rk):
    p = rk.Bruss2D(40)
    _, st, _ = run_gpu(rk, p, "DOPRI54", 0.0, 4.0, p.ode_start(), 1e-6, h0=1e-3, step_limit=37)
    assert st.steps == 37
