"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/rkx.h declares, and its host-side pieces (method tables, start
vectors, descriptors) equal the oracle's bitwise.  No device calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(REPO, "include", "rkx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|size_t|char\s*\*|const char\s*\*)\s*\**\s*(rkx_\w+)\s*\(",
                       src, flags=re.M)
    return sorted(set(names))


def test_library_exports_header():
    import rk_amd
    L = rk_amd.lib()
    names = header_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
    # and the Python declarations cover them all
    from rk_amd._lib import SIGNATURES
    assert sorted(n for n, _, _ in SIGNATURES) == names


@pytest.mark.parametrize("name", ["HEUN2", "RK4", "SSPRK3", "RKF23", "DOPRI54", "VERNER65",
                                  "DOPRI87"])
def test_method_tables_match_oracle(orc, name):
    import rk_amd
    A, b, bh, c, s, o = rk_amd.method_table(name)
    Ao, bo, bho, co, so, oo = orc.method_arrays(name)
    assert (s, o) == (so, oo)
    assert np.array_equal(A, Ao) and np.array_equal(b, bo)
    assert np.array_equal(bh, bho) and np.array_equal(c, co)
    assert rk_amd.method(name).has_emb == orc.method(name).has_emb


@pytest.mark.parametrize("N", [2, 3, 17, 128])
def test_bruss_start_matches_oracle(orc, N):
    import rk_amd
    p = rk_amd.Bruss2D(N)
    assert p.ode_size == 2 * N * N and p.ode_acc_dist() == 2 * N
    assert np.array_equal(p.ode_start(), orc.ode_start(orc.make_ode("bruss2d", N)))
    # zero noise amplitude reproduces the reference start vector
    assert np.array_equal(p.ode_start_seeded(7, 3, 0.0), p.ode_start())
    a = p.ode_start_seeded(7, 3, 0.01)
    assert np.abs(a - p.ode_start()).max() <= 0.01 and not np.array_equal(a, p.ode_start())
    assert np.array_equal(a, p.ode_start_seeded(7, 3, 0.01))


def test_lv_descriptors(orc):
    import rk_amd
    p = rk_amd.LV(5)
    assert p.ode_size == 10 and p.ode_acc_dist() == 1
    assert np.array_equal(p.ode_start(), orc.ode_start(orc.make_ode("lv", 5)))
    q = rk_amd.LVChain(9)
    assert q.ode_size == 18 and q.ode_acc_dist() == 2


@pytest.mark.parametrize("name", ["HEUN2", "RK4", "SSPRK3", "RKF23", "DOPRI54", "VERNER65",
                                  "DOPRI87"])
def test_footprints_per_form(name):
    """Resident vectors of n per step form: the fused forms two (BRUSS2D, LV
    pairs, LV chain), D, A and E s + 3, F s + 4.  The LV chain has no DOPRI87
    kernel: there a low-storage request runs the loop order with the fewest
    vector passes, A (120 against F 130, E 143, D 207 for DOPRI87's sparse
    table), and holds A's vectors."""
    import ctypes as C
    import rk_amd
    from rk_amd._lib import lib
    m = rk_amd.method(name)
    for p in (rk_amd.Bruss2D(40), rk_amd.LV(1000), rk_amd.LVChain(1000)):
        vec = p.n * 8
        fp = {impl: lib().rkx_solve_footprint(C.byref(p.desc), C.byref(m), code) // vec
              for impl, code in (("D", 0), ("PipeDls", 2), ("A", 3), ("E", 4), ("F", 5))}
        assert fp["D"] == fp["A"] == fp["E"] == m.s + 3 and fp["F"] == m.s + 4
        if isinstance(p, rk_amd.LVChain) and name == "DOPRI87":
            assert fp["PipeDls"] == m.s + 3
        else:
            assert fp["PipeDls"] == 2


def test_bad_arguments_raise():
    import rk_amd
    with pytest.raises(rk_amd.RkxError):
        rk_amd.Bruss2D(1)
    with pytest.raises(rk_amd.RkxError):
        rk_amd.method("NOPE")
    with pytest.raises(rk_amd.RkxError):
        rk_amd.LVChain(2)


def test_dropin_binaries_built():
    """The reference-style solver binaries exist, define solver() and the
    ode.h / methods.h symbols, and load librkx.so from the package."""
    import subprocess
    from rk_amd.build import BIN, DROPIN
    for name in DROPIN:
        path = os.path.join(BIN, name)
        assert os.path.exists(path), name
        syms = subprocess.run(["nm", path], capture_output=True, text=True).stdout
        for s in ("solver", "ode_start", "ode_eval_rng", "ode_eval_comp", "ode_acc_dist",
                  "ode_size", "DOPRI54", "VERNER65"):
            assert re.search(r" [TDB] %s$" % s, syms, re.M), (name, s)
        ldd = subprocess.run(["ldd", path], capture_output=True, text=True).stdout
        assert "rk_amd/bin/../librkx.so" in ldd


def test_reference_driver_binaries_link_the_gpu_solver():
    """oracle/_ref/refmain_*: main() and the ode_* plugin come from the
    reference's own objects, solver(), the method functions and
    rkx_dropin_problem() from rk_amd/csrc/dropin, over rk_amd/librkx.so."""
    import subprocess
    ref = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                       "_ref")
    names = [f for f in os.listdir(ref) if f.startswith("refmain_")] if os.path.isdir(ref) else []
    if not names:
        pytest.skip("reference-driver binaries not built (no reference tree at build time)")
    for name in names:
        path = os.path.join(ref, name)
        syms = subprocess.run(["nm", path], capture_output=True, text=True).stdout
        for s in ("main", "solver", "rkx_dropin_problem", "ode_start", "ode_eval_comp",
                  "ode_size", "DOPRI54", "VERNER65"):
            assert re.search(r" [TDB] %s$" % s, syms, re.M), (name, s)
        assert "rkx_solve_host" in syms
        ldd = subprocess.run(["ldd", path], capture_output=True, text=True).stdout
        assert "oracle/_ref/../../rk_amd/librkx.so" in ldd, ldd


def test_serving_summary_statistics():
    """Latency percentiles and throughput of the serving summary (host logic)."""
    from rk_amd.serve import Request, summarize
    rs = []
    for k in range(100):
        r = Request(k, float(k), None, "DOPRI54", 0.0, 1.0, np.zeros(2), 1e-6)
        r.finished = r.arrival + (k + 1) * 0.01
        rs.append(r)
    s = summarize(rs, 200.0)
    assert s["requests"] == 100 and s["solves_per_s"] == 0.5
    assert abs(s["p50_latency_s"] - 0.505) < 1e-12
    assert abs(s["p99_latency_s"] - np.percentile(np.arange(1, 101) * 0.01, 99)) < 1e-12


def test_bruss2d_size_limit():
    """n = 2N^2 must fit a C int: N = 32767 is the largest grid (descriptor only)."""
    import rk_amd
    assert rk_amd.Bruss2D(32767).n == 2 * 32767 * 32767
    with pytest.raises(rk_amd.RkxError):
        rk_amd.Bruss2D(32768)
    with pytest.raises(rk_amd.RkxError):
        rk_amd.Bruss2D(1)


def test_solve_argument_errors_before_any_device_call():
    """rkx_solve / rkx_batch_* validate their arguments on the host and return
    RKX_EINVAL (with a message) before touching the device: a non-explicit
    table, unknown impl / step-control modes, null vectors, a trace length
    without a buffer, a batch with an unsupported control mode."""
    import ctypes as C
    import rk_amd
    from rk_amd._lib import Method, Opts, Stats, lib
    L = lib()
    p = rk_amd.Bruss2D(16)
    m = rk_amd.method("DOPRI54")
    dummy = C.c_void_p(0x1000)          # never dereferenced: validation fails first
    st = Stats()

    def solve_rc(meth, opts, y0=dummy, y=dummy):
        return L.rkx_solve(C.byref(p.desc), C.byref(meth), 0.0, 1.0, y0, y, 1e-6,
                           C.byref(opts), C.byref(st))

    o = Opts()
    L.rkx_opts_default(C.byref(o))
    implicit = Method()
    C.memmove(C.byref(implicit), C.byref(m), C.sizeof(Method))
    implicit.A[2][2] = 0.5                       # diagonal entry: not explicit
    assert solve_rc(implicit, o) == -1 and b"explicit" in L.rkx_last_error()
    for field, bad in (("impl", 7), ("ctrl", 9)):
        ob = Opts()
        L.rkx_opts_default(C.byref(ob))
        setattr(ob, field, bad)
        assert solve_rc(m, ob) == -1
    assert solve_rc(m, o, y0=C.c_void_p(0)) == -1 and b"null" in L.rkx_last_error()
    ot = Opts()
    L.rkx_opts_default(C.byref(ot))
    ot.trace_cap = 10                             # no trace buffer
    assert solve_rc(m, ot) == -1 and b"trace" in L.rkx_last_error()
    from rk_amd.batch import BatchOpts
    bo = BatchOpts()
    L.rkx_batch_opts_default(C.byref(bo))
    bo.ctrl = rk_amd.RKX_CTRL_HOST                # per-step host control: not for batches
    h = C.c_void_p()
    assert L.rkx_batch_create(C.byref(h), C.byref(bo)) == -1
    assert L.rkx_batch_create(None, C.byref(bo)) == -1
    ens = lambda count, h0: L.rkx_lv_ensemble(count, dummy, dummy, dummy, 0.0, 1.0, C.byref(m),
                                              1e-6, h0, 0, dummy, dummy, None)
    assert ens(0, 1e-3) == -1 and ens(16, 0.0) == -1
    assert ens((1 << 30) + 1, 1e-3) == -1 and b"2^30" in L.rkx_last_error()


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of the header's structs (rk_amd/_lib.py, rk_amd/batch.py)
    have the C compiler's size and field offsets, so no field is read at the
    wrong place when the header changes."""
    import ctypes as C
    import subprocess
    from rk_amd._lib import Method, Ode, Opts, Stats
    from rk_amd.batch import BatchOpts
    structs = {"rkx_ode": Ode, "rkx_method": Method, "rkx_opts": Opts, "rkx_stats": Stats,
               "rkx_batch_opts": BatchOpts}
    lines = ['#include "rkx.h"', "#include <stddef.h>", "#include <stdio.h>", "int main(void) {"]
    want = []
    for cname, cls in structs.items():
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        want.append(C.sizeof(cls))
        for fname, _ in cls._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, fname))
            want.append(getattr(cls, fname).offset)
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(REPO, "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == want
