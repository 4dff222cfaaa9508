"""rkx_solve_noalloc (the solve that allocates nothing) against rkx_solve
through the C ABI: identical step counts and bitwise the same y, on two
calls in a row whose second is smaller (the scratch array keeps the first
call's contents), for BRUSS2D grids, LV chains and LV pairs, with device and host step
control, in place (y0 == y) and into a separate output."""
import ctypes as C

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


def call(fn, p, meth, te, y0, tol, h0, inplace=False, **kw):
    import torch
    from rk_amd._lib import Stats, check, lib
    from rk_amd.methods import method
    from rk_amd.solver import _opts
    y0d = torch.from_numpy(np.ascontiguousarray(y0)).cuda()
    yd = y0d if inplace else torch.empty_like(y0d)
    torch.cuda.synchronize()
    o, _ = _opts(h0=h0, **kw)
    st = Stats()
    m = method(meth)
    print("%s n=%d %s" % (fn, p.n, meth), flush=True)
    check(getattr(lib(), fn)(C.byref(p.desc), C.byref(m), 0.0, te, C.c_void_p(y0d.data_ptr()),
                             C.c_void_p(yd.data_ptr()), tol, C.byref(o), C.byref(st)))
    torch.cuda.synchronize()
    return yd.cpu().numpy(), st


@pytest.mark.parametrize("ctrl", ["device", "host"])
def test_noalloc_equals_solve_two_calls(rk, ctrl):
    cases = [(rk.Bruss2D(300, 2.5e-3), "DOPRI54", 0.1, 1e-6, False),
             (rk.Bruss2D(90, 2e-3), "VERNER65", 0.3, 1e-8, True),
             (rk.LVChain(60_000, diff=1e-3), "DOPRI54", 1.0, 1e-6, False),
             (rk.LVChain(7_001, diff=2e-3), "RKF23", 1.0, 1e-5, True),
             (rk.LV(50_000), "DOPRI54", 4.0, 1e-8, False),
             (rk.LV(3_001), "VERNER65", 2.0, 1e-8, True)]
    for p, meth, te, tol, inplace in cases:
        y0 = p.ode_start_seeded(7, 1, 0.01)
        ya, sa = call("rkx_solve_noalloc", p, meth, te, y0, tol, 1e-3, inplace=inplace, ctrl=ctrl)
        yb, sb = call("rkx_solve", p, meth, te, y0, tol, 1e-3, ctrl=ctrl)
        assert sa.steps > 3
        assert (sa.steps_acc, sa.steps_rej) == (sb.steps_acc, sb.steps_rej), (p.n, meth)
        assert np.array_equal(ya, yb), (p.n, meth)


def test_noalloc_limits(rk):
    """Beyond the scratch array, the D form, or a problem without a fused
    kernel: refused before any device work."""
    from rk_amd._lib import RkxError
    big = rk.Bruss2D(1447)                    # 2 * 1447^2 = 4 187 618 > 4 186 112
    with pytest.raises(RkxError) as e:
        call("rkx_solve_noalloc", big, "DOPRI54", 0.1, np.zeros(big.n), 1e-6, 1e-3)
    assert e.value.code == -4
    p = rk.Bruss2D(20)
    with pytest.raises(RkxError) as e:
        call("rkx_solve_noalloc", p, "DOPRI54", 0.1, p.ode_start(), 1e-6, 1e-3, impl="D")
    assert e.value.code == -1
    ch = rk.LVChain(100)                      # the chain kernel has no DOPRI87 form
    with pytest.raises(RkxError) as e:
        call("rkx_solve_noalloc", ch, "DOPRI87", 0.1, ch.ode_start(), 1e-6, 1e-3)
    assert e.value.code == -1
