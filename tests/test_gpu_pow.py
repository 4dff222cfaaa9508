"""The device step control's pow (rkx_pow_device) against the C library's pow
on this host, bitwise: 10^8 step-control arguments, the whole positive
double range, special values, and every (tol/err, 1/(ord+1)) pair the
reference's step control produced along C1 solves of three tables."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def devpow():
    import torch
    assert torch.cuda.is_available()
    from rk_amd._lib import check, lib
    L = lib()

    def f(x, y):
        X = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
        Y = torch.from_numpy(np.ascontiguousarray(y, dtype=np.float64)).cuda()
        P = torch.empty_like(X)
        torch.cuda.synchronize()
        check(L.rkx_pow_device(X.data_ptr(), Y.data_ptr(), P.data_ptr(), X.numel(), None))
        return P.cpu().numpy()
    return f


def test_1e8_step_control_arguments(devpow, orc):
    from test_pow import step_control_args
    rng = np.random.default_rng(12345)
    chunk, total, bad = 10_000_000, 0, 0
    for c in range(10):
        x, y = step_control_args(rng, chunk)
        got, want = devpow(x, y), orc.pow_libm(x, y)
        bad += int(np.count_nonzero(got.view(np.uint64) != want.view(np.uint64)))
        total += chunk
        print("chunk %d: %d mismatches so far of %d" % (c, bad, total), flush=True)
    assert bad == 0


def test_every_positive_double(devpow, orc):
    rng = np.random.default_rng(99)
    n = 10_000_000
    x = rng.integers(1, 0x7ff0000000000000, n, dtype=np.uint64).view(np.float64)
    y = rng.integers(1, 1 << 52, n).astype(np.float64) * 2.0 ** -53
    assert np.array_equal(devpow(x, y).view(np.uint64), orc.pow_libm(x, y).view(np.uint64))


def test_special_values(devpow, orc):
    x = np.array([0.0, -0.0, np.inf, 1.0, 5e-324, 2.2250738585072014e-308, 1e-6 / 1.7e308,
                  1e-10 / 1.7e308, np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0), 2.0 ** 1023,
                  1.7976931348623157e308])
    for y in (1.0 / 2, 1.0 / 3, 1.0 / 6, 1.0 / 7, 1.0 / 9, 2.0 ** -60):
        yy = np.full_like(x, y)
        assert np.array_equal(devpow(x, yy).view(np.uint64), orc.pow_libm(x, yy).view(np.uint64))


@pytest.mark.parametrize("meth,tol", [("DOPRI54", 1e-6), ("VERNER65", 1e-8), ("RKF23", 1e-6)])
def test_solve_arguments(devpow, orc, meth, tol):
    """The exact arguments of a C1 solve's step control (from the oracle's
    per-attempt error norms)."""
    po = orc.make_ode("bruss2d", 128)
    y0 = orc.ode_start(po)
    _, st, tr = orc.solve(po, meth, 0.0, 4.0, y0, tol, h0=1e-3, trace_cap=100000)
    m = orc.method(meth)
    x = tol / tr[:, 2]
    y = np.full_like(x, 1.0 / (m.ord + 1.0))
    assert np.array_equal(devpow(x, y).view(np.uint64), orc.pow_libm(x, y).view(np.uint64))
