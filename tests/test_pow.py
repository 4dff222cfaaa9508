"""The step control's pow (rk_amd/csrc/rkx_pow.hpp), evaluated on the host
CPU through rkx_pow_host, against the C library's pow (the reference's one
libm call, include/tools.h:226) bitwise.

The restated algorithm and tables must reproduce the C library's result in
every case, including those where the C library is not correctly rounded:
test_libm_is_not_correctly_rounded_somewhere finds such inputs, so agreement
there cannot come from computing pow merely accurately."""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "rk_amd", "csrc")


@pytest.fixture(scope="module")
def rkpow():
    from rk_amd._lib import check, lib
    L = lib()

    def f(x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        p = np.empty_like(x)
        dp = C.POINTER(C.c_double)
        check(L.rkx_pow_host(x.ctypes.data_as(dp), y.ctypes.data_as(dp), p.ctypes.data_as(dp),
                             x.size))
        return p
    return f


def step_control_args(rng, n):
    """x = tol/err over [2^-14, 2^16] with random mantissas, y = 1/(ord+1) for
    ord = 1..12 (every table's order and more)."""
    e = rng.integers(-14, 17, n)
    m = rng.integers(0, 1 << 52, n, dtype=np.uint64)
    bits = ((e + 1023).astype(np.uint64) << np.uint64(52)) | m
    x = bits.view(np.float64)
    y = 1.0 / (rng.integers(1, 13, n) + 1.0)
    return x, y


def test_step_control_domain(rkpow, orc):
    rng = np.random.default_rng(20261203)
    x, y = step_control_args(rng, 2_000_000)
    assert np.array_equal(rkpow(x, y).view(np.uint64), orc.pow_libm(x, y).view(np.uint64))


def test_every_positive_double(rkpow, orc):
    """x over all positive finite doubles (subnormals included), y uniform in
    (0, 1/2]."""
    rng = np.random.default_rng(7)
    n = 1_000_000
    bits = rng.integers(1, 0x7ff0000000000000, n, dtype=np.uint64)
    x = bits.view(np.float64)
    y = (rng.integers(1, 1 << 52, n).astype(np.float64)) * 2.0 ** -53
    assert np.array_equal(rkpow(x, y).view(np.uint64), orc.pow_libm(x, y).view(np.uint64))


def test_special_values(rkpow, orc):
    x = np.array([0.0, -0.0, np.inf, 1.0, 5e-324, 2.2250738585072014e-308, 1e-6 / 1.7e308,
                  1e-10 / 1.7e308, np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0), 2.0 ** 1023,
                  1.7976931348623157e308])
    for y in (1.0 / 2, 1.0 / 3, 1.0 / 6, 1.0 / 7, 1.0 / 9, 2.0 ** -60):
        yy = np.full_like(x, y)
        assert np.array_equal(rkpow(x, yy).view(np.uint64), orc.pow_libm(x, yy).view(np.uint64))


def test_libm_is_not_correctly_rounded_somewhere(rkpow, orc):
    """Among 60 000 step-control arguments the C library's pow differs from the
    correctly rounded value at least once; rkx_pow_host follows it there."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 200
    rng = np.random.default_rng(11)
    x, y = step_control_args(rng, 60_000)
    got, want = rkpow(x, y), orc.pow_libm(x, y)
    assert np.array_equal(got, want)
    misrounded = 0
    for xi, yi, wi in zip(x, y, want):
        cr = float(mpmath.power(mpmath.mpf(float(xi)), mpmath.mpf(float(yi))))
        misrounded += cr != wi
    print("C library pow not correctly rounded in %d of %d cases" % (misrounded, x.size))
    assert misrounded >= 1


def test_tables_regenerate_from_their_definitions(tmp_path):
    """rkx_pow_tables.inc equals what gen_pow_tables.py derives from the
    published table definitions."""
    pytest.importorskip("mpmath")
    import subprocess
    import sys
    out = tmp_path / "tables.inc"
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_pow_tables.py"), str(out)], check=True)
    assert out.read_text() == open(os.path.join(CSRC, "rkx_pow_tables.inc")).read()
