"""Pins of the oracle that do not need the reference: order conditions of the
method tables, the BRUSS2D and LV right-hand sides against brute force and
closed forms, and the step-size controller on synthetic error values."""
import math

import numpy as np
import pytest

EMB = ["RKF23", "DOPRI54", "VERNER65", "DOPRI87"]
ALL = ["HEUN2", "RK4", "SSPRK3"] + EMB
# order of the propagated weights b, and of the embedded weights b_hat
ORDERS = {"HEUN2": (2, None), "RK4": (4, None), "SSPRK3": (3, None), "RKF23": (2, 3),
          "DOPRI54": (5, 4), "VERNER65": (5, 6), "DOPRI87": (8, 7)}


def conditions(A, b, c, p):
    """Residuals of the Butcher order conditions up to order min(p, 4) plus the
    bushy trees up to order p."""
    res = []
    for k in range(1, p + 1):
        res.append(b @ c ** (k - 1) - 1.0 / k)
    if p >= 3:
        res.append(b @ (A @ c) - 1 / 6)
    if p >= 4:
        res.append(b @ (c * (A @ c)) - 1 / 8)
        res.append(b @ (A @ c ** 2) - 1 / 12)
        res.append(b @ (A @ (A @ c)) - 1 / 24)
    return np.abs(np.array(res))


@pytest.mark.parametrize("name", ALL)
def test_method_table(orc, name):
    A, b, bh, c, s, ordr = orc.method_arrays(name)
    if name == "DOPRI87":
        # The reference's DOPRI87 has A[6][3] = 77736583/692538347
        # (src/methods.c:630); the published Prince-Dormand value is
        # 77736538/692538347 (two digits transposed).  Parity keeps the
        # reference's value; the row-sum condition pins that this is the only
        # deviation.
        assert A[6][3] == 77736583.0 / 692538347.0
        assert abs(A[6].sum() - c[6]) > 6e-8
        A = A.copy()
        A[6][3] = 77736538.0 / 692538347.0
    assert np.allclose(A.sum(axis=1), c, atol=1e-14, rtol=0)   # row sums = c
    assert np.all(np.triu(A) == 0.0)                            # explicit
    pb, pbh = ORDERS[name]
    assert conditions(A, b, c, pb).max() < 1e-13
    if pbh is not None:
        assert conditions(A, bh, c, pbh).max() < 1e-13
        # the weight vectors must not coincide (else there is no error estimate)
        assert np.abs(b - bh).max() > 1e-3
    else:
        assert np.all(bh == 0.0)


def test_verner65_a65_is_zero(orc):
    A, *_ = orc.method_arrays("VERNER65")
    assert A[6][5] == 0.0 and abs(A[6].sum() - 8 / 9) < 1e-13


def dense_bruss(N, alpha, y):
    """Brute force: 5-point Neumann Laplacian by explicit mirror indexing."""
    u = y[0::2].reshape(N, N)
    v = y[1::2].reshape(N, N)

    def lap(z):
        zp = np.pad(z, 1, mode="reflect")   # mirror: ghost(-1) = z(1)
        return zp[:-2, 1:-1] + zp[2:, 1:-1] + zp[1:-1, :-2] + zp[1:-1, 2:] - 4 * z

    N1 = N - 1.0
    fu = 1.0 + u * u * v - 4.4 * u + alpha * N1 * N1 * lap(u)
    fv = 3.4 * u - u * u * v + alpha * N1 * N1 * lap(v)
    f = np.empty_like(y)
    f[0::2] = fu.ravel()
    f[1::2] = fv.ravel()
    return f


@pytest.mark.parametrize("N", [2, 3, 4, 6, 9])
def test_bruss_rhs_bruteforce(orc, N):
    rng = np.random.default_rng(N)
    p = orc.make_ode("bruss2d", N, 3e-3)
    y = rng.uniform(0.2, 4.0, p.n)
    f = orc.eval_rng(p, y)
    assert np.allclose(f, dense_bruss(N, 3e-3, y), rtol=1e-13, atol=1e-13)
    # range form == component form, bitwise, for every sub-range
    for i in range(p.n):
        for j in range(i, p.n, max(1, p.n // 7)):
            fr = orc.eval_rng(p, y, i, j)
            for q in (i, j):
                assert fr[q] == orc.eval_comp(p, y, q)


def test_bruss_steady_state(orc):
    p = orc.make_ode("bruss2d", 5)
    y = np.empty(p.n)
    y[0::2] = 1.0
    y[1::2] = 3.4
    assert np.abs(orc.eval_rng(p, y)).max() < 1e-15


def test_bruss_start_vector(orc):
    N = 7
    p = orc.make_ode("bruss2d", N)
    y = orc.ode_start(p)
    row, col = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    assert np.array_equal(y[0::2], (0.5 + col / (N - 1)).ravel())
    assert np.array_equal(y[1::2], (1.0 + 5.0 * row / (N - 1)).ravel())


def test_lv_rhs_closed_form(orc):
    p = orc.make_ode("lv", 4)
    rng = np.random.default_rng(1)
    y = rng.uniform(0.1, 2.0, p.n)
    f = orc.eval_rng(p, y)
    x, z = y[0::2], y[1::2]
    assert np.array_equal(f[0::2], 0.1 * x - 0.2 * x * z)
    assert np.array_equal(f[1::2], 0.4 * x * z - 0.2 * z)


def test_lv_invariant(orc):
    """V = d1 x - c1 ln x + d0 y - c0 ln y is conserved by the exact flow."""
    p = orc.make_ode("lv", 1)
    y0 = orc.ode_start(p)
    y, st, _ = orc.solve(p, "DOPRI54", 0.0, 20.0, y0, 1e-10, h0=1e-3)

    def V(v):
        x, z = v
        return 0.4 * x - 0.2 * math.log(x) + 0.2 * z - 0.1 * math.log(z)
    assert abs(V(y) - V(y0)) < 1e-8


# four distinct constants (the defaults have d0 == c1 == 0.2, so a swap of two
# of them would pass the checks above unnoticed)
LV_DISTINCT = (0.37, 1.21, 0.53, 1.77)


def test_lv_rhs_closed_form_distinct_constants(orc):
    c0, d0, d1, c1 = LV_DISTINCT
    p = orc.make_ode("lv", 5, lv=LV_DISTINCT)
    rng = np.random.default_rng(3)
    y = rng.uniform(0.1, 2.0, p.n)
    f = orc.eval_rng(p, y)
    x, z = y[0::2], y[1::2]
    assert np.array_equal(f[0::2], c0 * x - d0 * x * z)
    assert np.array_equal(f[1::2], d1 * x * z - c1 * z)
    for i in range(p.n):
        assert orc.eval_comp(p, y, i) == f[i]


def test_lv_invariant_distinct_constants(orc):
    """V = d1 x - c1 ln x + d0 y - c0 ln y is conserved for x' = c0 x - d0 x y,
    y' = d1 x y - c1 y; with any two constants swapped it is not."""
    c0, d0, d1, c1 = LV_DISTINCT
    p = orc.make_ode("lv", 1, lv=LV_DISTINCT)
    y0 = np.array([0.8, 0.45])
    y, st, _ = orc.solve(p, "VERNER65", 0.0, 15.0, y0, 1e-11, h0=1e-3)

    def V(v, a, b, c, d):
        x, z = v
        return c * x - d * math.log(x) + b * z - a * math.log(z)
    assert abs(V(y, c0, d0, d1, c1) - V(y0, c0, d0, d1, c1)) < 1e-8
    # the same check with d0 and c1 exchanged fails: the pin sees the swap
    assert abs(V(y, c0, c1, d1, d0) - V(y0, c0, c1, d1, d0)) > 1e-3


def test_chain_zero_diffusion_is_pairs(orc):
    """The LV chain with D = 0 is the independent LV pairs, bitwise."""
    pc = orc.make_ode("lvchain", 16, diff=0.0)
    pl = orc.make_ode("lv", 16)
    rng = np.random.default_rng(2)
    y = rng.uniform(0.1, 2.0, pc.n)
    assert np.array_equal(orc.eval_rng(pc, y), orc.eval_rng(pl, y))


def test_chain_diffusion_bruteforce(orc):
    pc = orc.make_ode("lvchain", 10, diff=0.37)
    rng = np.random.default_rng(3)
    y = rng.uniform(0.1, 2.0, pc.n)
    x, z = y[0::2], y[1::2]
    fx = 0.1 * x - 0.2 * x * z + 0.37 * (np.roll(x, 1) + np.roll(x, -1) - 2 * x)
    fz = 0.4 * x * z - 0.2 * z + 0.37 * (np.roll(z, 1) + np.roll(z, -1) - 2 * z)
    f = orc.eval_rng(pc, y)
    assert np.allclose(f[0::2], fx, rtol=1e-14, atol=1e-15)
    assert np.allclose(f[1::2], fz, rtol=1e-14, atol=1e-15)


def test_step_control_synthetic(orc):
    L = orc.lib()
    h = 1e-3
    assert L.orc_h_new(h, 0.0, 1e-6, 5) == 2 * h           # err = 0 -> AMAX
    assert L.orc_h_new(h, math.inf, 1e-6, 5) == 0.5 * h    # err = inf -> AMIN
    assert L.orc_h_new(h, 1e-6, 1e-6, 5) == h * 0.9
    assert L.orc_h_new(h, 5.0, 0.0, 5) == h                # fixed step
    want = h * min(2.0, max(0.5, 0.9 * (1e-6 / 3e-7) ** (1.0 / 6.0)))
    assert L.orc_h_new(h, 3e-7, 1e-6, 5) == want
    # relative max-norm with the 0/0 sentinel and NaN blindness (tools.h:280-288)
    assert L.orc_update_error_max(0.5, 1.0, 0.0, 0.0) == 1.7e308
    assert L.orc_update_error_max(0.5, math.nan, 1.0, 1.0) == 0.5
    assert L.orc_update_error_max(0.1, 1.0, -4.0, 2.0) == 0.25
    # initial step (src/tools.c:25-36)
    assert L.orc_initial_stepsize(0.1, 1e-9, 200, 0.0) == 1e-8
    assert L.orc_initial_stepsize(1.0, 1e-3, 200, 0.0) == 1.0 / 200
    assert L.orc_initial_stepsize(1.0, 1e-6, 0, 0.0) == 10.0 * 1e-6
    assert L.orc_initial_stepsize(1.0, 1e-6, 0, 1e-3) == 1e-3


def test_step_limit_and_final_clip(orc):
    p = orc.make_ode("bruss2d", 6)
    y0 = orc.ode_start(p)
    _, st, _ = orc.solve(p, "DOPRI54", 0.0, 4.0, y0, 1e-6, h0=1e-3, step_limit=7)
    assert st["steps"] == 7
    _, st, tr = orc.solve(p, "RK4", 0.0, 0.1, y0, 0.0, h0=0.03, trace_cap=10)
    assert st["steps"] == 4 and st["t"] == 0.1        # 3 full steps, then clipped
    assert tr[3, 1] == 0.1 - tr[3, 0]


def test_exponential_convergence(orc):
    """y' = lambda y has no reference counterpart here; instead the LV pair at a tight
    tolerance converges to a tighter one (accuracy, not parity)."""
    p = orc.make_ode("lv", 1)
    y0 = orc.ode_start(p)
    ya, *_ = orc.solve(p, "DOPRI87", 0.0, 5.0, y0, 1e-13, h0=1e-4)
    yb, *_ = orc.solve(p, "DOPRI54", 0.0, 5.0, y0, 1e-10, h0=1e-4)
    assert np.abs(ya - yb).max() < 1e-8


def test_long_double_reference_pinned(orc):
    """The long-double accuracy reference (orc_solve_ld) is pinned by two
    different tables agreeing at tight tolerances and by the LV pair's
    conserved quantity V = d1 x - c1 ln x + d0 y - c0 ln y."""
    p = orc.make_ode("lv", 1)
    y0 = orc.ode_start(p)
    ya, _ = orc.solve_ld(p, "DOPRI87", 0.0, 20.0, y0, 1e-16, h0=1e-4)
    yb, _ = orc.solve_ld(p, "DOPRI54", 0.0, 20.0, y0, 1e-14, h0=1e-4)
    assert np.abs(ya - yb).max() < 1e-11
    c0, d0, d1, c1 = 0.1, 0.2, 0.4, 0.2
    V = lambda x, y: d1 * x - c1 * np.log(x) + d0 * y - c0 * np.log(y)
    assert abs(V(*ya) - V(*y0)) < 1e-11
    pb = orc.make_ode("bruss2d", 16)
    z0 = orc.ode_start(pb)
    za, _ = orc.solve_ld(pb, "DOPRI54", 0.0, 4.0, z0, 1e-13, h0=1e-3)
    zb, _ = orc.solve_ld(pb, "VERNER65", 0.0, 4.0, z0, 1e-13, h0=1e-3)
    assert np.max(np.abs(za - zb) / np.abs(za)) < 5e-11


@pytest.mark.parametrize("meth", ["DOPRI54", "VERNER65"])
@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_fp64_accuracy_against_long_double(orc, meth, tol):
    """Global error of the FP64 solve (the parity path) against the
    long-double run at tol 1e-13: within 10 tol (relative, max over
    components) on BRUSS2D N=16, t in [0, 4]."""
    p = orc.make_ode("bruss2d", 16)
    y0 = orc.ode_start(p)
    yr, _ = orc.solve_ld(p, "DOPRI54", 0.0, 4.0, y0, 1e-13, h0=1e-3)
    yd, _, _ = orc.solve(p, meth, 0.0, 4.0, y0, tol, h0=1e-3)
    assert np.max(np.abs(yd - yr) / np.abs(yr)) <= 10 * tol
