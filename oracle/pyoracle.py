"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAX_S = 16


class Ode(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("grid", C.c_int), ("alpha", C.c_double),
                ("c0", C.c_double), ("d0", C.c_double), ("d1", C.c_double), ("c1", C.c_double),
                ("diff", C.c_double)]


class Method(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("s", C.c_int), ("ord", C.c_int), ("has_emb", C.c_int),
                ("A", (C.c_double * MAX_S) * MAX_S), ("b", C.c_double * MAX_S),
                ("b_hat", C.c_double * MAX_S), ("c", C.c_double * MAX_S)]


class Opts(C.Structure):
    _fields_ = [("h0", C.c_double), ("step_limit", C.c_int), ("block", C.c_int),
                ("trace", C.POINTER(C.c_double)), ("trace_cap", C.c_int), ("dump", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("steps_acc", C.c_int), ("steps_rej", C.c_int), ("t", C.c_double),
                ("h", C.c_double), ("kernel_time", C.c_double)]


BRUSS2D, LV, LV_CHAIN = 1, 2, 3
LV_CONST = (0.1, 0.2, 0.4, 0.2)   # c0, d0, d1, c1 of src/problems/lv.c:31-34

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_ode_init_bruss2d.argtypes = [C.POINTER(Ode), C.c_int, C.c_double]
        L.orc_ode_init_lv.argtypes = [C.POINTER(Ode), C.c_int] + [C.c_double] * 4
        L.orc_ode_init_lv_chain.argtypes = [C.POINTER(Ode), C.c_int] + [C.c_double] * 5
        L.orc_eval_comp.argtypes = [C.POINTER(Ode), C.c_int, C.c_double, dp]
        L.orc_eval_comp.restype = C.c_double
        L.orc_eval_rng.argtypes = [C.POINTER(Ode), C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_acc_dist.argtypes = [C.POINTER(Ode)]
        L.orc_ode_start.argtypes = [C.POINTER(Ode), C.c_double, dp]
        L.orc_method_by_name.argtypes = [C.c_char_p, C.POINTER(Method)]
        L.orc_h_new.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_h_new.restype = C.c_double
        L.orc_pow.argtypes = [dp, dp, dp, C.c_size_t]
        L.orc_pow.restype = None
        L.orc_update_error_max.argtypes = [C.c_double] * 4
        L.orc_update_error_max.restype = C.c_double
        L.orc_initial_stepsize.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double]
        L.orc_initial_stepsize.restype = C.c_double
        sig = [C.POINTER(Ode), C.POINTER(Method), C.c_double, C.c_double, dp, dp, C.c_double,
               C.POINTER(Opts)]
        for fn in ("orc_solve_D", "orc_solve_PipeD", "orc_solve_PipeDls", "orc_solve_A",
                   "orc_solve_E", "orc_solve_F"):
            getattr(L, fn).argtypes = sig + [C.POINTER(Stats)]
        L.orc_solve_pthreads_D.argtypes = sig + [C.c_int, C.POINTER(Stats)]
        ip = C.POINTER(C.c_int)
        L.orc_lv_ensemble.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double,
                                      C.POINTER(Method), C.c_double, C.c_double, C.c_int, ip, ip,
                                      C.c_int]
        L.orc_solve_ld.argtypes = [C.POINTER(Ode), C.c_char_p, C.c_double, C.c_double, dp, dp,
                                   C.c_double, C.c_double, C.c_int, C.POINTER(Stats)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_ode(problem, grid, alpha=2e-3, lv=LV_CONST, diff=1e-3):
    p = Ode()
    L = lib()
    if problem == "bruss2d":
        rc = L.orc_ode_init_bruss2d(C.byref(p), grid, alpha)
    elif problem == "lv":
        rc = L.orc_ode_init_lv(C.byref(p), grid, *lv)
    elif problem == "lvchain":
        rc = L.orc_ode_init_lv_chain(C.byref(p), grid, *lv, diff)
    else:
        raise ValueError(problem)
    if rc:
        raise ValueError("bad problem size")
    return p


def method(name):
    m = Method()
    if lib().orc_method_by_name(name.encode(), C.byref(m)):
        raise ValueError("unknown method " + name)
    return m


def method_arrays(name):
    m = method(name)
    s = m.s
    A = np.array([[m.A[i][j] for j in range(s)] for i in range(s)])
    return (A, np.array(m.b[:s]), np.array(m.b_hat[:s]), np.array(m.c[:s]), s, m.ord)


def ode_start(p, t0=0.0):
    y0 = np.empty(p.n)
    lib().orc_ode_start(C.byref(p), t0, _dp(y0))
    return y0


def eval_rng(p, y, i=0, j=None, t=0.0):
    j = p.n - 1 if j is None else j
    f = np.zeros(p.n)
    yy = np.ascontiguousarray(y, dtype=np.float64)
    lib().orc_eval_rng(C.byref(p), i, j, t, _dp(yy), _dp(f))
    return f


def eval_comp(p, y, i, t=0.0):
    yy = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_eval_comp(C.byref(p), i, t, _dp(yy))


def solve(p, meth, t0, te, y0, tol, h0=0.0, step_limit=0, impl="D", threads=1, trace_cap=0,
          block=0, dump_path=None):
    """Returns (y, stats dict, trace array (attempts x 3: t, h, err) or None).
    dump_path: write the DUMP trajectory (include/tools.h:239-256) there."""
    m = method(meth) if isinstance(meth, str) else meth
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    y = np.empty_like(y0)
    o = Opts()
    o.h0 = h0
    o.step_limit = step_limit
    o.block = block
    tr = None
    if trace_cap:
        tr = np.zeros((trace_cap, 3))
        o.trace = _dp(tr)
        o.trace_cap = trace_cap
    st = Stats()
    L = lib()
    fp = None
    if dump_path:
        libc = C.CDLL(None)
        libc.fopen.restype = C.c_void_p
        libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        libc.fclose.argtypes = [C.c_void_p]
        fp = libc.fopen(dump_path.encode(), b"w")
        o.dump = fp
    args = (C.byref(p), C.byref(m), t0, te, _dp(y0), _dp(y), tol, C.byref(o))
    if impl == "pthreads_D":
        rc = L.orc_solve_pthreads_D(*args, threads, C.byref(st))
    else:
        rc = getattr(L, "orc_solve_" + impl)(*args, C.byref(st))
    if fp:
        libc.fclose(fp)
    if rc:
        raise RuntimeError("oracle solve failed: %d" % rc)
    stats = dict(acc=st.steps_acc, rej=st.steps_rej, steps=st.steps_acc + st.steps_rej,
                 t=st.t, h=st.h, kernel_time=st.kernel_time)
    if tr is not None:
        tr = tr[:min(trace_cap, stats["steps"])]
    return y, stats, tr


def solve_ld(p, meth, t0, te, y0, tol, h0=0.0, step_limit=0):
    """The long-double accuracy reference (oracle_ld.c): (y rounded to double, stats)."""
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    y = np.empty_like(y0)
    st = Stats()
    rc = lib().orc_solve_ld(C.byref(p), meth.encode(), t0, te, _dp(y0), _dp(y), tol, h0,
                            step_limit, C.byref(st))
    if rc:
        raise RuntimeError("long-double oracle solve failed: %d" % rc)
    return y, dict(acc=st.steps_acc, rej=st.steps_rej, steps=st.steps_acc + st.steps_rej, t=st.t)


def lv_ensemble(params, y0, t0, te, meth, tol, h0, step_limit=0, threads=None):
    """Independent LV pairs, each solved by the oracle's seq D with its own
    constants (rows of params: c0, d0, d1, c1).  Returns (y, acc, rej)."""
    m = method(meth) if isinstance(meth, str) else meth
    params = np.ascontiguousarray(params, dtype=np.float64)
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    count = params.shape[0]
    y = np.empty_like(y0)
    acc = np.zeros(count, dtype=np.int32)
    rej = np.zeros(count, dtype=np.int32)
    threads = threads or min(16, os.cpu_count() or 1)
    ip = C.POINTER(C.c_int)
    rc = lib().orc_lv_ensemble(count, _dp(params), _dp(y0), _dp(y), t0, te, C.byref(m), tol, h0,
                               step_limit, acc.ctypes.data_as(ip), rej.ctypes.data_as(ip),
                               threads)
    if rc:
        raise RuntimeError("oracle ensemble failed: %d" % rc)
    return y, acc, rej


def pow_libm(x, y):
    """The C library's pow elementwise (the reference's step-control pow)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    p = np.empty_like(x)
    lib().orc_pow(_dp(x), _dp(y), _dp(p), x.size)
    return p


def footprint(y):
    return [float(y[0]), float(y[-1]), float(np.min(y)), float(np.max(y))]
