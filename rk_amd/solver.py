"""solver() of include/solver.h:47 over the C ABI.

`solve` takes device buffers (torch CUDA tensors or raw device pointers) so
the state stays resident in HBM; `solver` keeps the reference's signature
solver(t0, te, y0, y, tol) on host numpy arrays.
"""
import ctypes as C
import os

import numpy as np

from ._lib import (RKX_CTRL_DEVICE, RKX_CTRL_HOST, RKX_IMPL_A, RKX_IMPL_D, RKX_IMPL_E, RKX_IMPL_F,
                   RKX_IMPL_PIPED, RKX_IMPL_PIPEDLS, Opts, Stats, check, lib)
from .methods import method as _method

IMPLS = {"D": RKX_IMPL_D, "PipeD": RKX_IMPL_PIPED, "PipeDls": RKX_IMPL_PIPEDLS,
         "A": RKX_IMPL_A, "E": RKX_IMPL_E, "F": RKX_IMPL_F}
CTRLS = {"device": RKX_CTRL_DEVICE, "host": RKX_CTRL_HOST}


def _ptr(x):
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if not x.is_cuda or not x.is_contiguous():
            raise ValueError("device buffers must be contiguous CUDA tensors")
        return x.data_ptr()
    raise TypeError("expected a CUDA tensor or a device pointer")


def _opts(impl="PipeDls", ctrl="device", h0=0.0, step_limit=0, tile_rows=0, tile_waves=0,
          trace_cap=0, stream=None, dump_path=None):
    o = Opts()
    lib().rkx_opts_default(C.byref(o))
    o.impl = IMPLS[impl] if isinstance(impl, str) else impl
    o.ctrl = CTRLS[ctrl] if isinstance(ctrl, str) else ctrl
    o.h0 = h0
    o.step_limit = step_limit
    o.tile_rows = tile_rows
    o.tile_waves = tile_waves
    trace = None
    if trace_cap:
        trace = np.zeros((trace_cap, 3))
        o.trace = trace.ctypes.data_as(C.POINTER(C.c_double))
        o.trace_cap = trace_cap
    o.stream = stream or 0
    if dump_path is not None:
        o.dump_path = os.fsencode(dump_path)
    return o, trace


def solve(problem, meth, t0, te, y0_dev, y_dev, tol, **kw):
    """Integrate on device buffers.  Returns (Stats, trace or None); the trace
    holds (t, h, err) of every attempted step up to trace_cap.

    Ordering: without `stream` the solve runs on the library's own stream,
    which is not ordered after torch's; when y0_dev is a tensor, torch's
    current stream is synchronized first, so pending work that writes y0_dev
    has finished.  The call returns after y_dev holds the result."""
    m = _method(meth) if isinstance(meth, str) else meth
    if not kw.get("stream") and hasattr(y0_dev, "data_ptr"):
        import torch
        torch.cuda.current_stream(y0_dev.device).synchronize()
    o, trace = _opts(**kw)
    st = Stats()
    check(lib().rkx_solve(C.byref(problem.desc), C.byref(m), t0, te, C.c_void_p(_ptr(y0_dev)),
                          C.c_void_p(_ptr(y_dev)), tol, C.byref(o), C.byref(st)))
    if trace is not None:
        trace = trace[:min(len(trace), st.steps)]
    return st, trace


def solve_host(problem, meth, t0, te, y0, tol, **kw):
    """Integrate host arrays (copies in and out).  Returns (y, Stats, trace)."""
    m = _method(meth) if isinstance(meth, str) else meth
    o, trace = _opts(**kw)
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    y = np.empty_like(y0)
    st = Stats()
    dp = C.POINTER(C.c_double)
    check(lib().rkx_solve_host(C.byref(problem.desc), C.byref(m), t0, te, y0.ctypes.data_as(dp),
                               y.ctypes.data_as(dp), tol, C.byref(o), C.byref(st)))
    if trace is not None:
        trace = trace[:min(len(trace), st.steps)]
    return y, st, trace


def solver(t0, te, y0, y, tol, problem, meth="DOPRI54", **kw):
    """solver(t0, te, y0, y, tol) of include/solver.h:47: y0 is read, the result
    is written into y (host arrays); the problem and method, which the
    reference binds at compile/link time, are explicit arguments here."""
    out, st, _ = solve_host(problem, meth, t0, te, y0, tol, **kw)
    y[...] = out
    return st
