"""Ensembles of independent Lotka-Volterra pairs (rkx_lv_ensemble).

Each system has its own constants (x' = c0 x - d0 x y, y' = d1 x y - c1 y),
start value, step size and accept/reject; one GPU thread integrates one
system over the whole interval.
"""
import ctypes as C
import math

import numpy as np

from ._lib import check, lib
from .methods import method as _method


def lv_ensemble_inputs(count=262144, seed=1):
    """Seeded constants and start values of the batched LV configuration:
    a, b, c, d log-uniform on [0.1, 2] (x' = a x - b x y, y' = c x y - d y) and
    x0, y0 uniform on [0.5, 2]; params rows are (c0, d0, d1, c1) = (a, b, c, d)."""
    g = np.random.Generator(np.random.Philox(seed))
    lo, hi = math.log(0.1), math.log(2.0)
    params = np.exp(g.uniform(lo, hi, size=(count, 4)))
    y0 = g.uniform(0.5, 2.0, size=(count, 2))
    return np.ascontiguousarray(params), np.ascontiguousarray(y0)


def initial_step(te, t0, tol, step_limit=0, h0=0.0):
    """initial_stepsize (src/tools.c:25-36): H_INIT, else min(10 tol, H/STEP_LIMIT)."""
    if h0 > 0:
        return h0
    H = te - t0
    return min(10.0 * tol, H / step_limit if step_limit > 0 else math.inf)


def lv_ensemble(params, y0, t0, te, meth, tol, h0=0.0, step_limit=0, stream=None):
    """params: (count, 4) CUDA tensor or array; y0: (count, 2).  Returns
    (y, acc, rej) as CUDA tensors."""
    import torch
    m = _method(meth) if isinstance(meth, str) else meth
    P = torch.as_tensor(params, dtype=torch.float64, device="cuda").contiguous()
    Y0 = torch.as_tensor(y0, dtype=torch.float64, device="cuda").contiguous()
    count = P.shape[0]
    Y = torch.empty_like(Y0)
    acc = torch.zeros(count, dtype=torch.int32, device="cuda")
    rej = torch.zeros(count, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    h = initial_step(te, t0, tol, step_limit, h0)
    check(lib().rkx_lv_ensemble(count, C.c_void_p(P.data_ptr()), C.c_void_p(Y0.data_ptr()),
                                C.c_void_p(Y.data_ptr()), t0, te, C.byref(m), tol, h, step_limit,
                                C.c_void_p(acc.data_ptr()), C.c_void_p(rej.data_ptr()),
                                C.c_void_p(stream or 0)))
    return Y, acc, rej
