"""Batched solves over the C ABI (include/rkx.h, rkx_batch_*).

Many instances of uneven size advance together, one attempted step each per
batched step, each with its own step size and accept/reject.  Instances can
be added and removed between runs (continuous batching); an instance's
result is bitwise the same alone or in a mix.
"""
import ctypes as C

import numpy as np

from ._lib import RKX_CTRL_DEVICE, Stats, check, lib
from .methods import method as _method
from .solver import CTRLS

STORAGE = {"auto": 0, "low": 1, "full": 2}


class BatchOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("ctrl", C.c_int), ("tile_waves", C.c_int),
                ("tile_rows", C.c_int), ("stream", C.c_void_p)]


class Batch:
    """tile_waves / tile_rows: the fused kernel's strip width (wavefronts per
    component) and band height for every instance (0: chosen per size;
    rk_amd.autotune.batch_tiles picks them by timing the batch's own work)."""

    def __init__(self, hbm_budget=0, ctrl="device", tile_waves=0, tile_rows=0, stream=None):
        o = BatchOpts()
        lib().rkx_batch_opts_default(C.byref(o))
        o.hbm_budget = int(hbm_budget)
        o.ctrl = CTRLS[ctrl] if isinstance(ctrl, str) else ctrl
        o.tile_waves = tile_waves
        o.tile_rows = tile_rows
        o.stream = stream or 0
        h = C.c_void_p()
        check(lib().rkx_batch_create(C.byref(h), C.byref(o)))
        self._h = h
        self._n = {}

    def close(self):
        if self._h:
            lib().rkx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def footprint(problem, meth, storage="auto"):
        m = _method(meth) if isinstance(meth, str) else meth
        return lib().rkx_batch_footprint(C.byref(problem.desc), C.byref(m), STORAGE[storage])

    def add(self, problem, meth, t0, te, y0, tol, h0=0.0, step_limit=0, storage="auto"):
        """Returns the instance id, or None when the HBM budget is exhausted."""
        m = _method(meth) if isinstance(meth, str) else meth
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        rc = lib().rkx_batch_add(self._h, C.byref(problem.desc), C.byref(m), t0, te,
                                 y0.ctypes.data_as(C.POINTER(C.c_double)), tol, h0, step_limit,
                                 STORAGE[storage])
        if rc == -4:
            return None
        check(rc if rc < 0 else 0)
        self._n[rc] = problem.n
        return rc

    def run(self, max_steps):
        """Advance up to max_steps batched steps; returns the number still active."""
        a = C.c_int()
        check(lib().rkx_batch_run(self._h, max_steps, C.byref(a)))
        return a.value

    def status(self, i):
        st = Stats()
        d = C.c_int()
        check(lib().rkx_batch_status(self._h, i, C.byref(st), C.byref(d)))
        return st, bool(d.value)

    def result(self, i):
        y = np.empty(self._n[i])
        check(lib().rkx_batch_result(self._h, i, y.ctypes.data_as(C.POINTER(C.c_double))))
        return y

    def tile(self, i):
        """(strip waves, band rows) of instance i's row-sweep tile now ((0, 0)
        without the row sweep)."""
        w, r = C.c_int(), C.c_int()
        check(lib().rkx_batch_tile(self._h, i, C.byref(w), C.byref(r)))
        return w.value, r.value

    def remove(self, i):
        check(lib().rkx_batch_remove(self._h, i))
        self._n.pop(i, None)

    @property
    def bytes(self):
        return lib().rkx_batch_bytes(self._h)
