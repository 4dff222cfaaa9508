"""Problem descriptors mirroring include/ode.h:29-36.

Each object carries ode_size, ode_start(t) and ode_acc_dist() like the
reference's problem plugins (src/problems/bruss2d-mix.c, src/problems/lv.c);
the right-hand side itself runs on the device (eval_rng_device).
"""
import ctypes as C

import numpy as np

from ._lib import Ode, check, lib

LV_CONST = (0.1, 0.2, 0.4, 0.2)   # c0, d0, d1, c1 (src/problems/lv.c:31-34)


class Problem:
    def __init__(self):
        self.desc = Ode()

    @property
    def ode_size(self):
        return self.desc.n

    @property
    def n(self):
        return self.desc.n

    def ode_acc_dist(self):
        d = lib().rkx_ode_acc_dist(C.byref(self.desc))
        if d < 0:
            check(d)
        return d

    def ode_start(self, t=0.0):
        y0 = np.empty(self.n, dtype=np.float64)
        check(lib().rkx_ode_start(C.byref(self.desc), t, y0.ctypes.data_as(C.POINTER(C.c_double))))
        return y0

    def ode_start_seeded(self, seed, instance=0, amp=0.01):
        y0 = np.empty(self.n, dtype=np.float64)
        check(lib().rkx_ode_start_seeded(C.byref(self.desc), seed, instance, amp,
                                         y0.ctypes.data_as(C.POINTER(C.c_double))))
        return y0

    def eval_rng_device(self, y_dev, f_dev, i=0, j=None, t=0.0, stream=None):
        """ode_eval_rng(i, j, t, y, f) on device pointers (ints)."""
        j = self.n - 1 if j is None else j
        check(lib().rkx_eval_rng_device(C.byref(self.desc), i, j, t, C.c_void_p(y_dev),
                                        C.c_void_p(f_dev), C.c_void_p(stream or 0)))


class Bruss2D(Problem):
    """BRUSS2D-MIX on an N x N grid (src/problems/bruss2d-mix.c)."""

    def __init__(self, N, alpha=2e-3):
        super().__init__()
        check(lib().rkx_ode_bruss2d(C.byref(self.desc), N, alpha))
        self.N, self.alpha = N, alpha


class LV(Problem):
    """`pairs` uncoupled Lotka-Volterra pairs (src/problems/lv.c)."""

    def __init__(self, pairs, c0=LV_CONST[0], d0=LV_CONST[1], d1=LV_CONST[2], c1=LV_CONST[3]):
        super().__init__()
        check(lib().rkx_ode_lv(C.byref(self.desc), pairs, c0, d0, d1, c1))
        self.pairs = pairs


class LVChain(Problem):
    """Diffusive periodic predator-prey chain of `cells` cells (no reference counterpart)."""

    def __init__(self, cells, diff=1e-3, c0=LV_CONST[0], d0=LV_CONST[1], d1=LV_CONST[2],
                 c1=LV_CONST[3]):
        super().__init__()
        check(lib().rkx_ode_lv_chain(C.byref(self.desc), cells, c0, d0, d1, c1, diff))
        self.cells, self.diff = cells, diff
