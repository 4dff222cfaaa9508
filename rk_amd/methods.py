"""Method tables mirroring include/methods.h:92-118 (rk_method_t)."""
import ctypes as C

import numpy as np

from ._lib import Method, check, lib

METHODS = ("HEUN2", "RK4", "SSPRK3", "RKF23", "DOPRI54", "VERNER65", "DOPRI87")


def method(name):
    """The rkx_method struct of a named method (src/methods.c:62-724)."""
    m = Method()
    check(lib().rkx_method_get(name.encode(), C.byref(m)))
    return m


def method_table(name):
    """(A, b, b_hat, c, s, ord) -- the out-parameters of rk_method_t."""
    m = method(name)
    s = m.s
    A = np.array([[m.A[i][j] for j in range(s)] for i in range(s)])
    return A, np.array(m.b[:s]), np.array(m.b_hat[:s]), np.array(m.c[:s]), s, m.ord


def has_emb_sol(name):
    """HAS_EMB_SOL(M) (methods.h:95-118)."""
    return bool(method(name).has_emb)
