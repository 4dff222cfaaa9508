"""rk_amd: MI355X-native embedded Runge-Kutta solver (Python side of the C ABI).

The compute path is rk_amd/librkx.so (hand-written HIP kernels for gfx950 plus
the C ABI declared in include/rkx.h).  This package is a thin ctypes layer that
mirrors the reference's three plugin interfaces:

  * problems  -- include/ode.h:29-36   (ode_size, ode_start, ode_eval_rng, ode_acc_dist)
  * methods   -- include/methods.h:92  (rk_method_t: A, b, b_hat, c, s, ord)
  * solver    -- include/solver.h:47   (solver(t0, te, y0, y, tol))

There is no CPU fallback: importing a function that needs the library raises
if librkx.so is missing or no HIP device is present.
"""
from ._lib import (RKX_BRUSS2D, RKX_LV, RKX_LV_CHAIN, RKX_IMPL_D, RKX_IMPL_PIPED,
                   RKX_IMPL_PIPEDLS, RKX_IMPL_A, RKX_IMPL_E, RKX_IMPL_F, RKX_CTRL_DEVICE,
                   RKX_CTRL_HOST, RkxError, lib, lib_path)
from .problems import Bruss2D, LV, LVChain, Problem
from .methods import method, method_table, METHODS
from .solver import solve, solve_host, solver, Stats
from .batch import Batch

__all__ = [
    "RKX_BRUSS2D", "RKX_LV", "RKX_LV_CHAIN", "RKX_IMPL_D", "RKX_IMPL_PIPED",
    "RKX_IMPL_PIPEDLS", "RKX_IMPL_A", "RKX_IMPL_E", "RKX_IMPL_F", "RKX_CTRL_DEVICE", "RKX_CTRL_HOST", "RkxError", "lib", "lib_path",
    "Bruss2D", "LV", "LVChain", "Problem", "method", "method_table", "METHODS",
    "solve", "solve_host", "solver", "Stats", "Batch",
]
