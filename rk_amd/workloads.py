"""Seeded synthetic workloads of the batched and serving configurations.

The reference has no batched mode; these generators fix the parameters the
workload descriptions leave open, with a counter-based generator so every
value is reproducible:

* mixed BRUSS2D batch: grid N uniform on {32..512} (seed 0), alpha
  log-uniform on [1e-3, 5e-3], u0 = 0.5 + x + 0.01*U[-1,1) noise from
  rkx_ode_start_seeded, v0 = 1 + 5y; DOPRI54, tol 1e-6, t in [0, 4], h0 1e-3.
* LV chains: M uniform on [1e4, 1e6] cells, diffusion 1e-3, periodic, the
  reference LV constants (lv.c:31-34), x0 = 0.2 + 0.1*U, y0 = 1.1 + 0.1*U;
  DOPRI54, tol 1e-6, t in [0, 4], h0 1e-3.
* serving stream: Poisson arrivals at a fixed rate (seeded), 70 % BRUSS2D
  N uniform on [64, 1024], 30 % LV chains M uniform on [1e4, 5e5], tol from
  {1e-4, 1e-6, 1e-8}, t_end uniform on [0.5, 4].
"""
import math

import numpy as np

from .problems import Bruss2D, LVChain


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def mixed_bruss(count=512, seed=0, nmin=32, nmax=512):
    """[(N, alpha), ...] of the mixed BRUSS2D batch."""
    g = _rng(seed)
    Ns = g.integers(nmin, nmax + 1, size=count)
    la, lb = math.log(1e-3), math.log(5e-3)
    alphas = np.exp(g.uniform(la, lb, size=count))
    return [(int(n), float(a)) for n, a in zip(Ns, alphas)]


def bruss_instance(N, alpha, instance, seed=0, amp=0.01):
    p = Bruss2D(N, alpha)
    return p, p.ode_start_seeded(seed, instance, amp)


def lv_chains(count=64, seed=1, mmin=10_000, mmax=1_000_000):
    g = _rng(seed)
    return [int(m) for m in g.integers(mmin, mmax + 1, size=count)]


def chain_instance(M, instance, seed=1, diff=1e-3):
    p = LVChain(M, diff=diff)
    return p, p.ode_start_seeded(seed, instance, 0.0)


def serving_stream(rate_hz, duration_s=60.0, seed=2, nmax=1024, mmax=500_000):
    """[(arrival_s, kind, size, tol, t_end), ...] sorted by arrival."""
    g = _rng(seed)
    out = []
    t = 0.0
    while True:
        t += g.exponential(1.0 / rate_hz)
        if t >= duration_s:
            break
        if g.random() < 0.7:
            kind, size = "bruss2d", int(g.integers(64, nmax + 1))
        else:
            kind, size = "lvchain", int(g.integers(10_000, mmax + 1))
        tol = float(g.choice([1e-4, 1e-6, 1e-8]))
        te = float(g.uniform(0.5, 4.0))
        out.append((t, kind, size, tol, te))
    return out
