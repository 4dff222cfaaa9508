"""Autotuners: the tile shape for batches of fused BRUSS2D instances, and the
step form (loop order) of a single solve.

The fused kernel's strip width (1..4 wavefronts per component: 64..256 grid
columns per workgroup) and band height (rows per workgroup) trade wasted
lanes and halo rows against parallelism and occupancy, and the best trade
depends on the mix of grid sizes and on how many workgroups of other
instances share the CUs -- which a per-kernel cost model does not capture.
batch_tiles() therefore measures: it runs the same batched steps of a
sample of the batch under each candidate shape and returns the fastest.  The
sample is stratified by size: its launches then mix tiles of all lengths as
the whole batch does, and on the C3 mix its ranking of the shapes matches the
ranking of full runs, where a sample of the largest instances alone (the ones
that dominate the work) picked a shape 7 % slower over the whole batch.
Every candidate performs identical arithmetic (the tile shape never changes
results), so wall times compare directly.
"""
import time

CANDIDATES = ((1, 64), (2, 64), (2, 96), (2, 128), (2, 160), (2, 192), (4, 64), (4, 128))


def batch_tiles(instances, meth="DOPRI54", t0=0.0, te=4.0, tol=1e-6, h0=1e-3, sample=128,
                steps=32, candidates=CANDIDATES, stratified=True):
    """instances: [(problem, y0), ...] of BRUSS2D problems.  Returns
    ((tile_waves, tile_rows), {(K, R): seconds}) for the fastest candidate.
    The sample is stratified by size (every len/sample-th instance in order
    of size), so that it shares the tiles of its launches out as the whole
    mix does; stratified=False takes the largest instances instead."""
    import torch
    from .batch import Batch
    from .methods import method
    s = method(meth).s
    order = sorted(range(len(instances)), key=lambda k: -instances[k][0].n)
    if stratified and len(order) > sample:
        stride = len(order) / sample
        order = [order[int(j * stride)] for j in range(sample)]
    pick = [instances[k] for k in order[:sample]]
    times = {}
    for K, R in candidates:
        if 64 * K <= 2 * s:
            continue
        with Batch(tile_waves=K, tile_rows=R) as b:
            for p, y0 in pick:
                b.add(p, meth, t0, te, y0, tol, h0=h0)
            b.run(16)                          # tables, graph capture, warm caches
            torch.cuda.synchronize()
            t = time.perf_counter()
            b.run(steps)
            torch.cuda.synchronize()
            times[(K, R)] = time.perf_counter() - t
    best = min(times, key=times.get)
    return best, times


FORMS = ("PipeDls", "D", "A", "E", "F")


def solve_form(problem, meth, y0, tol, t0=0.0, te=1.0, h0=1e-3, window=(20, 120), forms=FORMS):
    """Time the step forms on one problem: the fused low-storage sweep
    (PipeDls; a problem without a fused kernel runs the full-storage loop
    order with the least traffic under that name) and the full-storage loop orders D, A, E and F of the
    reference (src/impl/seq/D.c, A.c, E.c, F.c).  All of them give the same
    steps and results bitwise, so per-step times compare directly; each is
    the difference of two solves capped at window[0] and window[1] attempted
    steps (set-up excluded).  y0: device tensor.  Returns (fastest form,
    {form: seconds per attempted step})."""
    import torch
    from .solver import solve
    y = torch.empty_like(y0)
    times = {}
    for f in forms:
        a, _ = solve(problem, meth, t0, te, y0, y, tol, h0=h0, impl=f, step_limit=window[0])
        b, _ = solve(problem, meth, t0, te, y0, y, tol, h0=h0, impl=f, step_limit=window[1])
        if b.steps > a.steps:
            times[f] = (b.kernel_time - a.kernel_time) / (b.steps - a.steps)
    if not times:
        raise ValueError("solve_form: the solve ends before window[0] attempted steps")
    return min(times, key=times.get), times
