"""Batched solves on the GPU: every instance of a mix gives, bitwise, what it
gives alone (plain rkx_solve), with its own step counts; the reference step
control is kept per instance (bitwise equal to the oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rk():
    import torch
    assert torch.cuda.is_available()
    import rk_amd
    rk_amd.lib()
    return rk_amd


def plain(rk, p, y0, meth, te, tol, h0, ctrl="device", step_limit=0):
    import torch
    y0d = torch.from_numpy(np.ascontiguousarray(y0)).cuda()
    yd = torch.empty_like(y0d)
    torch.cuda.synchronize()
    st, _ = rk.solve(p, meth, 0.0, te, y0d, yd, tol, h0=h0, ctrl=ctrl, step_limit=step_limit)
    torch.cuda.synchronize()
    return yd.cpu().numpy(), st


MIX = [(24, 2e-3), (40, 4.1e-3), (70, 1.3e-3), (33, 2.7e-3), (130, 1.1e-3), (17, 3e-3)]


def test_mix_equals_alone_bitwise(rk):
    ctrl = "device"
    from rk_amd.workloads import bruss_instance
    insts = [bruss_instance(N, a, k, seed=5) for k, (N, a) in enumerate(MIX)]
    with rk.Batch(ctrl=ctrl) as b:
        ids = [b.add(p, "DOPRI54", 0.0, 0.6, y0, 1e-6, h0=1e-3) for p, y0 in insts]
        left = b.run(100000)
        assert left == 0
        for (p, y0), i in zip(insts, ids):
            y_alone, st_alone = plain(rk, p, y0, "DOPRI54", 0.6, 1e-6, 1e-3, ctrl=ctrl)
            st, done = b.status(i)
            assert done
            assert (st.steps_acc, st.steps_rej) == (st_alone.steps_acc, st_alone.steps_rej)
            assert np.array_equal(b.result(i), y_alone)


@pytest.mark.parametrize("meth,grids", [("DOPRI54", (9, 40, 100)), ("VERNER65", (12, 30, 90))])
def test_short_and_tall_grids_share_a_launch(rk, meth, grids):
    """A grid too short for single-reflection ghost rows puts its whole launch
    group (same table and strip width) on the general reflection form; every
    instance still equals its alone run bitwise.  VERNER65 also runs whole
    ring trips, with the requested band height rounded up to fill them."""
    from rk_amd.workloads import bruss_instance
    insts = [bruss_instance(N, 2e-3, k, seed=3) for k, N in enumerate(grids)]
    with rk.Batch(tile_waves=1, tile_rows=24) as b:
        ids = [b.add(p, meth, 0.0, 0.3, y0, 1e-6, h0=1e-3) for p, y0 in insts]
        assert b.run(100000) == 0
        for (p, y0), i in zip(insts, ids):
            y_alone, st_alone = plain(rk, p, y0, meth, 0.3, 1e-6, 1e-3)
            st, done = b.status(i)
            assert done
            assert (st.steps_acc, st.steps_rej) == (st_alone.steps_acc, st_alone.steps_rej)
            assert np.array_equal(b.result(i), y_alone)


@pytest.mark.parametrize("waves,rows", [(4, 23), (2, 9)])
def test_paired_narrow_strips_in_a_batch(rk, waves, rows):
    """Grids whose columns left over after the full-width strips run as a
    narrow strip of paired bands (N = 40, 60, 260 at 4 waves; 40, 130, 242,
    260 at 2) share launches with grids that have none (130, 200, 242 at 4
    waves; 60, 200 at 2); the instances end at different times, so the batch
    shortens the bands of the ones left while it runs.  Every instance
    equals its alone run bitwise."""
    from rk_amd.workloads import bruss_instance
    grids = [(60, 0.3), (130, 0.5), (200, 0.2), (242, 0.4), (260, 0.6), (40, 0.25)]
    insts = [bruss_instance(N, 2e-3, k, seed=6) for k, (N, _) in enumerate(grids)]
    with rk.Batch(tile_waves=waves, tile_rows=rows) as b:
        ids = [b.add(p, "DOPRI54", 0.0, te, y0, 1e-6, h0=1e-3)
               for (p, y0), (_, te) in zip(insts, grids)]
        assert b.run(100000) == 0
        for (p, y0), (_, te), i in zip(insts, grids, ids):
            y_alone, st_alone = plain(rk, p, y0, "DOPRI54", te, 1e-6, 1e-3)
            st, done = b.status(i)
            assert done
            assert (st.steps_acc, st.steps_rej) == (st_alone.steps_acc, st_alone.steps_rej)
            assert np.array_equal(b.result(i), y_alone)


def test_dopri87_instances_in_a_batch(rk):
    """DOPRI87 instances run their fused kernel in a batch (strip width
    narrowed to the 2 waves it is built for when the options ask for 4) and
    equal their alone runs bitwise."""
    from rk_amd.workloads import bruss_instance
    insts = [bruss_instance(N, a, k, seed=4) for k, (N, a) in enumerate([(24, 2e-3), (70, 3e-3),
                                                                          (150, 1.5e-3)])]
    with rk.Batch(tile_waves=4, tile_rows=40) as b:
        ids = [b.add(p, "DOPRI87", 0.0, 0.3, y0, 1e-8, h0=1e-4) for p, y0 in insts]
        assert b.run(100000) == 0
        for (p, y0), i in zip(insts, ids):
            y_alone, st_alone = plain(rk, p, y0, "DOPRI87", 0.3, 1e-8, 1e-4)
            st, done = b.status(i)
            assert done and st.resident_bytes == 2 * p.n * 8
            assert (st.steps_acc, st.steps_rej) == (st_alone.steps_acc, st_alone.steps_rej)
            assert np.array_equal(b.result(i), y_alone)


def test_batch_matches_oracle(rk, orc):
    """Device step control with the pow service: each instance reproduces the
    oracle (seq D restated) bitwise."""
    from rk_amd.workloads import bruss_instance
    insts = [bruss_instance(N, a, k, seed=9) for k, (N, a) in enumerate(MIX[:3])]
    with rk.Batch() as b:
        ids = [b.add(p, "DOPRI54", 0.0, 0.5, y0, 1e-6, h0=1e-3) for p, y0 in insts]
        b.run(100000)
        for (p, y0), i in zip(insts, ids):
            po = orc.make_ode("bruss2d", p.N, p.alpha)
            yo, so, _ = orc.solve(po, "DOPRI54", 0.0, 0.5, y0, 1e-6, h0=1e-3)
            st, _ = b.status(i)
            assert (st.steps_acc, st.steps_rej) == (so["acc"], so["rej"])
            assert np.array_equal(b.result(i), yo)


def test_continuous_batching_and_budget(rk):
    """Instances join while others run, the HBM budget refuses what does not
    fit, and removal frees room; results still equal the alone runs."""
    from rk_amd.workloads import bruss_instance
    p1, y1 = bruss_instance(64, 2e-3, 1)
    p2, y2 = bruss_instance(48, 3e-3, 2)
    p3, y3 = bruss_instance(64, 1e-3, 3)
    fp = rk.Batch.footprint(p1, "DOPRI54")
    assert fp == 2 * p1.n * 8
    with rk.Batch(hbm_budget=fp + rk.Batch.footprint(p2, "DOPRI54")) as b:
        i1 = b.add(p1, "DOPRI54", 0.0, 1.0, y1, 1e-6, h0=1e-3)
        b.run(40)                                  # i1 is under way
        i2 = b.add(p2, "DOPRI54", 0.0, 1.0, y2, 1e-6, h0=1e-3)
        assert i2 is not None
        assert b.add(p3, "DOPRI54", 0.0, 1.0, y3, 1e-6, h0=1e-3) is None   # over budget
        b.run(100000)
        r1, r2 = b.result(i1), b.result(i2)
        b.remove(i1)
        i3 = b.add(p3, "DOPRI54", 0.0, 1.0, y3, 1e-6, h0=1e-3)
        assert i3 is not None
        b.run(100000)
        r3 = b.result(i3)
    for (p, y0), r in (((p1, y1), r1), ((p2, y2), r2), ((p3, y3), r3)):
        ya, _ = plain(rk, p, y0, "DOPRI54", 1.0, 1e-6, 1e-3)
        assert np.array_equal(r, ya)


def test_full_storage_and_mixed_kinds(rk):
    """Full-storage (D form) instances -- an LV chain, a VERNER65 grid, a
    DOPRI87 grid without a fused kernel, LV pairs -- in the same batch as fused
    ones, among them LV pairs through their own per-pair kernel."""
    from rk_amd.workloads import bruss_instance, chain_instance
    cases = [(*bruss_instance(30, 2e-3, 0), "DOPRI54", "low"),
             (*bruss_instance(30, 2e-3, 0), "DOPRI54", "full"),
             (*chain_instance(3000, 0), "DOPRI54", "auto"),
             (*chain_instance(3000, 0), "DOPRI54", "full"),
             (*chain_instance(2100, 1), "VERNER65", "low"),
             (*bruss_instance(20, 2e-3, 4), "VERNER65", "auto"),
             (*bruss_instance(12, 2e-3, 5), "DOPRI87", "auto")]
    lv = rk.LV(5000, 1.3, 0.7, 0.45, 0.9)      # LV pairs: the per-pair fused kernel (low)
    lv_y0 = 0.5 + np.random.default_rng(9).random(lv.n)
    cases += [(lv, lv_y0, "DOPRI54", "auto"), (lv, lv_y0, "VERNER65", "full"),
              (lv, lv_y0, "DOPRI87", "low")]
    with rk.Batch() as b:
        ids = [b.add(p, m, 0.0, 0.3, y0, 1e-7, h0=1e-3, storage=sto) for p, y0, m, sto in cases]
        assert b.run(100000) == 0
        for (p, y0, m, sto), i in zip(cases, ids):
            ya, sa = plain(rk, p, y0, m, 0.3, 1e-7, 1e-3)
            st, _ = b.status(i)
            assert (st.steps_acc, st.steps_rej) == (sa.steps_acc, sa.steps_rej)
            assert np.array_equal(b.result(i), ya), (m, sto)


def test_step_limit_per_instance(rk):
    from rk_amd.workloads import bruss_instance
    p, y0 = bruss_instance(40, 2e-3, 0)
    with rk.Batch() as b:
        i = b.add(p, "DOPRI54", 0.0, 4.0, y0, 1e-6, h0=1e-3, step_limit=13)
        j = b.add(p, "DOPRI54", 0.0, 4.0, y0, 1e-6, h0=1e-3, step_limit=29)
        b.run(1000)
        assert b.status(i)[0].steps == 13 and b.status(j)[0].steps == 29
