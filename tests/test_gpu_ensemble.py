"""LV ensemble on the GPU: each system equals a plain single solve of that
system bitwise, identical systems reproduce the reference's LV known answer
exactly, and every system of the seeded C4-A configuration (262 144 systems,
VERNER65, tol 1e-8, t in [0, 20]) equals the oracle's seq D solve of that
system: identical accepted/rejected counts and bitwise the same y."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def rk():
    import torch
    assert torch.cuda.is_available()
    import rk_amd
    rk_amd.lib()
    return rk_amd


def test_ensemble_equals_plain_solves(rk):
    import torch
    from rk_amd.ensemble import lv_ensemble, lv_ensemble_inputs
    params, y0 = lv_ensemble_inputs(300, seed=7)
    Y, acc, rej = lv_ensemble(params, y0, 0.0, 6.0, "VERNER65", 1e-8)
    Y, acc, rej = Y.cpu().numpy(), acc.cpu().numpy(), rej.cpu().numpy()
    for k in (0, 1, 17, 123, 299):
        p = rk.LV(1, *params[k])
        y0d = torch.from_numpy(y0[k].copy()).cuda()
        yd = torch.empty_like(y0d)
        torch.cuda.synchronize()
        st, _ = rk.solve(p, "VERNER65", 0.0, 6.0, y0d, yd, 1e-8, h0=1e-7, ctrl="device")
        assert (st.steps_acc, st.steps_rej) == (acc[k], rej[k])
        assert np.array_equal(yd.cpu().numpy(), Y[k])


def test_identical_systems_reproduce_reference_kat(rk):
    """10000 identical pairs share one trajectory in the reference
    (src/problems/lv.c:31-37 with LV_COUNT=10000); every system of the
    ensemble lands on the reference's step counts and footprint exactly."""
    case = [c for c in json.load(open(os.path.join(HERE, "golden", "kat.json")))["cases"]
            if c["name"] == "lv10000_verner65_t20"][0]
    from rk_amd.ensemble import lv_ensemble
    n = 10000
    params = np.tile([0.1, 0.2, 0.4, 0.2], (n, 1))
    y0 = np.tile([0.2, 1.1], (n, 1))
    Y, acc, rej = lv_ensemble(params, y0, 0.0, 20.0, "VERNER65", 1e-8, h0=1e-7)
    Y, acc, rej = Y.cpu().numpy(), acc.cpu().numpy(), rej.cpu().numpy()
    assert (acc == case["acc"]).all() and (rej == case["rej"]).all()
    flat = Y.reshape(-1)
    got = [float(flat[0]), float(flat[-1]), float(flat.min()), float(flat.max())]
    assert got == [float(x) for x in case["footprint"]]


def test_c4a_every_system_against_oracle(rk, orc):
    """The full C4-A ensemble against the oracle, system by system."""
    from rk_amd.ensemble import initial_step, lv_ensemble, lv_ensemble_inputs
    params, y0 = lv_ensemble_inputs(262144, seed=1)
    h0 = initial_step(20.0, 0.0, 1e-8)
    Y, acc, rej = lv_ensemble(params, y0, 0.0, 20.0, "VERNER65", 1e-8)
    Y, acc, rej = Y.cpu().numpy(), acc.cpu().numpy(), rej.cpu().numpy()
    print("GPU done; oracle over 262144 systems", flush=True)
    yo, ao, ro = orc.lv_ensemble(params, y0, 0.0, 20.0, "VERNER65", 1e-8, h0)
    assert np.array_equal(acc, ao) and np.array_equal(rej, ro)
    assert np.array_equal(Y.view(np.uint64), yo.view(np.uint64))
    print("systems: %d, attempted steps: mean %.1f, max %d"
          % (acc.size, float(np.mean(acc + rej)), int(np.max(acc + rej))))
