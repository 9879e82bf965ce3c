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


def test_shipped_network_gains_over_plain_at_equal_time(ct, make_scan):
    """The shipped network (ctbb_icd_amd/data/unet_config4.npz, trained on
    config4_slab phantoms 1000-1023: one-sweep image -> 50 sweeps from the
    FBP image) on held-out phantom 2001 of the same scan model.  Its start is
    accepted by the cost check and costs less than the one-sweep image it was
    made from; after equal wall time -- the plain run from zero gets every
    sweep that fits in the warm run's time, the network, rotations and
    projection included -- the warm run's FP64 cost is below the plain
    run's."""
    import time
    from pathlib import Path

    from ctbb_icd_amd import warmstart as ws
    ct.init(0)
    seed = 2001
    s = make_scan("config4_slab", synthetic_seed=seed)
    g = s.geom
    net, meta = ws.load_params(Path(ct.__file__).parent / "data" / "unet_config4.npz", 8,
                               g.ny, g.nx)
    assert meta["train_config"] == "config4_slab"
    assert not (meta["seeds"][0] <= seed < meta["seeds"][1])     # held out
    ssm = ct.SSM.build(g)
    raw, w = s.synthesize(use_gpu=True)
    n_after = 3

    # warm: one plain sweep, the network, the cost check, then n_after sweeps
    r = ct.Recon(ssm, raw, w, s.slice_index, s.opts())
    r.iterate(1)                                   # untimed: first-sweep set-up
    r.free()
    r = ct.Recon(ssm, raw, w, s.slice_index, s.opts())
    t0 = time.perf_counter()
    r.iterate(1)
    x1 = r.volume()
    c1 = r.cost_running()
    pred = ws.to_rotating(s, ws.predict_fixed(net, ws.to_fixed(s, x1)))
    r.set_volume(pred)
    c_start = r.cost_running()
    r.iterate(n_after)
    t_warm = time.perf_counter() - t0
    c_warm = r.cost()
    r.free()
    assert c_start < c1, (c_start, c1)            # accepted, and lower

    # plain from zero for the same wall time
    r = ct.Recon(ssm, raw, w, s.slice_index, s.opts())
    t0 = time.perf_counter()
    n_plain = 0
    while time.perf_counter() - t0 < t_warm:
        r.iterate(1)
        n_plain += 1
    t_plain = time.perf_counter() - t0
    c_plain = r.cost()
    r.free()
    ssm.free()
    net.free()
    print(f"[gain] seed {seed}: one-sweep cost {c1:.8e}, learned start {c_start:.8e}; "
          f"warm 1+{n_after} sweeps in {t_warm:.2f} s -> {c_warm:.8e}; plain {n_plain} "
          f"sweeps in {t_plain:.2f} s -> {c_plain:.8e} (rel gap "
          f"{(c_plain - c_warm) / c_warm:.2e})", flush=True)
    assert n_plain >= 1 + n_after
    assert c_warm < c_plain


def test_config3_pipeline_checks_the_start_and_reaches_target(ct, make_scan):
    """The pipeline at the reduced-size learned-start scan (config3: the
    clinical geometry cut to 256 x 256 x 32, 1.4 mm pixels): plain sweeps
    lower the cost; a network whose start is far off (+50 mu_water) is
    rejected by the cost check and the run continues from the plain image,
    reaching the target; the untrained network (the identity plus the
    de-rotate / rotate resampling) is judged by the same check and its run
    reaches the target too."""
    from ctbb_icd_amd import warmstart as ws
    ct.init(0)
    s = make_scan("config3")
    g = s.geom
    assert (g.nx, g.ny, g.nz) == (256, 256, 32)
    ssm = ct.SSM.build(g)
    raw, w = s.synthesize(use_gpu=True)
    r = ct.Recon(ssm, raw, w, s.slice_index, s.opts())
    costs = []
    for _ in range(8):
        r.iterate(1)
        costs.append(r.cost())
    r.free()
    assert all(b < a for a, b in zip(costs, costs[1:])), costs
    target = costs[-1] * (1 + 1e-3)
    bad = ct.UNet(3, 8, 8, g.ny, g.nx, seed=2)
    p = bad.params()
    p[-1] = 50.0
    bad.set_params(p)
    run = ws.warm_run(s, ssm, raw, w, bad, 1, target, 40)
    assert run.accepted is False
    assert run.costs[0] == pytest.approx(costs[0], rel=1e-9)
    assert run.reached
    ident = ct.UNet(3, 8, 8, g.ny, g.nx, seed=2)
    run2 = ws.warm_run(s, ssm, raw, w, ident, 1, target, 40)
    assert run2.accepted in (True, False)
    assert run2.reached
    print(f"[config3] plain costs {costs[0]:.6e} .. {costs[-1]:.6e}; bad start rejected, "
          f"{run.iterations} its; identity start "
          f"{'accepted' if run2.accepted else 'rejected'}, {run2.iterations} its", flush=True)
    bad.free()
    ident.free()
    ssm.free()


def test_threaded_pairs_match_one_at_a_time(ct, tmp_path):
    """generate_pairs with workers > 1 (several phantoms at once, one host
    thread and one reconstruction each, sharing a matrix whose ray-major
    copy the first reconstructions build concurrently) returns the pairs of
    one phantom at a time, bit for bit and in seed order."""
    from ctbb_icd_amd import configs, warmstart as ws
    ct.init(0)

    def make(seed):
        return ct.Scan(configs.write_prm("tiny_ffs", tmp_path / f"p{seed}.yaml",
                                         synthetic_seed=seed))

    g = make(1).geom
    seeds = [11, 12, 13, 14, 15]
    ssm1 = ct.SSM.build(g)
    serial = ws.generate_pairs(make, seeds, ssm1, 1, 3, lambda m: None, "fbp", workers=1)
    ssm1.free()
    ssm3 = ct.SSM.build(g)
    threaded = ws.generate_pairs(make, seeds, ssm3, 1, 3, lambda m: None, "fbp", workers=3)
    ssm3.free()
    assert [p.seed for p in threaded] == seeds
    for a, b in zip(serial, threaded):
        assert np.array_equal(a.early, b.early) and np.array_equal(a.target, b.target)
    assert not np.array_equal(serial[0].target, serial[1].target)
