"""The 26-neighbour penalty on the GPU against the oracle (the neighbourhood
through the slice rotation, pinned in tests/test_oracle_penalty26.py):
penalty steps of single voxels (FP64 volume: the same FP64 arithmetic,
1e-14; the FP32 volume the sweep holds: the oracle's step on the rounded
volume), the cost (1e-10), and ICD sweeps in the same order (FP32 state
1e-4 relative RMS, FP64 state 1e-9), with the cost never rising."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(ct):
    ct.init(0)
    return ct


@pytest.mark.parametrize("name", ["tiny_ffs", "config0"])
def test_penalty26_steps_match_oracle(gpu, make_scan, orc, name):
    s = make_scan(name)
    g = s.geom
    rng = np.random.default_rng(21)
    x = rng.uniform(0, 0.3, g.nvox)
    n = 400
    q = rng.integers(0, g.nvox, n)
    nxy = g.nx * g.ny
    q[:6] = [0, g.nx - 1, nxy - 1, nxy * (g.nz - 1), g.nvox - 1, nxy + g.nx * 5 + 7]
    alpha = rng.uniform(0, 40, n)
    alpha[7] = 0.0
    beta = rng.normal(0, 3, n)
    x32 = x.astype(np.float32).astype(np.float64)
    for pen in (0, 1):
        o = s.opts(penalty=pen, lam=0.7, delta=0.02, neighbours=26)
        du = gpu.penalty_steps(g, o, x, q, alpha, beta, slice_index=s.slice_index)
        du32 = gpu.penalty_steps(g, o, x, q, alpha, beta, volume_fp32=True,
                                 slice_index=s.slice_index)
        for i in range(n):
            ref = orc.penalty_step(g, s.slice_index, 26, pen, int(q[i]), alpha[i], beta[i],
                                   0.7, 0.02, x)
            ref32 = orc.penalty_step(g, s.slice_index, 26, pen, int(q[i]), alpha[i], beta[i],
                                     0.7, 0.02, x32)
            assert abs(du[i] - ref) <= 1e-14 * max(1.0, abs(ref)), (pen, i, du[i], ref)
            assert abs(du32[i] - ref32) <= 1e-14 * max(1.0, abs(ref32)), (pen, i)


@pytest.mark.parametrize("name,tile,fp64,n_iter", [
    ("tiny_ffs", 4, 0, 3), ("tiny_ffs", 4, 1, 3), ("tiny", 1024, 0, 2),
    ("clin_crop", 0, 0, 2), ("clin_crop", 0, 1, 1)])
def test_icd26_matches_oracle_same_order(gpu, make_scan, orc, name, tile, fp64, n_iter):
    s = make_scan(name)
    g = s.geom
    raw, w = s.synthesize(use_gpu=False)
    m = gpu.SSM.build(g)
    host = orc.ssm_build(g)
    pen, lam, delta = s.run.penalty, s.run.lambda_, s.run.delta
    r = gpu.Recon(m, raw, w, s.slice_index, s.opts(tile=tile, group_spacing=0, neighbours=26,
                                                    state_fp64=fp64))
    costs = [r.cost()]
    for _ in range(n_iter):
        st = r.iterate(1)
        costs.append(r.cost())
    xg = r.volume64()
    xo, _, nupd = orc.icd(g, host, raw, w, s.slice_index, pen, lam, delta, st.group_spacing,
                          n_iter, tile=st.tile, surrogate=True, neighbours=26)
    rms = np.sqrt(np.mean((xg - xo) ** 2)) / np.sqrt(np.mean(xo ** 2))
    c_orc = orc.cost(g, host, raw, w, s.slice_index, pen, lam, delta, xg, neighbours=26)
    print(f"[icd26] {name} tile {st.tile} spacing {st.group_spacing} fp64 {fp64}: rel RMS "
          f"{rms:.2e}, cost {costs[-1]:.12e} vs oracle {c_orc:.12e}", flush=True)
    assert rms <= (1e-9 if fp64 else 1e-4)
    assert abs(costs[-1] - c_orc) <= 1e-10 * c_orc
    assert all(b <= a * (1 + 1e-12) for a, b in zip(costs, costs[1:])), costs
    r.free()
    m.free()


def test_cost26_without_context_matches_oracle(gpu, make_scan, orc):
    s = make_scan("tiny_ffs")
    g = s.geom
    raw, w = s.synthesize(use_gpu=False)
    m = gpu.SSM.build(g)
    x = s.phantom().astype(np.float32) * 0.9
    for pen in (0, 1):
        o = s.opts(penalty=pen, neighbours=26)
        c = gpu.cost(m, raw, w, s.slice_index, o, x)
        ref = orc.cost(g, orc.ssm_build(g), raw, w, s.slice_index, pen, s.run.lambda_,
                       s.run.delta, x.astype(np.float64), neighbours=26)
        assert abs(c - ref) <= 1e-10 * abs(ref), (pen, c, ref)
        c8 = gpu.cost(m, raw, w, s.slice_index, s.opts(penalty=pen), x)
        assert c != c8
