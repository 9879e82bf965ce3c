"""GPU parity and properties at the clinical geometries.

  * The stored system matrix of the clinical scanner (64 x 736 detector,
    in-plane flying focal spot, 3000 SSM views; config 1 -- the same matrix
    serves configs 2 and 4 -- and the 256 x 256 image of config 3): a
    deterministic sample of columns spread over the field of view, every SSM
    view, bit-exact against the oracle's column-subset build.
  * The library's default schedule (tile 0, slice spacing 0: the spacing
    whose groups have overlapping footprints in at most 1 % of the columns)
    on the clinical cone geometry, where that spacing is below the
    conflict-free one: the volume within 1e-4 relative RMS of the oracle
    replaying the same schedule, and the FP64 cost never rises over sweeps.
  * Full-size runs (configs 1, 2, 4; sizes the CPU oracle cannot run):
    a sweep lowers the FP64 cost; after a resync the running FP32 error
    sinogram equals y - A x of the FP64 projection; the FP16 matrix with an
    FP64 resync every 2 sweeps (config 2) against the FP32 run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(ct):
    ct.init(0)
    return ct


def _sample_columns(g, n, seed):
    """n columns inside the reconstructed field, spread over it: the centre,
    the four axis ends, and seeded random ones."""
    rng = np.random.default_rng(seed)
    cx, cy = int(g.center_voxel_x), int(g.center_voxel_y)
    fixed = [cx + g.nx * cy, 0 + g.nx * cy, (g.nx - 1) + g.nx * cy, cx, cx + g.nx * (g.ny - 1),
             0, g.nx * g.ny - 1]
    rest = rng.choice(g.nx * g.ny, n - len(fixed), replace=False)
    return np.array(sorted(set(fixed) | set(int(c) for c in rest)), np.int64)


@pytest.mark.parametrize("name", ["config1", "config3"])
def test_clinical_ssm_columns_bit_exact(gpu, make_scan, orc, name):
    s = make_scan(name)
    g = s.geom
    print(f"[clinical-ssm] {name}: {g.nx}x{g.ny}, {g.n_rows}x{g.n_channels}, "
          f"{g.n_views_sm} views", flush=True)
    m = gpu.SSM.build(g)
    cols = _sample_columns(g, 64, 7)
    ptr, idx, val = orc.ssm_cols(g, cols)
    tot = 0
    for t, c in enumerate(cols):
        cp, gi, gv = m.download_cols(int(c), int(c) + 1)
        n = int(ptr[t + 1] - ptr[t])
        assert cp[-1] == n, (name, int(c), int(cp[-1]), n)
        assert np.array_equal(gi, idx[ptr[t]:ptr[t + 1]]), (name, int(c))
        assert np.array_equal(gv, val[ptr[t]:ptr[t + 1]]), (name, int(c))
        tot += n
    print(f"[clinical-ssm] {name}: {len(cols)} columns, {tot} entries bit-exact", flush=True)
    assert tot > 0
    m.free()


def test_default_schedule_matches_oracle_on_clinical_cone(gpu, make_scan, orc):
    s = make_scan("clin_crop")
    g = s.geom
    raw, w = s.synthesize(use_gpu=False)
    m = gpu.SSM.build(g)
    host = orc.ssm_build(g)
    sp_free = m.group_spacing(s.slice_index)
    sp_default = m.group_spacing(s.slice_index, max_conflict_fraction=0.01)
    print(f"[default] clin_crop spacing conflict-free {sp_free}, default {sp_default}",
          flush=True)
    assert sp_default < sp_free
    r = gpu.Recon(m, raw, w, s.slice_index, s.opts(tile=0, group_spacing=0))
    pen, lam, delta = s.run.penalty, s.run.lambda_, s.run.delta
    costs = [r.cost()]
    for _ in range(3):
        st = r.iterate(1)
        costs.append(r.cost())
    assert st.group_spacing == sp_default and st.tile == 4
    print(f"[default] costs {costs}", flush=True)
    assert all(b <= a for a, b in zip(costs, costs[1:])), costs
    xg = r.volume().astype(np.float64)
    xo, _, nupd = orc.icd(g, host, raw, w, s.slice_index, pen, lam, delta, st.group_spacing,
                          3, tile=st.tile, surrogate=True)
    rms = np.sqrt(np.mean((xg - xo) ** 2)) / np.sqrt(np.mean(xo ** 2))
    print(f"[default] rel RMS vs oracle {rms:.2e}", flush=True)
    assert rms <= 1e-4
    c_orc = orc.cost(g, host, raw, w, s.slice_index, pen, lam, delta, xg)
    assert abs(costs[-1] - c_orc) <= 1e-10 * abs(c_orc)
    r.free()


def _fullsize(gpu, make_scan, name, fp16=False, sweeps=1, resync=0, free_ssm=True):
    s = make_scan(name)
    g = s.geom
    raw, w = s.synthesize(use_gpu=True)
    m = gpu.SSM.build(g)
    if fp16:
        m.to_fp16()
    r = gpu.Recon(m, raw, w, s.slice_index, s.opts(resync_every=resync))
    c0 = r.cost()
    st = r.iterate(sweeps)
    return s, m, r, raw, w, c0, st


@pytest.mark.parametrize("name", ["config1", "config4"])
def test_fullsize_sweep_lowers_cost_and_resync_is_exact(gpu, make_scan, name):
    s, m, r, raw, w, c0, st = _fullsize(gpu, make_scan, name)
    c1 = r.cost()
    print(f"[fullsize] {name}: {st.voxel_updates} updates in {st.sweep_ms:.0f} ms, "
          f"cost {c0:.6e} -> {c1:.6e}, tile {st.tile}, spacing {st.group_spacing}",
          flush=True)
    assert c1 < c0
    run = r.cost_running()
    print(f"[fullsize] {name}: running cost {run:.10e} vs FP64 {c1:.10e} "
          f"(rel {abs(run - c1) / c1:.2e})", flush=True)
    r.resync()
    run2 = r.cost_running()
    assert abs(run2 - c1) <= 1e-6 * c1, (run2, c1)
    if name == "config1":
        # e = y - A x of the FP64 projection of the current volume, rounded
        e = r.error().astype(np.float64)
        y = m.forward_project(s.slice_index, r.volume().astype(np.float64))
        ref = raw.astype(np.float64) - y
        err = np.abs(e - ref).max()
        print(f"[fullsize] {name}: resync max |e - (y - Ax)| {err:.3e}", flush=True)
        assert err <= 1e-6 * np.abs(ref).max()
    r.free()
    m.free()


def test_config2_fp16_resync_against_fp32_and_fp64(gpu, make_scan):
    """Config 2's precision split (FP16 matrix with per-column scales, FP64
    resync every 2 sweeps) against the FP32 plain run of the same data and
    against the FP64 state (x, e and the sums in FP64: the GPU mode that
    matches the CPU FP64 oracle to ~1e-15 in tests/test_gpu_fp64.py, standing
    in for it at this size), all in the same order over 4 sweeps: volume
    max / RMS error and final cost."""
    def run(name, fp16=False, resync=0, fp64=0):
        s = make_scan(name)
        raw, w = s.synthesize(use_gpu=True)
        m = gpu.SSM.build(s.geom)
        if fp16:
            m.to_fp16()
        r = gpu.Recon(m, raw, w, s.slice_index,
                      s.opts(resync_every=resync, state_fp64=fp64))
        c0 = r.cost()
        st = r.iterate(4)
        x = r.volume64() if fp64 else r.volume().astype(np.float64)
        c = r.cost()
        r.free()
        m.free()
        return raw, x, c0, c, st

    raw64, x64, _, c64, _ = run("config1", fp64=1)
    raw32, x32, _, c32, _ = run("config1")
    raw16, x16, c0, c16, st16 = run("config2", fp16=True, resync=2)
    assert np.array_equal(raw32, raw16) and np.array_equal(raw32, raw64)
    hu = 1.93e-4
    res = {}
    for tag, x, c in (("FP32 plain", x32, c32), ("FP16 + resync", x16, c16)):
        d = x - x64
        res[tag] = (np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(x64 ** 2)), np.abs(d).max(),
                    (c - c64) / c64)
        print(f"[config2] {tag} vs FP64 state after 4 sweeps: rel RMS {res[tag][0]:.3e}, "
              f"max |dx| {res[tag][1]:.3e} /cm ({res[tag][1] / hu:.2f} HU), cost rel "
              f"{res[tag][2]:+.2e}", flush=True)
    d = x16 - x32
    rms = np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(x32 ** 2))
    mx = np.abs(d).max()
    print(f"[config2] FP16+resync vs FP32 after 4 sweeps: rel RMS {rms:.3e}, max "
          f"|dx| {mx:.3e} /cm ({mx / hu:.2f} HU), cost {c16:.8e} vs {c32:.8e} "
          f"(rel {(c16 - c32) / c32:+.2e}); resync {st16.resync_ms:.0f} ms", flush=True)
    assert c16 < c0
    assert res["FP32 plain"][0] <= 1e-4          # the FP32 state's bound
    assert abs(res["FP32 plain"][2]) <= 1e-5
    assert rms <= 1e-2 and res["FP16 + resync"][0] <= 1e-2
    assert abs(c16 - c32) <= 1e-3 * c32
