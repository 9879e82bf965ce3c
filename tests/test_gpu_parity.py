"""GPU parity: every device entry point against the CPU oracle, through the
C ABI.  Tolerances (written per test) follow the contract of the hot path:
SSM indices bit-exact and values within 1e-6 relative (FP32), rotated slices
within 1e-6, forward projection (FP64) within 1e-12, ICD volume within 1e-4
relative RMS of the oracle replaying the same update order, cost within 1e-5
relative of the oracle's reference order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SSM_CASES = ["micro", "tiny", "tiny_ffs", "tiny_zffs", "tiny_zphiffs", "config0"]


@pytest.fixture(scope="module")
def gpu(ct):
    ct.init(0)
    return ct


@pytest.fixture(scope="module")
def ssm_pair(gpu, make_scan, orc):
    cache = {}

    def _get(name):
        if name not in cache:
            s = make_scan(name)
            print(f"[ssm] {name}: {s.geom.nx}x{s.geom.ny}, {s.geom.n_views_sm} views", flush=True)
            dev = gpu.SSM.build(s.geom)
            host = orc.ssm_build(s.geom)
            cache[name] = (s, dev, host)
        return cache[name]
    return _get


@pytest.mark.parametrize("name", SSM_CASES)
def test_ssm_bit_exact(ssm_pair, name):
    s, dev, (cp, idx, val) = ssm_pair(name)
    gcp, gidx, gval = dev.download()
    assert gcp[-1] == cp[-1], f"nnz {gcp[-1]} vs oracle {cp[-1]}"
    assert np.array_equal(gcp, cp)
    assert np.array_equal(gidx, idx)
    # values: bit-exact expected; the contract allows 1e-6 relative
    rel = np.abs(gval.astype(np.float64) - val) / np.maximum(np.abs(val), 1e-30)
    assert rel.max() <= 1e-6, rel.max()
    assert np.array_equal(gval, val)


def test_ssm_fp16_values(gpu, make_scan):
    s = make_scan("tiny_ffs")
    m = gpu.SSM.build(s.geom)
    cp, idx, val = m.download()
    m.to_fp16()
    cp2, idx2, val2 = m.download()
    assert np.array_equal(cp, cp2) and np.array_equal(idx, idx2)
    for c in range(len(cp) - 1):
        a, b = val[cp[c]:cp[c + 1]], val2[cp[c]:cp[c + 1]]
        if len(a):
            # FP16 with a per-column scale: error <= half an FP16 ulp of the
            # column maximum mapped to 2^15 (11-bit mantissa) plus subnormals
            assert np.abs(a - b).max() <= np.abs(a).max() * 2.0 ** -11 + 1e-12


def test_forward_projection(gpu, ssm_pair, orc):
    s, dev, host = ssm_pair("tiny_ffs")
    x = np.random.default_rng(0).uniform(0, 0.3, s.geom.nvox)
    y_gpu = dev.forward_project(s.slice_index, x)
    y_cpu = orc.forward_project(s.geom, host, s.slice_index, x)
    assert np.abs(y_gpu - y_cpu).max() <= 1e-12 * np.abs(y_cpu).max()


def test_rotation(gpu, make_scan, orc):
    s = make_scan("config0")
    g = s.geom
    vol = np.random.default_rng(1).uniform(0, 0.3, g.nvox).astype(np.float32)
    for direction in (1, -1):
        a = gpu.rotate_slices(g, s.tube_angles, s.slice_index, vol, direction)
        b = orc.rotate_slices(g, s.tube_angles, s.slice_index, vol, direction)
        assert np.abs(a - b).max() <= 1e-6
        assert np.array_equal(a, b)


def test_penalty_steps(gpu, make_scan, orc):
    s = make_scan("tiny")
    g = s.geom
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 0.3, g.nvox)
    q = rng.integers(0, g.nvox, 500)
    q[:4] = [0, 1, g.nvox - 1, g.nvox - 2]
    alpha = rng.uniform(0, 40, 500)
    alpha[5] = 0.0
    beta = rng.normal(0, 3, 500)
    for pen in (0, 1):
        o = s.opts(penalty=pen, lam=0.7, delta=0.02)
        du = gpu.penalty_steps(g, o, x, q, alpha, beta)
        for i in range(500):
            if pen == 0:
                ref = orc.quadratic(g, int(q[i]), alpha[i], beta[i], 0.7, x)
            else:
                ref = orc.edge_preserving(g, int(q[i]), alpha[i], beta[i], 0.7, 0.02, x)
            assert du[i] == ref or abs(du[i] - ref) <= 1e-14 * max(1.0, abs(ref)), (pen, i)
        # the sweep's instantiation (the volume held in FP32): the reference's
        # step on the FP32-rounded volume
        x32 = x.astype(np.float32).astype(np.float64)
        du_w = gpu.penalty_steps(g, o, x, q, alpha, beta, volume_fp32=True)
        for i in range(500):
            if pen == 0:
                ref = orc.quadratic(g, int(q[i]), alpha[i], beta[i], 0.7, x32)
            else:
                ref = orc.edge_preserving(g, int(q[i]), alpha[i], beta[i], 0.7, 0.02, x32)
            assert du_w[i] == ref or abs(du_w[i] - ref) <= 1e-14 * max(1.0, abs(ref)), \
                (pen, i, du_w[i], ref)


@pytest.mark.parametrize("name", ["tiny", "tiny_ffs", "config0"])
def test_group_spacing_is_conflict_free(ssm_pair, orc, name):
    s, dev, host = ssm_pair(name)
    g = s.geom
    sp = dev.group_spacing(s.slice_index)
    print(f"[spacing] {name}: {sp}", flush=True)
    assert 2 <= sp <= g.nz
    si = s.slice_index
    fov = (g.acquisition_fov / 2) ** 2
    cols = [c for c in range(g.nx * g.ny)
            if ((c % g.nx - g.center_voxel_x) * g.voxel_size_x) ** 2
            + ((c // g.nx - g.center_voxel_y) * g.voxel_size_y) ** 2 < fov]
    shifts = sorted({abs(int(si[k2] - si[k])) for k in range(g.nz)
                     for k2 in range(k + sp, g.nz, sp)})
    shifts = [d for d in shifts if d < g.n_views_sm]
    step = max(1, len(cols) // 400)
    for c in cols[::step]:
        for dv in shifts:
            assert not orc.column_overlaps(g, host, c, dv), (c, dv)


def _icd_case(gpu, ssm_pair, orc, name, n_iter, fp16=False, resync=0, tile=8,
              plain=False):
    s, dev, host = ssm_pair(name)
    g = s.geom
    raw, w = s.synthesize(use_gpu=False)
    m = dev
    if fp16:
        m = gpu.SSM.build(g)
        m.to_fp16()
    sp = dev.group_spacing(s.slice_index)
    opts = s.opts(resync_every=resync, tile=tile, group_spacing=sp,
                  plain_curvature=int(plain))
    r = gpu.Recon(m, raw, w, s.slice_index, opts)
    st = r.iterate(n_iter)
    xg = r.volume().astype(np.float64)
    pen, lam, delta = s.run.penalty, s.run.lambda_, s.run.delta
    xo, _, nupd = orc.icd(g, host, raw, w, s.slice_index, pen, lam, delta, sp, n_iter,
                          tile=tile, surrogate=not plain)
    return s, r, st, xg, xo, nupd, raw, w, host, sp


@pytest.mark.parametrize("name,n_iter,tile,plain", [
    ("tiny", 3, 1024, False), ("tiny", 3, 4, False), ("tiny", 3, 4, True),
    ("tiny_ffs", 3, 1024, False), ("tiny_ffs", 3, 6, False), ("tiny_ffs", 3, 2, True),
    ("tiny_zphiffs", 2, 4, False),
    ("config0", 1, 16, False), ("config0", 1, 8, False), ("config0", 1, 1024, False)])
def test_icd_matches_oracle_same_order(gpu, ssm_pair, orc, name, n_iter, tile, plain):
    """tile 1024: single-column batches (the reference's column order with
    conflict-free slice groups); smaller tiles: multi-column batches with the
    batch surrogate curvature (default) or the plain one."""
    s, r, st, xg, xo, nupd, raw, w, host, sp = _icd_case(gpu, ssm_pair, orc, name,
                                                         n_iter, tile=tile, plain=plain)
    print(f"[icd] {name} tile {tile} plain {plain}: spacing {sp}, {st.voxel_updates} "
          f"updates, {st.sweep_ms:.1f} ms", flush=True)
    assert st.voxel_updates == nupd
    rms = np.sqrt(np.mean((xg - xo) ** 2)) / np.sqrt(np.mean(xo ** 2))
    print(f"[icd] {name}: rel RMS vs oracle {rms:.2e}", flush=True)
    assert rms <= 1e-4
    # FP64 cost of the GPU volume: device vs oracle evaluation of the same x
    g = s.geom
    c_gpu = r.cost()
    c_orc = orc.cost(g, host, raw, w, s.slice_index, s.run.penalty, s.run.lambda_,
                     s.run.delta, r.volume().astype(np.float64))
    assert abs(c_gpu - c_orc) <= 1e-10 * abs(c_orc)
    if not plain:
        # (plain Jacobi curvature may overshoot in weighted batches)
        assert c_gpu < orc.cost(g, host, raw, w, s.slice_index, s.run.penalty,
                                s.run.lambda_, s.run.delta, np.zeros(g.nvox))
    # final cost against the reference's own (z-innermost) order
    xr, _, _ = orc.icd(g, host, raw, w, s.slice_index, s.run.penalty, s.run.lambda_,
                       s.run.delta, 0, n_iter)
    c_ref = orc.cost(g, host, raw, w, s.slice_index, s.run.penalty, s.run.lambda_,
                     s.run.delta, xr)
    print(f"[icd] {name}: cost gpu {c_gpu:.10e} ref-order {c_ref:.10e} "
          f"rel {abs(c_gpu - c_ref) / abs(c_ref):.2e}", flush=True)
    if tile >= max(g.nx, g.ny):
        # single-column batches only permute slices inside a column
        assert abs(c_gpu - c_ref) <= 1e-5 * abs(c_ref)


def test_icd_fp16_and_resync(gpu, ssm_pair, orc):
    s, r, st, xg, xo, nupd, raw, w, host, sp = _icd_case(
        gpu, ssm_pair, orc, "tiny_ffs", 4, fp16=True, resync=2)
    rms = np.sqrt(np.mean((xg - xo) ** 2)) / np.sqrt(np.mean(xo ** 2))
    print(f"[icd-fp16] rel RMS vs FP64 oracle {rms:.2e}, resync {st.resync_ms:.2f} ms",
          flush=True)
    assert rms <= 5e-3
    # after a resync, e equals y - A x of the current volume (FP64 -> FP32)
    r.resync()
    e = r.error().astype(np.float64)
    y = orc.forward_project(s.geom, host, s.slice_index, r.volume().astype(np.float64))
    e_ref = raw - y
    assert np.abs(e - e_ref).max() <= 1e-3 * np.abs(e_ref).max()


def test_running_error_tracks_fresh_projection(gpu, ssm_pair, orc):
    s, r, st, xg, xo, nupd, raw, w, host, sp = _icd_case(gpu, ssm_pair, orc, "tiny", 2)
    e_run = r.error().astype(np.float64)
    y = orc.forward_project(s.geom, host, s.slice_index, r.volume().astype(np.float64))
    assert np.abs(e_run - (raw - y)).max() <= 1e-4 * np.abs(raw).max()


def test_weighted_running_error_tracks_fresh_projection(gpu, ssm_pair, orc):
    """Weighted sweeps keep w e current everywhere and e itself only where
    w == 0; the downloaded e (formed as w e / w where w > 0) equals y - A x of
    the FP64 projection on both kinds of samples."""
    s, dev, host = ssm_pair("tiny_ffs")
    raw, w = s.synthesize(use_gpu=False)
    w = w.copy()
    w[::7] = 0.0
    zero = w == 0.0
    r = gpu.Recon(dev, raw, w, s.slice_index, s.opts(tile=4))
    r.iterate(3)
    e_run = r.error().astype(np.float64)
    y = orc.forward_project(s.geom, host, s.slice_index, r.volume().astype(np.float64))
    ref = raw.astype(np.float64) - y
    scale = np.abs(raw).max()
    err_w, err_0 = (np.abs(e_run - ref)[~zero].max(), np.abs(e_run - ref)[zero].max())
    print(f"[weighted-error] max |e - (y - Ax)|: w > 0 {err_w:.2e}, w == 0 {err_0:.2e} "
          f"({zero.sum()} samples)", flush=True)
    # e started as y (x = 0): the sweeps moved it on the w == 0 samples too
    assert np.abs(e_run[zero] - raw[zero]).max() > 0.1 * scale
    assert err_w <= 1e-4 * scale and err_0 <= 1e-4 * scale
    # the running cost is formed from the same e
    c64, c32 = r.cost(), r.cost_running()
    assert abs(c32 - c64) <= 1e-5 * abs(c64)
    r.free()


def test_synthetic_generator_gpu_matches_cpu(gpu, make_scan):
    s = make_scan("tiny")
    a, _ = s.synthesize(use_gpu=False)
    b, _ = s.synthesize(use_gpu=True)
    assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
    s2 = make_scan("tiny_ffs")
    a, wa = s2.synthesize(use_gpu=False)
    b, wb = s2.synthesize(use_gpu=True)
    # Poisson counts come from the same counter-based streams; libm ulps may
    # flip a rare rejection step
    assert np.mean(wa == wb) > 0.999


def test_running_cost_matches_fp64_cost(gpu, ssm_pair, orc):
    """The cost from the kept FP32 error sinogram equals the FP64 cost up to
    the FP32 rounding of e: right after set_volume and after sweeps."""
    s, dev, host = ssm_pair("tiny_ffs")
    raw, w = s.synthesize(use_gpu=False)
    r = gpu.Recon(dev, raw, w, s.slice_index, s.opts())
    x0 = s.phantom()
    r.set_volume(x0)
    c64, c32 = r.cost(), r.cost_running()
    print(f"[running-cost] set_volume: fp64 {c64:.10e} running {c32:.10e}", flush=True)
    assert abs(c32 - c64) <= 1e-6 * abs(c64)
    r.iterate(2)
    c64, c32 = r.cost(), r.cost_running()
    print(f"[running-cost] 2 sweeps: fp64 {c64:.10e} running {c32:.10e}", flush=True)
    assert abs(c32 - c64) <= 1e-5 * abs(c64)
    r.free()


def test_repeated_runs_are_bit_identical(gpu, ssm_pair):
    """Sums are formed in a fixed order (no floating-point atomics): two
    reconstructions of the same scan with the default schedule and the
    batch surrogate -- first sweep included -- give the same bits, volume
    and error sinogram."""
    s, dev, _ = ssm_pair("tiny_ffs")
    raw, w = s.synthesize(use_gpu=False)
    out = []
    for _ in range(2):
        r = gpu.Recon(dev, raw, w, s.slice_index, s.opts())
        r.iterate(3)
        out.append((r.volume(), r.error()))
        r.free()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("name,over,tile", [
    ("micro", {}, 3),                                   # tiles cut the image raggedly
    ("tiny", dict(z_start=-1.5, z_end=-1.38), 4),       # a single slice
    ("tiny", dict(nx=23, ny=21), 4),                    # odd, non-square image
])
def test_icd_edge_shapes_match_oracle(gpu, make_scan, orc, name, over, tile):
    s = make_scan(name, **over)
    g = s.geom
    raw, w = s.synthesize(use_gpu=False)
    m = gpu.SSM.build(g)
    host = orc.ssm_build(g)
    sp = m.group_spacing(s.slice_index)
    r = gpu.Recon(m, raw, w, s.slice_index, s.opts(tile=tile, group_spacing=sp))
    st = r.iterate(2)
    xg = r.volume().astype(np.float64)
    xo, _, nupd = orc.icd(g, host, raw, w, s.slice_index, s.run.penalty, s.run.lambda_,
                          s.run.delta, sp, 2, tile=tile, surrogate=True)
    print(f"[edge] {name} {over} tile {tile}: {g.nx}x{g.ny}x{g.nz}, spacing {sp}, "
          f"{st.voxel_updates} updates", flush=True)
    assert st.voxel_updates == nupd
    rms = np.sqrt(np.mean((xg - xo) ** 2)) / max(np.sqrt(np.mean(xo ** 2)), 1e-30)
    assert rms <= 1e-4, rms
    r.free()


@pytest.mark.parametrize("name", ["tiny", "tiny_ffs", "tiny_zffs"])
def test_fbp_matches_oracle(gpu, make_scan, orc, name):
    """GPU FBP start image (ctbb_fbp) against the FP64 restatement."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
    import fbp_oracle
    s = make_scan(name)
    g = s.geom
    raw, w = s.synthesize(use_gpu=False)
    m = gpu.SSM.build(g)
    xg = m.fbp(raw, s.slice_index).astype(np.float64)
    xo = fbp_oracle.fbp(g, orc.ssm_build(g), raw, s.slice_index)
    rms = np.sqrt(np.mean((xg - xo) ** 2)) / np.sqrt(np.mean(xo ** 2))
    print(f"[fbp] {name}: rel RMS vs oracle {rms:.2e}", flush=True)
    assert rms <= 1e-4
    m16 = gpu.SSM.build(g)
    m16.to_fp16()
    x16 = m16.fbp(raw, s.slice_index).astype(np.float64)
    rms16 = np.sqrt(np.mean((x16 - xo) ** 2)) / np.sqrt(np.mean(xo ** 2))
    assert rms16 <= 5e-3, rms16
    # as a start image it is better than zeros (tiny scans of ~100 views per
    # turn leave FBP streaks, so the gain is moderate: measured 35 % on tiny)
    r = gpu.Recon(m, raw, w, s.slice_index, s.opts())
    c0 = r.cost()
    r.set_volume(np.maximum(xg, 0.0).astype(np.float32))
    c1 = r.cost()
    print(f"[fbp] {name}: cost zero start {c0:.6e}, FBP start {c1:.6e}", flush=True)
    assert c1 < 0.8 * c0
    r.free()


def test_fbp_wide_cone_matches_oracle(gpu, ssm_pair, orc):
    """The wide-cone scan that pins the FBP's cone weight (10 cm detector
    rows, cone half-angle up to 17.9 degrees): the device matrix equals the
    oracle's and the device FBP equals the restatement."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
    import fbp_oracle
    s, dev, (cp, idx, val) = ssm_pair("tiny_cone")
    gcp, gidx, gval = dev.download()
    assert np.array_equal(gcp, cp) and np.array_equal(gidx, idx)
    assert np.array_equal(gval, val)
    raw, _ = s.synthesize(use_gpu=False)
    xg = dev.fbp(raw, s.slice_index).astype(np.float64)
    xo = fbp_oracle.fbp(s.geom, (cp, idx, val), raw, s.slice_index)
    rms = np.sqrt(np.mean((xg - xo) ** 2)) / np.sqrt(np.mean(xo ** 2))
    print(f"[fbp] tiny_cone: rel RMS vs oracle {rms:.2e}", flush=True)
    assert rms <= 1e-4
