"""Pins of the oracle on the clinical 64-row detector and of its helpers.

  * The column-subset SSM (voxel-driven, for clinical geometries where the
    ray-driven build is out of reach) equals the ray-driven build column for
    column, bit for bit.
  * The in-plane flying focal spot deflection da = D R sin(dgamma) / (4 (D-R))
    (src/generate_system_matrix_ffs.cpp:67, applied +da e_u on even and -da
    e_u on odd views, :250-261) moves a central voxel's footprint by a quarter
    channel each way: the value-weighted channel centroids of even and odd
    views differ by half a channel (the interlace the deflection exists
    for).  A dropped, halved or sign-flipped da fails it.
  * The ray-driven build over a range of views (the CPU baseline of the
    matrix build times it on a view sample) gives, range by range, exactly
    the full build's entries of those views, and rejects ranges outside the
    matrix.
  * The footprint-overlap helper equals a brute-force set intersection.
  * The OpenMP CPU baseline (the reference's loop structure) computes what the
    sequential oracle computes in the reference order, with the matrix held in
    RAM or read from a matrix.bin file every iteration.
  * The batch surrogate on the clinical cone geometry with the library's
    default slice spacing (footprint conflicts inside <= 1 % of the columns):
    every step of a sweep lowers the cost although the step's voxels share
    data samples across columns and slices (counted, so the check is not
    vacuous).
"""
import numpy as np
import pytest


def _interlace(g, ptr, idx, val, t):
    q = idx[ptr[t]:ptr[t + 1]].astype(np.int64)
    v = val[ptr[t]:ptr[t + 1]].astype(np.float64)
    view = q // (g.n_rows * g.n_channels)
    ch = (q // g.n_rows) % g.n_channels
    sw = np.bincount(view, weights=v, minlength=g.n_views_sm)
    sc = np.bincount(view, weights=v * ch, minlength=g.n_views_sm)
    cen = np.where(sw > 0, sc / np.maximum(sw, 1e-30), np.nan)
    n = g.n_views_sm // 2
    return np.nanmean(cen[0:2 * n:2]) - np.nanmean(cen[1:2 * n:2])


@pytest.mark.parametrize("name", ["tiny", "tiny_ffs", "tiny_zphiffs", "clin_crop"])
def test_column_subset_ssm_equals_ray_driven_build(make_scan, orc, name):
    s = make_scan(name)
    g = s.geom
    cp, idx, val = orc.ssm_build(g)
    ncol = g.nx * g.ny
    cols = np.arange(ncol) if ncol <= 1024 else \
        np.random.default_rng(5).choice(ncol, 300, replace=False)
    ptr, cidx, cval = orc.ssm_cols(g, cols)
    for t, c in enumerate(cols):
        a, b = cp[c], cp[c + 1]
        assert ptr[t + 1] - ptr[t] == b - a, (name, c)
        assert np.array_equal(cidx[ptr[t]:ptr[t + 1]], idx[a:b]), (name, c)
        assert np.array_equal(cval[ptr[t]:ptr[t + 1]], val[a:b]), (name, c)


@pytest.mark.parametrize("name", ["tiny", "tiny_ffs", "tiny_zphiffs"])
def test_view_range_build_is_a_piece_of_the_full_build(make_scan, orc, name):
    g = make_scan(name).geom
    cp, idx, val = orc.ssm_build(g)
    n = g.n_views_sm
    cuts = [0, 1, n // 3, n // 3 + 2, n - 1, n]
    pieces = [orc.ssm_build_views(g, a, b) for a, b in zip(cuts[:-1], cuts[1:])]
    per_view = g.n_rows * g.n_channels
    for c in range(g.nx * g.ny):
        ref_i, ref_v = idx[cp[c]:cp[c + 1]], val[cp[c]:cp[c + 1]]
        got_i = np.concatenate([p[1][p[0][c]:p[0][c + 1]] for p in pieces])
        got_v = np.concatenate([p[2][p[0][c]:p[0][c + 1]] for p in pieces])
        assert np.array_equal(got_i, ref_i), (name, c)
        assert np.array_equal(got_v, ref_v), (name, c)
    for (a, b), p in zip(zip(cuts[:-1], cuts[1:]), pieces):
        q = p[1].astype(np.int64)
        assert np.all((q >= a * per_view) & (q < b * per_view)), (name, a, b)
    with pytest.raises(ValueError):
        orc.ssm_build_views(g, n - 1, n + 1)
    with pytest.raises(ValueError):
        orc.ssm_build_views(g, 2, 1)


@pytest.mark.parametrize("name,cols", [
    ("tiny_ffs", [11 + 24 * 11, 12 + 24 * 11, 11 + 24 * 12, 12 + 24 * 12]),
    ("clin_crop", [31 + 64 * 31, 32 + 64 * 32]),
    ("config1", [255 + 512 * 255, 256 + 512 * 256, 255 + 512 * 256]),
])
def test_inplane_ffs_interlaces_half_a_channel(make_scan, orc, name, cols):
    """A voxel at the isocentre seen from a source displaced by +-da along e_u
    lands da (D - R) / (R D) = sin(dgamma) / 4 away in fan angle, i.e. a
    quarter channel; even minus odd views: half a channel."""
    s = make_scan(name)
    g = s.geom
    assert g.phiffs == 1 and g.zffs == 0
    ptr, idx, val = orc.ssm_cols(g, cols)
    for t in range(len(cols)):
        il = _interlace(g, ptr, idx, val, t)
        assert abs(il - 0.5) <= 0.02, (name, cols[t], il)


def test_no_ffs_has_no_interlace(make_scan, orc):
    s = make_scan("tiny")
    g = s.geom
    cols = [11 + 24 * 11, 12 + 24 * 12]
    ptr, idx, val = orc.ssm_cols(g, cols)
    for t in range(len(cols)):
        assert abs(_interlace(g, ptr, idx, val, t)) <= 0.02


@pytest.mark.parametrize("name", ["tiny_ffs", "clin_crop"])
def test_column_overlaps_equals_brute_force(make_scan, orc, name):
    s = make_scan(name)
    g = s.geom
    ssm = orc.ssm_build(g)
    cp, idx, _ = ssm
    rng = np.random.default_rng(6)
    cols = rng.choice(g.nx * g.ny, 60, replace=False)
    plane = g.n_channels * g.n_rows
    hits = 0
    for c in cols:
        S = idx[cp[c]:cp[c + 1]].astype(np.int64)
        for dv in [0, 1, 2, 6, 12, 36, 66, 72, 78, g.n_views_sm // 2, g.n_views_sm - 1,
                   g.n_views_sm]:
            brute = len(S) > 0 and np.intersect1d(S, S + dv * plane).size > 0
            assert orc.column_overlaps(g, ssm, int(c), dv) == brute, (c, dv)
            hits += brute
    assert hits > 0


@pytest.mark.parametrize("name", ["tiny", "tiny_ffs"])
def test_openmp_baseline_equals_sequential_oracle(make_scan, orc, name):
    s = make_scan(name)
    g = s.geom
    ssm = orc.ssm_build(g)
    raw, w = s.synthesize(use_gpu=False)
    pen, lam, delta = s.run.penalty, s.run.lambda_, s.run.delta
    xs, ys, ns = orc.icd(g, ssm, raw, w, s.slice_index, pen, lam, delta, 0, 2)
    xo, yo, no = orc.icd_ref_omp(g, ssm, raw, w, s.slice_index, pen, lam, delta, 2, 4)
    assert ns == no
    # the OpenMP reductions add in another order: equal to ~1e-12
    assert np.abs(xo - xs).max() <= 1e-12 * max(1.0, np.abs(xs).max())
    assert np.abs(yo - ys).max() <= 1e-12 * max(1.0, np.abs(ys).max())


@pytest.mark.parametrize("name", ["tiny", "tiny_ffs"])
def test_streamed_baseline_equals_in_ram_baseline(make_scan, orc, name, tmp_path):
    """The CPU baseline with the matrix read from a matrix.bin file on every
    iteration (as the reference reads it) computes what the in-RAM form
    computes on the same band of rows; a file shorter than the band is an
    error."""
    s = make_scan(name)
    g = s.geom
    ssm = orc.ssm_build(g)
    raw, w = s.synthesize(use_gpu=False)
    pen, lam, delta = s.run.penalty, s.run.lambda_, s.run.delta
    r0, r1 = g.ny // 4, g.ny // 4 + 6
    f = tmp_path / "band.bin"
    orc.write_band_file(f, g, ssm, r0, r1)
    assert f.stat().st_size == 8 * (r1 - r0) * g.nx + 8 * int(ssm[0][r1 * g.nx]
                                                              - ssm[0][r0 * g.nx])
    xr, yr, nr = orc.icd_ref_omp(g, ssm, raw, w, s.slice_index, pen, lam, delta, 2, 3,
                                 r0, r1)
    xf, yf, nf = orc.icd_ref_omp_file(g, f, raw, w, s.slice_index, pen, lam, delta, 2, 3,
                                      r0, r1)
    assert nf == nr > 0
    assert np.abs(xf - xr).max() <= 1e-12 * max(1.0, np.abs(xr).max())
    assert np.abs(yf - yr).max() <= 1e-12 * max(1.0, np.abs(yr).max())
    short = tmp_path / "short.bin"
    short.write_bytes(f.read_bytes()[:-4])
    with pytest.raises(IOError):
        orc.icd_ref_omp_file(g, short, raw, w, s.slice_index, pen, lam, delta, 1, 3, r0, r1)


def _default_spacing(orc, g, ssm, slice_index, frac=0.01):
    """Smallest spacing whose slice groups have overlapping footprints in at
    most `frac` of the FOV columns (the library's default rule)."""
    cols = range(g.nx * g.ny)
    for sp in range(2, g.nz + 1):
        shifts = sorted({abs(int(slice_index[k2] - slice_index[k])) for k in range(g.nz)
                         for k2 in range(k + sp, g.nz, sp)})
        shifts = [d for d in shifts if d < g.n_views_sm]
        bad = sum(1 for c in cols if any(orc.column_overlaps(g, ssm, c, d) for d in shifts))
        if bad <= frac * g.nx * g.ny:
            return sp, bad
    return g.nz, 0


def test_batch_surrogate_lowers_cost_every_step_on_clinical_cone(make_scan, orc):
    s = make_scan("clin_crop")
    g = s.geom
    ssm = orc.ssm_build(g)
    raw, w = s.synthesize(use_gpu=False)
    assert w is not None
    sp, bad = _default_spacing(orc, g, ssm, s.slice_index)
    assert bad > 0, "the default spacing of this geometry should allow some conflicts"
    pen, lam, delta = s.run.penalty, s.run.lambda_, s.run.delta
    x, _, n, costs, shared = orc.icd_trace(g, ssm, raw, w, s.slice_index, pen, lam,
                                           delta, sp, 1, 4, surrogate=1, max_costs=4096)
    print(f"[surrogate] clin_crop spacing {sp} ({bad} conflicting columns), "
          f"{len(costs) - 1} steps, shared samples {shared}, "
          f"cost {costs[0]:.6e} -> {costs[-1]:.6e}")
    assert len(costs) - 1 == 16 * sp
    # voxels of one step share samples across columns and slices, and (the
    # default spacing) inside a column
    assert shared[0] > 0 and shared[1] > 0 and shared[2] > 0
    d = np.diff(costs)
    assert np.all(d <= 1e-13 * costs[0]), (d.max(), int(np.argmax(d)))
    assert costs[-1] < costs[0]
