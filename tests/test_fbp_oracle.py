"""Pins of the FBP start-image restatement (oracle/fbp_oracle.py).

The reference's own wFBP start is an external program, so nothing in the
reference fixes these numbers; the pins are closed-form properties: the FBP
of a uniform water cylinder returns its attenuation inside it, and zero
outside the acquisition FOV.  A wrong ramp-kernel constant or a wrong view
normalisation moves the interior value far from mu (small scans); on the
config 0 scanner a 10 cm cylinder also resolves the cos(gamma) pre-weight
(dropped: centre 0.6 % low), the 1 / L^2 weight and the redundancy weights
(dropped: 1.2-1.3 % RMS over the interior, against 0.47 %).  The FDK cone
weight D / sqrt(D^2 + v^2) moves these numbers by < 0.01 %; it is pinned on
a scan with 10 cm detector rows (cone half-angle up to 17.9 degrees): a
z-invariant cylinder's line integral along a ray at cone angle kappa is its
in-plane chord / cos(kappa), which the weight takes back to the chord, so
the interior returns mu to 0.5 % with it and 2.1 % high without it.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def fbp_oracle():
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
    import fbp_oracle
    return fbp_oracle


@pytest.mark.parametrize("name", ["tiny", "tiny_ffs", "tiny_zffs"])
def test_fbp_of_a_water_cylinder_returns_mu(make_scan, orc, fbp_oracle, name):
    s = make_scan(name, synthetic_n_ellipsoids=0, synthetic_I0=0.0, synthetic_weights=0)
    g = s.geom
    raw, _ = s.synthesize(use_gpu=False)
    x = fbp_oracle.fbp(g, orc.ssm_build(g), raw, s.slice_index).reshape(g.nz, g.ny, g.nx)
    truth = s.phantom().reshape(g.nz, g.ny, g.nx)
    mu = 0.193
    k = g.nz // 2
    X, Y = np.meshgrid((np.arange(g.nx) - g.center_voxel_x) * g.voxel_size_x,
                       (np.arange(g.ny) - g.center_voxel_y) * g.voxel_size_y)
    inner = X ** 2 + Y ** 2 < (0.7 * 2.0) ** 2          # cylinder radius 2.0 cm
    assert np.allclose(truth[k][inner], mu)
    ratio = x[k][inner].mean() / mu
    print(f"[fbp] {name}: interior mean / mu = {ratio:.3f}", flush=True)
    assert 0.9 < ratio < 1.1, ratio
    fov = X ** 2 + Y ** 2 < (g.acquisition_fov / 2) ** 2
    assert np.all(x[:, ~fov] == 0.0)


def test_fbp_of_a_large_cylinder_resolves_the_weights(make_scan, orc, fbp_oracle):
    """config 0 scanner, 10 cm water cylinder, central slice only."""
    s = make_scan("config0", synthetic_n_ellipsoids=0, synthetic_I0=0.0,
                  synthetic_weights=0, synthetic_cylinder_radius=10.0)
    g = s.geom
    raw, _ = s.synthesize(use_gpu=False)
    k = g.nz // 2
    x = fbp_oracle.fbp(g, orc.ssm_build(g), raw, s.slice_index,
                       slices=[k]).reshape(g.nz, g.ny, g.nx)[k]
    mu = 0.193
    X, Y = np.meshgrid((np.arange(g.nx) - g.center_voxel_x) * g.voxel_size_x,
                       (np.arange(g.ny) - g.center_voxel_y) * g.voxel_size_y)
    r = np.hypot(X, Y)
    inner, centre = r < 8.5, r < 3.0
    dev = x[inner] / mu - 1.0
    c = x[centre].mean() / mu
    rms = np.sqrt(np.mean(dev ** 2))
    print(f"[fbp] config0 cylinder: centre {c:.4f}, interior mean {dev.mean() + 1:.4f}, "
          f"RMS {rms:.4f}", flush=True)
    assert abs(c - 1.0) < 0.003, c
    assert abs(dev.mean()) < 0.005, dev.mean()
    assert rms < 0.008, rms


def test_redundancy_weights_count_every_angle_once(fbp_oracle):
    """Over any view range, the weights of the views at one source angle
    sum to 1, so the range's weights add up to its number of distinct
    angles, min(span, N_turn) (an off-by-one in n_v fails this)."""
    rng = np.random.default_rng(7)
    for _ in range(300):
        nturn = int(rng.integers(1, 40))
        first = int(rng.integers(0, 50))
        last = first + int(rng.integers(0, 4 * nturn + 3))
        v = np.arange(first, last + 1)
        w = fbp_oracle.redundancy_weights(v, first, last, nturn)
        for a in range(nturn):
            sel = (v % nturn) == a
            if sel.any():
                assert np.isclose(w[sel].sum(), 1.0), (nturn, first, last, a)
        assert np.isclose(w.sum(), min(last - first + 1, nturn))


def test_fbp_cone_weight_on_a_wide_cone(make_scan, orc, fbp_oracle):
    s = make_scan("tiny_cone", synthetic_I0=0.0, synthetic_weights=0)
    g = s.geom
    ap = (np.arange(g.n_rows) - g.center_row) * g.axial_detector_spacing
    kappa = np.degrees(np.arctan(np.abs(ap).max() / g.source_detector_distance))
    assert kappa > 15.0
    raw, _ = s.synthesize(use_gpu=False)
    k = g.nz // 2
    x = fbp_oracle.fbp(g, orc.ssm_build(g), raw, s.slice_index,
                       slices=[k]).reshape(g.nz, g.ny, g.nx)[k]
    X, Y = np.meshgrid((np.arange(g.nx) - g.center_voxel_x) * g.voxel_size_x,
                       (np.arange(g.ny) - g.center_voxel_y) * g.voxel_size_y)
    inner = X ** 2 + Y ** 2 < (0.7 * 2.0) ** 2          # cylinder radius 2.0 cm
    ratio = x[inner].mean() / 0.193
    print(f"[fbp] tiny_cone (cone half-angle {kappa:.1f} deg): interior mean / mu = "
          f"{ratio:.5f}", flush=True)
    assert abs(ratio - 1.0) < 5e-3, ratio
