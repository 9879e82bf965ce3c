"""Pins of the FBP start-image restatement (oracle/fbp_oracle.py).

The reference's own wFBP start is an external program, so nothing in the
reference fixes these numbers; the pins are closed-form properties: the FBP
of a uniform water cylinder returns its attenuation inside it, and zero
outside the acquisition FOV.  A dropped pre-weight, a wrong ramp-kernel
constant or a wrong view normalisation moves the interior value far from mu.
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
