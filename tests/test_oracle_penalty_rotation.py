"""Oracle penalties and slice rotation against golden vectors written by the
reference's own src/penalties.cpp and src/rotate_slices.cpp (tests/golden),
plus closed forms: root of the edge-preserving derivative, rho' = psi,
rotation by 0, constant images, rotation round trips."""
from pathlib import Path

import numpy as np
import pytest
from scipy.optimize import brentq

GOLD = Path(__file__).resolve().parent / "golden"


def geom_of(ct, nx, ny, nz):
    g = ct.Geom()
    g.nx, g.ny, g.nz = nx, ny, nz
    g.center_voxel_x, g.center_voxel_y = (nx - 1) / 2, (ny - 1) / 2
    return g


def test_penalties_match_reference_golden(ct, orc):
    d = np.load(GOLD / "penalties.npz")
    nx, ny, nz = map(int, d["dims"])
    g = geom_of(ct, nx, ny, nz)
    x = d["x"]
    for a, lam in enumerate(d["lam"]):
        for i, q in enumerate(d["q"]):
            got = orc.quadratic(g, int(q), d["alpha"][i], d["beta"][i], lam, x)
            assert got == d["quadratic"][a, i] or (np.isnan(got) and np.isnan(d["quadratic"][a, i]))
            for b, delta in enumerate(d["delta"]):
                got = orc.edge_preserving(g, int(q), d["alpha"][i], d["beta"][i], lam, delta, x)
                assert got == d["edge_preserving"][a, b, i]


def test_penalties_match_reference_live(ct, orc):
    ref = orc.ref()
    if ref is None:
        pytest.skip("reference sources not present (oracle/_ref not built)")
    rng = np.random.default_rng(5)
    nx, ny, nz = 9, 8, 4
    g = geom_of(ct, nx, ny, nz)
    x = np.ascontiguousarray(rng.uniform(0, 0.3, nx * ny * nz))
    for _ in range(300):
        q = int(rng.integers(0, nx * ny * nz))
        a, b = rng.uniform(0.1, 30), rng.normal(0, 3)
        lam, delta = rng.uniform(0, 10), rng.uniform(0.002, 0.2)
        assert orc.quadratic(g, q, a, b, lam, x) == ref.ref_quadratic(nx, ny, nz, q, a, b, lam, x.ctypes.data)
        assert orc.edge_preserving(g, q, a, b, lam, delta, x) == ref.ref_edge_preserving(
            nx, ny, nz, q, a, b, lam, delta, x.ctypes.data)


def _neighbours(q, nx, n):
    offs = [-nx - 1, -nx, -nx + 1, -1, 1, nx - 1, nx, nx + 1]
    w = [2 ** -0.5, 1, 2 ** -0.5, 1, 1, 2 ** -0.5, 1, 2 ** -0.5]
    return [(q + o, wi) for o, wi in zip(offs, w) if 0 <= q + o < n]


def test_quadratic_closed_form(ct, orc):
    rng = np.random.default_rng(1)
    nx, ny, nz = 6, 6, 2
    g = geom_of(ct, nx, ny, nz)
    n = nx * ny * nz
    x = rng.uniform(0, 1, n)
    W = 4 + 2 * 2 ** 0.5
    for q in [0, 7, 14, 35, 40, 71]:
        a, b, lam = 3.0, -0.7, 0.9
        s = sum(w * (x[m] - x[q]) for m, w in _neighbours(q, nx, n))
        assert orc.quadratic(g, q, a, b, lam, x) == pytest.approx((b + lam * s) / (a + lam * W), rel=1e-14)


def test_edge_preserving_root(ct, orc):
    rng = np.random.default_rng(2)
    nx, ny, nz = 6, 5, 3
    g = geom_of(ct, nx, ny, nz)
    n = nx * ny * nz
    x = rng.uniform(0, 0.3, n)
    for _ in range(100):
        q = int(rng.integers(0, n))
        a, b = rng.uniform(0.5, 20), rng.normal(0, 2)
        lam, delta = rng.uniform(0.01, 5), rng.uniform(0.005, 0.1)
        nb = _neighbours(q, nx, n)

        def f(c):
            s = 0.0
            for m, w in nb:
                t = c - (x[m] - x[q])
                s += w * np.sign(t if t != 0 else 1.0) * (1 - 1 / (1 + abs(t / delta)))
            return 2 * a * c - 2 * b + lam * s
        lo = min([b / a] + [x[m] - x[q] for m, _ in nb]) - 1e-9
        hi = max([b / a] + [x[m] - x[q] for m, _ in nb]) + 1e-9
        root = brentq(f, lo, hi, xtol=1e-14)
        assert abs(orc.edge_preserving(g, q, a, b, lam, delta, x) - root) < 2e-8


def test_rho_prime_is_psi():
    delta = 0.02
    t = np.linspace(-0.3, 0.3, 601)
    t = t[np.abs(t) > 1e-3]
    rho = lambda s: np.abs(s) - delta * np.log1p(np.abs(s) / delta)
    h = 1e-6
    fd = (rho(t + h) - rho(t - h)) / (2 * h)
    psi = np.sign(t) * (1 - 1 / (1 + np.abs(t) / delta))
    assert np.allclose(fd, psi, atol=1e-7)


def test_rotation_matches_reference_golden(ct, orc):
    d = np.load(GOLD / "rotation.npz")
    nx, nz, readings = map(int, d["n"])
    g = geom_of(ct, nx, nx, nz)
    g.readings = readings
    for direction, tag in [(1, "rotating2fixed"), (-1, "fixed2rotating")]:
        got = orc.rotate_slices(g, d["tube_angles"], d["slice_index"], d["volume"], direction)
        assert np.array_equal(got, d[tag])


def test_rotation_closed_forms(ct, orc):
    n, nz = 33, 2
    g = geom_of(ct, n, n, nz)
    g.readings = 4
    vol = np.random.default_rng(3).uniform(0, 1, n * n * nz).astype(np.float32)
    ident = orc.rotate_slices(g, np.zeros(4), np.array([0, 1]), vol, 1)
    v3, i3 = vol.reshape(nz, n, n), ident.reshape(nz, n, n)
    # angle 0: identity except the border rule zeroes the last row and column
    assert np.array_equal(i3[:, :n - 1, :n - 1], v3[:, :n - 1, :n - 1])
    assert np.all(i3[:, n - 1, :] == 0) and np.all(i3[:, :, n - 1] == 0)
    const = np.full(n * n * nz, 0.25, np.float32)
    rot = orc.rotate_slices(g, np.array([33.0, 71.0, 0, 0]), np.array([0, 1]), const, 1).reshape(nz, n, n)
    yy, xx = np.mgrid[0:n, 0:n] - (n - 1) / 2
    inside = xx ** 2 + yy ** 2 < ((n - 3) / 2) ** 2
    assert np.allclose(rot[:, inside], 0.25, atol=1e-6)
    # fixed -> rotating -> fixed on a smooth image (two bilinear passes blur a
    # sigma = 5 px Gaussian by ~1.5% at the peak; a wrong sign errs by O(1))
    smooth = np.exp(-(xx ** 2 + yy ** 2) / (2 * 5.0 ** 2)).astype(np.float32)
    sm = np.stack([smooth] * nz).reshape(-1)
    ang = np.array([40.0, 123.0, 0, 0])
    back = orc.rotate_slices(g, ang, np.array([0, 1]),
                             orc.rotate_slices(g, ang, np.array([0, 1]), sm, -1), 1)
    assert np.abs(back.reshape(nz, n, n)[:, inside] - np.stack([smooth] * nz)[:, inside]).max() < 0.02
