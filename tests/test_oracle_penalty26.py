"""Pins of the oracle's 26-neighbour penalty (no reference counterpart: the
reference stencil is 8 in-plane neighbours, src/penalties.cpp:39-59,
"limited to 2D", src/icd_iteration.cpp:140).

The slices of the rotating frame are rotated against each other by the tube
angle between their central views; the neighbours of voxel p in slice k+1
sit around pi_k(p), the three-shear integer rotation of p by
theta_k - theta_{k+1}.  Pinned here:
  * pi_k is a bijection of the grid (injective, exact inverse) within 1.5
    voxels of the true rotation, whose angle is the slice step in views
    times the tube angle increment;
  * the step of every voxel -- interior, slice ends, image edges -- is the
    exact coordinate minimiser of the 26-neighbour cost (quadratic: the
    cost is minimal at the step, compared at +-h; edge-preserving: the
    same within the bisection's 1e-8).  A partner list that is not
    symmetric (p lists q but q does not list p) makes the coordinate
    derivative of the cost differ from the step's, and fails this;
  * a constant image has a zero step;
  * sweeps never increase the 26-neighbour cost.
"""
import numpy as np
import pytest


def _angle(g, si, k):
    return -(int(si[k + 1]) - int(si[k])) * g.tube_angle_increment


@pytest.mark.parametrize("name", ["tiny_ffs", "config0"])
def test_slice_rotation_is_a_bijection_near_the_rotation(make_scan, orc, name):
    s = make_scan(name)
    g = s.geom
    si = s.slice_index
    k = 0
    phi = _angle(g, si, k)
    seen = set()
    worst = 0.0
    pad = 8
    for j in range(-pad, g.ny + pad, 1 if g.nx <= 32 else 3):
        for i in range(-pad, g.nx + pad, 1 if g.nx <= 32 else 3):
            io, jo = orc.rot26(g, si, k, i, j)
            assert (io, jo) not in seen
            seen.add((io, jo))
            assert orc.rot26(g, si, k, io, jo, inverse=True) == (i, j)
            x, y = i - g.center_voxel_x, j - g.center_voxel_y
            xr = np.cos(phi) * x - np.sin(phi) * y + g.center_voxel_x
            yr = np.sin(phi) * x + np.cos(phi) * y + g.center_voxel_y
            worst = max(worst, np.hypot(io - xr, jo - yr))
    print(f"[rot26] {name}: slice step {np.degrees(phi):.2f} deg, worst {worst:.2f} voxel")
    assert worst <= 1.5


def _dense(orc, s, ssm):
    g = s.geom
    A = np.zeros((g.data_size, g.nvox))
    e = np.zeros(g.nvox)
    for v in range(g.nvox):
        e[:] = 0.0
        e[v] = 1.0
        A[:, v] = orc.forward_project(g, ssm, s.slice_index, e)
    return A


@pytest.mark.parametrize("penalty", [0, 1])
def test_step26_is_the_coordinate_minimiser(make_scan, orc, penalty):
    s = make_scan("micro")
    g = s.geom
    ssm = orc.ssm_build(g)
    raw, _ = s.synthesize(use_gpu=False)
    rng = np.random.default_rng(9)
    wt = rng.uniform(0.5, 2.0, g.data_size).astype(np.float32)
    lam, delta = 0.3, 0.02
    A = _dense(orc, s, ssm)
    x = rng.uniform(0.0, 0.3, g.nvox)
    r = raw.astype(np.float64) - A @ x
    assert g.nz >= 3
    nxy = g.nx * g.ny
    # every slice (first, middle, last); centre, edge and corner columns
    for k in range(g.nz):
        for c in (0, g.nx - 1, 3 + g.nx * 4, nxy - 1, g.nx * 3):
            q = c + nxy * k
            a = A[:, q]
            alpha = float(np.sum(wt * a * a))
            beta = float(np.sum(wt * a * r))
            du = orc.penalty_step(g, s.slice_index, 26, penalty, q, alpha, beta, lam, delta, x)

            def phi(t):
                xt = x.copy()
                xt[q] += t
                return orc.cost(g, ssm, raw, wt, s.slice_index, penalty, lam, delta, xt,
                                neighbours=26)
            c0 = phi(du)
            h = 1e-4
            assert c0 <= phi(du + h) and c0 <= phi(du - h), (k, c, du)


def test_constant_image_has_zero_step26(make_scan, orc):
    s = make_scan("tiny_ffs")
    g = s.geom
    x = np.full(g.nvox, 0.19)
    for pen in (0, 1):
        for q in (0, 5, g.nx * g.ny * 3 + 77, g.nvox - 1):
            du = orc.penalty_step(g, s.slice_index, 26, pen, q, 2.0, 0.0, 0.5, 0.02, x)
            assert abs(du) <= 1e-8, (pen, q, du)


@pytest.mark.parametrize("spacing,tile", [(0, 8), (3, 4)])
def test_sweeps26_never_increase_cost(make_scan, orc, spacing, tile):
    s = make_scan("tiny_ffs")
    g = s.geom
    ssm = orc.ssm_build(g)
    raw, w = s.synthesize(use_gpu=False)
    pen, lam, delta = s.run.penalty, 50.0, s.run.delta
    x, yhat = None, None
    prev = orc.cost(g, ssm, raw, w, s.slice_index, pen, lam, delta, np.zeros(g.nvox),
                    neighbours=26)
    for _ in range(3):
        x, yhat, _ = orc.icd(g, ssm, raw, w, s.slice_index, pen, lam, delta, spacing, 1,
                             x=x, yhat=yhat, tile=tile, surrogate=True, neighbours=26)
        c = orc.cost(g, ssm, raw, w, s.slice_index, pen, lam, delta, x, neighbours=26)
        assert c <= prev * (1 + 1e-12)
        prev = c
