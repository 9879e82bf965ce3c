"""numpy restatement of the GPU filtered backprojection start image
(ctbb_fbp, include/ctbb_icd.h).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels.  FP64 throughout.  No reference counterpart: the reference starts
from an external wFBP program (src/initialize_recon_volume.cpp:82-178);
this is an equiangular fan-beam FBP (Kak & Slaney, Principles of
Computerized Tomographic Imaging, ch. 3.4.2) per rotated slice, with an FDK
cone weight, backprojected through the stored system matrix:

  R(view, ch, row) = y * R_src * cos(gamma_ch) * D / sqrt(D^2 + ap_row^2)
  Q(view, ch, row) = alpha * sum_k R(view, k, row) g(ch - k),
      g(0) = 1 / (8 alpha^2), g(n even) = 0,
      g(n odd) = -1 / (2 pi^2 sin^2(n alpha))
  x(c, k) = 2 pi * sum_v w_v P_v / sum_v w_v,
      P_v = (sum_e a_e Q[ray_e + off_k]) / (sum_e a_e) / L_v(c)^2
  over the SSM views v with entries e of column c whose samples lie in the
  sinogram; L_v(c): in-plane distance from the view's nominal source to the
  column.  Redundancy weight w_v = 1 / n_v, n_v = the number of views of the
  column's view range [first, last] at the same source angle as v (views
  N_turn apart), so every projection angle counts once: first / last = the
  views of the column's first / last entry, clipped to the views whose
  samples lie in the sinogram for slice k.

Pins (tests/test_fbp_oracle.py): water cylinders return mu inside and zero
outside the FOV; on the config 0 scanner a 10 cm cylinder resolves the ramp
constant, cos(gamma), 1/L^2 and the redundancy weights; on a scan with 10 cm
detector rows (cone half-angle up to 17.9 degrees) the z-invariant cylinder
resolves the FDK cone weight (without it the interior reads 2.1 % high).
"""
from __future__ import annotations

import numpy as np


def channel_angles(g):
    ffs = bool(g.zffs or g.phiffs)
    j = np.arange(g.n_channels, dtype=np.float64)
    if ffs:
        return (j - g.center_channel_non_ffs) * g.fan_angle_increment, g.fan_angle_increment
    a = g.transaxial_detector_spacing / g.source_detector_distance
    return (j - g.center_channel_non_ffs) * a, a


def filter_sinogram(g, raw):
    """Q [readings, n_channels, n_rows] from raw in the reference layout."""
    nv, nch, nr = g.readings, g.n_channels, g.n_rows
    p = np.asarray(raw, np.float64).reshape(nv, nch, nr)
    gam, alpha = channel_angles(g)
    ap = (np.arange(nr) - g.center_row) * g.axial_detector_spacing
    D = g.source_detector_distance
    w = (g.focal_spot_radius * np.cos(gam))[:, None] * (D / np.sqrt(D * D + ap * ap))[None, :]
    R = p * w[None, :, :]
    # H[j, k] = g(j - k); Q[v, j, r] = alpha * sum_k H[j, k] R[v, k, r]
    n = np.arange(nch)[:, None] - np.arange(nch)[None, :]
    H = np.zeros((nch, nch))
    H[n == 0] = 1.0 / (8.0 * alpha * alpha)
    odd = (n % 2) != 0
    H[odd] = -1.0 / (2.0 * np.pi ** 2 * np.sin(n[odd] * alpha) ** 2)
    Q = alpha * np.matmul(H, R)            # batched over views: (nch, nch) @ (nch, nr)
    return Q


def redundancy_weights(v, first, last, nturn):
    """1 / n_v for views v of the range [first, last]: n_v = the number of
    views of the range at v's source angle (views nturn apart)."""
    v = np.asarray(v, np.int64)
    return 1.0 / ((last - v) // nturn + (v - first) // nturn + 1)


def fbp(g, ssm, raw, slice_index, slices=None):
    """Rotating-frame FBP volume [nz * ny * nx] (index c + nx*ny*k); with
    `slices`, only those slices are formed (the others stay zero)."""
    cp, idx, val = ssm
    Q = filter_sinogram(g, raw).reshape(-1)
    ds = Q.size
    nsm, nch, nr = g.n_views_sm, g.n_channels, g.n_rows
    nffs = (1 << int(g.zffs)) * (1 << int(g.phiffs))
    offs = []
    for k in range(g.nz):
        ci = int(slice_index[k])
        if g.zffs or g.phiffs:
            ci -= ci % nffs
        offs.append((ci - nsm // 2) * nch * nr)
    ncol = g.nx * g.ny
    x = np.zeros(g.nz * ncol)
    fov = (g.acquisition_fov / 2.0) ** 2
    R = g.focal_spot_radius
    nturn = int(round(2.0 * np.pi / g.tube_angle_increment))
    for c in range(ncol):
        ix, iy = c % g.nx, c // g.nx
        px = (ix - g.center_voxel_x) * g.voxel_size_x
        py = (iy - g.center_voxel_y) * g.voxel_size_y
        if not (px * px + py * py < fov):
            continue
        q = idx[cp[c]:cp[c + 1]].astype(np.int64)
        a = val[cp[c]:cp[c + 1]].astype(np.float64)
        if q.size == 0:
            continue
        view = q // (nch * nr)
        th = (view - 0.5 * nsm) * g.tube_angle_increment
        L2 = (R * np.cos(th) - px) ** 2 + (R * np.sin(th) - py) ** 2
        starts = np.flatnonzero(np.r_[True, view[1:] != view[:-1]])
        for k in (range(g.nz) if slices is None else slices):
            r = q + offs[k]
            ok = (r >= 0) & (r < ds)
            aq = np.where(ok, a * Q[np.clip(r, 0, ds - 1)], 0.0)
            aa = np.where(ok, a, 0.0)
            sq = np.add.reduceat(aq, starts)
            sa = np.add.reduceat(aa, starts)
            l2 = L2[starts]
            m = sa > 0
            if m.any():
                offv = offs[k] // (nch * nr)
                first = max(int(view[0]), -offv)
                last = min(int(view[-1]), g.readings - 1 - offv)
                w = redundancy_weights(view[starts][m], first, last, nturn)
                x[c + ncol * k] = 2.0 * np.pi * np.sum(w * sq[m] / sa[m] / l2[m]) / np.sum(w)
    return x
