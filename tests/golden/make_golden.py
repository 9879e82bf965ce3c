"""Write golden vectors for the penalty steps and the slice rotation by running
the reference's own src/penalties.cpp and src/rotate_slices.cpp (+ interp.h),
compiled unmodified into oracle/_ref/libctbb_ref.so by oracle/Makefile.

Run in the container that holds the reference sources:
    make -C oracle && python tests/golden/make_golden.py
Output: tests/golden/penalties.npz, tests/golden/rotation.npz (inputs and
expected outputs only)."""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "oracle"))
import oracle  # noqa: E402


def main():
    ref = oracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libctbb_ref.so is not built")
    rng = np.random.default_rng(20261203)
    # penalties: a 7 x 6 x 3 volume (every voxel, boundaries included)
    nx, ny, nz = 7, 6, 3
    n = nx * ny * nz
    x = rng.uniform(0.0, 0.4, n)
    x[rng.random(n) < 0.15] = 0.0
    q = np.arange(n, dtype=np.int64)
    alpha = rng.uniform(0.0, 50.0, n)
    alpha[::17] = 0.0
    beta = rng.normal(0.0, 5.0, n)
    lam = np.array([0.0, 0.3, 25.0])
    delta = np.array([0.005, 0.02, 0.5])
    quad = np.zeros((len(lam), n))
    ep = np.zeros((len(lam), len(delta), n))
    xp = np.ascontiguousarray(x)
    for a, L in enumerate(lam):
        for i in range(n):
            quad[a, i] = ref.ref_quadratic(nx, ny, nz, int(q[i]), alpha[i], beta[i], L,
                                           xp.ctypes.data)
            for b, d in enumerate(delta):
                ep[a, b, i] = ref.ref_edge_preserving(nx, ny, nz, int(q[i]), alpha[i],
                                                      beta[i], L, d, xp.ctypes.data)
    np.savez(HERE / "penalties.npz", dims=np.array([nx, ny, nz]), x=x, q=q,
             alpha=alpha, beta=beta, lam=lam, delta=delta, quadratic=quad,
             edge_preserving=ep)
    # rotation: 17 x 17 x 4 (square, as the reference requires)
    nxr, nzr = 17, 4
    vol = rng.uniform(0.0, 1.0, nxr * nxr * nzr).astype(np.float32)
    readings = 40
    tube_angles = np.fmod(np.arange(readings) * (360.0 / 24.0) + 3.7, 360.0)
    slice_index = np.array([2, 11, 23, 37], dtype=np.int64)
    out = {}
    for direction, tag in [(1, "rotating2fixed"), (-1, "fixed2rotating")]:
        v = vol.copy()
        ref.ref_rotate_slices(nxr, nxr, nzr, tube_angles.ctypes.data,
                              slice_index.ctypes.data, v.ctypes.data, direction)
        out[tag] = v
    np.savez(HERE / "rotation.npz", n=np.array([nxr, nzr, readings]), volume=vol,
             tube_angles=tube_angles, slice_index=slice_index, **out)
    print("wrote", HERE / "penalties.npz", HERE / "rotation.npz")


if __name__ == "__main__":
    main()
