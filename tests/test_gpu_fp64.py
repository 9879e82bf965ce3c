"""The FP64 sweep state (state_fp64 = 1): volume, error sinogram, updates
and partial sums in FP64 like the reference's image and sinogram estimate
(src/icd_iteration.cpp:47-48,198-249).  It is the mode that produces
converged reference costs, so it is checked against the CPU FP64 oracle
replaying the same order, far tighter than the FP32 sweep's 1e-4:
quadratic penalty 1e-10 relative RMS, edge-preserving 1e-9 (the bisection
stops on a 1e-8 bracket, so a last-bit difference in a voxel's sums can
move that voxel's step by up to the bracket).  Also: the FP64 cost of the
FP64 state equals the oracle's cost of the same volume; FP64 runs the
single-column schedule to the reference order's converged cost."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(ct):
    ct.init(0)
    return ct


@pytest.mark.parametrize("name,tile,n_iter,tol", [
    ("tiny_ffs", 4, 3, 1e-9),        # edge-preserving, weighted, batched
    ("tiny", 1024, 3, 1e-10),        # quadratic, single columns
    ("config0", 8, 1, 1e-10),        # quadratic, weighted, batched
    ("clin_crop", 0, 2, 1e-9),       # clinical cone, default schedule
])
def test_fp64_state_matches_fp64_oracle(gpu, make_scan, orc, name, tile, n_iter, tol):
    s = make_scan(name)
    g = s.geom
    raw, w = s.synthesize(use_gpu=False)
    m = gpu.SSM.build(g)
    host = orc.ssm_build(g)
    r = gpu.Recon(m, raw, w, s.slice_index, s.opts(tile=tile, group_spacing=0, state_fp64=1))
    st = r.iterate(n_iter)
    xg = r.volume64()
    pen, lam, delta = s.run.penalty, s.run.lambda_, s.run.delta
    xo, _, nupd = orc.icd(g, host, raw, w, s.slice_index, pen, lam, delta, st.group_spacing,
                          n_iter, tile=st.tile, surrogate=True)
    assert st.voxel_updates == nupd
    rms = np.sqrt(np.mean((xg - xo) ** 2)) / np.sqrt(np.mean(xo ** 2))
    mx = np.abs(xg - xo).max()
    c_gpu = r.cost()
    c_orc = orc.cost(g, host, raw, w, s.slice_index, pen, lam, delta, xo)
    print(f"[fp64] {name} tile {st.tile} spacing {st.group_spacing}: rel RMS vs FP64 oracle "
          f"{rms:.2e}, max |dx| {mx:.2e}, cost {c_gpu:.14e} vs {c_orc:.14e} "
          f"(rel {abs(c_gpu - c_orc) / c_orc:.1e}), {st.sweep_ms:.1f} ms", flush=True)
    assert rms <= tol
    assert abs(c_gpu - c_orc) <= 1e-9 * c_orc
    # the running FP64 error sinogram is the cost's data term to FP64 rounding
    assert abs(r.cost_running() - c_gpu) <= 1e-9 * c_gpu
    r.free()
    m.free()


def _converge(r, tag, max_sweeps=400, rel=1e-9):
    """Sweeps until one lowers the FP64 cost by at most rel (relative)."""
    prev = r.cost()
    for it in range(max_sweeps):
        r.iterate(1)
        c = r.cost()
        if it % 10 == 9:
            print(f"[converge] {tag}: sweep {it + 1}, cost {c:.12e}", flush=True)
        if prev - c <= rel * c:
            return c, it + 1
        prev = c
    return c, max_sweeps


@pytest.mark.parametrize("name", ["tiny_ffs", "config0"])
def test_default_schedule_converges_to_reference_order_cost(gpu, make_scan, orc, name):
    """Converged costs do not depend on the update order (a convex cost with
    a unique minimiser): the library's default schedule (tile 4, batch
    surrogate), FP32 and FP64 state, run to convergence, against the
    reference order run to convergence (src/icd_iteration.cpp:157-185).
    tiny_ffs: the oracle's own reference order (raster columns, slices
    ascending), run here.  config0, where the oracle needs ~30 s per sweep
    and the GPU's single-column schedule minutes: the converged cost of the
    FP64 state with single-column batches -- the reference's raster column
    order, checked against the oracle per sweep in the FP64 parity test
    above -- from tests/golden/config0_converged.json
    (tests/golden/make_config0_converged.py), for the same inputs."""
    import json
    from pathlib import Path
    s = make_scan(name)
    g = s.geom
    raw, w = s.synthesize(use_gpu=False)
    m = gpu.SSM.build(g)
    pen, lam, delta = s.run.penalty, s.run.lambda_, s.run.delta
    if name == "tiny_ffs":
        host = orc.ssm_build(g)
        x, yhat, prev, c_ref = None, None, None, None
        for n_ref in range(1, 401):
            x, yhat, _ = orc.icd(g, host, raw, w, s.slice_index, pen, lam, delta, 0, 1,
                                 x=x, yhat=yhat)
            c_ref = orc.cost(g, host, raw, w, s.slice_index, pen, lam, delta, x)
            if prev is not None and prev - c_ref <= 1e-9 * c_ref:
                break
            prev = c_ref
    else:
        fx = json.loads((Path(__file__).parent / "golden" / "config0_converged.json")
                        .read_text())
        fp = fx["inputs"]
        assert float(np.sum(raw, dtype=np.float64)) == pytest.approx(fp["raw_sum"], rel=1e-12)
        assert float(np.sum(np.square(raw, dtype=np.float64))) == pytest.approx(
            fp["raw_sumsq"], rel=1e-12)
        assert float(np.sum(w, dtype=np.float64)) == pytest.approx(fp["w_sum"], rel=1e-12)
        c_ref, n_ref = fx["cost"], fx["sweeps"]
    out = {}
    for fp64 in (0, 1):
        r = gpu.Recon(m, raw, w, s.slice_index, s.opts(state_fp64=fp64))
        out[fp64] = _converge(r, f"default FP{64 if fp64 else 32}",
                              rel=1e-9 if name == "tiny_ffs" else 1e-8)
        r.free()
    print(f"[converged] {name}: reference order {c_ref:.12e} ({n_ref} sweeps); default "
          f"schedule FP32 {out[0][0]:.12e} ({out[0][1]} sweeps, rel "
          f"{abs(out[0][0] - c_ref) / c_ref:.2e}), FP64 {out[1][0]:.12e} ({out[1][1]} "
          f"sweeps, rel {abs(out[1][0] - c_ref) / c_ref:.2e})", flush=True)
    for fp64 in (0, 1):
        assert abs(out[fp64][0] - c_ref) <= 1e-5 * c_ref
    m.free()
