"""Write tests/golden/config0_converged.json: the converged cost of the
config0 scan (CPU-synthesised, seed 1) in the reference's update order --
raster columns one at a time (src/icd_iteration.cpp:157-185), FP64 image and
error sinogram like the reference's x and yhat (:47-48,198-249).

The GPU library runs that order as single-column batches in the FP64 state
(tile >= nx, state_fp64 = 1), whose sweeps match the CPU FP64 oracle in the
same order to ~1e-15 relative RMS (tests/test_gpu_fp64.py).  It is launch-
bound there (three launches per column and slice group) and needs minutes to
converge, so the converged cost is computed once, here, and the GPU test
compares the library's default schedule, run to convergence, against it.
The inputs are fingerprinted (sums of the raw data and weights) so a changed
generator fails the test instead of comparing different problems.

Run on a machine with the GPU:  python tests/golden/make_config0_converged.py
"""
import json
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
OUT = HERE / "config0_converged.json"


def fingerprint(raw, w):
    return {"raw_sum": float(np.sum(raw, dtype=np.float64)),
            "raw_sumsq": float(np.sum(np.square(raw, dtype=np.float64))),
            "w_sum": float(np.sum(w, dtype=np.float64)) if w is not None else None}


def main(rel=1e-9, max_sweeps=400, max_seconds=600.0):
    import ctbb_icd_amd as ct
    from ctbb_icd_amd import configs
    ct.set_verbose(False)
    ct.init(0)
    tmp = tempfile.TemporaryDirectory()
    s = ct.Scan(configs.write_prm("config0", Path(tmp.name) / "config0.yaml"))
    raw, w = s.synthesize(use_gpu=False)
    m = ct.SSM.build(s.geom)
    r = ct.Recon(m, raw, w, s.slice_index, s.opts(tile=1024, state_fp64=1))
    t0 = time.perf_counter()
    prev = r.cost()
    costs, n = [], 0
    while n < max_sweeps and time.perf_counter() - t0 < max_seconds:
        st = r.iterate(1)
        n += 1
        c = r.cost()
        costs.append(c)
        if n % 10 == 0:
            print(f"[config0 reference order] sweep {n}: cost {c:.13e} "
                  f"({time.perf_counter() - t0:.0f} s)", flush=True)
        if prev - c <= rel * c:
            break
        prev = c
    last = (costs[-2] - costs[-1]) / costs[-1] if len(costs) > 1 else None
    out = {"config": "config0", "seed": 1, "synthesis": "cpu",
           "order": "single-column batches (reference raster column order), FP64 state",
           "tile": st.tile, "group_spacing": st.group_spacing, "sweeps": n,
           "cost": costs[-1], "last_decrease_rel": last, "stop_rel": rel,
           "inputs": fingerprint(raw, w), "costs_every_10": costs[9::10]}
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}: cost {costs[-1]:.13e} after {n} sweeps (last decrease {last:.2e})")
    r.free()
    m.free()
    tmp.cleanup()


if __name__ == "__main__":
    main()
