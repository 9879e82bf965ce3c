"""Converged reference of a scan: the FP64 ICD (state_fp64) run until a sweep
lowers the cost by less than a relative tolerance, for the time-to-tolerance
measurement.

The converged cost does not depend on the start image or the sweep order
(the cost is convex with a unique minimiser), so the run starts from the GPU
filtered backprojection (ctbb_fbp) and uses the library's default schedule.
It records the FP64 cost after every sweep, an estimate of the distance
still to go (the last sweep decreases continued as a geometric series), and
the converged values of a seeded sample of voxels (rotated frame) for the
volume error in HU of a run measured against it.

  python -m ctbb_icd_amd.reference --config config4 --out ctbb_icd_amd/data/config4_seed1.npz
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np

from . import Recon, SSM, Scan, configs, init, set_verbose

SAMPLE = 1 << 18      # voxels whose converged values are kept
SAMPLE_SEED = 12345


def sample_indices(nvox: int, n: int = SAMPLE, seed: int = SAMPLE_SEED) -> np.ndarray:
    """The voxels (rotated frame, flat index) compared against the reference."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(nvox, size=min(n, nvox), replace=False))


def remaining_estimate(costs: list[float], k: int = 10) -> float:
    """Cost still to go after the last sweep: the last k sweep decreases,
    fitted as a geometric sequence and summed to infinity (0 if they do not
    shrink)."""
    d = -np.diff(np.asarray(costs, np.float64))[-k:]
    d = d[d > 0]
    if len(d) < 3:
        return float("nan")
    ratio = float(np.exp(np.polyfit(np.arange(len(d)), np.log(d), 1)[0]))
    if not ratio < 1.0:
        return float("inf")
    return float(d[-1] * ratio / (1.0 - ratio))


def converged(scan: Scan, ssm: SSM, raw, w, rel_tol: float = 1e-9, pre_rel: float = 1e-7,
              max_sweeps: int = 400, log=print) -> tuple[dict, np.ndarray]:
    """Converged cost and volume of one scan with a matrix already built:
    FP32 sweeps from the FBP image until a sweep lowers the running cost by
    less than pre_rel (the FP32 state's floor is near 1e-7), then the FP64
    state from that image until a sweep lowers the FP64 cost by less than
    rel_tol.  The minimiser is unique, so the FP32 leg only shortens the
    FP64 one.  Returns (summary, FP64 volume, rotated frame)."""
    t0 = time.perf_counter()
    x0 = np.maximum(ssm.fbp(raw, scan.slice_index), 0.0).astype(np.float32)
    r = Recon(ssm, raw, w, scan.slice_index, scan.opts())
    r.set_volume(x0)
    prev, n32 = r.cost_running(), 0
    while n32 < max_sweeps:
        r.iterate(1)
        n32 += 1
        c = r.cost_running()
        if n32 % 20 == 0:
            log(f"[reference] FP32 sweep {n32}: running cost {c:.12e}")
        if prev - c < pre_rel * c:
            break
        prev = c
    x32 = r.volume64()
    r.free()
    r = Recon(ssm, raw, w, scan.slice_index, scan.opts(state_fp64=1))
    r.set_volume64(x32)
    costs = [r.cost()]
    while len(costs) <= max_sweeps:
        r.iterate(1)
        costs.append(r.cost_running())
        rel = (costs[-2] - costs[-1]) / costs[-1]
        if (len(costs) - 1) % 10 == 0:
            log(f"[reference] FP64 sweep {len(costs) - 1}: cost {costs[-1]:.12e}, "
                f"decrease {rel:.2e}")
        if 0 <= rel < rel_tol:
            break
    final = r.cost()
    left = remaining_estimate(costs)
    x = r.volume64()
    r.free()
    out = dict(cost=final, fp32_sweeps=n32, fp64_sweeps=len(costs) - 1,
               last_decrease=(costs[-2] - costs[-1]) / costs[-1],
               remaining_estimate=left, remaining_rel=left / final,
               seconds=time.perf_counter() - t0)
    log(f"[reference] converged cost {final:.12e} ({n32} FP32 + {out['fp64_sweeps']} FP64 "
        f"sweeps, estimated remaining {out['remaining_rel']:.1e} relative, "
        f"{out['seconds']:.1f} s)")
    return out, x


def run(config: str, seed: int, minutes: float, rel_tol: float, log=print) -> dict:
    scan = Scan(configs.write_prm(config, Path("/tmp") / f"ref_{config}_{seed}.yaml",
                                  synthetic_seed=seed))
    g = scan.geom
    t0 = time.perf_counter()
    raw, w = scan.synthesize(use_gpu=True)
    ssm = SSM.build(g)
    x0 = np.maximum(ssm.fbp(raw, scan.slice_index), 0.0).astype(np.float32)
    r = Recon(ssm, raw, w, scan.slice_index, scan.opts(state_fp64=1))
    r.set_volume(x0)
    costs = [r.cost()]
    log(f"[reference] {config} seed {seed}: {g.nx}x{g.ny}x{g.nz}, FBP start cost "
        f"{costs[0]:.12e} ({time.perf_counter() - t0:.1f} s)")
    deadline = t0 + minutes * 60.0
    sweep_s = []
    while time.perf_counter() < deadline:
        t = time.perf_counter()
        st = r.iterate(1)
        sweep_s.append(st.sweep_ms * 1e-3)
        costs.append(r.cost_running())
        n = len(costs) - 1
        rel = (costs[-2] - costs[-1]) / costs[-1]
        if n % 5 == 0 or rel < rel_tol:
            log(f"[reference] sweep {n}: cost {costs[-1]:.12e}, decrease {rel:.2e} "
                f"({time.perf_counter() - t:.1f} s)")
        if 0 <= rel < rel_tol:
            break
    final = r.cost()          # fresh FP64 projection
    left = remaining_estimate(costs)
    x = r.volume64()
    idx = sample_indices(g.nvox)
    out = dict(config=config, seed=seed, sweeps=len(costs) - 1, cost=final,
               running_cost=costs[-1], last_decrease=(costs[-2] - costs[-1]) / costs[-1],
               remaining_estimate=left, remaining_rel=left / final,
               sweep_seconds_mean=float(np.mean(sweep_s)) if sweep_s else None,
               costs=costs, sample_seed=SAMPLE_SEED, sample_size=int(idx.size),
               tile=st.tile, group_spacing=st.group_spacing,
               wall_seconds=time.perf_counter() - t0)
    log(f"[reference] converged cost {final:.12e} after {out['sweeps']} sweeps; last "
        f"decrease {out['last_decrease']:.2e}, estimated remaining {left:.3e} "
        f"({out['remaining_rel']:.2e} relative)")
    r.free()
    ssm.free()
    return out, x[idx].astype(np.float32)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--config", default="config4")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--minutes", type=float, default=16.0, help="sweep time budget")
    ap.add_argument("--rel-tol", type=float, default=1e-9,
                    help="stop when a sweep lowers the cost by less than this (relative)")
    ap.add_argument("--out", required=True, help=".npz: summary (JSON) + sampled voxels")
    a = ap.parse_args()
    set_verbose(False)
    init(0)
    summary, values = run(a.config, a.seed, a.minutes, a.rel_tol,
                          log=lambda m: print(m, flush=True))
    np.savez_compressed(a.out, summary=json.dumps(summary), values=values)
    print(f"[reference] wrote {a.out}", flush=True)


if __name__ == "__main__":
    main()
