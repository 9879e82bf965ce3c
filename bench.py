"""Benchmark of the ICD reconstruction hot path on one MI355X.

A step is one ICD iteration (a full sweep over every voxel of the volume
inside the acquisition FOV) of the clinical-size plain configuration:
512x512x96 volume, 64 x 736 detector with the in-plane flying focal spot
(2304 views/turn, 3 rotations), FP32 stored system matrix, edge-preserving
penalty, batches of tile 4 with the batch surrogate curvature.  Inputs (synthetic sinogram, SSM) are resident in HBM before the
timed region.  Prints one JSON line.

  value            voxel updates per second over the timed iterations
  roofline         the sweep kernel against HBM bandwidth: algorithmic bytes
                   per launch (per in-range SSM entry 4 B error read + 4 B
                   error write [+ 4 B weight], plus each FOV column's SSM
                   entries once per sweep) over its average launch time,
                   measured with HIP events on the reconstruction's stream
  cpu_baseline     the reference's OpenMP ICD structure (oracle) timed on
                   this host's cores on a band of rows of the same workload
  time_to_tolerance  the metric's second half (plain FP32 vs FP16 matrix +
                   learned warm start, config4), from its committed summary
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="config1")
    ap.add_argument("--tile", type=int, default=0, help="batch tile (0: library default)")
    ap.add_argument("--fp16", action="store_true", help="FP16 SSM (config 2 precision split)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the cpu_baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the cpu_baseline (0: 11, the reference's team)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a device copy of this size first (counter calibration)")
    return ap.parse_args()


def copy_probe(ct, mib: int) -> None:
    ms = ct.copy_probe(mib * 1024 * 1024)
    print(f"[bench] copy probe (ctbb_copy_kernel): {mib} MiB read + {mib} MiB "
          f"written in {ms:.3f} ms ({2 * mib / 1024 / ms * 1e3:.0f} GiB/s)", file=sys.stderr)


def measured_traffic(workload_key: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE runs, each scaled by
    what the counter reads for a copy of known size), or None."""
    f = ROOT / "profiles" / f"icd_sweep_{workload_key}_counters.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def time_to_tolerance():
    """Second half of the metric -- wall time to reach (1 + 1e-4) x the
    30-sweep plain FP32 cost, plain FP32 from zero vs FP16 matrix + learned
    warm start, config4 -- from the committed summary of its run
    (python -m ctbb_icd_amd.warmstart ..., about ten minutes, so not part of
    the default bench), or None."""
    import re
    f = ROOT / "profiles" / "warmstart_config4.txt"
    if not f.exists():
        return None
    m = re.search(r"# mean: plain ([0-9.]+) sweeps / ([0-9.]+) s, warm ([0-9.]+) / "
                  r"([0-9.]+) s -> ([0-9.]+)x", f.read_text())
    if not m:
        return None
    out = {"workload": "config4 (512x512x256, 6 turns, count weights), 2 held-out "
                       "phantoms", "tolerance": 1e-4,
           "plain_fp32_s": float(m.group(2)), "plain_sweeps": float(m.group(1)),
           "fp16_learned_warm_s": float(m.group(4)), "warm_sweeps": float(m.group(3)),
           "speedup": float(m.group(5)), "source": "profiles/warmstart_config4.txt"}
    # the non-learned comparison: GPU FBP start on the FP16 matrix, same phantoms
    f2 = re.search(r"# FBP start: ([0-9]+) sweeps / ([0-9.]+) s .* -> ([0-9.]+)x",
                   f.read_text())
    if f2:
        out["fp16_fbp_start"] = {"sweeps": int(f2.group(1)), "seconds": float(f2.group(2)),
                                 "speedup": float(f2.group(3))}
    return out


def cpu_baseline(scan, ssm, raw, w, target_s: float, threads: int):
    """Time the reference-structured OpenMP ICD (oracle) on a band of rows."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    g = scan.geom
    nx, ny = g.nx, g.ny
    y0 = ny // 2
    rows = 1
    elapsed, upd = 0.0, 0
    ncol = nx * ny
    while True:
        c0, c1 = y0 * nx, (y0 + rows) * nx
        cp_band, idx, val = ssm.download_cols(c0, c1)
        cp = np.zeros(ncol + 1, np.int64)
        cp[c0:c1 + 1] = cp_band
        cp[c1 + 1:] = cp_band[-1]
        t = time.perf_counter()
        _, _, upd = orc.icd_ref_omp(g, (cp, idx, val), raw, w, scan.slice_index,
                                    scan.run.penalty, scan.run.lambda_, scan.run.delta,
                                    1, threads, y0, y0 + rows)
        elapsed = time.perf_counter() - t
        if elapsed >= 0.3 * target_s or y0 + 2 * rows > ny:
            break
        rows = min(ny - y0, max(rows + 1, int(rows * target_s / max(elapsed, 1e-3))))
    return {
        "value": upd / elapsed, "unit": "voxel updates/s", "cores": threads,
        "kind": "port",
        "sample": f"{rows} row(s) of voxel columns (iy {y0}..{y0 + rows - 1}, all x, "
                  f"all {g.nz} slices), first iteration from zero, "
                  f"{upd} updates in {elapsed:.1f} s; SSM held in RAM",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import ctbb_icd_amd as ct
    from ctbb_icd_amd import configs
    ct.set_verbose(False)
    ct.init(0)
    if args.copy_probe_mib:
        copy_probe(ct, args.copy_probe_mib)
    tmp = tempfile.TemporaryDirectory()
    over = {"ssm_fp16": 1} if args.fp16 else {}
    scan = ct.Scan(configs.write_prm(args.config, Path(tmp.name) / "bench.yaml", **over))
    g = scan.geom
    fp16 = bool(args.fp16 or scan.run.ssm_fp16)      # --fp16 or the prm's ssm_fp16
    t0 = time.perf_counter()
    raw, w = scan.synthesize(use_gpu=True)
    t1 = time.perf_counter()
    ssm = ct.SSM.build(g)
    nnz, _, build_ms = ssm.info()
    spacing = ssm.group_spacing(scan.slice_index, max_conflict_fraction=0.01)
    print(f"[bench] {args.config}: synth {t1 - t0:.1f}s, SSM nnz {nnz} "
          f"({nnz / (g.nx * g.ny):.0f}/column) built in {build_ms:.0f} ms, "
          f"slice spacing {spacing}", file=sys.stderr, flush=True)
    cpu = None
    if args.cpu_seconds > 0:
        threads = args.cpu_threads or 11
        cpu = cpu_baseline(scan, ssm, raw, w, args.cpu_seconds, threads)
        print(f"[bench] cpu baseline: {cpu['value']:.3e} updates/s ({cpu['sample']})",
              file=sys.stderr, flush=True)
    if fp16:
        ssm.to_fp16()
    r = ct.Recon(ssm, raw, w, scan.slice_index, scan.opts(tile=args.tile))
    if args.warmup:
        r.iterate(args.warmup)
    t = time.perf_counter()
    st = r.iterate(args.steps)
    wall = time.perf_counter() - t
    weighted = w is not None
    entry_bytes = 6 if fp16 else 8
    steps = st.sweep_launches // 3                 # gather, solve, scatter per step
    visits = st.entry_visits / steps               # in-range entries per step
    # algorithmic bytes per launch (per in-range entry: 4 B error read in the
    # gather, 4 B error write [+ 4 B weight read] in the scatter -- the
    # weighted gather reads the kept w * e, the scatter reads w to update it;
    # each FOV column's SSM entries once per sweep, charged to the gather)
    gather_bytes = visits * 4 + st.ssm_entries * entry_bytes * args.steps / steps
    scatter_bytes = visits * (8 if weighted else 4)
    # the solve: each voxel update reads and writes its FP32 voxel (the
    # partial sums it adds up are this design's intermediates, not counted)
    solve_bytes = st.voxel_updates / steps * 8
    kernels = {
        "icd_gather_kernel": (gather_bytes, st.gather_ms_avg),
        "icd_solve_kernel": (solve_bytes, st.solve_ms_avg),
        "icd_scatter_kernel": (scatter_bytes, st.scatter_ms_avg),
    }
    dom = max(kernels, key=lambda k: kernels[k][1])
    dom_bytes, dom_ms = kernels[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    value = st.voxel_updates / wall
    cost = r.cost()
    out = {
        "metric": "voxel updates/s",
        "value": value,
        "unit": "voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-ssm/f32" if fp16 else "f32",
        "data": "synthetic (seeded ellipsoid phantom, analytic line integrals, Poisson noise)",
        "config": {
            "workload": f"{args.config}: plain ICD {g.nx}x{g.ny}x{g.nz}, "
                        f"{g.n_rows}x{g.n_channels} detector, {scan.run.num_views_per_turn} "
                        f"views/turn (in-plane FFS), {g.readings} views, "
                        f"{'FP16' if fp16 else 'FP32'} SSM, "
                        f"{'FP64 resync every %d, ' % scan.run.resync_every if scan.run.resync_every else ''}"
                        f"{'edge-preserving' if scan.run.penalty else 'quadratic'} penalty",
            "voxels": g.nvox, "ssm_nnz": nnz, "slice_spacing": spacing,
            "tile": st.tile, "batch_curvature": "surrogate",
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm", "kernel": dom,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (measured_traffic(args.config, dom)
                        if not fp16 and args.tile == 0 else None),
            "algorithmic_bytes_per_launch": dom_bytes,
            "avg_launch_ms": dom_ms,
        },
        "cpu_baseline": cpu,
        "time_to_tolerance": time_to_tolerance(),
        "extra": {
            "kernels": {k: {"algorithmic_bytes_per_launch": b, "avg_launch_ms": ms,
                            "GB_per_s": b / (ms * 1e-3) / 1e9 if ms else None}
                        for k, (b, ms) in kernels.items()},
            "sweep_GB_per_s_algorithmic": (st.entry_visits * (12 if weighted else 8)
                                           + st.ssm_entries * entry_bytes * args.steps)
                                          / (st.sweep_ms * 1e-3) / 1e9,
            "launches_per_iteration": st.sweep_launches / args.steps,
            "sweep_ms_per_iteration": st.sweep_ms / args.steps,
            "entry_visits_per_iteration": st.entry_visits / args.steps,
            "ssm_build_ms": build_ms,
            "ssm_build_rays_per_s": g.n_views_sm * g.n_channels * g.n_rows / (build_ms * 1e-3),
            "cost_after": cost,
        },
    }
    print(json.dumps(out))
    tmp.cleanup()


if __name__ == "__main__":
    main()
