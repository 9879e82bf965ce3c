"""Benchmark of the ICD reconstruction hot path on one MI355X.

Workload (default): the full-size scan -- 512x512x256 volume, 64 x 736
detector with the in-plane flying focal spot (2304 views/turn, 6 turns,
13824 views), seeded ellipsoid phantom with a lung-like region, Poisson
noise at I0 = 1e5 with count weights, edge-preserving penalty.  A step is
one ICD iteration: a full sweep over every voxel inside the acquisition FOV
(library default schedule: tile 4, batch surrogate, FP32 state, FP32
matrix).  Inputs (synthetic sinogram, matrix) are resident in HBM before the
timed region.  Prints one JSON line.

  value              voxel updates per second over the K timed iterations
  roofline           the sweep kernel with the longest launches against HBM
                     bandwidth: algorithmic bytes per launch (per in-range
                     matrix entry 4 B error read [gather] or 4 B error write
                     + 4 B weight read [scatter], plus each FOV column's
                     matrix entries once per sweep) over its average launch
                     time, measured with HIP events on the reconstruction's
                     stream; traffic (HBM bytes per launch) from the
                     committed rocprofv3 counters, and the HBM rate it gives
                     at the measured launch time
  cpu_baseline       the reference's OpenMP ICD structure (the oracle's
                     restatement) on this host, 11 threads (the reference's
                     team), on a band of rows of the same workload
  time_to_tolerance  wall time to a cost within 1e-4 relative of the
                     converged FP64 cost of this scan (ctbb_icd_amd/data):
                     plain FP32 ICD from zero (the run the timed iterations
                     belong to, continued) against the FP16 matrix with the
                     learned warm start (ctbb_icd_amd/data network), each
                     with its volume error in HU against the converged run
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6
MU_WATER = 0.193          # 1/cm; 1 HU = MU_WATER / 1000
DATA = ROOT / "ctbb_icd_amd" / "data"
METRIC = ("voxel updates/s; time to cost ≤1e-4 rel., plain vs FP16-SSM+learned warm "
          "start")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="config4")
    ap.add_argument("--tile", type=int, default=0, help="batch tile (0: library default)")
    ap.add_argument("--fp16", action="store_true", help="FP16 matrix for the timed sweeps")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of each cpu_baseline sample (0 = skip)")
    ap.add_argument("--ttt", type=int, default=1,
                    help="1: measure time to tolerance (needs the committed reference)")
    ap.add_argument("--ttt-plain-cap", type=float, default=400.0,
                    help="plain leg: stop after this many seconds of sweeps")
    ap.add_argument("--ttt-warm-cap", type=float, default=150.0,
                    help="warm leg: stop after this many seconds")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a device copy of this size first (counter calibration)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def ssm_build_flop(workload_key: str):
    """FP64 FLOP of one matrix build (both passes) from the committed
    rocprofv3 counter summary of this workload, or None."""
    f = ROOT / "profiles" / f"ssm_build_{workload_key}_counters.txt"
    if not f.exists():
        return None
    total = 0.0
    for line in f.read_text().splitlines():
        parts = line.split()
        if parts and parts[0].startswith("ssm_build_kernel<") and len(parts) >= 6:
            total += float(parts[5])
    return total or None


def cpu_topology() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return "lscpu unavailable"
    keep = ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core")
    return "; ".join(line.split(":", 1)[0].strip() + ": " + line.split(":", 1)[1].strip()
                     for line in out.splitlines() if line.split(":", 1)[0].strip() in keep)


def cpu_baseline(scan, ssm, raw, w, target_s: float, threads: int, streamed: bool = False):
    """Time the reference-structured OpenMP ICD (oracle) on a band of rows,
    the matrix held in RAM; streamed: also on the same band with the matrix
    read from a matrix.bin file of those rows every iteration, as the
    reference reads it (src/icd_iteration.cpp:164-179)."""
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
    out = {
        "value": upd / elapsed, "unit": "voxel updates/s", "cores": threads,
        "kind": "port",
        "sample": f"{rows} row(s) of voxel columns (iy {y0}..{y0 + rows - 1}, all x, "
                  f"all {g.nz} slices), first iteration from zero, "
                  f"{upd} updates in {elapsed:.1f} s; matrix held in RAM",
    }
    if streamed:
        with tempfile.TemporaryDirectory() as d:
            f = Path(d) / "band_matrix.bin"
            orc.write_band_file(f, g, (cp, idx, val), y0, y0 + rows)
            t = time.perf_counter()
            _, _, upd_f = orc.icd_ref_omp_file(g, f, raw, w, scan.slice_index,
                                               scan.run.penalty, scan.run.lambda_,
                                               scan.run.delta, 1, threads, y0, y0 + rows)
            el_f = time.perf_counter() - t
            out["streamed"] = {
                "value": upd_f / el_f, "unit": "voxel updates/s",
                "sample": f"the same rows, matrix read from a matrix.bin file of them "
                          f"({f.stat().st_size / 2**20:.0f} MiB, just written: page "
                          f"cache) during the sweep, {upd_f} updates in {el_f:.1f} s"}
    return out


def cpu_ssm_build(g, target_s: float):
    """Time the reference's single-threaded ray-driven matrix build (the
    oracle's restatement of src/generate_system_matrix_ffs.cpp:104-230) on a
    sample of SSM views spread over the window; rays/s extrapolates to the
    whole matrix (every view traces the same n_channels x n_rows rays, each
    across the whole grid)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    n = g.n_views_sm
    per_view = g.n_channels * g.n_rows
    seen, elapsed, entries = set(), 0.0, 0
    for i in range(n):
        v = int(((i + 0.5) * 0.6180339887498949) % 1.0 * n)
        if v in seen:
            continue
        seen.add(v)
        _, idx, _, fill_s = orc.ssm_build_views(g, v, v + 1)
        elapsed += fill_s
        entries += len(idx)
        if elapsed >= target_s:
            break
    rate = len(seen) * per_view / elapsed
    return {
        "value": rate, "unit": "rays/s", "cores": 1, "kind": "port",
        "sample": f"{len(seen)} of {n} views (golden-ratio spread), "
                  f"{len(seen) * per_view} rays, {entries} entries written in "
                  f"{elapsed:.1f} s by the pass that traces and writes (the counting "
                  f"pass that sizes the arrays untimed); whole matrix extrapolated "
                  f"to {n * per_view / rate:.0f} s",
    }


def load_reference(config: str, seed: int):
    f = DATA / f"{config}_seed{seed}.npz"
    if not f.exists():
        return None, None, f
    z = np.load(f, allow_pickle=False)
    return json.loads(str(z["summary"])), z["values"].astype(np.float64), f


def rmse_hu(x, ref_values, idx, nxy, nz, ends=4):
    """Volume error against the converged reference over its sampled voxels,
    in HU (1 HU = mu_water / 1000): all of them, and the interior slices
    (the `ends` slices at either end of the volume, which absorb attenuation
    from outside the reconstructed slab and converge slowest, left out)."""
    d = x[idx].astype(np.float64) - ref_values
    k = idx // nxy
    inner = (k >= ends) & (k < nz - ends)
    hu = MU_WATER / 1000.0
    return {"all": float(np.sqrt(np.mean(d ** 2)) / hu),
            "interior": float(np.sqrt(np.mean(d[inner] ** 2)) / hu),
            "interior_slices": [ends, nz - ends],
            "end_slices_max_abs": float(np.abs(d[~inner]).max() / hu) if (~inner).any()
            else None}


def plain_to_tolerance(ct, r, target, sweep_wall, cap_s, n):
    """Continue the plain run (n sweeps so far, their wall times in
    sweep_wall) one sweep at a time until its FP64 cost is at or below target
    (the running cost from the kept error sinogram screens, the FP64
    projection confirms; neither is timed)."""
    reached, cost = False, None
    while sum(sweep_wall) < cap_s:
        run = r.cost_running()
        if run <= target * (1 + 1e-6):
            cost = r.cost()
            if cost <= target:
                reached = True
                break
        t = time.perf_counter()
        r.iterate(1)
        sweep_wall.append(time.perf_counter() - t)
        n += 1
        if n % 10 == 0:
            log(f"plain sweep {n}: running cost {run:.10e} (target {target:.10e}), "
                f"{sum(sweep_wall):.1f} s")
    if cost is None:
        cost = r.cost()
        reached = cost <= target
    return dict(seconds=float(sum(sweep_wall)), sweeps=n, reached=bool(reached), cost=cost)


def warm_to_tolerance(ct, scan, raw, w, target, cap_s, ref_values, idx, ssm32):
    """FP16 matrix + learned warm start to the target.  The time includes the
    early sweep, both rotations, the network, the FP64 projection of the
    start and the acceptance check, and the sweeps after it; the matrix build
    and conversion (once per scan geometry, as for the plain run) and the
    target checks (the running cost screens, the FP32 matrix's FP64 cost
    confirms) are not timed."""
    from ctbb_icd_amd import warmstart
    g = scan.geom
    netf = DATA / "unet_config4.npz"
    if not netf.exists():
        return dict(error=f"{netf.name} missing")
    net, meta = warmstart.load_params(netf, 8, g.ny, g.nx)
    opts = scan.opts()
    t0 = time.perf_counter()
    m16 = ct.SSM.build(g)
    m16.to_fp16()
    setup_s = time.perf_counter() - t0
    r = ct.Recon(m16, raw, w, scan.slice_index, opts)
    t1 = time.perf_counter()
    marks = [t1]
    r.iterate(1)
    marks.append(time.perf_counter())
    x_plain = r.volume()
    fixed = warmstart.to_fixed(scan, x_plain)
    pred = warmstart.predict_fixed(net, fixed)
    x_warm = warmstart.to_rotating(scan, pred)
    marks.append(time.perf_counter())
    c_plain = r.cost_running()
    r.set_volume(x_warm)
    c_warm = r.cost_running()
    accepted = c_warm <= c_plain
    if not accepted:
        r.set_volume(x_plain)
    marks.append(time.perf_counter())
    # the FP16 matrix, like the FP32 one, is set up once per scan geometry:
    # its build is reported, not counted
    spent = marks[-1] - t1
    sweeps = 1
    reached, cost = False, None
    while spent < cap_s:
        run = r.cost_running()
        if run <= target * (1 + 1e-4):
            cost = ct.cost(ssm32, raw, w, scan.slice_index, opts, r.volume())
            if cost <= target:
                reached = True
                break
        t = time.perf_counter()
        r.iterate(1)
        spent += time.perf_counter() - t
        sweeps += 1
        if sweeps % 10 == 0:
            log(f"warm sweep {sweeps}: running cost {run:.10e}, {spent:.1f} s")
    if cost is None:
        cost = ct.cost(ssm32, raw, w, scan.slice_index, opts, r.volume())
        reached = cost <= target
    x = r.volume()
    out = dict(seconds=float(spent), sweeps=sweeps, reached=bool(reached), cost=cost,
               accepted=bool(accepted), start_cost_plain=c_plain, start_cost_warm=c_warm,
               phases_s=dict(fp16_matrix=setup_s, early_sweep=marks[1] - marks[0],
                             network_and_rotations=marks[2] - marks[1],
                             projection_and_check=marks[3] - marks[2]),
               rmse_hu=rmse_hu(x, ref_values, idx, g.nx * g.ny, g.nz),
               network=dict(trained_on=meta.get("train_config"),
                            phantoms=meta.get("phantoms"),
                            target_its=meta.get("target_its"),
                            target_start=meta.get("target_start")))
    r.free()
    m16.free()
    net.free()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import ctbb_icd_amd as ct
    from ctbb_icd_amd import configs
    from ctbb_icd_amd.reference import sample_indices
    ct.set_verbose(False)
    ct.init(0)
    if args.copy_probe_mib:
        ms = ct.copy_probe(args.copy_probe_mib * 1024 * 1024)
        log(f"copy probe (ctbb_copy_kernel): {args.copy_probe_mib} MiB read + "
            f"{args.copy_probe_mib} MiB written in {ms:.3f} ms")
    tmp = tempfile.TemporaryDirectory()
    scan = ct.Scan(configs.write_prm(args.config, Path(tmp.name) / "bench.yaml"))
    g = scan.geom
    fp16 = bool(args.fp16 or scan.run.ssm_fp16)
    t0 = time.perf_counter()
    raw, w = scan.synthesize(use_gpu=True)
    t1 = time.perf_counter()
    ssm = ct.SSM.build(g)
    nnz, _, build_ms = ssm.info()
    rays = g.n_views_sm * g.n_channels * g.n_rows
    log(f"{args.config}: synth {t1 - t0:.1f}s, SSM nnz {nnz} ({nnz / (g.nx * g.ny):.0f}/column) "
        f"built in {build_ms:.0f} ms")
    cpu = cpu_all = cpu_build = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(scan, ssm, raw, w, args.cpu_seconds, 11, streamed=True)
        log(f"cpu baseline (11 threads): {cpu['value']:.3e} updates/s ({cpu['sample']})")
        log(f"cpu baseline (11 threads, matrix streamed from a file): "
            f"{cpu['streamed']['value']:.3e} updates/s ({cpu['streamed']['sample']})")
        n_all = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        cpu_all = cpu_baseline(scan, ssm, raw, w, args.cpu_seconds, n_all)
        cpu_all["topology"] = cpu_topology()
        log(f"cpu baseline ({n_all} threads): {cpu_all['value']:.3e} updates/s")
        cpu_build = cpu_ssm_build(g, 0.5 * args.cpu_seconds)
        log(f"cpu matrix build (1 thread): {cpu_build['value']:.3e} rays/s "
            f"({cpu_build['sample']})")
    m_run = ssm
    if fp16:
        m_run = ct.SSM.build(g)
        m_run.to_fp16()
    r = ct.Recon(m_run, raw, w, scan.slice_index, scan.opts(tile=args.tile))
    sweep_wall = []
    for _ in range(args.warmup):
        t = time.perf_counter()
        r.iterate(1)
        sweep_wall.append(time.perf_counter() - t)
    t = time.perf_counter()
    st = r.iterate(args.steps)
    wall = time.perf_counter() - t
    sweep_wall.append(wall)
    weighted = w is not None
    entry_bytes = 6 if fp16 else 8
    steps = st.sweep_launches // 3                 # gather, solve, scatter per step
    visits = st.entry_visits / steps               # in-range entries per step
    # algorithmic bytes per launch (per in-range entry: 4 B error read in the
    # gather, 4 B error write [+ 4 B weight read] in the scatter -- the
    # weighted gather reads the kept w * e, the scatter reads w to update it;
    # each FOV column's matrix entries once per sweep, charged to the gather)
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
    log(f"timed: {args.steps} sweeps in {wall:.2f} s, {value:.3e} updates/s")

    ttt = None
    if args.ttt:
        ref, ref_values, ref_file = load_reference(args.config, 1)
        if ref is None:
            ttt = {"error": f"no converged reference {ref_file.name}"}
        else:
            idx = sample_indices(g.nvox, ref["sample_size"], ref["sample_seed"])
            target = ref["cost"] * (1 + 1e-4)
            plain = plain_to_tolerance(ct, r, target, sweep_wall, args.ttt_plain_cap,
                                       args.warmup + args.steps)
            plain["rmse_hu"] = rmse_hu(r.volume(), ref_values, idx, g.nx * g.ny, g.nz)
            log(f"plain FP32: {plain}")
            r.free()
            warm = warm_to_tolerance(ct, scan, raw, w, target, args.ttt_warm_cap, ref_values,
                                     idx, ssm)
            log(f"FP16 + learned start: {warm}")
            ttt = {
                "source": "measured",
                "workload": f"{args.config} (seed 1): plain FP32 from zero vs FP16 matrix + "
                            "learned warm start, library default schedule",
                "tolerance": 1e-4,
                "reference": {"cost": ref["cost"], "sweeps": ref["sweeps"],
                              "state": "FP64", "remaining_rel": ref["remaining_rel"],
                              "file": str(ref_file.relative_to(ROOT))},
                "target_cost": target,
                "plain_fp32": plain,
                "fp16_learned": warm,
                "speedup": (plain["seconds"] / warm["seconds"]
                            if "seconds" in warm and warm["seconds"] > 0 else None),
                "speedup_is_lower_bound": bool(not plain["reached"]),
            }
    if r.h:
        cost = r.cost()
        r.free()
    else:
        cost = (ttt or {}).get("plain_fp32", {}).get("cost")
    # measured HBM bytes per launch of the dominant kernel (committed counter
    # summary of this workload at the default schedule and FP32 matrix)
    traffic = (measured_traffic(args.config, dom)
               if not fp16 and args.tile == 0 else None)
    build_bytes = nnz * 8 + g.nx * g.ny * g.n_views_sm * 4
    build_flop = ssm_build_flop(args.config)
    out = {
        "metric": METRIC,
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
        "data": "synthetic (seeded ellipsoid phantom with lung texture, analytic line "
                "integrals, Poisson noise, count weights)",
        "config": {
            "workload": f"{args.config}: plain ICD {g.nx}x{g.ny}x{g.nz}, "
                        f"{g.n_rows}x{g.n_channels} detector, {scan.run.num_views_per_turn} "
                        f"views/turn (in-plane FFS), {g.readings} views, "
                        f"{'FP16' if fp16 else 'FP32'} matrix, "
                        f"{'edge-preserving' if scan.run.penalty else 'quadratic'} penalty, "
                        f"{'count weights' if weighted else 'unweighted'}",
            "voxels": g.nvox, "ssm_nnz": nnz, "slice_spacing": st.group_spacing,
            "tile": st.tile, "batch_curvature": "surrogate", "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm", "kernel": dom,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_GB_per_s": (traffic / (dom_ms * 1e-3) / 1e9
                                 if traffic and dom_ms else None),
            "algorithmic_bytes_per_launch": dom_bytes,
            "avg_launch_ms": dom_ms,
        },
        "cpu_baseline": cpu,
        "time_to_tolerance": ttt,
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
            "ssm_build": {
                "ms": build_ms, "rays": rays, "rays_per_s": rays / (build_ms * 1e-3),
                "bytes_written": build_bytes,
                "hbm_write_frac": build_bytes / (build_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "fp64_flop": build_flop,
                "fp64_tflops": build_flop / (build_ms * 1e-3) / 1e12 if build_flop else None,
                "fp64_valu_frac": (build_flop / (build_ms * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS
                                   if build_flop else None),
                "fp64_flop_source": "profiles/ssm_build_%s_counters.txt" % args.config
                                    if build_flop else None,
                "cpu_baseline": cpu_build,
                "gpu_over_cpu": (rays / (build_ms * 1e-3) / cpu_build["value"]
                                 if cpu_build else None),
            },
            "cpu_baseline_all_cores": cpu_all,
            "cost_after": cost,
        },
    }
    print(json.dumps(out))
    tmp.cleanup()


if __name__ == "__main__":
    main()
