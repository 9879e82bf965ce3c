"""Learned warm start for the ICD reconstruction.

The start image of initialize_recon_volume (src/initialize_recon_volume.cpp:
51-76,180-183: zeros or a file) is replaced by the output of a 2.5D U-Net
(include/ctbb_unet.h) applied to the image after a few plain ICD
iterations.  The network works in the fixed frame: the rotated slices are
de-rotated (ctbb_rotate_slices, rotating -> fixed), stacked as (k-1, k, k+1)
with the end slices repeated, scaled by 1/mu_water, mapped, scaled back,
clamped at 0 and rotated back.  The ICD cost decides: a warm start whose
FP64 cost exceeds that of the plain image it came from is rejected and the
plain image is used instead.

Training pairs come from the GPU ICD on seeded phantoms of one scan
geometry: (image after `n_early` iterations from zero, image after
`n_target` iterations).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import Recon, SSM, Scan, UNet, current_device, init, rotate_slices

MU_WATER = 0.193   # 1/cm; inputs and outputs are scaled by 1/MU_WATER
MU_MAX = 1.0       # 1/cm; network inputs and training images are clipped to [0, MU_MAX]


def slice_stacks(vol: np.ndarray) -> np.ndarray:
    """[nz, ny, nx] -> [nz, 3, ny, nx]: slices (k-1, k, k+1), ends repeated."""
    prev = np.concatenate([vol[:1], vol[:-1]])
    nxt = np.concatenate([vol[1:], vol[-1:]])
    return np.stack([prev, vol, nxt], axis=1)


def to_fixed(scan: Scan, x_rot: np.ndarray) -> np.ndarray:
    g = scan.geom
    v = rotate_slices(g, scan.tube_angles, scan.slice_index, x_rot, +1)
    return v.reshape(g.nz, g.ny, g.nx)


def to_rotating(scan: Scan, vol: np.ndarray) -> np.ndarray:
    g = scan.geom
    return rotate_slices(g, scan.tube_angles, scan.slice_index,
                         np.ascontiguousarray(vol, np.float32).reshape(-1), -1)


def predict_fixed(net: UNet, vol: np.ndarray) -> np.ndarray:
    """Network output for every slice of a fixed-frame volume [nz, ny, nx]
    (inputs clipped to [0, MU_MAX] as in training)."""
    inp = (slice_stacks(np.clip(vol, 0.0, MU_MAX)) / MU_WATER).astype(np.float32)
    out = np.empty(vol.shape, np.float32)
    for k0 in range(0, vol.shape[0], net.max_batch):
        k1 = min(k0 + net.max_batch, vol.shape[0])
        out[k0:k1] = net.infer(inp[k0:k1])
    return np.maximum(out * MU_WATER, 0.0)


@dataclass
class Pair:
    seed: int
    early: np.ndarray     # fixed frame [nz, ny, nx]
    target: np.ndarray


def generate_pairs(make_scan, seeds, ssm: SSM, n_early: int, n_target: int,
                   log=print, target_start: str = "zero", workers: int = 1) -> list[Pair]:
    """make_scan(seed) -> Scan of the shared geometry with that phantom.
    Input: the image after n_early sweeps from zero.  Target: the image after
    n_target sweeps from zero (target_start "zero", the early sweeps
    included) or from the GPU FBP image ("fbp": closer to convergence for
    the same number of sweeps).  workers > 1: that many phantoms at a time,
    each from its own host thread on its own reconstruction (and so its own
    HIP stream), which keeps the GPU busy where one small scan's launches
    cannot; every run is deterministic, so the pairs are those of one
    phantom at a time, returned in seed order."""
    def one(s):
        t = time.perf_counter()
        scan = make_scan(s)
        raw, w = scan.synthesize(use_gpu=True)
        r = Recon(ssm, raw, w, scan.slice_index, scan.opts())
        r.iterate(n_early)
        early = to_fixed(scan, r.volume())
        if target_start == "fbp":
            r.set_volume(np.maximum(ssm.fbp(raw, scan.slice_index), 0.0).astype(np.float32))
            r.iterate(n_target)
        else:
            r.iterate(n_target - n_early)
        target = to_fixed(scan, r.volume())
        r.free()
        log(f"[warmstart] pair seed {s}: {time.perf_counter() - t:.1f} s")
        return Pair(s, early, target)

    seeds = list(seeds)
    if workers <= 1:
        return [one(s) for s in seeds]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=workers, initializer=init,
                            initargs=(current_device(),)) as ex:
        return list(ex.map(one, seeds))


def save_params(path, net: UNet, meta: dict) -> None:
    """The trained network as shipped: float16 parameters plus its shape and
    training description (npz)."""
    import json
    np.savez_compressed(path, params=net.params().astype(np.float16),
                        meta=json.dumps(dict(meta, levels=net.levels, base=net.base)))


def load_params(path, max_batch: int, height: int, width: int) -> tuple[UNet, dict]:
    import json
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    net = UNet(meta["levels"], meta["base"], max_batch, height, width, seed=11)
    net.set_params(z["params"].astype(np.float32))
    return net, meta



def train(net: UNet, pairs: list[Pair], epochs: int, lr: float, seed: int = 0,
          log=print, drop_ends: int = 2) -> list[float]:
    """Adam on the mean squared error (scaled units), slices shuffled; cosine
    learning-rate decay to lr/20.  The `drop_ends` slices at either end of a
    volume (which absorb attenuation from outside the reconstructed slab) are
    not trained on, and images are clipped to [0, MU_MAX] so that the few
    poorly determined voxels of weakly regularised runs do not dominate the
    loss.  Returns the mean loss of every epoch."""
    rng = np.random.default_rng(seed)

    def clip(v):
        return np.clip(v, 0.0, MU_MAX)

    sl = slice(drop_ends, -drop_ends if drop_ends else None)
    inp = np.concatenate([slice_stacks(clip(p.early))[sl] for p in pairs]) / MU_WATER
    tgt = np.concatenate([clip(p.target)[sl] for p in pairs]) / MU_WATER
    inp = inp.astype(np.float32)
    tgt = tgt.astype(np.float32)
    n = inp.shape[0]
    steps = epochs * ((n + net.max_batch - 1) // net.max_batch)
    hist, k = [], 0
    for ep in range(epochs):
        order = rng.permutation(n)
        losses = []
        for b0 in range(0, n, net.max_batch):
            sel = order[b0:b0 + net.max_batch]
            losses.append(net.grad(inp[sel], tgt[sel]))
            frac = k / max(1, steps - 1)
            net.adam_step(lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * frac))))
            k += 1
        hist.append(float(np.mean(losses)))
        log(f"[warmstart] epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e}")
    return hist


@dataclass
class Run:
    """One reconstruction to a cost target."""
    start: str
    iterations: int = 0
    seconds: float = 0.0
    reached: bool = False
    costs: list = field(default_factory=list)
    accepted: bool | None = None    # warm starts: passed the cost check
    running: list = field(default_factory=list)   # running cost after every sweep
    volume: np.ndarray | None = None               # final volume (keep=True)


def run_to_target(r: Recon, target: float, max_iter: int, t0: float, run: Run,
                  log=print, cost_fn=None, screen: float = 0.0, keep: bool = False) -> Run:
    """Iterate until the FP64 cost is <= target (checked after every
    iteration; the cost evaluation is not counted in the time).  cost_fn(r)
    evaluates the cost (default: the reconstruction's own).  screen > 0:
    the FP64 cost is evaluated only once the running cost (kept error
    sinogram, no projection) is within target * (1 + screen).  keep: the
    final volume is kept in run.volume."""
    cost_fn = cost_fn or (lambda rr: rr.cost())
    spent = time.perf_counter() - t0
    while run.iterations < max_iter:
        t = time.perf_counter()
        r.iterate(1)
        spent += time.perf_counter() - t
        run.iterations += 1
        if screen > 0:
            run.running.append(r.cost_running())
            if run.iterations % 10 == 0:
                log(f"[warmstart] {run.start} iteration {run.iterations}: running cost "
                    f"{run.running[-1]:.10e} ({spent:.2f} s)")
            if run.running[-1] > target * (1 + screen):
                continue
        c = cost_fn(r)
        run.costs.append(c)
        if screen <= 0:
            log(f"[warmstart] {run.start} iteration {run.iterations}: cost {c:.10e} "
                f"({spent:.2f} s)")
        if c <= target:
            run.reached = True
            break
    run.seconds = spent
    if keep:
        run.volume = r.volume64()
    log(f"[warmstart] {run.start}: {run.iterations} iterations, {run.seconds:.2f} s, "
        f"reached={run.reached}")
    return run


def plain_run(scan: Scan, ssm: SSM, raw, w, target: float, max_iter: int,
              log=print, screen: float = 0.0, keep: bool = False) -> Run:
    r = Recon(ssm, raw, w, scan.slice_index, scan.opts())
    try:
        return run_to_target(r, target, max_iter, time.perf_counter(), Run("plain"), log,
                             screen=screen, keep=keep)
    finally:
        r.free()


def warm_run(scan: Scan, ssm: SSM, raw, w, net: UNet, n_early: int, target: float,
             max_iter: int, log=print, cost_fn=None, screen: float = 0.0,
             keep: bool = False) -> Run:
    """n_early plain iterations, the network, the cost check, then ICD to the
    target.  The time includes the early iterations, the rotations, the
    network, the projection of the start image and the two cost evaluations
    of the check (from the kept error sinogram: the plain image's after its
    last sweep, the learned start's right after its FP64 projection); the
    iteration count includes the early iterations.  Target checks (cost_fn,
    default the FP64 cost) are not timed."""
    run = Run("warm")
    r = Recon(ssm, raw, w, scan.slice_index, scan.opts())
    try:
        t0 = time.perf_counter()
        marks = [t0]
        r.iterate(n_early)
        marks.append(time.perf_counter())
        x_plain = r.volume()
        fixed = to_fixed(scan, x_plain)
        marks.append(time.perf_counter())
        pred = predict_fixed(net, fixed)
        marks.append(time.perf_counter())
        x_warm = to_rotating(scan, pred)
        marks.append(time.perf_counter())
        c_plain = r.cost_running()
        r.set_volume(x_warm)
        c_warm = r.cost_running()
        marks.append(time.perf_counter())
        run.accepted = bool(c_warm <= c_plain)
        if not run.accepted:
            r.set_volume(x_plain)
        run.iterations = n_early
        spent = time.perf_counter() - t0
        ph = np.diff(marks)
        log(f"[warmstart] start cost plain {c_plain:.8e} warm {c_warm:.8e} -> "
            f"{'accepted' if run.accepted else 'rejected'} ({spent:.2f} s: early "
            f"{ph[0]:.2f}, to fixed {ph[1]:.2f}, network {ph[2]:.2f}, to rotating "
            f"{ph[3]:.2f}, projection + checks {ph[4]:.2f})")
        run.costs = [(cost_fn or (lambda rr: rr.cost()))(r)]
        if run.costs[0] <= target:
            run.reached = True
            run.seconds = spent
            if keep:
                run.volume = r.volume64()
            return run
        return run_to_target(r, target, max_iter, time.perf_counter() - spent, run, log,
                             cost_fn, screen, keep)
    finally:
        r.free()


def fbp_run(scan: Scan, ssm: SSM, raw, w, target: float, max_iter: int, log=print,
            cost_fn=None, screen: float = 0.0, keep: bool = False) -> Run:
    """ICD from the GPU filtered backprojection start (ctbb_fbp, clamped at
    0); the time includes the FBP and the projection of the start image."""
    run = Run("fbp")
    r = Recon(ssm, raw, w, scan.slice_index, scan.opts())
    try:
        t0 = time.perf_counter()
        r.set_volume(np.maximum(ssm.fbp(raw, scan.slice_index), 0.0).astype(np.float32))
        spent = time.perf_counter() - t0
        run.accepted = True
        run.costs = [(cost_fn or (lambda rr: rr.cost()))(r)]
        log(f"[warmstart] FBP start cost {run.costs[0]:.8e} ({spent:.2f} s)")
        if run.costs[0] <= target:
            run.reached = True
            run.seconds = spent
            if keep:
                run.volume = r.volume64()
            return run
        return run_to_target(r, target, max_iter, time.perf_counter() - spent, run, log,
                             cost_fn, screen, keep)
    finally:
        r.free()


def _main():
    import argparse
    import json
    from pathlib import Path
    from . import configs, cost as model_cost, set_verbose

    ap = argparse.ArgumentParser(description="train the warm-start network and "
                                 "measure time to cost tolerance, plain vs warm")
    ap.add_argument("--phase", choices=["all", "train", "eval"], default="all")
    ap.add_argument("--config", default="config3", help="evaluation scan")
    ap.add_argument("--train-config", default="", help="training scan (default: --config)")
    ap.add_argument("--train", type=int, default=16, help="training phantoms")
    ap.add_argument("--test", type=int, default=3, help="held-out phantoms")
    ap.add_argument("--early", type=int, default=1, help="plain iterations before the network")
    ap.add_argument("--target-its", type=int, default=40, help="iterations of the training targets")
    ap.add_argument("--ref-its", type=int, default=80, help="iterations of the reference runs")
    ap.add_argument("--dense-from", type=int, default=0,
                    help="reference runs: cost after every iteration from this one on "
                         "(every 5th before it)")
    ap.add_argument("--tol", type=float, default=1e-4, help="relative cost tolerance")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--drop-ends", type=int, default=2,
                    help="slices at either end of a training volume left out of the loss")
    ap.add_argument("--fbp", action="store_true",
                    help="also time ICD from the GPU FBP start (same matrix as the warm run)")
    ap.add_argument("--fp16-warm", action="store_true",
                    help="warm runs on the FP16 SSM, judged with the FP32 model's cost")
    ap.add_argument("--fbp-target", type=float, default=0.0,
                    help="reference cost of phantom --fbp-seed from an earlier run: time "
                         "only the FBP start against it (no training, reference or "
                         "learned-start runs)")
    ap.add_argument("--fbp-seed", type=int, default=2000, help="phantom of --fbp-target")
    ap.add_argument("--workdir", default="/tmp/ctbb_warmstart")
    ap.add_argument("--out", default="", help="JSON summary path")
    ap.add_argument("--params-out", default="", help="trained parameters (.npy)")
    ap.add_argument("--params-in", default="", help="parameters to evaluate (.npy)")
    ap.add_argument("--target-start", choices=["zero", "fbp"], default="zero",
                    help="training targets: sweeps from zero or from the FBP image")
    ap.add_argument("--train-seed0", type=int, default=1000, help="first training phantom")
    ap.add_argument("--workers", type=int, default=1,
                    help="training pairs generated this many phantoms at a time")
    ap.add_argument("--save", default="", help="trained network as shipped (.npz)")
    ap.add_argument("--net", default="", help="network to evaluate (.npz from --save)")
    ap.add_argument("--converged", action="store_true",
                    help="targets from converged FP64 references (reference.converged) "
                         "instead of a plain run of --ref-its sweeps; --ref-its caps the "
                         "plain and warm runs")
    ap.add_argument("--ref-rel", type=float, default=1e-9,
                    help="converged references: stop when an FP64 sweep lowers the cost "
                         "by less than this (relative)")
    ap.add_argument("--test-seed0", type=int, default=2000, help="first held-out phantom")
    a = ap.parse_args()

    set_verbose(False)
    init(0)
    wd = Path(a.workdir)
    wd.mkdir(parents=True, exist_ok=True)

    def make_scan(cfg, seed):
        return Scan(configs.write_prm(cfg, wd / f"{cfg}_{seed}.yaml", synthetic_seed=seed))

    def log(msg):
        print(msg, flush=True)

    summary = dict(config=a.config, train_config=a.train_config or a.config,
                   fp16_warm=a.fp16_warm, early=a.early, tol=a.tol)
    if a.fbp_target > 0:
        scan = make_scan(a.config, a.fbp_seed)
        g = scan.geom
        ssm = SSM.build(g)
        ssm_w = ssm
        if a.fp16_warm:
            ssm_w = SSM.build(g)
            ssm_w.to_fp16()
        raw, w = scan.synthesize(use_gpu=True)
        opts = scan.opts()
        cost_fn = None
        if a.fp16_warm:
            def cost_fn(rr):
                return model_cost(ssm, raw, w, scan.slice_index, opts, rr.volume())
        target = a.fbp_target * (1 + a.tol)
        fr = fbp_run(scan, ssm_w, raw, w, target, a.ref_its, log, cost_fn)
        log(f"[warmstart] seed {a.fbp_seed}: FBP start reaches {a.tol:g} of {a.fbp_target:.10e}: "
            f"{fr.reached} after {fr.iterations} iterations ({fr.seconds:.2f} s)")
        summary.update(fbp_only=dict(seed=a.fbp_seed, reference_cost=a.fbp_target, target=target,
                                     iterations=fr.iterations, seconds=fr.seconds,
                                     reached=fr.reached, costs=fr.costs))
        if a.out:
            Path(a.out).write_text(json.dumps(summary, indent=1))
        return
    net = None
    if a.phase in ("all", "train"):
        tcfg = a.train_config or a.config
        s0 = make_scan(tcfg, 1)
        g = s0.geom
        t = time.perf_counter()
        ssm = SSM.build(g)
        log(f"[warmstart] train on {tcfg}: {g.nx}x{g.ny}x{g.nz}, SSM in "
            f"{time.perf_counter() - t:.1f} s")
        t = time.perf_counter()
        pairs = generate_pairs(lambda sd: make_scan(tcfg, sd),
                               range(a.train_seed0, a.train_seed0 + a.train),
                               ssm, a.early, a.target_its, log, a.target_start,
                               a.workers)
        t_pairs = time.perf_counter() - t
        log(f"[warmstart] {len(pairs)} pairs in {t_pairs:.1f} s ({a.workers} at a time)")
        ssm.free()
        net = UNet(4, 32, a.batch, g.ny, g.nx, seed=11)
        t = time.perf_counter()
        hist = train(net, pairs, a.epochs, a.lr, log=log, drop_ends=a.drop_ends)
        t_train = time.perf_counter() - t
        ms, flops = net.last_timing()
        log(f"[warmstart] training {t_train:.1f} s; last step {ms:.2f} ms, "
            f"{flops / (ms * 1e-3) / 1e12:.1f} TFLOP/s in the 3x3 convolutions")
        summary.update(train=a.train, target_its=a.target_its, epochs=a.epochs,
                       pairs_seconds=t_pairs, workers=a.workers, train_seconds=t_train, loss_history=hist, unet_step_ms=ms,
                       unet_step_tflops=flops / (ms * 1e-3) / 1e12)
        if a.params_out:
            np.save(a.params_out, net.params())
        if a.save:
            save_params(a.save, net, dict(train_config=tcfg, phantoms=a.train,
                                          seeds=[a.train_seed0, a.train_seed0 + a.train],
                                          early=a.early, target_its=a.target_its,
                                          target_start=a.target_start, epochs=a.epochs,
                                          drop_ends=a.drop_ends,
                                          lr=a.lr, final_loss=hist[-1],
                                          train_seconds=t_train))
        del pairs
    if a.phase in ("all", "eval"):
        s0 = make_scan(a.config, 1)
        g = s0.geom
        t = time.perf_counter()
        ssm = SSM.build(g)
        log(f"[warmstart] evaluate on {a.config}: {g.nx}x{g.ny}x{g.nz}, SSM in "
            f"{time.perf_counter() - t:.1f} s")
        ssm_w = ssm
        if a.fp16_warm:
            ssm_w = SSM.build(g)
            ssm_w.to_fp16()
        if a.net and net is None:
            net, meta = load_params(a.net, a.batch, g.ny, g.nx)
            summary.update(network=dict(meta, file=Path(a.net).name))
        if net is None or net.height != g.ny or net.width != g.nx:
            p = net.params() if net is not None else np.load(a.params_in)
            net = UNet(4, 32, a.batch, g.ny, g.nx, seed=11)
            net.set_params(p)
        if a.converged:
            _eval_converged(a, summary, make_scan, ssm, ssm_w, net, log, model_cost)
            return
        results = []
        for s in range(a.test_seed0, a.test_seed0 + a.test):
            scan = make_scan(a.config, s)
            raw, w = scan.synthesize(use_gpu=True)
            opts = scan.opts()
            # reference: plain FP32 ICD from zero; target = its final cost
            r = Recon(ssm, raw, w, scan.slice_index, opts)
            secs, costs = [], {}
            for it in range(1, a.ref_its + 1):
                t = time.perf_counter()
                r.iterate(1)
                secs.append(time.perf_counter() - t)
                if it >= a.dense_from or it % 5 == 0 or it == a.ref_its:
                    costs[it] = r.cost()
                    log(f"[warmstart] reference iteration {it}: cost {costs[it]:.10e} "
                        f"({np.sum(secs):.2f} s)")
            r.free()
            final = costs[a.ref_its]
            target = final * (1 + a.tol)
            n_plain = min(i for i, c in costs.items() if c <= target)
            plain = dict(iterations=n_plain, seconds=float(np.sum(secs[:n_plain])))
            log(f"[warmstart] seed {s}: reference cost {final:.10e}; plain reaches {a.tol:g} "
                f"after {n_plain} iterations ({plain['seconds']:.2f} s)")
            cost_fn = None
            if a.fp16_warm:
                def cost_fn(rr, raw=raw, w=w, scan=scan, opts=opts):
                    return model_cost(ssm, raw, w, scan.slice_index, opts, rr.volume())
            warm = warm_run(scan, ssm_w, raw, w, net, a.early, target, a.ref_its, log, cost_fn)
            fbp = None
            if a.fbp:
                fr = fbp_run(scan, ssm_w, raw, w, target, a.ref_its, log, cost_fn)
                fbp = dict(iterations=fr.iterations, seconds=fr.seconds, reached=fr.reached,
                           costs=fr.costs)
            results.append(dict(seed=s, reference_cost=final, target=target, fbp=fbp,
                                plain_iterations=n_plain, plain_seconds=plain["seconds"],
                                plain_seconds_per_iteration=float(np.mean(secs)),
                                warm_iterations=warm.iterations, warm_seconds=warm.seconds,
                                warm_reached=warm.reached, warm_accepted=warm.accepted,
                                reference_costs=costs, warm_costs=warm.costs))
        pi = float(np.mean([x["plain_iterations"] for x in results]))
        wi = float(np.mean([x["warm_iterations"] for x in results]))
        ps = float(np.mean([x["plain_seconds"] for x in results]))
        wsec = float(np.mean([x["warm_seconds"] for x in results]))
        summary.update(test=a.test, ref_its=a.ref_its, results=results,
                       mean_plain_iterations=pi, mean_warm_iterations=wi,
                       mean_plain_seconds=ps, mean_warm_seconds=wsec,
                       speedup=ps / wsec if wsec > 0 else None)
        log(f"[warmstart] mean: plain {pi:.1f} its / {ps:.2f} s, warm {wi:.1f} its / "
            f"{wsec:.2f} s, speedup {ps / wsec:.2f}x")
    if a.out:
        Path(a.out).write_text(json.dumps(summary, indent=1))


def _eval_converged(a, summary, make_scan, ssm, ssm_w, net, log, model_cost):
    """Held-out phantoms a.test_seed0 ...: the converged FP64 reference
    (reference.converged), then plain FP32 from zero and the learned start,
    each to (1 + a.tol) x the converged cost; iterations, wall time and the
    volume error in HU against the converged volume.  The summary (a.out) is
    rewritten after every phantom."""
    import json
    from pathlib import Path
    from . import reference
    hu = MU_WATER / 1000.0
    results = []
    for s in range(a.test_seed0, a.test_seed0 + a.test):
        scan = make_scan(a.config, s)
        raw, w = scan.synthesize(use_gpu=True)
        opts = scan.opts()
        ref, x_ref = reference.converged(scan, ssm, raw, w, rel_tol=a.ref_rel, log=log)
        target = ref["cost"] * (1 + a.tol)
        cost_fn, slack = None, 1e-6
        if a.fp16_warm:
            def cost_fn(rr, raw=raw, w=w, scan=scan, opts=opts):
                return model_cost(ssm, raw, w, scan.slice_index, opts, rr.volume())
            slack = 1e-4
        plain = plain_run(scan, ssm, raw, w, target, a.ref_its, log, screen=1e-6, keep=True)
        warm = warm_run(scan, ssm_w, raw, w, net, a.early, target, a.ref_its, log, cost_fn,
                        screen=slack, keep=True)

        nxy, nz = scan.geom.nx * scan.geom.ny, scan.geom.nz

        def rmse(v, ends=4):
            """HU error against the converged volume: all voxels, and the
            slices away from the slab ends (which converge slowest)."""
            d = (v - x_ref).reshape(nz, nxy)
            return dict(all=float(np.sqrt(np.mean(d ** 2)) / hu),
                        interior=float(np.sqrt(np.mean(d[ends:nz - ends] ** 2)) / hu))
        res = dict(seed=s, reference=ref, target=target,
                   plain=dict(iterations=plain.iterations, seconds=plain.seconds,
                              reached=plain.reached, cost=plain.costs[-1] if plain.costs
                              else None, rmse_hu=rmse(plain.volume)),
                   warm=dict(iterations=warm.iterations, seconds=warm.seconds,
                             reached=warm.reached, accepted=warm.accepted,
                             cost=warm.costs[-1] if warm.costs else None,
                             rmse_hu=rmse(warm.volume)))
        if a.fbp:
            fr = fbp_run(scan, ssm_w, raw, w, target, a.ref_its, log, cost_fn, screen=slack,
                         keep=True)
            res["fbp"] = dict(iterations=fr.iterations, seconds=fr.seconds,
                              reached=fr.reached, rmse_hu=rmse(fr.volume))
        results.append(res)
        log(f"[warmstart] seed {s}: converged {ref['cost']:.10e}; plain {plain.iterations} "
            f"its / {plain.seconds:.2f} s ({res['plain']['rmse_hu']['interior']:.2f} HU), warm "
            f"{warm.iterations} its / {warm.seconds:.2f} s "
            f"({res['warm']['rmse_hu']['interior']:.2f} HU, "
            f"{'accepted' if warm.accepted else 'rejected'})")
        ps = float(np.mean([x["plain"]["seconds"] for x in results]))
        wsec = float(np.mean([x["warm"]["seconds"] for x in results]))
        summary.update(test=len(results), test_seed0=a.test_seed0, cap_iterations=a.ref_its,
                       results=results,
                       mean_plain_iterations=float(np.mean([x["plain"]["iterations"]
                                                            for x in results])),
                       mean_warm_iterations=float(np.mean([x["warm"]["iterations"]
                                                           for x in results])),
                       mean_plain_seconds=ps, mean_warm_seconds=wsec,
                       speedup=ps / wsec if wsec > 0 else None,
                       all_reached=all(x["plain"]["reached"] and x["warm"]["reached"]
                                       for x in results))
        if a.out:
            Path(a.out).write_text(json.dumps(summary, indent=1))
    log(f"[warmstart] mean over {len(results)}: plain {summary['mean_plain_iterations']:.1f} "
        f"its / {summary['mean_plain_seconds']:.2f} s, warm "
        f"{summary['mean_warm_iterations']:.1f} its / {summary['mean_warm_seconds']:.2f} s, "
        f"speedup {summary['speedup']:.2f}x")


if __name__ == "__main__":
    _main()
