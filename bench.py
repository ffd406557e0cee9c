#!/usr/bin/env python3
"""Benchmark: poses/sec of DRR + patch gradient-NCC at the 384^2 detector, batched vs plain.

Workload (CMA-ES population batching, single view): seeded 256x256x200 float32 ellipsoid
phantom at 1.0 mm, 384x384 detector at 0.776 mm, focal 1020 mm, ray step 0.5 mm, lambda = 100
poses per generation (sigma 10 deg / 15 mm about the ground truth), patch gradient-NCC
(11x11 patches, Sobel after the default 5x5 smoothing, mean over patches). `--config c4`
runs the 512x512x400 @0.5 mm volume at 768^2 with 0.25 mm steps instead.

One timed step = one generation: the lambda poses go to the device, one batched submission
(graph-replayed) casts all DRRs and scores them, lambda similarity values come back. The plain
leg evaluates poses one at a time (DRR launch, similarity launch, synchronous read-back).

The served legs run registrations arriving as a Poisson process: `--serve-jobs` of the C2
workload (default), and with `--serve-mix-jobs` a mixed stream of three resident phantoms,
detectors of 192^2 to 768^2, lambda 20-200, one or three views, batched and then served
serially for comparison.

Prints one JSON line (see the keys below). `--gpus N` with N > 1 exits: this build runs on one
GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--lam", type=int, default=100)
    ap.add_argument("--config", default="c2", choices=["c2", "c4"],
                    help="c2: 256x256x200 @1 mm, 384^2, 0.5 mm steps (default); "
                         "c4: 512x512x400 @0.5 mm, 768^2, 0.25 mm steps")
    ap.add_argument("--plain-poses", type=int, default=30)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--geom", default="fp64", choices=["fp32", "fp64"])
    ap.add_argument("--smooth", type=int, default=5)
    ap.add_argument("--no-skip", action="store_true", help="evaluate every sample (no empty-space map)")
    ap.add_argument("--serve-jobs", type=int, default=24, help="served-latency leg: number of jobs (0 skips)")
    ap.add_argument("--serve-rate", type=float, default=1.0, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--serve-gens", type=int, default=50, help="CMA-ES generations per job (upper bound)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default DRR launch variant instead of tuning it on the first generation")
    ap.add_argument("--serve-policy", default="srpt", choices=["srpt", "fifo"],
                    help="served legs: order in which jobs fill a submission (least remaining work first, "
                         "or arrival order)")
    ap.add_argument("--serve-step-samples", type=float, default=2.0e10,
                    help="served legs: estimated trilinear samples per submission (0: no limit)")
    ap.add_argument("--serve-mix-jobs", type=int, default=0,
                    help="mixed served stream (three resident phantoms, 192/384/768 detectors, lambda 20-200, "
                         "30-80 generations, one or three views): number of jobs, run batched and serially")
    ap.add_argument("--serve-mix-rate", type=float, default=0.25, help="mixed stream: Poisson rate, jobs/s")
    ap.add_argument("--c3-regi", action="store_true",
                    help="also time a three-view, three-level registration (96^2 CMA-ES, then the refiner "
                         "at 192^2 and 384^2)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run plain device copies of this many MiB (counter calibration under --pmc)")
    return ap.parse_args()


def generation_poses(w, lam, n_gen, seed=11):
    from xreg_amd import synth

    out = []
    for g in range(n_gen):
        out.append(synth.random_poses(w.gt, w.center, lam, 10.0, 15.0, seed=seed + g, normal=True))
    return out


def cpu_baseline(w, fixed, poses, budget_s, smooth):
    """The CPU restatement of RayCasterLineIntCPU + patch gradient-NCC (oracle), all host
    threads, on as many poses of the same workload as fit in about budget_s seconds."""
    from oracle import oracle as O

    O.lib()
    threads = min(len(os.sched_getaffinity(0)), 16)
    # one untimed pose first (page faults, thread start-up)
    drr, _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, np.asarray(poses[0], np.float32)[None], step=w.step,
                       threads=threads)
    O.patch_grad_ncc_f32(fixed, drr, smooth_ksize=smooth, threads=threads)
    t0 = time.perf_counter()
    done = 0
    t_drr = t_sim = 0.0
    for p in poses:
        a = time.perf_counter()
        drr, _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, np.asarray(p, np.float32)[None], step=w.step,
                           threads=threads)
        b = time.perf_counter()
        O.patch_grad_ncc_f32(fixed, drr, smooth_ksize=smooth, threads=threads)
        c = time.perf_counter()
        t_drr += b - a
        t_sim += c - b
        done += 1
        if c - t0 >= budget_s and done >= 2:
            break
    el = time.perf_counter() - t0
    # one pose on one thread, for the single-core figure
    a = time.perf_counter()
    drr, _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, np.asarray(poses[0], np.float32)[None], step=w.step,
                       threads=1)
    b = time.perf_counter()
    O.patch_grad_ncc_f32(fixed, drr, smooth_ksize=smooth, threads=1)
    c = time.perf_counter()
    return {"value": done / el, "unit": "poses/s", "cores": threads, "kind": "port",
            "sample": f"{done} poses of the timed generations ({w.cams[0].num_det_rows}^2 patch-grad-NCC, same "
                      f"volume and fixed image) after one untimed pose, "
                      f"DRR {t_drr:.2f}s + similarity {t_sim:.2f}s, oracle FP32 restatement, {threads} threads",
            "single_thread": {"value": 1.0 / (c - a), "drr_s": round(b - a, 3), "similarity_s": round(c - b, 3)}}


def profiled_traffic(key: str):
    """HBM-side traffic per DRR launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated
    against a plain copy) for this exact configuration, from the committed profile summary; None
    when the configuration was not profiled."""
    for p in sorted((ROOT / "profiles").glob("drr_kernel_*.md")):
        for line in p.read_text().splitlines():
            if line.startswith("traffic: ") and f" {key} " in line + " ":
                for tok in line.split():
                    if tok.startswith("bytes_per_launch="):
                        return float(tok.split("=", 1)[1])
    return None


def serve_leg(ctx, vol, w, n_jobs, rate, max_gens, lam, geom, smooth, policy, step_samples, seed=21):
    """Served registration latency: n_jobs single-view CMA-ES registrations of the same workload
    (lambda poses per generation, up to max_gens generations) arriving as a Poisson process of
    `rate` jobs/s. Each job has its own ground truth (the reference pose moved by U(+-3 deg,
    +-5 mm)), its own noisy fixed image cast on the device, and starts from U(+-5 deg, +-10 mm)
    about it. Latency = finish - arrival, on the wall clock."""
    from xreg_amd import _xrh as X
    from xreg_amd import serve, synth
    from xreg_amd import geometry as G

    rng = np.random.default_rng(seed)
    cam = w.cams[0]
    P = cam.num_det_rows * cam.num_det_cols
    buf = ctx.alloc(4 * P)
    specs, t = [], 0.0
    for j in range(n_jobs):
        gt = G.perturbed_pose(w.gt, w.center, np.deg2rad(rng.uniform(-3, 3, 3)), rng.uniform(-5, 5, 3))
        X.line_int(ctx, vol, [cam], gt.astype(np.float32)[None], np.zeros(1, np.int64), w.step, buf, geom=geom)
        fixed = synth.noisy_fixed(buf.download(np.float32).reshape(cam.num_det_rows, cam.num_det_cols),
                                  seed=1000 + j)
        init = G.perturbed_pose(gt, w.center, np.deg2rad(rng.uniform(-5, 5, 3)), rng.uniform(-10, 10, 3))
        specs.append(serve.JobSpec(job_id=j, arrival=t, volume="c2", cams=[cam], fixed=[fixed], init_pose=init,
                                   center=w.center, step=w.step, lam=lam, max_gens=max_gens, seed=5000 + j,
                                   sigma=(np.deg2rad(5.0),) * 3 + (10.0,) * 3, smooth_ksize=smooth))
        t += float(rng.exponential(1.0 / rate))
    buf.free()
    srv = serve.RegistrationServer(ctx, {"c2": vol}, hbm_budget=200 << 30, max_items=max(4096, lam),
                                   max_views=64, max_det=max(cam.num_det_rows, cam.num_det_cols), geom=geom,
                                   policy=policy, step_samples=step_samples)
    t0 = time.perf_counter()
    res = srv.run(specs, realtime=True)
    wall = time.perf_counter() - t0
    srv.close()
    lat = [r.latency for r in res]
    return {"jobs": len(res), "arrival_rate_jobs_per_s": rate, "lambda": lam, "max_generations": max_gens,
            "mean_generations": round(float(np.mean([r.generations for r in res])), 2),
            "p50_s": round(serve.percentile(lat, 50), 4), "p99_s": round(serve.percentile(lat, 99), 4),
            "max_s": round(max(lat), 4), "wall_s": round(wall, 3), "scheduler_steps": srv.steps,
            "policy": policy, "step_samples": step_samples,
            "poses_per_s": round(srv.poses_evaluated / wall, 1),
            "max_poses_per_submission": srv.max_items_in_step}


# Mixed served stream: three resident phantoms (name, voxels x y z, spacing mm) and a 298 mm
# detector read out at 192^2, 384^2 or 768^2 pixels.
MIX_PHANTOMS = (("ph192", (192, 192, 192), 1.5), ("ph256", (256, 256, 200), 1.0), ("ph512", (512, 512, 400), 0.5))
MIX_DETS = (192, 384, 768)
MIX_DET_MM = 192 * 1.552


def serve_mix_leg(ctx, n_jobs, rate, geom, smooth, policy, step_samples, seed=31):
    """Mixed served stream: n_jobs registrations arriving as a Poisson process of `rate` jobs/s,
    each drawn uniformly: phantom (three resident volumes), detector size, one or three views
    (30 degrees apart), CMA-ES lambda in [20, 200] and a generation cap in [30, 80], patch
    gradient-NCC, ray step half a voxel. Each job has its own ground truth (U(+-3 deg, +-5 mm)
    about the phantom's reference pose), noisy fixed images cast on the device, and a start
    U(+-5 deg, +-10 mm) about it. The same jobs run twice on the wall clock: continuously
    batched, then served one at a time in arrival order (the serial baseline)."""
    from xreg_amd import _xrh as X
    from xreg_amd import serve, synth
    from xreg_amd import geometry as G

    rng = np.random.default_rng(seed)
    vols, meta = {}, {}
    for name, size, sp in MIX_PHANTOMS:
        vols[name] = X.Volume(ctx, synth.phantom(size), G.idx_to_phys(sp))
        meta[name] = (sp, G.vol_center_phys(size, sp))
    buf = ctx.alloc(4 * max(MIX_DETS) ** 2)
    specs, t = [], 0.0
    for j in range(n_jobs):
        name = MIX_PHANTOMS[int(rng.integers(len(MIX_PHANTOMS)))][0]
        sp, center = meta[name]
        det = int(MIX_DETS[int(rng.integers(len(MIX_DETS)))])
        nv = 1 if rng.random() < 0.5 else 3
        lam, gens = int(rng.integers(20, 201)), int(rng.integers(30, 81))
        cams = synth.multiview_cams(1020.0, det, det, MIX_DET_MM / det, 3, 30.0)[:nv]
        gt = G.perturbed_pose(G.gt_pose(center, 510.0), center, np.deg2rad(rng.uniform(-3, 3, 3)),
                              rng.uniform(-5, 5, 3))
        fixed = []
        for v, cam in enumerate(cams):
            X.line_int(ctx, vols[name], [cam], gt.astype(np.float32)[None], np.zeros(1, np.int64), sp / 2, buf,
                       geom=geom)
            fixed.append(synth.noisy_fixed(buf.download(np.float32, count=det * det).reshape(det, det),
                                           seed=2000 + 3 * j + v))
        init = G.perturbed_pose(gt, center, np.deg2rad(rng.uniform(-5, 5, 3)), rng.uniform(-10, 10, 3))
        specs.append(serve.JobSpec(job_id=j, arrival=t, volume=name, cams=cams, fixed=fixed, init_pose=init,
                                   center=center, step=sp / 2, lam=lam, max_gens=gens, seed=9000 + j,
                                   sigma=(np.deg2rad(5.0),) * 3 + (10.0,) * 3, smooth_ksize=smooth))
        t += float(rng.exponential(1.0 / rate))
    buf.free()

    def run(max_active):
        srv = serve.RegistrationServer(ctx, vols, hbm_budget=200 << 30, max_items=2048, max_views=64,
                                       max_det=max(MIX_DETS), geom=geom, max_active=max_active, policy=policy,
                                       step_samples=step_samples)
        t0 = time.perf_counter()
        mode = "serial" if max_active == 1 else "batched"
        res = srv.run(specs, realtime=True,
                      on_finish=lambda r: print(f"served mix ({mode}): job {r.job_id} done, {r.generations} "
                                                f"generations, latency {r.latency:.2f} s", file=sys.stderr,
                                                flush=True))
        wall = time.perf_counter() - t0
        srv.close()
        lat = [r.latency for r in res]
        out = {"wall_s": round(wall, 3), "jobs_per_s": round(len(res) / wall, 4),
               "poses_per_s": round(srv.poses_evaluated / wall, 1),
               "p50_s": round(serve.percentile(lat, 50), 4), "p99_s": round(serve.percentile(lat, 99), 4),
               "max_s": round(max(lat), 4), "scheduler_steps": srv.steps,
               "max_poses_per_submission": srv.max_items_in_step}
        print(f"served mix ({mode}): {out}", file=sys.stderr, flush=True)
        return out, srv.poses_evaluated

    batched, n_b = run(None)
    serial, n_s = run(1)
    for v in vols.values():
        v.close()
    return {"jobs": n_jobs, "arrival_rate_jobs_per_s": rate,
            "phantoms": [f"{n}:{s[0]}x{s[1]}x{s[2]}@{sp}mm" for n, s, sp in MIX_PHANTOMS],
            "detectors": list(MIX_DETS), "lambda_range": [20, 200], "generation_cap_range": [30, 80],
            "views": [1, 3], "poses_evaluated": n_b, "policy": policy, "step_samples": step_samples,
            "batched": batched, "serial": serial,
            "same_work": n_b == n_s}


def c3_leg(ctx, geom, smooth, seed=41):
    """Multi-view, multi-resolution registration: three views 30 degrees apart of the
    256x256x200 phantom with 384^2 detectors, gradient-NCC averaged over the views. CMA-ES
    (lambda 50, up to 50 generations) on the 96^2 level, then the BOBYQA-pattern refiner (up to
    20 x 13 single-pose evaluations) at 192^2 and at 384^2. The fixed images are the
    ground-truth DRRs plus 1% noise; the start is U(+-5 deg, +-10 mm) away. Reports each
    level's wall-clock time (set-up included), evaluations and the mean projection distance."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import regi, synth

    rng = np.random.default_rng(seed)
    w = synth.config("c3", n_poses=1)
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    fixed = []
    for v, cam in enumerate(w.cams):
        buf = ctx.alloc(4 * cam.num_det_rows * cam.num_det_cols)
        X.line_int(ctx, vol, [cam], w.gt.astype(np.float32)[None], np.zeros(1, np.int64), w.step, buf, geom=geom)
        fixed.append(synth.noisy_fixed(buf.download(np.float32).reshape(cam.num_det_rows, cam.num_det_cols),
                                       seed=40 + v))
        buf.free()
    init = G.perturbed_pose(w.gt, w.center, np.deg2rad(rng.uniform(-5, 5, 3)), rng.uniform(-10, 10, 3))
    cam0 = w.cams[0]

    def ds(img, f):
        return X.downsample_image(ctx, img, f, spacing=(cam0.det_row_spacing, cam0.det_col_spacing))

    levels = [regi.Level(0.25, "cmaes", dict(lam=50, sigma=[np.deg2rad(5.0)] * 3 + [10.0] * 3, max_gens=50,
                                             seed=seed)),
              regi.Level(0.5, "refine", dict(rho_beg=[np.deg2rad(0.5)] * 3 + [1.0] * 3, max_evals=260)),
              regi.Level(1.0, "refine", dict(rho_beg=[np.deg2rad(0.25)] * 3 + [0.5] * 3, max_evals=260))]
    t0 = time.perf_counter()
    pose, res = regi.run_multi_level(ctx, vol, w.cams, fixed, init, w.center, w.step, levels, ds,
                                     metric=X.XRH_SIM_GRAD_NCC, smooth_ksize=smooth, geom=geom)
    wall = time.perf_counter() - t0
    vol.close()
    ext = np.array(w.vol.shape[::-1]) * w.spacing
    pts = [w.center + 0.25 * ext * np.array(c) for c in [(-1, -1, -1), (1, 1, 1), (1, -1, 1), (-1, 1, -1),
                                                          (0, 0, 0)]]
    mpd0 = G.mean_projection_distance(init, w.gt, pts, 1020.0)
    out = {"views": 3, "total_wall_s": round(wall, 3), "mpd_start_mm": round(mpd0, 3), "levels": []}
    for lvl, r in zip(levels, res):
        det = int(round(cam0.num_det_rows * lvl.ds_factor))
        out["levels"].append({"detector": det, "method": lvl.method, "evals": int(r.evals),
                              "wall_s": round(float(r.extra["wall_s"]), 3),
                              "mpd_mm": round(G.mean_projection_distance(r.pose, w.gt, pts, 1020.0), 3)})
    print(f"c3 registration: {out}", file=sys.stderr, flush=True)
    return out


def main():
    args = parse()
    if args.serve_step_samples <= 0:
        args.serve_step_samples = None
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.no_skip:
        os.environ["XRH_SKIP_EMPTY"] = "0"
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    storage = {"f32": X.XRH_F32, "f16": X.XRH_F16, "bf16": X.XRH_BF16}[args.storage]
    geom = X.XRH_GEOM_FP64 if args.geom == "fp64" else X.XRH_GEOM_FP32
    lam = args.lam
    w = synth.config(args.config, n_poses=1)
    cam = w.cams[0]
    P = cam.num_det_rows * cam.num_det_cols
    det = f"{cam.num_det_rows}"
    workload = f"cmaes-lambda{lam}-{det}sq-patch-grad-ncc"
    X_, Y_, Z_ = w.size_xyz

    ctx = X.Context(0)
    vol = X.Volume(ctx, w.vol, w.idx_to_phys, storage=storage)
    # fixed image: DRR at the ground truth + 1% noise, computed on the device
    buf = ctx.alloc(4 * P)
    X.line_int(ctx, vol, w.cams, w.gt.astype(np.float32)[None], np.zeros(1, np.int64), w.step, buf, geom=geom)
    fixed = synth.noisy_fixed(buf.download(np.float32).reshape(cam.num_det_rows, cam.num_det_cols))
    fx = X.Fixed(ctx, fixed, X.XRH_SIM_PATCH_GRAD_NCC, smooth_ksize=args.smooth, patch_radius=5)
    batch = X.Batch(ctx, [X.make_view(vol, fx, cam, w.step)], max_items=lam, geom=geom)
    gens = generation_poses(w, lam, args.warmup + args.steps)
    iv = np.zeros(lam, np.int32)

    # ---- DRR launch variant: tuned on the first generation (untimed) ----
    tune = None if args.no_autotune else batch.autotune(iv, gens[0], reps=2)

    # ---- batched: warmup, then exactly K generations between synchronisations ----
    for g in range(args.warmup):
        batch.eval(iv, gens[g])
    ctx.synchronize()
    t0 = time.perf_counter()
    sims = None
    for g in range(args.warmup, args.warmup + args.steps):
        sims = batch.eval(iv, gens[g])
    ctx.synchronize()
    t_batched = time.perf_counter() - t0
    value = args.steps * lam / t_batched

    # ---- DRR kernel duration (HIP events on the library stream, eager submissions) ----
    batch.set_options(use_graph=False, time_kernels=True)
    drr_ms = []
    for g in range(min(5, len(gens))):
        batch.eval(iv, gens[g])
        drr_ms.append(batch.last_timing()[0])
    batch.set_options(use_graph=True, time_kernels=False)
    drr_ms_avg = float(np.mean(drr_ms[1:])) if len(drr_ms) > 1 else float(drr_ms[0])
    # the same generations as the timed DRR launches above: algorithmic and evaluated samples
    counts = [batch.sample_count(iv, gens[g], evaluated=True) for g in range(min(5, len(gens)))]
    S = float(np.mean([c[0] for c in counts[1:]])) if len(counts) > 1 else float(counts[0][0])
    S_eval = float(np.mean([c[1] for c in counts[1:]])) if len(counts) > 1 else float(counts[0][1])
    bv = {X.XRH_F32: 4, X.XRH_F16: 2, X.XRH_BF16: 2}[storage]
    bytes_launch = 8.0 * bv * S + 4.0 * P * lam  # gathers + DRR write
    achieved = bytes_launch / (drr_ms_avg * 1e-3) / 1e9

    # ---- plain: one pose per launch, synchronous read-back ----
    n_plain = min(args.plain_poses, lam)
    plain_value = None
    if n_plain > 0:
        pl = gens[0][:n_plain]
        for i in range(min(2, n_plain)):
            X.line_int(ctx, vol, w.cams, pl[i].astype(np.float32)[None], np.zeros(1, np.int64), w.step, buf,
                       geom=geom)
            fx.compute(buf, 1)
        ctx.synchronize()
        t0 = time.perf_counter()
        for i in range(n_plain):
            X.line_int(ctx, vol, w.cams, pl[i].astype(np.float32)[None], np.zeros(1, np.int64), w.step, buf,
                       geom=geom)
            fx.compute(buf, 1)
        ctx.synchronize()
        plain_value = n_plain / (time.perf_counter() - t0)

    # ---- drop-in call pattern: what Intensity2D3DRegi gets through RayCasterLineIntHIP and
    # ImgSimMetric2D*HIP -- one caster compute() over all lambda projections (xrh_line_int),
    # then one metric compute() (xrh_sim_compute) per generation, no graph replay
    dbuf = ctx.alloc(4 * P * lam)
    zeros = np.zeros(lam, np.int64)
    for g in range(args.warmup):
        X.line_int(ctx, vol, w.cams, gens[g].astype(np.float32), zeros, w.step, dbuf, geom=geom)
        fx.compute(dbuf, lam)
    ctx.synchronize()
    t0 = time.perf_counter()
    for g in range(args.warmup, args.warmup + args.steps):
        X.line_int(ctx, vol, w.cams, gens[g].astype(np.float32), zeros, w.step, dbuf, geom=geom)
        fx.compute(dbuf, lam)
    ctx.synchronize()
    dropin_value = args.steps * lam / (time.perf_counter() - t0)
    dbuf.free()

    # STREAM-style copy probe with the library's copy kernel (xrh_copy_kernel): measured HBM
    # bandwidth next to the published peak; under --pmc it is also the counters' calibration
    # (a dispatch that reads and writes exactly copy_mib MiB)
    copy_mib = args.calib_copy_mib if args.calib_copy_mib > 0 else 1024
    src = ctx.alloc(copy_mib << 20)
    dst = ctx.alloc(copy_mib << 20)
    src.upload(np.ones((copy_mib << 20) // 4, np.float32))
    dst.copy_from(src)
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_from(src)
    copy_gbs = 5 * 2.0 * (copy_mib << 20) / (time.perf_counter() - t0) / 1e9
    src.free()
    dst.free()

    served = None
    if args.serve_jobs > 0:
        served = serve_leg(ctx, vol, w, args.serve_jobs, args.serve_rate, args.serve_gens, lam, geom, args.smooth,
                           args.serve_policy, args.serve_step_samples)

    served_mix = None
    if args.serve_mix_jobs > 0:
        served_mix = serve_mix_leg(ctx, args.serve_mix_jobs, args.serve_mix_rate, geom, args.smooth,
                                   args.serve_policy, args.serve_step_samples)

    c3 = c3_leg(ctx, geom, args.smooth) if args.c3_regi else None

    cpu = None
    if not args.no_cpu:
        # poses of the timed generations, in order, until the time budget is spent
        cpu = cpu_baseline(w, fixed, np.concatenate(gens[args.warmup:]), args.cpu_seconds, args.smooth)

    line = {
        "metric": f"poses/sec (DRR+grad-NCC, {det}^2 det) batched vs plain; p50/p99 served regi latency",
        "value": round(value, 2),
        "unit": "poses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_batched / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage.replace("f32", "fp32"),
        "data": "synthetic seeded ellipsoid phantom",
        "config": {"workload": workload, "volume": f"{X_}x{Y_}x{Z_}@{w.spacing:g}mm",
                   "detector": f"{det}x{cam.num_det_cols}@{cam.det_col_spacing:.3f}mm", "step_mm": w.step,
                   "lambda": lam,
                   "similarity": f"patch-grad-ncc r5 smooth{args.smooth}", "geometry": args.geom,
                   "storage": args.storage, "parallelism": "single-gpu"},
        "plain_poses_per_s": round(plain_value, 2) if plain_value else None,
        "batched_over_plain": round(value / plain_value, 2) if plain_value else None,
        "dropin_poses_per_s": round(dropin_value, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(f"workload={workload} storage={args.storage} "
                                                 f"geometry={args.geom} skip={0 if args.no_skip else 1} "
                                                 f"smooth={args.smooth}"),
                     "kernel": "drr_kernel", "kernel_ms": round(drr_ms_avg, 4),
                     "samples_per_launch": int(S), "samples_evaluated_per_launch": int(S_eval),
                     "algorithmic_bytes_per_launch": int(bytes_launch),
                     "evaluated_gbs": round((8.0 * bv * S_eval + 4.0 * P * lam) / (drr_ms_avg * 1e-3) / 1e9, 1),
                     "measured_copy_gbs": round(copy_gbs, 1)},
        "empty_space_skipping": not args.no_skip,
        "autotune": tune,
        "served": served,
        "served_mix": served_mix,
        "c3_registration": c3,
        "cpu_baseline": cpu,
        "sim_checksum": float(np.sum(sims)),
    }
    print(json.dumps(line))
    for o in (batch, fx, vol):
        o.close()
    buf.free()
    ctx.close()


if __name__ == "__main__":
    main()
