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

Beside the headline the line carries: the same generations with the other ray-geometry mode
(FP32 set-up, whose step counts equal RayCasterLineIntCPU's, when the headline uses FP64); the
plain and drop-in call patterns; a three-view, three-level registration (`c3_registration`:
gradient-NCC, CMA-ES at 96^2, the refiner at 192^2 and 384^2); the 512x512x400 @0.5 mm volume at
768^2 stored as FP32, FP16 and BF16 (`c4`); and two served legs, registrations arriving as a
Poisson process: `served` (the C2 workload) and `served_mix` (three resident phantoms, detectors
of 192^2 to 768^2, lambda 20-200, one or three views; batched, then served one job at a time for
comparison; latency percentiles per job class).

Prints one JSON line (see the keys below). `--gpus N` with N > 1 exits: this build runs on one
GPU.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PEAK_CLOCK_GHZ = 2.4      # MI355X peak engine clock
MI355X_CUS = 256


def device_cus() -> int:
    """Compute units of the first GPU node in the KFD topology (simd_count / 4 SIMDs per CU),
    else the MI355X's 256."""
    base = Path("/sys/class/kfd/kfd/topology/nodes")
    try:
        for node in sorted(base.iterdir()):
            props = {}
            for line in (node / "properties").read_text().splitlines():
                k, _, v = line.partition(" ")
                props[k] = v
            simd = int(props.get("simd_count", "0"))
            if simd > 0:
                return simd // 4
    except (OSError, ValueError):
        pass
    return MI355X_CUS


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
    ap.add_argument("--serve-mix-jobs", type=int, default=24,
                    help="mixed served stream (three resident phantoms, 192/384/768 detectors, lambda 20-200, "
                         "30-80 generations, one or three views): number of jobs, run batched and serially")
    ap.add_argument("--serve-mix-rate", type=float, default=0.5, help="mixed stream: Poisson rate, jobs/s")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the three-view, three-level registration (96^2 CMA-ES, then the refiner at "
                         "192^2 and 384^2)")
    ap.add_argument("--c4-steps", type=int, default=2,
                    help="reduced-precision storage leg (512x512x400 @0.5 mm, 768^2): timed generations per "
                         "storage type (0 skips)")
    ap.add_argument("--interp-poses", type=int, default=0,
                    help="interpolator leg: poses of the 128x128x96 @2 mm volume at 192^2 cast with each "
                         "interpolator and with 4-ray anti-aliasing (0 skips)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run plain device copies of this many MiB (counter calibration under --pmc)")
    return ap.parse_args()


def generation_poses(w, lam, n_gen, seed=11):
    from xreg_amd import synth

    out = []
    for g in range(n_gen):
        out.append(synth.random_poses(w.gt, w.center, lam, 10.0, 15.0, seed=seed + g, normal=True))
    return out


def cpu_info() -> dict:
    """The host the CPU baseline ran on: logical CPUs, the affinity mask, the thread budget the
    environment grants (OMP_NUM_THREADS) and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": int(omp) if omp.isdigit() else None, "model": model}


def cpu_baseline(w, fixed, poses, smooth, reps=5, per_rep=16):
    """The CPU restatement of RayCasterLineIntCPU + patch gradient-NCC (oracle) on the GPU host.
    Every CPU of the affinity mask is used unless OMP_NUM_THREADS grants fewer (the host's share
    for this GPU). The DRR and similarity stages are timed separately: one untimed repetition,
    then the median of `reps` repetitions, each over `per_rep` poses of the timed generations
    (one caster call over all of them, as RayCasterLineIntCPU::compute gets a generation, then
    one metric call); the single-thread figure is the median of `reps` one-pose repetitions."""
    from oracle import oracle as O

    O.lib()
    info = cpu_info()
    threads = info["affinity_cpus"]
    if info["omp_num_threads"]:
        threads = min(threads, info["omp_num_threads"])

    def stages(ps, th):
        a = time.perf_counter()
        drr, _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, np.asarray(ps, np.float32), step=w.step, threads=th)
        b = time.perf_counter()
        O.patch_grad_ncc_f32(fixed, drr, smooth_ksize=smooth, threads=th)
        return b - a, time.perf_counter() - b

    def med(ps_of_rep, th):
        stages(ps_of_rep(0), th)  # untimed: page faults, thread start-up
        runs = [stages(ps_of_rep(r), th) for r in range(reps)]
        return float(np.median([r[0] for r in runs])), float(np.median([r[1] for r in runs]))

    n = len(poses)
    t_drr, t_sim = med(lambda r: [poses[(r * per_rep + k) % n] for k in range(per_rep)], threads)
    t1_drr, t1_sim = med(lambda r: [poses[r % n]], 1)
    return {"value": per_rep / (t_drr + t_sim), "unit": "poses/s", "cores": threads, "kind": "port",
            "sample": f"median of {reps} repetitions (after one untimed) of {per_rep} poses of the timed "
                      f"generations ({w.cams[0].num_det_rows}^2 patch-grad-NCC, same volume and fixed image): "
                      f"DRR {t_drr:.3f}s + similarity {t_sim:.3f}s per repetition, oracle FP32 restatement, "
                      f"{threads} threads",
            "host": info,
            "drr_s_per_pose": round(t_drr / per_rep, 4), "similarity_s_per_pose": round(t_sim / per_rep, 4),
            "single_thread": {"value": 1.0 / (t1_drr + t1_sim), "drr_s": round(t1_drr, 3),
                              "similarity_s": round(t1_sim, 3)}}


def kernel_source_key() -> str:
    """Hash of the DRR kernel's sources: profile figures are looked up under it, so a figure
    measured on other kernel code is never reported."""
    h = hashlib.sha256()
    for f in ("xrh_drr.hip", "xrh_ray.h", "xrh_internal.h"):
        h.update((ROOT / "xreg_amd" / "csrc" / f).read_bytes())
    return h.hexdigest()[:12]


def profiled(key: str, folder: Path | None = None) -> dict:
    """Per-launch counter figures of the DRR kernel measured by rocprofv3 for this exact
    configuration and kernel source, from the committed profile summaries (lines `counters: <key>
    name=value ...`): valu_insts (SQ_INSTS_VALU), vmem_insts (SQ_INSTS_VMEM_RD), traffic_bytes
    (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy), ta_busy_frac (TA_TA_BUSY per CU
    over GRBM_GUI_ACTIVE per XCD: the share of the kernel's cycles in which the texture address
    unit, which takes every active lane's fetch address, is busy). Empty when not profiled."""
    out = {}
    for p in sorted((folder or ROOT / "profiles").glob("*.md")):
        for line in p.read_text().splitlines():
            if line.startswith("counters: ") and f" {key} " in line + " ":
                for tok in line.split():
                    if "=" in tok and tok.split("=", 1)[0] in ("valu_insts", "vmem_insts", "traffic_bytes",
                                                                 "ta_busy_frac"):
                        k, v = tok.split("=", 1)
                        out[k] = float(v)
    return out


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
        out["classes"] = by_class(res)
        print(f"served mix ({mode}): {out}", file=sys.stderr, flush=True)
        return out, srv.poses_evaluated

    def by_class(res):
        """p50 / p99 / max latency per job class (phantom, detector, views), and whether its p99
        stays under the 2 s service-level objective."""
        cls = {}
        for r in res:
            sp = specs[r.job_id]
            det = int(sp.cams[0].num_det_rows)
            cls.setdefault(f"{sp.volume}/{det}sq/{len(sp.cams)}v", []).append(r.latency)
        out = {}
        for k in sorted(cls):
            lat = cls[k]
            p99 = serve.percentile(lat, 99)
            out[k] = {"jobs": len(lat), "p50_s": round(serve.percentile(lat, 50), 3), "p99_s": round(p99, 3),
                      "max_s": round(max(lat), 3), "p99_under_2s": bool(p99 < 2.0)}
        return out

    batched, n_b = run(None)
    serial, n_s = run(1)
    for v in vols.values():
        v.close()
    return {"jobs": n_jobs, "arrival_rate_jobs_per_s": rate,
            "phantoms": [f"{n}:{s[0]}x{s[1]}x{s[2]}@{sp}mm" for n, s, sp in MIX_PHANTOMS],
            "detectors": list(MIX_DETS), "lambda_range": [20, 200], "generation_cap_range": [30, 80],
            "views": [1, 3], "poses_evaluated": n_b, "policy": policy, "step_samples": step_samples,
            "batched": batched, "serial": serial,
            "classes_meeting_2s_p99": sorted(k for k, v in batched["classes"].items() if v["p99_under_2s"]),
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
                              "mpd_mm": round(G.mean_projection_distance(r.pose, w.gt, pts, 1020.0), 3),
                              "fun_start": float(r.history[0]), "fun_end": float(r.fun)})
    print(f"c3 registration: {out}", file=sys.stderr, flush=True)
    return out


def batched_leg(ctx, batch, lam, gens, warmup, steps, autotune):
    """Times exactly `steps` generations (after `warmup`) of one batched, graph-replayed
    submission per generation: the lambda poses go to the device, all DRRs are cast and scored,
    lambda similarity values come back. Returns poses/s, the tuning record, the last
    generation's similarity values, the DRR kernel's average duration (HIP events on the
    library stream, eager submissions of the same generations) and the sample counts of those
    launches (reference algorithm, and evaluated after empty-space skipping)."""
    iv = np.zeros(lam, np.int32)
    tune = batch.autotune(iv, gens[0], reps=2) if autotune else None
    for g in range(warmup):
        batch.eval(iv, gens[g])
    ctx.synchronize()
    t0 = time.perf_counter()
    sims = None
    for g in range(warmup, warmup + steps):
        sims = batch.eval(iv, gens[g])
    ctx.synchronize()
    t = time.perf_counter() - t0
    batch.set_options(use_graph=False, time_kernels=True)
    n_t = min(5, len(gens))
    drr_ms = []
    for g in range(n_t):
        batch.eval(iv, gens[g])
        drr_ms.append(batch.last_timing()[0])
    batch.set_options(use_graph=True, time_kernels=False)
    counts = [batch.sample_count(iv, gens[g], evaluated=True) for g in range(n_t)]
    k0 = 1 if n_t > 1 else 0   # the first eager launch is not averaged
    return {"poses_per_s": steps * lam / t, "seconds": t, "tune": tune, "sims": sims,
            "drr_ms": float(np.mean(drr_ms[k0:])), "S": float(np.mean([c[0] for c in counts[k0:]])),
            "S_eval": float(np.mean([c[1] for c in counts[k0:]]))}


def interp_leg(ctx, n_poses, reps=3):
    """The CPU caster's other sampling options (xrh_line_int_interp) on the 128x128x96 @2 mm
    phantom, 192^2 detector, 1 mm steps: n_poses poses per call, wall clock around synchronised
    calls (minimum of `reps` after one untimed call that also builds the volume's grids), against
    the trilinear cast of the same call."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    w = synth.config("c1", n_poses=n_poses)
    cam = w.cams[0]
    poses = w.poses.astype(np.float32).reshape(-1, 16)
    c4p = np.zeros(n_poses, np.int64)
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    buf = ctx.alloc(4 * n_poses * cam.num_det_rows * cam.num_det_cols)
    cases = (("linear", X.XRH_INTERP_LINEAR, 0), ("nn", X.XRH_INTERP_NN, 0), ("bspline", X.XRH_INTERP_BSPLINE, 0),
             ("sinc", X.XRH_INTERP_SINC, 0), ("linear_aa4", X.XRH_INTERP_LINEAR, 4), ("nn_aa4", X.XRH_INTERP_NN, 4))
    out = {"volume": "128x128x96@2mm", "detector": "192x192@1.552mm", "step_mm": w.step, "poses": n_poses}
    for name, interp, aa in cases:
        best = None
        for r in range(reps + 1):
            ctx.synchronize()
            t0 = time.perf_counter()
            X.line_int_interp(ctx, vol, [cam], poses, c4p, w.step, buf, interp, aa_rays=aa, aa_seed=5)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            if r > 0:
                best = dt if best is None else min(best, dt)
        out[name] = {"ms_per_call": round(best * 1e3, 3), "poses_per_s": round(n_poses / best, 1)}
        print(f"interp {name}: {out[name]}", file=sys.stderr, flush=True)
    out["grid_bytes"] = vol.nbytes
    buf.free()
    vol.close()
    return out


def c4_leg(ctx, lam, steps, geom, smooth):
    """Reduced-precision volume storage: the seeded 512x512x400 phantom at 0.5 mm stored as FP32,
    FP16 and BF16 (identical contents), 768^2 detector at 0.388 mm, 0.25 mm steps, patch
    gradient-NCC against the FP32 ground-truth DRR plus 1% noise; lambda poses per generation,
    one untimed generation, then `steps` timed generations per storage type."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    w = synth.config("c4", n_poses=1)
    cam = w.cams[0]
    P = cam.num_det_rows * cam.num_det_cols
    gens = generation_poses(w, lam, 1 + steps, seed=71)
    out = {"volume": "512x512x400@0.5mm", "detector": "768x768@0.388mm", "step_mm": w.step, "lambda": lam,
           "poses_timed": lam * steps, "geometry": "fp64" if geom == X.XRH_GEOM_FP64 else "fp32"}
    fixed = None
    for name, st in (("f32", X.XRH_F32), ("f16", X.XRH_F16), ("bf16", X.XRH_BF16)):
        vol = X.Volume(ctx, w.vol, w.idx_to_phys, storage=st)
        if fixed is None:
            buf = ctx.alloc(4 * P)
            X.line_int(ctx, vol, w.cams, w.gt.astype(np.float32)[None], np.zeros(1, np.int64), w.step, buf,
                       geom=geom)
            fixed = synth.noisy_fixed(buf.download(np.float32).reshape(cam.num_det_rows, cam.num_det_cols))
            buf.free()
        fx = X.Fixed(ctx, fixed, X.XRH_SIM_PATCH_GRAD_NCC, smooth_ksize=smooth, patch_radius=5)
        batch = X.Batch(ctx, [X.make_view(vol, fx, cam, w.step)], max_items=lam, geom=geom)
        r = batched_leg(ctx, batch, lam, gens, 1, steps, autotune=False)
        out[name] = {"poses_per_s": round(r["poses_per_s"], 1), "drr_ms_per_launch": round(r["drr_ms"], 3),
                     "samples_per_launch": int(r["S"]), "samples_evaluated_per_launch": int(r["S_eval"]),
                     "volume_bytes": vol.nbytes, "sim_checksum": float(np.sum(r["sims"]))}
        print(f"c4 {name}: {out[name]}", file=sys.stderr, flush=True)
        for o in (batch, fx, vol):
            o.close()
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

    def note(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

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
    gens = generation_poses(w, lam, args.warmup + args.steps)

    # ---- batched (the headline): DRR launch variant tuned on the first generation (untimed),
    # warmup, then exactly K generations between synchronisations ----
    batch = X.Batch(ctx, [X.make_view(vol, fx, cam, w.step)], max_items=lam, geom=geom)
    main_leg = batched_leg(ctx, batch, lam, gens, args.warmup, args.steps, not args.no_autotune)
    value, tune, sims = main_leg["poses_per_s"], main_leg["tune"], main_leg["sims"]
    t_batched = main_leg["seconds"]
    drr_ms_avg, S, S_eval = main_leg["drr_ms"], main_leg["S"], main_leg["S_eval"]
    note(f"batched {value:.1f} poses/s, DRR {drr_ms_avg:.3f} ms per launch")

    # ---- the same generations with FP32 ray geometry (the reference's float set-up, step
    # counts identical to RayCasterLineIntCPU's) ----
    other_geom = X.XRH_GEOM_FP32 if geom == X.XRH_GEOM_FP64 else X.XRH_GEOM_FP64
    b2 = X.Batch(ctx, [X.make_view(vol, fx, cam, w.step)], max_items=lam, geom=other_geom)
    alt = batched_leg(ctx, b2, lam, gens, args.warmup, args.steps, not args.no_autotune)
    b2.close()
    note(f"batched, {'fp32' if other_geom == X.XRH_GEOM_FP32 else 'fp64'} geometry: {alt['poses_per_s']:.1f} poses/s")

    bv = {X.XRH_F32: 4, X.XRH_F16: 2, X.XRH_BF16: 2}[storage]
    bytes_launch = 8.0 * bv * S + 4.0 * P * lam  # gathers + DRR write
    algorithmic_gbs = bytes_launch / (drr_ms_avg * 1e-3) / 1e9

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
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_from(src)
    ctx.synchronize()
    copy_gbs = 5 * 2.0 * (copy_mib << 20) / (time.perf_counter() - t0) / 1e9
    src.free()
    dst.free()
    note(f"plain {plain_value}, drop-in {dropin_value:.1f} poses/s, copy {copy_gbs:.0f} GB/s")

    served = None
    if args.serve_jobs > 0:
        served = serve_leg(ctx, vol, w, args.serve_jobs, args.serve_rate, args.serve_gens, lam, geom, args.smooth,
                           args.serve_policy, args.serve_step_samples)
        note(f"served C2 class: p50 {served['p50_s']} s, p99 {served['p99_s']} s")

    c3 = None if args.no_c3 else c3_leg(ctx, geom, args.smooth)

    c4 = None
    if args.c4_steps > 0 and args.config == "c2":
        c4 = c4_leg(ctx, lam, args.c4_steps, geom, args.smooth)

    interp = interp_leg(ctx, args.interp_poses) if args.interp_poses > 0 else None

    served_mix = None
    if args.serve_mix_jobs > 0:
        served_mix = serve_mix_leg(ctx, args.serve_mix_jobs, args.serve_mix_rate, geom, args.smooth,
                                   args.serve_policy, args.serve_step_samples)

    cpu = None
    if not args.no_cpu:
        note("cpu baseline")
        cpu = cpu_baseline(w, fixed, np.concatenate(gens[args.warmup:]), args.smooth)

    # ---- roofline of the DRR kernel (the dominant kernel) ----
    # The DRR's gathers are served mostly by L1/L2/Infinity Cache (the C2 volume is cache
    # resident), so the algorithmic gather bytes of the reference algorithm exceed what HBM
    # moves; the bound that the counters show is vector-ALU issue. `achieved` is the kernel's
    # VALU wave-instructions per launch (SQ_INSTS_VALU, rocprofv3, same configuration and kernel
    # source) over its live-measured duration; `peak` is one VALU wave-instruction per CU per
    # clock at the 2.4 GHz peak engine clock. `hbm_frac` is the measured HBM-side traffic
    # (FETCH_SIZE + WRITE_SIZE, calibrated) over the same duration against 8 TB/s.
    key = (f"workload={workload} storage={args.storage} geometry={args.geom} skip={0 if args.no_skip else 1} "
           f"smooth={args.smooth} src={kernel_source_key()}")
    prof = profiled(key)
    cus = device_cus()
    valu_peak = cus * PEAK_CLOCK_GHZ  # G wave-instructions / s
    valu_rate = prof["valu_insts"] / (drr_ms_avg * 1e-3) / 1e9 if "valu_insts" in prof else None
    traffic = prof.get("traffic_bytes")
    roofline = {"bound": "valu", "achieved": round(valu_rate, 1) if valu_rate else None, "peak": valu_peak,
                "unit": "G VALU wave-instructions/s", "frac": round(valu_rate / valu_peak, 4) if valu_rate else None,
                "traffic": traffic,
                "hbm_gbs": round(traffic / (drr_ms_avg * 1e-3) / 1e9, 1) if traffic else None,
                "hbm_frac": round(traffic / (drr_ms_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                "hbm_peak_gbs": HBM_PEAK_GBS, "profile_key": key,
                "kernel": "drr_kernel", "kernel_ms": round(drr_ms_avg, 4), "cus": cus,
                "valu_insts_per_launch": prof.get("valu_insts"), "vmem_insts_per_launch": prof.get("vmem_insts"),
                "fetch_unit_busy_frac": prof.get("ta_busy_frac"),
                "samples_per_launch": int(S), "samples_evaluated_per_launch": int(S_eval),
                "valu_lane_ops_per_evaluated_sample": (round(64.0 * prof["valu_insts"] / S_eval, 2)
                                                       if "valu_insts" in prof else None),
                "algorithmic_bytes_per_launch": int(bytes_launch),
                "algorithmic_gbs": round(algorithmic_gbs, 1),
                "evaluated_gbs": round((8.0 * bv * S_eval + 4.0 * P * lam) / (drr_ms_avg * 1e-3) / 1e9, 1),
                "measured_copy_gbs": round(copy_gbs, 1)}

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
        f"{'fp32' if other_geom == X.XRH_GEOM_FP32 else 'fp64'}_geometry_poses_per_s": round(alt["poses_per_s"], 2),
        f"{'fp32' if other_geom == X.XRH_GEOM_FP32 else 'fp64'}_geometry_drr_ms": round(alt["drr_ms"], 4),
        "roofline": roofline,
        "empty_space_skipping": not args.no_skip,
        "autotune": tune,
        "served": served,
        "served_mix": served_mix,
        "c3_registration": c3,
        "c4": c4,
        "interp": interp,
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
