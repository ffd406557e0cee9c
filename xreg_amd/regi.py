"""Registration drivers over the batched objective: CMA-ES, exhaustive (grid) search, a
BOBYQA-pattern local refiner, and a multi-resolution driver.

The objective follows Intensity2D3DRegi::obj_fn (lib/regi/interfaces_2d_3d/
xregIntensity2D3DRegi.cpp:571-742): optimiser vectors -> poses (an SE3OptVars parameterisation
inside the inter-frame sandwich, optvars.py; by default the se(3) Lie algebra about a rotation
centre on the volume side of a starting pose), all views x all candidates cast and scored in
one batched submission
(view-major, :605-621), then the per-view values combined by ImgSimMetric2DCombineMean
(lib/regi/sim_metrics_2d/xregImgSimMetric2DCombine.cpp:79-98: float sum in view order, divided
by the number of views).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _xrh as X
from . import geometry as G
from .cmaes import CMAES
from .optvars import InterFramePose, SE3OptVars


def combine_mean(per_view: np.ndarray) -> np.ndarray:
    """ImgSimMetric2DCombineMean: float32 sum over views (axis 0, in order) / number of views."""
    per_view = np.asarray(per_view, dtype=np.float32)
    acc = np.zeros(per_view.shape[1], dtype=np.float32)
    for v in range(per_view.shape[0]):
        acc = (acc + per_view[v]).astype(np.float32)
    return (acc / np.float32(per_view.shape[0])).astype(np.float32)


class PoseParam(InterFramePose):
    """The usual set-up of the hip-surgery drivers: optimiser vector (default 6: se(3) Lie
    algebra, rotation [rad] then translation [mm]) about an intermediate frame at `center`
    (volume axes), given w.r.t. the volume, around the guess `init_pose`:
    pose = C * T(x) * C^-1 * init_pose."""

    def __init__(self, init_pose: np.ndarray, center, opt_vars: SE3OptVars | None = None):
        self.init = np.asarray(init_pose, dtype=np.float64)
        self.center = np.asarray(center, dtype=np.float64)
        super().__init__(self.init, opt_vars, G.rigid(None, self.center), inter_frame_wrt_vol=True)


class Objective:
    """Multi-view similarity of candidate parameter vectors, one batched submission per call.

    views: list of (camera, fixed image) pairs sharing one volume. Fixed images are processed
    once (xrh_fixed_create); every call casts len(X) x V DRRs and scores them."""

    def __init__(self, ctx: X.Context, vol: X.Volume, views, param: PoseParam, step: float,
                 metric: int = X.XRH_SIM_PATCH_GRAD_NCC, smooth_ksize: int = 5, patch_radius: int = 5,
                 max_batch: int = 256, geom: int = X.XRH_GEOM_FP32, use_graph: bool = True,
                 autotune: bool = False):
        """autotune: tune the DRR launch variant (xrh_batch_autotune) on each new submission
        size the first time it is evaluated (values are unchanged; the first call of each size
        then costs 26 x 3 extra submissions)."""
        self.ctx, self.vol, self.param, self.step = ctx, vol, param, float(step)
        self.autotune = bool(autotune)
        self.tuned_sizes = {}     # poses per submission -> xrh_batch_autotune result
        self.fixed = [X.Fixed(ctx, img, metric, smooth_ksize=smooth_ksize, patch_radius=patch_radius)
                      for _, img in views]
        self.cams = [c for c, _ in views]
        self.V = len(views)
        self.max_batch = int(max_batch)
        self.batch = X.Batch(ctx, [X.make_view(vol, f, c, step) for c, f in zip(self.cams, self.fixed)],
                             max_items=self.V * self.max_batch, geom=geom)
        self.batch.set_options(use_graph=use_graph, time_kernels=False)
        self.evals = 0        # candidate poses evaluated
        self.calls = 0        # batched submissions
        self.metric = int(metric)

    def sims_aux(self):
        """The metrics' debug records (ImgSimMetric2D::aux_info, xregImgSimMetric2D.h:134-136):
        per view, for patch metrics, the patch list each compute call used -- one call per view
        per submission here, as one ImgSimMetric2D::compute per obj_fn call in the reference --
        else None (the reference's NCC and gradient-NCC metrics keep none)."""
        return metric_aux_records(self.fixed, self.metric, self.calls)

    def eval_poses(self, poses: np.ndarray) -> np.ndarray:
        poses = np.asarray(poses, dtype=np.float64).reshape(-1, 4, 4)
        out = np.empty(len(poses), dtype=np.float32)
        for s in range(0, len(poses), self.max_batch):
            p = poses[s:s + self.max_batch]
            k = len(p)
            iv = np.repeat(np.arange(self.V, dtype=np.int32), k)
            ip = np.tile(p, (self.V, 1, 1))
            if self.autotune and k not in self.tuned_sizes:
                self.tuned_sizes[k] = self.batch.autotune(iv, ip, reps=2)
            sims = self.batch.eval(iv, ip)
            out[s:s + k] = combine_mean(sims.reshape(self.V, k))
            self.calls += 1
        self.evals += len(poses)
        return out

    def __call__(self, Xp: np.ndarray) -> np.ndarray:
        Xp = np.atleast_2d(np.asarray(Xp, dtype=np.float64))
        return self.eval_poses(self.param(Xp))

    def close(self):
        self.batch.close()
        for f in self.fixed:
            f.close()


def metric_aux_records(fixed, metric: int, calls: int):
    """Per fixed image: for the patch metrics the record ImgSimMetric2DPatchNCCCPU keeps with
    save_aux_info (xregImgSimMetric2DPatchNCCCPU.cpp:290-298: every compute call appends the
    patch list and the indices of the patches used -- all of them, random-patch mode being off),
    and for patch gradient-NCC the same for the x and y gradients
    (xregImgSimMetric2DPatchGradNCCCPU.cpp:246-252); None for the other metrics."""
    out = []
    for f in fixed:
        if metric not in (X.XRH_SIM_PATCH_NCC, X.XRH_SIM_PATCH_GRAD_NCC):
            out.append(None)
            continue
        pl = f.patches()
        rec = {"type": "patch-ncc-aux", "calls": int(calls), "patches": pl,
               "patch_inds": np.arange(len(pl["weights"]), dtype=np.uint64)}
        if metric == X.XRH_SIM_PATCH_GRAD_NCC:
            rec = {"type": "patch-grad-ncc-aux", "x": rec, "y": dict(rec)}
        out.append(rec)
    return out


class MultiParam:
    """Several objects' parameterisations stacked into one optimiser vector (object o owns the
    next params[o].num_params entries), as Intensity2D3DRegi concatenates the per-object
    SE3OptVars. Called on one vector it returns the list of poses; on k vectors a list of
    (k, 4, 4) arrays, one per object."""

    def __init__(self, params):
        self.params = list(params)
        self.sizes = [p.num_params for p in self.params]
        self.num_params = int(sum(self.sizes))

    def __call__(self, Xp: np.ndarray):
        Xp = np.asarray(Xp, dtype=np.float64)
        out, o = [], 0
        for p, n in zip(self.params, self.sizes):
            out.append(p(Xp[..., o:o + n]))
            o += n
        return out


class MultiObjective:
    """Intensity2D3DRegi::obj_fn with several objects (xregIntensity2D3DRegi.cpp:571-696): per
    candidate, every object's DRR is cast into the same projections (view-major, one per view
    and candidate), the first object with the REPLACE store -- on top of the static volumes'
    background images when `bg` is given (has_a_static_vol_, :593-625) -- and the others with
    ACCUM; then each view is scored and the views are averaged (ImgSimMetric2DCombineMean).
    All cameras share one detector size, as RayCasterCPU requires (xregRayCastBaseCPU.cpp:
    54-70).

    batched=True (default): one xrh_batch_eval_objects submission per call covers every view,
    candidate and object (the objects of a candidate are cast in order inside one launch,
    replayed from a graph when the shape repeats). batched=False: one xrh_line_int launch per
    object over all views and candidates, then one metric call per view -- the reference's
    call pattern. Both give the same DRRs and values."""

    def __init__(self, ctx: X.Context, vols, views, params, step: float,
                 metric: int = X.XRH_SIM_PATCH_GRAD_NCC, smooth_ksize: int = 5, patch_radius: int = 5,
                 max_batch: int = 64, geom: int = X.XRH_GEOM_FP32, bg=None, batched: bool = True):
        if len(vols) != len(params) or not vols:
            raise ValueError("one parameterisation per object")
        self.ctx, self.vols, self.step, self.geom = ctx, list(vols), float(step), int(geom)
        self.param = MultiParam(params)
        self.cams = [c for c, _ in views]
        self.V = len(views)
        self.O = len(self.vols)
        rows, cols = self.cams[0].num_det_rows, self.cams[0].num_det_cols
        if any((c.num_det_rows, c.num_det_cols) != (rows, cols) for c in self.cams):
            raise ValueError("all views must share one detector size")
        self.npx = rows * cols
        self.fixed = [X.Fixed(ctx, img, metric, smooth_ksize=smooth_ksize, patch_radius=patch_radius)
                      for _, img in views]
        self.max_batch = int(max_batch)
        self.bg = None
        if bg is not None:
            b = np.ascontiguousarray(bg, dtype=np.float32).reshape(self.V, rows, cols)
            self.bg = ctx.alloc(b.nbytes)
            self.bg.upload(b)
        self.batched = bool(batched)
        self.drr = None
        self.batch = None
        if self.batched:
            # view table object-major: view o * V + v binds object o's volume to camera v (the
            # fixed image of view v; only object 0's views are scored and carry backgrounds)
            tbl = [X.make_view(vol, self.fixed[v], self.cams[v], self.step) for vol in self.vols for v in range(self.V)]
            self.batch = X.Batch(ctx, tbl, max_items=self.O * self.V * self.max_batch, geom=self.geom)
        else:
            self.drr = ctx.alloc(4 * self.V * self.max_batch * self.npx)
        self.evals = 0
        self.calls = 0
        self.metric = int(metric)

    def sims_aux(self):
        """Per view, the metric's debug record (see Objective.sims_aux)."""
        return metric_aux_records(self.fixed, self.metric, self.calls)

    def eval_poses(self, poses_per_obj, drr: X.DeviceBuffer | None = None) -> np.ndarray:
        """Objective values of n candidates given each object's n poses. drr (batched form only):
        optional device buffer receiving the view-major DRRs of the last chunk."""
        poses_per_obj = [np.asarray(p, dtype=np.float64).reshape(-1, 4, 4) for p in poses_per_obj]
        n = len(poses_per_obj[0])
        out = np.empty(n, dtype=np.float32)
        for s in range(0, n, self.max_batch):
            k = min(self.max_batch, n - s)
            if self.batched:
                # items view-major (view v, candidate c) -> entries (object o): view o * V + v
                ev = (np.arange(self.O)[None, None, :] * self.V + np.arange(self.V)[:, None, None]
                      + np.zeros((1, k, 1), np.int64)).reshape(self.V * k, self.O)
                ep = np.stack([np.broadcast_to(poses_per_obj[o][s:s + k][None], (self.V, k, 4, 4))
                               for o in range(self.O)], axis=2).reshape(self.V * k, self.O, 4, 4)
                sims = self.batch.eval_objects(ev, ep, bg=self.bg, drr=drr).reshape(self.V, k)
            else:
                c4p = np.repeat(np.arange(self.V, dtype=np.int64), k)
                for o, vol in enumerate(self.vols):
                    xf = np.tile(poses_per_obj[o][s:s + k], (self.V, 1, 1)).astype(np.float32)
                    X.line_int(self.ctx, vol, self.cams, xf, c4p, self.step, self.drr,
                               store=X.XRH_STORE_REPLACE if o == 0 else X.XRH_STORE_ACCUM,
                               bg=self.bg if o == 0 else None, geom=self.geom)
                sims = np.stack([self.fixed[v].compute(self.drr, k, offset_imgs=v * k) for v in range(self.V)])
            out[s:s + k] = combine_mean(sims)
            self.calls += 1
        self.evals += n
        return out

    def __call__(self, Xp: np.ndarray) -> np.ndarray:
        Xp = np.atleast_2d(np.asarray(Xp, dtype=np.float64))
        return self.eval_poses(self.param(Xp))

    def close(self):
        if self.batch is not None:
            self.batch.close()
        if self.drr is not None:
            self.drr.free()
        if self.bg is not None:
            self.bg.free()
        for f in self.fixed:
            f.close()


# ------------------------------------------------------------------------------------------
# drivers


@dataclass
class RegiResult:
    x: np.ndarray
    fun: float
    pose: np.ndarray
    evals: int
    history: list = field(default_factory=list)
    extra: dict = field(default_factory=dict)


def run_cmaes(obj: Objective, lam: int, sigma, max_gens: int, seed: int, x0=None, bounds=None,
              tol_fun: float = 1e-12, tol_x: float = 1e-11) -> RegiResult:
    """Intensity2D3DRegiCMAES::run (xregIntensity2D3DRegiCMAES.cpp:75-261): lambda candidates per
    generation evaluated in one submission, bounds by resampling. As in the reference the
    estimate is the distribution mean after the last generation (update_regi_xforms(
    xforms_at_cur_mean()), :248); `fun` is the smallest value seen (its debug_sim_val_, :219).
    Per generation, extra["iter_means"] holds the mean after the update and extra["iter_best"]
    the smallest value so far -- the iter_vars / sims of the debug record (:340-384).
    extra["best_x"] / ["best_pose"] keep the best candidate evaluated."""
    x0 = np.zeros(obj.param.num_params) if x0 is None else np.asarray(x0, dtype=np.float64)
    es = CMAES(x0, np.asarray(sigma, dtype=np.float64), lam, seed,
               bounds=None if bounds is None else np.asarray(bounds, dtype=np.float64), tol_fun=tol_fun, tol_x=tol_x)
    best_f, best_x = np.inf, x0.copy()
    iter_means, iter_best = [], []
    aux = {"name": "CMA-ES", "pop_size": int(lam), "pop_params": [], "sim_vals": [], "cov_mats": [],
           "num_rejects": [], "obj_dims": list(getattr(obj.param, "sizes", [obj.param.num_params]))}
    e0 = obj.evals
    for _ in range(max_gens):
        r0 = es.num_rejects
        Xc = es.ask()
        aux["num_rejects"].append(es.num_rejects - r0)
        f = obj(np.asarray(Xc)).astype(np.float64)
        k = int(np.argmin(f))
        if f[k] < best_f:
            best_f, best_x = float(f[k]), np.asarray(Xc[k]).copy()
        es.tell(f)
        # the CMA-ES opt-aux record (xregIntensity2D3DRegiCMAES.cpp:225-243): population,
        # values and the covariance after the update
        aux["pop_params"].append(np.asarray(Xc, np.float64).copy())
        aux["sim_vals"].append(f.copy())
        aux["cov_mats"].append(es.C.copy())
        iter_means.append(es.m.copy())
        iter_best.append(best_f)
        if es.stop():
            break
    xm = es.m.copy()
    return RegiResult(xm, best_f, obj.param(xm), obj.evals - e0, list(es.history),
                      extra={"best_x": best_x, "best_pose": obj.param(best_x),
                             "iter_means": np.array(iter_means), "iter_best": np.array(iter_best),
                             "opt_aux": aux})


def mesh_grid(axes) -> np.ndarray:
    """All combinations of per-parameter sample vectors, first parameter most significant
    (row-major, as the reference's mesh-grid iterator)."""
    mg = np.meshgrid(*[np.asarray(a, dtype=np.float64) for a in axes], indexing="ij")
    return np.stack([m.reshape(-1) for m in mg], axis=1)


def run_exhaustive(obj: Objective, candidates: np.ndarray, batch: int | None = None,
                   save_all: bool = True) -> RegiResult:
    """Intensity2D3DRegiExhaustive::run (xregIntensity2D3DRegiExhaustive.cpp:77-300): the candidate
    list is evaluated in consecutive batches of `batch` (the objective's batch size by
    default); the minimum is updated only on a strictly smaller value, so the first of equal
    minima wins; optionally every value is kept."""
    cand = np.asarray(candidates, dtype=np.float64).reshape(-1, obj.param.num_params)
    b = obj.max_batch if batch is None else int(batch)
    all_vals = []
    best_f, best_i = np.float32(np.finfo(np.float32).max), -1
    e0 = obj.evals
    for s in range(0, len(cand), b):
        vals = obj(cand[s:s + b])
        for j, v in enumerate(vals):
            if best_f > v:
                best_f, best_i = v, s + j
        if save_all:
            all_vals.append(vals)
    x = cand[best_i] if best_i >= 0 else np.zeros(obj.param.num_params)
    all_v = np.concatenate(all_vals) if all_vals else np.zeros(0, np.float32)
    # the Exhaustive opt-aux record (xregIntensity2D3DRegiExhaustive.cpp:397-440): the candidate
    # poses per object, the minimum and, with every value kept, all values and the index
    poses = obj.param(cand) if len(cand) else []
    aux = {"name": "Exhaustive", "cam_wrt_vols": poses if isinstance(poses, list) else [poses],
           "has_min": best_i >= 0, "min_sim_val": float(best_f), "has_all_sim_vals": bool(save_all),
           "min_sim_val_idx": int(best_i), "all_sim_vals": all_v}
    return RegiResult(x, float(best_f), obj.param(x), obj.evals - e0,
                      extra={"all_sim_vals": all_v, "best_index": best_i, "opt_aux": aux})


def run_local_refine(obj: Objective, x0, rho_beg, rho_end: float = 1e-3, max_evals: int = 200,
                     bounds=None) -> RegiResult:
    """A derivative-free trust-region refiner with BOBYQA's evaluation pattern
    (Intensity2D3DRegiBOBYQA, xregIntensity2D3DRegiBOBYQA.cpp:29-39, through NLopt): 2n+1
    initial single-pose evaluations (x0, x0 +- rho e_i), then one single-pose evaluation per
    iteration. The model is a quadratic with diagonal curvature fitted (least squares) to the
    best 2n+1..4n points seen; the step minimises it within a box trust region (scaled by
    rho_beg per parameter) and the region follows the ratio of actual to predicted decrease.
    NLopt's BOBYQA (Powell's full quadratic model, Powell 2009) is not reproduced; this
    refiner keeps its calling pattern and bound handling."""
    x0 = np.asarray(x0, dtype=np.float64)
    n = x0.size
    scale = np.broadcast_to(np.asarray(rho_beg, dtype=np.float64), (n,)).copy()
    lo = hi = None
    if bounds is not None:
        bnd = np.asarray(bounds, dtype=np.float64).reshape(n, 2)
        lo, hi = bnd[:, 0], bnd[:, 1]

    def clip(x):
        return x if lo is None else np.minimum(np.maximum(x, lo), hi)

    e0 = obj.evals
    pts, vals = [], []

    def f(x):
        v = float(obj(x[None])[0])
        pts.append(x.copy())
        vals.append(v)
        return v

    # initial stencil, one pose per evaluation
    fx = f(clip(x0))
    for i in range(n):
        for sgn in (1.0, -1.0):
            e = np.zeros(n)
            e[i] = sgn * scale[i]
            f(clip(x0 + e))
    delta = 1.0           # trust region radius in units of scale
    best = int(np.argmin(vals))
    xb, fb = pts[best].copy(), vals[best]
    hist = [fx]
    while obj.evals - e0 < max_evals and delta > rho_end:
        # fit g, diag(H) about xb on the nearest points (scaled coordinates)
        P = (np.asarray(pts) - xb) / scale
        d2 = np.sum(P * P, axis=1)
        m = min(len(pts), 4 * n)
        idx = np.argsort(d2)[:m]
        A = np.concatenate([np.ones((m, 1)), P[idx], 0.5 * P[idx] ** 2], axis=1)
        coef, *_ = np.linalg.lstsq(A, np.asarray(vals)[idx] - fb, rcond=None)
        g, h = coef[1:n + 1], coef[n + 1:]
        s = np.where(h > 1e-12, -g / np.maximum(h, 1e-12), -np.sign(g) * delta)
        s = np.clip(s, -delta, delta)
        pred = -(g @ s + 0.5 * np.sum(h * s * s))
        xn = clip(xb + s * scale)
        if not np.any(np.abs(xn - xb) > 0) or pred <= 0:
            delta *= 0.5
            continue
        fn = f(xn)
        ratio = (fb - fn) / pred
        if fn < fb:
            xb, fb = xn, fn
        delta = min(2.0 * delta, 2.0) if ratio > 0.7 else (delta if ratio > 0.1 else 0.5 * delta)
        hist.append(fb)
    return RegiResult(xb, fb, obj.param(xb), obj.evals - e0, hist)


# ------------------------------------------------------------------------------------------
# multi-resolution


@dataclass
class Level:
    ds_factor: float                   # detector downsampling factor (1 = full resolution)
    method: str = "cmaes"              # "cmaes" | "exhaustive" | "refine"
    opts: dict = field(default_factory=dict)


def run_multi_level(ctx, vol, cams, fixed_imgs, init_pose, center, step, levels, downsample_image,
                    metric: int = X.XRH_SIM_PATCH_GRAD_NCC, smooth_ksize: int = 5, patch_radius: int = 5,
                    geom: int = X.XRH_GEOM_FP32, max_batch: int = 256, autotune: bool = False):
    """MultiLevelMultiObjRegi::run (lib/regi/interfaces_2d_3d/xregMultiObjMultiLevel2D3DRegi.cpp:
    164-531), single object: at each level the cameras are downsampled (DownsampleCameraModel)
    and the fixed images with `downsample_image(img, ds)`; the level's registration starts
    from the previous level's estimate, with its own objective (caster + metrics). Each
    result's extra["wall_s"] is the level's wall-clock time, set-up included. autotune: each
    level's objective tunes the DRR launch variant for its detector size (Objective)."""
    pose = np.asarray(init_pose, dtype=np.float64)
    results = []
    for lvl in levels:
        t0 = time.perf_counter()
        ds = float(lvl.ds_factor)
        lcams = [c if ds == 1.0 else G.downsample_camera(c, ds) for c in cams]
        limgs = [img if ds == 1.0 else downsample_image(img, ds) for img in fixed_imgs]
        for c, im in zip(lcams, limgs):
            if im.shape != (c.num_det_rows, c.num_det_cols):
                raise ValueError(f"level {ds}: image {im.shape} does not match camera "
                                 f"{c.num_det_rows}x{c.num_det_cols}")
        obj = Objective(ctx, vol, list(zip(lcams, limgs)), PoseParam(pose, center), step, metric=metric,
                        smooth_ksize=smooth_ksize, patch_radius=patch_radius, geom=geom, max_batch=max_batch,
                        autotune=autotune)
        try:
            o = dict(lvl.opts)
            if lvl.method == "cmaes":
                r = run_cmaes(obj, **o)
            elif lvl.method == "exhaustive":
                r = run_exhaustive(obj, **o)
            elif lvl.method == "refine":
                r = run_local_refine(obj, np.zeros(obj.param.num_params), **o)
            else:
                raise ValueError(lvl.method)
        finally:
            obj.close()
        r.extra["wall_s"] = time.perf_counter() - t0
        if autotune:
            r.extra["autotune"] = dict(obj.tuned_sizes)
        results.append(r)
        pose = r.pose
    return pose, results


def debug_results(res: RegiResult, param: InterFramePose, vol_index: int = 0, projs_used=(),
                  regi_time_secs: float | None = None, sims_aux=None):
    """The SingleRegiDebugResults record of one CMA-ES run (xregIntensity2D3DRegiDebug.h:45-141,
    filled as Intensity2D3DRegi::run and the CMA-ES debug hooks fill it): parameterisation,
    guess, final pose, intermediate frame, per-generation means and best-so-far values, the
    optimiser's record (population, values, covariance and rejections per generation) and, when
    given, the metrics' records (Objective.sims_aux()). Write it with h5io.write_regi_debug."""
    from . import h5io

    means = res.extra.get("iter_means")
    best = res.extra.get("iter_best")
    if means is None or best is None:
        raise ValueError("debug record needs a CMA-ES result (per-generation means and values)")
    return h5io.RegiDebugResults(
        se3_params=param.vars, vols_used=[int(vol_index)], init_poses=[np.asarray(param.guess, np.float32)],
        final_poses=[np.asarray(res.pose, np.float32)], inter_frames=[np.asarray(param.inter_frame, np.float32)],
        inter_frames_wrt_vol=[param.inter_frame_wrt_vol], iter_vars=[np.asarray(means, np.float32)],
        sims=np.asarray(best, np.float32), projs_used=list(projs_used), regi_time_secs=regi_time_secs,
        opt_aux=res.extra.get("opt_aux"), sims_aux=list(sims_aux) if sims_aux is not None else [])
