"""Continuous-batching 2D/3D registration server on one GPU.

Jobs (single- or multi-view intensity-based registrations with CMA-ES, the loop of
Intensity2D3DRegiCMAES::run, lib/regi/interfaces_2d_3d/xregIntensity2D3DRegiCMAES.cpp:75-261)
arrive over time. Volumes stay resident; each admitted job owns its fixed images and CMA-ES
state. At every scheduling step the server packs the pending items of admitted jobs (a job's
generation is lambda poses x its views, any mix of volumes, detector sizes and lambda; a large
generation may be spread over several steps) into one batched submission
(xrh_batch_set_views + xrh_batch_eval); once a job's generation is complete it gets its lambda
objective values (mean over its views, ImgSimMetric2DCombineMean) and its optimiser advances.
Jobs are admitted while the context's HBM use stays under the budget; a job's results do not
depend on what it was batched with (every item is evaluated independently with fixed-order
reductions).
"""
from __future__ import annotations

import heapq
import time
from dataclasses import dataclass, field

import numpy as np

from . import _xrh as X
from . import geometry as G
from .cmaes import CMAES
from .regi import combine_mean


@dataclass
class JobSpec:
    job_id: int
    arrival: float                 # seconds since the server clock started
    volume: str                    # key of a resident volume
    cams: list                     # xrh Camera per view
    fixed: list                    # float32 fixed image per view
    init_pose: np.ndarray          # 4x4 float64 cam->phys starting point
    center: np.ndarray             # rotation centre (physical)
    step: float
    lam: int
    max_gens: int
    seed: int
    sigma: tuple = (np.deg2rad(10.0),) * 3 + (15.0,) * 3
    bounds: tuple | None = None
    metric: int = X.XRH_SIM_PATCH_GRAD_NCC
    smooth_ksize: int = 5
    patch_radius: int = 5
    max_wall_s: float | None = None  # finish with the current estimate once this long in service


@dataclass
class JobResult:
    job_id: int
    arrival: float
    start: float
    finish: float
    generations: int
    best_sim: float
    best_params: np.ndarray
    best_pose: np.ndarray
    history: list = field(default_factory=list)
    params: np.ndarray | None = None   # final estimate: the CMA-ES mean
    pose: np.ndarray | None = None
    status: str = "max_gens"           # "max_gens", "converged" (CMA-ES stop test), "deadline", or
                                       # "rejected" (can never be admitted: see `reason`)
    reason: str = ""
    first_poses: np.ndarray | None = None  # the first generation's lambda poses (lambda, 4, 4)
    first_sims: np.ndarray | None = None   # and their similarity values per view (views, lambda)

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def pose_from_params(init_pose: np.ndarray, center: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Lie-algebra delta about the rotation centre, applied on the volume side of the starting
    pose (SE3OptVarsLieAlg, lib/regi/xregSE3OptVars.cpp:127-137)."""
    c = np.asarray(center, dtype=np.float64)
    return G.rigid(None, c) @ G.exp_se3(x) @ G.rigid(None, -c) @ init_pose


def poses_from_params(init_pose: np.ndarray, center: np.ndarray, X: np.ndarray) -> np.ndarray:
    """pose_from_params over the rows of X, (n, 4, 4)."""
    c = np.asarray(center, dtype=np.float64)
    return G.rigid(None, c) @ G.exp_se3_batch(X) @ (G.rigid(None, -c) @ init_pose)


class _Active:
    """An admitted job: its fixed images and views on the device, its CMA-ES state, and how far
    the current generation's items (view-major, lambda per view) have been evaluated."""

    def __init__(self, spec: JobSpec, fixeds, views, start: float, extent_mm: float = 1.0, layout_key=(0, 0)):
        self.spec = spec
        self.layout_key = layout_key   # (storage, layout) of the job's volume
        self.fixeds = fixeds
        self.views = views
        self.start = start
        # trilinear samples per item, estimated: the view's detector pixels x (volume extent /
        # ray step); a generation is lambda items per view
        per_ray = max(1.0, extent_mm / float(spec.step))
        self.view_work = [float(int(c.num_det_rows) * int(c.num_det_cols)) * per_ray for c in spec.cams]
        self.gen_work = float(spec.lam) * sum(self.view_work)
        self.n_items = spec.lam * len(spec.cams)
        self.es = CMAES(np.zeros(6), np.asarray(spec.sigma), spec.lam, spec.seed,
                        bounds=None if spec.bounds is None else np.asarray(spec.bounds))
        self.gen = 0
        self.first = (None, None)
        self.best = (np.inf, None)
        self.status = "max_gens"
        self._new_generation(self.es.ask())

    def _new_generation(self, X):
        self.X = X
        self.poses = poses_from_params(self.spec.init_pose, self.spec.center, np.asarray(self.X))
        self.cursor = 0                # next item of this generation to evaluate
        self.sims = np.empty(self.n_items, np.float32)

    def items(self):
        return self.poses

    def item_work(self, i: int) -> float:
        return self.view_work[i // self.spec.lam]

    def chunk(self, budget: float | None, max_items: int, first: bool) -> int:
        """How many of the generation's remaining items fit in `budget` estimated samples and
        `max_items` (at least one when this job opens the submission)."""
        k, work = 0, 0.0
        while self.cursor + k < self.n_items and k < max_items:
            w = self.item_work(self.cursor + k)
            if budget is not None and work + w > budget and not (first and k == 0):
                break
            work += w
            k += 1
        return k

    def remaining_work(self) -> float:
        done = sum(self.item_work(i) for i in range(self.cursor)) if self.cursor else 0.0
        return self.gen_work * (self.spec.max_gens - self.gen) - done

    def close(self):
        for f in self.fixeds:
            f.close()


class RegistrationServer:
    def __init__(self, ctx: X.Context, volumes: dict, hbm_budget: int, max_items: int = 4096,
                 max_views: int = 256, max_det: int = 768, geom: int = X.XRH_GEOM_FP32,
                 device=None, max_active: int | None = None, policy: str = "srpt",
                 step_samples: float | None = 2.0e10):
        """device: the object providing Batch / Fixed / make_view (the xrh binding by default;
        tests substitute a host-side fake to exercise the scheduler without a GPU).
        max_active: at most this many jobs in service at once (None: as many as fit; 1 serves
        the jobs one after another in arrival order, the serial baseline).
        policy: the order in which admitted jobs' items fill a submission. "fifo": admission
        order. "srpt": least estimated remaining work first (samples left in the job's
        generations; ties by arrival). Jobs contribute items in that order while the
        submission's estimated samples stay within step_samples (None: no limit; the first job
        always gets at least one item) and its items within max_items; one submission holds
        only jobs whose volumes share the first job's storage type and layout. "srpt" serves
        short jobs ahead of long ones (lower median latency, and a large generation no longer
        holds up jobs that arrive during it), "fifo" keeps arrival order."""
        if policy not in ("fifo", "srpt"):
            raise ValueError(f"unknown packing policy {policy!r}")
        self.dev = device if device is not None else X
        self.ctx = ctx
        self.volumes = volumes      # name -> resident volume handle
        self.budget = int(hbm_budget)
        self.max_items = max_items
        self.max_det = max_det
        self.batch = self.dev.Batch(ctx, None, max_items, geom=geom, capacity=(max_views, max_det, max_det))
        self.batch.set_options(use_graph=False, time_kernels=False)
        self.max_views = max_views
        self.max_active = max_active
        self.policy = policy
        self.step_samples = step_samples
        self.steps = 0
        self.poses_evaluated = 0
        self.max_items_in_step = 0

    def close(self):
        self.batch.close()

    def _fixed_bytes(self, spec: JobSpec) -> int:
        per = 0
        for f in spec.fixed:
            per += f.size * 4 * (6 if spec.metric in (X.XRH_SIM_GRAD_NCC, X.XRH_SIM_PATCH_GRAD_NCC) else 3)
        return per

    def _admit(self, spec: JobSpec, now: float):
        if self.ctx.bytes_in_use() + self._fixed_bytes(spec) > self.budget:
            return None
        fixeds, views = [], []
        vol = self.volumes[spec.volume]
        try:
            for cam, img in zip(spec.cams, spec.fixed):
                if cam.num_det_rows > self.max_det or cam.num_det_cols > self.max_det:
                    raise ValueError(f"job {spec.job_id}: detector exceeds the server's capacity")
                fx = self.dev.Fixed(self.ctx, img, spec.metric, smooth_ksize=spec.smooth_ksize,
                                    patch_radius=spec.patch_radius)
                fixeds.append(fx)
                views.append(self.dev.make_view(vol, fx, cam, spec.step))
        except X.XrhError as e:
            for f in fixeds:
                f.close()
            if e.status == X.XRH_ERR_BUDGET:
                return None
            raise
        return _Active(spec, fixeds, views, now, float(getattr(vol, "extent_mm", 1.0)),
                       (getattr(vol, "storage", 0), getattr(vol, "layout", 0)))

    def _can_ever_fit(self, spec: JobSpec) -> str:
        """Why a job can never be admitted by this server ("" if it can)."""
        if len(spec.cams) > self.max_views:
            return f"{len(spec.cams)} views exceed the server's {self.max_views}"
        if any(c.num_det_rows > self.max_det or c.num_det_cols > self.max_det for c in spec.cams):
            return f"a detector exceeds the server's {self.max_det}^2 capacity"
        return ""

    def run(self, jobs: list[JobSpec], realtime: bool = True, on_finish=None) -> list[JobResult]:
        """Serves the jobs (sorted by arrival). With realtime=True arrivals follow the wall clock;
        otherwise every job is available at once (throughput mode). on_finish(result) is called
        as each job completes. A job that can never be admitted (more views or a larger detector
        than the server holds, or fixed images larger than the whole HBM budget with nothing
        else in service) is returned at once with status "rejected" instead of blocking the
        queue.

        Each scheduling step fills one submission in policy order: every job contributes the
        next items of its current generation (view-major) while the step's estimated samples
        stay within step_samples and its items within max_items; the job that opens the step
        always gets at least one item. A generation larger than the step budget is therefore
        spread over several steps, so that short jobs arriving meanwhile wait for one step, not
        for a whole large generation. A job's optimiser advances once all of its generation's
        items are in."""
        pending = [(j.arrival, j.job_id, j) for j in jobs]
        heapq.heapify(pending)
        active: list[_Active] = []
        done: list[JobResult] = []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))

        def reject(spec: JobSpec, why: str):
            t = time.perf_counter() - t0
            done.append(JobResult(spec.job_id, spec.arrival if realtime else 0.0, t, t, 0, float("inf"),
                                  None, None, [], None, None, "rejected", why))
            if on_finish is not None:
                on_finish(done[-1])

        while pending or active:
            now = clock()
            # admission: arrived jobs, in arrival order, while HBM and views allow
            while pending and pending[0][0] <= now:
                spec = pending[0][2]
                why = self._can_ever_fit(spec)
                if why:
                    heapq.heappop(pending)
                    reject(spec, why)
                    continue
                if self.max_active is not None and len(active) >= self.max_active:
                    break
                if sum(len(a.views) for a in active) + len(spec.cams) > self.max_views:
                    break
                a = self._admit(spec, time.perf_counter() - t0)
                if a is None:
                    if not active:
                        # nothing in service holds memory that could be released for it
                        heapq.heappop(pending)
                        reject(spec, "fixed images exceed the HBM budget")
                        continue
                    break
                heapq.heappop(pending)
                active.append(a)
            if not active:
                if pending and realtime:
                    time.sleep(max(0.0, min(0.002, pending[0][0] - clock())))
                continue
            if self.policy == "srpt":
                order = sorted(active, key=lambda a: (a.remaining_work(), a.spec.arrival, a.spec.job_id))
            else:
                order = active
            budget = self.step_samples
            packed, n_items = [], 0
            for a in order:
                if packed and a.layout_key != packed[0][0].layout_key:
                    continue  # one submission casts volumes of one storage type and layout
                k = a.chunk(budget, self.max_items - n_items, first=not packed)
                if k == 0:
                    continue
                packed.append((a, a.cursor, k))
                n_items += k
                if budget is not None:
                    budget = max(0.0, budget - sum(a.item_work(a.cursor + i) for i in range(k)))
                if n_items >= self.max_items:
                    break
            views, item_view, item_pose = [], [], []
            for a, c0, k in packed:
                base_view = len(views)
                views.extend(a.views)
                lam = a.spec.lam
                for i in range(c0, c0 + k):              # view-major, as xReg lays out projections
                    item_view.append(base_view + i // lam)
                    item_pose.append(a.poses[i % lam])
            self.batch.set_views(views)
            sims = self.batch.eval(np.array(item_view, np.int32), np.stack(item_pose))
            self.steps += 1
            self.poses_evaluated += len(item_view)
            self.max_items_in_step = max(self.max_items_in_step, len(item_view))
            finished = []
            pos = 0
            for a, c0, k in packed:
                a.sims[c0:c0 + k] = sims[pos:pos + k]
                pos += k
                a.cursor = c0 + k
                if a.cursor < a.n_items:
                    continue  # the rest of this generation goes in later steps
                s = a.sims.reshape(len(a.views), a.spec.lam)
                if a.gen == 0:
                    a.first = (a.poses.copy(), s.copy())
                f = combine_mean(s).astype(np.float64)  # ImgSimMetric2DCombineMean over views
                j = int(np.argmin(f))
                if f[j] < a.best[0]:
                    a.best = (float(f[j]), a.X[j].copy())
                a.es.tell(f)
                a.gen += 1
                in_service = (time.perf_counter() - t0) - a.start
                if a.gen >= a.spec.max_gens:
                    a.status = "max_gens"
                elif a.es.stop():
                    a.status = "converged"
                elif a.spec.max_wall_s is not None and in_service >= a.spec.max_wall_s:
                    a.status = "deadline"
                else:
                    a._new_generation(a.es.ask())
                    continue
                finished.append(a)
            for a in finished:
                t = time.perf_counter() - t0
                bx = a.best[1]
                xm = a.es.m.copy()  # the estimate is the distribution mean, as Intensity2D3DRegiCMAES
                done.append(JobResult(a.spec.job_id, a.spec.arrival if realtime else 0.0, a.start, t, a.gen,
                                      a.best[0], bx, pose_from_params(a.spec.init_pose, a.spec.center, bx),
                                      list(a.es.history), xm,
                                      pose_from_params(a.spec.init_pose, a.spec.center, xm), a.status,
                                      first_poses=a.first[0], first_sims=a.first[1]))
                a.close()
                active.remove(a)
                if on_finish is not None:
                    on_finish(done[-1])
        return done


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if xs else float("nan")
