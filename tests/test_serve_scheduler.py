"""Scheduler logic of the registration server on the CPU, with a host-side fake device whose
objective is a closed-form function of the pose (no GPU): admission under the HBM budget,
packing whole generations under the item capacity, completion of every job, and identical
results alone and in the mix."""
import types

import numpy as np
import pytest

from xreg_amd import geometry as G
from xreg_amd import serve


class FakeCtx:
    def __init__(self):
        self.used = 0

    def bytes_in_use(self):
        return self.used


class FakeFixed:
    def __init__(self, ctx, img, metric, smooth_ksize=5, patch_radius=5):
        self.ctx, self.nbytes = ctx, img.size * 4 * 6
        self.target = img  # the fake objective's target pose lives here
        ctx.used += self.nbytes

    def close(self):
        self.ctx.used -= self.nbytes


class FakeBatch:
    steps = []

    def __init__(self, ctx, views, max_items, geom=0, capacity=None):
        self.max_items = max_items
        self.views = None

    def set_options(self, **kw):
        pass

    def set_views(self, views):
        self.views = list(views)

    def eval(self, item_view, item_pose):
        assert len(item_view) <= self.max_items
        FakeBatch.steps.append(len(item_view))
        out = np.empty(len(item_view), np.float32)
        for i, (v, p) in enumerate(zip(item_view, item_pose)):
            t = self.views[v].fixed.target
            out[i] = np.float32(np.sum((p[:3, 3] - t[:3, 3]) ** 2) + 100.0 * np.sum((p[:3, :3] - t[:3, :3]) ** 2))
        return out


FAKE = types.SimpleNamespace(Batch=FakeBatch, Fixed=FakeFixed,
                             make_view=lambda vol, fx, cam, step: types.SimpleNamespace(fixed=fx, cam=cam))


def _specs(n, lam=10, views=1, gens=15):
    cam = G.camera(1020.0, 48, 48, 3.0, 3.0)
    center = np.array([64.0, 64.0, 50.0])
    gt = G.gt_pose(center, 510.0)
    rng = np.random.default_rng(0)
    out = []
    for j in range(n):
        init = G.perturbed_pose(gt, center, np.deg2rad(rng.uniform(-4, 4, 3)), rng.uniform(-5, 5, 3))
        out.append(serve.JobSpec(job_id=j, arrival=0.0, volume="v", cams=[cam] * views,
                                 fixed=[np.asarray(gt, np.float32)] * views, init_pose=init, center=center,
                                 step=1.0, lam=lam + j % 3, max_gens=gens, seed=j))
    return out


def test_all_jobs_complete_and_converge():
    ctx = FakeCtx()
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=64, max_views=8, max_det=64,
                                   device=FAKE)
    specs = _specs(6, gens=60)
    res = srv.run(specs, realtime=False)
    assert sorted(r.job_id for r in res) == list(range(6))
    for r in res:
        assert r.best_sim < r.history[0]
    assert ctx.used == 0  # every job released its fixed images
    assert srv.max_items_in_step <= 64


def test_budget_limits_concurrency():
    ctx = FakeCtx()
    specs = _specs(5, lam=10, gens=3)
    per_job = specs[0].fixed[0].size * 4 * 6
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=2 * per_job, max_items=1000, max_views=8,
                                   max_det=64, device=FAKE)
    assert srv._fixed_bytes(specs[0]) == per_job
    FakeBatch.steps = []
    res = srv.run(specs, realtime=False)
    assert len(res) == 5
    # never more than two jobs (<= 2 * 12 items) in one submission
    assert max(FakeBatch.steps) <= 24


def test_alone_equals_mixed_scheduler():
    ctx = FakeCtx()
    specs = _specs(4, views=2, gens=8)
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=40, max_views=8, max_det=64,
                                   device=FAKE)
    alone = {s.job_id: srv.run([s], realtime=False)[0] for s in specs}
    mixed = {r.job_id: r for r in srv.run(specs, realtime=False)}
    for j in alone:
        assert alone[j].history == mixed[j].history


def test_serial_mode_serves_one_job_at_a_time_with_the_same_results():
    """max_active=1 (the serial baseline): every submission holds one job's generation, jobs
    finish in arrival order, and each job's run equals its continuously batched run."""
    ctx = FakeCtx()
    specs = _specs(4, views=2, gens=6)
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=200, max_views=8, max_det=64,
                                   device=FAKE, max_active=1)
    FakeBatch.steps = []
    serial = srv.run(specs, realtime=False)
    assert [r.job_id for r in serial] == [0, 1, 2, 3]
    assert max(FakeBatch.steps) <= 2 * max(s.lam for s in specs)
    assert len(FakeBatch.steps) == sum(r.generations for r in serial)
    for a, b in zip(serial[:-1], serial[1:]):
        assert a.finish <= b.start
    srv2 = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=200, max_views=8, max_det=64,
                                    device=FAKE)
    mixed = {r.job_id: r for r in srv2.run(specs, realtime=False)}
    for r in serial:
        assert r.history == mixed[r.job_id].history


def test_srpt_packing_runs_short_jobs_first_within_the_step_budget():
    """policy="srpt": the job with the least remaining work goes first; a submission holds
    further jobs only while the estimated samples stay within step_samples; a long job then runs
    alone; every job's run equals its run alone."""
    ctx = FakeCtx()
    specs = _specs(4, lam=10, gens=6)
    specs[0].max_gens, specs[0].lam = 30, 40        # the long job arrives first
    gen = [s.lam * 48 * 48 for s in specs]            # samples per generation (extent/step = 1)
    # room for job 0 alone, or for the three short jobs together
    budget = float(max(gen[0], sum(gen[1:])))
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=1000, max_views=8, max_det=64,
                                   device=FAKE, policy="srpt", step_samples=budget)
    FakeBatch.steps = []
    res = {r.job_id: r for r in srv.run(specs, realtime=False)}
    short_items = sum(s.lam for s in specs[1:])
    # the short jobs go first; the long job fills what is left of each step's budget (7 of its
    # 40 items), so its generations spread over several steps until the short jobs are done
    assert FakeBatch.steps[0] == short_items + (specs[0].lam - short_items)
    assert FakeBatch.steps[-1] == specs[0].lam
    assert max(res[j].finish for j in (1, 2, 3)) <= res[0].finish
    assert srv.max_items_in_step == specs[0].lam
    for s in specs:
        alone = srv.run([s], realtime=False)[0]
        assert alone.history == res[s.job_id].history
    with pytest.raises(ValueError):
        serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1, device=FAKE, policy="lifo")


def test_fifo_policy_without_limit_packs_every_admitted_job():
    ctx = FakeCtx()
    specs = _specs(3, lam=10, gens=4)
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=1000, max_views=8, max_det=64,
                                   device=FAKE, policy="fifo", step_samples=None)
    FakeBatch.steps = []
    srv.run(specs, realtime=False)
    assert FakeBatch.steps[0] == sum(s.lam for s in specs)


def test_fifo_policy_with_limit_keeps_arrival_order():
    """policy="fifo" with a step limit: the long job that arrived first runs alone until it
    finishes; the short jobs then share submissions."""
    ctx = FakeCtx()
    specs = _specs(4, lam=10, gens=6)
    specs[0].max_gens, specs[0].lam = 30, 40
    gen = [s.lam * 48 * 48 for s in specs]
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=1000, max_views=8, max_det=64,
                                   device=FAKE, policy="fifo", step_samples=float(max(gen[0], sum(gen[1:]))))
    FakeBatch.steps = []
    res = {r.job_id: r for r in srv.run(specs, realtime=False)}
    n0 = res[0].generations
    assert FakeBatch.steps[:n0] == [specs[0].lam] * n0
    assert FakeBatch.steps[n0] == sum(s.lam for s in specs[1:])
    assert res[0].finish <= min(res[j].finish for j in (1, 2, 3))


def test_jobs_on_volumes_of_different_layouts_never_share_a_submission():
    """A batch casts volumes of one storage type and layout: jobs on an FP16 volume and on an
    FP32 one alternate between submissions instead of failing, with the same results."""
    ctx = FakeCtx()
    vols = {"a": types.SimpleNamespace(storage=0, layout=4), "b": types.SimpleNamespace(storage=1, layout=2)}
    specs = _specs(4, lam=10, gens=5)
    for j, sp in enumerate(specs):
        sp.volume = "ab"[j % 2]
    seen = []

    class LayoutBatch(FakeBatch):
        def eval(self, item_view, item_pose):
            keys = {(self.views[v].vol.storage, self.views[v].vol.layout) for v in set(item_view)}
            seen.append(keys)
            return super().eval(item_view, item_pose)

    dev = types.SimpleNamespace(Batch=LayoutBatch, Fixed=FakeFixed,
                                make_view=lambda vol, fx, cam, step: types.SimpleNamespace(fixed=fx, cam=cam, vol=vol))
    srv = serve.RegistrationServer(ctx, vols, hbm_budget=1 << 40, max_items=1000, max_views=8, max_det=64,
                                   device=dev, policy="fifo", step_samples=None)
    res = {r.job_id: r for r in srv.run(specs, realtime=False)}
    assert len(res) == 4 and all(len(k) == 1 for k in seen)
    assert {next(iter(k)) for k in seen} == {(0, 4), (1, 2)}
    for sp in specs:
        alone = srv.run([sp], realtime=False)[0]
        assert alone.history == res[sp.job_id].history


def test_deadline_finishes_jobs_with_their_current_estimate():
    """A job past its in-service limit ends after the generation in flight, with the CMA-ES mean
    as its estimate and status "deadline"; jobs without a limit run their generations."""
    ctx = FakeCtx()
    specs = _specs(3, gens=40)
    specs[0].max_wall_s = 0.0
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=64, max_views=8, max_det=64,
                                   device=FAKE)
    res = {r.job_id: r for r in srv.run(specs, realtime=False)}
    assert res[0].status == "deadline" and res[0].generations == 1
    assert res[0].pose is not None and np.all(np.isfinite(res[0].pose))
    for j in (1, 2):
        assert res[j].status in ("max_gens", "converged")
        assert res[j].generations > 1
    assert ctx.used == 0


def test_generation_larger_than_a_submission_is_split_with_the_same_results():
    """lambda = 20 with 2 views through submissions of at most 8 items: each generation spans
    five steps, and the run equals the one with room for whole generations."""
    ctx = FakeCtx()
    spec = _specs(1, lam=20, views=2, gens=5)[0]
    small = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=8, max_views=8, max_det=64,
                                     device=FAKE)
    FakeBatch.steps = []
    a = small.run([spec], realtime=False)[0]
    assert max(FakeBatch.steps) == 8 and len(FakeBatch.steps) == 5 * a.generations
    big = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=64, max_views=8, max_det=64,
                                   device=FAKE)
    b = big.run([spec], realtime=False)[0]
    assert a.history == b.history and a.generations == b.generations


def test_step_budget_splits_a_large_generation_and_short_jobs_wait_one_step():
    """A large job is in service when a short one arrives: with a step budget smaller than the
    large generation, the short job's generations go into the next submission instead of
    waiting for the large generation to end."""
    ctx = FakeCtx()
    specs = _specs(2, lam=10, gens=4)
    specs[0].lam, specs[0].max_gens = 60, 4          # 60 items of 48^2 rays per generation
    per_item = 48 * 48
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=1000, max_views=8, max_det=64,
                                   device=FAKE, policy="srpt", step_samples=float(15 * per_item))
    FakeBatch.steps = []
    res = {r.job_id: r for r in srv.run(specs, realtime=False)}
    assert max(FakeBatch.steps) == 15
    # the short job (lambda 11) finishes within its first generations' steps, well before the
    # large job's 4 x 60 items are through
    assert res[1].finish < res[0].finish
    for sp in specs:
        assert srv.run([sp], realtime=False)[0].history == res[sp.job_id].history


def test_job_with_too_many_views_is_rejected_without_blocking():
    ctx = FakeCtx()
    specs = _specs(3, views=1, gens=3)
    specs[1].cams = specs[1].cams * 9
    specs[1].fixed = specs[1].fixed * 9
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=64, max_views=8, max_det=64,
                                   device=FAKE)
    res = {r.job_id: r for r in srv.run(specs, realtime=False)}
    assert res[1].status == "rejected" and "views" in res[1].reason
    assert res[0].generations == 3 and res[2].generations == 3


def test_job_whose_fixed_images_exceed_the_budget_is_rejected_without_blocking():
    ctx = FakeCtx()
    specs = _specs(3, views=1, gens=3)
    specs[1].fixed = [np.zeros((2048, 2048), np.float32)]   # 96 MiB of fixed-image buffers
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=16 << 20, max_items=64, max_views=8, max_det=64,
                                   device=FAKE)
    res = {r.job_id: r for r in srv.run(specs, realtime=True)}
    assert res[1].status == "rejected" and "budget" in res[1].reason
    assert res[0].generations == 3 and res[2].generations == 3
    assert ctx.used == 0


def test_job_with_too_large_a_detector_is_rejected():
    ctx = FakeCtx()
    specs = _specs(2, gens=2)
    specs[0].cams = [G.camera(1020.0, 96, 96, 1.5, 1.5)]
    srv = serve.RegistrationServer(ctx, {"v": None}, hbm_budget=1 << 40, max_items=64, max_views=8, max_det=64,
                                   device=FAKE)
    res = {r.job_id: r for r in srv.run(specs, realtime=False)}
    assert res[0].status == "rejected" and res[1].generations == 2


def test_exp_se3_matches_matrix_exponential():
    """exp_se3 and its batched form against scipy's matrix exponential of the twist."""
    from scipy.linalg import expm

    rng = np.random.default_rng(5)
    X = np.concatenate([rng.normal(0, 0.3, (16, 3)), rng.normal(0, 10, (16, 3))], axis=1)
    X[0, :3] = 0.0  # pure translation takes the small-angle branch
    B = G.exp_se3_batch(X)
    for x, b in zip(X, B):
        tw = np.zeros((4, 4))
        tw[:3, :3] = [[0, -x[2], x[1]], [x[2], 0, -x[0]], [-x[1], x[0], 0]]
        tw[:3, 3] = x[3:]
        np.testing.assert_allclose(G.exp_se3(x), expm(tw), atol=1e-10)
        np.testing.assert_allclose(b, G.exp_se3(x), atol=1e-12)
    init = G.gt_pose(np.array([10.0, 20.0, 30.0]), 510.0)
    P = serve.poses_from_params(init, np.array([10.0, 20.0, 30.0]), X)
    for x, p in zip(X, P):
        np.testing.assert_allclose(p, serve.pose_from_params(init, np.array([10.0, 20.0, 30.0]), x), atol=1e-9)
