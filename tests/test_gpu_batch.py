"""Batched generation evaluation (xrh_batch_*): same values as the per-pose plain path and as
the CPU reference, across graph replays, uneven detector sizes and mixed metrics."""
import numpy as np
import pytest

from tests.helpers import DRR_REL_TOL, SIM_ABS_TOL, drr_rel_err, poses_f32

pytestmark = pytest.mark.gpu


def _setup(ctx, oracle, name, metric, n=6):
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    w = synth.config(name, n_poses=n, small=True)
    gt_drr, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams[:1], poses_f32(w.gt[None]), step=w.step)
    fixed = synth.noisy_fixed(gt_drr[0])
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    fx = X.Fixed(ctx, fixed, metric)
    return w, vol, fx, fixed


@pytest.mark.parametrize("metric", [0, 1, 3])
def test_batch_matches_reference_and_replays(ctx, oracle, metric):
    from xreg_amd import _xrh as X

    w, vol, fx, fixed = _setup(ctx, oracle, "c2", metric)
    b = X.Batch(ctx, [X.make_view(vol, fx, w.cams[0], w.step)], max_items=16)
    poses = w.poses
    iv = np.zeros(poses.shape[0], np.int32)
    drr_ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(poses), step=w.step)
    ref = oracle.sim_f32(metric, fixed, drr_ref)
    first = b.eval(iv, poses)
    assert np.abs(first - ref).max() <= SIM_ABS_TOL
    # graph replays with host work between them give the same values
    for _ in range(3):
        _ = np.linalg.svd(np.random.default_rng(0).normal(size=(64, 64)))
        again = b.eval(iv, poses)
        assert np.array_equal(again, first)
    # a different generation through the same graph
    poses2 = poses[::-1].copy()
    got2 = b.eval(iv, poses2)
    assert np.array_equal(got2, first[::-1])
    # eager, timed submission gives the same values
    b.set_options(use_graph=False, time_kernels=True)
    eager = b.eval(iv, poses)
    assert np.array_equal(eager, first)
    drr_ms, tot_ms = b.last_timing()
    assert 0.0 < drr_ms <= tot_ms
    b.close()
    fx.close()
    vol.close()


def test_batch_equals_plain_path(ctx, oracle):
    from xreg_amd import _xrh as X

    w, vol, fx, fixed = _setup(ctx, oracle, "c2", 3)
    n = w.poses.shape[0]
    rows, cols = w.cams[0].num_det_rows, w.cams[0].num_det_cols
    plain = np.empty(n, np.float32)
    buf = ctx.alloc(4 * rows * cols)
    for i in range(n):
        X.line_int(ctx, vol, w.cams, poses_f32(w.poses[i:i + 1]), np.zeros(1, np.int64), w.step, buf)
        plain[i] = fx.compute(buf, 1)[0]
    b = X.Batch(ctx, [X.make_view(vol, fx, w.cams[0], w.step)], max_items=n)
    drr = ctx.alloc(4 * n * rows * cols)
    batched = b.eval(np.zeros(n, np.int32), w.poses, drr=drr)
    assert np.array_equal(batched, plain)
    # DRRs written by the batch equal the plain DRRs
    X.line_int(ctx, vol, w.cams, poses_f32(w.poses[-1:]), np.zeros(1, np.int64), w.step, buf)
    last = buf.download(np.float32)
    all_drr = drr.download(np.float32).reshape(n, -1)
    assert np.array_equal(all_drr[-1], last)
    for o in (b, fx, vol):
        o.close()
    buf.free()
    drr.free()


def test_batch_uneven_sizes_and_mixed_metrics(ctx, oracle):
    """Views with different detector sizes, volumes and metrics in one batch; each item equals
    its own single-view evaluation."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import synth

    w1 = synth.config("c2", n_poses=4, small=True)
    w2 = synth.config("c1", n_poses=4, small=True)
    cam_big = w1.cams[0]
    cam_small = G.downsample_camera(cam_big, 0.5)
    views, objs, refs = [], [], []
    specs = [(w1, cam_big, 3), (w1, cam_small, 1), (w2, w2.cams[0], 0)]
    for w, cam, metric in specs:
        vol = X.Volume(ctx, w.vol, w.idx_to_phys)
        g, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses_f32(w.gt[None]), step=w.step)
        fixed = synth.noisy_fixed(g[0])
        fx = X.Fixed(ctx, fixed, metric)
        views.append(X.make_view(vol, fx, cam, w.step))
        objs += [vol, fx]
        d, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses_f32(w.poses), step=w.step)
        refs.append(oracle.sim_f32(metric, fixed, d))
    b = X.Batch(ctx, views, max_items=32)
    iv = np.array([0, 1, 2, 0, 1, 2, 2, 1, 0, 0, 1, 2], np.int32)
    pi = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3])
    poses = np.stack([specs[v][0].poses[p] for v, p in zip(iv, pi)])
    got = b.eval(iv, poses)
    want = np.array([refs[v][p] for v, p in zip(iv, pi)], np.float32)
    assert np.abs(got - want).max() <= SIM_ABS_TOL
    b.close()
    for o in objs:
        o.close()


def test_async_launch_and_wait(ctx, oracle):
    """xrh_batch_launch queues a generation and returns; xrh_batch_wait delivers the same values
    as xrh_batch_eval. A second launch before the wait, and a wait with nothing launched, are
    refused without touching the queued work."""
    from xreg_amd import _xrh as X

    print("case: batch launch / wait", flush=True)
    w, vol, fx, fixed = _setup(ctx, oracle, "c2", 3)
    b = X.Batch(ctx, [X.make_view(vol, fx, w.cams[0], w.step)], max_items=16)
    iv = np.zeros(w.poses.shape[0], np.int32)
    want = b.eval(iv, w.poses)
    assert X.lib().xrh_ctx_stream(ctx.h) is not None
    with pytest.raises(X.XrhError):
        b.wait()  # nothing launched
    for use_graph in (True, False):
        b.set_options(use_graph=use_graph, time_kernels=False)
        b.launch(iv, w.poses)
        _ = np.linalg.svd(np.random.default_rng(1).normal(size=(96, 96)))  # host work meanwhile
        with pytest.raises(X.XrhError):
            b.launch(iv, w.poses)
        got = b.wait()
        assert np.array_equal(got, want)
        with pytest.raises(X.XrhError):
            b.wait()
    b.close()
    fx.close()
    vol.close()


def test_graph_cache_evicts_and_rebuilds(ctx, oracle):
    """More distinct submission shapes than the batch keeps graphs for (32): every submission,
    including one whose graph was evicted and is captured again, gives the eager values."""
    from xreg_amd import _xrh as X

    print("case: graph cache eviction (40 submission shapes)", flush=True)
    w, vol, fx, fixed = _setup(ctx, oracle, "c2", 3, n=8)
    b = X.Batch(ctx, [X.make_view(vol, fx, w.cams[0], w.step)], max_items=40)
    poses = np.concatenate([w.poses] * 5)[:40]
    iv = np.zeros(40, np.int32)
    b.set_options(use_graph=False, time_kernels=False)
    eager = b.eval(iv, poses)
    b.set_options(use_graph=True, time_kernels=False)
    for n in list(range(1, 41)) + [1, 2, 40, 3]:
        got = b.eval(iv[:n], poses[:n])
        assert np.array_equal(got, eager[:n]), n
    b.close()
    fx.close()
    vol.close()


@pytest.mark.parametrize("kernel,use_bg", [(0, False), (0, True), (1, False)])
def test_batch_objects_equal_replace_then_accum_casts(ctx, oracle, kernel, use_bg):
    """Three objects per item in one submission (xrh_batch_eval_objects) against the reference's
    sequence -- a REPLACE cast of the first volume (onto the background image when given), ACCUM
    casts of the others, then the metric: the DRRs are equal bit for bit, the similarity values
    equal those of xrh_sim_compute on them, and both match the oracle's REPLACE + ACCUM DRRs."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    print(f"case: batch objects kernel={kernel} bg={use_bg}", flush=True)
    w = synth.config("c3", n_poses=4, small=True)
    vols_np = [np.ascontiguousarray(w.vol, np.float32), np.ascontiguousarray(w.vol[::-1] * 0.5, np.float32),
               np.ascontiguousarray(w.vol[:, ::-1, ::-1] * 0.25, np.float32)]
    vols = [X.Volume(ctx, v, w.idx_to_phys) for v in vols_np]
    cams = w.cams[:2]
    rows, cols = cams[0].num_det_rows, cams[0].num_det_cols
    P = rows * cols
    fixed = [synth.noisy_fixed(np.ones((rows, cols), np.float32), seed=90 + v) for v in range(2)]
    fxs = [X.Fixed(ctx, f, X.XRH_SIM_PATCH_GRAD_NCC) for f in fixed]
    V, O, k = 2, 3, 4
    tbl = [X.make_view(vol, fxs[v], cams[v], w.step, kernel=kernel) for vol in vols for v in range(V)]
    b = X.Batch(ctx, tbl, max_items=V * O * k)
    poses = [w.poses, np.roll(w.poses, 1, axis=0), np.roll(w.poses, 2, axis=0)]
    ev = np.array([[o * V + v for o in range(O)] for v in range(V) for _ in range(k)], np.int32)
    ep = np.stack([np.stack([poses[o][c] for o in range(O)]) for v in range(V) for c in range(k)])
    bgimg = np.random.default_rng(4).uniform(0, 0.3, size=(V, rows, cols)).astype(np.float32)
    bgb = None
    if use_bg:
        bgb = ctx.alloc(bgimg.nbytes)
        bgb.upload(bgimg)
    drr = ctx.alloc(4 * V * k * P)
    sims = b.eval_objects(ev, ep, bg=bgb, drr=drr)
    got = drr.download(np.float32).reshape(V * k, rows, cols)
    # the reference's sequence, view-major projections
    seq = ctx.alloc(4 * V * k * P)
    c4p = np.repeat(np.arange(V, dtype=np.int64), k)
    for o in range(O):
        X.line_int(ctx, vols[o], cams, poses_f32(np.concatenate([poses[o]] * V)), c4p, w.step, seq, kernel=kernel,
                   store=X.XRH_STORE_REPLACE if o == 0 else X.XRH_STORE_ACCUM, bg=bgb if o == 0 else None)
    want = seq.download(np.float32).reshape(V * k, rows, cols)
    assert np.array_equal(got, want)
    want_sims = np.concatenate([fxs[v].compute(seq, k, offset_imgs=v * k) for v in range(V)])
    assert np.array_equal(sims, want_sims)
    # the oracle's REPLACE + ACCUM casts
    for v in range(V):
        d, _ = oracle.drr_f32(vols_np[0], w.idx_to_phys, [cams[v]], poses_f32(poses[0]), step=w.step, kernel=kernel,
                              bg=bgimg[v][None] if use_bg else None)
        for o in range(1, O):
            d, _ = oracle.drr_f32(vols_np[o], w.idx_to_phys, [cams[v]], poses_f32(poses[o]), step=w.step,
                                  kernel=kernel, store=1, out=d)
        assert drr_rel_err(got[v * k:(v + 1) * k], d).max() <= DRR_REL_TOL
    for x in (drr, seq) + ((bgb,) if bgb else ()):
        x.free()
    b.close()
    for o in fxs + vols:
        o.close()
