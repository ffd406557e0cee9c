"""GPU parity of the general patch combine (xrh_fixed_set_patch_weights: patch weights, the
patches used, weight-in-combine and mean flags) against the oracle, through the C-ABI, on the
specialised stride-1 kernel and the general kernel, with and without a mask, plain and batched."""
import numpy as np
import pytest

from tests.helpers import SIM_ABS_TOL, poses_f32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def images(oracle):
    from xreg_amd import synth

    w = synth.config("c2", n_poses=4, small=True)
    poses = np.concatenate([w.gt[None], w.poses])
    drr, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(poses), step=w.step)
    fixed = synth.noisy_fixed(drr[0])
    return fixed, drr


def _mask(shape):
    m = np.ones(shape, np.uint8)
    m[:15, :25] = 0
    m[40:52, 10:70] = 0
    return m


def _tol(ref):
    # a plain sum of many patch scores grows with their count; the float reference rounds in
    # proportion to it
    return SIM_ABS_TOL * np.maximum(1.0, np.abs(ref))


CASES = [
    # (radius, stride, masked): radius 5 stride 1 unmasked runs the specialised kernel
    (5, 1, False),
    (4, 2, False),
    (5, 1, True),
    (3, 1, True),
]


@pytest.mark.parametrize("radius,stride,masked", CASES)
@pytest.mark.parametrize("metric", [2, 3])
def test_weighted_combines_match_oracle(ctx, oracle, images, metric, radius, stride, masked):
    from xreg_amd import _xrh as X

    fixed, mov = images
    mask = _mask(fixed.shape) if masked else None
    print(f"case: patch weights metric={metric} r={radius} s={stride} masked={masked}", flush=True)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, metric, smooth_ksize=5, patch_radius=radius, patch_stride=stride, mask=mask)
    default = f.compute(buf, mov.shape[0])
    n_p = f.patches()["weights"].size
    rng = np.random.default_rng(radius * 10 + stride)
    w = rng.uniform(0.05, 2.0, n_p).astype(np.float32)
    w[rng.integers(0, n_p, 20)] = 0.0
    inds = rng.integers(0, n_p, 150)
    inds[:10] = inds[10]  # repeats
    settings = [
        dict(weights=w),
        dict(weights=w, patch_inds=inds),
        dict(weights=w, patch_inds=inds, compute_mean=True),
        dict(weights=w, patch_inds=inds, weight_in_combine=False),
        dict(weights=w, weight_in_combine=False, compute_mean=True),
        dict(patch_inds=inds),  # the creation weights (unit, or the mask's)
    ]
    for st in settings:
        f.set_patch_weights(**st)
        got = f.compute(buf, mov.shape[0])
        kw = dict(weights=st.get("weights"), inds=st.get("patch_inds"),
                  weight_in_combine=st.get("weight_in_combine", True), compute_mean=st.get("compute_mean", False),
                  smooth_ksize=5, radius=radius, stride=stride, mask=mask)
        if kw["weights"] is None and masked:
            kw["weights"] = oracle.patch_weights_f32(fixed.shape, radius, stride, mask=mask)
        ref = oracle.patch_sim_weighted_f32(metric, fixed, mov, **kw)
        ref64 = oracle.patch_sim_weighted_f64(metric, fixed.astype(np.float64), mov.astype(np.float64), **kw)
        names = sorted(k for k in st if k != "weights")
        print(f"  {names}: max |gpu-cpu| {np.abs(got - ref).max():.3e} |gpu-fp64| {np.abs(got - ref64).max():.3e}"
              f" |cpu-fp64| {np.abs(ref - ref64).max():.3e}", flush=True)
        assert np.all(np.abs(got - ref) <= _tol(ref) + np.abs(ref - ref64))
        assert np.all(np.abs(got - ref64) <= _tol(ref64))
    # the default restored: unmasked bit-identical, masked within rounding of the folded tables
    f.set_patch_weights()
    again = f.compute(buf, mov.shape[0])
    if masked:
        assert np.abs(again - default).max() <= 1e-6
    else:
        assert np.array_equal(again, default)
    f.close()
    buf.free()


def test_skipped_and_single_patches(ctx, oracle, images):
    from xreg_amd import _xrh as X

    fixed, mov = images
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, 3, smooth_ksize=5, patch_radius=5)
    n_p = f.patches()["weights"].size
    w = np.ones(n_p, np.float32)
    q = n_p // 3 + 7
    w[q] = 5e-7
    # a patch of weight <= 1e-6 used alone: skipped, so the score is exactly 0
    f.set_patch_weights(weights=w, patch_inds=[q])
    assert np.array_equal(f.compute(buf, mov.shape[0]), np.zeros(mov.shape[0], np.float32))
    # one patch scored on its own equals the oracle's score of that patch
    for wic in (True, False):
        f.set_patch_weights(weights=w, patch_inds=[q + 1], weight_in_combine=wic)
        got = f.compute(buf, mov.shape[0])
        ref = oracle.patch_sim_weighted_f32(3, fixed, mov, inds=[q + 1], smooth_ksize=5, radius=5)
        assert np.abs(got - ref).max() <= SIM_ABS_TOL
    # the reported weights are the ones last set
    assert np.array_equal(f.patches()["weights"], w)
    f.close()
    buf.free()


def test_rejected_settings_change_nothing(ctx, images):
    from xreg_amd import _xrh as X

    fixed, mov = images
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, 2, patch_radius=5)
    n_p = f.patches()["weights"].size
    w = np.linspace(0.5, 1.0, n_p).astype(np.float32)
    f.set_patch_weights(weights=w, patch_inds=np.arange(0, n_p, 7))
    before = f.compute(buf, mov.shape[0])
    bad = [
        dict(patch_inds=[0, n_p]),
        dict(patch_inds=[-1]),
        dict(patch_inds=np.zeros(0, np.int64)),
        dict(weights=np.zeros(n_p, np.float32)),             # weights of the used patches sum to 0
        dict(weights=np.full(n_p, np.nan, np.float32)),
    ]
    for st in bad:
        with pytest.raises(X.XrhError):
            f.set_patch_weights(**st)
        assert np.array_equal(f.compute(buf, mov.shape[0]), before)
    with pytest.raises(ValueError):
        f.set_patch_weights(weights=np.ones(n_p - 1, np.float32))
    # zero weights are fine when they do not divide: plain sum
    f.set_patch_weights(weights=np.zeros(n_p, np.float32), weight_in_combine=False, compute_mean=True)
    f.close()
    g = X.Fixed(ctx, fixed, 0)
    with pytest.raises(X.XrhError):
        g.set_patch_weights(patch_inds=[0])
    g.close()
    buf.free()


def test_batch_and_graph_replays_follow_the_weights(ctx, oracle):
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    w = synth.config("c2", n_poses=5, small=True)
    gt_drr, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams[:1], poses_f32(w.gt[None]), step=w.step)
    fixed = synth.noisy_fixed(gt_drr[0])
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    fx = X.Fixed(ctx, fixed, 3)
    b = X.Batch(ctx, [X.make_view(vol, fx, w.cams[0], w.step)], max_items=8)
    iv = np.zeros(w.poses.shape[0], np.int32)
    drr_ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(w.poses), step=w.step)
    first = b.eval(iv, w.poses)  # builds and replays the graph
    n_p = fx.patches()["weights"].size
    rng = np.random.default_rng(9)
    wt = rng.uniform(0.1, 1.0, n_p).astype(np.float32)
    inds = rng.integers(0, n_p, 100)
    for k in range(3):
        fx.set_patch_weights(weights=wt, patch_inds=inds[k:], compute_mean=(k == 2))
        got = b.eval(iv, w.poses)
        ref = oracle.patch_sim_weighted_f32(3, fixed, drr_ref, weights=wt, inds=inds[k:], compute_mean=(k == 2))
        print(f"case: batch weights {k}: max |gpu-cpu| {np.abs(got - ref).max():.3e}", flush=True)
        assert np.abs(got - ref).max() <= SIM_ABS_TOL
    fx.set_patch_weights()
    assert np.array_equal(b.eval(iv, w.poses), first)
    for o in (b, fx, vol):
        o.close()


def test_per_direction_weights_and_fixed_gradients(ctx, oracle, images):
    """Patch-Grad-NCC with separate x and y weights (each direction's own fixed gradient
    variances, as use_variances_in_grad_imgs_as_wgts) against the oracle composed per
    direction; the fixed gradient images the device holds equal the oracle's Sobel."""
    from xreg_amd import _xrh as X

    fixed, mov = images
    print("case: per-direction patch weights", flush=True)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    for radius, stride in ((5, 1), (4, 2)):
        f = X.Fixed(ctx, fixed, 3, smooth_ksize=5, patch_radius=radius, patch_stride=stride)
        gx, gy = f.image(0), f.image(1)
        rx, ry = oracle.sobel_f32(fixed, 5)
        assert np.abs(gx - rx).max() <= 1e-5 * max(1.0, np.abs(rx).max())
        assert np.abs(gy - ry).max() <= 1e-5 * max(1.0, np.abs(ry).max())
        wx = oracle.patch_weights_f32(fixed.shape, radius, stride, wgt_img=oracle.patch_variance_image_f32(
            gx, radius, stride))
        wy = oracle.patch_weights_f32(fixed.shape, radius, stride, wgt_img=oracle.patch_variance_image_f32(
            gy, radius, stride))
        assert np.abs(wx - wy).max() > 1e-6
        f.set_patch_weights(weights=wx, weights_y=wy)
        got = f.compute(buf, mov.shape[0])
        ref = np.empty(mov.shape[0], np.float32)
        for m in range(mov.shape[0]):
            mx, my = oracle.sobel_f32(mov[m], 5)
            sx = oracle.patch_sim_weighted_f32(2, rx, mx[None], weights=wx, radius=radius, stride=stride)[0]
            sy = oracle.patch_sim_weighted_f32(2, ry, my[None], weights=wy, radius=radius, stride=stride)[0]
            ref[m] = np.float32(0.5 * float(np.float32(sx + sy)))
        print(f"  r={radius} s={stride}: max |gpu-cpu| {np.abs(got - ref).max():.3e}", flush=True)
        assert np.abs(got - ref).max() <= 2 * SIM_ABS_TOL
        # the same weights in both directions equal the single-weight call bit for bit
        f.set_patch_weights(weights=wx, weights_y=wx)
        a = f.compute(buf, mov.shape[0])
        f.set_patch_weights(weights=wx)
        assert np.array_equal(a, f.compute(buf, mov.shape[0]))
        f.close()
    g = X.Fixed(ctx, fixed, 2, patch_radius=5)
    n_p = g.patches()["weights"].size
    with pytest.raises(X.XrhError):
        g.set_patch_weights(weights=np.ones(n_p, np.float32), weights_y=np.ones(n_p, np.float32))
    with pytest.raises(X.XrhError):
        g.image(1)
    assert np.array_equal(g.image(0), fixed.astype(np.float32))
    g.close()
    buf.free()


@pytest.mark.parametrize("radius,stride", [(5, 1), (3, 2)])
@pytest.mark.parametrize("metric", [2, 3])
def test_mask_stats_match_oracle(ctx, oracle, images, metric, radius, stride):
    """use_mask_for_patch_stats (xrh_sim_params.patch_mask_stats): patch statistics over the
    unmasked pixels, with the mask's weights and with a weighted patch list, against the oracle;
    patches with no unmasked pixel included."""
    from xreg_amd import _xrh as X

    fixed, mov = images
    mask = _mask(fixed.shape)
    mask[60:75, 60:75] = 0     # whole patches without an unmasked pixel
    print(f"case: mask stats metric={metric} r={radius} s={stride}", flush=True)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, metric, smooth_ksize=5, patch_radius=radius, patch_stride=stride, mask=mask,
                patch_mask_stats=True)
    wm = oracle.patch_weights_f32(fixed.shape, radius, stride, mask=mask)
    n_p = wm.size
    inds = np.random.default_rng(3).integers(0, n_p, 200)
    for st in (dict(), dict(patch_inds=inds, weight_in_combine=False, compute_mean=True)):
        f.set_patch_weights(**st)
        got = f.compute(buf, mov.shape[0])
        kw = dict(weights=wm, inds=st.get("patch_inds"), weight_in_combine=st.get("weight_in_combine", True),
                  compute_mean=st.get("compute_mean", False), smooth_ksize=5, radius=radius, stride=stride,
                  mask=mask, mask_stats=True)
        ref = oracle.patch_sim_weighted_f32(metric, fixed, mov, **kw)
        ref64 = oracle.patch_sim_weighted_f64(metric, fixed.astype(np.float64), mov.astype(np.float64), **kw)
        print(f"  {sorted(st)}: max |gpu-cpu| {np.abs(got - ref).max():.3e} |gpu-fp64| {np.abs(got - ref64).max():.3e}",
              flush=True)
        assert np.all(np.abs(got - ref64) <= _tol(ref64) + np.abs(ref - ref64))
        assert np.all(np.abs(got - ref) <= _tol(ref) + np.abs(ref - ref64))
    plain = oracle.patch_sim_weighted_f32(metric, fixed, mov, weights=wm, smooth_ksize=5, radius=radius,
                                          stride=stride, mask=mask)
    f.set_patch_weights()
    assert np.abs(f.compute(buf, mov.shape[0]) - plain).max() > 1e-4  # the option changes the metric
    f.close()
    # without a mask the option changes nothing
    g0 = X.Fixed(ctx, fixed, metric, smooth_ksize=5, patch_radius=radius, patch_stride=stride)
    g1 = X.Fixed(ctx, fixed, metric, smooth_ksize=5, patch_radius=radius, patch_stride=stride, patch_mask_stats=True)
    assert np.array_equal(g0.compute(buf, mov.shape[0]), g1.compute(buf, mov.shape[0]))
    g0.close()
    g1.close()
    buf.free()


@pytest.mark.parametrize("radius,stride,masked", [(5, 1, False), (3, 2, True)])
@pytest.mark.parametrize("metric", [2, 3])
def test_per_patch_scores_match_oracle(ctx, oracle, images, metric, radius, stride, masked):
    """xrh_sim_patch_scores (save_all_per_patch_scores): every grid patch's unweighted score
    against the oracle, and the similarity values it returns against the plain call."""
    from xreg_amd import _xrh as X

    fixed, mov = images
    mask = _mask(fixed.shape) if masked else None
    print(f"case: per-patch scores metric={metric} r={radius} s={stride} masked={masked}", flush=True)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, metric, smooth_ksize=5, patch_radius=radius, patch_stride=stride, mask=mask,
                patch_mask_stats=masked)
    plain = f.compute(buf, mov.shape[0])
    sims, scores = f.patch_scores(buf, mov.shape[0])
    ref = oracle.patch_scores_f32(metric, fixed, mov, smooth_ksize=5, radius=radius, stride=stride, mask=mask,
                                  mask_stats=masked)
    print(f"  max |gpu-cpu| per patch {np.abs(scores - ref).max():.3e}; sims vs plain {np.abs(sims - plain).max():.3e}",
          flush=True)
    assert scores.shape == ref.shape
    assert np.abs(scores - ref).max() <= 2e-5
    assert np.abs(sims - plain).max() <= 1e-6
    # weights do not change the per-patch scores
    f.set_patch_weights(weights=np.linspace(0.1, 1.0, ref.shape[1]).astype(np.float32))
    _, again = f.patch_scores(buf, mov.shape[0])
    assert np.array_equal(again, scores)
    f.close()
    buf.free()


@pytest.mark.parametrize("shape,radius,stride", [((11, 11), 5, 1), ((11, 40), 5, 1), ((9, 9), 4, 3), ((23, 12), 3, 2)])
@pytest.mark.parametrize("metric", [2, 3])
def test_small_images_and_single_patches(ctx, oracle, metric, shape, radius, stride):
    """Images holding one patch, one row of patches, or a few strided patches: the specialised
    and the general kernel, unit and explicit weights, masked statistics, per-patch scores."""
    from xreg_amd import _xrh as X

    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    fixed = rng.random(shape).astype(np.float32)
    mov = (0.5 * fixed[None] + 0.5 * rng.random((3, *shape))).astype(np.float32)
    mask = (rng.random(shape) > 0.25).astype(np.uint8)
    print(f"case: small image {shape} metric={metric} r={radius} s={stride}", flush=True)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    for mk, ms in ((None, False), (mask, False), (mask, True)):
        f = X.Fixed(ctx, fixed, metric, smooth_ksize=3, patch_radius=radius, patch_stride=stride, mask=mk,
                    patch_mask_stats=ms)
        n_p = f.patches()["weights"].size
        w = oracle.patch_weights_f32(shape, radius, stride, mask=mk)
        for st in (dict(), dict(weights=np.linspace(0.5, 1.5, n_p).astype(np.float32), compute_mean=True)):
            f.set_patch_weights(**st)
            got = f.compute(buf, 3)
            ref = oracle.patch_sim_weighted_f32(metric, fixed, mov, weights=st.get("weights", w),
                                                compute_mean=st.get("compute_mean", False), smooth_ksize=3,
                                                radius=radius, stride=stride, mask=mk, mask_stats=ms)
            ref64 = oracle.patch_sim_weighted_f64(metric, fixed.astype(np.float64), mov.astype(np.float64),
                                                  weights=st.get("weights", w),
                                                  compute_mean=st.get("compute_mean", False), smooth_ksize=3,
                                                  radius=radius, stride=stride, mask=mk, mask_stats=ms)
            assert np.all(np.abs(got - ref) <= _tol(ref) + np.abs(ref - ref64)), (mk is not None, ms, sorted(st))
        sims, scores = f.patch_scores(buf, 3)
        sref = oracle.patch_scores_f32(metric, fixed, mov, smooth_ksize=3, radius=radius, stride=stride, mask=mk,
                                       mask_stats=ms)
        assert np.abs(scores - sref).max() <= 2e-5
        f.close()
    buf.free()
