"""GPU parity of the similarity metrics (xrh_sim_compute) against the CPU restatements of
ImgSimMetric2D{NCC,GradNCC,PatchNCC,PatchGradNCC}CPU, through the C-ABI."""
import numpy as np
import pytest

from tests.helpers import SIM_ABS_TOL, poses_f32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def images(oracle):
    from xreg_amd import synth

    w = synth.config("c2", n_poses=5, small=True)
    poses = np.concatenate([w.gt[None], w.poses])
    drr, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(poses), step=w.step)
    fixed = synth.noisy_fixed(drr[0])
    return fixed, drr


@pytest.mark.parametrize("metric,ksize", [(0, 0), (1, 0), (1, 3), (1, 5), (1, 7), (2, 0), (3, 0), (3, 5), (3, 3),
                                          (3, 7)])
def test_sim_matches_cpu_reference(ctx, oracle, images, metric, ksize):
    from xreg_amd import _xrh as X

    fixed, mov = images
    print(f"case: metric={metric} smooth={ksize} shape={fixed.shape} n={mov.shape[0]}", flush=True)
    ref = oracle.sim_f32(metric, fixed, mov, smooth_ksize=ksize)
    ref64 = oracle.sim_f64(metric, fixed.astype(np.float64), mov.astype(np.float64), smooth_ksize=ksize)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, metric, smooth_ksize=ksize)
    got = f.compute(buf, mov.shape[0])
    f.close()
    buf.free()
    print(f"max |gpu-cpu| {np.abs(got - ref).max():.3e}; |gpu-fp64| {np.abs(got - ref64).max():.3e}; "
          f"|cpu-fp64| {np.abs(ref - ref64).max():.3e}")
    assert np.abs(got - ref).max() <= SIM_ABS_TOL
    assert np.abs(got - ref64).max() <= SIM_ABS_TOL


@pytest.mark.parametrize("radius,stride", [(2, 1), (3, 1), (4, 2), (5, 2), (6, 1), (7, 3), (8, 1)])
@pytest.mark.parametrize("metric", [2, 3])
def test_patch_radius_and_stride(ctx, oracle, images, metric, radius, stride):
    """Patch diameters with a compiled kernel (7 ... 15) and without one (5, 17), strides 1-3."""
    from xreg_amd import _xrh as X

    fixed, mov = images
    print(f"case: patch metric={metric} radius={radius} stride={stride}", flush=True)
    ref = oracle.sim_f32(metric, fixed, mov, smooth_ksize=5, radius=radius, stride=stride)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, metric, smooth_ksize=5, patch_radius=radius, patch_stride=stride)
    got = f.compute(buf, mov.shape[0])
    f.close()
    buf.free()
    print(f"max |gpu-cpu| {np.abs(got - ref).max():.3e}")
    assert np.abs(got - ref).max() <= SIM_ABS_TOL


def test_ncc_mask(ctx, oracle, images):
    from xreg_amd import _xrh as X

    fixed, mov = images
    mask = np.zeros(fixed.shape, np.uint8)
    mask[5:-7, 3:-4] = 1
    ref = oracle.ncc_f32(fixed, mov, mask=mask)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, 0, mask=mask)
    got = f.compute(buf, mov.shape[0])
    f.close()
    buf.free()
    assert np.abs(got - ref).max() <= SIM_ABS_TOL


def test_sim_does_not_modify_moving_images(ctx, images):
    from xreg_amd import _xrh as X

    fixed, mov = images
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    for metric in range(4):
        f = X.Fixed(ctx, fixed, metric)
        f.compute(buf, mov.shape[0])
        f.close()
    after = buf.download(np.float32).reshape(mov.shape)
    buf.free()
    assert np.array_equal(after, mov)


@pytest.mark.parametrize("metric", [2, 3])
def test_patch_metrics_with_mask(ctx, oracle, images, metric):
    """Masked Patch-NCC / Patch-Grad-NCC (mask weights, skipped patches, correlation over
    unmasked pixels) against the oracle, through the plain and the batched paths."""
    from xreg_amd import _xrh as X

    print(f"case: masked patch metric {metric}", flush=True)
    fixed, mov = images
    mask = np.ones(fixed.shape, np.uint8)
    mask[:20, :30] = 0
    mask[50:60, 10:80] = 0
    mask[70:, 70:] = 0
    ref = oracle.sim_f32(metric, fixed, mov, smooth_ksize=5, radius=5, stride=1, mask=mask)
    ref64 = (oracle.patch_ncc_f64 if metric == 2 else
             (lambda f, m, **k: oracle.patch_grad_ncc_f64(f, m, smooth_ksize=5, **k)))(
        fixed.astype(np.float64), mov.astype(np.float64), radius=5, stride=1, mask=mask)
    unmasked = oracle.sim_f32(metric, fixed, mov, smooth_ksize=5, radius=5, stride=1)
    assert np.abs(ref - unmasked).max() > 1e-4  # the mask matters for these images
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, metric, smooth_ksize=5, patch_radius=5, mask=mask)
    got = f.compute(buf, mov.shape[0])
    buf.free()
    cpu_vs_64 = float(np.abs(ref - ref64).max())
    print(f"max |gpu-cpu| {np.abs(got - ref).max():.3e}; |gpu-fp64| {np.abs(got - ref64).max():.3e}; "
          f"|cpu-fp64| {cpu_vs_64:.3e}")
    assert np.abs(got - ref).max() <= SIM_ABS_TOL
    # the reference divides by a float running sum of ~7e3 normalised weights, whose rounding
    # moves its own result ~1e-5 from the exact value; the device keeps that float divisor
    assert np.abs(got - ref64).max() <= SIM_ABS_TOL + 2.0 * cpu_vs_64
    f.close()


def test_patch_mask_without_unmasked_patch_is_rejected(ctx, images):
    from xreg_amd import _xrh as X

    fixed, _ = images
    with pytest.raises(X.XrhError):
        X.Fixed(ctx, fixed, 2, patch_radius=5, mask=np.zeros(fixed.shape, np.uint8))


@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_sim_special_images(ctx, oracle, images, metric):
    """Moving image = fixed image (perfect correlation), a constant moving image (standard
    deviation clamped at 1e-6), and a negated fixed image, against the oracle."""
    from xreg_amd import _xrh as X

    fixed, _ = images
    print(f"case: special moving images, metric {metric}", flush=True)
    mov = np.stack([fixed, np.full_like(fixed, 3.25), -fixed]).astype(np.float32)
    ref = oracle.sim_f32(metric, fixed, mov, smooth_ksize=5)
    ref64 = oracle.sim_f64(metric, fixed.astype(np.float64), mov.astype(np.float64), smooth_ksize=5)
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    f = X.Fixed(ctx, fixed, metric, smooth_ksize=5)
    got = f.compute(buf, mov.shape[0])
    f.close()
    buf.free()
    print(f"gpu {got}, oracle {ref}, fp64 {ref64}")
    assert np.abs(got - ref64).max() <= SIM_ABS_TOL
    # The reference adds the ~7e3 patch scores in a float running sum
    # (xregImgSimMetric2DPatchNCCCPU.cpp:262-266); with scores near 2 (negated image) that sum
    # alone moves its result ~2e-4 from the exact value. The device sums in FP64, so it is held
    # to the FP32 restatement only up to the restatement's own distance from FP64.
    assert np.all(np.abs(got - ref) <= SIM_ABS_TOL + np.abs(ref - ref64))
    # closed forms: the correlation of identical images is (n - 1) / n (sample deviations over n
    # products); NCC scores (1 -+ that) / 2, the patch scores 1 -+ that per 11x11 patch
    if metric < 2:
        n = fixed.size
        want = [1.0 / (2 * n), 0.5, 1.0 - 1.0 / (2 * n)]
    else:
        want = [1.0 / 121, 1.0, 2.0 - 1.0 / 121]
    assert np.abs(got - np.array(want)).max() <= SIM_ABS_TOL
