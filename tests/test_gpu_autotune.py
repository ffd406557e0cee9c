"""DRR launch variants and their autotuner (xrh_batch_set_drr_variant, xrh_batch_autotune):
every variant casts identical DRRs, the tuner picks one of them and later submissions keep
their values."""
import itertools

import numpy as np
import pytest

from tests.helpers import SIM_ABS_TOL, poses_f32

pytestmark = pytest.mark.gpu


def _mixed_batch(ctx, oracle, metrics=(0, 3)):
    """Two views of different detector sizes (metrics: NCC and patch gradient-NCC by default),
    5 poses each."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import synth

    w = synth.config("c2", n_poses=5, small=True)
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    cams = [w.cams[0], G.downsample_camera(w.cams[0], 0.5)]
    fixeds, views, fixed_imgs = [], [], []
    for cam, metric in zip(cams, metrics):
        g, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses_f32(w.gt[None]), step=w.step)
        img = synth.noisy_fixed(g[0])
        fx = X.Fixed(ctx, img, metric)
        fixeds.append(fx)
        fixed_imgs.append(img)
        views.append(X.make_view(vol, fx, cam, w.step))
    n = w.poses.shape[0]
    iv = np.repeat(np.arange(2, dtype=np.int32), n)
    ip = np.concatenate([w.poses, w.poses])
    npx = sum(c.num_det_rows * c.num_det_cols for c in cams) * n
    return w, vol, cams, fixeds, fixed_imgs, views, iv, ip, npx


@pytest.mark.parametrize("metrics", [(0, 3), (1, 3)])
def test_every_variant_casts_the_same_drrs(ctx, oracle, metrics):
    """With and without an NCC view (whose sums the DRR epilogue fuses)."""
    from xreg_amd import _xrh as X

    print(f"case: drr variants metrics={metrics}", flush=True)
    w, vol, cams, fixeds, imgs, views, iv, ip, npx = _mixed_batch(ctx, oracle, metrics)
    b = X.Batch(ctx, views, max_items=16)
    out = ctx.alloc(4 * npx)
    ref_sims = b.eval(iv, ip, drr=out)
    ref_drr = out.download(np.float32)
    n = w.poses.shape[0]
    # the default variant matches the CPU reference
    for v, cam in enumerate(cams):
        d, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses_f32(w.poses), step=w.step)
        ref = oracle.sim_f32(metrics[v], imgs[v], d)
        assert np.abs(ref_sims[v * n:(v + 1) * n] - ref).max() <= SIM_ABS_TOL
    for reuse, shape, bands in itertools.product((0, 1), (0, 1, 2), (1, 4, 8, 16)):
        b.set_drr_variant(reuse, shape, bands)
        sims = b.eval(iv, ip, drr=out)
        assert np.array_equal(out.download(np.float32), ref_drr), (reuse, shape, bands)
        # metrics that read the DRRs back: identical; fused NCC sums change order
        assert np.array_equal(sims[n:], ref_sims[n:]), (reuse, shape, bands)
        if metrics[0] == 0:
            assert np.abs(sims[:n] - ref_sims[:n]).max() <= 1e-6, (reuse, shape, bands)
        else:
            assert np.array_equal(sims[:n], ref_sims[:n]), (reuse, shape, bands)
    b.set_drr_variant(None)
    assert np.array_equal(b.eval(iv, ip), ref_sims)
    for bad in ((2, 0, 8), (0, 3, 8), (0, 0, 0), (0, 0, 17)):
        with pytest.raises(X.XrhError):
            b.set_drr_variant(*bad)
    out.free()
    b.close()
    for f in fixeds:
        f.close()
    vol.close()


def test_autotune_keeps_values(ctx, oracle):
    from xreg_amd import _xrh as X

    print("case: autotune", flush=True)
    w, vol, cams, fixeds, imgs, views, iv, ip, npx = _mixed_batch(ctx, oracle)
    b = X.Batch(ctx, views, max_items=16)
    before = b.eval(iv, ip)
    r = b.autotune(iv, ip, reps=1)
    assert r["variants_timed"] == 20
    assert r["cell_reuse"] in (0, 1) and r["wave_shape"] in (0, 1, 2) and r["bands"] in (4, 8, 16)
    assert 0.0 < r["chosen_ms"] <= r["default_ms"]
    n = w.poses.shape[0]
    # graph replays after tuning
    for _ in range(3):
        after = b.eval(iv, ip)
        assert np.array_equal(after[n:], before[n:])
        assert np.abs(after[:n] - before[:n]).max() <= 1e-6
    # a plain cast of the same shape (it may pick the tuned variant from the context's table)
    buf = ctx.alloc(4 * cams[0].num_det_rows * cams[0].num_det_cols * n)
    X.line_int(ctx, vol, [cams[0]], poses_f32(w.poses), np.zeros(n, np.int64), w.step, buf)
    d, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cams[0]], poses_f32(w.poses), step=w.step)
    got = buf.download(np.float32).reshape(d.shape)
    from tests.helpers import DRR_REL_TOL, drr_rel_err

    assert drr_rel_err(got, d).max() <= DRR_REL_TOL
    buf.free()
    b.close()
    for f in fixeds:
        f.close()
    vol.close()
