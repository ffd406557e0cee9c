"""The non-trilinear interpolators of the CPU caster (RayCaster::set_interp_method,
lib/ray_cast/xregRayCastInterface.h:165-191; xregRayCastLineIntCPU.cpp:115-147) on the device,
through xrh_line_int_interp, against the oracle on the same volume, rays and poses:
  * nearest neighbour: bit for bit (the device casts the reference's float rays, so every
    sample reads the same voxel and the float sums run in the same order);
  * cubic B-spline and Lanczos-4 windowed sinc: per pixel within 1e-4 relative (both sides
    evaluate a sample in double and round it to float; the float line sums match).
Also: MAX projections, ACCUM onto background images, 16-bit storage (the grid holds the stored
16-bit values), every FP32 layout giving the same DRRs, the grids' bytes, and anti-aliasing
(jittered rays averaged per pixel) against the oracle with the same jitter seed."""
import numpy as np
import pytest

from tests.helpers import DRR_REL_TOL, drr_rel_err, poses_f32

pytestmark = pytest.mark.gpu

NAMES = {1: "nn", 2: "sinc", 3: "bspline"}


@pytest.fixture(scope="module")
def c1i():
    from xreg_amd import geometry as G
    from xreg_amd import synth

    w = synth.config("c1", n_poses=3)
    return w, G.downsample_camera(w.cams[0], 0.5)


def _cast(ctx, vol, cam, poses, step, interp, kernel=0, store=0, bg=None, default_bg=0.0, prefill=None, aa_rays=0,
          aa_seed=0):
    from xreg_amd import _xrh as X

    n = poses.shape[0]
    npx = cam.num_det_rows * cam.num_det_cols
    buf = ctx.alloc(4 * n * npx)
    if prefill is not None:
        buf.upload(np.ascontiguousarray(prefill, np.float32))
    X.line_int_interp(ctx, vol, [cam], poses, np.zeros(n, np.int64), step, buf, interp, kernel=kernel,
                      store=store, bg=bg, default_bg=default_bg, aa_rays=aa_rays, aa_seed=aa_seed)
    out = buf.download(np.float32).reshape(n, cam.num_det_rows, cam.num_det_cols)
    buf.free()
    return out


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("interp", [1, 3, 2])
def test_interp_drr_matches_oracle(ctx, oracle, c1i, interp, kernel):
    from xreg_amd import _xrh as X

    w, cam = c1i
    print(f"case: interp {NAMES[interp]} kernel {kernel}, c1 volume, {cam.num_det_rows}^2, 3 poses", flush=True)
    poses = poses_f32(w.poses)
    fill = 0.25 if kernel == 0 else 0.0      # REPLACE pre-fill; MAX projections stay below 0.25
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    got = _cast(ctx, vol, cam, poses, w.step, interp, kernel=kernel, default_bg=fill)
    vol.close()
    ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses, step=w.step, kernel=kernel, default_bg=fill,
                            interp=interp)
    err = drr_rel_err(got, ref)
    print(f"max rel err {err.max():.3e}; pixels differing {int((got != ref).sum())} of {got.size}; "
          f"image max {ref.max():.4g}", flush=True)
    assert ref.max() > (1.0 if kernel == 0 else 0.01)
    if interp == 1:
        assert np.array_equal(got, ref)
    else:
        assert err.max() <= DRR_REL_TOL


@pytest.mark.parametrize("interp", [1, 3])
def test_interp_accumulates_onto_background(ctx, oracle, c1i, interp):
    from xreg_amd import _xrh as X

    w, cam = c1i
    print(f"case: interp {NAMES[interp]} ACCUM onto background images", flush=True)
    poses = poses_f32(w.poses[:2])
    rng = np.random.default_rng(2)
    bg = rng.uniform(0.0, 3.0, size=(1, cam.num_det_rows, cam.num_det_cols)).astype(np.float32)
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    bgd = ctx.alloc(bg.nbytes)
    bgd.upload(bg)
    got = _cast(ctx, vol, cam, poses, w.step, interp, store=1, bg=bgd)
    # ACCUM without background images adds onto what the buffer holds
    prev = rng.uniform(0.0, 1.0, size=(2, cam.num_det_rows, cam.num_det_cols)).astype(np.float32)
    got2 = _cast(ctx, vol, cam, poses, w.step, interp, store=1, prefill=prev)
    bgd.free()
    vol.close()
    ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses, step=w.step, store=1, bg=bg, interp=interp)
    ref2 = prev.copy()
    oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses, step=w.step, store=1, out=ref2, interp=interp)
    if interp == 1:
        assert np.array_equal(got, ref) and np.array_equal(got2, ref2)
    else:
        assert drr_rel_err(got, ref).max() <= DRR_REL_TOL
        assert drr_rel_err(got2, ref2).max() <= DRR_REL_TOL


@pytest.mark.parametrize("storage", [1, 2])
def test_interp_reads_the_stored_16bit_voxels(ctx, oracle, c1i, storage):
    from xreg_amd import _xrh as X

    w, cam = c1i
    print(f"case: interp nn on {('f16', 'bf16')[storage - 1]} storage", flush=True)
    import torch

    t = torch.from_numpy(np.ascontiguousarray(w.vol, np.float32))
    rounded = t.to(torch.float16 if storage == 1 else torch.bfloat16).to(torch.float32).numpy()
    poses = poses_f32(w.poses[:2])
    vol = X.Volume(ctx, w.vol, w.idx_to_phys, storage=storage)
    got = _cast(ctx, vol, cam, poses, w.step, 1)
    vol.close()
    ref, _ = oracle.drr_f32(rounded, w.idx_to_phys, [cam], poses, step=w.step, interp=1)
    assert np.array_equal(got, ref)


def test_interp_grids_are_the_same_for_every_fp32_layout(ctx, c1i):
    from xreg_amd import _xrh as X

    w, cam = c1i
    print("case: interp nn + bspline over the linear, quad, coef, face and brick layouts", flush=True)
    poses = poses_f32(w.poses[:2])
    outs = {}
    for layout in (X.LAYOUT_LINEAR, X.LAYOUT_QUAD, X.LAYOUT_COEF, X.LAYOUT_FACE, X.LAYOUT_BRICK):
        vol = X.Volume(ctx, w.vol, w.idx_to_phys, layout=layout)
        outs[layout] = (_cast(ctx, vol, cam, poses, w.step, 1), _cast(ctx, vol, cam, poses, w.step, 3))
        vol.close()
    first = outs[X.LAYOUT_LINEAR]
    for layout, (nn, bs) in outs.items():
        assert np.array_equal(nn, first[0]), layout
        assert np.array_equal(bs, first[1]), layout


def test_interp_grid_bytes_and_linear_passthrough(ctx, c1i):
    from xreg_amd import _xrh as X

    w, cam = c1i
    print("case: interp grid bytes, linear = xrh_line_int with FP32 geometry", flush=True)
    poses = poses_f32(w.poses[:2])
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    b0 = vol.nbytes
    lin = _cast(ctx, vol, cam, poses, w.step, 0)
    assert vol.nbytes == b0
    buf = ctx.alloc(lin.nbytes)
    X.line_int(ctx, vol, [cam], poses, np.zeros(2, np.int64), w.step, buf)
    assert np.array_equal(lin, buf.download(np.float32).reshape(lin.shape))
    buf.free()
    Z, Y, X_ = w.vol.shape
    _cast(ctx, vol, cam, poses, w.step, 1)
    b1 = vol.nbytes
    assert b1 - b0 == 4 * (X_ + 8) * (Y + 8) * (Z + 8)
    _cast(ctx, vol, cam, poses, w.step, 2)      # sinc shares the nearest-neighbour grid
    assert vol.nbytes == b1
    _cast(ctx, vol, cam, poses, w.step, 3)
    assert vol.nbytes - b1 == 8 * (X_ + 4) * (Y + 4) * (Z + 4)
    vol.close()


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("interp", [0, 1, 3])
def test_antialiased_drr_matches_oracle(ctx, oracle, c1i, interp, kernel):
    """RayCasterCPU::set_anti_alias_factor (xregRayCastBaseCPU.h:113-130): 4 jittered rays per
    pixel, averaged; the same seed gives the oracle the same jitter. Linear (ITK's double lerp on
    both sides) and nearest neighbour are bit for bit; B-spline within 1e-4."""
    from xreg_amd import _xrh as X

    w, cam = c1i
    print(f"case: anti-aliasing 4 rays, interp {interp}, kernel {kernel}, 2 poses", flush=True)
    poses = poses_f32(w.poses[:2])
    seed = 0xC0FFEE
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    got = _cast(ctx, vol, cam, poses, w.step, interp, kernel=kernel, aa_rays=4, aa_seed=seed)
    again = _cast(ctx, vol, cam, poses, w.step, interp, kernel=kernel, aa_rays=4, aa_seed=seed)
    other = _cast(ctx, vol, cam, poses, w.step, interp, kernel=kernel, aa_rays=4, aa_seed=seed + 1)
    vol.close()
    ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses, step=w.step, kernel=kernel, interp=interp,
                            aa_rays=4, aa_seed=seed)
    err = drr_rel_err(got, ref)
    print(f"max rel err {err.max():.3e}; pixels differing {int((got != ref).sum())} of {got.size}", flush=True)
    assert np.array_equal(got, again)
    assert not np.array_equal(got, other)
    if interp in (0, 1):
        assert np.array_equal(got, ref)
    else:
        assert err.max() <= DRR_REL_TOL


def test_interp_rejects_bad_arguments_before_any_work(ctx, c1i):
    """An unknown interpolator, an anti-aliasing factor outside [0, XRH_MAX_AA_RAYS] and a volume of
    another context fail with XRH_ERR_INVALID and a message; nothing is cast and no grid is built."""
    from xreg_amd import _xrh as X

    w, cam = c1i
    print("case: interp argument checks", flush=True)
    poses = poses_f32(w.poses[:1])
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    b0 = vol.nbytes
    npx = cam.num_det_rows * cam.num_det_cols
    buf = ctx.alloc(4 * npx)
    buf.upload(np.full(npx, 7.0, np.float32))
    for kw, what in (({"interp": 4}, "interpolation"), ({"interp": -1}, "interpolation"),
                     ({"interp": 1, "aa_rays": -1}, "anti-aliasing"),
                     ({"interp": 1, "aa_rays": X.XRH_MAX_AA_RAYS + 1}, "anti-aliasing")):
        with pytest.raises(X.XrhError, match=what):
            X.line_int_interp(ctx, vol, [cam], poses, np.zeros(1, np.int64), w.step, buf, **kw)
    other = X.Context(0)
    try:
        with pytest.raises(X.XrhError, match="another context"):
            X.line_int_interp(other, vol, [cam], poses, np.zeros(1, np.int64), w.step, buf, 1)
    finally:
        other.close()
    assert vol.nbytes == b0
    assert np.all(buf.download(np.float32) == 7.0)
    buf.free()
    vol.close()


@pytest.mark.parametrize("shape", [(5, 3, 2), (9, 7, 4)])
def test_interp_tiny_volumes_and_ragged_detectors(ctx, oracle, shape):
    """Volumes with sides of 2-9 voxels (the B-spline support passes both mirror boundaries of a
    side at once; sides under 4 take the clamp) seen by ragged detectors (3 x 17, 1 x 1):
    nearest neighbour bit for bit, B-spline and sinc within 1e-4, with and without
    anti-aliasing."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import synth

    print(f"case: interp tiny volume {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    vol = rng.uniform(0.5, 1.5, size=shape).astype(np.float32)
    sp = 2.0
    i2p = G.idx_to_phys(sp)
    center = G.vol_center_phys(shape[::-1], sp)
    gt = G.gt_pose(center, 510.0)
    poses = poses_f32(np.concatenate([gt[None], synth.random_poses(gt, center, 2, 5.0, 2.0, seed=3)]))
    v = X.Volume(ctx, vol, i2p)
    for cam in (G.camera(1020.0, 3, 17, 2.0, 1.0), G.camera(1020.0, 1, 1, 1.0, 1.0)):
        for interp in (1, 3, 2):
            for aa in (0, 2):
                got = _cast(ctx, v, cam, poses, 0.5, interp, aa_rays=aa, aa_seed=11)
                ref, _ = oracle.drr_f32(vol, i2p, [cam], poses, step=0.5, interp=interp, aa_rays=aa, aa_seed=11)
                if interp == 1:
                    assert np.array_equal(got, ref), (cam.num_det_rows, interp, aa)
                else:
                    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL, (cam.num_det_rows, interp, aa)
    v.close()
