"""Full-size workloads (384^2 / 768^2 detectors, 256x256x200 and 512x512x400 volumes): parity
with the oracle on a few poses, and size-independent properties on whole generations
(exact scaling, batch = plain, half-precision storage bounds)."""
import numpy as np
import pytest

from tests.helpers import DRR_REL_TOL, SIM_ABS_TOL, drr_rel_err, poses_f32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2full():
    from xreg_amd import synth

    return synth.config("c2", n_poses=100)


def _drr(ctx, w, vol, poses, geom=0):
    from xreg_amd import _xrh as X

    cam = w.cams[0]
    n = len(poses)
    buf = ctx.alloc(4 * n * cam.num_det_rows * cam.num_det_cols)
    X.line_int(ctx, vol, [cam], poses_f32(poses), np.zeros(n, np.int64), w.step, buf, geom=geom)
    out = buf.download(np.float32).reshape(n, cam.num_det_rows, cam.num_det_cols)
    buf.free()
    return out


def test_c2_full_drr_matches_oracle(ctx, oracle, c2full):
    from xreg_amd import _xrh as X

    print("case: c2 full size DRR vs oracle, 3 poses", flush=True)
    w = c2full
    poses = np.concatenate([w.gt[None], w.poses[:2]])
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    got = _drr(ctx, w, vol, poses)
    vol.close()
    ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(poses), step=w.step)
    err = drr_rel_err(got, ref)
    print(f"max rel err {err.max():.3e}")
    assert err.max() <= DRR_REL_TOL


def test_c2_full_scaling_is_exact(ctx, c2full):
    """Doubling every voxel doubles every sample and every sum exactly (powers of two commute
    with rounding), so DRR(2V) == 2 DRR(V) bit for bit over a whole generation."""
    from xreg_amd import _xrh as X

    print("case: c2 full size DRR(2V) == 2 DRR(V), 100 poses", flush=True)
    w = c2full
    v1 = X.Volume(ctx, w.vol, w.idx_to_phys)
    a = _drr(ctx, w, v1, w.poses)
    v1.close()
    v2 = X.Volume(ctx, 2.0 * w.vol, w.idx_to_phys)
    b = _drr(ctx, w, v2, w.poses)
    v2.close()
    assert a.max() > 0
    assert np.array_equal(b, 2.0 * a)


def test_c2_full_batch_equals_plain(ctx, oracle, c2full):
    """A lambda = 100 generation through the batched submission equals DRR + similarity one pose
    at a time, and the similarity of the first poses matches the oracle."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    print("case: c2 full size batch vs plain, 100 poses", flush=True)
    w = c2full
    cam = w.cams[0]
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    g = _drr(ctx, w, vol, w.gt[None])
    fixed = synth.noisy_fixed(g[0])
    fx = X.Fixed(ctx, fixed, X.XRH_SIM_PATCH_GRAD_NCC, smooth_ksize=5, patch_radius=5)
    batch = X.Batch(ctx, [X.make_view(vol, fx, cam, w.step)], max_items=100)
    sims = batch.eval(np.zeros(100, np.int32), w.poses)
    buf = ctx.alloc(4 * cam.num_det_rows * cam.num_det_cols)
    plain = np.empty(100, np.float32)
    for i in range(100):
        X.line_int(ctx, vol, [cam], poses_f32(w.poses[i:i + 1]), np.zeros(1, np.int64), w.step, buf)
        plain[i] = fx.compute(buf, 1)[0]
    buf.free()
    assert np.array_equal(sims, plain)
    drr, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(w.poses[:2]), step=w.step)
    want = oracle.patch_grad_ncc_f32(fixed, drr, smooth_ksize=5)
    print(f"sim vs oracle {np.abs(sims[:2] - want).max():.3e}")
    assert np.abs(sims[:2] - want).max() <= SIM_ABS_TOL
    batch.close()
    fx.close()
    vol.close()


@pytest.fixture(scope="module")
def c4full():
    from xreg_amd import synth

    return synth.config("c4", n_poses=4)


def test_c4_full_drr_matches_oracle(ctx, oracle, c4full):
    from xreg_amd import _xrh as X

    print("case: c4 full size DRR vs oracle, 1 pose", flush=True)
    w = c4full
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    got = _drr(ctx, w, vol, w.poses[:1])
    vol.close()
    ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(w.poses[:1]), step=w.step)
    err = drr_rel_err(got, ref)
    print(f"max rel err {err.max():.3e}")
    assert err.max() <= DRR_REL_TOL


@pytest.mark.parametrize("storage", [1, 2])
def test_c4_full_half_storage_bound(ctx, c4full, storage):
    """FP16 / BF16 storage against FP32 storage on the same device geometry: the error of a line
    integral is bounded by the voxel rounding (2^-11 / 2^-8) times the image maximum."""
    from xreg_amd import _xrh as X

    print(f"case: c4 full size storage={storage}, 4 poses", flush=True)
    w = c4full
    v32 = X.Volume(ctx, w.vol, w.idx_to_phys)
    a = _drr(ctx, w, v32, w.poses, geom=1)
    v32.close()
    vh = X.Volume(ctx, w.vol, w.idx_to_phys, storage=storage)
    b = _drr(ctx, w, vh, w.poses, geom=1)
    vh.close()
    err = np.abs(b.astype(np.float64) - a) / float(a.max())
    bound = 2.0 ** -11 if storage == 1 else 2.0 ** -8
    print(f"max |err| / max {err.max():.3e} (bound {bound:.2e})")
    assert err.max() < bound


def test_c4_768_precision_against_fp64_oracle(ctx, oracle):
    """C4 at its own 768^2 detector and 0.25 mm steps: 6 poses of the C4 distribution cast from
    FP32, FP16 and BF16 storage (FP64 ray set-up, FP32 accumulation) against the FP64 oracle
    (FP64 volume, geometry and accumulation), and the NCC of each against the FP64 NCC of a
    fixed image made from the ground-truth pose."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    print("case: c4 768^2 precision vs FP64 oracle, 3 storages x 6 poses", flush=True)
    w = synth.config("c4", n_poses=6)
    cam = w.cams[0]
    poses = np.asarray(w.poses, np.float64)
    ref, _ = oracle.drr_f64(w.vol, w.idx_to_phys, [cam], poses, step=w.step)
    gt, _ = oracle.drr_f64(w.vol, w.idx_to_phys, [cam], w.gt[None], step=w.step)
    fixed = synth.noisy_fixed(gt[0].astype(np.float32))
    ncc_ref = oracle.ncc_f64(fixed, ref)
    scale = float(ref.max())
    bounds = {0: (3e-5, 1e-5), 1: (2.0 ** -11, 1e-4), 2: (2.0 ** -8, 1e-3)}
    fx = X.Fixed(ctx, fixed, X.XRH_SIM_NCC)
    n = len(poses)
    for storage in (0, 1, 2):
        v = X.Volume(ctx, w.vol, w.idx_to_phys, storage=storage)
        buf = ctx.alloc(4 * n * cam.num_det_rows * cam.num_det_cols)
        X.line_int(ctx, v, [cam], poses_f32(poses), np.zeros(n, np.int64), w.step, buf, geom=1)
        got = buf.download(np.float32).reshape(ref.shape)
        ncc = fx.compute(buf, n)
        buf.free()
        v.close()
        derr = np.abs(got.astype(np.float64) - ref) / scale
        nerr = np.abs(ncc.astype(np.float64) - ncc_ref)
        print(f"storage {('fp32', 'fp16', 'bf16')[storage]}: DRR |err|/max max {derr.max():.3e} "
              f"median {np.median(derr):.3e}; NCC |err| max {nerr.max():.3e}", flush=True)
        assert derr.max() < bounds[storage][0]
        assert nerr.max() < bounds[storage][1]
    fx.close()


def test_c4_200_poses_half_storage_bound(ctx):
    """All 200 poses of C4 at 768^2: FP16 and BF16 storage against FP32 storage on the same
    device geometry, within the voxel rounding (2^-11 / 2^-8) of the image maximum, and the FP32
    DRRs finite and non-negative (the phantom is non-negative)."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    print("case: c4 768^2, 200 poses, fp16 / bf16 vs fp32 storage", flush=True)
    w = synth.config("c4")
    assert len(w.poses) == 200
    vols = [X.Volume(ctx, w.vol, w.idx_to_phys, storage=s) for s in (0, 1, 2)]
    worst = {1: 0.0, 2: 0.0}
    for c in range(0, 200, 50):
        chunk = w.poses[c:c + 50]
        a = _drr(ctx, w, vols[0], chunk, geom=1)
        assert np.all(np.isfinite(a)) and a.min() >= 0.0
        scale = float(a.max())
        for s in (1, 2):
            b = _drr(ctx, w, vols[s], chunk, geom=1)
            worst[s] = max(worst[s], float((np.abs(b.astype(np.float64) - a) / scale).max()))
    for v in vols:
        v.close()
    print(f"max |err| / max: fp16 {worst[1]:.3e}, bf16 {worst[2]:.3e}", flush=True)
    assert worst[1] < 2.0 ** -11
    assert worst[2] < 2.0 ** -8


def test_c2_full_cell_reuse_is_bit_identical(ctx, c2full, monkeypatch):
    from xreg_amd import _xrh as X

    print("case: c2 full size cell reuse vs per-sample loads, 100 poses", flush=True)
    w = c2full
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    monkeypatch.setenv("XRH_CELL_REUSE", "1")
    a = _drr(ctx, w, vol, w.poses, geom=1)
    monkeypatch.setenv("XRH_CELL_REUSE", "0")
    b = _drr(ctx, w, vol, w.poses, geom=1)
    vol.close()
    assert np.array_equal(a, b)


def test_c3_full_multiview_pyramid_registration(ctx, oracle):
    """C3 as specified, through the same code bench.py reports (`c3_leg`): three views (30
    degrees apart) of the 256x256x200 volume at 384^2, gradient-NCC averaged over the views,
    registered over a three-level pyramid (detector 96^2, 192^2, 384^2): CMA-ES (lambda 50, up
    to 50 generations) at the coarse level, then the BOBYQA-pattern refiner (13 single-pose
    evaluations, then one per iteration, at most 20 x 13) at 192^2 and 384^2, FP64 ray
    geometry. The estimate ends near the ground truth and every level improves on its start."""
    import bench

    print("case: c3 full size 3 views x 3 levels registration (bench c3 leg)", flush=True)
    out = bench.c3_leg(ctx, 1, 5)
    lv = out["levels"]
    print(f"mPD {out['mpd_start_mm']} -> " + " -> ".join(str(x["mpd_mm"]) for x in lv) +
          f"; evals {[x['evals'] for x in lv]}; wall {out['total_wall_s']} s")
    assert [x["method"] for x in lv] == ["cmaes", "refine", "refine"]
    assert [x["detector"] for x in lv] == [96, 192, 384]
    assert lv[0]["evals"] % 50 == 0 and 13 <= lv[1]["evals"] <= 260 and 13 <= lv[2]["evals"] <= 260
    assert lv[-1]["mpd_mm"] < 0.1 * out["mpd_start_mm"]
    for x in lv:
        assert x["fun_end"] <= x["fun_start"]


def test_volume_past_4gib_packed_layout_falls_back_to_linear(ctx, oracle):
    """A 700x700x560 FP32 volume: its 16-byte-per-voxel packed layout would pass the 4 GiB
    reach of 32-bit buffer offsets, so the automatic choice keeps the padded linear layout
    (1.1 GB), and the DRRs still match the oracle."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G

    print("case: 700x700x560 volume, linear fallback", flush=True)
    X_, Y_, Z_ = 700, 700, 560
    sp = 0.5
    vol = np.zeros((Z_, Y_, X_), np.float32)
    vol[200:360, 250:450, 250:450] = 0.02
    vol[250:300, 300:350, 320:420] = 0.05
    vol[280:282, 100:600, 349:351] = 0.08
    i2p = G.idx_to_phys(sp)
    center = G.vol_center_phys((X_, Y_, Z_), sp)
    gt = G.gt_pose(center, 510.0)
    cam = G.camera(1020.0, 64, 64, 4.656, 4.656)
    poses = np.stack([gt, G.perturbed_pose(gt, center, np.deg2rad([4.0, -3.0, 5.0]), [5.0, -4.0, 3.0])])
    v = X.Volume(ctx, vol, i2p)
    assert v.nbytes < 2 << 30  # 4 bytes per padded voxel, not 16
    buf = ctx.alloc(4 * 2 * 64 * 64)
    X.line_int(ctx, v, [cam], poses_f32(poses), np.zeros(2, np.int64), 0.25, buf)
    got = buf.download(np.float32).reshape(2, 64, 64)
    buf.free()
    v.close()
    ref, _ = oracle.drr_f32(vol, i2p, [cam], poses_f32(poses), step=0.25)
    assert ref.max() > 0.5
    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL


def test_c4_precision_against_fp64_oracle(ctx, oracle, c4full):
    """The 512x512x400 @0.5 mm volume stored as FP32, FP16 and BF16 (same seeded content; FP32
    accumulation, FP64 ray set-up), against the FP64 oracle (FP64 volume and accumulation):
    per-pixel DRR error relative to the image maximum and NCC error. The detector is the C4
    one read out at 192^2 (same field of view) so the FP64 oracle finishes in seconds."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import synth

    print("case: c4 precision vs FP64 oracle, 3 storages x 2 poses", flush=True)
    w = c4full
    cam = G.downsample_camera(w.cams[0], 0.25)
    poses = w.poses[:2]
    ref, _ = oracle.drr_f64(w.vol, w.idx_to_phys, [cam], poses, step=w.step)
    gt, _ = oracle.drr_f64(w.vol, w.idx_to_phys, [cam], w.gt[None], step=w.step)
    fixed = synth.noisy_fixed(gt[0].astype(np.float32))
    ncc_ref = oracle.ncc_f64(fixed, ref)
    scale = float(ref.max())
    # FP32: float accumulation over ~1,600 samples per ray; FP16 / BF16: voxel rounding
    # (2^-11 / 2^-8 relative) on top
    bounds = {0: (3e-5, 1e-5), 1: (2.0 ** -11, 1e-4), 2: (2.0 ** -8, 1e-3)}
    fx = X.Fixed(ctx, fixed, X.XRH_SIM_NCC)
    for storage in (0, 1, 2):
        v = X.Volume(ctx, w.vol, w.idx_to_phys, storage=storage)
        buf = ctx.alloc(4 * 2 * cam.num_det_rows * cam.num_det_cols)
        X.line_int(ctx, v, [cam], poses_f32(poses), np.zeros(2, np.int64), w.step, buf, geom=1)
        got = buf.download(np.float32).reshape(ref.shape)
        ncc = fx.compute(buf, 2)
        buf.free()
        v.close()
        derr = np.abs(got.astype(np.float64) - ref) / scale
        nerr = np.abs(ncc.astype(np.float64) - ncc_ref)
        print(f"storage {('fp32', 'fp16', 'bf16')[storage]}: DRR |err|/max max {derr.max():.3e} "
              f"median {np.median(derr):.3e}; NCC |err| max {nerr.max():.3e}", flush=True)
        assert derr.max() < bounds[storage][0]
        assert nerr.max() < bounds[storage][1]
    fx.close()


def test_c1_full_drr_and_ncc_match_oracle(ctx, oracle):
    """The plain single-view case at full size: 128x128x96 @2 mm, 192^2 detector, 1 mm steps,
    100 poses (U(+-15 deg, +-20 mm)), fixed image = ground-truth DRR + 1% noise: the 100 DRRs
    and 100 NCC values against the FP32 restatement of the CPU reference, through the plain
    path (one launch for the DRRs, one for the similarities)."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    print("case: c1 full size, 100 poses DRR + NCC", flush=True)
    w = synth.config("c1", n_poses=100)
    cam = w.cams[0]
    n = w.poses.shape[0]
    g, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(w.gt[None]), step=w.step)
    fixed = synth.noisy_fixed(g[0])
    ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(w.poses), step=w.step)
    ref_ncc = oracle.ncc_f32(fixed, ref)
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    buf = ctx.alloc(4 * n * cam.num_det_rows * cam.num_det_cols)
    X.line_int(ctx, vol, [cam], poses_f32(w.poses), np.zeros(n, np.int64), w.step, buf)
    got = buf.download(np.float32).reshape(ref.shape)
    fx = X.Fixed(ctx, fixed, X.XRH_SIM_NCC)
    ncc = fx.compute(buf, n)
    fx.close()
    buf.free()
    vol.close()
    err = drr_rel_err(got, ref)
    print(f"max DRR rel err {err.max():.3e}; max NCC err {np.abs(ncc - ref_ncc).max():.3e}")
    assert err.max() <= DRR_REL_TOL
    assert np.abs(ncc - ref_ncc).max() <= SIM_ABS_TOL
    assert np.abs(ncc - oracle.ncc_f64(fixed, ref)).max() <= SIM_ABS_TOL
    # against the FP64 oracle (FP64 volume, geometry and accumulation): the FP32 algorithm's own
    # distance, led by rare step-count flips on grazing rays, stated here
    r64, _ = oracle.drr_f64(w.vol, w.idx_to_phys, w.cams, w.poses, step=w.step)
    e64 = np.abs(got.astype(np.float64) - r64) / float(r64.max())
    n64 = np.abs(ncc.astype(np.float64) - oracle.ncc_f64(fixed, r64))
    print(f"vs FP64 oracle: DRR |err|/max max {e64.max():.3e} median {np.median(e64):.3e} "
          f"99.9% {np.quantile(e64, 0.999):.3e}; NCC |err| max {n64.max():.3e}")
    assert np.median(e64) < 1e-5
    assert n64.max() < 1e-4
