"""Known-answer tests that pin the oracle's DRR (the restatement of RayCasterLineIntCPU)
independently of its own code: chord lengths from an independent double-precision slab
computation, arithmetic series for fields linear in the index, store/background semantics."""
import numpy as np
import pytest

from tests.helpers import poses_f32


def _geometry(odd=True):
    from xreg_amd import geometry as G

    X, Y, Z, sp = 32, 28, 24, 2.0
    det = 33 if odd else 32
    px = 4.0
    cam = G.camera(1020.0, det, det, px, px)
    center = G.vol_center_phys((X, Y, Z), sp)
    gt = G.gt_pose(center, 510.0)
    return X, Y, Z, sp, cam, gt, center


def _rays_double(cam, pose, sp, dims):
    """Independent double-precision source/detector points in index space for every pixel."""
    rows, cols = cam.num_det_rows, cam.num_det_cols
    cs, rs = cam.det_col_spacing, cam.det_row_spacing
    c, r = np.meshgrid(np.arange(cols), np.arange(rows))
    det = np.stack([(c - (cols - 1) / 2) * cs, (r - (rows - 1) / 2) * rs, -1020.0 * np.ones_like(c, float)], -1)
    M = np.diag([1 / sp, 1 / sp, 1 / sp, 1.0]) @ pose
    src = M[:3, 3]
    detw = det @ M[:3, :3].T + M[:3, 3]
    return src, detw - src


def _chords(src, d, dims):
    mx = np.array(dims, float) - 1
    t0 = np.zeros(d.shape[:-1])
    t1 = np.ones(d.shape[:-1])
    for k in range(3):
        with np.errstate(divide="ignore", invalid="ignore"):
            a = (0 - src[k]) / d[..., k]
            b = (mx[k] - src[k]) / d[..., k]
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        t0, t1 = np.maximum(t0, lo), np.minimum(t1, hi)
    hit = t1 - t0 > 2e-3
    t0n, t1n = t0 + 1e-3, t1 - 1e-3
    return hit, t0n, t1n


def test_constant_volume_chord_counts(oracle):
    """mu constant: DRR = mu * step * (n + 1), n = floor(chord / step_idx) from independent
    double geometry; only rays whose chord/step sits within 1e-4 of an integer may differ."""
    X, Y, Z, sp, cam, gt, center = _geometry()
    mu = 0.02
    vol = np.full((Z, Y, X), mu, np.float32)
    from xreg_amd import synth

    poses = np.concatenate([gt[None], synth.random_poses(gt, center, 3, 15.0, 20.0)])
    step = 1.0
    drr, _ = oracle.drr_f32(vol, np.diag([sp, sp, sp, 1.0]), [cam], poses_f32(poses), step=step)
    checked = 0
    for p in range(poses.shape[0]):
        src, d = _rays_double(cam, poses_f32(poses[p:p + 1]).reshape(4, 4).astype(np.float64), sp, (X, Y, Z))
        hit, t0, t1 = _chords(src, d, (X, Y, Z))
        L = (t1 - t0) * np.linalg.norm(d, axis=-1)
        q = L / (step / sp)
        n = np.floor(q)
        want = np.where(hit, mu * step * (n + 1), 0.0)
        safe = np.abs(q - np.round(q)) > 1e-4
        got = drr[p].astype(np.float64)
        err = np.abs(got - want) / np.maximum(want, 1e-12)
        assert np.all(err[hit & safe] < 1e-5), err[hit & safe].max()
        assert np.all(got[~hit] == 0.0)
        checked += int((hit & safe).sum())
    assert checked > 1000


def test_linear_field_arithmetic_series(oracle):
    """v = a i + b j + c k + e: trilinear is exact, the ray sum is an arithmetic series
    step * (n+1) * (v(start) + n/2 * grad . delta). A transposed axis or a wrong step vector
    breaks it (a, b, c differ)."""
    X, Y, Z, sp, cam, gt, center = _geometry()
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    a, b, c, e = 0.003, 0.001, -0.002, 0.1
    vol = (a * i + b * j + c * k + e).astype(np.float32)
    step = 1.0
    from xreg_amd import synth

    poses = np.concatenate([gt[None], synth.random_poses(gt, center, 2, 10.0, 10.0, seed=9)])
    drr, _ = oracle.drr_f32(vol, np.diag([sp, sp, sp, 1.0]), [cam], poses_f32(poses), step=step)
    for p in range(poses.shape[0]):
        src, d = _rays_double(cam, poses_f32(poses[p:p + 1]).reshape(4, 4).astype(np.float64), sp, (X, Y, Z))
        hit, t0, t1 = _chords(src, d, (X, Y, Z))
        dl = np.linalg.norm(d, axis=-1)
        L = (t1 - t0) * dl
        q = L / (step / sp)
        n = np.floor(q)
        start = src + t0[..., None] * d
        delta = d * ((step / sp) / dl)[..., None]
        g = np.array([a, b, c])
        v0 = start @ g + e
        want = step * (n + 1) * (v0 + 0.5 * n * (delta @ g))
        safe = hit & (np.abs(q - np.round(q)) > 1e-4)
        got = drr[p].astype(np.float64)
        assert np.all(np.abs(got[safe] - want[safe]) <= 2e-5 * np.abs(want[safe]) + 1e-6)


def test_axis_aligned_central_ray(oracle):
    """The principal ray of the ground-truth pose runs along +y through index (X/2, ., Z/2):
    the DRR pixel is step * sum of the y-profile linearly interpolated at the sample points."""
    X, Y, Z, sp, cam, gt, center = _geometry(odd=True)
    prof = np.random.default_rng(3).uniform(0.01, 0.05, Y)
    vol = np.broadcast_to(prof[None, :, None], (Z, Y, X)).astype(np.float32).copy()
    step = 0.7
    drr, _ = oracle.drr_f32(vol, np.diag([sp, sp, sp, 1.0]), [cam], poses_f32(gt[None]), step=step)
    src, d = _rays_double(cam, gt, sp, (X, Y, Z))
    r = c = 16
    hit, t0, t1 = _chords(src, d, (X, Y, Z))
    dl = np.linalg.norm(d[r, c])
    n = int(np.floor((t1[r, c] - t0[r, c]) * dl / (step / sp)))
    y0 = src[1] + t0[r, c] * d[r, c, 1]
    ys = y0 + np.arange(n + 1) * (step / sp)
    want = step * np.interp(ys, np.arange(Y), prof).sum()
    assert abs(float(drr[0, r, c]) - want) < 1e-5 * want


def test_pos_z_frame_is_the_neg_z_image_upside_down(oracle):
    """A POS_Z camera (positive-focal intrinsics, detector at z = +f) with the pose turned half a
    turn about x sees the volume from the same source along mirrored directions: its DRR is the
    NEG_Z camera's with the rows reversed (the rays and their samples coincide)."""
    from xreg_amd import geometry as G

    X, Y, Z, sp, cam, gt, center = _geometry(odd=True)
    vol = np.random.default_rng(5).uniform(0.0, 0.05, (Z, Y, X)).astype(np.float32)
    rows, cols = cam.num_det_rows, cam.num_det_cols
    K = G.naive_intrins(1020.0, rows, cols, cam.det_row_spacing, cam.det_col_spacing)
    K[0, 0] *= -1
    K[1, 1] *= -1
    pz = G.camera_ke(K, np.eye(4), rows, cols, cam.det_row_spacing, cam.det_col_spacing, coord_frame=G.FRAME_POS_Z)
    pose = G.perturbed_pose(gt, center, np.deg2rad([4.0, -3.0, 2.0]), [2.0, -1.0, 3.0])
    i2p = np.diag([sp, sp, sp, 1.0])
    neg, _ = oracle.drr_f32(vol, i2p, [cam], poses_f32(pose[None]), step=0.9)
    pos, _ = oracle.drr_f32(vol, i2p, [pz], poses_f32((pose @ G.rigid(G.rot_x(np.pi)))[None]), step=0.9)
    assert neg.max() > 0
    assert np.abs(pos[0] - neg[0][::-1]).max() <= 1e-5 * neg.max()


def test_store_methods_background_and_max(oracle):
    X, Y, Z, sp, cam, gt, center = _geometry()
    from xreg_amd import synth

    vol = synth.phantom((X, Y, Z))
    rows = cols = cam.num_det_rows
    M = np.diag([sp, sp, sp, 1.0])
    poses = poses_f32(gt[None])
    base, _ = oracle.drr_f32(vol, M, [cam], poses, step=1.0)
    init = np.full((1, rows, cols), 0.5, np.float32)
    acc, _ = oracle.drr_f32(vol, M, [cam], poses, step=1.0, store=1, out=init.copy())
    assert np.allclose(acc, base + 0.5, atol=1e-6)
    rep, _ = oracle.drr_f32(vol, M, [cam], poses, step=1.0, store=0, out=init.copy())
    assert np.array_equal(rep, base)
    bg = np.random.default_rng(0).uniform(0, 1, (1, rows, cols)).astype(np.float32)
    withbg, _ = oracle.drr_f32(vol, M, [cam], poses, step=1.0, bg=bg)
    assert np.allclose(withbg, base + bg, atol=1e-6)
    dflt, _ = oracle.drr_f32(vol, M, [cam], poses, step=1.0, default_bg=2.0)
    assert np.allclose(dflt, base + 2.0, atol=1e-6)
    # MAX of a constant volume is mu * step wherever the ray hits
    mu = np.float32(0.04)
    cvol = np.full((Z, Y, X), mu, np.float32)
    mx, _ = oracle.drr_f32(cvol, M, [cam], poses, step=0.5, kernel=1)
    s, _ = oracle.drr_f32(cvol, M, [cam], poses, step=0.5)
    assert np.all(mx[s > 0] == np.float32(mu * np.float32(0.5)))
    assert np.all(mx[s == 0] == 0.0)  # default background 0 beats lowest()


def test_linearity_and_fp64_agreement(oracle):
    X, Y, Z, sp, cam, gt, center = _geometry()
    from xreg_amd import synth

    vol = synth.phantom((X, Y, Z))
    M = np.diag([sp, sp, sp, 1.0])
    poses = poses_f32(np.concatenate([gt[None], synth.random_poses(gt, center, 2, 15, 20)]))
    d1, s1 = oracle.drr_f32(vol, M, [cam], poses, step=1.0)
    d3, _ = oracle.drr_f32(vol * np.float32(4.0), M, [cam], poses, step=1.0)
    assert np.allclose(d3, 4.0 * d1, rtol=1e-6, atol=1e-7)
    d64, s64 = oracle.drr_f64(vol, M, [cam], poses.astype(np.float64), step=1.0)
    # FP32 sample positions drift ~1e-4 voxel from exact ones, which shows at grazing rays
    # through the phantom's sharp outer boundary; measured against the image maximum the two
    # agree to 1e-5 except where the step count flips (one sample, ~1/n of the pixel)
    err = np.abs(d1 - d64) / d64.max()
    assert np.median(err) < 1e-7
    assert (err > 1e-5).mean() < 5e-3
    assert abs(s1 - s64) / s64 < 1e-3


def test_thread_count_does_not_change_results(oracle):
    X, Y, Z, sp, cam, gt, center = _geometry()
    from xreg_amd import synth

    vol = synth.phantom((X, Y, Z))
    M = np.diag([sp, sp, sp, 1.0])
    poses = poses_f32(gt[None])
    a, _ = oracle.drr_f32(vol, M, [cam], poses, step=1.0, threads=1)
    b, _ = oracle.drr_f32(vol, M, [cam], poses, step=1.0, threads=7)
    assert np.array_equal(a, b)


def test_planted_mistake_is_caught(oracle):
    """The constant-volume KAT fails for an oracle whose step is off by one sample: a check that
    these closed forms bite (the mistake is planted in the expected count on the CPU)."""
    X, Y, Z, sp, cam, gt, center = _geometry()
    vol = np.full((Z, Y, X), 0.02, np.float32)
    drr, _ = oracle.drr_f32(vol, np.diag([sp, sp, sp, 1.0]), [cam], poses_f32(gt[None]), step=1.0)
    src, d = _rays_double(cam, gt, sp, (X, Y, Z))
    hit, t0, t1 = _chords(src, d, (X, Y, Z))
    n = np.floor((t1 - t0) * np.linalg.norm(d, axis=-1) / (1.0 / sp))
    wrong = np.where(hit, 0.02 * (n + 2), 0.0)  # one sample too many
    assert not np.allclose(drr[0][hit], wrong[hit], rtol=1e-5)


@pytest.mark.parametrize("fp64", [False, True])
def test_ray_steps_match_the_drr_of_a_unit_volume(oracle, fp64):
    """oracle.ray_steps (set-up only) against what the DRR itself took: in a volume of ones every
    sample is exactly 1, so a pixel is step * (n + 1) and recovers the ray's n; a miss is 0 and
    -1. The per-ray counts also add up to the DRR's own sample total."""
    from xreg_amd import synth

    X, Y, Z, sp, cam, gt, center = _geometry()
    vol = np.ones((Z, Y, X), np.float32)
    poses = np.concatenate([gt[None], synth.random_poses(gt, center, 4, 25.0, 30.0, seed=17)])
    step = 0.75
    i2p = np.diag([sp, sp, sp, 1.0])
    if fp64:
        pf = poses_f32(poses).astype(np.float64)
        drr, total = oracle.drr_f64(vol, i2p, [cam], pf, step=step)
    else:
        pf = poses_f32(poses)
        drr, total = oracle.drr_f32(vol, i2p, [cam], pf, step=step)
    steps = oracle.ray_steps(vol.shape, i2p, [cam], pf, step=step, fp64=fp64)
    n_from_drr = np.rint(np.asarray(drr, np.float64) / step).astype(np.int64) - 1
    assert (steps == -1).any() and (steps > 10).any()
    assert np.array_equal(steps, n_from_drr)
    assert int((steps + 1).sum()) == total
