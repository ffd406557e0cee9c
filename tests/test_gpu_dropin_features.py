"""The rest of the drop-in classes' interface, driven as xReg's callers drive it
(tests/cpp/dropin_features.cpp): MAX projections, two objects cast REPLACE then ACCUM, a static
volume's DRR as background, an external host buffer, host-bound moving images, the exhaustive
driver's short last batch, a masked metric and the patch metrics' options (weight image, mask
weighting, mean and unweighted combines, listed and random patches), each against the oracle."""
import subprocess

import numpy as np
import pytest

from tests.helpers import DRR_REL_TOL, SIM_ABS_TOL, drr_rel_err, poses_f32
from tests.test_gpu_dropin import DRIVER_DIR

pytestmark = pytest.mark.gpu

FEATURES = DRIVER_DIR / "dropin_features"


def _records(raw: bytes) -> dict:
    out, off = {}, 0
    while off < len(raw):
        n = int(np.frombuffer(raw, np.uint32, 1, off)[0])
        name = raw[off + 4:off + 4 + n].decode()
        nb = int(np.frombuffer(raw, np.uint64, 1, off + 4 + n)[0])
        out[name] = raw[off + 12 + n:off + 12 + n + nb]
        off += 12 + n + nb
    return out


@pytest.mark.parametrize("metric,smooth", [(0, 0), (3, 5), (2, 0)])
def test_dropin_features(tmp_path, oracle, metric, smooth):
    from xreg_amd import synth

    w = synth.config("c2", n_poses=6, small=True)
    cam = w.cams[0]
    rows, cols = cam.num_det_rows, cam.num_det_cols
    X, Y, Z = w.size_xyz
    npose = w.poses.shape[0]
    vol_a = np.ascontiguousarray(w.vol, np.float32)
    vol_b = np.ascontiguousarray(w.vol[::-1, :, ::-1] * 0.7, np.float32)  # a second object, same grid
    pa = w.poses
    pb = np.roll(w.poses, 1, axis=0)
    g, _ = oracle.drr_f32(vol_a, w.idx_to_phys, w.cams, poses_f32(w.gt[None]), step=w.step)
    fixed = synth.noisy_fixed(g[0])
    rng = np.random.default_rng(11)
    bg = rng.uniform(0.0, 0.5 * float(g.max()), size=(rows, cols)).astype(np.float32)
    rr, cc = np.mgrid[0:rows, 0:cols]
    mask = (((rr - rows / 2) ** 2 + (cc - cols / 2) ** 2) <= (0.4 * min(rows, cols)) ** 2).astype(np.uint8)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([X, Y, Z, rows, cols, npose, metric, smooth], np.int32).tofile(f)
        np.array([w.spacing, w.step, 1020.0, cam.det_col_spacing], np.float32).tofile(f)
        vol_a.tofile(f)
        vol_b.tofile(f)
        fixed.astype(np.float32).tofile(f)
        bg.tofile(f)
        mask.tofile(f)
        poses_f32(pa).tofile(f)
        poses_f32(pb).tofile(f)
    out = tmp_path / "out.bin"
    print(f"case: dropin features metric={metric} poses={npose}", flush=True)
    r = subprocess.run([str(FEATURES), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = _records(out.read_bytes())

    def f32(name, shape=None):
        a = np.frombuffer(rec[name], np.float32)
        return a.reshape(shape) if shape else a

    def u64(name):
        return int(np.frombuffer(rec[name], np.uint64)[0])

    shp = (npose, rows, cols)
    ref_a, _ = oracle.drr_f32(vol_a, w.idx_to_phys, w.cams, poses_f32(pa), step=w.step)
    # MAX kernel (REPLACE, default background 0)
    ref_max, _ = oracle.drr_f32(vol_a, w.idx_to_phys, w.cams, poses_f32(pa), step=w.step, kernel=1)
    assert drr_rel_err(f32("max_drr", shp), ref_max).max() <= DRR_REL_TOL
    # nearest-neighbour (bit for bit), B-spline and windowed-sinc interpolation through
    # set_interp_method / use_*_interp
    for name, interp in (("nn_drr", 1), ("bspline_drr", 3), ("sinc_drr", 2)):
        ref_i, _ = oracle.drr_f32(vol_a, w.idx_to_phys, w.cams, poses_f32(pa), step=w.step, interp=interp)
        if interp == 1:
            assert np.array_equal(f32(name, shp), ref_i)
        else:
            assert drr_rel_err(f32(name, shp), ref_i).max() <= DRR_REL_TOL, name
    # anti-aliasing: 3 jittered rays per pixel with the seed the driver fixed
    ref_aa, _ = oracle.drr_f32(vol_a, w.idx_to_phys, w.cams, poses_f32(pa), step=w.step, aa_rays=3, aa_seed=20240611)
    assert np.array_equal(f32("aa_drr", shp), ref_aa)
    # external host buffer: raw_host_pixel_buf() is the caller's memory and proj(i) reads it
    assert u64("ext_bad") == 0
    assert drr_rel_err(f32("ext_drr", shp), ref_a).max() <= DRR_REL_TOL
    # device-bound and host-bound moving images give the same values, and the oracle's
    dev = f32("dev_sims")
    assert np.array_equal(f32("host_sims"), dev)
    ref_sims = oracle.sim_f32(metric, fixed, ref_a, smooth_ksize=smooth)
    assert np.abs(dev - ref_sims).max() <= SIM_ABS_TOL
    # a short batch: the first n - 2 poses only
    assert np.array_equal(f32("short_sims"), dev[:npose - 2])
    # masked metric
    ref_mask = oracle.sim_f32(metric, fixed, ref_a, smooth_ksize=smooth, mask=mask)
    assert np.abs(f32("mask_sims") - ref_mask).max() <= SIM_ABS_TOL
    assert np.abs(ref_mask - ref_sims).max() > 1e-4  # the mask matters for these images
    # two objects: REPLACE with A, then ACCUM with B (pre_compute keeps the buffer)
    ref_two, _ = oracle.drr_f32(vol_b, w.idx_to_phys, w.cams, poses_f32(pb), step=w.step, store=1,
                                out=ref_a.copy())
    assert drr_rel_err(f32("two_obj_drr", shp), ref_two).max() <= DRR_REL_TOL
    assert np.abs(f32("two_obj_sims") - oracle.sim_f32(metric, fixed, ref_two, smooth_ksize=smooth)).max() \
        <= SIM_ABS_TOL
    # object B over a static volume: background image + B
    ref_st, _ = oracle.drr_f32(vol_b, w.idx_to_phys, w.cams, poses_f32(pb), step=w.step, bg=bg[None])
    assert drr_rel_err(f32("static_drr", shp), ref_st).max() <= DRR_REL_TOL
    assert np.abs(f32("static_sims") - oracle.sim_f32(metric, fixed, ref_st, smooth_ksize=smooth)).max() \
        <= SIM_ABS_TOL
    assert u64("bg_restored") == 1
    assert u64("iface_bad") == 0  # to_host_buf / to_ocl_buf throw, aux_info is empty before saving
    # with save_aux_info the patch metrics record each compute call's patch list (r = 5,
    # stride 1: (rows - 10) x (cols - 10) patches, rows of centres outer, weights 1)
    if metric in (2, 3):
        assert u64("aux_calls") == 1 and u64("aux_patches") == (rows - 10) * (cols - 10)
    else:
        assert u64("aux_calls") == 0
    assert u64("aux_bad") == 0
    assert u64("max_num_projs_possible") >= npose
    if metric in (2, 3):
        _check_patch_options(oracle, rec, metric, smooth, fixed, ref_a, mask)


def _check_patch_options(oracle, rec, metric, smooth, fixed, drr, mask):
    """The patch options the driver set (ImgSimMetric2DPatchCommon), each against the oracle's
    general patch combine with the weights compute_weights gives."""
    rows, cols = fixed.shape
    rr, cc = np.mgrid[0:rows, 0:cols]
    wimg = (1.0 + (7 * rr + 3 * cc) % 11).astype(np.float32)
    n_p = (rows - 10) * (cols - 10)

    def check(name, **kw):
        got = np.frombuffer(rec[name], np.float32)
        ref = oracle.patch_sim_weighted_f32(metric, fixed, drr, smooth_ksize=smooth, **kw)
        ref64 = oracle.patch_sim_weighted_f64(metric, fixed.astype(np.float64), drr.astype(np.float64),
                                              smooth_ksize=smooth, **kw)
        print(f"  {name}: max |gpu-cpu| {np.abs(got - ref).max():.3e}", flush=True)
        assert np.all(np.abs(got - ref) <= SIM_ABS_TOL + np.abs(ref - ref64)), name
        return got

    w_img = oracle.patch_weights_f32(fixed.shape, 5, 1, wgt_img=wimg)
    assert np.array_equal(np.frombuffer(rec["popt_wimg_weights"], np.float32), w_img)
    check("popt_wimg_sims", weights=w_img)
    check("popt_mean_sims", weight_in_combine=False, compute_mean=True)
    lst = list(range(3, n_p, 37))
    lst.append(lst[1])
    assert np.array_equal(np.frombuffer(rec["popt_list_inds"], np.uint64), np.array(lst, np.uint64))
    assert int(np.frombuffer(rec["popt_list_num"], np.uint64)[0]) == len(lst)
    list_sims = check("popt_list_sims", weights=w_img, inds=lst)
    # save_all_per_patch_scores: per used patch (local index), every moving image's score
    got = np.frombuffer(rec["popt_list_scores"], np.float32).reshape(len(lst), drr.shape[0])
    ref = oracle.patch_scores_f32(metric, fixed, drr, smooth_ksize=smooth)[:, lst].T
    assert np.abs(got - ref).max() <= 2e-5
    assert list_sims.shape == (drr.shape[0],)
    # random patches: 40 distinct patches whose centres lie at least sqrt(2) r apart
    ri = np.frombuffer(rec["popt_rand_inds"], np.uint64).astype(np.int64)
    assert ri.size == 40 and np.unique(ri).size == 40 and ri.max() < n_p
    ppr = cols - 10
    ctr = np.stack([ri // ppr, ri % ppr], axis=1).astype(np.float64)
    dist = np.sqrt(((ctr[:, None] - ctr[None]) ** 2).sum(-1)) + np.eye(40) * 1e9
    assert dist.min() >= np.sqrt(2.0) * 5
    check("popt_rand_sims", weights=w_img, inds=ri)
    w_mask = oracle.patch_weights_f32(fixed.shape, 5, 1, wgt_img=wimg, mask=mask, normalize=False)
    assert np.any(w_mask == 0) and np.any(w_mask > 1)
    check("popt_wimg_mask_sims", weights=w_mask, mask=mask)
    check("popt_nomaskwgt_sims", mask=mask)
    assert np.array_equal(np.frombuffer(rec["popt_default_sims"], np.float32),
                          np.frombuffer(rec["mask_sims"], np.float32))
    # fixed-image patch variances as weights, over the unmasked pixels, zero at masked centres
    assert int(np.frombuffer(rec["popt_mov_var_refused"], np.uint64)[0]) == 1
    w_var = oracle.patch_weights_f32(fixed.shape, 5, 1, wgt_img=oracle.patch_variance_image_f32(fixed, 5, 1, mask),
                                     mask=mask)
    check("popt_fvar_mask_sims", weights=w_var, mask=mask)
    if metric == 3:
        # each direction weighted by its own fixed gradient image's variances
        fx_, fy_ = oracle.sobel_f32(fixed, smooth)
        wx = oracle.patch_weights_f32(fixed.shape, 5, 1, wgt_img=oracle.patch_variance_image_f32(fx_, 5, 1, mask),
                                      mask=mask)
        wy = oracle.patch_weights_f32(fixed.shape, 5, 1, wgt_img=oracle.patch_variance_image_f32(fy_, 5, 1, mask),
                                      mask=mask)
        got = np.frombuffer(rec["popt_gvar_mask_sims"], np.float32)
        ref = np.empty(drr.shape[0], np.float32)
        for m in range(drr.shape[0]):
            mx, my = oracle.sobel_f32(drr[m], smooth)
            sx = oracle.patch_sim_weighted_f32(2, fx_, mx[None], weights=wx, mask=mask)[0]
            sy = oracle.patch_sim_weighted_f32(2, fy_, my[None], weights=wy, mask=mask)[0]
            ref[m] = np.float32(0.5 * float(np.float32(sx + sy)))
        print(f"  popt_gvar_mask_sims: max |gpu-cpu| {np.abs(got - ref).max():.3e}", flush=True)
        assert np.abs(got - ref).max() <= 2 * SIM_ABS_TOL
        assert np.abs(got - np.frombuffer(rec["popt_fvar_mask_sims"], np.float32)).max() > 1e-5
    # patch statistics over the unmasked pixels (use_mask_for_patch_stats), then off again: the
    # masked metric's values
    w_mask = oracle.patch_weights_f32(fixed.shape, 5, 1, mask=mask)
    check("popt_mstats_sims", weights=w_mask, mask=mask, mask_stats=True)
    assert np.array_equal(np.frombuffer(rec["popt_mstats_off_sims"], np.float32),
                          np.frombuffer(rec["mask_sims"], np.float32))
