"""Writes the committed golden fixtures under tests/golden/ (run from the repo root):

    python tests/golden/make_golden.py

The reference ships no vectors for this path, so these fixtures come from this repository's
oracle (FP32 restatement and FP64 oracle), whose functions are pinned by closed forms and
brute force in tests/test_oracle_*.py. They freeze small, fully specified cases (inputs
included) so that any later change of the oracle or of the GPU path shows up as a diff.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from xreg_amd import geometry as G  # noqa: E402
from xreg_amd import synth  # noqa: E402

OUT = Path(__file__).resolve().parent


def cam_arrays(cam):
    kinv, einv, ph = G.camera_arrays(cam)
    meta = np.array([cam.num_det_rows, cam.num_det_cols, cam.coord_frame], np.int32)
    return meta, np.float32(cam.focal_len), kinv, einv, ph


def drr_case():
    w = synth.config("c1", n_poses=5, small=True)
    poses = np.concatenate([w.gt[None], w.poses]).astype(np.float32).reshape(-1, 16)
    d32, s32 = O.drr_f32(w.vol, w.idx_to_phys, w.cams, poses, step=w.step, threads=1)
    d64, s64 = O.drr_f64(w.vol, w.idx_to_phys, w.cams, poses.astype(np.float64), step=w.step, threads=1)
    dmax, _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, poses, step=w.step, kernel=1, threads=1)
    meta, focal, kinv, einv, ph = cam_arrays(w.cams[0])
    np.savez_compressed(OUT / "drr_c1_small.npz", vol=w.vol, idx_to_phys=w.idx_to_phys, step=np.float32(w.step),
                        poses=poses, cam_meta=meta, cam_focal=focal, cam_kinv=kinv, cam_einv=einv, cam_pinhole=ph,
                        drr_f32=d32, drr_f64=d64, drr_max_f32=dmax, samples=np.int64(s32),
                        samples_f64=np.int64(s64))


def sim_case():
    w = synth.config("c2", n_poses=3, small=True)
    poses = np.concatenate([w.gt[None], w.poses]).astype(np.float32).reshape(-1, 16)
    drr, _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, poses, step=w.step, threads=1)
    fixed = synth.noisy_fixed(drr[0])
    out = {"fixed": fixed, "mov": drr}
    for metric in range(4):
        for k in (0, 5):
            out[f"sim{metric}_k{k}_f32"] = O.sim_f32(metric, fixed, drr, smooth_ksize=k, threads=1)
            out[f"sim{metric}_k{k}_f64"] = O.sim_f64(metric, fixed.astype(np.float64), drr.astype(np.float64),
                                                     smooth_ksize=k, threads=1)
    np.savez_compressed(OUT / "sim_c2_small.npz", **out)


def interp_case():
    """The CPU caster's other sampling options on the small C1 case: nearest-neighbour,
    windowed-sinc and B-spline DRRs, and linear / nearest-neighbour DRRs with 2 anti-aliasing
    rays per pixel (jitter seed 7)."""
    w = synth.config("c1", n_poses=2, small=True)
    poses = np.concatenate([w.gt[None], w.poses]).astype(np.float32).reshape(-1, 16)
    out = {"vol": w.vol, "idx_to_phys": w.idx_to_phys, "step": np.float32(w.step), "poses": poses}
    meta, focal, kinv, einv, ph = cam_arrays(w.cams[0])
    out.update(cam_meta=meta, cam_focal=focal, cam_kinv=kinv, cam_einv=einv, cam_pinhole=ph)
    for name, interp in (("nn", 1), ("sinc", 2), ("bspline", 3)):
        out[f"drr_{name}"], _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, poses, step=w.step, interp=interp, threads=1)
    for name, interp in (("linear", 0), ("nn", 1)):
        out[f"drr_{name}_aa2"], _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, poses, step=w.step, interp=interp,
                                             aa_rays=2, aa_seed=7, threads=1)
    np.savez_compressed(OUT / "interp_c1_small.npz", **out)


PATCHW_SETTINGS = [
    # (name, weights key or None, indices key or None, weight_in_combine, compute_mean, masked)
    ("var_all", "w_var", None, True, False, False),
    ("var_list_mean", "w_var", "inds", True, True, False),
    ("unit_list_sum", None, "inds", False, False, False),
    ("mask_var_list", "w_var_mask", "inds", True, False, True),
]


def patchw_case():
    """Patch metrics with the general combine on the small C2 case: weights from the fixed
    image's patch variances (with and without a mask), a list of patches with repeats, the
    mean and unweighted combines."""
    w = synth.config("c2", n_poses=3, small=True)
    poses = np.concatenate([w.gt[None], w.poses]).astype(np.float32).reshape(-1, 16)
    drr, _ = O.drr_f32(w.vol, w.idx_to_phys, w.cams, poses, step=w.step, threads=1)
    fixed = synth.noisy_fixed(drr[0])
    mask = np.ones(fixed.shape, np.uint8)
    mask[:12, :20] = 0
    mask[30:40, 5:60] = 0
    r = 3
    n_p = (fixed.shape[0] - 2 * r) * (fixed.shape[1] - 2 * r)
    out = {"fixed": fixed, "mov": drr, "mask": mask, "radius": np.int32(r),
           "w_var": O.patch_weights_f32(fixed.shape, r, 1, wgt_img=O.patch_variance_image_f32(fixed, r, 1)),
           "w_var_mask": O.patch_weights_f32(fixed.shape, r, 1, wgt_img=O.patch_variance_image_f32(fixed, r, 1, mask),
                                             mask=mask),
           "inds": np.concatenate([np.arange(5, n_p, 53), [17, 17, 17]]).astype(np.int64)}
    for metric in (2, 3):
        for name, wk, ik, wic, mean, masked in PATCHW_SETTINGS:
            kw = dict(weights=out[wk] if wk else None, inds=out[ik] if ik else None, weight_in_combine=wic,
                      compute_mean=mean, smooth_ksize=5, radius=r, mask=mask if masked else None, threads=1)
            out[f"m{metric}_{name}_f32"] = O.patch_sim_weighted_f32(metric, fixed, drr, **kw)
            out[f"m{metric}_{name}_f64"] = O.patch_sim_weighted_f64(metric, fixed.astype(np.float64),
                                                                    drr.astype(np.float64), **kw)
    np.savez_compressed(OUT / "patchw_c2_small.npz", **out)


if __name__ == "__main__":
    drr_case()
    sim_case()
    interp_case()
    patchw_case()
    print("wrote", sorted(p.name for p in OUT.glob("*.npz")))
