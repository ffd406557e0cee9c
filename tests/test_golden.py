"""The oracle reproduces the committed golden fixtures (CPU), and the GPU path matches them
(GPU-marked)."""
from pathlib import Path

import numpy as np
import pytest

from tests.helpers import DRR_REL_TOL, SIM_ABS_TOL, drr_rel_err

GOLD = Path(__file__).resolve().parent / "golden"


def _cam(z):
    from xreg_amd._xrh import Camera

    c = Camera()
    c.num_det_rows, c.num_det_cols, c.coord_frame = (int(v) for v in z["cam_meta"])
    c.focal_len = float(z["cam_focal"])
    for i, v in enumerate(z["cam_kinv"].reshape(-1)):
        c.intrins_inv[i] = float(v)
    for i, v in enumerate(z["cam_einv"].reshape(-1)):
        c.extrins_inv[i] = float(v)
    for i, v in enumerate(z["cam_pinhole"]):
        c.pinhole[i] = float(v)
    return c


@pytest.fixture(scope="module")
def gdrr():
    return np.load(GOLD / "drr_c1_small.npz")


@pytest.fixture(scope="module")
def gsim():
    return np.load(GOLD / "sim_c2_small.npz")


@pytest.fixture(scope="module")
def ginterp():
    return np.load(GOLD / "interp_c1_small.npz")


INTERP_CASES = (("drr_nn", 1, 0), ("drr_sinc", 2, 0), ("drr_bspline", 3, 0), ("drr_linear_aa2", 0, 2),
                ("drr_nn_aa2", 1, 2))


def test_oracle_reproduces_golden_interp(oracle, ginterp):
    z = ginterp
    cam = _cam(z)
    for name, interp, aa in INTERP_CASES:
        got, _ = oracle.drr_f32(z["vol"], z["idx_to_phys"], [cam], z["poses"], step=float(z["step"]), interp=interp,
                                aa_rays=aa, aa_seed=7)
        assert np.array_equal(got, z[name]), name


def test_oracle_reproduces_golden_drr(oracle, gdrr):
    z = gdrr
    cam = _cam(z)
    d32, s = oracle.drr_f32(z["vol"], z["idx_to_phys"], [cam], z["poses"], step=float(z["step"]))
    assert np.array_equal(d32, z["drr_f32"]) and s == int(z["samples"])
    d64, _ = oracle.drr_f64(z["vol"], z["idx_to_phys"], [cam], z["poses"].astype(np.float64),
                            step=float(z["step"]))
    assert np.allclose(d64, z["drr_f64"], rtol=1e-13, atol=1e-15)
    dmax, _ = oracle.drr_f32(z["vol"], z["idx_to_phys"], [cam], z["poses"], step=float(z["step"]), kernel=1)
    assert np.array_equal(dmax, z["drr_max_f32"])


def test_oracle_reproduces_golden_sims(oracle, gsim):
    z = gsim
    for metric in range(4):
        for k in (0, 5):
            got = oracle.sim_f32(metric, z["fixed"], z["mov"], smooth_ksize=k)
            assert np.array_equal(got, z[f"sim{metric}_k{k}_f32"]), (metric, k)
            got64 = oracle.sim_f64(metric, z["fixed"].astype(np.float64), z["mov"].astype(np.float64),
                                   smooth_ksize=k)
            assert np.allclose(got64, z[f"sim{metric}_k{k}_f64"], rtol=0, atol=1e-13)


@pytest.mark.gpu
def test_gpu_matches_golden_drr(ctx, gdrr):
    from xreg_amd import _xrh as X

    z = gdrr
    cam = _cam(z)
    n = z["poses"].shape[0]
    rows, cols = cam.num_det_rows, cam.num_det_cols
    vol = X.Volume(ctx, z["vol"], z["idx_to_phys"])
    buf = ctx.alloc(4 * n * rows * cols)
    X.line_int(ctx, vol, [cam], z["poses"], np.zeros(n, np.int64), float(z["step"]), buf)
    got = buf.download(np.float32).reshape(n, rows, cols)
    assert drr_rel_err(got, z["drr_f32"]).max() <= DRR_REL_TOL
    s = X.line_int_sample_count(ctx, vol, [cam], z["poses"], np.zeros(n, np.int64), float(z["step"]))
    assert s == int(z["samples"])
    buf.free()
    vol.close()


@pytest.mark.gpu
def test_gpu_matches_golden_sims(ctx, gsim):
    from xreg_amd import _xrh as X

    z = gsim
    mov = z["mov"]
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    for metric in range(4):
        for k in (0, 5):
            f = X.Fixed(ctx, z["fixed"], metric, smooth_ksize=k)
            got = f.compute(buf, mov.shape[0])
            f.close()
            assert np.abs(got - z[f"sim{metric}_k{k}_f32"]).max() <= SIM_ABS_TOL, (metric, k)
            assert np.abs(got - z[f"sim{metric}_k{k}_f64"]).max() <= SIM_ABS_TOL, (metric, k)
    buf.free()


@pytest.mark.gpu
def test_gpu_matches_golden_interp(ctx, ginterp):
    """Nearest neighbour and the anti-aliased linear / nearest-neighbour casts bit for bit, sinc
    and B-spline within the DRR tolerance."""
    from xreg_amd import _xrh as X

    z = ginterp
    cam = _cam(z)
    n = z["poses"].shape[0]
    rows, cols = cam.num_det_rows, cam.num_det_cols
    vol = X.Volume(ctx, z["vol"], z["idx_to_phys"])
    buf = ctx.alloc(4 * n * rows * cols)
    for name, interp, aa in INTERP_CASES:
        X.line_int_interp(ctx, vol, [cam], z["poses"], np.zeros(n, np.int64), float(z["step"]), buf, interp,
                          aa_rays=aa, aa_seed=7)
        got = buf.download(np.float32).reshape(n, rows, cols)
        if interp in (0, 1):
            assert np.array_equal(got, z[name]), name
        else:
            assert drr_rel_err(got, z[name]).max() <= DRR_REL_TOL, name
    buf.free()
    vol.close()


@pytest.fixture(scope="module")
def gpatchw():
    return np.load(GOLD / "patchw_c2_small.npz")


# (name, weights key or None, indices key or None, weight_in_combine, compute_mean, masked); the
# settings tests/golden/make_golden.py wrote
PATCHW_SETTINGS = (("var_all", "w_var", None, True, False, False),
                   ("var_list_mean", "w_var", "inds", True, True, False),
                   ("unit_list_sum", None, "inds", False, False, False),
                   ("mask_var_list", "w_var_mask", "inds", True, False, True))


def test_oracle_reproduces_golden_patch_weights(oracle, gpatchw):
    z = gpatchw
    r = int(z["radius"])
    assert np.array_equal(oracle.patch_weights_f32(z["fixed"].shape, r, 1, wgt_img=oracle.patch_variance_image_f32(
        z["fixed"], r, 1)), z["w_var"])
    for metric in (2, 3):
        for name, wk, ik, wic, mean, masked in PATCHW_SETTINGS:
            got = oracle.patch_sim_weighted_f32(metric, z["fixed"], z["mov"], weights=z[wk] if wk else None,
                                                inds=z[ik] if ik else None, weight_in_combine=wic, compute_mean=mean,
                                                smooth_ksize=5, radius=r, mask=z["mask"] if masked else None)
            assert np.array_equal(got, z[f"m{metric}_{name}_f32"]), (metric, name)


@pytest.mark.gpu
def test_gpu_matches_golden_patch_weights(ctx, gpatchw):
    from xreg_amd import _xrh as X

    z = gpatchw
    r = int(z["radius"])
    mov = z["mov"]
    buf = ctx.alloc(mov.nbytes)
    buf.upload(mov)
    for metric in (2, 3):
        for name, wk, ik, wic, mean, masked in PATCHW_SETTINGS:
            print(f"case: golden patch weights metric={metric} {name}", flush=True)
            f = X.Fixed(ctx, z["fixed"], metric, smooth_ksize=5, patch_radius=r, mask=z["mask"] if masked else None)
            f.set_patch_weights(weights=z[wk] if wk else None, patch_inds=z[ik] if ik else None,
                                weight_in_combine=wic, compute_mean=mean)
            got = f.compute(buf, mov.shape[0])
            f.close()
            ref32, ref64 = z[f"m{metric}_{name}_f32"], z[f"m{metric}_{name}_f64"]
            tol = SIM_ABS_TOL * np.maximum(1.0, np.abs(ref64))
            assert np.all(np.abs(got - ref64) <= tol), (metric, name)
            assert np.all(np.abs(got - ref32) <= tol + np.abs(ref32 - ref64)), (metric, name)
    buf.free()
