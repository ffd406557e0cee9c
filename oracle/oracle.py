"""TEST INFRASTRUCTURE ONLY: numpy wrappers of liboracle.so (the CPU restatement of the
reference's DRR + similarity path, see xrh_oracle.h). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg load this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


class OrcCamF(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("frame", C.c_int32), ("focal", C.c_float),
                ("kinv", C.c_float * 9), ("einv", C.c_float * 16), ("pinhole", C.c_float * 3)]


class OrcCamD(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("frame", C.c_int32), ("pad_", C.c_int32),
                ("focal", C.c_double), ("kinv", C.c_double * 9), ("einv", C.c_double * 16),
                ("pinhole", C.c_double * 3)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def cam_f(cam) -> OrcCamF:
    """From an xrh Camera (same float members)."""
    o = OrcCamF()
    o.rows, o.cols, o.frame, o.focal = cam.num_det_rows, cam.num_det_cols, cam.coord_frame, cam.focal_len
    for i in range(9):
        o.kinv[i] = cam.intrins_inv[i]
    for i in range(16):
        o.einv[i] = cam.extrins_inv[i]
    for i in range(3):
        o.pinhole[i] = cam.pinhole[i]
    return o


def cam_d(cam) -> OrcCamD:
    o = OrcCamD()
    o.rows, o.cols, o.frame, o.focal = cam.num_det_rows, cam.num_det_cols, cam.coord_frame, cam.focal_len
    for i in range(9):
        o.kinv[i] = cam.intrins_inv[i]
    for i in range(16):
        o.einv[i] = cam.extrins_inv[i]
    for i in range(3):
        o.pinhole[i] = cam.pinhole[i]
    return o


def cam_naive(focal, rows, cols, row_sp, col_sp) -> OrcCamF:
    o = OrcCamF()
    lib().orc_cam_naive_f32(C.c_float(focal), rows, cols, C.c_float(row_sp), C.c_float(col_sp), C.byref(o))
    return o


def p2i_f32(idx_to_phys: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(idx_to_phys, dtype=np.float32).reshape(16)
    out = np.zeros(16, dtype=np.float32)
    lib().orc_affine_inverse_f32(_f(m), _f(out))
    return out


def p2i_f64(idx_to_phys: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(idx_to_phys, dtype=np.float64).reshape(16)
    out = np.zeros(16, dtype=np.float64)
    lib().orc_affine_inverse_f64(_d(m), _d(out))
    return out


def drr_f32(vol_zyx, idx_to_phys, cams, xforms, cam_for_proj=None, step=1.0, kernel=0, store=0, bg=None,
            default_bg=0.0, out=None, threads=0, interp=0, aa_rays=0, aa_seed=0):
    """RayCasterLineIntCPU restatement. Returns (drr [n, rows, cols] float32, samples).
    interp: RayCaster::InterpMethod (0 linear, 1 nearest neighbour, 2 windowed sinc, 3 B-spline);
    aa_rays: RayCasterCPU's anti-aliasing factor (0 off), jitter drawn by aa_jitter(aa_seed, ...)."""
    vol = np.ascontiguousarray(vol_zyx, dtype=np.float32)
    dims = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], dtype=np.int64)
    x = np.ascontiguousarray(np.asarray(xforms, dtype=np.float32).reshape(-1, 16))
    n = x.shape[0]
    c4p = np.zeros(n, dtype=np.int64) if cam_for_proj is None else np.ascontiguousarray(cam_for_proj, np.int64)
    ca = (OrcCamF * len(cams))(*[cam_f(c) for c in cams])
    rows, cols = cams[0].num_det_rows, cams[0].num_det_cols
    if out is None:
        out = np.zeros((n, rows, cols), dtype=np.float32)
    bgp = None
    if bg is not None:
        bga = np.ascontiguousarray(bg, dtype=np.float32)
        bgp = _f(bga)
    s = C.c_int64(0)
    p2i = p2i_f32(idx_to_phys)
    lib().orc_drr_interp_f32(_f(vol), _i64(dims), _f(p2i), ca, len(cams), _f(x), _i64(c4p), C.c_int64(n),
                             C.c_float(step), kernel, store, bgp, C.c_float(default_bg), int(interp),
                             int(aa_rays), C.c_uint64(int(aa_seed) & (2**64 - 1)), _f(out), threads, C.byref(s))
    return out, int(s.value)


def drr_f64(vol_zyx, idx_to_phys, cams, xforms, cam_for_proj=None, step=1.0, kernel=0, store=0, bg=None,
            default_bg=0.0, out=None, threads=0):
    """FP64 oracle (FP64 volume, geometry and accumulation). xforms may be float64."""
    vol = np.ascontiguousarray(vol_zyx, dtype=np.float64)
    dims = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], dtype=np.int64)
    x = np.ascontiguousarray(np.asarray(xforms, dtype=np.float64).reshape(-1, 16))
    n = x.shape[0]
    c4p = np.zeros(n, dtype=np.int64) if cam_for_proj is None else np.ascontiguousarray(cam_for_proj, np.int64)
    ca = (OrcCamD * len(cams))(*[cam_d(c) for c in cams])
    rows, cols = cams[0].num_det_rows, cams[0].num_det_cols
    if out is None:
        out = np.zeros((n, rows, cols), dtype=np.float64)
    bgp = None
    if bg is not None:
        bga = np.ascontiguousarray(bg, dtype=np.float64)
        bgp = _d(bga)
    s = C.c_int64(0)
    p2i = p2i_f64(np.asarray(idx_to_phys, dtype=np.float64))
    lib().orc_drr_f64(_d(vol), _i64(dims), _d(p2i), ca, len(cams), _d(x), _i64(c4p), C.c_int64(n),
                      C.c_double(step), kernel, store, bgp, C.c_double(default_bg), _d(out), threads, C.byref(s))
    return out, int(s.value)


def ray_steps(vol_shape_zyx, idx_to_phys, cams, xforms, cam_for_proj=None, step=1.0, fp64=False, threads=0):
    """Per-ray last sample index n (-1: miss) of drr_f32 (or drr_f64 with fp64=True): the ray
    set-up alone, [n_proj, rows, cols] int32."""
    dims = np.array([vol_shape_zyx[2], vol_shape_zyx[1], vol_shape_zyx[0]], dtype=np.int64)
    dt = np.float64 if fp64 else np.float32
    x = np.ascontiguousarray(np.asarray(xforms, dtype=dt).reshape(-1, 16))
    n = x.shape[0]
    c4p = np.zeros(n, dtype=np.int64) if cam_for_proj is None else np.ascontiguousarray(cam_for_proj, np.int64)
    rows, cols = cams[0].num_det_rows, cams[0].num_det_cols
    out = np.empty((n, rows, cols), dtype=np.int32)
    op = out.ctypes.data_as(C.POINTER(C.c_int32))
    if fp64:
        ca = (OrcCamD * len(cams))(*[cam_d(c) for c in cams])
        p2i = p2i_f64(np.asarray(idx_to_phys, dtype=np.float64))
        lib().orc_ray_steps_f64(_i64(dims), _d(p2i), ca, _d(x), _i64(c4p), C.c_int64(n), C.c_double(step), op,
                                threads)
    else:
        ca = (OrcCamF * len(cams))(*[cam_f(c) for c in cams])
        p2i = p2i_f32(idx_to_phys)
        lib().orc_ray_steps_f32(_i64(dims), _f(p2i), ca, _f(x), _i64(c4p), C.c_int64(n), C.c_float(step), op,
                                threads)
    return out


def ncc_f32(fixed, mov, mask=None):
    f = np.ascontiguousarray(fixed, dtype=np.float32)
    m = np.ascontiguousarray(mov, dtype=np.float32).reshape(-1, f.size)
    out = np.zeros(m.shape[0], dtype=np.float32)
    mk = None
    if mask is not None:
        mka = np.ascontiguousarray(mask, dtype=np.uint8)
        mk = mka.ctypes.data_as(C.POINTER(C.c_uint8))
    lib().orc_ncc_f32(_f(f), _f(m), C.c_int64(f.size), C.c_int64(m.shape[0]), mk, _f(out))
    return out


def ncc_f64(fixed, mov, mask=None):
    f = np.ascontiguousarray(fixed, dtype=np.float64)
    m = np.ascontiguousarray(mov, dtype=np.float64).reshape(-1, f.size)
    out = np.zeros(m.shape[0], dtype=np.float64)
    mk = None
    if mask is not None:
        mka = np.ascontiguousarray(mask, dtype=np.uint8)
        mk = mka.ctypes.data_as(C.POINTER(C.c_uint8))
    lib().orc_ncc_f64(_d(f), _d(m), C.c_int64(f.size), C.c_int64(m.shape[0]), mk, _d(out))
    return out


def sobel_f32(img, smooth_ksize=0):
    im = np.ascontiguousarray(img, dtype=np.float32)
    gx = np.zeros_like(im)
    gy = np.zeros_like(im)
    lib().orc_sobel_f32(_f(im), im.shape[0], im.shape[1], smooth_ksize, _f(gx), _f(gy))
    return gx, gy


def sobel_f64(img, smooth_ksize=0):
    im = np.ascontiguousarray(img, dtype=np.float64)
    gx = np.zeros_like(im)
    gy = np.zeros_like(im)
    lib().orc_sobel_f64(_d(im), im.shape[0], im.shape[1], smooth_ksize, _d(gx), _d(gy))
    return gx, gy


def grad_ncc_f32(fixed, mov, smooth_ksize=5, mask=None):
    f = np.ascontiguousarray(fixed, dtype=np.float32)
    m = np.ascontiguousarray(mov, dtype=np.float32).reshape(-1, *f.shape)
    if mask is not None:
        # ImgSimMetric2DGradNCCCPU with a mask: the masked NCC of the x and of the y gradients
        fgx, fgy = sobel_f32(f, smooth_ksize)
        g = [sobel_f32(mi, smooth_ksize) for mi in m]
        sx = ncc_f32(fgx, np.stack([a for a, _ in g]), mask=mask)
        sy = ncc_f32(fgy, np.stack([b for _, b in g]), mask=mask)
        return (np.float32(0.5) * (sx + sy)).astype(np.float32)
    out = np.zeros(m.shape[0], dtype=np.float32)
    lib().orc_grad_ncc_f32(_f(f), _f(m), f.shape[0], f.shape[1], C.c_int64(m.shape[0]), smooth_ksize, _f(out))
    return out


def grad_ncc_f64(fixed, mov, smooth_ksize=5):
    f = np.ascontiguousarray(fixed, dtype=np.float64)
    m = np.ascontiguousarray(mov, dtype=np.float64).reshape(-1, *f.shape)
    out = np.zeros(m.shape[0], dtype=np.float64)
    lib().orc_grad_ncc_f64(_d(f), _d(m), f.shape[0], f.shape[1], C.c_int64(m.shape[0]), smooth_ksize, _d(out))
    return out


def _mask_ptr(mask, shape):
    if mask is None:
        return None, None
    mk = np.ascontiguousarray(mask, dtype=np.uint8).reshape(shape)
    return mk, mk.ctypes.data_as(C.POINTER(C.c_uint8))


def patch_ncc_f32(fixed, mov, radius=5, stride=1, threads=0, mask=None):
    f = np.ascontiguousarray(fixed, dtype=np.float32)
    m = np.ascontiguousarray(mov, dtype=np.float32).reshape(-1, *f.shape)
    out = np.zeros(m.shape[0], dtype=np.float32)
    mk, mp = _mask_ptr(mask, f.shape)
    lib().orc_patch_ncc_masked_f32(_f(f), _f(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]), radius,
                                   stride, _f(out), threads)
    return out


def patch_ncc_f64(fixed, mov, radius=5, stride=1, threads=0, mask=None):
    f = np.ascontiguousarray(fixed, dtype=np.float64)
    m = np.ascontiguousarray(mov, dtype=np.float64).reshape(-1, *f.shape)
    out = np.zeros(m.shape[0], dtype=np.float64)
    mk, mp = _mask_ptr(mask, f.shape)
    lib().orc_patch_ncc_masked_f64(_d(f), _d(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]), radius,
                                   stride, _d(out), threads)
    return out


def patch_grad_ncc_f32(fixed, mov, smooth_ksize=5, radius=5, stride=1, threads=0, mask=None):
    f = np.ascontiguousarray(fixed, dtype=np.float32)
    m = np.ascontiguousarray(mov, dtype=np.float32).reshape(-1, *f.shape)
    out = np.zeros(m.shape[0], dtype=np.float32)
    mk, mp = _mask_ptr(mask, f.shape)
    lib().orc_patch_grad_ncc_masked_f32(_f(f), _f(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]),
                                        smooth_ksize, radius, stride, _f(out), threads)
    return out


def patch_grad_ncc_f64(fixed, mov, smooth_ksize=5, radius=5, stride=1, threads=0, mask=None):
    f = np.ascontiguousarray(fixed, dtype=np.float64)
    m = np.ascontiguousarray(mov, dtype=np.float64).reshape(-1, *f.shape)
    out = np.zeros(m.shape[0], dtype=np.float64)
    mk, mp = _mask_ptr(mask, f.shape)
    lib().orc_patch_grad_ncc_masked_f64(_d(f), _d(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]),
                                        smooth_ksize, radius, stride, _d(out), threads)
    return out


def patch_weights_f32(shape, radius=5, stride=1, wgt_img=None, mask=None, use_mask_for_weighting=True,
                      normalize=True):
    """ImgSimMetric2DPatchCommon::compute_weights: per-patch weights, rows of centres outer."""
    rows, cols = shape
    ppr = (cols - 1 - 2 * radius) // stride + 1
    ppc = (rows - 1 - 2 * radius) // stride + 1
    out = np.zeros(ppr * ppc, dtype=np.float32)
    wi = None if wgt_img is None else np.ascontiguousarray(wgt_img, dtype=np.float32).reshape(shape)
    mk, mp = _mask_ptr(mask, shape)
    lib().orc_patch_weights_f32(None if wi is None else _f(wi), mp, rows, cols, radius, stride,
                                int(bool(use_mask_for_weighting)), int(bool(normalize)), _f(out))
    return out


def patch_variance_image_f32(img, radius=5, stride=1, mask=None):
    """Fixed-image patch variances at the patch centres (use_fixed_img_patch_variances_as_wgts)."""
    im = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros_like(im)
    mk, mp = _mask_ptr(mask, im.shape)
    lib().orc_patch_variance_image_f32(_f(im), mp, im.shape[0], im.shape[1], radius, stride, _f(out))
    return out


def _inds(inds):
    if inds is None:
        return None, None, 0
    a = np.ascontiguousarray(inds, dtype=np.int64).ravel()
    return a, _i64(a), a.size


def patch_sim_weighted_f32(metric, fixed, mov, weights=None, inds=None, weight_in_combine=True, compute_mean=False,
                           smooth_ksize=5, radius=5, stride=1, mask=None, threads=0, mask_stats=False):
    """Patch-NCC (metric 2) / Patch-Grad-NCC (3) with the general combine of
    ImgSimMetric2DPatchNCCCPU::compute: patch weights, the used patch indices, and the
    weight-in-combine / mean flags."""
    f = np.ascontiguousarray(fixed, dtype=np.float32)
    m = np.ascontiguousarray(mov, dtype=np.float32).reshape(-1, *f.shape)
    out = np.zeros(m.shape[0], dtype=np.float32)
    mk, mp = _mask_ptr(mask, f.shape)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32).ravel()
    ia, ip, n = _inds(inds)
    wp = None if w is None else _f(w)
    if metric == 2:
        lib().orc_patch_ncc_weighted_f32(_f(f), _f(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]), radius,
                                         stride, wp, ip, C.c_int64(n), int(bool(weight_in_combine)),
                                         int(bool(compute_mean)), int(bool(mask_stats)), _f(out), threads)
    else:
        lib().orc_patch_grad_ncc_weighted_f32(_f(f), _f(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]),
                                              smooth_ksize, radius, stride, wp, ip, C.c_int64(n),
                                              int(bool(weight_in_combine)), int(bool(compute_mean)),
                                              int(bool(mask_stats)), _f(out), threads)
    return out


def patch_sim_weighted_f64(metric, fixed, mov, weights=None, inds=None, weight_in_combine=True, compute_mean=False,
                           smooth_ksize=5, radius=5, stride=1, mask=None, threads=0, mask_stats=False):
    f = np.ascontiguousarray(fixed, dtype=np.float64)
    m = np.ascontiguousarray(mov, dtype=np.float64).reshape(-1, *f.shape)
    out = np.zeros(m.shape[0], dtype=np.float64)
    mk, mp = _mask_ptr(mask, f.shape)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64).ravel()
    ia, ip, n = _inds(inds)
    wp = None if w is None else _d(w)
    if metric == 2:
        lib().orc_patch_ncc_weighted_f64(_d(f), _d(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]), radius,
                                         stride, wp, ip, C.c_int64(n), int(bool(weight_in_combine)),
                                         int(bool(compute_mean)), int(bool(mask_stats)), _d(out), threads)
    else:
        lib().orc_patch_grad_ncc_weighted_f64(_d(f), _d(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]),
                                              smooth_ksize, radius, stride, wp, ip, C.c_int64(n),
                                              int(bool(weight_in_combine)), int(bool(compute_mean)),
                                              int(bool(mask_stats)), _d(out), threads)
    return out


def patch_scores_f32(metric, fixed, mov, smooth_ksize=5, radius=5, stride=1, mask=None, mask_stats=False,
                     threads=0):
    """Every grid patch's unweighted score per moving image (num_mov, num_patches)."""
    f = np.ascontiguousarray(fixed, dtype=np.float32)
    m = np.ascontiguousarray(mov, dtype=np.float32).reshape(-1, *f.shape)
    ppr = (f.shape[1] - 1 - 2 * radius) // stride + 1
    ppc = (f.shape[0] - 1 - 2 * radius) // stride + 1
    out = np.zeros((m.shape[0], ppr * ppc), dtype=np.float32)
    mk, mp = _mask_ptr(mask, f.shape)
    lib().orc_patch_scores_f32(int(metric), _f(f), _f(m), mp, f.shape[0], f.shape[1], C.c_int64(m.shape[0]),
                               smooth_ksize, radius, stride, int(bool(mask_stats)), _f(out), threads)
    return out


def sim_f32(metric, fixed, mov, smooth_ksize=5, radius=5, stride=1, threads=0, mask=None):
    if metric == 0:
        return ncc_f32(fixed, mov, mask=mask)
    if metric == 1:
        return grad_ncc_f32(fixed, mov, smooth_ksize, mask=mask)
    if metric == 2:
        return patch_ncc_f32(fixed, mov, radius, stride, threads, mask=mask)
    return patch_grad_ncc_f32(fixed, mov, smooth_ksize, radius, stride, threads, mask=mask)


def sim_f64(metric, fixed, mov, smooth_ksize=5, radius=5, stride=1, threads=0):
    if metric == 0:
        return ncc_f64(fixed, mov)
    if metric == 1:
        return grad_ncc_f64(fixed, mov, smooth_ksize)
    if metric == 2:
        return patch_ncc_f64(fixed, mov, radius, stride, threads)
    return patch_grad_ncc_f64(fixed, mov, smooth_ksize, radius, stride, threads)


def itk_interp(vol_zyx, idx):
    vol = np.ascontiguousarray(vol_zyx, dtype=np.float32)
    dims = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], dtype=np.int64)
    p = np.ascontiguousarray(idx, dtype=np.float32)
    lib().orc_itk_linear_interp.restype = C.c_double
    return float(lib().orc_itk_linear_interp(_f(vol), _i64(dims), _f(p)))


def itk_interp_points(vol_zyx, idx, interp):
    """The reference's interpolator `interp` (0 linear, 1 nearest neighbour, 2 windowed sinc,
    3 B-spline) at each float continuous index (x, y, z) of idx [n, 3]; returns float64 [n]."""
    vol = np.ascontiguousarray(vol_zyx, dtype=np.float32)
    dims = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], dtype=np.int64)
    pts = np.ascontiguousarray(np.asarray(idx, dtype=np.float32).reshape(-1, 3))
    out = np.empty(pts.shape[0], dtype=np.float64)
    L = lib()
    if interp == 3:
        coef = np.empty(vol.size, dtype=np.float64)
        L.orc_bspline_coefficients(_f(vol), _i64(dims), _d(coef))
        fn = L.orc_itk_bspline_interp
        fn.restype = C.c_double
        for i, p in enumerate(pts):
            out[i] = fn(_d(coef), _i64(dims), _f(np.ascontiguousarray(p)))
        return out
    fn = {0: L.orc_itk_linear_interp, 1: L.orc_itk_nn_interp, 2: L.orc_itk_sinc_interp}[interp]
    fn.restype = C.c_double
    for i, p in enumerate(pts):
        out[i] = fn(_f(vol), _i64(dims), _f(np.ascontiguousarray(p)))
    return out


def aa_jitter(seed, pixel, k, comp):
    """orc_aa_jitter: the jitter of ray k of range index `pixel`, component 0 (column) or 1 (row)."""
    fn = lib().orc_aa_jitter
    fn.restype = C.c_float
    fn.argtypes = [C.c_uint64, C.c_uint64, C.c_int32, C.c_int32]
    return float(fn(int(seed) & (2**64 - 1), int(pixel), int(k), int(comp)))


def bspline_coefficients(vol_zyx):
    """orc_bspline_coefficients: the cubic B-spline coefficients of the volume (float64, zyx)."""
    vol = np.ascontiguousarray(vol_zyx, dtype=np.float32)
    dims = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], dtype=np.int64)
    coef = np.empty(vol.shape, dtype=np.float64)
    lib().orc_bspline_coefficients(_f(vol), _i64(dims), _d(coef))
    return coef


def ray_rect_f32(mn, mx, p0, d, limit=True):
    a = [np.ascontiguousarray(v, dtype=np.float32) for v in (mn, mx, p0, d)]
    t0, t1 = C.c_float(0), C.c_float(0)
    hit = lib().orc_ray_rect_f32(_f(a[0]), _f(a[1]), _f(a[2]), _f(a[3]), int(limit), C.byref(t0), C.byref(t1))
    return bool(hit), float(t0.value), float(t1.value)


def eigen_sum_f32(x):
    a = np.ascontiguousarray(x, dtype=np.float32)
    lib().orc_eigen_sum_f32.restype = C.c_float
    return float(lib().orc_eigen_sum_f32(_f(a), C.c_int64(a.size)))


def cam_ind_to_det_f32(ocam: OrcCamF, col, row):
    out = np.zeros(3, dtype=np.float32)
    lib().orc_cam_ind_to_det_f32(C.byref(ocam), C.c_float(col), C.c_float(row), _f(out))
    return out


def hu_to_lin_att(hu, hu_lower=-1000.0, mu_water=0.02683 * 1.0, mu_air=0.02485 * 0.0001):
    """HUToLinAtt restated (orc_hu_to_lin_att); defaults of xregHUToLinAtt.h:73-76."""
    a = np.ascontiguousarray(hu, dtype=np.float32)
    out = np.empty_like(a)
    lib().orc_hu_to_lin_att(_f(a), C.c_int64(a.size), C.c_double(hu_lower), C.c_double(mu_water),
                            C.c_double(mu_air), _f(out))
    return out


def gaussian_kernel(variance, max_error=0.01, max_width=32):
    """itk::GaussianOperator coefficients restated (orc_gaussian_kernel)."""
    k = np.zeros(1025, dtype=np.float64)
    n = lib().orc_gaussian_kernel(C.c_double(variance), C.c_double(max_error), C.c_int32(max_width),
                                  k.ctypes.data_as(C.POINTER(C.c_double)), C.c_int32(k.size))
    assert n > 0
    return k[:n].copy()


def downsample_image(img, factor, sigma=-1.0, spacing=(1.0, 1.0)):
    """DownsampleImage restated (orc_downsample_image); spacing is (row, col)."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    orows = int((a.shape[0] * factor) + 0.5)
    ocols = int((a.shape[1] * factor) + 0.5)
    out = np.empty((orows, ocols), dtype=np.float32)
    lib().orc_downsample_image(_f(a), C.c_int32(a.shape[0]), C.c_int32(a.shape[1]), C.c_double(spacing[0]),
                               C.c_double(spacing[1]), C.c_double(factor), C.c_double(sigma), _f(out),
                               C.c_int32(orows), C.c_int32(ocols))
    return out
