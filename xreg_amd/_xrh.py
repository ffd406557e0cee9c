"""ctypes binding of libxrh.so (include/xrh.h).

Loading fails loudly when the library is missing or a symbol is absent: there is no CPU
fallback behind these classes.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libxrh.so"

XRH_OK = 0
XRH_ERR_INVALID = 1
XRH_ERR_HIP = 2
XRH_ERR_OOM = 3
XRH_ERR_BUDGET = 4
XRH_ERR_UNSUPPORTED = 5
XRH_ERR_STATE = 6

XRH_F32, XRH_F16, XRH_BF16 = 0, 1, 2
XRH_LINE_INT_SUM, XRH_LINE_INT_MAX = 0, 1
XRH_STORE_REPLACE, XRH_STORE_ACCUM = 0, 1
XRH_GEOM_FP32, XRH_GEOM_FP64 = 0, 1
XRH_INTERP_LINEAR, XRH_INTERP_NN, XRH_INTERP_SINC, XRH_INTERP_BSPLINE = 0, 1, 2, 3
XRH_SIM_NCC, XRH_SIM_GRAD_NCC, XRH_SIM_PATCH_NCC, XRH_SIM_PATCH_GRAD_NCC = 0, 1, 2, 3
LAYOUT_LINEAR, LAYOUT_QUAD, LAYOUT_OCT, LAYOUT_COEF, LAYOUT_FACE, LAYOUT_BRICK = 0, 1, 2, 3, 4, 5


class XrhError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"xrh status {status}: {msg}")
        self.status = status


class Camera(C.Structure):
    """xrh_camera: CameraModel's float members (xregPerspectiveXform.h:108-273)."""

    _fields_ = [
        ("num_det_rows", C.c_int32),
        ("num_det_cols", C.c_int32),
        ("coord_frame", C.c_int32),
        ("focal_len", C.c_float),
        ("intrins_inv", C.c_float * 9),
        ("extrins_inv", C.c_float * 16),
        ("pinhole", C.c_float * 3),
        ("det_row_spacing", C.c_float),
        ("det_col_spacing", C.c_float),
    ]


class SimParams(C.Structure):
    _fields_ = [
        ("metric", C.c_int32),
        ("smooth_ksize", C.c_int32),
        ("patch_radius", C.c_int32),
        ("patch_stride", C.c_int32),
        ("patch_mask_stats", C.c_int32),
    ]


class View(C.Structure):
    _fields_ = [
        ("vol", C.c_void_p),
        ("fixed", C.c_void_p),
        ("cam", Camera),
        ("step", C.c_float),
        ("kernel", C.c_int32),
    ]


_lib = None

_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_U64 = C.c_uint64
_F = C.c_float
_PF = C.POINTER(C.c_float)
_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)
_PI32 = C.POINTER(C.c_int32)
_PU64 = C.POINTER(C.c_uint64)
_PU8 = C.POINTER(C.c_uint8)

# name -> (restype, argtypes); the full exported surface of include/xrh.h
class DrrVariant(C.Structure):
    _fields_ = [("cell_reuse", C.c_int32), ("wave_shape", C.c_int32), ("bands", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("chosen", DrrVariant), ("chosen_ms", C.c_float), ("default_ms", C.c_float),
                ("variants_timed", C.c_int32)]


SIGNATURES = {
    "xrh_last_error": (C.c_char_p, []),
    "xrh_abi_version": (_I32, []),
    "xrh_ctx_create": (_I32, [_I32, _U64, C.POINTER(_P)]),
    "xrh_ctx_destroy": (_I32, [_P]),
    "xrh_ctx_synchronize": (_I32, [_P]),
    "xrh_ctx_stream": (_P, [_P]),
    "xrh_ctx_bytes_in_use": (_U64, [_P]),
    "xrh_device_alloc": (_I32, [_P, _U64, C.POINTER(_P)]),
    "xrh_device_free": (_I32, [_P, _P]),
    "xrh_copy_h2d": (_I32, [_P, _P, _P, _U64]),
    "xrh_copy_d2h": (_I32, [_P, _P, _P, _U64]),
    "xrh_copy_d2d": (_I32, [_P, _P, _P, _U64]),
    "xrh_downsample_size": (_I32, [_I32, _I32, C.c_double, C.POINTER(_I32), C.POINTER(_I32)]),
    "xrh_downsample_image": (_I32, [_P, _PF, _I32, _I32, C.c_double, C.c_double, C.c_double, C.c_double, _PF,
                                    _I32, _I32]),
    "xrh_volume_create": (_I32, [_P, _PF, _PI64, _PF, _I32, C.POINTER(_P)]),
    "xrh_volume_create_layout": (_I32, [_P, _PF, _PI64, _PF, _I32, _I32, C.POINTER(_P)]),
    "xrh_volume_create_hu": (_I32, [_P, _PF, _PI64, _PF, _I32, _I32, C.c_double, C.c_double, C.c_double,
                                    C.POINTER(_P)]),
    "xrh_volume_destroy": (_I32, [_P]),
    "xrh_volume_bytes": (_U64, [_P]),
    "xrh_volume_info": (_I32, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "xrh_line_int": (_I32, [_P, _P, C.POINTER(Camera), _I32, _PF, _PI64, _I64, _F, _I32, _I32, _P, _F,
                            _I32, _P]),
    "xrh_line_int_noalloc": (_I32, [_P, _P, C.POINTER(Camera), _I32, _PF, _PI64, _I64, _F, _I32, _I32, _P, _F,
                                    _I32, _P]),
    "xrh_line_int_interp": (_I32, [_P, _P, C.POINTER(Camera), _I32, _PF, _PI64, _I64, _F, _I32, _I32, _P, _F,
                                   _I32, _I32, _U64, _P]),
    "xrh_line_int_sample_count": (_I32, [_P, _P, C.POINTER(Camera), _I32, _PF, _PI64, _I64, _F, _I32,
                                         _PU64, _PU64]),
    "xrh_fixed_create": (_I32, [_P, _PF, _I32, _I32, C.POINTER(SimParams), _PU8, C.POINTER(_P)]),
    "xrh_fixed_destroy": (_I32, [_P]),
    "xrh_fixed_patches": (_I32, [_P, _I64, _PI32, _PI32, _PI32, _PI32, _PF, _PI64]),
    "xrh_fixed_set_patch_weights": (_I32, [_P, _PF, _PF, _PI64, _I64, _I32, _I32]),
    "xrh_fixed_image": (_I32, [_P, _I32, _PF]),
    "xrh_sim_patch_scores": (_I32, [_P, _P, C.c_void_p, _I64, _PF, _PF]),
    "xrh_sim_compute": (_I32, [_P, _P, _P, _I64, _P, _PF]),
    "xrh_batch_create": (_I32, [_P, C.POINTER(View), _I32, _I64, _I32, C.POINTER(_P)]),
    "xrh_batch_destroy": (_I32, [_P]),
    "xrh_batch_create_empty": (_I32, [_P, _I32, _I64, _I32, _I32, _I32, C.POINTER(_P)]),
    "xrh_batch_set_views": (_I32, [_P, C.POINTER(View), _I32]),
    "xrh_batch_set_options": (_I32, [_P, _I32, _I32]),
    "xrh_batch_autotune": (_I32, [_P, _PI32, C.POINTER(C.c_double), C.c_int64, _I32, C.POINTER(TuneResult)]),
    "xrh_batch_set_drr_variant": (_I32, [_P, C.POINTER(DrrVariant)]),
    "xrh_batch_eval": (_I32, [_P, _PI32, _PD, _I64, _PF, _P]),
    "xrh_batch_eval_objects": (_I32, [_P, _PI32, _PD, _I64, _I32, _P, _PF, _P]),
    "xrh_batch_launch": (_I32, [_P, _PI32, _PD, _I64]),
    "xrh_batch_wait": (_I32, [_P, _PF]),
    "xrh_batch_last_timing": (_I32, [_P, _PF, _PF]),
    "xrh_batch_sample_count": (_I32, [_P, _PI32, _PD, _I64, _PU64, _PU64]),
    "xrh_batch_ray_steps": (_I32, [_P, _PI32, _PD, _I64, _P]),
}


def lib() -> C.CDLL:
    """Loads libxrh.so once; raises if it is missing (build with xreg_amd.build)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not found: build it with `python -m xreg_amd.build`")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)  # AttributeError if not exported
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(status: int) -> None:
    if status != XRH_OK:
        msg = lib().xrh_last_error().decode(errors="replace")
        raise XrhError(status, msg)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_PF)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_PD)


class DeviceBuffer:
    """Device memory owned by a Context (charged to its HBM budget)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _P()
        check(lib().xrh_device_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().xrh_copy_h2d(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, dtype=np.float32, count: int | None = None, offset_bytes: int = 0) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        if count is None:
            count = (self.nbytes - offset_bytes) // itemsize
        out = np.empty(count, dtype=dtype)
        check(lib().xrh_copy_d2h(self.ctx.h, out.ctypes.data, self.ptr + offset_bytes, out.nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None, dst_offset: int = 0,
                  src_offset: int = 0) -> None:
        """Device-to-device copy (xrh_copy_d2d)."""
        n = min(self.nbytes - dst_offset, other.nbytes - src_offset) if nbytes is None else int(nbytes)
        assert 0 <= n and dst_offset + n <= self.nbytes and src_offset + n <= other.nbytes
        check(lib().xrh_copy_d2d(self.ctx.h, self.ptr + dst_offset, other.ptr + src_offset, n))

    def free(self) -> None:
        if self.ptr:
            check(lib().xrh_device_free(self.ctx.h, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                lib().xrh_device_free(self.ctx.h, self.ptr)
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        p = _P()
        check(lib().xrh_ctx_create(int(device), int(hbm_budget_bytes), C.byref(p)))
        self.h = p.value

    def synchronize(self) -> None:
        check(lib().xrh_ctx_synchronize(self.h))

    def bytes_in_use(self) -> int:
        return int(lib().xrh_ctx_bytes_in_use(self.h))

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def close(self) -> None:
        if self.h:
            check(lib().xrh_ctx_destroy(self.h))
            self.h = None


# HUToLinAttFilter defaults (lib/image/xregHUToLinAtt.h:73-76)
HU_MU_WATER = 0.02683 * 1.0
HU_MU_AIR = 0.02485 * 0.0001
HU_LOWER = -1000.0


class Volume:
    def __init__(self, ctx: Context, vox_zyx: np.ndarray, idx_to_phys: np.ndarray, storage: int = XRH_F32,
                 layout: int | None = None, hu: dict | None = None):
        """hu: None for attenuation voxels, or a dict (keys hu_lower, mu_water, mu_air, any
        subset; the reference's defaults otherwise) for Hounsfield-unit voxels converted on the
        device (xrh_volume_create_hu)."""
        vox = np.ascontiguousarray(vox_zyx, dtype=np.float32)
        assert vox.ndim == 3
        size = np.array([vox.shape[2], vox.shape[1], vox.shape[0]], dtype=np.int64)
        m = np.ascontiguousarray(idx_to_phys, dtype=np.float32).reshape(16)
        p = _P()
        if hu is not None:
            check(lib().xrh_volume_create_hu(ctx.h, _fp(vox), size.ctypes.data_as(_PI64), _fp(m), int(storage),
                                             -1 if layout is None else int(layout),
                                             float(hu.get("hu_lower", HU_LOWER)),
                                             float(hu.get("mu_water", HU_MU_WATER)),
                                             float(hu.get("mu_air", HU_MU_AIR)), C.byref(p)))
        elif layout is None:
            check(lib().xrh_volume_create(ctx.h, _fp(vox), size.ctypes.data_as(_PI64), _fp(m), int(storage),
                                          C.byref(p)))
        else:
            check(lib().xrh_volume_create_layout(ctx.h, _fp(vox), size.ctypes.data_as(_PI64), _fp(m),
                                                 int(storage), int(layout), C.byref(p)))
        self.ctx = ctx
        self.h = p.value
        self.size = tuple(int(s) for s in size)
        # longest physical edge of the volume box (mm): the schedulers' per-ray work estimate
        axes = np.asarray(idx_to_phys, dtype=np.float64).reshape(4, 4)[:3, :3]
        self.extent_mm = float(max(self.size[i] * np.linalg.norm(axes[:, i]) for i in range(3)))
        st, ly = C.c_int32(0), C.c_int32(0)
        check(lib().xrh_volume_info(self.h, C.byref(st), C.byref(ly)))
        self.storage, self.layout = int(st.value), int(ly.value)

    @property
    def nbytes(self) -> int:
        return int(lib().xrh_volume_bytes(self.h))

    def close(self) -> None:
        if self.h:
            check(lib().xrh_volume_destroy(self.h))
            self.h = None


def _cams_array(cams) -> C.Array:
    arr = (Camera * len(cams))()
    for i, c in enumerate(cams):
        arr[i] = c
    return arr


def line_int(ctx: Context, vol: Volume, cams, xforms: np.ndarray, cam_for_proj: np.ndarray, step: float,
             drr: DeviceBuffer, kernel: int = XRH_LINE_INT_SUM, store: int = XRH_STORE_REPLACE,
             bg: DeviceBuffer | None = None, default_bg: float = 0.0, geom: int = XRH_GEOM_FP32) -> None:
    x = np.ascontiguousarray(xforms, dtype=np.float32).reshape(-1, 16)
    c4p = np.ascontiguousarray(cam_for_proj, dtype=np.int64)
    ca = _cams_array(cams)
    check(lib().xrh_line_int(ctx.h, vol.h, ca, len(cams), _fp(x), c4p.ctypes.data_as(_PI64), x.shape[0],
                             float(step), int(kernel), int(store), bg.ptr if bg else None, float(default_bg),
                             int(geom), drr.ptr))


XRH_NOALLOC_MAX_PROJS = 1024
XRH_NOALLOC_MAX_CAMS = 64


def line_int_noalloc(ctx: Context, vol: Volume, cams, xforms: np.ndarray, cam_for_proj: np.ndarray, step: float,
                     drr: DeviceBuffer, kernel: int = XRH_LINE_INT_SUM, store: int = XRH_STORE_REPLACE,
                     bg: DeviceBuffer | None = None, default_bg: float = 0.0, geom: int = XRH_GEOM_FP32) -> None:
    """line_int through the fixed staging area (xrh_line_int_noalloc): nothing is allocated;
    at most XRH_NOALLOC_MAX_PROJS projections and XRH_NOALLOC_MAX_CAMS cameras; synchronous."""
    x = np.ascontiguousarray(xforms, dtype=np.float32).reshape(-1, 16)
    c4p = np.ascontiguousarray(cam_for_proj, dtype=np.int64)
    ca = _cams_array(cams)
    check(lib().xrh_line_int_noalloc(ctx.h, vol.h, ca, len(cams), _fp(x), c4p.ctypes.data_as(_PI64), x.shape[0],
                                     float(step), int(kernel), int(store), bg.ptr if bg else None,
                                     float(default_bg), int(geom), drr.ptr))


XRH_MAX_AA_RAYS = 4096


def line_int_interp(ctx: Context, vol: Volume, cams, xforms: np.ndarray, cam_for_proj: np.ndarray, step: float,
                    drr: DeviceBuffer, interp: int, kernel: int = XRH_LINE_INT_SUM, store: int = XRH_STORE_REPLACE,
                    bg: DeviceBuffer | None = None, default_bg: float = 0.0, aa_rays: int = 0,
                    aa_seed: int = 0) -> None:
    """line_int with RayCaster::InterpMethod `interp` (XRH_INTERP_*) and RayCasterCPU's
    anti-aliasing factor `aa_rays` (jitter reproducible from `aa_seed`), FP32 ray set-up
    (xrh_line_int_interp). The first call on a volume that needs an interpolation grid builds it
    (synchronous, counted in the volume's bytes)."""
    x = np.ascontiguousarray(xforms, dtype=np.float32).reshape(-1, 16)
    c4p = np.ascontiguousarray(cam_for_proj, dtype=np.int64)
    ca = _cams_array(cams)
    check(lib().xrh_line_int_interp(ctx.h, vol.h, ca, len(cams), _fp(x), c4p.ctypes.data_as(_PI64), x.shape[0],
                                    float(step), int(kernel), int(store), bg.ptr if bg else None,
                                    float(default_bg), int(interp), int(aa_rays), int(aa_seed) & (2**64 - 1),
                                    drr.ptr))


def line_int_sample_count(ctx: Context, vol: Volume, cams, xforms, cam_for_proj, step: float,
                          geom: int = XRH_GEOM_FP32, evaluated: bool = False):
    """Samples of the reference algorithm (and, with evaluated=True, also the number the
    kernel evaluates after empty-space skipping)."""
    x = np.ascontiguousarray(xforms, dtype=np.float32).reshape(-1, 16)
    c4p = np.ascontiguousarray(cam_for_proj, dtype=np.int64)
    ca = _cams_array(cams)
    out, ev = C.c_uint64(0), C.c_uint64(0)
    check(lib().xrh_line_int_sample_count(ctx.h, vol.h, ca, len(cams), _fp(x), c4p.ctypes.data_as(_PI64),
                                          x.shape[0], float(step), int(geom), C.byref(out), C.byref(ev)))
    return (int(out.value), int(ev.value)) if evaluated else int(out.value)


class Fixed:
    def __init__(self, ctx: Context, img: np.ndarray, metric: int, smooth_ksize: int = 5, patch_radius: int = 5,
                 patch_stride: int = 1, mask: np.ndarray | None = None, patch_mask_stats: bool = False):
        im = np.ascontiguousarray(img, dtype=np.float32)
        prm = SimParams(int(metric), int(smooth_ksize), int(patch_radius), int(patch_stride),
                        int(bool(patch_mask_stats)))
        mk = None
        if mask is not None:
            self._mask = np.ascontiguousarray(mask, dtype=np.uint8)
            mk = self._mask.ctypes.data_as(_PU8)
        p = _P()
        check(lib().xrh_fixed_create(ctx.h, _fp(im), im.shape[0], im.shape[1], C.byref(prm), mk, C.byref(p)))
        self.ctx = ctx
        self.h = p.value
        self.shape = im.shape
        self.metric = metric

    def compute(self, mov: DeviceBuffer, num_mov: int, offset_imgs: int = 0) -> np.ndarray:
        out = np.empty(num_mov, dtype=np.float32)
        npix = self.shape[0] * self.shape[1]
        check(lib().xrh_sim_compute(self.ctx.h, self.h, mov.ptr + 4 * npix * offset_imgs, int(num_mov), None,
                                    _fp(out)))
        return out

    def patch_scores(self, mov: DeviceBuffer, num_mov: int, offset_imgs: int = 0):
        """Similarity values and every grid patch's unweighted score per moving image
        (xrh_sim_patch_scores): (num_mov,) and (num_mov, num_patches) float32."""
        n_p = C.c_int64(0)
        check(lib().xrh_fixed_patches(self.h, 0, None, None, None, None, None, C.byref(n_p)))
        sims = np.empty(num_mov, np.float32)
        scores = np.empty((num_mov, n_p.value), np.float32)
        npix = self.shape[0] * self.shape[1]
        check(lib().xrh_sim_patch_scores(self.ctx.h, self.h, mov.ptr + 4 * npix * offset_imgs, int(num_mov),
                                         _fp(sims), _fp(scores)))
        return sims, scores

    def patches(self) -> dict:
        """The patch list of a patch metric (xrh_fixed_patches): start / stop rows and columns
        (inclusive) and weights, rows of centres outer."""
        n = C.c_int64(0)
        check(lib().xrh_fixed_patches(self.h, 0, None, None, None, None, None, C.byref(n)))
        a = {k: np.empty(n.value, np.int32) for k in ("start_rows", "start_cols", "stop_rows", "stop_cols")}
        wt = np.empty(n.value, np.float32)
        check(lib().xrh_fixed_patches(self.h, n.value, a["start_rows"].ctypes.data_as(_PI32),
                                      a["start_cols"].ctypes.data_as(_PI32), a["stop_rows"].ctypes.data_as(_PI32),
                                      a["stop_cols"].ctypes.data_as(_PI32), _fp(wt), C.byref(n)))
        a["weights"] = wt
        return a

    def set_patch_weights(self, weights=None, patch_inds=None, weight_in_combine: bool = True,
                          compute_mean: bool = False, weights_y=None) -> None:
        """How a patch metric combines its patch scores (xrh_fixed_set_patch_weights): per-patch
        weights (None: the creation weights; weights_y: Patch-Grad-NCC's y direction when it
        differs), the patch indices used in order (None: all), and ImgSimMetric2DPatchCommon's
        weight_patch_sims_in_combine / compute_mean_of_patch_sims."""
        n_p = C.c_int64(0)
        check(lib().xrh_fixed_patches(self.h, 0, None, None, None, None, None, C.byref(n_p)))
        keep = []

        def as_ptr(a):
            if a is None:
                return None
            v = np.ascontiguousarray(a, dtype=np.float32).ravel()
            if v.size != n_p.value:
                raise ValueError(f"{v.size} weights for {n_p.value} patches")
            keep.append(v)
            return _fp(v)

        wp, wyp = as_ptr(weights), as_ptr(weights_y)
        ip, n = None, 0
        if patch_inds is not None:
            ind = np.ascontiguousarray(patch_inds, dtype=np.int64)
            ip, n = ind.ctypes.data_as(_PI64), ind.size
        check(lib().xrh_fixed_set_patch_weights(self.h, wp, wyp, ip, n, int(bool(weight_in_combine)),
                                                int(bool(compute_mean))))

    def image(self, which: int = 0) -> np.ndarray:
        """The fixed image as the device holds it for a patch metric (xrh_fixed_image): 0 the
        image (Patch-NCC) or its x gradient (Patch-Grad-NCC), 1 the y gradient."""
        out = np.empty(self.shape, np.float32)
        check(lib().xrh_fixed_image(self.h, int(which), _fp(out)))
        return out

    def close(self) -> None:
        if self.h:
            check(lib().xrh_fixed_destroy(self.h))
            self.h = None


class Batch:
    def __init__(self, ctx: Context, views, max_items: int, geom: int = XRH_GEOM_FP32,
                 capacity: tuple[int, int, int] | None = None):
        """views: initial view table. capacity=(max_views, max_rows, max_cols) creates an
        empty, resizable batch instead (set_views before evaluating)."""
        p = _P()
        self.ctx = ctx
        if capacity is not None:
            mv, mr, mc = capacity
            check(lib().xrh_batch_create_empty(ctx.h, int(mv), int(max_items), int(mr), int(mc), int(geom),
                                               C.byref(p)))
            self.h = p.value
            self._views = None
            if views:
                self.set_views(views)
            return
        arr = (View * len(views))()
        for i, v in enumerate(views):
            arr[i] = v
        check(lib().xrh_batch_create(ctx.h, arr, len(views), int(max_items), int(geom), C.byref(p)))
        self.h = p.value
        self._views = arr

    def set_views(self, views) -> None:
        arr = (View * len(views))()
        for i, v in enumerate(views):
            arr[i] = v
        check(lib().xrh_batch_set_views(self.h, arr, len(views)))
        self._views = arr

    def set_options(self, use_graph: bool = True, time_kernels: bool = False) -> None:
        check(lib().xrh_batch_set_options(self.h, int(use_graph), int(time_kernels)))

    def eval(self, item_view: np.ndarray, item_pose: np.ndarray, drr: DeviceBuffer | None = None) -> np.ndarray:
        iv = np.ascontiguousarray(item_view, dtype=np.int32)
        ip = np.ascontiguousarray(item_pose, dtype=np.float64).reshape(-1, 16)
        out = np.empty(iv.shape[0], dtype=np.float32)
        check(lib().xrh_batch_eval(self.h, iv.ctypes.data_as(_PI32), _dp(ip), iv.shape[0], _fp(out),
                                   drr.ptr if drr else None))
        return out

    def eval_objects(self, entry_view: np.ndarray, entry_pose: np.ndarray, bg: DeviceBuffer | None = None,
                     drr: DeviceBuffer | None = None) -> np.ndarray:
        """Several objects per item (xrh_batch_eval_objects): entry_view is (n_items, num_objs)
        view indices, entry_pose (n_items, num_objs, 4, 4) poses; object 0 of an item is cast onto
        its view's background image in `bg` (one image per view, back to back), the others added
        in order. Returns n_items similarity values."""
        ev = np.ascontiguousarray(entry_view, dtype=np.int32)
        if ev.ndim != 2:
            raise ValueError("entry_view must be (n_items, num_objs)")
        n, k = ev.shape
        ep = np.ascontiguousarray(entry_pose, dtype=np.float64).reshape(n * k, 16)
        out = np.empty(n, dtype=np.float32)
        check(lib().xrh_batch_eval_objects(self.h, ev.ctypes.data_as(_PI32), _dp(ep), n, k,
                                           bg.ptr if bg else None, _fp(out), drr.ptr if drr else None))
        return out

    def launch(self, item_view: np.ndarray, item_pose: np.ndarray) -> None:
        iv = np.ascontiguousarray(item_view, dtype=np.int32)
        ip = np.ascontiguousarray(item_pose, dtype=np.float64).reshape(-1, 16)
        check(lib().xrh_batch_launch(self.h, iv.ctypes.data_as(_PI32), _dp(ip), iv.shape[0]))
        self._launched = iv.shape[0]  # only a launch the library accepted sizes the wait

    def wait(self) -> np.ndarray:
        out = np.empty(getattr(self, "_launched", 0), dtype=np.float32)
        check(lib().xrh_batch_wait(self.h, _fp(out)))
        return out

    def autotune(self, item_view: np.ndarray, item_pose: np.ndarray, reps: int = 2) -> dict:
        """Times every DRR launch variant on this submission and keeps the fastest
        (xrh_batch_autotune)."""
        iv = np.ascontiguousarray(item_view, dtype=np.int32)
        ip = np.ascontiguousarray(item_pose, dtype=np.float64).reshape(-1, 16)
        r = TuneResult()
        check(lib().xrh_batch_autotune(self.h, iv.ctypes.data_as(_PI32), _dp(ip), iv.shape[0], int(reps),
                                       C.byref(r)))
        return {"cell_reuse": r.chosen.cell_reuse, "wave_shape": r.chosen.wave_shape, "bands": r.chosen.bands,
                "chosen_ms": float(r.chosen_ms), "default_ms": float(r.default_ms),
                "variants_timed": int(r.variants_timed)}

    def set_drr_variant(self, cell_reuse: int | None = None, wave_shape: int = 0, bands: int = 8) -> None:
        """Forces a DRR launch variant; cell_reuse=None returns to the tuned / default choice."""
        if cell_reuse is None:
            check(lib().xrh_batch_set_drr_variant(self.h, None))
        else:
            v = DrrVariant(int(cell_reuse), int(wave_shape), int(bands))
            check(lib().xrh_batch_set_drr_variant(self.h, C.byref(v)))

    def last_timing(self) -> tuple[float, float]:
        a, t = C.c_float(0), C.c_float(0)
        check(lib().xrh_batch_last_timing(self.h, C.byref(a), C.byref(t)))
        return float(a.value), float(t.value)

    def sample_count(self, item_view: np.ndarray, item_pose: np.ndarray, evaluated: bool = False):
        iv = np.ascontiguousarray(item_view, dtype=np.int32)
        ip = np.ascontiguousarray(item_pose, dtype=np.float64).reshape(-1, 16)
        out, ev = C.c_uint64(0), C.c_uint64(0)
        check(lib().xrh_batch_sample_count(self.h, iv.ctypes.data_as(_PI32), _dp(ip), iv.shape[0],
                                           C.byref(out), C.byref(ev)))
        return (int(out.value), int(ev.value)) if evaluated else int(out.value)

    def ray_steps(self, item_view: np.ndarray, item_pose: np.ndarray, out: DeviceBuffer) -> None:
        """Every ray's last sample index n (-1: miss) into `out` (int32, item order, rows * cols per
        item), under the batch's geometry mode (xrh_batch_ray_steps)."""
        iv = np.ascontiguousarray(item_view, dtype=np.int32)
        ip = np.ascontiguousarray(item_pose, dtype=np.float64).reshape(-1, 16)
        check(lib().xrh_batch_ray_steps(self.h, iv.ctypes.data_as(_PI32), _dp(ip), iv.shape[0], out.ptr))

    def close(self) -> None:
        if self.h:
            check(lib().xrh_batch_destroy(self.h))
            self.h = None


def make_view(vol: Volume, fixed: Fixed, cam: Camera, step: float, kernel: int = XRH_LINE_INT_SUM) -> View:
    return View(vol.h, fixed.h, cam, float(step), int(kernel))


def gpu_available() -> bool:
    """True when a HIP device is visible (checked without initialising the GPU runtime in this
    process more than the library itself would)."""
    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "":
        return False
    return Path("/dev/kfd").exists()


def downsample_size(rows: int, cols: int, factor: float) -> tuple[int, int]:
    r, c = _I32(), _I32()
    check(lib().xrh_downsample_size(int(rows), int(cols), float(factor), C.byref(r), C.byref(c)))
    return r.value, c.value


def downsample_image(ctx: Context, img: np.ndarray, factor: float, sigma: float = -1.0,
                     spacing: tuple[float, float] = (1.0, 1.0)) -> np.ndarray:
    """DownsampleImage of a fixed image on the device (xrh_downsample_image); spacing is
    (row, col) pixel spacing."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    orows, ocols = downsample_size(a.shape[0], a.shape[1], factor)
    out = np.empty((orows, ocols), dtype=np.float32)
    check(lib().xrh_downsample_image(ctx.h, _fp(a), a.shape[0], a.shape[1], float(spacing[0]), float(spacing[1]),
                                     float(factor), float(sigma), _fp(out), orows, ocols))
    return out
