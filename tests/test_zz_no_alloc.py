"""xrh_line_int_noalloc (the DRR entry point that allocates nothing: fixed staging compiled
into the library) against xrh_line_int through the C-ABI: identical DRRs on two calls in a row,
the second smaller, no change in the context's device memory, and the limits refused before
any work is queued. The oracle pins xrh_line_int itself (test_gpu_drr.py); equality with it
carries that parity over."""
import numpy as np
import pytest

from tests.helpers import poses_f32

pytestmark = pytest.mark.gpu


def _cast(X, ctx, fn, vol, cams, poses, c4p, step, rows, cols, geom, kernel, store, init=None, bg=None):
    n = poses.shape[0]
    buf = ctx.alloc(4 * n * rows * cols)
    if init is not None:
        buf.upload(init)
    fn(ctx, vol, cams, poses_f32(poses), c4p, step, buf, kernel=kernel, store=store, bg=bg, geom=geom)
    ctx.synchronize()
    out = buf.download(np.float32).reshape(n, rows, cols)
    buf.free()
    return out


def test_noalloc_equals_line_int(ctx):
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    w = synth.config("c3", n_poses=6, small=True)
    rows, cols = w.cams[0].num_det_rows, w.cams[0].num_det_cols
    own = X.Context(0)  # a context of its own: its memory accounting sees only this test
    try:
        vol = X.Volume(own, w.vol, w.idx_to_phys)
        # call 1: two views x 6 poses, view-major, FP64 geometry, SUM, REPLACE
        cams = w.cams[:2]
        poses1 = np.concatenate([w.poses, w.poses])
        c4p1 = np.repeat(np.arange(2, dtype=np.int64), len(w.poses))
        print(f"case: noalloc call 1: {len(poses1)} projections, {len(cams)} cameras, {rows}x{cols}", flush=True)
        before = own.bytes_in_use()
        buf1 = own.alloc(4 * len(poses1) * rows * cols)
        X.line_int_noalloc(own, vol, cams, poses_f32(poses1), c4p1, w.step, buf1, geom=X.XRH_GEOM_FP64)
        got1 = buf1.download(np.float32).reshape(len(poses1), rows, cols)
        buf1.free()
        assert own.bytes_in_use() == before  # nothing charged: no staging was allocated
        want1 = _cast(X, own, X.line_int, vol, cams, poses1, c4p1, w.step, rows, cols, X.XRH_GEOM_FP64,
                      X.XRH_LINE_INT_SUM, X.XRH_STORE_REPLACE)
        assert got1.max() > 0
        assert np.array_equal(got1, want1)

        # call 2, smaller: one view x 3 poses, FP32 geometry, MAX, ACCUM onto a prefilled buffer
        poses2 = w.poses[:3]
        c4p2 = np.zeros(3, np.int64)
        init = np.full((3, rows, cols), 0.01, np.float32)
        print(f"case: noalloc call 2: {len(poses2)} projections, 1 camera, MAX + ACCUM", flush=True)
        got2 = _cast(X, own, X.line_int_noalloc, vol, w.cams[:1], poses2, c4p2, w.step, rows, cols,
                     X.XRH_GEOM_FP32, X.XRH_LINE_INT_MAX, X.XRH_STORE_ACCUM, init=init)
        want2 = _cast(X, own, X.line_int, vol, w.cams[:1], poses2, c4p2, w.step, rows, cols,
                      X.XRH_GEOM_FP32, X.XRH_LINE_INT_MAX, X.XRH_STORE_ACCUM, init=init)
        assert np.array_equal(got2, want2)
        assert (got2 > 0.01).any() and (got2 >= 0.01).all()

        # limits: refused with a message before anything is queued
        big = np.repeat(poses_f32(w.poses[:1]), X.XRH_NOALLOC_MAX_PROJS + 1, axis=0)
        dummy = own.alloc(4)
        with pytest.raises(X.XrhError) as ei:
            X.check(X.lib().xrh_line_int_noalloc(own.h, vol.h, X._cams_array(w.cams[:1]), 1, X._fp(big),
                                                 np.zeros(len(big), np.int64).ctypes.data_as(X._PI64), len(big),
                                                 float(w.step), 0, 0, None, 0.0, 0, dummy.ptr))
        assert ei.value.status == X.XRH_ERR_INVALID and "1024" in str(ei.value)
        dummy.free()
        vol.close()
    finally:
        own.close()
