"""xrh_downsample_image (device) against the oracle's DownsampleImage restatement."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,factor,sigma", [((384, 384), 0.5, -1.0), ((384, 384), 0.25, -1.0),
                                                ((97, 131), 0.3, -1.0), ((96, 96), 0.5, 0.0),
                                                ((64, 48), 0.5, 1.7), ((40, 40), 1.0, -1.0),
                                                ((1, 17), 0.5, -1.0), ((2, 2), 0.5, -1.0)])
def test_downsample_matches_oracle(ctx, oracle, shape, factor, sigma):
    from xreg_amd import _xrh as X
    from scipy import ndimage

    print(f"case: downsample {shape} factor={factor} sigma={sigma}", flush=True)
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    img = (ndimage.gaussian_filter(rng.random(shape), 1.0) + 0.05).astype(np.float32)
    got = X.downsample_image(ctx, img, factor, sigma=sigma, spacing=(0.776, 0.776))
    want = oracle.downsample_image(img, factor, sigma=sigma, spacing=(0.776, 0.776))
    assert got.shape == want.shape
    err = np.abs(got.astype(np.float64) - want) / max(float(np.abs(want).max()), 1e-30)
    print(f"max err / max {err.max():.3e}")
    assert err.max() <= 1e-6


def test_downsample_rejects_wrong_output_size(ctx):
    import ctypes as C

    from xreg_amd import _xrh as X

    img = np.ones((10, 10), np.float32)
    out = np.empty((4, 4), np.float32)
    st = X.lib().xrh_downsample_image(ctx.h, img.ctypes.data_as(C.POINTER(C.c_float)), 10, 10, 1.0, 1.0, 0.5, -1.0,
                                      out.ctypes.data_as(C.POINTER(C.c_float)), 4, 4)
    assert st == X.XRH_ERR_INVALID
